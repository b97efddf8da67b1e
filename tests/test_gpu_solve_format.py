"""The solve phase's two kernel sets give the same bits (through the C-ABI).

Format 1 (default) runs the SpMV and l1-Jacobi sweeps on sliced-ELL copies of
the level operators and forms each coarse level's first zero-guess sweep in
the restriction's epilogue; format 0 runs the windowed CSR kernels and one
launch per sweep, the plain kernel sequence of hypre's cycle
(src/parcsr_ls/par_cycle.c:23).  Format 0 is the path the oracle parity tests
pin; here a V-cycle (FP64 and FP32) must agree bit for bit between the two,
also with per-level relaxation weights and two sweeps on coarse levels, and
PCG must take the same number of iterations.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-solve-format]", *a, flush=True)
    sys.stdout.flush()


PROBLEMS = {
    "lap27_40": lambda hx: hx.ParCSRMatrix.laplacian27(40, 40, 40),
    "lap7_48": lambda hx: hx.ParCSRMatrix.laplacian7(48, 48, 48),
    "jump27_32": lambda hx: hx.ParCSRMatrix.var_diffusion(32, 27, 8, -4.0, 4.0, 3),
    "aniso7_40": lambda hx: hx.ParCSRMatrix.aniso7(40, 0.01, 0.7),
}


def cycle(hx, A, fmt, f, fp32=False, level_params=False):
    amg = hx.BoomerAMG(max_iter=1, tol=0.0, fp32_cycle=1 if fp32 else 0, solve_format=fmt)
    if level_params:
        amg.set_level_params(omega=[0.9, 1.1, 0.8, 1.2, 0.95, 1.05], sweeps=[1, 2, 2, 1, 2, 2])
    amg.setup(A)
    u = amg.cycle_host(f)
    amg.destroy()
    return u


@pytest.mark.parametrize("name", sorted(PROBLEMS))
@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("level_params", [False, True])
def test_vcycle_bits_equal_across_formats(hx, oracle, name, fp32, level_params):
    say(name, "fp32" if fp32 else "fp64", "level params" if level_params else "defaults")
    A = PROBLEMS[name](hx)
    n, _ = A.dims()
    f = 2 * oracle.rand_stream(7, n) - 1
    u0 = cycle(hx, A, 0, f, fp32, level_params)
    u1 = cycle(hx, A, 1, f, fp32, level_params)
    assert np.isfinite(u0).all() and np.abs(u0).max() > 0
    bad = np.nonzero(u0.view(np.int64) != u1.view(np.int64))[0]
    assert bad.size == 0, f"{bad.size} entries differ, first at {bad[:5]}"


@pytest.mark.parametrize("name", ["lap27_40", "jump27_32"])
def test_pcg_iterations_equal_across_formats(hx, oracle, name):
    say("pcg", name)
    A = PROBLEMS[name](hx)
    n, _ = A.dims()
    bh = 2 * oracle.rand_stream(11, n) - 1
    out = []
    for fmt in (0, 1):
        b = hx.ParVector(n, bh)
        x = hx.ParVector(n)
        amg = hx.BoomerAMG(solve_format=fmt)
        pcg = hx.PCG(amg, tol=1e-8)
        pcg.setup(A, b, x)
        pcg.solve(A, b, x)
        out.append((pcg.num_iterations(), np.array(pcg.history()), x.get()))
    (i0, h0, x0), (i1, h1, x1) = out
    assert i0 == i1
    # the fused PCG dot sums its partials in a different order: last bits only
    np.testing.assert_allclose(h1, h0, rtol=1e-9)
    np.testing.assert_allclose(x1, x0, rtol=1e-8, atol=1e-12 * np.abs(x0).max())


def mixed_mode_matrix(n=80_000, far=70_000):
    """Symmetric, diagonally dominant: a chain, plus couplings i <-> i + far
    for even i < n - far, i <-> i + 2 for odd i and i <-> i + 1000 for every
    third i in [20001, 30000); rows sorted diagonal first, then ascending.  Its 64-row slices include steps whose col - row
    span is 0 (8-bit offsets), up to a few hundred (16-bit) and about 70000
    (32-bit columns)."""
    import scipy.sparse as sp
    rows, cols = [], []
    i = np.arange(n - 1)
    rows += [i, i + 1]
    cols += [i + 1, i]
    e = np.arange(0, n - far, 2)
    rows += [e, e + far]
    cols += [e + far, e]
    o = np.arange(1, n - 2, 2)
    rows += [o, o + 2]
    cols += [o + 2, o]
    m = np.arange(20_001, 30_000, 3)
    rows += [m, m + 1000]
    cols += [m + 1000, m]
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    M = sp.csr_matrix((-np.ones(len(r)), (r, c)), shape=(n, n))
    M.sum_duplicates()
    d = -np.asarray(M.sum(axis=1)).ravel() + 1.0
    M = (M + sp.diags(d)).tocsr()
    M.sort_indices()
    ia, ja, a = [0], [], []
    for k in range(n):
        cs = M.indices[M.indptr[k]:M.indptr[k + 1]]
        vs = M.data[M.indptr[k]:M.indptr[k + 1]]
        order = sorted(range(len(cs)), key=lambda q: (cs[q] != k, cs[q]))
        ja += [int(cs[q]) for q in order]
        a += [float(vs[q]) for q in order]
        ia.append(len(ja))
    return np.array(ia, np.int32), np.array(ja, np.int32), np.array(a)


def slice_modes(ia, ja):
    """Each 64-row slice's column format as the setup chooses it: 2 (8-bit)
    when every step spans <= 255 in col - row, 1 (16-bit) when <= 65535,
    else 0."""
    n = len(ia) - 1
    modes = []
    for s0 in range(0, n, 64):
        rr = range(s0, min(n, s0 + 64))
        w = max(ia[r + 1] - ia[r] for r in rr)
        span = 0
        for k in range(w):
            offs = [int(ja[ia[r] + k]) - r for r in rr if k < ia[r + 1] - ia[r]]
            span = max(span, max(offs) - min(offs))
        modes.append(2 if span <= 255 else (1 if span <= 65535 else 0))
    return np.array(modes)


def test_vcycle_bits_equal_across_formats_all_column_widths(hx, oracle):
    """A matrix whose fine-level slices use all three column formats of the
    sliced-ELL copy: the V-cycle is bit-equal to the CSR kernels'."""
    say("mixed column widths")
    ia, ja, a = mixed_mode_matrix()
    modes = slice_modes(ia, ja)
    assert {0, 1, 2} <= set(modes.tolist()), np.bincount(modes)
    A = hx.ParCSRMatrix.from_csr(ia, ja, a)
    f = 2 * oracle.rand_stream(9, len(ia) - 1) - 1
    for fp32 in (False, True):
        u0 = cycle(hx, A, 0, f, fp32)
        u1 = cycle(hx, A, 1, f, fp32)
        assert np.isfinite(u0).all() and np.abs(u0).max() > 0
        bad = np.nonzero(u0.view(np.int64) != u1.view(np.int64))[0]
        assert bad.size == 0, f"{bad.size} entries differ (fp32={fp32}), first at {bad[:5]}"


def banded_zones_matrix(n=20_000):
    """Symmetric, diagonally dominant: row i couples to i + d for d = 1 in
    [0, 6400), d = 1..40 in [6400, 12800) and d = 1..90 beyond, so that its
    64-row slices are at most 32 wide, between 33 and 128, and wider than
    128."""
    import scipy.sparse as sp
    rows, cols = [], []
    for lo, hi, dmax in ((0, 6400, 1), (6400, 12800, 40), (12800, n, 90)):
        for d in range(1, dmax + 1):
            i = np.arange(lo, min(hi, n - d))
            rows += [i, i + d]
            cols += [i + d, i]
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    M = sp.csr_matrix((-np.ones(len(r)), (r, c)), shape=(n, n))
    M.sum_duplicates()
    d = -np.asarray(M.sum(axis=1)).ravel() + 1.0
    M = (M + sp.diags(d)).tocsr()
    M.sort_indices()
    # diagonal first, then ascending columns
    ia = M.indptr.astype(np.int32)
    ja = M.indices.astype(np.int32).copy()
    a = M.data.copy()
    for k in range(n):
        s, e = ia[k], ia[k + 1]
        cs = ja[s:e]
        p = int(np.nonzero(cs == k)[0][0])
        order = np.r_[p, np.arange(0, p), np.arange(p + 1, e - s)]
        ja[s:e] = cs[order]
        a[s:e] = a[s:e][order]
    return ia, ja, a


def test_vcycle_bits_equal_across_formats_all_slice_widths(hx, oracle):
    """Slices up to 32 wide, 33 to 128 and wider than 128 at the fine level:
    the V-cycle is bit-equal to the CSR kernels'."""
    say("slice widths")
    ia, ja, a = banded_zones_matrix()
    lens = np.diff(ia)
    w = np.array([lens[s:s + 64].max() for s in range(0, len(lens), 64)])
    assert (w <= 32).any() and ((w > 32) & (w <= 128)).any() and (w > 128).any()
    A = hx.ParCSRMatrix.from_csr(ia, ja, a)
    f = 2 * oracle.rand_stream(13, len(ia) - 1) - 1
    for fp32 in (False, True):
        u0 = cycle(hx, A, 0, f, fp32)
        u1 = cycle(hx, A, 1, f, fp32)
        assert np.isfinite(u0).all() and np.abs(u0).max() > 0
        bad = np.nonzero(u0.view(np.int64) != u1.view(np.int64))[0]
        assert bad.size == 0, f"{bad.size} entries differ (fp32={fp32}), first at {bad[:5]}"
