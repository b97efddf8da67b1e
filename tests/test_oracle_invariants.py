"""Pin the oracle where the reference's recorded outputs leave it open (CPU only):
published constants, closed forms, brute-force products, invariants.
Each check is chosen so that a plausible slip (wrong sign, dropped term,
transposed operand, wrong index) fails it."""
import numpy as np
import pytest
import scipy.sparse as sp


# ---------------------------------------------------------------- RNG
def test_park_miller_minimal_standard(oracle):
    # Park & Miller (CACM 31(10), 1988): from seed 1 the 10000th value is 1043618065
    u = oracle.rand_stream(1, 10000)
    assert int(round(u[-1] * 2147483647)) == 1043618065
    assert int(round(u[0] * 2147483647)) == 16807


def test_park_miller_jump_ahead_closed_form(oracle):
    m = 2147483647
    seed = 2747
    u = oracle.rand_stream(seed, 2000)
    for k in (0, 1, 7, 100, 1999):
        x = pow(16807, k + 1, m) * seed % m
        assert u[k] == x / m          # the division hypre_Rand performs


def test_seed_clamp(oracle):
    assert np.array_equal(oracle.rand_stream(0, 5), oracle.rand_stream(1, 5))


# ---------------------------------------------------------------- generators
@pytest.mark.parametrize("N", [3, 7, 12])
def test_laplace_closed_forms(oracle, N):
    A7 = oracle.gen_laplace7(N, N, N)
    assert A7.nnz == 7 * N ** 3 - 6 * N ** 2
    A27 = oracle.gen_laplace27(N, N, N)
    assert A27.nnz == (3 * N - 2) ** 3
    for A, d in ((A7, 6.0), (A27, 26.0)):
        assert np.all(A.ja[A.ia[:-1]] == np.arange(A.nrows))          # diagonal first
        assert np.all(A.a[A.ia[:-1]] == d)
        M = A.to_scipy()
        assert (M - M.T).nnz == 0
        for i in range(A.nrows):                                       # then ascending
            cols = A.ja[A.ia[i] + 1: A.ia[i + 1]]
            assert np.all(np.diff(cols) > 0)
    # interior row sums vanish, corner 7pt row sum = 3
    rs = A7.to_scipy().sum(axis=1).A1
    assert rs[(N // 2) * (N * N + N + 1)] == 0.0
    assert rs[0] == 3.0


def test_vardiff_constant_coefficient_reduces_to_laplacian(oracle):
    A = oracle.gen_vardiff27(6, 3, 0.0, 0.0, 5)       # kappa == 1 everywhere
    B = oracle.gen_laplace27(6, 6, 6)
    assert np.array_equal(A.ia, B.ia) and np.array_equal(A.ja, B.ja)
    assert np.array_equal(A.a, B.a)
    A7 = oracle.gen_vardiff7(6, 3, 0.0, 0.0, 5)
    B7 = oracle.gen_laplace7(6, 6, 6)
    assert np.array_equal(A7.ja, B7.ja) and np.array_equal(A7.a, B7.a)


# ---------------------------------------------------------------- strength
def test_strength_closed_forms(oracle):
    A = oracle.gen_laplace7(5, 5, 5)
    S = oracle.strength(A, 0.25, 1.0)
    off = sum(1 for i in range(A.nrows) for j in A.ja[A.ia[i] + 1:A.ia[i + 1]])
    assert S.nnz == off                           # all Laplacian couplings are strong
    # anisotropic: -0.1 couplings are weak against max |a_ij| = 1 at theta 0.25
    B = oracle.gen_laplace7(5, 5, 5, 1.0, 0.1, 1.0)
    S = oracle.strength(B, 0.25, 1.0)
    for i in range(B.nrows):
        cols = set(S.ja[S.ia[i]:S.ia[i + 1]])
        expect = {j for j in B.ja[B.ia[i] + 1:B.ia[i + 1]] if abs(j - i) != 5}
        assert cols == expect
    # theta 0.05 keeps the y couplings too
    S2 = oracle.strength(B, 0.05, 1.0)
    assert S2.nnz == off


def test_strength_row_sum_weakening(oracle):
    # with max_row_sum < 1 a row with |row sum| > max_row_sum*|diag| has no strong entries
    A = oracle.gen_laplace7(4, 4, 4)
    S_full = oracle.strength(A, 0.25, 1.0)
    S_w = oracle.strength(A, 0.25, 0.3)   # corner rows: |3| > 0.3*6 -> weak
    rs = A.to_scipy().sum(axis=1).A1
    for i in range(A.nrows):
        n_w = S_w.ia[i + 1] - S_w.ia[i]
        if abs(rs[i]) > 0.3 * 6.0:
            assert n_w == 0
        else:
            assert n_w == S_full.ia[i + 1] - S_full.ia[i]


# ---------------------------------------------------------------- PMIS
@pytest.mark.parametrize("gen", ["l7", "l27", "aniso"])
def test_pmis_invariants(oracle, gen):
    A = {"l7": lambda: oracle.gen_laplace7(9, 9, 9),
         "l27": lambda: oracle.gen_laplace27(8, 8, 8),
         "aniso": lambda: oracle.gen_laplace7(9, 9, 9, 1.0, 0.05, 1.0)}[gen]()
    S = oracle.strength(A, 0.25, 0.9)
    m = oracle.pmis_measure(S)
    cf, rounds = oracle.pmis(S)
    Ss = S.to_scipy()
    G = (Ss + Ss.T).tocsr()
    C = cf == 1
    # independence of C in S + S^T
    for i in np.nonzero(C)[0]:
        nb = G.indices[G.indptr[i]:G.indptr[i + 1]]
        assert not np.any(C[nb])
    # every point decided; empty S rows are SF
    assert set(np.unique(cf)) <= {1, -1, -3}
    assert np.all((cf == -3) == (np.diff(S.ia) == 0))
    # F points that influence someone (measure > 1) depend on a C point
    for i in np.nonzero(cf == -1)[0]:
        if m[i] > 1:
            assert np.any(C[S.ja[S.ia[i]:S.ia[i + 1]]])
    # measures = column counts of S + Park-Miller draws of seed 2747
    cnt = np.bincount(S.ja, minlength=S.nrows).astype(float)
    assert np.array_equal(m, cnt + oracle.rand_stream(2747, S.nrows))
    assert rounds >= 1


def test_pmis_given_measure_tiny_bruteforce(oracle):
    # a path 0-1-2-3-4 with symmetric strength and measures making 1 and 3 local maxima
    ia = np.array([0, 1, 3, 5, 7, 8], np.int32)
    ja = np.array([1, 0, 2, 1, 3, 2, 4, 3], np.int32)
    S = oracle.Csr(5, 5, ia, ja, None)
    m = np.array([1.2, 2.9, 2.1, 2.8, 1.1])
    cf, rounds = oracle.pmis(S, m)
    assert list(cf) == [-1, 1, -1, 1, -1]


# ---------------------------------------------------------------- interpolation
def test_qsort2_abs_order(oracle):
    rng = np.random.default_rng(3)
    for _ in range(50):
        n = rng.integers(1, 40)
        w = rng.choice([-2.0, -1.0, 0.5, 1.0, 3.0], size=n)
        v = np.arange(n, dtype=np.int32)
        v2, w2 = oracle.qsort2_abs(v, w)
        assert np.all(np.diff(np.abs(w2)) <= 0)                 # descending |w|
        assert sorted(v2.tolist()) == list(range(n))              # a permutation
        assert np.array_equal(w2, w[v2])                          # pairs stay together
    # hand trace of qsort.c:452 on four ties: swap(0,1) -> [1,0,2,3]; the right
    # part [1..3] swaps its middle (2) to the front -> [1,2,0,3]; nothing else moves
    v2, _ = oracle.qsort2_abs(np.arange(4, dtype=np.int32), np.ones(4))
    assert v2.tolist() == [1, 2, 0, 3]


def test_extpi_invariants(oracle):
    A = oracle.gen_laplace7(9, 9, 9)
    S = oracle.strength(A, 0.25, 0.9)
    cf, _ = oracle.pmis(S)
    P = oracle.extpi_interp_raw(A, S, cf)
    Pm = P.to_scipy()
    f2c = np.cumsum(cf == 1) - 1
    rs_A = A.to_scipy().sum(axis=1).A1
    for i in range(A.nrows):
        cols = P.ja[P.ia[i]:P.ia[i + 1]]
        vals = P.a[P.ia[i]:P.ia[i + 1]]
        if cf[i] == 1:
            assert cols.tolist() == [f2c[i]] and vals.tolist() == [1.0]
        elif cf[i] == -1:
            # C-hat within distance two of i in the strength graph
            d1 = set(S.ja[S.ia[i]:S.ia[i + 1]])
            d2 = set(d1)
            for k in d1:
                d2 |= set(S.ja[S.ia[k]:S.ia[k + 1]])
            cpts = {f2c[j] for j in d2 if cf[j] == 1}
            assert set(cols.tolist()) <= cpts
            if rs_A[i] == 0.0 and len(cols):
                assert abs(vals.sum() - 1.0) < 1e-12   # constants are interpolated exactly
    # truncation keeps <= 4 entries and the row sums
    Pt, cf2 = oracle.extpi_interp(A, S, cf, 0.0, 4)
    assert np.diff(Pt.ia).max() <= 4
    np.testing.assert_allclose(Pt.to_scipy().sum(axis=1).A1, Pm.sum(axis=1).A1, rtol=1e-13, atol=1e-15)
    # kept columns are those of the 4 largest magnitudes of the untruncated row
    for i in range(A.nrows):
        cols = P.ja[P.ia[i]:P.ia[i + 1]]
        mag = np.abs(P.a[P.ia[i]:P.ia[i + 1]])
        if len(cols) > 4:
            t = np.sort(mag)[::-1][3]
            kept = set(Pt.ja[Pt.ia[i]:Pt.ia[i + 1]].tolist())
            assert len(kept) == 4
            assert {c for c, w in zip(cols, mag) if w > t} <= kept
            assert kept <= {c for c, w in zip(cols, mag) if w >= t}


def test_transpose_and_rap_bruteforce(oracle):
    A = oracle.gen_laplace27(7, 7, 7)
    S = oracle.strength(A, 0.25, 0.9)
    cf, _ = oracle.pmis(S)
    P, _ = oracle.extpi_interp(A, S, cf, 0.0, 4)
    R = oracle.transpose(P)
    Rs, Ps, As = R.to_scipy(), P.to_scipy(), A.to_scipy()
    assert (Rs - Ps.T).nnz == 0
    for i in range(R.nrows):
        assert np.all(np.diff(R.ja[R.ia[i]:R.ia[i + 1]]) > 0)   # stable: fine rows ascending
    Ac = oracle.rap(R, A, P)
    ref = (Ps.T @ As @ Ps).toarray()
    np.testing.assert_allclose(Ac.to_scipy().toarray(), ref, rtol=1e-12, atol=1e-12)
    # pattern = symbolic product (no dropping), diagonal first
    sym = (abs(Ps).T @ abs(As) @ abs(Ps)).tocsr()
    sym.sum_duplicates()
    assert Ac.nnz == sym.nnz
    assert np.all(Ac.ja[Ac.ia[:-1]] == np.arange(Ac.nrows))


def test_l1_norms(oracle):
    A = oracle.gen_vardiff7(6, 2, -2.0, 2.0, 11)
    l1 = oracle.l1_norms(A)
    np.testing.assert_allclose(l1, abs(A.to_scipy()).sum(axis=1).A1, rtol=1e-15)


# ---------------------------------------------------------------- solve side
def test_coarse_gauss_elimination_exact(oracle):
    A = oracle.gen_laplace7(2, 2, 2)                   # 8 unknowns <= max_coarse_size 9
    H = oracle.Hierarchy(A, max_levels=1)
    b = np.arange(1.0, 9.0)
    u = H.cycle(b)
    np.testing.assert_allclose(u, np.linalg.solve(A.to_scipy().toarray(), b), rtol=1e-13)
    # PCG with an exact preconditioner stops after one iteration
    x, it, rr, hist, err = H.pcg(b, tol=1e-8)
    assert it == 1 and rr < 1e-12


def test_vcycle_is_symmetric_and_contracts(oracle):
    A = oracle.gen_laplace7(10, 10, 10)
    H = oracle.Hierarchy(A)
    rng = np.random.default_rng(0)
    f, g = rng.standard_normal(A.nrows), rng.standard_normal(A.nrows)
    Mf, Mg = H.cycle(f), H.cycle(g)
    # V(1,1) with Jacobi-type smoothing and Galerkin coarse grids is symmetric
    assert abs(f @ Mg - g @ Mf) < 1e-10 * abs(f @ Mg)
    x, it, rr, conv, err = H.amg_solve(oracle.matvec(1.0, A, np.ones(A.nrows)), tol=1e-8)
    assert err == 0 and conv < 0.5
    np.testing.assert_allclose(x, 1.0, atol=1e-6)


def test_pcg_residual_history_is_true_residual(oracle):
    A = oracle.gen_laplace27(9, 9, 9)
    H = oracle.Hierarchy(A)
    b = 2 * oracle.rand_stream(1, A.nrows) - 1
    x, it, rr, hist, err = H.pcg(b, tol=1e-8)
    r = b - A.to_scipy() @ x
    assert np.linalg.norm(r) / np.linalg.norm(b) == pytest.approx(rr, rel=1e-4)
    assert rr < 1e-8 and err == 0
