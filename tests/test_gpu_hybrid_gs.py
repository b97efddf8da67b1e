"""Hybrid Gauss-Seidel smoothers (relax 3/4/6/8/13/14) on the GPU against the
oracle: one relaxation bit for bit (every type, weights, partition counts from
one partition to one row per partition and beyond, zero initial guess), the
V-cycle and PCG with the reference's host default pairing (13 down, 14 up)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-hgs]", *a, flush=True)
    sys.stdout.flush()


def to_host(hx, M):
    return hx.HostCsr(M.nrows, M.ncols, np.ascontiguousarray(M.ia, np.int32),
                      np.ascontiguousarray(M.ja, np.int32), np.ascontiguousarray(M.a, np.float64))


TYPES = {3: (1, False, 1, False), 4: (-1, False, 1, False), 6: (1, True, 1, False),
         13: (1, False, 0, True), 14: (-1, False, 0, True), 8: (1, True, 0, True)}


def oracle_sweep(oracle, A, f, u, rt, parts, l1, w, om):
    direction, symm, skip, uses_l1 = TYPES[rt]
    if uses_l1 and not (w == 1.0 and om == 1.0):
        skip = 1
    return oracle.hybrid_gs(A, f, u, parts, direction, symm, skip, l1 if uses_l1 else None, w, om)


def problems(O):
    return {
        "lap7_14": lambda: O.gen_laplace7(14, 14, 14),
        "lap27_12": lambda: O.gen_laplace27(12, 12, 12),
        "vardiff27_12": lambda: O.gen_vardiff27(12, 3, -3.0, 3.0, 7),
    }


@pytest.mark.parametrize("rt", [13, 14, 8, 3, 4, 6])
def test_sweep_bit_exact(hx, oracle, rt):
    for name, gen in problems(oracle).items():
        A = gen()
        n = A.nrows
        l1 = oracle.l1_norms_diag(A)
        f = 2 * oracle.rand_stream(11, n) - 1
        u = 2 * oracle.rand_stream(12, n) - 1
        for parts in (1, 3, 64, 997, n, n + 5):
            for (w, om) in ((1.0, 1.0), (0.7, 1.0), (0.8, 1.3)):
                for zero in (False, True):
                    say("sweep", rt, name, parts, w, om, zero)
                    u0 = np.zeros(n) if zero else u
                    want = oracle_sweep(oracle, A, f, u0, rt, parts, l1, w, om)
                    got = hx.hybrid_gs(to_host(hx, A), f, u, rt, parts, l1, w, om, zero)
                    bad = np.nonzero(got != want)[0]
                    assert bad.size == 0, (rt, name, parts, w, om, zero, bad[:5],
                                           got[bad[:5]], want[bad[:5]])


def test_sweep_edge_cases(hx, oracle):
    say("edge cases")
    # one row; a zero divisor row is left unchanged; rows without a diagonal
    A = oracle.Csr(1, 1, np.array([0, 1], np.int32), np.array([0], np.int32), np.array([4.0]))
    got = hx.hybrid_gs(to_host(hx, A), np.array([2.0]), np.array([1.0]), 13, 1, np.array([4.0]))
    assert got[0] == oracle.hybrid_gs(A, [2.0], [1.0], 1, 1, False, 0, np.array([4.0]))[0]
    B = oracle.gen_laplace7(6, 5, 4)
    l1 = oracle.l1_norms_diag(B)
    l1[[3, 17, 40]] = 0.0
    f = 2 * oracle.rand_stream(4, B.nrows) - 1
    u = 2 * oracle.rand_stream(5, B.nrows) - 1
    for rt in (13, 14, 8):
        for parts in (1, 7, 200):
            got = hx.hybrid_gs(to_host(hx, B), f, u, rt, parts, l1)
            want = oracle_sweep(oracle, B, f, u, rt, parts, l1, 1.0, 1.0)
            assert np.array_equal(got, want), (rt, parts)
            assert got[3] == u[3] and got[17] == u[17] and got[40] == u[40]


def test_sweep_bit_exact_large(hx, oracle):
    say("sweep 27pt 64^3")
    A = oracle.gen_laplace27(64, 64, 64)
    n = A.nrows
    l1 = oracle.l1_norms_diag(A)
    f = 2 * oracle.rand_stream(21, n) - 1
    u = 2 * oracle.rand_stream(22, n) - 1
    for rt, parts in ((13, 4096), (14, 4096), (8, (n + 63) // 64)):
        got = hx.hybrid_gs(to_host(hx, A), f, u, rt, parts, l1)
        want = oracle_sweep(oracle, A, f, u, rt, parts, l1, 1.0, 1.0)
        assert np.array_equal(got, want), rt


@pytest.mark.parametrize("down,up,parts", [(13, 14, 0), (13, 14, 16), (8, 8, 0), (3, 4, 5)])
def test_vcycle_matches_oracle(hx, oracle, down, up, parts):
    say("v-cycle", down, up, parts)
    A_or = oracle.gen_vardiff27(14, 4, -2.0, 2.0, 5)
    A_dev = hx.ParCSRMatrix.from_csr(A_or.ia, A_or.ja, A_or.a)
    amg = hx.BoomerAMG(max_iter=1, tol=0.0, relax_type=down, relax_up=up,
                       hybrid_partitions=(parts, 8))
    amg.setup(A_dev)
    H = oracle.Hierarchy(A_or, relax_type=down, relax_up_type=up, hgs_parts=parts, hgs_rows=8)
    assert amg.num_levels() == H.num_levels
    f = 2 * oracle.rand_stream(3, A_or.nrows) - 1
    ug = amg.cycle_host(f)
    ur = H.cycle(f)
    np.testing.assert_allclose(ug, ur, rtol=1e-11, atol=1e-13 * np.abs(ur).max())
    for lev in range(H.num_levels - 1):
        if down in (8, 13, 14):
            assert np.array_equal(amg.level_l1(lev), H.l1(lev)), lev


@pytest.mark.parametrize("gen", ["lap27_30", "vardiff7_24"])
def test_pcg_iterations_match_oracle(hx, oracle, gen):
    say("pcg hybrid GS", gen)
    if gen == "lap27_30":
        A_or = oracle.gen_laplace27(30, 30, 30)
    else:
        A_or = oracle.gen_vardiff7(24, 4, -3.0, 3.0, 9)
    n = A_or.nrows
    b_h = oracle.ij_rhsrand(n)
    A = hx.ParCSRMatrix.from_csr(A_or.ia, A_or.ja, A_or.a)
    b = hx.ParVector(n, b_h)
    x = hx.ParVector(n)
    amg = hx.BoomerAMG(max_row_sum=1.0, relax_type=13, relax_up=14)
    pcg = hx.PCG(amg, tol=1e-8, max_iter=200, two_norm=1)
    pcg.setup(A, b, x)
    pcg.solve(A, b, x)
    H = oracle.Hierarchy(A_or, max_row_sum=1.0, relax_type=13, relax_up_type=14)
    _, it_or, rr_or, hist_or, err = H.pcg(b_h, 1e-8, 200)
    assert err == 0
    assert abs(pcg.num_iterations() - it_or) <= 1, (pcg.num_iterations(), it_or)
    np.testing.assert_allclose(pcg.history()[:3], hist_or[:3], rtol=1e-8)
    r = b_h - A_or.to_scipy() @ x.get()
    assert np.linalg.norm(r) / np.linalg.norm(b_h) < 1e-8


@pytest.mark.parametrize("down,up", [(13, 14), (8, 8)])
def test_fp32_cycle_with_hybrid_gs(hx, oracle, down, up):
    # the FP32 V-cycle (hybrid GS sweeps in FP32, row-reversed FP32 copies)
    # inside FP64 PCG reaches the FP64 tolerance within one iteration of FP64
    say("fp32 cycle hybrid GS", down, up)
    A_or = oracle.gen_laplace27(30, 30, 30)
    n = A_or.nrows
    b_h = oracle.ij_rhsrand(n)
    A = hx.ParCSRMatrix.from_csr(A_or.ia, A_or.ja, A_or.a)
    its = {}
    for fp32 in (0, 1):
        b = hx.ParVector(n, b_h)
        x = hx.ParVector(n)
        amg = hx.BoomerAMG(max_row_sum=1.0, relax_type=down, relax_up=up, fp32_cycle=fp32)
        pcg = hx.PCG(amg, tol=1e-8, max_iter=200, two_norm=1)
        pcg.setup(A, b, x)
        pcg.solve(A, b, x)
        r = b_h - A_or.to_scipy() @ x.get()
        assert np.linalg.norm(r) / np.linalg.norm(b_h) < 1e-8
        its[fp32] = pcg.num_iterations()
    assert its[1] <= its[0] + 1, its
