"""Pin the CPU oracle to the reference's own known answers (CPU only).

The oracle is the checker for every GPU parity test, so it is itself checked
against what the reference printed for the same inputs: per-level sizes,
nonzeros, row-length extremes and row sums (the near-zero coarse row sums only
match if the Galerkin values are bit-identical), interpolation row sums,
complexities and the PCG residual history.
"""
import os

import numpy as np
import pytest


def seq_row_sums(M):
    out = np.empty(M.nrows)
    for i in range(M.nrows):
        s = 0.0
        for j in range(M.ia[i], M.ia[i + 1]):
            s += M.a[j]
        out[i] = s
    return out


def test_ij_30_7pt_hierarchy_and_pcg(oracle, golden):
    g = golden["ij_30_7pt"]
    A = oracle.gen_laplace7(30, 30, 30)
    H = oracle.Hierarchy(A, max_row_sum=g["max_row_sum"])
    assert H.num_levels == len(g["levels"])
    for lev, ref in enumerate(g["levels"]):
        Al = H.A(lev)
        rl = np.diff(Al.ia)
        assert Al.nrows == ref["rows"]
        assert Al.nnz == ref["nnz"]
        assert rl.min() == ref["min_row"] and rl.max() == ref["max_row"]
        rs = seq_row_sums(Al)
        # printed with 4 significant digits (par_stats.c); sums in storage order
        assert float("%.3e" % rs.min()) == pytest.approx(ref["rowsum_min"], abs=1e-300)
        assert float("%.3e" % rs.max()) == pytest.approx(ref["rowsum_max"])
    for lev, ref in enumerate(g["interp"]):
        P = H.P(lev)
        assert (P.nrows, P.ncols) == (ref["rows"], ref["cols"])
        rl = np.diff(P.ia)
        assert rl.min() == ref["min_row"] and rl.max() == ref["max_row"]
        rs = seq_row_sums(P)
        assert float("%.3e" % rs.min()) == pytest.approx(ref["rowsum_min"])
        assert float("%.3e" % rs.max()) == pytest.approx(ref["rowsum_max"])
    gc, oc = H.complexities()
    assert round(gc, 6) == g["grid_complexity"]
    assert round(oc, 6) == g["operator_complexity"]
    b = oracle.ij_rhsrand(A.nrows)
    x, it, rr, hist, err = H.pcg(b, tol=1e-8)
    assert err == 0
    assert it == g["iterations"]
    assert float("%.6e" % rr) == g["final_rel_res"]
    np.testing.assert_allclose(hist, g["history"], rtol=5e-7)


def test_default_saved_jacobi_amg_solver(oracle, golden):
    """TEST_ij/default.jobs:11 -- PMIS, ext+i without truncation, Jacobi, AMG as solver."""
    g = golden["default_saved"]
    A = oracle.gen_laplace7(10, 10, 10)
    H = oracle.Hierarchy(A, max_row_sum=1.0, P_max_elmts=0, relax_type=0)
    gc, oc = H.complexities()
    assert round(gc, 6) == g["grid_complexity"]
    assert round(oc, 6) == g["operator_complexity"]
    b = oracle.matvec(1.0, A, np.ones(A.nrows))
    x, it, rr, conv, err = H.amg_solve(b, tol=1e-8, max_iter=100)
    assert round(conv, 6) == g["avg_convergence_factor"]


def test_coarsening_27pt_loose(oracle, golden):
    """TEST_ij/coarsening.jobs:38 ran on 4 ranks (rank-dependent PMIS seeds), so
    a single-rank hierarchy only lands near its complexities."""
    g = golden["coarsening_27pt_4ranks"]
    A = oracle.gen_laplace27(15, 15, 10)
    H = oracle.Hierarchy(A, max_row_sum=1.0)
    gc, oc = H.complexities()
    assert abs(gc - g["grid_complexity"]) < 0.05
    assert abs(oc - g["operator_complexity"]) < 0.08


@pytest.mark.skipif(not os.environ.get("HX_RUN_SLOW"), reason="~2 min serial setup; set HX_RUN_SLOW=1")
def test_ij_200_27pt_hierarchy_and_pcg(oracle, golden):
    g = golden["ij_200_27pt"]
    A = oracle.gen_laplace27(200, 200, 200)
    H = oracle.Hierarchy(A, max_row_sum=g["max_row_sum"])
    assert H.num_levels == len(g["levels"])
    for lev, ref in enumerate(g["levels"]):
        Al = H.A(lev)
        rl = np.diff(Al.ia)
        assert (Al.nrows, Al.nnz, rl.min(), rl.max()) == (ref["rows"], ref["nnz"], ref["min_row"], ref["max_row"])
    gc, oc = H.complexities()
    assert (round(gc, 6), round(oc, 6)) == (g["grid_complexity"], g["operator_complexity"])
    x, it, rr, hist, err = H.pcg(oracle.ij_rhsrand(A.nrows), tol=1e-8)
    assert it == g["iterations"]
    np.testing.assert_allclose(hist, g["history"], rtol=5e-7)
