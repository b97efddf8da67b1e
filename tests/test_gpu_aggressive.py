"""Aggressive coarsening on the device against the oracle (through the C-ABI).

Levels below agg_num_levels coarsen the first PMIS split's C points again
over the distance-two strength graph (PMIS with CF_init 3) and interpolate
with modified multipass (agg_interp_type 8), src/parcsr_ls/par_amg_setup.c:
1311-1352 and :1799-1824.  Bar: every level's C/F split, P (pattern, order
and values), R and A_c bit-exact; PCG iterations within +-1 of the oracle and
the FP64 true residual below the tolerance.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-aggressive]", *a, flush=True)
    sys.stdout.flush()


GEN = {
    "l7_20": lambda O: O.gen_laplace7(20, 20, 20),
    "l27_16": lambda O: O.gen_laplace27(16, 16, 16),
    "aniso_18": lambda O: O.gen_aniso7(18, 0.02, 0.9),
    "jump27_16": lambda O: O.gen_vardiff27(16, 4, -3.0, 3.0, 5),
}


@pytest.mark.parametrize("agg,pmx", [(1, 0), (2, 0), (1, 4)])
@pytest.mark.parametrize("gen", sorted(GEN))
def test_aggressive_hierarchy_bit_exact(hx, oracle, gen, agg, pmx):
    say(gen, "agg_num_levels", agg, "agg_P_max_elmts", pmx)
    A_or = GEN[gen](oracle)
    A = hx.ParCSRMatrix.from_csr(A_or.ia, A_or.ja, A_or.a)
    H = oracle.Hierarchy(A_or, agg_num_levels=agg, agg_interp_type=8, agg_P_max_elmts=pmx)
    amg = hx.BoomerAMG(keep_setup_data=1, agg_num_levels=agg, agg_interp_type=8,
                       agg_P_max_elmts=pmx)
    amg.setup(A)
    assert amg.num_levels() == H.num_levels
    for lev in range(H.num_levels):
        G, R = amg.level_matrix(lev, "A"), H.A(lev)
        assert np.array_equal(G.ia, R.ia) and np.array_equal(G.ja, R.ja), f"A{lev} pattern"
        assert np.array_equal(G.a, R.a), f"A{lev} values"
        if lev < H.num_levels - 1:
            assert np.array_equal(amg.level_cf(lev), H.CF(lev)), f"CF{lev}"
            for kind, ref in (("P", H.P(lev)), ("R", H.R(lev))):
                M = amg.level_matrix(lev, kind)
                assert np.array_equal(M.ia, ref.ia) and np.array_equal(M.ja, ref.ja), f"{kind}{lev}"
                assert np.array_equal(M.a, ref.a), f"{kind}{lev} values"
    amg.destroy()


@pytest.mark.parametrize("gen", ["l7_20", "jump27_16"])
def test_aggressive_pcg(hx, oracle, gen):
    say("pcg", gen)
    A_or = GEN[gen](oracle)
    A = hx.ParCSRMatrix.from_csr(A_or.ia, A_or.ja, A_or.a)
    n = A_or.nrows
    bh = 2 * oracle.rand_stream(7, n) - 1
    b, x = hx.ParVector(n, bh), hx.ParVector(n)
    amg = hx.BoomerAMG(agg_num_levels=1, agg_interp_type=8)
    pcg = hx.PCG(amg, tol=1e-8)
    pcg.setup(A, b, x)
    pcg.solve(A, b, x)
    H = oracle.Hierarchy(A_or, agg_num_levels=1, agg_interp_type=8)
    _, it_ref, _, _, _ = H.pcg(bh, tol=1e-8)
    assert abs(pcg.num_iterations() - it_ref) <= 1, (pcg.num_iterations(), it_ref)
    r = bh - A_or.to_scipy() @ x.get()
    assert np.linalg.norm(r) / np.linalg.norm(bh) < 1e-8


def test_aggressive_options_outside_the_build_are_refused(hx):
    """agg_interp_type other than 8 (hypre's default 4 included) and HMIS
    with aggressive levels return HYPRE_ERROR_ARG at setup, not a different
    method."""
    A = hx.ParCSRMatrix.laplacian7(8, 8, 8)
    for kw in ({"agg_num_levels": 1, "agg_interp_type": 4},
               {"agg_num_levels": 1, "agg_interp_type": 8, "coarsen_type": 10}):
        amg = hx.BoomerAMG(**kw)
        with pytest.raises(hx.HypreError) as e:
            amg.setup(A)
        assert e.value.code & hx.ERROR_ARG
        hx.clear_errors()
        amg.destroy()
