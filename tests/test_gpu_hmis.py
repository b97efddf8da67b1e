"""HMIS coarsening (coarsen_type 10) on the device against the oracle.

The Ruge-Stueben first pass runs on the host (as the reference's own device
builds run it, src/parcsr_ls/par_coarsen.c:2842-2880), the PMIS rounds with
CF_init 1 on the device.  Bar: C/F splits, round counts and whole
hierarchies bit-exact, PCG iterations within +-1.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-hmis]", *a, flush=True)
    sys.stdout.flush()


def host_s(hx, S):
    return hx.HostCsr(S.nrows, S.ncols, np.ascontiguousarray(S.ia, np.int32),
                      np.ascontiguousarray(S.ja, np.int32), None)


GEN = {
    "l7": lambda O: O.gen_laplace7(20, 19, 18),
    "l27": lambda O: O.gen_laplace27(18, 18, 18),
    "aniso": lambda O: O.gen_aniso7(20, 0.01, 1.1),
    "jump": lambda O: O.gen_vardiff27(20, 4, -4.0, 4.0, 7),
}


@pytest.mark.parametrize("gen", sorted(GEN))
def test_hmis_split_bit_exact(hx, oracle, gen):
    say("split", gen)
    S = oracle.strength(GEN[gen](oracle))
    cf_r, rounds_r = oracle.hmis(S)
    cf_g, rounds_g = hx.hmis(host_s(hx, S))
    assert np.array_equal(cf_g, cf_r)
    assert rounds_g == rounds_r
    m = oracle.pmis_measure(S, 11)
    cf_r, rounds_r = oracle.hmis(S, m)
    cf_g, rounds_g = hx.hmis(host_s(hx, S), m)
    assert np.array_equal(cf_g, cf_r) and rounds_g == rounds_r


@pytest.mark.parametrize("gen", sorted(GEN))
def test_hmis_hierarchy_bit_exact_and_pcg(hx, oracle, gen):
    say("hierarchy", gen)
    A_or = GEN[gen](oracle)
    A = hx.ParCSRMatrix.from_csr(A_or.ia, A_or.ja, A_or.a)
    H = oracle.Hierarchy(A_or, coarsen_type=10)
    n = A_or.nrows
    bh = 2 * oracle.rand_stream(5, n) - 1
    b, x = hx.ParVector(n, bh), hx.ParVector(n)
    amg = hx.BoomerAMG(coarsen_type=10)
    pcg = hx.PCG(amg, tol=1e-8)
    pcg.setup(A, b, x)
    assert amg.num_levels() == H.num_levels
    for lev in range(H.num_levels):
        G, R = amg.level_matrix(lev, "A"), H.A(lev)
        assert np.array_equal(G.ia, R.ia) and np.array_equal(G.ja, R.ja), f"A{lev}"
        assert np.array_equal(G.a, R.a), f"A{lev} values"
        if lev < H.num_levels - 1:
            assert np.array_equal(amg.level_cf(lev), H.CF(lev)), f"CF{lev}"
            P, Pr = amg.level_matrix(lev, "P"), H.P(lev)
            assert np.array_equal(P.ja, Pr.ja) and np.array_equal(P.a, Pr.a), f"P{lev}"
    pcg.solve(A, b, x)
    _, it_ref, _, _, _ = H.pcg(bh, tol=1e-8)
    assert abs(pcg.num_iterations() - it_ref) <= 1
    r = bh - A_or.to_scipy() @ x.get()
    assert np.linalg.norm(r) / np.linalg.norm(bh) < 1e-8
