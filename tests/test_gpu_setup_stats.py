"""The setup statistics table printed by the library (print level 1) against
the oracle's (src/parcsr_ls/par_stats.c:20), through the C-ABI.

HYPRE_BoomerAMGSetup with print level 1 prints the reference's table after
setup (par_amg_setup.c:3830).  The hierarchies are bit-exact with the
oracle's, so the text must be the same byte for byte: the device pass that
forms each level's entries per row, row sums and interpolation weights, and
the host formatting, are both checked.  The oracle's text is itself pinned to
the reference's recorded ij run (tests/test_oracle_setup_stats.py).
"""
import sys

import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-setup-stats]", *a, flush=True)
    sys.stdout.flush()


CASES = {
    "lap7_30_ij": (lambda O: O.gen_laplace7(30, 30, 30), {"max_row_sum": 1.0}),
    "lap27_24": (lambda O: O.gen_laplace27(24, 24, 24), {}),
    "jump27_20": (lambda O: O.gen_vardiff27(20, 4, -4.0, 4.0, 1), {}),
    "hmis_lap7_20": (lambda O: O.gen_laplace7(20, 20, 20), {"coarsen_type": 10}),
    "agg1_lap27_18": (lambda O: O.gen_laplace27(18, 18, 18),
                      {"agg_num_levels": 1, "agg_interp_type": 8}),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_setup_stats_text_equals_oracle(hx, oracle, capfd, name):
    gen, prm = CASES[name]
    say(name, prm)
    A_or = gen(oracle)
    H = oracle.Hierarchy(A_or, **prm)
    expect = H.setup_stats(max_iter=1, tol=0.0)
    A = hx.ParCSRMatrix.from_csr(A_or.ia, A_or.ja, A_or.a)
    amg = hx.BoomerAMG(max_iter=1, tol=0.0, **prm)
    capfd.readouterr()
    amg.set(print_level=1)
    amg.setup(A)
    out = capfd.readouterr().out
    amg.destroy()
    assert out == expect, "\n--- device ---\n" + out + "\n--- oracle ---\n" + expect


def test_setup_stats_silent_at_print_level_0(hx, oracle, capfd):
    say("print level 0")
    A_or = oracle.gen_laplace7(12, 12, 12)
    A = hx.ParCSRMatrix.from_csr(A_or.ia, A_or.ja, A_or.a)
    amg = hx.BoomerAMG()
    capfd.readouterr()
    amg.setup(A)
    out = capfd.readouterr().out
    amg.destroy()
    assert out == ""
