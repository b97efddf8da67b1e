"""Bit-exact hierarchies at full size (through the C-ABI).

tests/golden/hierarchy_digests.json holds, per level, SHA-256 digests of the
oracle's S, C/F marker, P, R, A_c and l1 arrays for the benchmark's settings
(written by tests/golden/make_hierarchy_digests.py).  The device hierarchy of
the same matrix must produce the same digests: every C/F split, sparsity
pattern (with its within-row order), truncation tie-break and value equal bit
for bit, up to the 27-point 200^3 benchmark problem.
"""
import json
import os
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from hierarchy_digest import hierarchy_digests  # noqa: E402

with open(os.path.join(HERE, "golden", "hierarchy_digests.json")) as f:
    DIGESTS = json.load(f)

def _host(hx, A):
    return hx.ParCSRMatrix.from_csr(A.ia, A.ja, A.a)


# The Laplacian and anisotropic generators give the oracle's bits on the
# device; the jumping-coefficient generator's device pow() may differ from the
# host's in the last bit, so those matrices are built by the oracle and
# uploaded, which makes level 0 identical by construction.
MAKE = {
    "lap27_64": lambda hx, O: hx.ParCSRMatrix.laplacian27(64, 64, 64),
    "jump27_96": lambda hx, O: _host(hx, O.gen_vardiff27(96, 16, -4.0, 4.0, 1)),
    "aniso7_96": lambda hx, O: hx.ParCSRMatrix.aniso7(96, 0.01, 0.6),
    "lap27_200": lambda hx, O: hx.ParCSRMatrix.laplacian27(200, 200, 200),
    "lap27_64_hmis": lambda hx, O: hx.ParCSRMatrix.laplacian27(64, 64, 64),
    "jump27_64_hmis": lambda hx, O: _host(hx, O.gen_vardiff27(64, 16, -4.0, 4.0, 1)),
}
COARSEN = {"lap27_64_hmis": 10, "jump27_64_hmis": 10}


def first_difference(got, ref):
    for lev, (g, r) in enumerate(zip(got, ref)):
        for key in r:
            if g.get(key) != r[key]:
                return f"level {lev} {key}: {g.get(key)} != {r[key]}"
    return f"level counts {len(got)} != {len(ref)}"


@pytest.mark.parametrize("name", sorted(n for n in MAKE if n in DIGESTS))
def test_hierarchy_digests_match_oracle(hx, oracle, name):
    print(f"[gpu-digests] {name}", flush=True)
    A = MAKE[name](hx, oracle)
    amg = hx.BoomerAMG(max_row_sum=0.9, keep_setup_data=1, coarsen_type=COARSEN.get(name, 8))
    amg.setup(A)
    nl = amg.num_levels()

    def get(kind, lev):
        if kind == "CF":
            return amg.level_cf(lev)
        if kind == "l1":
            return amg.level_l1(lev)
        return amg.level_matrix(lev, kind)

    got = hierarchy_digests(nl, get)
    ref = DIGESTS[name]["levels"]
    amg.destroy()
    assert got == ref, first_difference(got, ref)
