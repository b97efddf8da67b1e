"""The oracle reproduces its committed full-size hierarchy digests (CPU only).

tests/golden/hierarchy_digests.json was written by
tests/golden/make_hierarchy_digests.py; recomputing the smaller cases here
checks that the oracle's multi-threaded loops give the serial loops' arrays
bit for bit, whatever the thread count (run with 1 and several threads).
"""
import json
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

SCRIPT = r'''
import json, os, sys
sys.path.insert(0, {root!r}); sys.path.insert(0, os.path.join({root!r}, "tests", "golden"))
from hierarchy_digest import hierarchy_digests
from make_hierarchy_digests import CASES, PARAMS
from oracle import oracle as O
A = CASES[{name!r}]()
H = O.Hierarchy(A, max_row_sum=0.9, **PARAMS.get({name!r}, {{}}))
get = lambda kind, lev: {{"A": H.A, "P": H.P, "R": H.R, "S": H.S, "CF": H.CF, "l1": H.l1}}[kind](lev)
print(json.dumps(hierarchy_digests(H.num_levels, get)))
'''


@pytest.mark.parametrize("name,threads", [("lap27_64", 1), ("lap27_64", 5), ("jump27_96", 8),
                                          ("aniso7_96", 3), ("jump27_64_hmis", 4)])
def test_oracle_matches_committed_digests(oracle, name, threads):
    with open(os.path.join(HERE, "golden", "hierarchy_digests.json")) as f:
        ref = json.load(f)[name]["levels"]
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    out = subprocess.run([sys.executable, "-c", SCRIPT.format(root=ROOT, name=name)], env=env,
                         capture_output=True, text=True, check=True).stdout
    assert json.loads(out.strip().splitlines()[-1]) == ref
