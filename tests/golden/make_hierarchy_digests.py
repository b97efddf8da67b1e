"""Write hierarchy_digests.json: the oracle's hierarchies at sizes too large
to commit as arrays, as per-level SHA-256 digests (hierarchy_digest.py).

    python tests/golden/make_hierarchy_digests.py          # about 3 minutes

Settings are the benchmark's (PMIS, or HMIS for the *_hmis cases, with the
Park-Miller measure stream seed 2747, ext+i with P_max_elmts 4, theta 0.25, max_row_sum 0.9, l1-Jacobi norms,
max_coarse_size 9).  The oracle (oracle/amg_oracle.c) is pinned by the
reference's recorded runs; these digests carry its exact arrays to the GPU
parity test tests/test_gpu_full_size_digests.py.
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from hierarchy_digest import hierarchy_digests  # noqa: E402
from oracle import oracle as O  # noqa: E402

CASES = {
    "lap27_64": lambda: O.gen_laplace27(64, 64, 64),
    "jump27_96": lambda: O.gen_vardiff27(96, 16, -4.0, 4.0, 1),
    "aniso7_96": lambda: O.gen_aniso7(96, 0.01, 0.6),
    "lap27_200": lambda: O.gen_laplace27(200, 200, 200),
    "lap27_64_hmis": lambda: O.gen_laplace27(64, 64, 64),
    "jump27_64_hmis": lambda: O.gen_vardiff27(64, 16, -4.0, 4.0, 1),
}
# settings beyond the benchmark's, per case
PARAMS = {"lap27_64_hmis": {"coarsen_type": 10}, "jump27_64_hmis": {"coarsen_type": 10}}


def main(names):
    out_path = os.path.join(HERE, "hierarchy_digests.json")
    res = {}
    if os.path.exists(out_path):
        with open(out_path) as f:
            res = json.load(f)
    for name in names:
        t0 = time.time()
        A = CASES[name]()
        extra = PARAMS.get(name, {})
        H = O.Hierarchy(A, max_row_sum=0.9, **extra)

        def get(kind, lev):
            return {"A": H.A, "P": H.P, "R": H.R, "S": H.S, "CF": H.CF, "l1": H.l1}[kind](lev)

        coarsen = "HMIS" if extra.get("coarsen_type") == 10 else "PMIS"
        res[name] = {"settings": f"{coarsen}, ext+i Pmx 4, theta 0.25, max_row_sum 0.9, l1 norms, "
                                 "max_coarse_size 9",
                     "levels": hierarchy_digests(H.num_levels, get)}
        print(f"{name}: {H.num_levels} levels, {time.time() - t0:.1f} s", flush=True)
        del H
        with open(out_path, "w") as f:
            json.dump(res, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:] or list(CASES))
