"""Benchmark: AMG-PCG setup+solve seconds to 1e-8 FP64 relative residual and
the HBM bandwidth of the fine-level SpMV, on one MI355X.

Workload (the single-GPU configuration the metric is quoted on): 3D 27-point
Laplacian on a 200^3 grid (8.0M unknowns, 213.8M nonzeros), RHS uniform[-1,1]
from the Park-Miller stream of seed 0, x0 = 0; PMIS coarsening, extended+i
interpolation (P_max_elmts 4), strong threshold 0.25, max_row_sum 0.9,
l1-Jacobi 1 pre / 1 post sweep, V-cycle, Gaussian elimination on the coarsest
grid (<= 9 rows), two-norm PCG to tol 1e-8.  One step = HYPRE_ParCSRPCGSetup
+ HYPRE_ParCSRPCGSolve with A and b already resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "AMG-PCG setup+solve sec to 1e-8 FP64 rel. residual; SpMV HBM GB/s"


def cpu_baseline(sample_n: int = 100):
    """The CPU oracle (serial C port of the reference's host path) on a bounded
    sample of the same workload: the 27-point Laplacian at sample_n^3, setup +
    PCG solve, scaled to 200^3 by the nonzero ratio."""
    from oracle import oracle as O
    A = O.gen_laplace27(sample_n, sample_n, sample_n)
    b = 2.0 * O.rand_stream(0, A.nrows) - 1.0
    t0 = time.perf_counter()
    H = O.Hierarchy(A)
    x, it, rr, hist, err = H.pcg(b, tol=1e-8)
    t = time.perf_counter() - t0
    nnz_full = (3 * 200 - 2) ** 3
    return {
        "value": round(t * nnz_full / A.nnz, 3),
        "unit": "s (setup+solve, scaled to 200^3 by nnz)",
        "cores": 1,
        "kind": "port",
        "sample": f"oracle setup+PCG on 27pt {sample_n}^3 ({A.nrows} rows, {A.nnz} nnz): "
                  f"{t:.2f} s, {it} iterations, scaled x{nnz_full / A.nnz:.3f}",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=200, help="grid edge (27-point stencil)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fp32-cycle", action="store_true", help="FP32 V-cycle inside FP64 PCG")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU; --gpus {args.gpus} is not "
              "supported (the multi-GPU design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from hypre_amd import api as hx

    hx.initialize()
    N = args.n
    A = hx.ParCSRMatrix.laplacian27(N, N, N)
    n, nnz = A.dims()
    b = hx.ParVector(n)
    b.set_random(0)                       # 2*hypre_Rand()-1, hypre_SeedRand(0) -> seed 1
    x = hx.ParVector(n)
    amg = hx.BoomerAMG(coarsen_type=8, interp_type=6, strong_threshold=0.25, max_row_sum=0.9,
                       P_max_elmts=4, relax_type=18, num_sweeps=1, max_coarse_size=9,
                       fp32_cycle=1 if args.fp32_cycle else 0)
    pcg = hx.PCG(amg, tol=1e-8, max_iter=1000, two_norm=1)

    def step():
        x.fill(0.0)
        pcg.setup(A, b, x)
        pcg.solve(A, b, x)

    for _ in range(args.warmup):
        step()
    hx.kernel_timing(True, only="spmv_A0")
    hx.reset_kernel_timing()
    setup_t, solve_t = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        s1, s2 = pcg.times()
        setup_t.append(s1)
        solve_t.append(s2)
    t1 = time.perf_counter()
    ms_spmv, launches, bytes_spmv = hx.get_kernel_timing("spmv_A0")
    hx.kernel_timing(False)
    sec_per_step = (t1 - t0) / args.steps

    its = pcg.num_iterations()
    rel = pcg.final_rel_res()
    # true residual in FP64 on the device: r = b - A x
    r = hx.ParVector(n)
    r.set(b.get())
    A.matvec(-1.0, x, 1.0, r)
    true_rel = (r.dot(r) / b.dot(b)) ** 0.5
    nlev = amg.num_levels()
    rows = [amg.level_matrix(l, "A").nrows for l in range(nlev)]
    nnzs = [amg.level_matrix(l, "A").nnz for l in range(nlev)]

    avg_ms = ms_spmv / max(launches, 1)
    bytes_per = bytes_spmv / max(launches, 1)
    achieved = bytes_per / (avg_ms * 1e-3) / 1e9 if launches else None
    out = {
        "metric": METRIC,
        "value": round(sec_per_step, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec_per_step * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-cycle/f64" if args.fp32_cycle else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"3D 27-point Laplacian {N}^3 ({n} rows, {nnz} nnz), Dirichlet; PMIS, "
                        "ext+i (Pmx 4), theta 0.25, max_row_sum 0.9, l1-Jacobi V(1,1), GE coarsest, "
                        "two-norm PCG to 1e-8, x0 = 0, RHS uniform[-1,1] seed 0",
            "grid": N,
            "rows": n,
            "nnz": nnz,
            "parallelism": "single GPU",
        },
        "setup_s": round(float(np.median(setup_t)), 5),
        "solve_s": round(float(np.median(solve_t)), 5),
        "iterations": its,
        "final_rel_res": rel,
        "true_rel_res_fp64": true_rel,
        "levels": nlev,
        "level_rows": rows,
        "grid_complexity": round(sum(rows) / rows[0], 6),
        "operator_complexity": round(sum(nnzs) / nnzs[0], 6),
        "spmv_gbs": round(achieved, 1) if achieved else None,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_spmv<32,double,double,DOT> (PCG s = A p on the fine level)",
            "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": None,
            "algorithmic_bytes_per_launch": bytes_per,
            "avg_launch_ms": round(avg_ms, 5),
            "launches": launches,
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
