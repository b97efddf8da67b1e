"""The full form at the largest single-GPU size: 27-point jumping-coefficient
diffusion on a 320^3 grid (32,768,000 rows, 879,217,912 nonzeros; kappa
log-uniform in [1e-4, 1e4] on 16^3-node cells, seed 1; RHS uniform[-1,1],
seed 2), plain FP64 AMG-PCG against the FP32 V-cycle inside FP64 PCG with the
learned per-level parameters (the committed model, kept only if it converges
faster than the defaults on this matrix).

Both runs must reach the 1e-8 relative residual recomputed in FP64
(r = b - A x with the FP64 matrix), the mixed run within 2 iterations of
the plain one; the difference of the two solutions is reported.  The mixed
+ learned method must be faster than plain FP64, each timed as the best of
two setup+solve steps after a warm-up."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 320


def say(*a):
    print("[gpu-north-star]", *a, flush=True)


def run(hx, A, b, amg):
    pcg = hx.PCG(amg, tol=1e-8, max_iter=1000, two_norm=1)
    x = hx.ParVector(b.n)
    best = None
    for k in range(3):
        x.fill(0.0)
        t0 = time.perf_counter()
        pcg.setup(A, b, x)
        pcg.solve(A, b, x)
        t = time.perf_counter() - t0
        if k > 0:
            best = t if best is None else min(best, t)
    r = hx.ParVector(b.n)
    b.copy_to(r)
    A.matvec(-1.0, x, 1.0, r)
    rel = (r.dot(r) / b.dot(b)) ** 0.5
    its = pcg.num_iterations()
    xs = x.get()
    for h in (pcg, x, r):
        h.destroy()
    return best, its, rel, xs


def test_fp32_cycle_with_learned_parameters_beats_plain_fp64_at_320(hx):
    from hypre_amd import learned as hl
    say(f"jump27 {N}^3")
    A = hx.ParCSRMatrix.var_diffusion(N, 27, 16, -4.0, 4.0, 1)
    n, nnz = A.dims()
    assert (n, nnz) == (32_768_000, 879_217_912)
    b = hx.ParVector(n)
    b.set_random(2)

    plain = hx.BoomerAMG(max_row_sum=0.9)
    t64, it64, rel64, x64 = run(hx, A, b, plain)
    plain.destroy()
    say(f"plain FP64: {t64:.4f} s, {it64} iterations, FP64 true rel. residual {rel64:.3e}")

    model = hl.ParamModel.load()
    amg = hx.BoomerAMG(max_row_sum=0.9, fp32_cycle=1)
    which, lst, dst, decide_s = hl.apply_with_fallback(amg, model, A, b, fp32=True)
    say(f"learned decision: {which} ({lst.iterations} vs {dst.iterations} iterations, "
        f"{decide_s:.3f} s once per matrix)")
    t32, it32, rel32, x32 = run(hx, A, b, amg)
    amg.destroy()
    model.destroy()
    err = float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64))
    say(f"FP32 cycle + learned: {t32:.4f} s, {it32} iterations, FP64 true rel. residual "
        f"{rel32:.3e}, ||x - x_fp64|| / ||x_fp64|| = {err:.3e}, speedup {t64 / t32:.3f}")
    A.destroy()
    b.destroy()

    assert rel64 < 1e-8 and rel32 < 1e-8, (rel64, rel32)
    assert it32 <= it64 + 2, (it32, it64)
    # both iterates satisfy the same residual bar; their difference is bounded
    # by the conditioning of a 1e8-contrast operator, so it is only sanity-checked
    assert err < 1e-2, err
    assert t32 < t64, (t32, t64)
