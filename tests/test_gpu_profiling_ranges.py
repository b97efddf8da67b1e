"""Profiling ranges named as the reference's (src/utilities/device_markers.c)
through roctx, through the C-ABI.

With ranges on, setup and solve open and close them in pairs (none left
open), and the results are the same bits as with ranges off: closing a range
only waits for the stream.  The range names and their nesting are recorded
with `rocprofv3 --marker-trace` (profiles/amg_pcg_27pt_roctx_ranges.md).
"""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-ranges]", *a, flush=True)
    sys.stdout.flush()


def solve(hx, A, b):
    amg = hx.BoomerAMG()
    pcg = hx.PCG(amg, tol=1e-8)
    x = hx.ParVector(b.n)
    pcg.setup(A, b, x)
    pcg.solve(A, b, x)
    out = (pcg.num_iterations(), np.array(pcg.history()), x.get())
    pcg.destroy()
    amg.destroy()
    x.destroy()
    return out


def test_ranges_pair_up_and_leave_results_unchanged(hx):
    say("ranges on/off, 27-point 24^3")
    A = hx.ParCSRMatrix.laplacian27(24, 24, 24)
    n, _ = A.dims()
    b = hx.ParVector(n)
    b.set_random(3)
    it0, h0, x0 = solve(hx, A, b)
    hx.call("HYPRE_MI355X_SetProfilingRanges", 1)
    pushed, open_ = C.c_int(), C.c_int()
    hx.call("HYPRE_MI355X_GetProfilingRanges", C.byref(pushed), C.byref(open_))
    before = pushed.value
    it1, h1, x1 = solve(hx, A, b)
    hx.call("HYPRE_MI355X_GetProfilingRanges", C.byref(pushed), C.byref(open_))
    hx.call("HYPRE_MI355X_SetProfilingRanges", 0)
    say("ranges opened", pushed.value - before, "still open", open_.value)
    # PCG-Setup, AMGsetup, per level and per cycle ranges
    assert pushed.value - before > 4 * it1
    assert open_.value == 0
    assert it0 == it1
    assert np.array_equal(h0, h1)
    assert np.array_equal(x0.view(np.int64), x1.view(np.int64))
    b.destroy()
    A.destroy()
