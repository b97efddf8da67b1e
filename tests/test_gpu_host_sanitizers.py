"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer.

`python -m hypre_amd.build --sanitize` builds every source with the host half
instrumented (device code as usual) and the C driver `ij_amg_pcg` against
it.  The driver goes through hypre's public API only: IJ assembly row by
row, BoomerAMG + PCG setup and solve, destruction.  Any heap overflow,
use after free or undefined behaviour in host code ends the run with an
error report; the run must converge as the plain build does.  Skipped when
the sanitizer build is absent.
"""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "hypre_amd", "examples", "ij_amg_pcg_asan")


def say(*a):
    print("[gpu-host-sanitizers]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.skipif(not os.path.exists(EXE), reason="sanitizer build absent (python -m hypre_amd.build --sanitize)")
@pytest.mark.parametrize("args", [("20", "7", "1.0"), ("16", "27", "0.9")])
def test_c_driver_clean_under_host_sanitizers(args):
    say("ij_amg_pcg_asan", *args)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:exitcode=23",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=24")
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=180, env=env)
    out = r.stdout + r.stderr
    say(out.strip().splitlines()[-1] if out.strip() else "(no output)")
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    assert "Iterations" in out or "iterations" in out, out[-2000:]
