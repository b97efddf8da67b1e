"""The C-ABI library loads and exports every entry point include/*.h declares
(CPU only: no compute calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "hypre_amd", "libhypre_mi355x.so")


def declared_functions():
    names = set()
    pat = re.compile(r"^\s*(?:HYPRE_Int|HYPRE_ParCSRMatrix|void)\s+\*?\s*((?:HYPRE_|Generate)\w+)\s*\(", re.M)
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            with open(os.path.join(ROOT, "include", fn)) as f:
                names.update(pat.findall(f.read()))
    return names


def test_headers_declare_the_hot_path():
    names = declared_functions()
    for must in ("HYPRE_IJMatrixCreate", "HYPRE_IJMatrixSetValues", "HYPRE_IJMatrixAssemble",
                 "HYPRE_IJMatrixGetObject", "HYPRE_IJVectorCreate", "HYPRE_BoomerAMGCreate",
                 "HYPRE_BoomerAMGSetup", "HYPRE_BoomerAMGSolve", "HYPRE_ParCSRPCGCreate",
                 "HYPRE_ParCSRPCGSetup", "HYPRE_ParCSRPCGSolve", "HYPRE_ParCSRPCGSetPrecond",
                 "HYPRE_PCGSetTol", "HYPRE_ParCSRMatrixMatvec", "GenerateLaplacian27pt"):
        assert must in names
    assert len(names) > 150


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build with python -m hypre_amd.build"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_match_exports():
    from hypre_amd.lib import SIGNATURES, lib
    L = lib()
    for name in SIGNATURES:
        assert hasattr(L, name), name


EXAMPLE = os.path.join(ROOT, "hypre_amd", "examples", "ij_amg_pcg")


def test_c_driver_uses_only_the_declared_api():
    """The plain-C driver (hypre's public API only) links against the library,
    and every library symbol it imports is a hypre entry point in include/*.h."""
    import subprocess
    assert os.path.exists(EXAMPLE), "build with python -m hypre_amd.build"
    out = subprocess.run(["nm", "-D", "--undefined-only", EXAMPLE], capture_output=True,
                         text=True, check=True).stdout
    imported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    hypre = {s for s in imported if s.startswith("HYPRE_") or s.startswith("Generate")}
    assert {"HYPRE_IJMatrixSetValues", "HYPRE_ParCSRPCGSolve", "HYPRE_BoomerAMGSetup"} <= hypre
    assert hypre <= declared_functions()
    assert not any(s.startswith("HYPRE_MI355X") for s in hypre)


def test_drop_in_headers_declare_only_reference_names():
    """Every hypre function name that include/HYPRE*.h declares exists in the
    reference's public headers of this path (names listed by
    tests/golden/list_reference_api.py); extensions carry the HYPRE_MI355X_
    prefix and live in include/hypre_mi355x.h."""
    with open(os.path.join(ROOT, "tests", "golden", "reference_api_names.txt")) as f:
        ref = set(f.read().split())
    pat = re.compile(r"^\s*(?:HYPRE_Int|HYPRE_ParCSRMatrix|void)\s+\*?\s*((?:HYPRE_|Generate)\w+)\s*\(", re.M)
    invented = {}
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        with open(os.path.join(ROOT, "include", fn)) as f:
            names = set(pat.findall(f.read()))
        if fn.startswith("HYPRE"):
            assert not any(n.startswith("HYPRE_MI355X") for n in names), fn
            extra = sorted(names - ref)
            if extra:
                invented[fn] = extra
        else:
            assert all(n.startswith("HYPRE_MI355X_") for n in names), fn
    assert not invented, invented


def test_c_driver_source_is_plain_hypre():
    """The example passes its communicator as the integer handle of hypre's
    sequential headers (src/utilities/HYPRE_utilities.h:175-177) and names no
    macro or function that those headers do not provide."""
    with open(EXAMPLE + ".c") as f:
        src = f.read()
    assert "MPI_COMM_WORLD" not in src and "HYPRE_MI355X" not in src
    with open(os.path.join(ROOT, "tests", "golden", "reference_api_names.txt")) as f:
        ref = set(f.read().split())
    called = set(re.findall(r"\b(HYPRE_[A-Za-z0-9_]+)\s*\(", src))
    assert called and called <= ref, sorted(called - ref)
