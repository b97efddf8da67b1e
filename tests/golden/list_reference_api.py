"""Write reference_api_names.txt: every function name that hypre's public
headers for this path declare (the HYPRE_*.h files listed in HEADERS, which
include the Generate* problem builders of src/parcsr_ls/HYPRE_parcsr_ls.h),
one per line, sorted.

Run in a container that holds the reference tree:

    python tests/golden/list_reference_api.py /root/reference

The list is data (names only); tests/test_abi.py checks that every hypre
name declared by include/HYPRE*.h appears in it, so the drop-in headers
declare no hypre name the reference does not have.
"""
import os
import re
import sys

HEADERS = {"HYPRE.h", "HYPRE_utilities.h", "HYPRE_seq_mv.h", "HYPRE_parcsr_mv.h",
           "HYPRE_IJ_mv.h", "HYPRE_krylov.h", "HYPRE_parcsr_ls.h"}
DECL = re.compile(r"\b((?:HYPRE_|Generate)[A-Za-z0-9_]*)\s*\(")


def public_names(ref: str) -> set:
    names = set()
    for root, _, files in os.walk(os.path.join(ref, "src")):
        if "/test" in root or "/examples" in root or "/docs" in root:
            continue
        for fn in files:
            if fn not in HEADERS:
                continue
            with open(os.path.join(root, fn), errors="replace") as f:
                text = f.read()
            text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
            text = re.sub(r"//[^\n]*", " ", text)
            names.update(DECL.findall(text))
    return names


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_api_names.txt")
    names = sorted(public_names(ref))
    with open(out, "w") as f:
        f.write("\n".join(names) + "\n")
    print(f"{len(names)} names -> {out}")
