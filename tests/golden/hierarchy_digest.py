"""SHA-256 digests of a BoomerAMG hierarchy, level by level.

For each level: the row count and the digests of A's CSR arrays (row
pointers, columns, values as little-endian int32 / float64 bytes), of S's
pattern, the C/F marker, the l1 norms, P and R (= P^T), the last four on every
level but the coarsest.  Equal digests mean bit-identical arrays, so the comparison
pins the C/F split, every sparsity pattern with its within-row order and
every value, at sizes where shipping the arrays themselves is impractical.
(The coarsest level is solved by Gaussian elimination and has no l1 norms.)

`get(kind, lev)` returns: for "A", "P", "R" an object with nrows, ia, ja, a;
for "S" one with nrows, ia, ja; for "CF" an int32 array; for "l1" a float64
array.
"""
import hashlib

import numpy as np


def _h(x, dtype) -> str:
    return hashlib.sha256(np.ascontiguousarray(np.asarray(x), dtype=dtype).tobytes()).hexdigest()


def csr_digest(M, values=True) -> dict:
    n = int(M.nrows)
    ia = np.asarray(M.ia)[: n + 1]
    nnz = int(ia[n])
    d = {"rows": n, "nnz": nnz, "ia": _h(ia, "<i4"), "ja": _h(np.asarray(M.ja)[:nnz], "<i4")}
    if values:
        d["a"] = _h(np.asarray(M.a)[:nnz], "<f8")
    return d


def hierarchy_digests(num_levels: int, get) -> list:
    out = []
    for lev in range(num_levels):
        d = {"A": csr_digest(get("A", lev))}
        if lev < num_levels - 1:
            d["S"] = csr_digest(get("S", lev), values=False)
            d["CF"] = _h(get("CF", lev), "<i4")
            d["P"] = csr_digest(get("P", lev))
            d["R"] = csr_digest(get("R", lev))
            d["l1"] = _h(get("l1", lev), "<f8")
        out.append(d)
    return out
