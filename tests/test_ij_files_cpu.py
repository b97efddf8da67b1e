"""IJ file formats on the CPU: the reference's own I/O fixtures parsed with a
numpy restatement of hypre's binary and text layouts, checked against what
the mathematics fixes (the 10^3 7-point Laplacian of the oracle's generator,
b = 1, x0 = 0) and against the recorded result of the reference's io.jobs
(diagonally scaled PCG, 11 iterations, relative residual 6.733697e-03)."""
import os
import struct

import numpy as np
import pytest

IO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "io")


def read_bin_matrix(path):
    """src/IJ_mv/IJMatrix.c:252-500: 11 uint64 header words, then rows, cols, values."""
    raw = open(path, "rb").read()
    h = struct.unpack("<11Q", raw[:88])
    nnz = h[6]
    it = np.uint32 if h[1] == 4 else np.uint64
    vt = np.float32 if h[2] == 4 else np.float64
    off = 88
    rows = np.frombuffer(raw, it, nnz, off).astype(np.int64)
    off += nnz * h[1]
    cols = np.frombuffer(raw, it, nnz, off).astype(np.int64)
    off += nnz * h[1]
    vals = np.frombuffer(raw, vt, nnz, off).astype(np.float64)
    assert off + nnz * h[2] == len(raw)
    return h, rows, cols, vals


def read_bin_vector(path):
    """src/IJ_mv/IJVector.c:87-245: 8 uint64 header words, then the values."""
    raw = open(path, "rb").read()
    h = struct.unpack("<8Q", raw[:64])
    vt = np.float32 if h[1] == 4 else np.float64
    v = np.frombuffer(raw, vt, h[5], 64).astype(np.float64)
    assert 64 + h[5] * h[1] == len(raw)
    return h, v


def read_ij_text(path):
    with open(path) as f:
        il, iu, jl, ju = (int(t) for t in f.readline().split())
        ent = {}
        for line in f:
            t = line.split()
            if t:
                ent[(int(t[0]), int(t[1]))] = float(t[2])
    return (il, iu, jl, ju), ent


def ds_pcg(A, b, x, tol, max_iter=1000):
    """Diagonally scaled PCG with the two-norm relative criterion of
    src/krylov/pcg.c (||r|| / ||b|| < tol), returning (iterations, rel)."""
    d = A.diagonal()
    r = b - A @ x
    bb = float(b @ b)
    p = r / d
    gamma = float(r @ p)
    for it in range(1, max_iter + 1):
        s = A @ p
        alpha = gamma / float(s @ p)
        x = x + alpha * p
        r = r - alpha * s
        rel = np.sqrt(float(r @ r) / bb)
        if rel < tol:
            return it, rel
        z = r / d
        gnew = float(r @ z)
        p = z + (gnew / gamma) * p
        gamma = gnew
    return max_iter, rel


@pytest.mark.parametrize("tag", ["i4f8", "i8f4"])
def test_binary_fixture_is_the_laplacian(oracle, tag):
    h, rows, cols, vals = read_bin_matrix(os.path.join(IO, f"poisson.IJ.A.{tag}.00000.bin"))
    assert h[0] == 1 and h[1] == int(tag[1]) and h[2] == int(tag[3])
    assert h[3:7] == (1000, 1000, 6400, 6400) and h[7:] == (0, 999, 0, 999)
    L = oracle.gen_laplace7(10, 10, 10).to_scipy().tocoo()
    got = {(int(r), int(c)): v for r, c, v in zip(rows, cols, vals)}
    want = {(int(r), int(c)): v for r, c, v in zip(L.row, L.col, L.data)}
    assert got == want
    # each row written diagonal first, then its other columns in order
    first = np.r_[True, rows[1:] != rows[:-1]]
    assert np.all(rows[first] == cols[first])


def test_binary_vectors():
    for name, val in (("b.i4f8", 1.0), ("b.i4f4", 1.0), ("x0.i4f8", 0.0), ("x0.i8f4", 0.0)):
        h, v = read_bin_vector(os.path.join(IO, f"poisson.IJ.{name}.00000.bin"))
        assert h[0] == 1 and h[2:] == (0, 1000, 1000, 1000, 1, 0)
        assert np.all(v == val)


def test_io_jobs_known_answer_with_numpy(oracle, golden):
    g = golden["io_ds_pcg"]
    h, rows, cols, vals = read_bin_matrix(os.path.join(IO, "poisson.IJ.A.i4f8.00000.bin"))
    import scipy.sparse as sp
    A = sp.csr_matrix((vals, (rows, cols)), shape=(1000, 1000))
    it, rel = ds_pcg(A, np.ones(1000), np.zeros(1000), 1e-2)
    assert it == g["iterations"]
    assert abs(rel - g["final_rel_res"]) <= 1e-6 * g["final_rel_res"] + 5e-10


def test_text_fixture_layout():
    (il, iu, jl, ju), ent = read_ij_text(os.path.join(IO, "beam.A.IJ.00000"))
    assert (il, iu, jl, ju) == (0, 458, 0, 458)
    assert all(il <= r <= iu and jl <= c <= ju for r, c in ent)
    with open(os.path.join(IO, "beam.b.IJ.00000")) as f:
        lo, hi = (int(t) for t in f.readline().split())
        lines = [l.split() for l in f if l.strip()]
    assert (lo, hi) == (0, 458) and len(lines) == 459
