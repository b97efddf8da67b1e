"""IJ file formats through the C-ABI (HYPRE_IJMatrixRead/ReadMM/ReadBinary/
Print/PrintBinary, HYPRE_IJVectorRead/ReadBinary/Print/PrintBinary): the
reference's own I/O fixtures, its io.jobs known answer (diagonally scaled
PCG, 11 iterations), MatrixMarket files checked against scipy's reader, and
print -> read round trips whose text is checked against the reference's
print format."""
import os
import shutil
import struct
import sys

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

IO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "io")


def say(*a):
    print("[gpu-ij-files]", *a, flush=True)
    sys.stdout.flush()


def coo_dict(M):
    """{(row, col): value} of a HostCsr or scipy matrix."""
    if hasattr(M, "tocoo"):
        c = M.tocoo()
        return {(int(r), int(k)): float(v) for r, k, v in zip(c.row, c.col, c.data)}
    out = {}
    for i in range(M.nrows):
        for q in range(M.ia[i], M.ia[i + 1]):
            out[(i, int(M.ja[q]))] = float(M.a[q])
    return out


def host_of(ij):
    return ij.get_object().to_host()


@pytest.fixture()
def workdir(tmp_path):
    for f in os.listdir(IO):
        shutil.copyfile(os.path.join(IO, f), tmp_path / f)
    return tmp_path


@pytest.mark.parametrize("tag,btag,xtag", [("i4f8", "i4f8", "i4f8"), ("i8f4", "i4f4", "i8f4")])
def test_binary_fixtures_and_io_jobs_known_answer(hx, oracle, golden, workdir, tag, btag, xtag):
    say("binary fixtures", tag)
    g = golden["io_ds_pcg"]
    A_ij = hx.ij_matrix_read(str(workdir / f"poisson.IJ.A.{tag}"), "bin")
    b_ij = hx.ij_vector_read(str(workdir / f"poisson.IJ.b.{btag}"), binary=True)
    x_ij = hx.ij_vector_read(str(workdir / f"poisson.IJ.x0.{xtag}"), binary=True)
    G = host_of(A_ij)
    L = oracle.gen_laplace7(10, 10, 10)
    assert np.array_equal(G.ia, L.ia) and np.array_equal(G.ja, L.ja)   # diagonal first, in order
    assert np.array_equal(G.a, L.a)
    A, b, x = A_ij.get_object(), b_ij.get_object(), x_ij.get_object()
    assert np.all(b.get() == 1.0) and np.all(x.get() == 0.0)
    pcg = hx.PCG("diag_scale", tol=1e-2, max_iter=1000, two_norm=1)
    pcg.setup(A, b, x)
    pcg.solve(A, b, x)
    assert pcg.num_iterations() == g["iterations"]
    assert abs(pcg.final_rel_res() - g["final_rel_res"]) <= 1e-6 * g["final_rel_res"] + 5e-10


def test_text_fixture(hx, workdir):
    say("text fixture")
    A_ij = hx.ij_matrix_read(str(workdir / "beam.A.IJ"), "ij")
    G = host_of(A_ij)
    want = {}
    with open(workdir / "beam.A.IJ.00000") as f:
        f.readline()
        for line in f:
            t = line.split()
            if t:
                want[(int(t[0]), int(t[1]))] = float(t[2])
    assert G.nrows == 459 and coo_dict(G) == want
    # every row stored diagonal first where the row has its diagonal
    for i in range(G.nrows):
        if (i, i) in want:
            assert G.ja[G.ia[i]] == i
    b = hx.ij_vector_read(str(workdir / "beam.b.IJ")).get_object().get()
    with open(workdir / "beam.b.IJ.00000") as f:
        f.readline()
        vals = np.array([float(l.split()[1]) for l in f if l.strip()])
    assert np.array_equal(b, vals)


def test_print_read_round_trips(hx, oracle, tmp_path):
    say("print / read round trips")
    A_or = oracle.gen_vardiff27(6, 2, -2.0, 2.0, 3)
    A_ij = hx.IJMatrix(0, A_or.nrows - 1)
    A_ij.initialize()
    A_ij.set_csr(A_or.ia, A_or.ja, A_or.a)
    A_ij.assemble()
    n = A_or.nrows
    base = str(tmp_path / "M")
    A_ij.print(base)
    # the reference's text layout: range line, then "%d %d %.14e" per entry
    with open(base + ".00000") as f:
        assert f.readline().split() == ["0", str(n - 1), "0", str(n - 1)]
        first = f.readline().split()
    assert first[0] == "0" and first[1] == "0" and first[2] == "%.14e" % A_or.a[0]
    R = host_of(hx.ij_matrix_read(base, "ij"))
    assert np.array_equal(R.ia, A_or.ia) and np.array_equal(R.ja, A_or.ja)
    np.testing.assert_allclose(R.a, A_or.a, rtol=1e-14, atol=0)   # 15 significant digits
    A_ij.print(base + "b", binary=True)
    raw = open(base + "b.00000.bin", "rb").read()
    h = struct.unpack("<11Q", raw[:88])
    assert h == (1, 4, 8, n, n, A_or.nnz, A_or.nnz, 0, n - 1, 0, n - 1)
    B = host_of(hx.ij_matrix_read(base + "b", "bin"))
    assert np.array_equal(B.ia, A_or.ia) and np.array_equal(B.ja, A_or.ja)
    assert np.array_equal(B.a, A_or.a)                              # binary is exact
    v = hx.IJVector(0, n - 1)
    vals = 2 * oracle.rand_stream(9, n) - 1
    v.set_values(vals)
    v.assemble()
    v.print(base + "v")
    v.print(base + "vb", binary=True)
    rt = hx.ij_vector_read(base + "v").get_object().get()
    np.testing.assert_allclose(rt, vals, rtol=1e-14, atol=0)
    rb = hx.ij_vector_read(base + "vb", binary=True).get_object().get()
    assert np.array_equal(rb, vals)


@pytest.mark.parametrize("symmetry", ["general", "symmetric"])
def test_matrix_market_against_scipy(hx, tmp_path, symmetry):
    say("matrix market", symmetry)
    rng = np.random.default_rng(4)
    M = sp.random(40, 40, density=0.15, random_state=rng, format="csr")
    M = M + sp.diags(rng.uniform(5, 6, 40))
    if symmetry == "symmetric":
        M = (M + M.T) * 0.5
    path = str(tmp_path / "m.mtx")
    scipy.io.mmwrite(path, M.tocoo(), symmetry=symmetry)
    with open(path) as f:
        assert symmetry in f.readline()
    G = host_of(hx.ij_matrix_read(path, "mm"))
    want = coo_dict(scipy.io.mmread(path).tocsr())
    got = coo_dict(G)
    assert got.keys() == want.keys()
    for k in want:
        assert got[k] == want[k]


def test_read_errors(hx, tmp_path):
    say("read errors")
    with pytest.raises(hx.HypreError):
        hx.ij_matrix_read(str(tmp_path / "missing"), "ij")
    bad = tmp_path / "bad.00000"
    bad.write_text("0 1 0 1\n0 0 1.0\n1 x 2.0\n")
    with pytest.raises(hx.HypreError):
        hx.ij_matrix_read(str(tmp_path / "bad"), "ij")
    mm = tmp_path / "c.mtx"
    mm.write_text("%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 1 1.0 0.0\n")
    with pytest.raises(hx.HypreError):
        hx.ij_matrix_read(str(mm), "mm")
    hx.clear_errors()
