"""SpMV through the C-ABI on matrices that exercise the streamed row windows:
wave ranges whose rows are all empty, rows longer than many windows, blocks
of rows ending inside a window, every window size (256/512/1024 entries) and
rows-per-block width (long average rows), sizes around the 64-row blocks.

The GPU sums each row's products in storage order from 0 and forms
alpha*t + beta*b, as src/seq_mv/csr_matvec.c does (restated by the oracle's
orc_matvec), so results are compared bit for bit.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-spmv]", *a, flush=True)
    sys.stdout.flush()


def ragged(n, ncols, lens, seed):
    rng = np.random.default_rng(seed)
    lens = np.asarray(lens, np.int64)
    ia = np.zeros(n + 1, np.int64)
    ia[1:] = np.cumsum(lens)
    nnz = int(ia[-1])
    ja = rng.integers(0, ncols, nnz, dtype=np.int64)
    a = rng.uniform(-1.0, 1.0, nnz)
    return ia.astype(np.int32), ja.astype(np.int32), a


def cases():
    rng = np.random.default_rng(5)
    out = {}
    out["one_row"] = (1, 1, [1])
    out["all_empty_300"] = (300, 300, [0] * 300)
    out["n65_some_empty"] = (65, 40, rng.integers(0, 4, 65) * (np.arange(65) % 3 != 0))
    # an empty band covering many whole waves, ragged rows either side
    lens = rng.integers(0, 40, 200_000)
    lens[60_000:75_000] = 0
    out["empty_band_200k"] = (200_000, 200_000, lens)
    # one row spanning ~20 windows among short rows, and a long last row
    lens = rng.integers(1, 8, 30_000)
    lens[12_345] = 20_000
    lens[-1] = 3_000
    out["long_rows"] = (30_000, 30_000, lens)
    # long average rows: fewer rows per block (RW 4), 1024-entry windows
    out["avg300_rw4"] = (5_000, 5_000, rng.integers(250, 350, 5_000))
    # short rows: 256- and 512-entry windows
    out["avg3_sw256"] = (70_000, 70_000, rng.integers(1, 6, 70_000))
    out["avg7_sw512"] = (70_001, 70_001, rng.integers(4, 11, 70_001))
    # 27-point-like rows, size just past a multiple of the 64-row block
    out["avg27_sw1024"] = (64 * 4097 + 7, 64 * 4097 + 7, rng.integers(20, 28, 64 * 4097 + 7))
    return out


@pytest.mark.parametrize("name", list(cases().keys()))
def test_spmv_stream_bit_exact(hx, oracle, name):
    n, ncols, lens = cases()[name]
    say("spmv stream", name, n)
    ia, ja, a = ragged(n, ncols, lens, sum(name.encode()))
    A = oracle.Csr(n, ncols, ia, ja, a)
    H = hx.HostCsr(n, ncols, ia, ja, a)
    x = 2 * oracle.rand_stream(7, ncols) - 1
    b = 2 * oracle.rand_stream(8, n) - 1
    for alpha, beta in ((1.0, 0.0), (-1.0, 1.0), (0.7, -0.3)):
        got = hx.spmv(H, x, alpha, beta, b)
        want = oracle.matvec(alpha, A, x, beta, b)
        bad = np.nonzero(got != want)[0]
        assert bad.size == 0, (name, alpha, beta, bad[:5], got[bad[:5]], want[bad[:5]])
