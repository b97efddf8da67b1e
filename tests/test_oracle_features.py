"""Pin the oracle's matrix features (oracle/param_model.py matrix_features) by
closed forms on matrices whose features are known exactly (CPU only).

Features: [log10(n)/8, mean row length/27, std(row length)/27, mean min/max
|a_ij|, mean 2nd/1st largest |a_ij|, mean strong fraction at theta 0.25,
fraction of rows with a >= 10x diagonal jump to a neighbour, same for 1000x].
"""
import math

import numpy as np
import pytest

from oracle import param_model as PM


def chain(diag_exp):
    """1D chain: diagonal 10^k_i, -1 to the left and right neighbours."""
    n = len(diag_exp)
    ia, ja, a = [0], [], []
    for i in range(n):
        ja.append(i)
        a.append(10.0 ** diag_exp[i])
        for j in (i - 1, i + 1):
            if 0 <= j < n:
                ja.append(j)
                a.append(-1.0)
        ia.append(len(ja))
    return ia, ja, a


def csr(oracle, ia, ja, a):
    n = len(ia) - 1
    return oracle.Csr(n, n, np.asarray(ia, np.int32), np.asarray(ja, np.int32),
                      np.asarray(a, np.float64))


@pytest.mark.parametrize("N", [4, 7])
def test_laplace27_closed_forms(oracle, N):
    f = PM.matrix_features(oracle.gen_laplace27(N, N, N))
    # row length = product over x, y, z of 2 (boundary) or 3 (interior)
    mean_len = ((3 * N - 2) / N) ** 3
    mean_sq = ((9 * N - 10) / N) ** 3
    want = [math.log10(N ** 3) / 8, mean_len / 27, math.sqrt(mean_sq - mean_len ** 2) / 27,
            1.0, 1.0, 1.0, 0.0, 0.0]
    np.testing.assert_allclose(f, want, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("eps", [0.1, 0.5])
def test_laplace7_anisotropic_closed_forms(oracle, eps):
    N = 6
    f = PM.matrix_features(oracle.gen_laplace7(N, N, N, 1.0, eps, 1.0))
    p1 = 2.0 / N                          # a dimension has 1 neighbour with prob. 2/N, else 2
    cnt = {1: p1, 2: 1.0 - p1}
    # strong: the x and z couplings (-1) always; the y coupling (-eps) only
    # when eps > 0.25 (threshold 0.25 * max |a_ij| = 0.25)
    strong = 0.0
    for cx, px in cnt.items():
        for cy, py in cnt.items():
            for cz, pz in cnt.items():
                s = cx + cz + (cy if eps > 0.25 else 0)
                strong += px * py * pz * s / (cx + cy + cz)
    assert f[1] == pytest.approx((7 * N ** 3 - 6 * N ** 2) / N ** 3 / 27, rel=1e-14)
    assert f[3] == pytest.approx(eps, rel=1e-15)        # min/max off-diagonal
    assert f[4] == 1.0                                  # two unit couplings in every row
    assert f[5] == pytest.approx(strong, rel=1e-13)
    assert f[6] == 0.0 and f[7] == 0.0                  # constant diagonal


def test_jump_histogram_hand_count(oracle):
    # diagonals 1,1,1,100,100,100,1e5,1e5: rows 2,3 see a 100x jump, rows 5,6 a
    # 1000x one, so 4 of 8 rows jump by >= 10 and 2 of 8 by >= 1000
    f = PM.matrix_features(csr(oracle, *chain([0, 0, 0, 2, 2, 2, 5, 5])))
    assert f[6] == 4 / 8 and f[7] == 2 / 8
    # row lengths 2,3,3,3,3,3,3,2: mean 2.75, mean square 7.75
    assert f[1] == pytest.approx(2.75 / 27, rel=1e-15)
    assert f[2] == pytest.approx(math.sqrt(7.75 - 2.75 ** 2) / 27, rel=1e-14)
    # a jump of exactly 10 counts in the first bin (rows 0 and 1), not in the second
    f = PM.matrix_features(csr(oracle, *chain([0, 1, 1])))
    assert f[6] == 2 / 3 and f[7] == 0.0


def test_sign_and_isolated_rows(oracle):
    A = oracle.gen_vardiff27(6, 2, -3.0, 3.0, 5)
    B = csr(oracle, A.ia, A.ja, -np.asarray(A.a))
    np.testing.assert_allclose(PM.matrix_features(B), PM.matrix_features(A), rtol=0, atol=0)
    # no off-diagonal entries: only the row-length features are non-zero
    f = PM.matrix_features(csr(oracle, [0, 1, 2, 3, 4], [0, 1, 2, 3], [1.0, 100.0, 1.0, 100.0]))
    np.testing.assert_array_equal(f, [math.log10(4) / 8, 1 / 27, 0, 0, 0, 0, 0, 0])
