"""Pin the oracle's problem generators by closed forms and reductions (CPU only).

gen_laplace7(nx, ny, nz, cx, cy, cz) follows the stencil of the reference's
GenerateLaplacian (src/parcsr_ls/par_laplace.c: diagonal 2(cx+cy+cz), -cx /
-cy / -cz to the x / y / z neighbours, out-of-domain neighbours dropped).
gen_aniso7(N, eps, theta) has no reference counterpart: it is the 7-point
stencil of the diffusion tensor R(theta) diag(1, eps) R(theta)^T in the xy
plane (cross term dropped) and 1 along z, so it must reduce to gen_laplace7
at theta = 0 and pi/2, keep the trace 1 + eps of the in-plane tensor, give
equal x and y couplings at pi/4 and be pi-periodic in theta.
"""
import math

import numpy as np
import pytest


def _coupling(A, N, row, d):
    """Value of A at (row, row + d), or None when that entry is absent."""
    s, e = int(A.ia[row]), int(A.ia[row + 1])
    hit = np.nonzero(A.ja[s:e] == row + d)[0]
    return float(A.a[s + hit[0]]) if hit.size else None


@pytest.mark.parametrize("c", [(1.0, 1.0, 1.0), (1.0, 0.01, 1.0), (0.3, 2.0, 5.0)])
def test_laplace7_coefficients(oracle, c):
    N = 5
    A = oracle.gen_laplace7(N, N, N, *c)
    cx, cy, cz = c
    interior = 2 + N * (2 + N * 2)
    assert A.a[A.ia[interior]] == 2.0 * (cx + cy + cz)
    for d, v in ((1, cx), (-1, cx), (N, cy), (-N, cy), (N * N, cz), (-N * N, cz)):
        assert _coupling(A, N, interior, d) == -v
    # corner row 0: the diagonal keeps 2(cx+cy+cz) (Dirichlet), 3 neighbours remain
    assert A.a[A.ia[0]] == 2.0 * (cx + cy + cz)
    assert A.ia[1] - A.ia[0] == 4
    assert _coupling(A, N, 0, -1) is None


@pytest.mark.parametrize("eps", [1.0, 0.1, 1e-3])
def test_aniso7_axis_aligned_reductions(oracle, eps):
    N = 6
    A = oracle.gen_aniso7(N, eps, 0.0)
    B = oracle.gen_laplace7(N, N, N, 1.0, eps, 1.0)
    assert np.array_equal(A.ia, B.ia) and np.array_equal(A.ja, B.ja)
    assert np.array_equal(A.a, B.a)
    # a quarter turn swaps the strong and the weak in-plane direction
    A = oracle.gen_aniso7(N, eps, math.pi / 2)
    B = oracle.gen_laplace7(N, N, N, eps, 1.0, 1.0)
    assert np.array_equal(A.ia, B.ia) and np.array_equal(A.ja, B.ja)
    assert np.array_equal(A.a, B.a)


@pytest.mark.parametrize("eps,theta", [(0.05, 0.3), (1e-3, 2.0), (0.5, 1.1)])
def test_aniso7_rotation_invariants(oracle, eps, theta):
    N = 5
    row = 2 + N * (2 + N * 2)
    A = oracle.gen_aniso7(N, eps, theta)
    kx = -_coupling(A, N, row, 1)
    ky = -_coupling(A, N, row, N)
    kz = -_coupling(A, N, row, N * N)
    assert kz == 1.0
    assert abs((kx + ky) - (1.0 + eps)) <= 4e-16 * (1.0 + eps)        # trace
    assert min(kx, ky) >= eps * (1 - 1e-15) and max(kx, ky) <= 1.0 + 1e-15
    assert A.a[A.ia[row]] == 2.0 * (kx + ky + kz)
    # pi-periodic: theta and theta + pi give the same tensor (up to rounding of cos/sin)
    B = oracle.gen_aniso7(N, eps, theta + math.pi)
    assert np.array_equal(A.ja, B.ja)
    assert np.allclose(A.a, B.a, rtol=1e-14, atol=0)
    # at pi/4 the x and y couplings are equal, (1 + eps) / 2
    C = oracle.gen_aniso7(N, eps, math.pi / 4)
    kx4, ky4 = -_coupling(C, N, row, 1), -_coupling(C, N, row, N)
    assert abs(kx4 - ky4) <= 1e-15 and abs(kx4 - 0.5 * (1 + eps)) <= 1e-15


def test_aniso7_orientation(oracle):
    """A rotation by theta turns the strong direction (1, 0) into (cos, sin):
    at small theta the x coupling dominates, near pi/2 the y coupling."""
    N, row = 5, 2 + 5 * (2 + 5 * 2)
    A = oracle.gen_aniso7(N, 1e-2, 0.2)
    assert -_coupling(A, N, row, 1) > -_coupling(A, N, row, N)
    A = oracle.gen_aniso7(N, 1e-2, 1.4)
    assert -_coupling(A, N, row, 1) < -_coupling(A, N, row, N)
