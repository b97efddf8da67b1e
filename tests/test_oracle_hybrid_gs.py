"""Pin the oracle's hybrid Gauss-Seidel smoother (relax 3/4/6/8/13/14) on the
CPU with closed forms that do not share its loop: block Gauss-Seidel written
as triangular solves per partition, Jacobi coupling between partitions.
A wrong partition bound, a swapped sweep order, a dropped diagonal term or a
copy taken at the wrong time fails one of them."""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp


def _parts(oracle, n, p):
    return [oracle.partition1d(n, p, j) for j in range(p)]


@pytest.mark.parametrize("n,p", [(10, 3), (7, 7), (5, 9), (1000, 64), (17, 1)])
def test_partition1d_layout(oracle, n, p):
    b = _parts(oracle, n, p)
    assert b[0][0] == 0 and b[-1][1] == n
    for (s0, e0), (s1, e1) in zip(b, b[1:]):
        assert e0 == s1                                  # contiguous, in order
    sizes = [e - s for s, e in b]
    assert max(sizes) - min(sizes) <= (0 if n % p == 0 else 1)
    assert sizes == sorted(sizes, reverse=True)          # the larger ones first


def test_hgs_parts_rule(oracle):
    assert oracle.hgs_parts(1000, 7, 64) == 7            # explicit count wins
    assert oracle.hgs_parts(8, 0, 64) == 1               # small levels: 8-row partitions
    assert oracle.hgs_parts(65, 0, 64) == 9
    assert oracle.hgs_parts(6400, 0, 64) == 800
    assert oracle.hgs_parts(6400, 0, 4) == 1600          # rows below 8 are kept
    assert oracle.hgs_parts(8000000, 0, 64) == 129033    # 62 rows: ~128K partitions
    assert oracle.hgs_parts(256 ** 3, 0, 64) == 262144   # 64 rows


def _matrix(rng, n, dense=0.3, nonsym=True):
    M = sp.random(n, n, density=dense, random_state=rng, format="csr")
    M.data = -rng.uniform(0.1, 1.0, M.nnz)
    if not nonsym:
        M = (M + M.T) * 0.5
    M = M.tolil()
    M.setdiag(0)
    M = M.tocsr()
    d = np.abs(M).sum(axis=1).A1 + rng.uniform(0.5, 2.0, n)
    M = M + sp.diags(d)
    return M.tocsr()


def _to_oracle_csr(oracle_mod, M):
    """CSR with the diagonal first and the other columns in a shuffled order
    (the smoother must follow the stored order, the closed form does not care)."""
    n = M.shape[0]
    ia = [0]
    ja, a = [], []
    rng = np.random.default_rng(5)
    for i in range(n):
        cols = M.indices[M.indptr[i]:M.indptr[i + 1]]
        vals = M.data[M.indptr[i]:M.indptr[i + 1]]
        off = [(c, v) for c, v in zip(cols, vals) if c != i]
        rng.shuffle(off)
        ja.append(i)
        a.append(M[i, i])
        for c, v in off:
            ja.append(c)
            a.append(v)
        ia.append(len(ja))
    return oracle_mod.Csr(n, n, np.array(ia, np.int32), np.array(ja, np.int32), np.array(a))


def _closed_form(M, f, u, bounds, lam, direction, symm, skip, w, om):
    """Block form of one hybrid sweep.  lam: divisor per row (l1 norm or diagonal)."""
    A = M.toarray()
    n = A.shape[0]
    v = u.copy()
    out = u.copy()
    for s, e in bounds:
        if e <= s:
            continue
        p = slice(s, e)
        App = A[p, p]
        D = np.diag(np.diag(App))
        Lo = np.tril(App, -1)
        Up = np.triu(App, 1)
        mask = np.ones(n, bool)
        mask[p] = False
        ext = A[p][:, mask] @ v[mask]          # other partitions: the copy
        Lam = np.diag(lam[p])
        fp = f[p]
        up_ = v[p]
        orders = [1, -1] if symm else [direction]
        cur = up_.copy()
        for k, o in enumerate(orders):
            Tri, Oth = (Lo, Up) if o > 0 else (Up, Lo)
            lower = o > 0
            if w == 1.0 and om == 1.0:
                if skip:
                    # D x_new = f - Tri x_new - Oth x_cur - ext
                    rhs = fp - Oth @ cur - ext
                    cur = sla.solve_triangular(D + Tri, rhs, lower=lower)
                else:
                    # Lam (x_new - x_cur) = f - Tri x_new - (D + Oth) x_cur - ext
                    rhs = Lam @ cur + fp - (D + Oth) @ cur - ext
                    cur = sla.solve_triangular(Lam + Tri, rhs, lower=lower)
            else:
                prod = 1.0 - w * om
                # in-partition: res0 = -(Tri x_new + Oth x_cur) (+ diag if not skip),
                # res2 = (Tri + Oth) v (+ diag if not skip); other partitions: res
                Dk = np.zeros_like(D) if skip else D
                base = (prod * Lam @ cur) if skip else (Lam @ cur)
                rhs = base + w * (om * (fp - ext) - Oth @ cur - Dk @ cur
                                  + (1.0 - om) * (Tri + Oth + Dk) @ up_)
                cur = sla.solve_triangular(Lam + w * Tri, rhs, lower=lower)
        out[p] = cur
    return out


CASES = [
    # (direction, symm, skip, uses_l1, w, omega)  relax type
    (1, False, 0, True, 1.0, 1.0),     # 13
    (-1, False, 0, True, 1.0, 1.0),    # 14
    (1, True, 0, True, 1.0, 1.0),      # 8
    (1, False, 1, False, 1.0, 1.0),    # 3
    (-1, False, 1, False, 1.0, 1.0),   # 4
    (1, True, 1, False, 1.0, 1.0),     # 6
    (1, False, 1, True, 0.7, 1.0),     # 13 with a relaxation weight
    (-1, False, 1, True, 0.8, 1.3),    # 14 with weight and omega
    (1, True, 1, False, 1.1, 0.9),     # 6 with weight and omega
]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("parts", [1, 3, 8, 40])
def test_hybrid_gs_matches_block_form(oracle, case, parts):
    direction, symm, skip, uses_l1, w, om = case
    rng = np.random.default_rng(parts * 31 + len(str(case)))
    n = 40
    M = _matrix(rng, n)
    A = _to_oracle_csr(oracle, M)
    f = rng.standard_normal(n)
    u = rng.standard_normal(n)
    if uses_l1:
        l1 = np.abs(M).sum(axis=1).A1 * rng.uniform(1.0, 1.5, n)   # any divisor works
    else:
        l1 = None
    lam = l1 if uses_l1 else M.diagonal()
    got = oracle.hybrid_gs(A, f, u, parts, direction, symm, skip, l1, w, om)
    want = _closed_form(M, f, u, _parts(oracle, n, parts), lam, direction, symm, skip, w, om)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_one_row_partitions_are_jacobi(oracle):
    rng = np.random.default_rng(3)
    M = _matrix(rng, 30)
    A = _to_oracle_csr(oracle, M)
    f, u = rng.standard_normal(30), rng.standard_normal(30)
    got = oracle.hybrid_gs(A, f, u, 30, 1, False, 1, None)
    D = M.diagonal()
    want = (f - (M @ u - D * u)) / D
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)


def test_zero_divisor_rows_are_skipped(oracle):
    rng = np.random.default_rng(4)
    M = _matrix(rng, 12)
    A = _to_oracle_csr(oracle, M)
    l1 = np.abs(M).sum(axis=1).A1
    l1[[2, 7]] = 0.0
    f, u = rng.standard_normal(12), rng.standard_normal(12)
    got = oracle.hybrid_gs(A, f, u, 3, 1, False, 0, l1)
    assert got[2] == u[2] and got[7] == u[7]


def test_l1_norms_option4_is_signed_diagonal(oracle):
    A = oracle.gen_laplace27(4, 4, 3)
    assert np.array_equal(oracle.l1_norms_diag(A), A.a[A.ia[:-1]])
    B = oracle.Csr(A.nrows, A.ncols, A.ia, A.ja, -A.a)
    assert np.array_equal(oracle.l1_norms_diag(B), -A.a[A.ia[:-1]])


def _vcycle_operator(oracle, H, n):
    B = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        B[:, k] = H.cycle(e)
    return B


@pytest.mark.parametrize("down,up", [(13, 14), (3, 4), (8, 8)])
def test_gs_vcycle_is_symmetric(oracle, down, up):
    # forward sweeps down and backward sweeps up (or symmetric sweeps on both)
    # make the V-cycle a symmetric operator for a symmetric A
    A = oracle.gen_laplace7(6, 5, 4)
    H = oracle.Hierarchy(A, relax_type=down, relax_up_type=up, hgs_parts=5,
                         max_row_sum=1.0)
    assert H.num_levels >= 2
    B = _vcycle_operator(oracle, H, A.nrows)
    np.testing.assert_allclose(B, B.T, atol=1e-12)
    # and contracts: the error propagator I - B A has spectral radius < 1
    E = np.eye(A.nrows) - B @ A.to_scipy().toarray()
    assert np.max(np.abs(np.linalg.eigvals(E))) < 0.5


def test_gs_pcg_converges_faster_than_l1_jacobi(oracle):
    A = oracle.gen_laplace27(14, 14, 14)
    b = oracle.rand_stream(7, A.nrows) * 2.0 - 1.0
    its = {}
    for name, kw in (("l1jac", dict(relax_type=18)),
                     ("hgs", dict(relax_type=13, relax_up_type=14))):
        H = oracle.Hierarchy(A, max_row_sum=1.0, **kw)
        _, it, rel, _, err = H.pcg(b, 1e-8, 100)
        assert err == 0 and rel < 1e-8
        its[name] = it
    assert its["hgs"] <= its["l1jac"]
