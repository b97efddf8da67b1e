"""Pin the oracle's aggressive coarsening (CPU only).

Aggressive levels (src/parcsr_ls/par_amg_setup.c:1311-1352, :1799-1824)
coarsen the C points of a first PMIS split again over the distance-two
strength graph (hypre_BoomerAMGCreate2ndS, par_strength.c:1795; PMIS with
CF_init 3), correct the split (par_strength.c:3068) and interpolate with the
modified multipass method (agg_interp_type 8, par_mod_multi_interp.c:17),
whose later passes multiply their one-step weights into the previous pass's
(hypre_ParMatmul, par_csr_matop.c:348).  The reference has no single-rank
aggressive-coarsening run on record (its aggressive jobs in
src/test/TEST_ij/agg_interp.jobs and coarsening.jobs run on 2-8 ranks, apart
from a 2^3 grid that never coarsens), so the oracle is pinned by:

* the second strength matrix against a brute-force product (S + S S on the
  C points, minus the diagonal);
* hand-derived weights of the multipass interpolation on 1-D chains, where
  every pass, the boundary scaling and the chaining through the product are
  fixed by the stencil;
* exact row sums of 1 on a zero-row-sum (graph Laplacian) operator, which
  every pass preserves;
* the product against scipy, and its first-seen column order against a
  brute-force restatement;
* invariants of the second coarsening (C independent in S2, F points that
  influence someone depend on a C point, rows without connections are C);
* Galerkin coarse operators of aggressive hierarchies against a dense
  P^T A P.
"""
import numpy as np
import pytest
import scipy.sparse as sp


def csr_from_scipy(oracle, M, diag_first=True):
    """Oracle Csr of a square scipy matrix, diagonal first in each row, the
    off-diagonals by ascending column (the generators' layout)."""
    M = sp.csr_matrix(M)
    M.sort_indices()
    n = M.shape[0]
    ia, ja, a = [0], [], []
    for i in range(n):
        cols = M.indices[M.indptr[i]:M.indptr[i + 1]]
        vals = M.data[M.indptr[i]:M.indptr[i + 1]]
        order = sorted(range(len(cols)), key=lambda k: (cols[k] != i, cols[k])) if diag_first \
            else range(len(cols))
        for k in order:
            ja.append(cols[k])
            a.append(vals[k])
        ia.append(len(ja))
    return oracle.Csr(n, n, np.array(ia, np.int32), np.array(ja, np.int32), np.array(a))


def chain(n):
    """1-D Dirichlet Laplacian tridiag(-1, 2, -1)."""
    return sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])


def graph_laplacian_3d(N):
    """7-point graph Laplacian of an N^3 grid (diagonal = number of
    neighbours, so every row sums to 0)."""
    e = np.ones(N)
    T = sp.diags([-e[:-1], -e[:-1]], [-1, 1])
    I = sp.identity(N)
    L = sp.kron(sp.kron(T, I), I) + sp.kron(sp.kron(I, T), I) + sp.kron(sp.kron(I, I), T)
    L = L.tocsr()
    d = -np.asarray(L.sum(axis=1)).ravel()
    return (L + sp.diags(d)).tocsr()


def dense(M):
    return M.to_scipy().toarray()


# ------------------------------------------------------------ second strength
@pytest.mark.parametrize("gen", ["l7", "l27", "aniso", "jump"])
def test_second_strength_is_distance_two_on_c_points(oracle, gen):
    A = {"l7": lambda: oracle.gen_laplace7(10, 9, 8),
         "l27": lambda: oracle.gen_laplace27(9, 9, 9),
         "aniso": lambda: oracle.gen_aniso7(10, 0.02, 0.8),
         "jump": lambda: oracle.gen_vardiff27(10, 3, -3.0, 3.0, 4)}[gen]()
    S = oracle.strength(A, 0.25, 0.9)
    cf, _ = oracle.pmis(S)
    S2, cf2 = oracle.create_2nd_s(S, cf)
    Cpts = np.nonzero(cf == 1)[0]
    assert S2.nrows == S2.ncols == len(Cpts)
    Ss = S.to_scipy()
    Ss.data[:] = 1.0
    R = (Ss + Ss @ Ss).tocsr()[Cpts][:, Cpts].tolil()
    R.setdiag(0)
    R = R.tocsr()
    R.eliminate_zeros()
    got = {(r, c) for r, c in zip(*S2.sorted_coo()[:2])}
    ref = set(zip(*R.nonzero()))
    assert got == ref
    # no duplicates within a row
    assert len(got) == S2.nnz
    # C points with an empty row are marked 2, the others keep 1
    empty = np.diff(S2.ia) == 0
    assert np.array_equal(cf2[Cpts] == 2, empty)
    assert np.array_equal(cf2[cf != 1], cf[cf != 1])


def test_second_strength_first_seen_order(oracle):
    """Row order of the reference's marker loop on a small random graph,
    restated literally (par_strength.c:2433-2466)."""
    rng = np.random.default_rng(5)
    n = 40
    M = sp.random(n, n, density=0.08, random_state=rng, format="csr")
    M.setdiag(0)
    M.eliminate_zeros()
    S = oracle.Csr(n, n, M.indptr.astype(np.int32), M.indices.astype(np.int32), None)
    cf = np.where(rng.uniform(size=n) < 0.4, 1, -1).astype(np.int32)
    S2, _ = oracle.create_2nd_s(S, cf)
    f2c = np.cumsum(cf == 1) - 1
    rows = []
    for ic, i1 in enumerate(np.nonzero(cf == 1)[0]):
        row = []
        for i2 in S.ja[S.ia[i1]:S.ia[i1 + 1]]:
            if cf[i2] == 1 and f2c[i2] not in row:
                row.append(f2c[i2])
            for i3 in S.ja[S.ia[i2]:S.ia[i2 + 1]]:
                if cf[i3] == 1 and f2c[i3] != ic and f2c[i3] not in row:
                    row.append(f2c[i3])
        rows.append(row)
    assert [list(S2.ja[S2.ia[r]:S2.ia[r + 1]]) for r in range(S2.nrows)] == rows


# ----------------------------------------------------------- second coarsening
@pytest.mark.parametrize("gen", ["l7", "l27", "jump"])
def test_second_coarsening_invariants(oracle, gen):
    A = {"l7": lambda: oracle.gen_laplace7(12, 12, 12),
         "l27": lambda: oracle.gen_laplace27(10, 10, 10),
         "jump": lambda: oracle.gen_vardiff27(10, 3, -3.0, 3.0, 4)}[gen]()
    S = oracle.strength(A, 0.25, 0.9)
    cf, _ = oracle.pmis(S)
    S2, cf2 = oracle.create_2nd_s(S, cf)
    m = oracle.pmis_measure(S2)
    cfn, rounds = oracle.pmis_cf3(S2)
    cfn_m, rounds_m = oracle.pmis_cf3(S2, m)
    assert np.array_equal(cfn, cfn_m) and rounds == rounds_m
    assert set(np.unique(cfn)) <= {1, -1}
    Cn = cfn == 1
    # rows without connections are C (CF_init 3)
    assert np.all(Cn[np.diff(S2.ia) == 0])
    G = (S2.to_scipy() + S2.to_scipy().T).tocsr()
    for i in np.nonzero(Cn & (np.diff(S2.ia) > 0))[0]:
        nb = G.indices[G.indptr[i]:G.indptr[i + 1]]
        assert not np.any(Cn[nb])
    for i in np.nonzero(cfn == -1)[0]:
        if m[i] > 1:
            assert np.any(Cn[S2.ja[S2.ia[i]:S2.ia[i + 1]]])
    # the correction: first-pass C points take the second split's marks
    cfc = oracle.correct_cf_marker(cf2, cfn)
    Cpts = np.nonzero(cf == 1)[0]
    assert np.array_equal(cfc[Cpts], cfn)
    assert np.array_equal(cfc[cf != 1], cf[cf != 1])


def test_second_coarsening_skips_the_first_pick(oracle):
    """CF_init 3 on a path 0-1-2-3-4 (symmetric strength): round 0 only marks
    points of measure < 1 as F, so a point whose measure would win the first
    pick of CF_init 0 is decided one round later; with these measures the
    splits agree, and the round count is one larger."""
    ia = np.array([0, 1, 3, 5, 7, 8], np.int32)
    ja = np.array([1, 0, 2, 1, 3, 2, 4, 3], np.int32)
    S = oracle.Csr(5, 5, ia, ja, None)
    m = np.array([1.2, 2.9, 2.1, 2.8, 1.1])
    cf0, r0 = oracle.pmis(S, m)
    cf3, r3 = oracle.pmis_cf3(S, m)
    assert list(cf0) == [-1, 1, -1, 1, -1] == list(cf3)
    assert r3 == r0 + 1
    # a measure below 1 makes a point F in round 0; an isolated point is C
    ia = np.array([0, 1, 2, 2], np.int32)
    ja = np.array([1, 0], np.int32)
    S = oracle.Csr(3, 3, ia, ja, None)
    cf3, _ = oracle.pmis_cf3(S, np.array([0.5, 1.7, 0.3]))
    assert list(cf3) == [-1, 1, 1]


# ------------------------------------------------------------------ product
def test_matmul_against_scipy_and_first_seen_order(oracle):
    rng = np.random.default_rng(9)
    for (n, k, m) in [(30, 20, 15), (17, 17, 17)]:
        A = sp.random(n, k, density=0.2, random_state=rng, format="csr")
        B = sp.random(k, m, density=0.25, random_state=rng, format="csr")
        Ao = oracle.Csr(n, k, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.copy())
        Bo = oracle.Csr(k, m, B.indptr.astype(np.int32), B.indices.astype(np.int32), B.data.copy())
        Co = oracle.matmul(Ao, Bo)
        np.testing.assert_allclose(dense(Co), (A @ B).toarray(), rtol=1e-14, atol=1e-15)
        for i in range(n):
            seen = [i] if n == m else []
            for k2 in A.indices[A.indptr[i]:A.indptr[i + 1]]:
                for c in B.indices[B.indptr[k2]:B.indptr[k2 + 1]]:
                    if c not in seen:
                        seen.append(c)
            assert list(Co.ja[Co.ia[i]:Co.ia[i + 1]]) == seen
        if n == m:
            # square products open each row with its diagonal (value 0 unless a
            # product lands there)
            assert all(Co.ja[Co.ia[i]] == i for i in range(n))


# --------------------------------------------------------------- multipass
def test_multipass_chain_hand_weights(oracle):
    """7-node chain, one C point (node 3).  Passes: {3}, {2,4}, {1,5}, {0,6}.
    Interior rows: row sum of the off-diagonals -2, one strong neighbour in
    the previous pass (-1), diagonal 2, so the one-step weight is
    -(-1) * (-2) / ((-1) * 2) = 1; the end rows have off-diagonal sum -1 and
    weight 1/2.  Chained: P = [1/2, 1, 1, 1, 1, 1, 1/2]^T."""
    A = csr_from_scipy(oracle, chain(7))
    S = oracle.strength(A, 0.25, 1.0)
    cf = np.array([-1, -1, -1, 1, -1, -1, -1], np.int32)
    P, cfa = oracle.mod_multipass(A, S, cf)
    assert P.ncols == 1 and list(np.diff(P.ia)) == [1] * 7
    assert np.array_equal(dense(P).ravel(), [0.5, 1, 1, 1, 1, 1, 0.5])


def test_multipass_chain_two_c_points(oracle):
    """9-node chain, C at 2 and 6.  Pass 2: {1, 3, 5, 7}; node 4 has two
    pass-2 neighbours (3 and 5) and becomes pass 3 with one-step weights 1/2
    each; the ends 0 and 8 are pass 3 with weight 1/2 on 1 (resp. 7).  Node 3
    interpolates from C 2 only (its other strong neighbour, 4, is not C):
    factor (-2)/((-1)(2)) = 1, weight 1."""
    A = csr_from_scipy(oracle, chain(9))
    S = oracle.strength(A, 0.25, 1.0)
    cf = np.array([-1, -1, 1, -1, -1, -1, 1, -1, -1], np.int32)
    P, _ = oracle.mod_multipass(A, S, cf)
    ref = np.zeros((9, 2))
    ref[[0, 1, 2, 3], 0] = [0.5, 1, 1, 1]
    ref[4] = [0.5, 0.5]
    ref[[5, 6, 7, 8], 1] = [1, 1, 1, 0.5]
    assert np.array_equal(dense(P), ref)
    # the pass-3 row of node 4 lists its columns in first-seen order of
    # the product (through node 3 first)
    assert list(P.ja[P.ia[4]:P.ia[5]]) == [0, 1]


def test_multipass_preserves_constants_on_zero_row_sums(oracle):
    """On a graph Laplacian every row sums to zero, each one-step weight row
    then sums to 1 and so does every product: P 1 = 1 on all rows reached."""
    A = csr_from_scipy(oracle, graph_laplacian_3d(9))
    S = oracle.strength(A, 0.25, 1.0)
    cf, _ = oracle.pmis(S)
    S2, cf2 = oracle.create_2nd_s(S, cf)
    cfn, _ = oracle.pmis_cf3(S2)
    cfc = oracle.correct_cf_marker(cf2, cfn)
    P, _ = oracle.mod_multipass(A, S, cfc)
    rs = np.asarray(P.to_scipy().sum(axis=1)).ravel()
    assert np.all(np.diff(P.ia) > 0)
    np.testing.assert_allclose(rs, 1.0, rtol=0, atol=1e-14)
    # C rows are unit rows
    Cpts = np.nonzero(cfc == 1)[0]
    assert np.all(np.diff(P.ia)[Cpts] == 1)
    assert np.array_equal(P.ja[P.ia[Cpts]], np.arange(len(Cpts)))
    assert np.all(P.a[P.ia[Cpts]] == 1.0)


def test_multipass_truncation_is_the_interp_truncation(oracle):
    A = oracle.gen_laplace27(14, 14, 14)
    S = oracle.strength(A, 0.25, 0.9)
    cf, _ = oracle.pmis(S)
    S2, cf2 = oracle.create_2nd_s(S, cf)
    cfn, _ = oracle.pmis_cf3(S2)
    cfc = oracle.correct_cf_marker(cf2, cfn)
    P0, _ = oracle.mod_multipass(A, S, cfc)
    P4, _ = oracle.mod_multipass(A, S, cfc, 0.0, 4)
    ref = oracle.Csr(P0.nrows, P0.ncols, P0.ia.copy(), P0.ja.copy(), P0.a.copy())
    oracle.truncate(ref, 0.0, 4)
    assert np.array_equal(P4.ia, ref.ia) and np.array_equal(P4.ja, ref.ja)
    assert np.array_equal(P4.a, ref.a)
    assert np.diff(P4.ia).max() <= 4 < np.diff(P0.ia).max()


# --------------------------------------------------------------- hierarchy
@pytest.mark.parametrize("gen,agg", [("l7", 1), ("l27", 1), ("jump", 2)])
def test_aggressive_hierarchy_galerkin_and_solve(oracle, gen, agg):
    A = {"l7": lambda: oracle.gen_laplace7(14, 14, 14),
         "l27": lambda: oracle.gen_laplace27(12, 12, 12),
         "jump": lambda: oracle.gen_vardiff27(12, 3, -3.0, 3.0, 4)}[gen]()
    H = oracle.Hierarchy(A, agg_num_levels=agg)
    H0 = oracle.Hierarchy(A)
    # aggressive levels coarsen much more
    for lev in range(min(agg, H.num_levels - 1)):
        assert H.A(lev + 1).nrows < 0.5 * H0.A(lev + 1).nrows
    for lev in range(H.num_levels - 1):
        P, Ac, Af = dense(H.P(lev)), dense(H.A(lev + 1)), dense(H.A(lev))
        np.testing.assert_allclose(Ac, P.T @ Af @ P, rtol=1e-12, atol=1e-12 * np.abs(Af).max())
        assert np.array_equal(dense(H.R(lev)), P.T)
    b = 2 * oracle.rand_stream(3, A.nrows) - 1
    x, its, rr, hist, err = H.pcg(b, tol=1e-8)
    r = b - A.to_scipy() @ x
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-8
    assert its < 60
