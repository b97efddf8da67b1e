"""Pin the oracle's HMIS coarsening (CPU only).

HMIS (src/parcsr_ls/par_coarsen.c:2842) is the Ruge-Stueben first pass of
hypre_BoomerAMGCoarsenRuge (par_coarsen.c:911, coarsen_type 10: points nobody
depends on become Z_PT, the routine stops after the first pass) followed by
PMIS with CF_init 1 (the first pass's C points are the first independent
set).  The reference has no single-rank HMIS run on record, so the oracle is
pinned by:

* hand-traced first passes (1D chain, 3x3 five-point grid);
* a second, literal restatement of the first pass in Python below, with the
  reference's list of lists (src/utilities/amg_linklist.c) kept as a Python
  list of [measure, points] pairs, compared on random strength graphs;
* invariants of the split.
"""
import numpy as np
import pytest
import scipy.sparse as sp

C, F, Z, SF = 1, -1, -2, -3


def first_pass_py(n, Si, Sj):
    """par_coarsen.c:1060-1391 for one rank, measure_type 0, coarsen_type 10."""
    ST = [[] for _ in range(n)]
    for i in range(n):
        for k in range(Si[i], Si[i + 1]):
            ST[Sj[k]].append(i)
    measure = [len(ST[i]) for i in range(n)]
    lol = []                                        # [[measure, [points...]], ...] descending

    def enter(m, i):
        for pos, (mm, pts) in enumerate(lol):
            if m > mm:
                lol.insert(pos, [m, [i]])
                return
            if m == mm:
                pts.append(i)
                return
        lol.append([m, [i]])

    def remove(m, i):
        for pos, (mm, pts) in enumerate(lol):
            if mm == m:
                pts.remove(i)
                if not pts:
                    del lol[pos]
                return
        raise AssertionError("no such list")

    CF = [0] * n
    num_left = 0
    for j in range(n):
        if Si[j + 1] == Si[j]:
            CF[j], measure[j] = SF, 0
        else:
            num_left += 1
    for j in range(n):
        if CF[j] == SF:
            continue
        if measure[j] > 0:
            enter(measure[j], j)
            continue
        CF[j] = Z
        for k in range(Si[j], Si[j + 1]):
            nb = Sj[k]
            if CF[nb] == SF:
                continue
            if nb < j:
                if measure[nb] > 0:
                    remove(measure[nb], nb)
                measure[nb] += 1
                enter(measure[nb], nb)
            else:
                measure[nb] += 1
        num_left -= 1
    while num_left > 0:
        index = lol[0][1][0]
        CF[index] = C
        remove(measure[index], index)
        measure[index] = 0
        num_left -= 1
        for nb in ST[index]:
            if CF[nb] != 0:
                continue
            CF[nb] = F
            remove(measure[nb], nb)
            num_left -= 1
            for k in range(Si[nb], Si[nb + 1]):
                nb2 = Sj[k]
                if CF[nb2] == 0:
                    remove(measure[nb2], nb2)
                    measure[nb2] += 1
                    enter(measure[nb2], nb2)
        for k in range(Si[index], Si[index + 1]):
            nb = Sj[k]
            if CF[nb] != 0:
                continue
            remove(measure[nb], nb)
            measure[nb] -= 1
            if measure[nb] > 0:
                enter(measure[nb], nb)
            else:
                CF[nb] = F
                num_left -= 1
                for k2 in range(Si[nb], Si[nb + 1]):
                    nb2 = Sj[k2]
                    if CF[nb2] == 0:
                        remove(measure[nb2], nb2)
                        measure[nb2] += 1
                        enter(measure[nb2], nb2)
    return np.array(CF, dtype=np.int32)


def test_first_pass_hand_traces(oracle):
    # 1D chain: the ends have measure 1, the interior 2; point 1 is taken
    # first and every other point follows
    for n, want in ((6, [F, C, F, C, F, C]), (7, [F, C, F, C, F, C, F])):
        S = oracle.strength(oracle.gen_laplace7(n, 1, 1))
        assert list(oracle.rs_first_pass(S)) == want
    # 3x3 five-point grid: the centre (measure 4) first, then the corners in
    # the order they reach measure 4 (0, 2, 6, 8)
    S = oracle.strength(oracle.gen_laplace7(3, 3, 1))
    assert list(oracle.rs_first_pass(S)) == [C, F, C, F, C, F, C, F, C]


def random_strength(oracle, n, density, seed, symmetric):
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=density, random_state=rng, format="csr")
    if symmetric:
        M = M + M.T
    M = M.tolil()
    M.setdiag(0)
    M = M.tocsr()
    M.eliminate_zeros()
    M.sort_indices()
    return oracle.Csr(n, n, M.indptr.astype(np.int32), M.indices.astype(np.int32), None)


@pytest.mark.parametrize("seed", range(12))
def test_first_pass_matches_literal_restatement(oracle, seed):
    n = 20 + 7 * seed
    S = random_strength(oracle, n, 0.08 + 0.01 * (seed % 4), seed, symmetric=seed % 2 == 0)
    want = first_pass_py(n, list(S.ia), list(S.ja))
    assert np.array_equal(oracle.rs_first_pass(S), want)


@pytest.mark.parametrize("gen", ["l7", "l27", "aniso", "jump"])
def test_first_pass_matches_literal_restatement_on_stencils(oracle, gen):
    A = {"l7": lambda: oracle.gen_laplace7(7, 6, 5),
         "l27": lambda: oracle.gen_laplace27(6, 6, 6),
         "aniso": lambda: oracle.gen_aniso7(7, 0.02, 0.4),
         "jump": lambda: oracle.gen_vardiff27(7, 3, -3.0, 3.0, 4)}[gen]()
    S = oracle.strength(A)
    assert np.array_equal(oracle.rs_first_pass(S), first_pass_py(S.nrows, list(S.ia), list(S.ja)))


@pytest.mark.parametrize("gen", ["l7", "l27", "jump"])
def test_hmis_invariants(oracle, gen):
    A = {"l7": lambda: oracle.gen_laplace7(12, 12, 12),
         "l27": lambda: oracle.gen_laplace27(10, 10, 10),
         "jump": lambda: oracle.gen_vardiff27(12, 4, -3.0, 3.0, 2)}[gen]()
    S = oracle.strength(A)
    n = S.nrows
    first = oracle.rs_first_pass(S)
    Sm = sp.csr_matrix((np.ones(len(S.ja)), S.ja, S.ia), shape=(n, n))
    dependents = np.asarray(Sm.sum(axis=0)).ravel()
    rows = [set(S.ja[S.ia[i]:S.ia[i + 1]]) for i in range(n)]
    for i in range(n):
        if first[i] == F:                    # F points depend strongly on a C point
            assert any(first[j] == C for j in rows[i])
        if first[i] == Z:                    # nobody depends on a Z point
            assert dependents[i] == 0
        if first[i] == SF:
            assert not rows[i]
    assert set(np.unique(first)) <= {C, F, Z, SF}
    cf, rounds = oracle.hmis(S)
    # the first pass's C points stay C (each has a dependent, so its PMIS
    # measure exceeds 1); every point with strong connections that ends F has
    # a strong C neighbour; Z points without strong connections end F
    assert np.all(cf[first == C] == C)
    assert set(np.unique(cf)) <= {C, F, SF}
    for i in range(n):
        if cf[i] == F and rows[i]:
            assert any(cf[j] == C for j in rows[i])
    assert rounds >= 1


@pytest.mark.parametrize("gen", ["l7", "l27", "aniso", "jump", "rand"])
def test_library_first_pass_matches_oracle(oracle, gen):
    """The library's host first pass (bucket FIFOs, no GPU involved) gives the
    oracle's split (the reference's list of lists) bit for bit."""
    from hypre_amd import api as hx
    if gen == "rand":
        S = random_strength(oracle, 300, 0.03, 99, symmetric=False)
    else:
        A = {"l7": lambda: oracle.gen_laplace7(14, 13, 12),
             "l27": lambda: oracle.gen_laplace27(12, 12, 12),
             "aniso": lambda: oracle.gen_aniso7(14, 0.02, 0.4),
             "jump": lambda: oracle.gen_vardiff27(13, 3, -3.0, 3.0, 4)}[gen]()
        S = oracle.strength(A)
    hs = hx.HostCsr(S.nrows, S.ncols, np.ascontiguousarray(S.ia, np.int32),
                    np.ascontiguousarray(S.ja, np.int32), None)
    assert np.array_equal(hx.rs_first_pass(hs), oracle.rs_first_pass(S))
