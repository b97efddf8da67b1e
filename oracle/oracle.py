"""ctypes front end of the CPU oracle (test infrastructure only).

Loads ``oracle/liboracle.so`` (built from ``oracle/amg_oracle.c`` by
``oracle/Makefile``).  Only tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use this module, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class _Csr(C.Structure):
    _fields_ = [("nrows", C.c_int), ("ncols", C.c_int),
                ("ia", C.POINTER(C.c_int)), ("ja", C.POINTER(C.c_int)),
                ("a", C.POINTER(C.c_double))]


class _Params(C.Structure):
    _fields_ = [("strong_threshold", C.c_double), ("max_row_sum", C.c_double),
                ("trunc_factor", C.c_double), ("P_max_elmts", C.c_int),
                ("max_levels", C.c_int), ("max_coarse_size", C.c_int),
                ("relax_type", C.c_int), ("coarse_relax_type", C.c_int),
                ("relax_weight", C.c_double), ("num_sweeps", C.c_int),
                ("relax_up_type", C.c_int), ("outer_weight", C.c_double),
                ("hgs_parts", C.c_int), ("hgs_rows", C.c_int), ("coarsen_type", C.c_int),
                ("agg_num_levels", C.c_int), ("agg_interp_type", C.c_int),
                ("agg_trunc_factor", C.c_double), ("agg_P_max_elmts", C.c_int)]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER
        L.orc_rand_stream.argtypes = [C.c_int, C.c_int, P(C.c_double)]
        for name in ("orc_gen_laplace7",):
            getattr(L, name).argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                         C.c_double, P(_Csr)]
        L.orc_gen_laplace27.argtypes = [C.c_int, C.c_int, C.c_int, P(_Csr)]
        for name in ("orc_gen_vardiff7", "orc_gen_vardiff27"):
            getattr(L, name).argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                         P(_Csr)]
        L.orc_gen_aniso7.argtypes = [C.c_int, C.c_double, C.c_double, P(_Csr)]
        L.orc_csr_free.argtypes = [P(_Csr)]
        L.orc_strength.argtypes = [P(_Csr), C.c_double, C.c_double, P(_Csr)]
        L.orc_pmis_measure.argtypes = [P(_Csr), C.c_int, P(C.c_double)]
        L.orc_pmis_with_measure.argtypes = [P(_Csr), P(C.c_double), P(C.c_int)]
        L.orc_pmis.argtypes = [P(_Csr), P(C.c_int)]
        L.orc_rs_first_pass.argtypes = [P(_Csr), P(C.c_int)]
        L.orc_hmis_with_measure.argtypes = [P(_Csr), P(C.c_double), P(C.c_int)]
        L.orc_hmis.argtypes = [P(_Csr), P(C.c_int)]
        L.orc_create_2nd_s.argtypes = [P(_Csr), P(C.c_int), P(_Csr)]
        L.orc_pmis_cf3_with_measure.argtypes = [P(_Csr), P(C.c_double), P(C.c_int)]
        L.orc_pmis_cf3.argtypes = [P(_Csr), P(C.c_int)]
        L.orc_correct_cf_marker.argtypes = [C.c_int, P(C.c_int), P(C.c_int)]
        L.orc_correct_cf_marker.restype = None
        L.orc_matmul.argtypes = [P(_Csr), P(_Csr), P(_Csr)]
        L.orc_mod_multipass.argtypes = [P(_Csr), P(_Csr), P(C.c_int), C.c_double, C.c_int,
                                        P(_Csr)]
        L.orc_extpi_interp.argtypes = [P(_Csr), P(_Csr), P(C.c_int), C.c_double, C.c_int,
                                       P(_Csr)]
        L.orc_extpi_interp_raw.argtypes = [P(_Csr), P(_Csr), P(C.c_int), P(_Csr)]
        L.orc_truncate.argtypes = [P(_Csr), C.c_double, C.c_int]
        L.orc_qsort2_abs.argtypes = [P(C.c_int), P(C.c_double), C.c_int, C.c_int]
        L.orc_transpose.argtypes = [P(_Csr), P(_Csr)]
        L.orc_rap.argtypes = [P(_Csr), P(_Csr), P(_Csr), P(_Csr)]
        L.orc_l1_norms.argtypes = [P(_Csr), P(C.c_double)]
        L.orc_l1_norms_diag.argtypes = [P(_Csr), P(C.c_double)]
        L.orc_partition1d.argtypes = [C.c_int, C.c_int, C.c_int, P(C.c_int), P(C.c_int)]
        L.orc_hgs_parts.argtypes = [C.c_int, C.c_int, C.c_int]
        L.orc_hybrid_gs.argtypes = [P(_Csr), P(C.c_double), P(C.c_double), P(C.c_double),
                                    C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
        L.orc_matvec.argtypes = [C.c_double, P(_Csr), P(C.c_double), C.c_double,
                                 P(C.c_double), P(C.c_double)]
        L.orc_dot.argtypes = [C.c_int, P(C.c_double), P(C.c_double)]
        L.orc_dot.restype = C.c_double
        L.orc_amg_default_params.argtypes = [P(_Params)]
        L.orc_amg_setup.argtypes = [P(_Csr), P(_Params)]
        L.orc_amg_setup.restype = C.c_void_p
        L.orc_amg_free.argtypes = [C.c_void_p]
        L.orc_amg_num_levels.argtypes = [C.c_void_p]
        for name in ("orc_amg_A", "orc_amg_P", "orc_amg_R", "orc_amg_S"):
            getattr(L, name).argtypes = [C.c_void_p, C.c_int]
            getattr(L, name).restype = P(_Csr)
        L.orc_amg_CF.argtypes = [C.c_void_p, C.c_int]
        L.orc_amg_CF.restype = P(C.c_int)
        L.orc_amg_l1.argtypes = [C.c_void_p, C.c_int]
        L.orc_amg_l1.restype = P(C.c_double)
        L.orc_amg_pmis_rounds.argtypes = [C.c_void_p, C.c_int]
        L.orc_amg_coarse_relax.argtypes = [C.c_void_p]
        L.orc_amg_setup_stats.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_char_p, C.c_int]
        L.orc_amg_cycle.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double)]
        L.orc_amg_solve.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), C.c_double,
                                    C.c_int, P(C.c_int), P(C.c_double), P(C.c_double)]
        L.orc_pcg_solve.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), C.c_double,
                                    C.c_int, P(C.c_int), P(C.c_double), P(C.c_double)]
        _lib = L
    return _lib


def _dp(x):
    return x.ctypes.data_as(C.POINTER(C.c_double))


def _ip(x):
    return x.ctypes.data_as(C.POINTER(C.c_int))


@dataclass
class Csr:
    """CSR matrix with int32 row pointers / columns and float64 values."""
    nrows: int
    ncols: int
    ia: np.ndarray
    ja: np.ndarray
    a: np.ndarray | None

    @property
    def nnz(self) -> int:
        return int(self.ia[self.nrows])

    def to_c(self):
        self.ia = np.ascontiguousarray(self.ia, dtype=np.int32)
        self.ja = np.ascontiguousarray(self.ja, dtype=np.int32)
        s = _Csr()
        s.nrows, s.ncols = self.nrows, self.ncols
        s.ia, s.ja = _ip(self.ia), _ip(self.ja)
        if self.a is not None:
            self.a = np.ascontiguousarray(self.a, dtype=np.float64)
            s.a = _dp(self.a)
        else:
            s.a = C.POINTER(C.c_double)()
        return s

    def to_scipy(self):
        import scipy.sparse as sp
        data = self.a if self.a is not None else np.ones(self.nnz)
        return sp.csr_matrix((data, self.ja, self.ia), shape=(self.nrows, self.ncols))

    def sorted_coo(self):
        """(rows, cols, vals) sorted by (row, col): the order-free pattern."""
        rows = np.repeat(np.arange(self.nrows, dtype=np.int64), np.diff(self.ia))
        cols = self.ja[: self.nnz].astype(np.int64)
        order = np.lexsort((cols, rows))
        vals = None if self.a is None else self.a[: self.nnz][order]
        return rows[order], cols[order], vals


def _from_c(s: _Csr, free: bool = True, with_values: bool = True) -> Csr:
    n, ncols = s.nrows, s.ncols
    ia = np.ctypeslib.as_array(s.ia, shape=(n + 1,)).copy()
    nnz = int(ia[n])
    ja = np.ctypeslib.as_array(s.ja, shape=(max(nnz, 1),))[:nnz].copy() if nnz else np.zeros(0, np.int32)
    a = None
    if with_values and bool(s.a):
        a = np.ctypeslib.as_array(s.a, shape=(max(nnz, 1),))[:nnz].copy() if nnz else np.zeros(0)
    if free:
        lib().orc_csr_free(C.byref(s))
    return Csr(n, ncols, ia, ja, a)


# ---------------------------------------------------------------- generators
def num_threads() -> int:
    """Threads of the oracle's OpenMP row loops."""
    return int(lib().orc_num_threads())


def rand_stream(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.float64)
    lib().orc_rand_stream(seed, n, _dp(out))
    return out


def gen_laplace7(nx, ny, nz, cx=1.0, cy=1.0, cz=1.0) -> Csr:
    s = _Csr()
    lib().orc_gen_laplace7(nx, ny, nz, cx, cy, cz, C.byref(s))
    return _from_c(s)


def gen_laplace27(nx, ny, nz) -> Csr:
    s = _Csr()
    lib().orc_gen_laplace27(nx, ny, nz, C.byref(s))
    return _from_c(s)


def gen_vardiff7(N, cell, lo, hi, seed) -> Csr:
    s = _Csr()
    lib().orc_gen_vardiff7(N, cell, lo, hi, seed, C.byref(s))
    return _from_c(s)


def gen_vardiff27(N, cell, lo, hi, seed) -> Csr:
    s = _Csr()
    lib().orc_gen_vardiff27(N, cell, lo, hi, seed, C.byref(s))
    return _from_c(s)


def gen_aniso7(N, eps, theta) -> Csr:
    s = _Csr()
    lib().orc_gen_aniso7(N, eps, theta, C.byref(s))
    return _from_c(s)


# ---------------------------------------------------------------- setup steps
def strength(A: Csr, theta=0.25, max_row_sum=0.9) -> Csr:
    ca, s = A.to_c(), _Csr()
    lib().orc_strength(C.byref(ca), theta, max_row_sum, C.byref(s))
    return _from_c(s, with_values=False)


def pmis_measure(S: Csr, seed=2747) -> np.ndarray:
    m = np.empty(S.nrows, dtype=np.float64)
    cs = S.to_c()
    lib().orc_pmis_measure(C.byref(cs), seed, _dp(m))
    return m


def pmis(S: Csr, measure: np.ndarray | None = None):
    """Returns (CF, rounds); CF uses 1 = C, -1 = F, -3 = isolated F."""
    cf = np.empty(S.nrows, dtype=np.int32)
    cs = S.to_c()
    if measure is None:
        r = lib().orc_pmis(C.byref(cs), _ip(cf))
    else:
        m = np.ascontiguousarray(measure, dtype=np.float64)
        r = lib().orc_pmis_with_measure(C.byref(cs), _dp(m), _ip(cf))
    return cf, r


def rs_first_pass(S: Csr) -> np.ndarray:
    """The Ruge-Stueben first pass of HMIS: 1 = C, -1 = F, -2 = Z, -3 = SF."""
    cf = np.empty(S.nrows, dtype=np.int32)
    cs = S.to_c()
    lib().orc_rs_first_pass(C.byref(cs), _ip(cf))
    return cf


def hmis(S: Csr, measure: np.ndarray | None = None):
    """HMIS (first pass + PMIS with CF_init 1). Returns (CF, PMIS rounds)."""
    cf = np.empty(S.nrows, dtype=np.int32)
    cs = S.to_c()
    if measure is None:
        r = lib().orc_hmis(C.byref(cs), _ip(cf))
    else:
        m = np.ascontiguousarray(measure, dtype=np.float64)
        r = lib().orc_hmis_with_measure(C.byref(cs), _dp(m), _ip(cf))
    return cf, r


def create_2nd_s(S: Csr, CF: np.ndarray):
    """Second strength matrix of aggressive coarsening over the C points of
    CF.  Returns (S2 pattern, CF with 2 on C points whose S2 row is empty)."""
    cf = np.ascontiguousarray(CF, dtype=np.int32).copy()
    cs, out = S.to_c(), _Csr()
    lib().orc_create_2nd_s(C.byref(cs), _ip(cf), C.byref(out))
    return _from_c(out, with_values=False), cf


def pmis_cf3(S: Csr, measure: np.ndarray | None = None):
    """PMIS with CF_init 3 (rows without strong connections become C).
    Returns (CF, rounds)."""
    cf = np.empty(S.nrows, dtype=np.int32)
    cs = S.to_c()
    if measure is None:
        r = lib().orc_pmis_cf3(C.byref(cs), _ip(cf))
    else:
        m = np.ascontiguousarray(measure, dtype=np.float64)
        r = lib().orc_pmis_cf3_with_measure(C.byref(cs), _dp(m), _ip(cf))
    return cf, r


def correct_cf_marker(CF: np.ndarray, CFN: np.ndarray) -> np.ndarray:
    cf = np.ascontiguousarray(CF, dtype=np.int32).copy()
    cfn = np.ascontiguousarray(CFN, dtype=np.int32)
    lib().orc_correct_cf_marker(len(cf), _ip(cf), _ip(cfn))
    return cf


def matmul(A: Csr, B: Csr) -> Csr:
    ca, cb, out = A.to_c(), B.to_c(), _Csr()
    lib().orc_matmul(C.byref(ca), C.byref(cb), C.byref(out))
    return _from_c(out)


def mod_multipass(A: Csr, S: Csr, CF: np.ndarray, trunc_factor=0.0, max_elmts=0):
    """Modified multipass interpolation (agg_interp_type 8).  Returns
    (P, CF_after) with -3 mapped to -1."""
    cf = np.ascontiguousarray(CF, dtype=np.int32).copy()
    ca, cs, p = A.to_c(), S.to_c(), _Csr()
    rc = lib().orc_mod_multipass(C.byref(ca), C.byref(cs), _ip(cf), trunc_factor, max_elmts,
                                 C.byref(p))
    if rc != 0:
        raise RuntimeError("multipass interpolation: more than 9 passes")
    return _from_c(p), cf


def extpi_interp(A: Csr, S: Csr, CF: np.ndarray, trunc_factor=0.0, max_elmts=4):
    """Returns (P, CF_after) -- CF_after has -3 mapped to -1 like the reference."""
    cf = np.ascontiguousarray(CF, dtype=np.int32).copy()
    ca, cs, p = A.to_c(), S.to_c(), _Csr()
    lib().orc_extpi_interp(C.byref(ca), C.byref(cs), _ip(cf), trunc_factor, max_elmts,
                           C.byref(p))
    return _from_c(p), cf


def extpi_interp_raw(A: Csr, S: Csr, CF: np.ndarray) -> Csr:
    cf = np.ascontiguousarray(CF, dtype=np.int32)
    ca, cs, p = A.to_c(), S.to_c(), _Csr()
    lib().orc_extpi_interp_raw(C.byref(ca), C.byref(cs), _ip(cf), C.byref(p))
    return _from_c(p)


def truncate(P: Csr, trunc_factor=0.0, max_elmts=4) -> Csr:
    """hypre_ParCSRMatrixTruncate (rescale 1) of P, in place; returns P."""
    libc = C.CDLL(None)
    libc.malloc.restype = C.c_void_p
    libc.malloc.argtypes = [C.c_size_t]
    cp = _Csr()
    cp.nrows, cp.ncols = P.nrows, P.ncols
    # orc_truncate replaces the arrays with malloc'd ones and frees the old:
    # hand it malloc'd copies
    bufs = []
    for arr, ty in ((np.ascontiguousarray(P.ia, np.int32), C.c_int),
                    (np.ascontiguousarray(P.ja, np.int32), C.c_int),
                    (np.ascontiguousarray(P.a, np.float64), C.c_double)):
        ptr = libc.malloc(max(arr.nbytes, 8))
        C.memmove(ptr, arr.ctypes.data, arr.nbytes)
        bufs.append(C.cast(ptr, C.POINTER(ty)))
    cp.ia, cp.ja, cp.a = bufs
    lib().orc_truncate(C.byref(cp), trunc_factor, max_elmts)
    out = _from_c(cp, free=True)
    P.ia, P.ja, P.a = out.ia, out.ja, out.a
    return P


def qsort2_abs(v: np.ndarray, w: np.ndarray):
    v = np.ascontiguousarray(v, dtype=np.int32).copy()
    w = np.ascontiguousarray(w, dtype=np.float64).copy()
    if len(v):
        lib().orc_qsort2_abs(_ip(v), _dp(w), 0, len(v) - 1)
    return v, w


def transpose(P: Csr) -> Csr:
    cp, r = P.to_c(), _Csr()
    lib().orc_transpose(C.byref(cp), C.byref(r))
    return _from_c(r)


def rap(R: Csr, A: Csr, P: Csr) -> Csr:
    cr, ca, cp, out = R.to_c(), A.to_c(), P.to_c(), _Csr()
    lib().orc_rap(C.byref(cr), C.byref(ca), C.byref(cp), C.byref(out))
    return _from_c(out)


def l1_norms(A: Csr) -> np.ndarray:
    out = np.empty(A.nrows, dtype=np.float64)
    ca = A.to_c()
    lib().orc_l1_norms(C.byref(ca), _dp(out))
    return out


def l1_norms_diag(A: Csr) -> np.ndarray:
    out = np.empty(A.nrows, dtype=np.float64)
    ca = A.to_c()
    lib().orc_l1_norms_diag(C.byref(ca), _dp(out))
    return out


def partition1d(n: int, p: int, j: int):
    s, e = C.c_int(), C.c_int()
    lib().orc_partition1d(n, p, j, C.byref(s), C.byref(e))
    return s.value, e.value


def hgs_parts(n: int, parts: int = 0, rows: int = 64) -> int:
    return lib().orc_hgs_parts(n, parts, rows)


def hybrid_gs(A: Csr, f, u, parts: int, direction: int = 1, symm: bool = False,
              skip_diag: int = 0, l1=None, relax_weight: float = 1.0,
              omega: float = 1.0) -> np.ndarray:
    """One hybrid Gauss-Seidel relaxation; returns the new u (input kept)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.ascontiguousarray(u, dtype=np.float64).copy()
    ca = A.to_c()
    l1p = C.POINTER(C.c_double)()
    if l1 is not None:
        l1 = np.ascontiguousarray(l1, dtype=np.float64)
        l1p = _dp(l1)
    lib().orc_hybrid_gs(C.byref(ca), _dp(f), l1p, _dp(out), parts, direction, int(symm),
                        skip_diag, relax_weight, omega)
    return out


def matvec(alpha, A: Csr, x, beta=0.0, b=None) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(A.nrows, dtype=np.float64)
    bb = np.ascontiguousarray(b if b is not None else y, dtype=np.float64)
    ca = A.to_c()
    lib().orc_matvec(alpha, C.byref(ca), _dp(x), beta, _dp(bb), _dp(y))
    return y


# ---------------------------------------------------------------- full AMG
def default_params(**kw) -> _Params:
    p = _Params()
    lib().orc_amg_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Hierarchy:
    """AMG hierarchy built by the oracle (owns the C object)."""

    def __init__(self, A: Csr, **params):
        self._A = A
        self._ca = A.to_c()
        self.params = default_params(**params)
        self._h = lib().orc_amg_setup(C.byref(self._ca), C.byref(self.params))
        self.num_levels = lib().orc_amg_num_levels(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_amg_free(self._h)
            self._h = None

    def A(self, lev) -> Csr:
        return _from_c(lib().orc_amg_A(self._h, lev).contents, free=False)

    def P(self, lev) -> Csr:
        return _from_c(lib().orc_amg_P(self._h, lev).contents, free=False)

    def R(self, lev) -> Csr:
        return _from_c(lib().orc_amg_R(self._h, lev).contents, free=False)

    def S(self, lev) -> Csr:
        return _from_c(lib().orc_amg_S(self._h, lev).contents, free=False, with_values=False)

    def CF(self, lev) -> np.ndarray:
        n = self.A(lev).nrows
        return np.ctypeslib.as_array(lib().orc_amg_CF(self._h, lev), shape=(n,)).copy()

    def l1(self, lev) -> np.ndarray:
        n = self.A(lev).nrows
        return np.ctypeslib.as_array(lib().orc_amg_l1(self._h, lev), shape=(n,)).copy()

    def pmis_rounds(self, lev) -> int:
        return lib().orc_amg_pmis_rounds(self._h, lev)

    def complexities(self):
        rows = [self.A(l).nrows for l in range(self.num_levels)]
        nnz = [self.A(l).nnz for l in range(self.num_levels)]
        return sum(rows) / rows[0], sum(nnz) / nnz[0]

    def setup_stats(self, max_iter=20, tol=1e-9) -> str:
        """The setup statistics text (src/parcsr_ls/par_stats.c:20) for this
        hierarchy, with the solver's max_iter and tol in its last part."""
        need = lib().orc_amg_setup_stats(self._h, max_iter, tol, None, 0)
        buf = C.create_string_buffer(need + 1)
        lib().orc_amg_setup_stats(self._h, max_iter, tol, buf, need + 1)
        return buf.value.decode()

    def cycle(self, f, u=None):
        f = np.ascontiguousarray(f, dtype=np.float64)
        u = np.zeros_like(f) if u is None else np.ascontiguousarray(u, dtype=np.float64).copy()
        lib().orc_amg_cycle(self._h, _dp(f), _dp(u))
        return u

    def pcg(self, b, tol=1e-8, max_iter=1000, x0=None):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).copy()
        it, rr = C.c_int(0), C.c_double(0)
        hist = np.zeros(max_iter + 1)
        err = lib().orc_pcg_solve(self._h, _dp(b), _dp(x), tol, max_iter, C.byref(it),
                                  C.byref(rr), _dp(hist))
        return x, it.value, rr.value, hist[1: it.value + 1].copy(), err

    def amg_solve(self, b, tol=1e-8, max_iter=100, x0=None):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).copy()
        it, rr, cv = C.c_int(0), C.c_double(0), C.c_double(0)
        err = lib().orc_amg_solve(self._h, _dp(b), _dp(x), tol, max_iter, C.byref(it),
                                  C.byref(rr), C.byref(cv))
        return x, it.value, rr.value, cv.value, err


def ij_rhsrand(n: int) -> np.ndarray:
    """RHS of the reference driver's -rhsrand option: 2*rand-1 from seed 22775,
    scaled to unit 2-norm (src/test/ij.c:3857-3860)."""
    b = 2.0 * rand_stream(22775, n) - 1.0
    nrm = float(np.dot(b, b))
    return b * (1.0 / np.sqrt(nrm))
