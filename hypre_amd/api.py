"""Python wrappers over the C-ABI of libhypre_mi355x.so.

Names and argument meanings follow hypre's C API (HYPRE_IJMatrix*,
HYPRE_BoomerAMG*, HYPRE_ParCSRPCG*); each call goes straight to the library.
Errors raise ``HypreError`` carrying hypre's error bitmask.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .lib import fnptr, lib

HYPRE_PARCSR = 5555
ERROR_GENERIC, ERROR_MEMORY, ERROR_ARG, ERROR_CONV = 1, 2, 4, 256


class HypreError(RuntimeError):
    def __init__(self, code: int, where: str):
        buf = C.create_string_buffer(256)
        lib().HYPRE_DescribeError(code, buf)
        super().__init__(f"{where}: hypre error {code} {buf.value.decode()}")
        self.code = code


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def check(ret: int, where: str, allow: int = 0):
    """Raise on a non-zero error flag (bits in `allow` are cleared and ignored)."""
    if ret:
        lib().HYPRE_ClearAllErrors()      # hypre's flag is sticky: clear it either way
        if ret & ~allow:
            raise HypreError(ret, where)
    return ret


def call(name: str, *args, allow: int = 0):
    return check(getattr(lib(), name)(*args), name, allow)


def clear_errors():
    lib().HYPRE_ClearAllErrors()


def get_error() -> int:
    return lib().HYPRE_GetError()


def initialize():
    call("HYPRE_Initialize")


# ------------------------------------------------------------------ matrices
@dataclass
class HostCsr:
    nrows: int
    ncols: int
    ia: np.ndarray
    ja: np.ndarray
    a: np.ndarray | None

    @property
    def nnz(self) -> int:
        return int(self.ia[self.nrows])


class ParCSRMatrix:
    """Owning wrapper of a HYPRE_ParCSRMatrix (unless borrowed from an IJ object)."""

    def __init__(self, handle: int, owned: bool = True, keepalive=None):
        if not handle:
            raise HypreError(get_error(), "matrix construction")
        self.h = C.c_void_p(handle)
        self.owned = owned
        self._keep = keepalive

    @classmethod
    def from_csr(cls, ia, ja, a, ncols=None):
        ia = np.ascontiguousarray(ia, dtype=np.int32)
        ja = np.ascontiguousarray(ja, dtype=np.int32)
        a = np.ascontiguousarray(a, dtype=np.float64)
        n = len(ia) - 1
        h = lib().HYPRE_MI355X_ParCSRMatrixFromHostCSR(n, n if ncols is None else ncols, _ip(ia),
                                                       _ip(ja), _dp(a))
        return cls(h)

    @classmethod
    def laplacian7(cls, nx, ny, nz, cx=1.0, cy=1.0, cz=1.0):
        v0 = 0.0
        if nx > 1:
            v0 += 2.0 * cx
        if ny > 1:
            v0 += 2.0 * cy
        if nz > 1:
            v0 += 2.0 * cz
        vals = np.array([v0, -cx, -cy, -cz], dtype=np.float64)
        return cls(lib().GenerateLaplacian(0, nx, ny, nz, 1, 1, 1, 0, 0, 0, _dp(vals)))

    @classmethod
    def laplacian27(cls, nx, ny, nz):
        v0 = 26.0
        if nx == 1 or ny == 1 or nz == 1:
            v0 = 8.0
        if nx * ny == 1 or nx * nz == 1 or ny * nz == 1:
            v0 = 2.0
        vals = np.array([v0, -1.0], dtype=np.float64)
        return cls(lib().GenerateLaplacian27pt(0, nx, ny, nz, 1, 1, 1, 0, 0, 0, _dp(vals)))

    @classmethod
    def var_diffusion(cls, N, stencil, cell, lo, hi, seed):
        return cls(lib().HYPRE_MI355X_GenerateVarDiffusion(N, stencil, cell, lo, hi, seed))

    @classmethod
    def aniso7(cls, N, eps, theta):
        return cls(lib().HYPRE_MI355X_GenerateAniso7(N, eps, theta))

    def dims(self):
        n, nnz = C.c_int(), C.c_int()
        call("HYPRE_MI355X_ParCSRMatrixGetDevice", self.h, C.byref(n), C.byref(nnz), None, None, None)
        return n.value, nnz.value

    def to_host(self) -> HostCsr:
        n, nnz = self.dims()
        ia = np.empty(n + 1, np.int32)
        ja = np.empty(max(nnz, 1), np.int32)
        a = np.empty(max(nnz, 1), np.float64)
        call("HYPRE_MI355X_ParCSRMatrixCopyToHost", self.h, _ip(ia), _ip(ja), _dp(a))
        return HostCsr(n, n, ia, ja[:nnz], a[:nnz])

    def matvec(self, alpha, x: "ParVector", beta, y: "ParVector"):
        call("HYPRE_ParCSRMatrixMatvec", alpha, self.h, x.h, beta, y.h)

    def destroy(self):
        if self.owned and self.h:
            lib().HYPRE_ParCSRMatrixDestroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class ParVector:
    def __init__(self, n: int, values=None):
        h = C.c_void_p()
        call("HYPRE_ParVectorCreate", 0, n, None, C.byref(h))
        call("HYPRE_ParVectorInitialize", h)
        self.h = h
        self.n = n
        if values is not None:
            self.set(values)

    def set(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert v.shape == (self.n,)
        call("HYPRE_MI355X_ParVectorCopyFromHost", self.h, _dp(v))

    def get(self) -> np.ndarray:
        out = np.empty(self.n, np.float64)
        call("HYPRE_MI355X_ParVectorCopyToHost", self.h, _dp(out))
        return out

    def set_random(self, seed: int):
        call("HYPRE_ParVectorSetRandomValues", self.h, seed)

    def fill(self, value: float):
        call("HYPRE_ParVectorSetConstantValues", self.h, value)

    def scale(self, value: float):
        call("HYPRE_ParVectorScale", value, self.h)

    def dot(self, other: "ParVector") -> float:
        d = C.c_double()
        call("HYPRE_ParVectorInnerProd", self.h, other.h, C.byref(d))
        return d.value

    def copy_to(self, other: "ParVector"):
        """other <- self (HYPRE_ParVectorCopy)."""
        call("HYPRE_ParVectorCopy", self.h, other.h)

    def destroy(self):
        if self.h:
            lib().HYPRE_ParVectorDestroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ------------------------------------------------------------------ IJ
class IJMatrix:
    @classmethod
    def _wrap(cls, h):
        m = cls.__new__(cls)
        m.h = h
        return m

    def print(self, path: str, binary: bool = False):
        call("HYPRE_IJMatrixPrintBinary" if binary else "HYPRE_IJMatrixPrint", self.h,
             path.encode())

    def __init__(self, ilower, iupper, jlower=None, jupper=None):
        h = C.c_void_p()
        call("HYPRE_IJMatrixCreate", 0, ilower, iupper, ilower if jlower is None else jlower,
             iupper if jupper is None else jupper, C.byref(h))
        self.h = h
        call("HYPRE_IJMatrixSetObjectType", h, HYPRE_PARCSR)

    def initialize(self):
        call("HYPRE_IJMatrixInitialize", self.h)

    def set_diag_offd_sizes(self, diag_sizes):
        d = np.ascontiguousarray(diag_sizes, dtype=np.int32)
        o = np.zeros_like(d)
        call("HYPRE_IJMatrixSetDiagOffdSizes", self.h, _ip(d), _ip(o))

    def _values(self, fname, rows, ncols, cols, values):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        ncols = np.ascontiguousarray(ncols, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        values = np.ascontiguousarray(values, dtype=np.float64)
        call(fname, self.h, len(rows), _ip(ncols), _ip(rows), _ip(cols), _dp(values))

    def set_values(self, rows, ncols, cols, values):
        self._values("HYPRE_IJMatrixSetValues", rows, ncols, cols, values)

    def add_to_values(self, rows, ncols, cols, values):
        self._values("HYPRE_IJMatrixAddToValues", rows, ncols, cols, values)

    def set_csr(self, ia, ja, a):
        ia = np.asarray(ia)
        n = len(ia) - 1
        self.set_values(np.arange(n), np.diff(ia), ja, a)

    def assemble(self):
        call("HYPRE_IJMatrixAssemble", self.h)

    def get_object(self) -> ParCSRMatrix:
        obj = C.c_void_p()
        call("HYPRE_IJMatrixGetObject", self.h, C.byref(obj))
        return ParCSRMatrix(obj.value, owned=False, keepalive=self)

    def destroy(self):
        if self.h:
            lib().HYPRE_IJMatrixDestroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class IJVector:
    @classmethod
    def _wrap(cls, h):
        v = cls.__new__(cls)
        v.h = h
        lo, hi = C.c_int(), C.c_int()
        call("HYPRE_IJVectorGetLocalRange", h, C.byref(lo), C.byref(hi))
        v.n = hi.value - lo.value + 1
        return v

    def print(self, path: str, binary: bool = False):
        call("HYPRE_IJVectorPrintBinary" if binary else "HYPRE_IJVectorPrint", self.h,
             path.encode())

    def __init__(self, jlower, jupper):
        h = C.c_void_p()
        call("HYPRE_IJVectorCreate", 0, jlower, jupper, C.byref(h))
        self.h = h
        self.n = jupper - jlower + 1
        call("HYPRE_IJVectorSetObjectType", h, HYPRE_PARCSR)
        call("HYPRE_IJVectorInitialize", h)

    def set_values(self, values, indices=None):
        v = np.ascontiguousarray(values, dtype=np.float64)
        idx = None if indices is None else np.ascontiguousarray(indices, dtype=np.int32)
        call("HYPRE_IJVectorSetValues", self.h, len(v), None if idx is None else _ip(idx), _dp(v))

    def add_to_values(self, values, indices):
        v = np.ascontiguousarray(values, dtype=np.float64)
        idx = np.ascontiguousarray(indices, dtype=np.int32)
        call("HYPRE_IJVectorAddToValues", self.h, len(v), _ip(idx), _dp(v))

    def assemble(self):
        call("HYPRE_IJVectorAssemble", self.h)

    def get_values(self, indices=None) -> np.ndarray:
        n = self.n if indices is None else len(indices)
        out = np.empty(n, np.float64)
        idx = None if indices is None else np.ascontiguousarray(indices, dtype=np.int32)
        call("HYPRE_IJVectorGetValues", self.h, n, None if idx is None else _ip(idx), _dp(out))
        return out

    def get_object(self) -> "ParVector":
        obj = C.c_void_p()
        call("HYPRE_IJVectorGetObject", self.h, C.byref(obj))
        return _BorrowedVector(obj, self.n, self)   # owned by the IJ vector

    def destroy(self):
        if self.h:
            lib().HYPRE_IJVectorDestroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class _BorrowedVector(ParVector):
    def __init__(self, handle, n, owner):
        self.h = handle
        self.n = n
        self._owner = owner

    def destroy(self):
        self.h = None


# ------------------------------------------------------------------ solvers
class BoomerAMG:
    """HYPRE_BoomerAMG* with the hot-path settings as defaults."""

    def __init__(self, **params):
        h = C.c_void_p()
        call("HYPRE_BoomerAMGCreate", C.byref(h))
        self.h = h
        defaults = dict(coarsen_type=8, interp_type=6, strong_threshold=0.25, max_row_sum=0.9,
                        trunc_factor=0.0, P_max_elmts=4, relax_type=18, num_sweeps=1,
                        max_levels=25, max_coarse_size=9, relax_wt=1.0)
        defaults.update(params)
        self.set(**defaults)

    _SETTERS = {
        "coarsen_type": "HYPRE_BoomerAMGSetCoarsenType",
        "interp_type": "HYPRE_BoomerAMGSetInterpType",
        "strong_threshold": "HYPRE_BoomerAMGSetStrongThreshold",
        "max_row_sum": "HYPRE_BoomerAMGSetMaxRowSum",
        "trunc_factor": "HYPRE_BoomerAMGSetTruncFactor",
        "P_max_elmts": "HYPRE_BoomerAMGSetPMaxElmts",
        "relax_type": "HYPRE_BoomerAMGSetRelaxType",
        "num_sweeps": "HYPRE_BoomerAMGSetNumSweeps",
        "max_iter": "HYPRE_BoomerAMGSetMaxIter",
        "tol": "HYPRE_BoomerAMGSetTol",
        "max_levels": "HYPRE_BoomerAMGSetMaxLevels",
        "max_coarse_size": "HYPRE_BoomerAMGSetMaxCoarseSize",
        "relax_wt": "HYPRE_BoomerAMGSetRelaxWt",
        "outer_wt": "HYPRE_BoomerAMGSetOuterWt",
        "print_level": "HYPRE_BoomerAMGSetPrintLevel",
        "cycle_type": "HYPRE_BoomerAMGSetCycleType",
        "agg_num_levels": "HYPRE_BoomerAMGSetAggNumLevels",
        "agg_interp_type": "HYPRE_BoomerAMGSetAggInterpType",
        "agg_trunc_factor": "HYPRE_BoomerAMGSetAggTruncFactor",
        "agg_P_max_elmts": "HYPRE_BoomerAMGSetAggPMaxElmts",
        "num_paths": "HYPRE_BoomerAMGSetNumPaths",
        "fp32_cycle": "HYPRE_MI355X_BoomerAMGSetCyclePrecision",
        "keep_setup_data": "HYPRE_MI355X_BoomerAMGSetKeepSetupData",
        "solve_format": "HYPRE_MI355X_BoomerAMGSetSolveFormat",
    }

    def set(self, **params):
        for k, v in params.items():
            if k in ("relax_down", "relax_up", "relax_coarse"):
                which = {"relax_down": 1, "relax_up": 2, "relax_coarse": 3}[k]
                call("HYPRE_BoomerAMGSetCycleRelaxType", self.h, v, which)
            elif k == "hybrid_partitions":
                parts, rows = v if isinstance(v, tuple) else (v, 0)
                call("HYPRE_MI355X_BoomerAMGSetHybridPartitions", self.h, parts, rows)
            else:
                call(self._SETTERS[k], self.h, v)
        return self

    def set_level_params(self, theta=None, omega=None, sweeps=None):
        call("HYPRE_MI355X_BoomerAMGClearLevelParams", self.h)
        for lev, t in enumerate(theta or []):
            call("HYPRE_MI355X_BoomerAMGSetLevelStrongThreshold", self.h, lev, float(t))
        for lev, w in enumerate(omega or []):
            call("HYPRE_MI355X_BoomerAMGSetLevelRelaxWt", self.h, lev, float(w))
        for lev, s in enumerate(sweeps or []):
            call("HYPRE_MI355X_BoomerAMGSetLevelNumSweeps", self.h, lev, int(s))

    def setup(self, A: ParCSRMatrix, b: ParVector | None = None, x: ParVector | None = None):
        call("HYPRE_BoomerAMGSetup", self.h, A.h, b.h if b else None, x.h if x else None)

    def solve(self, A: ParCSRMatrix, b: ParVector, x: ParVector):
        return call("HYPRE_BoomerAMGSolve", self.h, A.h, b.h, x.h, allow=ERROR_CONV)

    def num_levels(self) -> int:
        v = C.c_int()
        call("HYPRE_MI355X_BoomerAMGGetNumLevels", self.h, C.byref(v))
        return v.value

    def num_iterations(self) -> int:
        v = C.c_int()
        call("HYPRE_BoomerAMGGetNumIterations", self.h, C.byref(v))
        return v.value

    def final_rel_res(self) -> float:
        v = C.c_double()
        call("HYPRE_BoomerAMGGetFinalRelativeResidualNorm", self.h, C.byref(v))
        return v.value

    def level_dims(self, lev: int, which: str):
        """(rows, cols, nnz) of a level matrix without copying it."""
        n, nc, nnz = C.c_int(), C.c_int(), C.c_int()
        call("HYPRE_MI355X_BoomerAMGGetLevelMatrix", self.h, lev, C.c_char(which.encode()),
             C.byref(n), C.byref(nc), C.byref(nnz), None, None, None)
        return n.value, nc.value, nnz.value

    def level_matrix(self, lev: int, which: str) -> HostCsr:
        n, nc, nnz = C.c_int(), C.c_int(), C.c_int()
        w = C.c_char(which.encode())
        call("HYPRE_MI355X_BoomerAMGGetLevelMatrix", self.h, lev, w, C.byref(n), C.byref(nc),
             C.byref(nnz), None, None, None)
        ia = np.empty(n.value + 1, np.int32)
        ja = np.empty(max(nnz.value, 1), np.int32)
        a = np.empty(max(nnz.value, 1), np.float64) if which != "S" else None
        call("HYPRE_MI355X_BoomerAMGGetLevelMatrix", self.h, lev, w, None, None, None, _ip(ia),
             _ip(ja), None if a is None else _dp(a))
        return HostCsr(n.value, nc.value, ia, ja[: nnz.value], None if a is None else a[: nnz.value])

    def level_rows(self, lev: int) -> int:
        n, nc, nnz = C.c_int(), C.c_int(), C.c_int()
        call("HYPRE_MI355X_BoomerAMGGetLevelMatrix", self.h, lev, C.c_char(b"A"), C.byref(n),
             C.byref(nc), C.byref(nnz), None, None, None)
        return n.value

    def level_cf(self, lev: int) -> np.ndarray:
        n = self.level_rows(lev)
        out = np.empty(n, np.int32)
        call("HYPRE_MI355X_BoomerAMGGetLevelCF", self.h, lev, _ip(out))
        return out

    def level_l1(self, lev: int) -> np.ndarray:
        n = self.level_rows(lev)
        out = np.empty(n, np.float64)
        call("HYPRE_MI355X_BoomerAMGGetLevelL1", self.h, lev, _dp(out))
        return out

    def level_measure(self, lev: int) -> np.ndarray:
        n = self.level_rows(lev)
        out = np.empty(n, np.float64)
        call("HYPRE_MI355X_BoomerAMGGetLevelMeasure", self.h, lev, _dp(out))
        return out

    def cycle_host(self, f: np.ndarray, u: np.ndarray | None = None) -> np.ndarray:
        f = np.ascontiguousarray(f, dtype=np.float64)
        u = np.zeros_like(f) if u is None else np.ascontiguousarray(u, dtype=np.float64).copy()
        call("HYPRE_MI355X_BoomerAMGCycleHost", self.h, _dp(f), _dp(u))
        return u

    def destroy(self):
        if self.h:
            lib().HYPRE_BoomerAMGDestroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class PCG:
    """HYPRE_ParCSRPCG* preconditioned by a BoomerAMG V-cycle."""

    def __init__(self, precond: BoomerAMG | str | None = None, tol=1e-8, max_iter=1000, two_norm=1,
                 print_level=0):
        h = C.c_void_p()
        call("HYPRE_ParCSRPCGCreate", 0, C.byref(h))
        self.h = h
        call("HYPRE_ParCSRPCGSetTol", h, tol)
        call("HYPRE_ParCSRPCGSetMaxIter", h, max_iter)
        call("HYPRE_ParCSRPCGSetTwoNorm", h, two_norm)
        call("HYPRE_ParCSRPCGSetPrintLevel", h, print_level)
        self.precond = precond
        if precond == "diag_scale":
            # HYPRE_ParCSRDiagScale with no solver object, as the reference's driver does
            call("HYPRE_ParCSRPCGSetPrecond", h, fnptr("HYPRE_ParCSRDiagScale"),
                 fnptr("HYPRE_ParCSRDiagScaleSetup"), None)
        elif precond is not None:
            precond.set(max_iter=1, tol=0.0)
            call("HYPRE_ParCSRPCGSetPrecond", h, fnptr("HYPRE_BoomerAMGSolve"),
                 fnptr("HYPRE_BoomerAMGSetup"), precond.h)

    def setup(self, A, b, x):
        call("HYPRE_ParCSRPCGSetup", self.h, A.h, b.h, x.h)

    def solve(self, A, b, x):
        return call("HYPRE_ParCSRPCGSolve", self.h, A.h, b.h, x.h, allow=ERROR_CONV)

    def num_iterations(self) -> int:
        v = C.c_int()
        call("HYPRE_ParCSRPCGGetNumIterations", self.h, C.byref(v))
        return v.value

    def final_rel_res(self) -> float:
        v = C.c_double()
        call("HYPRE_ParCSRPCGGetFinalRelativeResidualNorm", self.h, C.byref(v))
        return v.value

    def history(self) -> np.ndarray:
        n = C.c_int()
        buf = np.empty(10000, np.float64)
        call("HYPRE_MI355X_PCGGetResidualHistory", self.h, len(buf), _dp(buf), C.byref(n))
        return buf[: n.value].copy()

    def times(self):
        a, b = C.c_double(), C.c_double()
        call("HYPRE_MI355X_PCGGetTimes", self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def destroy(self):
        if self.h:
            lib().HYPRE_ParCSRPCGDestroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ------------------------------------------------------------------ single kernels
def create_s(A: HostCsr, theta=0.25, max_row_sum=0.9) -> HostCsr:
    n = A.nrows
    S_ia = np.empty(n + 1, np.int32)
    S_ja = np.empty(max(A.nnz, 1), np.int32)
    call("HYPRE_MI355X_CreateS", n, _ip(A.ia), _ip(A.ja), _dp(A.a), theta, max_row_sum, _ip(S_ia),
         _ip(S_ja))
    return HostCsr(n, n, S_ia, S_ja[: S_ia[n]].copy(), None)


def pmis(S: HostCsr, measure: np.ndarray | None = None):
    n = S.nrows
    cf = np.empty(n, np.int32)
    rounds = C.c_int()
    m = None if measure is None else np.ascontiguousarray(measure, dtype=np.float64)
    ja = S.ja if len(S.ja) else np.zeros(1, np.int32)
    call("HYPRE_MI355X_PMIS", n, _ip(S.ia), _ip(ja), None if m is None else _dp(m), _ip(cf),
         C.byref(rounds))
    return cf, rounds.value


def hmis(S: HostCsr, measure: np.ndarray | None = None):
    n = S.nrows
    cf = np.empty(n, np.int32)
    rounds = C.c_int()
    m = None if measure is None else np.ascontiguousarray(measure, dtype=np.float64)
    ja = S.ja if len(S.ja) else np.zeros(1, np.int32)
    call("HYPRE_MI355X_HMIS", n, _ip(S.ia), _ip(ja), None if m is None else _dp(m), _ip(cf),
         C.byref(rounds))
    return cf, rounds.value


def rs_first_pass(S: HostCsr) -> np.ndarray:
    n = S.nrows
    cf = np.empty(n, np.int32)
    ja = S.ja if len(S.ja) else np.zeros(1, np.int32)
    call("HYPRE_MI355X_RSFirstPass", n, _ip(S.ia), _ip(ja), _ip(cf))
    return cf


def extpi_interp(A: HostCsr, S: HostCsr, cf: np.ndarray, trunc_factor=0.0, max_elmts=4):
    n = A.nrows
    cf = np.ascontiguousarray(cf, dtype=np.int32).copy()
    ncols, nnz = C.c_int(), C.c_int()
    sja = S.ja if len(S.ja) else np.zeros(1, np.int32)
    call("HYPRE_MI355X_ExtPIInterp", n, _ip(A.ia), _ip(A.ja), _dp(A.a), _ip(S.ia), _ip(sja),
         _ip(cf), trunc_factor, max_elmts, C.byref(ncols), C.byref(nnz), None, None, None)
    ia = np.empty(n + 1, np.int32)
    ja = np.empty(max(nnz.value, 1), np.int32)
    a = np.empty(max(nnz.value, 1), np.float64)
    call("HYPRE_MI355X_ExtPIInterp", n, _ip(A.ia), _ip(A.ja), _dp(A.a), _ip(S.ia), _ip(sja),
         _ip(cf), trunc_factor, max_elmts, None, None, _ip(ia), _ip(ja), _dp(a))
    return HostCsr(n, ncols.value, ia, ja[: nnz.value], a[: nnz.value]), cf


def rap(A: HostCsr, P: HostCsr) -> HostCsr:
    n, nc = A.nrows, P.ncols
    nnz = C.c_int()
    call("HYPRE_MI355X_RAP", n, nc, _ip(A.ia), _ip(A.ja), _dp(A.a), _ip(P.ia), _ip(P.ja), _dp(P.a),
         C.byref(nnz), None, None, None)
    ia = np.empty(nc + 1, np.int32)
    ja = np.empty(max(nnz.value, 1), np.int32)
    a = np.empty(max(nnz.value, 1), np.float64)
    call("HYPRE_MI355X_RAP", n, nc, _ip(A.ia), _ip(A.ja), _dp(A.a), _ip(P.ia), _ip(P.ja), _dp(P.a),
         None, _ip(ia), _ip(ja), _dp(a))
    return HostCsr(nc, nc, ia, ja[: nnz.value], a[: nnz.value])


def spmv(A: HostCsr, x, alpha=1.0, beta=0.0, y=None) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(A.nrows) if y is None else np.ascontiguousarray(y, dtype=np.float64).copy()
    call("HYPRE_MI355X_SpMV", A.nrows, A.ncols, _ip(A.ia), _ip(A.ja), _dp(A.a), alpha, _dp(x), beta,
         _dp(y))
    return y


def hybrid_gs(A: HostCsr, f, u, relax_type: int, parts: int, l1=None, relax_weight=1.0,
              omega=1.0, zero_guess=False) -> np.ndarray:
    """One hybrid Gauss-Seidel relaxation on the GPU; returns the new u."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.ascontiguousarray(u, dtype=np.float64).copy()
    l1a = None if l1 is None else np.ascontiguousarray(l1, dtype=np.float64)
    l1p = None if l1a is None else _dp(l1a)
    call("HYPRE_MI355X_HybridGS", A.nrows, _ip(A.ia), _ip(A.ja), _dp(A.a), _dp(f), l1p, _dp(out),
         relax_type, relax_weight, omega, parts, 1 if zero_guess else 0)
    return out


# ------------------------------------------------------------------ IJ files
def ij_matrix_read(path: str, fmt: str = "ij") -> "IJMatrix":
    """HYPRE_IJMatrixRead / ReadMM / ReadBinary; fmt in {"ij", "mm", "bin"}."""
    fn = {"ij": "HYPRE_IJMatrixRead", "mm": "HYPRE_IJMatrixReadMM",
          "bin": "HYPRE_IJMatrixReadBinary"}[fmt]
    h = C.c_void_p()
    call(fn, path.encode(), 0, HYPRE_PARCSR, C.byref(h))
    return IJMatrix._wrap(h)


def ij_vector_read(path: str, binary: bool = False) -> "IJVector":
    h = C.c_void_p()
    call("HYPRE_IJVectorReadBinary" if binary else "HYPRE_IJVectorRead", path.encode(), 0,
         HYPRE_PARCSR, C.byref(h))
    return IJVector._wrap(h)


# ------------------------------------------------------------------ timing
def kernel_timing(on: bool = True, only: str | None = None):
    call("HYPRE_MI355X_SetKernelTimingClass", (only or "").encode())
    call("HYPRE_MI355X_SetKernelTiming", 1 if on else 0)


def reset_kernel_timing():
    call("HYPRE_MI355X_ResetKernelTiming")


def get_kernel_timing(name: str):
    ms, cnt, by = C.c_double(), C.c_int(), C.c_double()
    call("HYPRE_MI355X_GetKernelTiming", name.encode(), C.byref(ms), C.byref(cnt), C.byref(by))
    return ms.value, cnt.value, by.value
