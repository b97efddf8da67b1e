"""Learned per-level BoomerAMG parameters (no counterpart in hypre).

A small perceptron on the device (8 matrix features -> 64 -> 64 -> 3 x 25
outputs, FP32, see hypre_amd/csrc/hx_mlp.hip) predicts, for every level, the
strong threshold theta, the l1-Jacobi weight omega and the number of sweeps.
It is trained on sweeps of plain settings: for each training matrix every
(theta, omega, sweeps) of a small grid is run through AMG-PCG and the setting
with the least work to 1e-8 becomes the target.

Work of one solve (deterministic, in fine-matrix nonzero passes):
    cycle = sum_l [ (2 s_l + 1) nnz(A_l) + 2 nnz(P_l) ] / nnz(A_0)
    work  = iterations * (1 + cycle) + setup_weight * operator_complexity
The guard `apply_with_fallback` keeps the defaults whenever the predicted
settings need more work per digit of residual reduction on the matrix at hand.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
from dataclasses import dataclass

import numpy as np

from . import api as hx
from .api import call
from .lib import lib

NIN, NH1, NH2, LEVELS = 8, 64, 64, 25
NOUT = 3 * LEVELS
SETUP_WEIGHT = 4.0
WEIGHTS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "learned_params.json")
DEFAULT = (0.25, 1.0, 1)

THETAS = (0.25, 0.4, 0.55, 0.7)
OMEGAS = (0.8, 1.0, 1.2)
SWEEPS = (1, 2)


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class ParamModel:
    """HYPRE_MI355X_ParamModel* through the C ABI."""

    def __init__(self, nin=NIN, nh1=NH1, nh2=NH2, nout=NOUT, params=None, seed=1):
        h = C.c_void_p()
        call("HYPRE_MI355X_ParamModelCreate", nin, nh1, nh2, nout, C.byref(h))
        self.h = h
        self.dims = (nin, nh1, nh2, nout)
        npar = C.c_int()
        call("HYPRE_MI355X_ParamModelNumParams", h, C.byref(npar))
        self.np = npar.value
        self.set_params(init_params(self.dims, seed) if params is None else params)

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (self.np,)
        call("HYPRE_MI355X_ParamModelSetParams", self.h, _fp(p))

    def get_params(self) -> np.ndarray:
        p = np.empty(self.np, np.float32)
        call("HYPRE_MI355X_ParamModelGetParams", self.h, _fp(p))
        return p

    def forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.empty((X.shape[0], self.dims[3]), np.float32)
        call("HYPRE_MI355X_ParamModelForward", self.h, X.shape[0], _fp(X), _fp(Y))
        return Y

    def gradient(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.empty(self.np, np.float32)
        loss = np.empty(1, np.float32)
        call("HYPRE_MI355X_ParamModelGradient", self.h, X.shape[0], _fp(X), _fp(T), _fp(g), _fp(loss))
        return g, float(loss[0])

    def train(self, X, T, steps, lr=1e-2) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        hist = np.empty(max(steps, 1), np.float32)
        call("HYPRE_MI355X_ParamModelTrain", self.h, X.shape[0], _fp(X), _fp(T), steps,
             C.c_float(lr), _fp(hist))
        return hist[:steps]

    def apply(self, amg: hx.BoomerAMG, A: hx.ParCSRMatrix):
        call("HYPRE_MI355X_BoomerAMGSetParamModel", amg.h, self.h, A.h)

    def save(self, path=WEIGHTS_FILE, meta=None):
        with open(path, "w") as f:
            json.dump({"dims": list(self.dims), "params": [float(v) for v in self.get_params()],
                       "meta": meta or {}}, f)

    @classmethod
    def load(cls, path=WEIGHTS_FILE) -> "ParamModel":
        with open(path) as f:
            d = json.load(f)
        return cls(*d["dims"], params=np.asarray(d["params"], np.float32))

    def destroy(self):
        if self.h:
            lib().HYPRE_MI355X_ParamModelDestroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def init_params(dims, seed=1) -> np.ndarray:
    """Glorot-uniform weights from a seeded generator, zero biases."""
    rng = np.random.default_rng(seed)
    out = []
    nin, nh1, nh2, nout = dims
    for a, b in ((nin, nh1), (nh1, nh2), (nh2, nout)):
        lim = math.sqrt(6.0 / (a + b))
        out += [rng.uniform(-lim, lim, size=a * b), np.zeros(b)]
    return np.concatenate(out).astype(np.float32)


def matrix_features(A: hx.ParCSRMatrix) -> np.ndarray:
    f = np.empty(8, np.float64)
    call("HYPRE_MI355X_MatrixFeatures", A.h, f.ctypes.data_as(C.POINTER(C.c_double)))
    return f


def encode(theta, omega, sweeps, levels=LEVELS) -> np.ndarray:
    """Targets in sigmoid space for one setting used on every level."""
    t = np.empty(3 * levels, np.float32)
    t[0::3] = (theta - 0.1) / 0.8
    t[1::3] = (omega - 0.6) / 0.8
    t[2::3] = (sweeps - 1) / 2.0
    return t


def decode(y, levels=LEVELS):
    return [(0.1 + 0.8 * float(y[3 * l]), 0.6 + 0.8 * float(y[3 * l + 1]),
             1 + int(math.floor(2.0 * float(y[3 * l + 2]) + 0.5))) for l in range(levels)]


@dataclass
class SolveStats:
    iterations: int
    rel_res: float
    cycle: float
    op_complexity: float

    @property
    def work(self) -> float:
        return self.iterations * (1.0 + self.cycle) + SETUP_WEIGHT * self.op_complexity

    @property
    def work_per_digit(self) -> float:
        digits = -math.log10(max(self.rel_res, 1e-300))
        return self.work / max(digits, 1e-3)


def solve_stats(A, b, level_params=None, model=None, tol=1e-8, max_iter=200) -> SolveStats:
    """AMG-PCG with uniform defaults, explicit per-level (theta, omega,
    sweeps), or a model's prediction; deterministic work count."""
    amg = hx.BoomerAMG()
    if model is not None:
        model.apply(amg, A)
    elif level_params is not None:
        th, om, sw = zip(*level_params)
        amg.set_level_params(theta=th, omega=om, sweeps=sw)
    pcg = hx.PCG(amg, tol=tol, max_iter=max_iter)
    x = hx.ParVector(b.n)
    x.fill(0.0)
    pcg.setup(A, b, x)
    pcg.solve(A, b, x)
    its, rr = pcg.num_iterations(), pcg.final_rel_res()
    nlev = amg.num_levels()
    nnzA = [amg.level_dims(l, "A")[2] for l in range(nlev)]
    nnzP = [amg.level_dims(l, "P")[2] for l in range(nlev - 1)]
    if model is not None:
        lp = decode(model.forward(matrix_features(A)[None, :])[0])
    elif level_params is not None:
        lp = list(level_params) + [DEFAULT] * LEVELS
    else:
        lp = [DEFAULT] * LEVELS
    cyc = sum((2 * lp[l][2] + 1) * nnzA[l] + (2 * nnzP[l] if l < nlev - 1 else 0)
              for l in range(nlev)) / nnzA[0]
    pcg.destroy()
    amg.destroy()
    return SolveStats(its, rr, cyc, sum(nnzA) / nnzA[0])


def best_setting(A, b):
    """Sweep the grid of uniform settings; return (best setting, its stats, default stats)."""
    base = solve_stats(A, b)
    best, best_stats = DEFAULT, base
    for th in THETAS:
        for om in OMEGAS:
            for sw in SWEEPS:
                if (th, om, sw) == DEFAULT:
                    continue
                st = solve_stats(A, b, level_params=[(th, om, sw)] * LEVELS)
                if st.rel_res < 1e-8 and st.work < best_stats.work:
                    best, best_stats = (th, om, sw), st
    return best, best_stats, base


def aniso_problem(N, seed):
    """Anisotropic 7-point diffusion of the training family: eps = 10^(-3U),
    rotation pi U about z, from a seeded generator."""
    rng = np.random.default_rng(seed)
    eps = 10.0 ** (-3.0 * rng.uniform())
    theta = math.pi * rng.uniform()
    A = hx.ParCSRMatrix.aniso7(N, eps, theta)
    b = hx.ParVector(N ** 3)
    b.set_random(seed + 1)
    return A, b, eps, theta


def choose(learned: SolveStats, default: SolveStats, tol: float = 1e-8) -> str:
    """"learned" only if the predicted settings converged and need less work
    per digit of residual reduction than the defaults."""
    if learned.rel_res < tol and learned.work_per_digit < default.work_per_digit:
        return "learned"
    return "default"


def apply_with_fallback(amg: hx.BoomerAMG, model: ParamModel, A, b):
    """Predict per-level parameters for A; keep them only when AMG-PCG with
    them needs less work per digit than the defaults.  Returns
    ("learned" | "default", learned stats, default stats)."""
    learned = solve_stats(A, b, model=model)
    default = solve_stats(A, b)
    which = choose(learned, default)
    if which == "learned":
        model.apply(amg, A)
    else:
        call("HYPRE_MI355X_BoomerAMGClearLevelParams", amg.h)
    return which, learned, default


def train_main(argv=None):
    """python -m hypre_amd.learned [--n 64] [--count 64] [--steps 3000]

    Sweeps the settings grid on `count` anisotropic problems (seeds 1..count),
    trains the device model on the best settings and writes WEIGHTS_FILE."""
    import argparse
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--out", default=WEIGHTS_FILE)
    args = ap.parse_args(argv)
    hx.initialize()
    X, T, rows = [], [], []
    t0 = time.perf_counter()
    for s in range(1, args.count + 1):
        A, b, eps, th = aniso_problem(args.n, s)
        f = matrix_features(A)
        best, bst, base = best_setting(A, b)
        X.append(f.astype(np.float32))
        T.append(encode(*best))
        rows.append({"seed": s, "eps": eps, "rotation": th, "best": best,
                     "work_best": round(bst.work, 2), "work_default": round(base.work, 2),
                     "its_best": bst.iterations, "its_default": base.iterations})
        print(json.dumps(rows[-1]), flush=True)
        A.destroy()
        b.destroy()
    X = np.stack(X)
    T = np.stack(T)
    model = ParamModel(seed=1)
    hist = model.train(X, T, args.steps, args.lr)
    meta = {"family": f"aniso7 {args.n}^3, eps=10^(-3U), rotation pi*U, seeds 1..{args.count}",
            "grid": {"theta": THETAS, "omega": OMEGAS, "sweeps": SWEEPS},
            "steps": args.steps, "lr": args.lr, "loss_first": float(hist[0]),
            "loss_last": float(hist[-1]), "sweep_seconds": round(time.perf_counter() - t0, 1),
            "mean_work_ratio_best_over_default":
                float(np.mean([r["work_best"] / r["work_default"] for r in rows])),
            "rows": rows}
    model.save(args.out, meta)
    print(json.dumps({k: v for k, v in meta.items() if k != "rows"}), flush=True)


if __name__ == "__main__":
    train_main()
