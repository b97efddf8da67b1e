"""Learned per-level BoomerAMG parameters (no counterpart in hypre).

A small perceptron on the device (8 matrix features -> 64 -> 64 -> 48
outputs, FP32, see hypre_amd/csrc/hx_mlp.hip) is a cost model: for each of
the 24 settings (strong threshold theta, l1-Jacobi weight omega, sweeps) of
GRID it predicts the setup+solve time relative to the defaults when the
setting is used on the fine level, and when it is used on the coarse
levels.  The solver takes the predicted fastest setting of each group
(level 0, levels 1-24).  The features (HYPRE_MI355X_MatrixFeatures) are the
row-length distribution, off-diagonal anisotropy ratios, the strong fraction
and a coefficient-jump histogram.

Training data come from sweeps of the plain method itself: for each training
matrix, AMG-PCG is timed (setup + solve to 1e-8, best of two runs) for every
setting of GRID on the fine level with the coarse levels at the defaults,
then for every setting on the coarse levels with the fine level at its best;
the targets are those times over the defaults' (the measured convergence per
second of each setting).  Learning the whole table, rather than only its
fastest entry, averages the timing noise over similar matrices: where the
features carry no information the model reduces to the best constant rule,
and where they do it can beat it.  The best constant rule of the training
sweep is the reading's prior: a group leaves it only for a setting predicted
faster by more than `prior_margin` (0.03, chosen by 8-fold cross-validation
on the training sweep, where the plain argmin lost to the constant rule out
of sample: the per-problem best setting is only 0.6% faster than the
constant rule on these families, less than the timing noise the targets
carry).  The guard `apply_with_fallback` keeps the defaults whenever the
predicted settings reach the tolerance more slowly (less convergence per
second) on the matrix at hand.
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
WEIGHTS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "learned_params.json")
DEFAULT = (0.25, 1.0, 1)

# settings of one level group (the fine level, or all coarse levels)
GRID = tuple((th, om, sw) for th in (0.25, 0.4, 0.55) for om in (1.0, 1.2, 1.4, 1.6)
             for sw in (1, 2))
NOUT = 2 * len(GRID)          # cost model: fine-level group, then coarse-level group
NOUT_LEVELS = 3 * LEVELS      # per-level reading (theta, omega, sweeps of each level)
# relative times r map to model targets (r - 0.5), clipped to (0, 1); a
# setting that did not converge is the top of the range
RATIO_LO, RATIO_HI = 0.5, 1.5


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class ParamModel:
    """HYPRE_MI355X_ParamModel* through the C ABI."""

    def __init__(self, nin=NIN, nh1=NH1, nh2=NH2, nout=NOUT_LEVELS, params=None, seed=1,
                 grid=None):
        h = C.c_void_p()
        call("HYPRE_MI355X_ParamModelCreate", nin, nh1, nh2, nout, C.byref(h))
        self.h = h
        self.dims = (nin, nh1, nh2, nout)
        npar = C.c_int()
        call("HYPRE_MI355X_ParamModelNumParams", h, C.byref(npar))
        self.np = npar.value
        self.set_params(init_params(self.dims, seed) if params is None else params)
        self.grid = None
        self.prior = None
        if grid is not None:
            self.set_grid(grid)

    def set_grid(self, grid):
        """Cost-model reading over `grid` (nout = 2 len(grid))."""
        grid = [tuple(g) for g in grid]
        th = np.array([g[0] for g in grid], np.float64)
        om = np.array([g[1] for g in grid], np.float64)
        sw = np.array([g[2] for g in grid], np.int32)
        call("HYPRE_MI355X_ParamModelSetGrid", self.h, len(grid),
             th.ctypes.data_as(C.POINTER(C.c_double)), om.ctypes.data_as(C.POINTER(C.c_double)),
             sw.ctypes.data_as(C.POINTER(C.c_int)))
        self.grid = grid
        self.prior = None

    def set_prior(self, fine, coarse, margin):
        """Keep `fine` on level 0 and `coarse` on the coarse levels unless
        another grid setting is predicted faster by more than `margin`."""
        call("HYPRE_MI355X_ParamModelSetPrior", self.h, self.grid.index(tuple(fine)),
             self.grid.index(tuple(coarse)), float(margin))
        self.prior = ((tuple(fine), tuple(coarse)), float(margin))

    def predict(self, A: hx.ParCSRMatrix):
        """(fine setting, coarse setting) the cost model picks for A."""
        return self.predict_from_features(matrix_features(A))

    def predict_from_features(self, f):
        y = self.forward(np.asarray(f, np.float32)[None, :])[0]
        return decode_costs(y, self.grid, self.prior)

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
                       "grid": [list(g) for g in self.grid] if self.grid else None,
                       "meta": meta or {}}, f)

    @classmethod
    def load(cls, path=WEIGHTS_FILE) -> "ParamModel":
        with open(path) as f:
            d = json.load(f)
        m = cls(*d["dims"], params=np.asarray(d["params"], np.float32), grid=d.get("grid"))
        meta = d.get("meta", {})
        if d.get("grid") and meta.get("prior_margin") is not None and meta.get("constant_rule"):
            fine, coarse = meta["constant_rule"]
            m.set_prior(fine, coarse, meta["prior_margin"])
        return m

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


def per_level(fine, coarse, levels=LEVELS):
    """Per-level settings: `fine` on level 0, `coarse` on levels 1.."""
    return [tuple(fine)] + [tuple(coarse)] * (levels - 1)


def encode_costs(fine_ratios, coarse_ratios) -> np.ndarray:
    """Cost-model targets from the relative times of the GRID settings on the
    fine level and on the coarse levels (None: did not converge)."""
    r = [RATIO_HI if v is None else v for v in list(fine_ratios) + list(coarse_ratios)]
    return np.clip(np.asarray(r, np.float64) - RATIO_LO, 1e-3, 1 - 1e-3).astype(np.float32)


def decode_costs(y, grid=GRID, prior=None):
    """The first predicted minimum of each group: (fine setting, coarse
    setting).  With prior = ((fine, coarse), margin), a group keeps its prior
    setting unless the minimum is predicted faster than it by more than the
    margin (HYPRE_MI355X_ParamModelSetPrior)."""
    G = len(grid)
    y = np.asarray(y, np.float64)
    gf, gc = int(np.argmin(y[:G])), int(np.argmin(y[G:2 * G]))
    if prior is not None:
        (pf, pc), margin = prior
        fi, ci = list(grid).index(tuple(pf)), list(grid).index(tuple(pc))
        if y[gf] > y[fi] - margin:
            gf = fi
        if y[G + gc] > y[G + ci] - margin:
            gc = ci
    return tuple(grid[gf]), tuple(grid[gc])


def encode(level_params) -> np.ndarray:
    """Targets in sigmoid space for per-level (theta, omega, sweeps)."""
    t = np.empty(3 * len(level_params), np.float32)
    for l, (th, om, sw) in enumerate(level_params):
        t[3 * l] = (th - 0.1) / 0.8
        t[3 * l + 1] = (om - 0.6) / 0.8
        t[3 * l + 2] = (sw - 1) / 2.0
    return t


def decode(y, levels=LEVELS):
    return [(0.1 + 0.8 * float(y[3 * l]), 0.6 + 0.8 * float(y[3 * l + 1]),
             1 + int(math.floor(2.0 * float(y[3 * l + 2]) + 0.5))) for l in range(levels)]


@dataclass
class SolveStats:
    iterations: int
    rel_res: float
    seconds: float          # setup + solve, best of the timed runs

    @property
    def converged(self) -> bool:
        return self.rel_res < 1e-8

    @property
    def seconds_per_digit(self) -> float:
        """Inverse convergence speed: seconds per decimal digit of residual
        reduction (the measured convergence factor per second)."""
        digits = -math.log10(max(self.rel_res, 1e-300))
        return self.seconds / max(digits, 1e-3)


def timed_solve(A, b, level_params=None, model=None, fp32=False, tol=1e-8, max_iter=200,
                repeats=2) -> SolveStats:
    """AMG-PCG from x0 = 0 with the defaults, explicit per-level settings or a
    model's prediction; HYPRE_ParCSRPCGSetup + Solve timed by the library,
    best of `repeats` runs."""
    amg = hx.BoomerAMG(fp32_cycle=1 if fp32 else 0)
    if model is not None:
        model.apply(amg, A)
    elif level_params is not None:
        th, om, sw = zip(*level_params)
        amg.set_level_params(theta=th, omega=om, sweeps=sw)
    pcg = hx.PCG(amg, tol=tol, max_iter=max_iter)
    x = hx.ParVector(b.n)
    best = None
    for _ in range(repeats):
        x.fill(0.0)
        pcg.setup(A, b, x)
        pcg.solve(A, b, x)
        t = sum(pcg.times())
        best = t if best is None else min(best, t)
    st = SolveStats(pcg.num_iterations(), pcg.final_rel_res(), best)
    pcg.destroy()
    amg.destroy()
    x.destroy()
    return st


def best_per_level(A, b, repeats=2):
    """Coordinate search over GRID: the fine level with coarse defaults, then
    the coarse levels with the best fine setting.  Returns (fine, coarse),
    their stats, the default stats and every timing."""
    base = timed_solve(A, b, repeats=repeats)
    fine, fine_st = DEFAULT, base
    table = {"fine": {}, "coarse": {}}
    for g in GRID:
        st = base if g == DEFAULT else timed_solve(A, b, per_level(g, DEFAULT), repeats=repeats)
        table["fine"][str(g)] = round(st.seconds, 6) if st.converged else None
        if st.converged and st.seconds < fine_st.seconds:
            fine, fine_st = g, st
    coarse, best_st = DEFAULT, fine_st
    for g in GRID:
        st = fine_st if g == DEFAULT else timed_solve(A, b, per_level(fine, g), repeats=repeats)
        table["coarse"][str(g)] = round(st.seconds, 6) if st.converged else None
        if st.converged and st.seconds < best_st.seconds:
            coarse, best_st = g, st
    return (fine, coarse), best_st, base, table


def aniso_problem(N, seed):
    """Anisotropic 7-point diffusion (the family of the 64-solve batch):
    eps = 10^(-3U), rotation pi U about z, from a seeded generator."""
    rng = np.random.default_rng(seed)
    eps = 10.0 ** (-3.0 * rng.uniform())
    theta = math.pi * rng.uniform()
    A = hx.ParCSRMatrix.aniso7(N, eps, theta)
    b = hx.ParVector(N ** 3)
    b.set_random(seed + 1)
    return A, b, {"family": "aniso7", "eps": eps, "rotation": theta}


def jump_problem(N, seed):
    """Jumping-coefficient diffusion: 7- or 27-point, kappa piecewise constant
    on cells of 4, 8 or 16 nodes, log-uniform over 2 to 8 decades."""
    rng = np.random.default_rng(10_000 + seed)
    stencil = 27 if rng.uniform() < 0.5 else 7
    cell = int(rng.choice([4, 8, 16]))
    dec = float(rng.choice([2.0, 4.0, 6.0, 8.0]))
    A = hx.ParCSRMatrix.var_diffusion(N, stencil, cell, -dec / 2, dec / 2, seed)
    b = hx.ParVector(N ** 3)
    b.set_random(seed + 1)
    return A, b, {"family": f"jump{stencil}", "cell": cell, "decades": dec}


def choose(learned: SolveStats, default: SolveStats) -> str:
    """"learned" only if the predicted settings converged and reduce the
    residual faster (fewer seconds per digit) than the defaults."""
    if learned.converged and learned.seconds_per_digit < default.seconds_per_digit:
        return "learned"
    return "default"


def apply_with_fallback(amg: hx.BoomerAMG, model: ParamModel, A, b, fp32=False, repeats=2):
    """Predict per-level parameters for A; keep them only when AMG-PCG with
    them converges faster than the defaults (best of `repeats` timed solves
    each: the first solve of a new hierarchy also pays for growing the
    allocator's cache, which would bias a single run against whichever
    candidate goes first).  Returns ("learned" | "default", learned stats,
    default stats, seconds spent deciding)."""
    import time
    t0 = time.perf_counter()
    learned = timed_solve(A, b, model=model, fp32=fp32, repeats=repeats)
    default = timed_solve(A, b, fp32=fp32, repeats=repeats)
    which = choose(learned, default)
    if which == "learned":
        model.apply(amg, A)
    else:
        call("HYPRE_MI355X_BoomerAMGClearLevelParams", amg.h)
    return which, learned, default, time.perf_counter() - t0


def constant_rules(rows, k=4):
    """Candidate constant rules (one (fine, coarse) pair for every matrix):
    the k most frequent training targets."""
    from collections import Counter
    cnt = Counter((tuple(r["fine"]), tuple(r["coarse"])) for r in rows)
    return [list(p) for p, _ in cnt.most_common(k)]


def train_main(argv=None):
    """python -m hypre_amd.learned [--aniso-n 128] [--aniso-count 64]
                                   [--jump-n 96] [--jump-count 32] [--steps 4000]

    Sweeps GRID on the training problems (anisotropic seeds 1..aniso-count,
    jumping-coefficient seeds 1..jump-count), trains the device cost model on
    the relative times of every setting, picks the best constant rule on the
    same problems, and writes WEIGHTS_FILE."""
    import argparse
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("--aniso-n", type=int, default=128)
    ap.add_argument("--aniso-count", type=int, default=64)
    ap.add_argument("--jump-n", type=int, default=96)
    ap.add_argument("--jump-count", type=int, default=32)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--prior-margin", type=float, default=0.03,
                    help="predicted gain (relative time) needed to leave the constant rule")
    ap.add_argument("--out", default=WEIGHTS_FILE)
    args = ap.parse_args(argv)
    hx.initialize()
    X, T, rows = [], [], []
    problems = ([("aniso", args.aniso_n, s) for s in range(1, args.aniso_count + 1)] +
                [("jump", args.jump_n, s) for s in range(1, args.jump_count + 1)])
    t0 = time.perf_counter()
    for kind, N, s in problems:
        A, b, info = (aniso_problem if kind == "aniso" else jump_problem)(N, s)
        f = matrix_features(A)
        (fine, coarse), bst, base, table = best_per_level(A, b)
        X.append(f.astype(np.float32))
        fr = [None if table["fine"][str(g)] is None else table["fine"][str(g)] / base.seconds
              for g in GRID]
        cr = [None if table["coarse"][str(g)] is None else table["coarse"][str(g)] / base.seconds
              for g in GRID]
        T.append(encode_costs(fr, cr))
        rows.append({"kind": kind, "n": N, "seed": s, **info, "features": [round(v, 6) for v in f],
                     "fine": fine, "coarse": coarse, "seconds_best": round(bst.seconds, 6),
                     "seconds_default": round(base.seconds, 6), "its_best": bst.iterations,
                     "its_default": base.iterations, "timings": table})
        print(json.dumps({k: v for k, v in rows[-1].items() if k != "timings"}), flush=True)
        A.destroy()
        b.destroy()
    X = np.stack(X)
    T = np.stack(T)
    model = ParamModel(nout=NOUT, seed=1, grid=GRID)
    hist = model.train(X, T, args.steps, args.lr)
    # the best constant rule on the training problems (mean time over default)
    cands = constant_rules(rows)
    scores = []
    for fine, coarse in cands:
        r = []
        for (kind, N, s), row in zip(problems, rows):
            A, b, _ = (aniso_problem if kind == "aniso" else jump_problem)(N, s)
            st = timed_solve(A, b, per_level(fine, coarse))
            r.append(st.seconds / row["seconds_default"] if st.converged else 10.0)
            A.destroy()
            b.destroy()
        scores.append(float(np.mean(r)))
        print(json.dumps({"constant_rule": [fine, coarse], "mean_time_over_default": scores[-1]}),
              flush=True)
    best_rule = cands[int(np.argmin(scores))]
    # the rule is the reading's prior: the network moves a group off it only
    # when it predicts a gain above the sweep's timing noise
    model.set_prior(best_rule[0], best_rule[1], args.prior_margin)
    picks = [model.predict_from_features(x) for x in X]
    meta = {"families": {"aniso7": f"{args.aniso_n}^3, eps=10^(-3U), rotation pi*U, seeds 1..{args.aniso_count}",
                         "jump": f"{args.jump_n}^3, 7/27-point, cells 4/8/16, 2-8 decades, seeds 1..{args.jump_count}"},
            "grid": [list(g) for g in GRID], "steps": args.steps, "lr": args.lr,
            "loss_first": float(hist[0]), "loss_last": float(hist[-1]),
            "sweep_seconds": round(time.perf_counter() - t0, 1),
            "mean_time_best_over_default":
                float(np.mean([r["seconds_best"] / r["seconds_default"] for r in rows])),
            "reading": "cost model: relative time of each GRID setting on the fine level "
                       "and on the coarse levels; the predicted fastest setting of each is used "
                       "when it beats the constant rule's setting by more than prior_margin",
            "prior_margin": args.prior_margin,
            "training_picks": [[list(a), list(b)] for a, b in picks],
            "constant_rule": best_rule, "constant_rule_candidates":
                [{"rule": c, "mean_time_over_default": sc} for c, sc in zip(cands, scores)],
            "rows": rows}
    model.save(args.out, meta)
    print(json.dumps({k: v for k, v in meta.items() if k != "rows"}), flush=True)


if __name__ == "__main__":
    train_main()
