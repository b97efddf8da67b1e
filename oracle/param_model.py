"""CPU oracle of the learned per-level parameter model (test infrastructure).

Only tests/ may import this module.  It restates, in float64 numpy, what
hypre_amd/csrc/hx_mlp.hip computes in FP32 on the device:

    Z1 = X W1 + b1, H1 = relu(Z1); Z2 = H1 W2 + b2, H2 = relu(Z2);
    Z3 = H2 W3 + b3, Y = sigmoid(Z3); loss = mean((Y - T)^2)

with the parameters packed as W1[nin][nh1], b1, W2[nh1][nh2], b2,
W3[nh2][nout], b3, and Adam (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected
moments).  The model has no counterpart in hypre, so there is no reference
file to cite; the restatement is pinned by central finite differences of the
loss and by Adam's closed-form first step (tests/test_param_model_oracle.py).
"""
from __future__ import annotations

import numpy as np


def shapes(nin, nh1, nh2, nout):
    return [("W1", (nin, nh1)), ("b1", (nh1,)), ("W2", (nh1, nh2)), ("b2", (nh2,)),
            ("W3", (nh2, nout)), ("b3", (nout,))]


def num_params(nin, nh1, nh2, nout) -> int:
    return sum(int(np.prod(s)) for _, s in shapes(nin, nh1, nh2, nout))


def unpack(p, dims):
    out, o = {}, 0
    for name, s in shapes(*dims):
        k = int(np.prod(s))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(s)
        o += k
    return out


def pack(parts, dims):
    return np.concatenate([np.asarray(parts[name], np.float64).ravel() for name, _ in shapes(*dims)])


def init_params(dims, seed=1):
    """Glorot-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = {}
    for name, s in shapes(*dims):
        if name.startswith("W"):
            lim = np.sqrt(6.0 / (s[0] + s[1]))
            parts[name] = rng.uniform(-lim, lim, size=s)
        else:
            parts[name] = np.zeros(s)
    return pack(parts, dims)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(p, dims, X):
    w = unpack(p, dims)
    Z1 = X @ w["W1"] + w["b1"]
    H1 = np.maximum(Z1, 0.0)
    Z2 = H1 @ w["W2"] + w["b2"]
    H2 = np.maximum(Z2, 0.0)
    Z3 = H2 @ w["W3"] + w["b3"]
    return sigmoid(Z3), (Z1, H1, Z2, H2, Z3)


def loss(p, dims, X, T) -> float:
    Y, _ = forward(p, dims, X)
    return float(np.mean((Y - T) ** 2))


def gradient(p, dims, X, T):
    """(grad, loss) of mean((Y - T)^2) by the chain rule."""
    w = unpack(p, dims)
    Y, (Z1, H1, Z2, H2, Z3) = forward(p, dims, X)
    n = Y.size
    D3 = 2.0 * (Y - T) / n * Y * (1.0 - Y)
    g = {"W3": H2.T @ D3, "b3": D3.sum(axis=0)}
    D2 = (D3 @ w["W3"].T) * (Z2 > 0)
    g["W2"], g["b2"] = H1.T @ D2, D2.sum(axis=0)
    D1 = (D2 @ w["W2"].T) * (Z1 > 0)
    g["W1"], g["b1"] = X.T @ D1, D1.sum(axis=0)
    return pack(g, dims), float(np.mean((Y - T) ** 2))


def adam_steps(p, dims, X, T, steps, lr, b1=0.9, b2=0.999, eps=1e-8):
    p = np.array(p, dtype=np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    hist = []
    for t in range(1, steps + 1):
        g, L = gradient(p, dims, X, T)
        hist.append(L)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, np.array(hist)


def level_params(y, levels=None):
    """Map sigmoid outputs of one sample to per-level (theta, omega, sweeps)."""
    L = len(y) // 3 if levels is None else levels
    out = []
    for l in range(L):
        out.append((0.1 + 0.8 * float(y[3 * l]), 0.6 + 0.8 * float(y[3 * l + 1]),
                    1 + int(np.floor(2.0 * float(y[3 * l + 2]) + 0.5))))
    return out


def matrix_features(A) -> np.ndarray:
    """The 8 features of hypre_amd's HYPRE_MI355X_MatrixFeatures on an oracle Csr
    (diagonal first in every row): log10(n)/8; mean and standard deviation of
    the row lengths, over 27; mean min/max and mean 2nd/1st largest |a_ij| over
    the off-diagonal entries; mean fraction of strong off-diagonal entries at
    theta 0.25 (the rule of src/parcsr_ls/par_strength.c:75 without the row-sum
    test); fractions of rows with a neighbour j whose |a_jj| differs from
    |a_ii| by a factor of at least 10, and of at least 1000 (zero diagonals
    count as 1)."""
    n = A.nrows
    q = np.zeros(7)
    ia, ja, a = np.asarray(A.ia), np.asarray(A.ja), np.asarray(A.a, dtype=np.float64)
    diag = np.abs(a[ia[:-1]])
    diag = np.where(diag > 0, diag, 1.0)
    for i in range(n):
        s, e = int(ia[i]), int(ia[i + 1])
        d = float(a[s])
        ad = diag[i]
        off = a[s + 1:e]
        av = np.abs(off)
        no = e - s - 1
        mx = av.max() if no else 0.0
        srt = np.sort(av)[::-1]
        mx2 = srt[1] if no > 1 else 0.0
        mn = av.min() if no else 0.0
        scale = 0.0
        for v in off:
            scale = (max(scale, v) if d < 0 else min(scale, v))
        thr = 0.25 * scale
        strong = int(np.sum(off > thr)) if d < 0 else int(np.sum(off < thr))
        dj = diag[ja[s + 1:e]]
        j1 = bool(np.any((dj >= 10.0 * ad) | (ad >= 10.0 * dj)))
        j3 = bool(np.any((dj >= 1000.0 * ad) | (ad >= 1000.0 * dj)))
        L = float(e - s)
        q += [L, L * L, (mn / mx) if (no and mx > 0) else 0.0,
              (mx2 / mx) if (no > 1 and mx > 0) else 0.0, (strong / no) if no else 0.0,
              1.0 if j1 else 0.0, 1.0 if j3 else 0.0]
    q /= n
    var = max(0.0, q[1] - q[0] ** 2)
    return np.array([np.log10(n) / 8.0, q[0] / 27.0, np.sqrt(var) / 27.0, q[2], q[3], q[4],
                     q[5], q[6]])
