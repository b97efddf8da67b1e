"""Pin the CPU restatement of the learned-parameter model (CPU only).

The model has no reference counterpart; its oracle is pinned by the
mathematics: central finite differences of the loss for every parameter
block, Adam's closed-form first step (each parameter moves by lr * sign(g)
up to eps), and the output-to-parameter mapping at its end points."""
import numpy as np

from oracle import param_model as PM

DIMS = (8, 16, 12, 9)


def data(seed=3, B=7):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(B, DIMS[0]))
    T = rng.uniform(size=(B, DIMS[3]))
    return X, T


def test_gradient_matches_central_differences():
    p = PM.init_params(DIMS, seed=5)
    # shift biases so that no pre-activation sits on the ReLU kink
    p = p + 0.01
    X, T = data()
    g, L = PM.gradient(p, DIMS, X, T)
    assert L == PM.loss(p, DIMS, X, T)
    h = 1e-6
    rng = np.random.default_rng(0)
    idx = rng.choice(len(p), size=60, replace=False)
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        fd = (PM.loss(p + e, DIMS, X, T) - PM.loss(p - e, DIMS, X, T)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (k, fd, g[k])


def test_adam_first_step_is_signed_lr():
    p = PM.init_params(DIMS, seed=2)
    X, T = data(4)
    g, _ = PM.gradient(p, DIMS, X, T)
    lr = 1e-2
    p1, hist = PM.adam_steps(p, DIMS, X, T, 1, lr)
    # bias-corrected moments after one step are g and g^2: step = lr g / (|g| + eps)
    np.testing.assert_allclose(p - p1, lr * g / (np.abs(g) + 1e-8), rtol=1e-9, atol=1e-15)
    mask = np.abs(g) > 1e-4
    np.testing.assert_allclose((p - p1)[mask], lr * np.sign(g[mask]), rtol=1e-3)


def test_training_reduces_loss():
    p = PM.init_params(DIMS, seed=7)
    X, T = data(9, 16)
    _, hist = PM.adam_steps(p, DIMS, X, T, 200, 1e-2)
    assert hist[-1] < 0.3 * hist[0]


def test_level_parameter_mapping():
    y = np.array([0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 0.1875, 0.5, 0.5])
    lp = PM.level_params(y)
    assert lp[0] == (0.1, 0.6, 1)
    assert lp[1] == (0.9, 1.4, 3)
    assert abs(lp[2][0] - 0.25) < 1e-15 and abs(lp[2][1] - 1.0) < 1e-15 and lp[2][2] == 2


def test_fallback_rule_keeps_defaults_unless_faster():
    """The guard of hypre_amd.learned: learned settings only when converged
    and faster per digit of residual reduction (seconds per digit)."""
    from hypre_amd.learned import SolveStats, choose
    d = SolveStats(iterations=20, rel_res=1e-9, seconds=0.9)          # 0.1 s per digit
    assert choose(SolveStats(15, 1e-9, 0.8), d) == "learned"
    assert choose(SolveStats(25, 1e-9, 1.0), d) == "default"
    assert choose(SolveStats(10, 2e-7, 0.3), d) == "default"           # did not converge
    # fewer seconds but fewer digits too: 0.85 s for 8 digits is slower per digit
    assert choose(SolveStats(12, 9.9e-9, 0.85), d) == "default"
    assert SolveStats(10, 1e-9, 0.9).seconds_per_digit == 0.1
