"""The learned per-level parameter model on the device (FP32 GEMMs on the
matrix cores, ReLU MLP, Adam) against its float64 CPU oracle, and its effect
on AMG-PCG.  Tolerances are FP32 ones: relative 1e-4 for forward outputs,
2e-3 of the largest gradient entry, 1e-3 on the loss history."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS = (8, 64, 64, 75)


def say(*a):
    print("[gpu-param-model]", *a, flush=True)
    sys.stdout.flush()


def batch(seed, B=64):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(B, DIMS[0])).astype(np.float32)
    T = rng.uniform(size=(B, DIMS[3])).astype(np.float32)
    return X, T


@pytest.fixture(scope="module")
def L(hx):
    from hypre_amd import learned
    return learned


@pytest.fixture(scope="module")
def PM():
    from oracle import param_model
    return param_model


def test_forward_matches_oracle(L, PM):
    say("forward")
    for shape in (DIMS, (8, 20, 36, 9)):
        p = PM.init_params(shape, seed=3).astype(np.float32)
        m = L.ParamModel(*shape, params=p)
        X, _ = batch(1, 37)
        Y = m.forward(X)
        Yr, _ = PM.forward(p.astype(np.float64), shape, X.astype(np.float64))
        np.testing.assert_allclose(Y, Yr, rtol=1e-4, atol=1e-6)
        m.destroy()


def test_gradient_matches_oracle(L, PM):
    say("gradient")
    p = (PM.init_params(DIMS, seed=4) + 0.01).astype(np.float32)
    m = L.ParamModel(*DIMS, params=p)
    X, T = batch(2)
    g, loss = m.gradient(X, T)
    gr, lr = PM.gradient(p.astype(np.float64), DIMS, X.astype(np.float64), T.astype(np.float64))
    assert abs(loss - lr) <= 1e-5 * abs(lr)
    scale = np.abs(gr).max()
    assert np.abs(g - gr).max() <= 2e-3 * scale
    m.destroy()


def test_adam_training_matches_oracle(L, PM):
    say("adam")
    p = PM.init_params(DIMS, seed=5).astype(np.float32)
    m = L.ParamModel(*DIMS, params=p)
    X, T = batch(3)
    hist = m.train(X, T, 50, 1e-3)
    pr, hr = PM.adam_steps(p.astype(np.float64), DIMS, X.astype(np.float64), T.astype(np.float64), 50, 1e-3)
    np.testing.assert_allclose(hist, hr, rtol=1e-3)
    assert hist[-1] < hist[0]
    np.testing.assert_allclose(m.get_params(), pr, rtol=0, atol=2e-4)
    m.destroy()


def test_training_is_reproducible(L, PM):
    say("reproducible")
    X, T = batch(6)
    out = []
    for _ in range(2):
        m = L.ParamModel(*DIMS, params=PM.init_params(DIMS, seed=8).astype(np.float32))
        h = m.train(X, T, 20, 3e-3)
        out.append((h, m.get_params()))
        m.destroy()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_matrix_features_match_oracle(hx, oracle, L, PM):
    say("features")
    for A in (oracle.gen_aniso7(12, 0.01, 0.7), oracle.gen_vardiff7(10, 2, -2.0, 2.0, 3),
              oracle.gen_laplace27(7, 7, 7)):
        G = hx.ParCSRMatrix.from_csr(A.ia, A.ja, A.a)
        np.testing.assert_allclose(L.matrix_features(G), PM.matrix_features(A), rtol=1e-12, atol=1e-14)


def test_model_sets_per_level_parameters(hx, L, PM):
    """BoomerAMGSetParamModel gives the hierarchy of the decoded settings."""
    say("apply")
    p = PM.init_params(DIMS, seed=9).astype(np.float32)
    m = L.ParamModel(*DIMS, params=p)
    A = hx.ParCSRMatrix.aniso7(20, 0.05, 0.4)
    y = m.forward(L.matrix_features(A)[None, :].astype(np.float32))[0]
    lp = L.decode(y)
    a1 = hx.BoomerAMG(keep_setup_data=1)
    m.apply(a1, A)
    a1.setup(A)
    a2 = hx.BoomerAMG(keep_setup_data=1)
    th, om, sw = zip(*lp)
    a2.set_level_params(theta=th, omega=om, sweeps=sw)
    a2.setup(A)
    assert a1.num_levels() == a2.num_levels()
    for l in range(a1.num_levels()):
        M1, M2 = a1.level_matrix(l, "A"), a2.level_matrix(l, "A")
        assert np.array_equal(M1.ja, M2.ja) and np.array_equal(M1.a, M2.a)
    m.destroy()


def test_learned_parameters_reduce_work_on_held_out_problems(hx, L):
    """The committed model (trained on seeds 1..64 of the anisotropic family)
    on held-out seeds: with the fallback guard the chosen settings never need
    more work per digit than the defaults, and on average they need less."""
    say("held-out gain")
    model = L.ParamModel.load()
    ratios, picks = [], []
    for seed in (1001, 1002, 1003, 1004, 1005, 1006):
        A, b, eps, th = L.aniso_problem(64, seed)
        amg = hx.BoomerAMG()
        which, learned, default = L.apply_with_fallback(amg, model, A, b)
        chosen = learned if which == "learned" else default
        assert chosen.rel_res < 1e-8
        assert chosen.work_per_digit <= default.work_per_digit
        ratios.append(chosen.work_per_digit / default.work_per_digit)
        picks.append(which)
        say(seed, f"eps={eps:.3g} rot={th:.3f}", which, f"its {learned.iterations} vs {default.iterations}",
            f"ratio {ratios[-1]:.3f}")
        amg.destroy()
    assert np.mean(ratios) < 0.97, ratios
    assert "learned" in picks
