"""Pins of the surrogate's CPU oracle (oracle/mlp.py): exact gradients by
central finite differences, the closed-form baseline against the C ray
integrator with the anomaly switched off, and bfloat16 rounding against
torch's conversion."""
import numpy as np
import torch

from oracle import mlp as M
from oracle import oracle as O
from q_psdm_amd import synth
from tests.helpers import o_model


def _batch(n, seed=3):
    rng = np.random.default_rng(seed)
    lo = np.array([0.0, 0.0, -50.0, -50.0, 0.0])
    hi = np.array([200.0, 160.0, 250.0, 210.0, 240.0])
    X = lo + (hi - lo) * rng.random((n, 5))
    return X, lo, hi


def test_gradient_matches_finite_differences():
    cfg = synth.tiny_layers()
    X, lo, hi = _batch(24)
    rng = np.random.default_rng(8)
    T = M.base_leg(cfg.model, X) + rng.normal(0, [0.003, 0.0005, 0.02], (24, 3))
    p = M.init_params(5)
    p += rng.normal(0, 0.05, p.size)          # non-zero biases, generic point
    scale = (0.01, 0.002, 0.05)
    loss, g = M.forward_backward(p, X, T, lo, hi, cfg.model, scale, bf16_round=False)
    idx = np.concatenate([rng.choice(np.arange(0, 640), 6), rng.choice(np.arange(640, 768), 3),
                          rng.choice(np.arange(768, 50304), 24), rng.choice(np.arange(50304, 50691), 8)])
    for k in idx:
        e = 1e-6 * max(1.0, abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] += e
        pm[k] -= e
        lp, _ = M.forward_backward(pp, X, T, lo, hi, cfg.model, scale, bf16_round=False)
        lm, _ = M.forward_backward(pm, X, T, lo, hi, cfg.model, scale, bf16_round=False)
        fd = (lp - lm) / (2 * e)
        assert abs(fd - g[k]) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(g).max()), (k, fd, g[k])


def test_baseline_equals_ray_integrals_without_anomaly():
    X, _, _ = _batch(200, seed=11)
    X[:5, 4] = 0.0  # surface points
    for cfg in (synth.tiny(), synth.tiny_layers(), synth.config5()):
        m = dict(cfg.model, av=0.0, aq=0.0)
        om = o_model(m)
        b = M.base_leg(cfg.model, X * (20.0 if cfg.name == "config5" else 1.0))
        Xs = X * (20.0 if cfg.name == "config5" else 1.0)
        for i in range(0, 200, 7):
            t, ts = O.ray_times(om, *Xs[i])
            sg = O.ray_sigma(om, *Xs[i])
            assert abs(b[i, 0] - t) <= 1e-12 * t and abs(b[i, 1] - ts) <= 1e-12 * ts
            assert abs(b[i, 2] + np.log(sg)) <= 1e-12 * abs(np.log(sg))


def test_sigma_closed_form_and_homogeneous_limit():
    # linear v(z): sigma = L (v0 + k z / 2); one layer, no anomaly: v L
    om = O.model_vz(1500.0, 0.6, 100.0)
    assert abs(O.ray_sigma(om, 0, 0, 300, 400, 1000) - np.sqrt(1.25e6) * 1800.0) < 1e-6
    lay = O.model_layers([2500.0], [80.0], 500.0)
    L = np.sqrt(100 ** 2 + 50 ** 2 + 700 ** 2)
    assert abs(O.ray_sigma(lay, 10, 20, 110, 70, 700) - 2500.0 * L) < 1e-6 * 2500 * L


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-3, 5000), [1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8,
                                                                            -1.0 - 2 ** -8, 0.0]])
    ref = torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(M.bf16(x), ref)


def test_bf16_forward_close_to_exact():
    cfg = synth.tiny_layers()
    X, lo, hi = _batch(64)
    p = M.init_params(1)
    a = M.forward(p, X, lo, hi, cfg.model, (0.01, 0.002, 0.05), bf16_round=False)
    b = M.forward(p, X, lo, hi, cfg.model, (0.01, 0.002, 0.05), bf16_round=True)
    assert np.abs(a - b).max() < 0.05 * 0.01 * 5 and not np.array_equal(a, b)
