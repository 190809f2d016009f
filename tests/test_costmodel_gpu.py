"""GPU checks of the cost model's hand-written HIP kernels (FP32 MLP, include/symcost.h)
against a float64 numpy restatement of the same network: forward pass, backpropagated
gradient, one Adam step, and fitting a known smooth function.

Tolerances: forward 1e-5 relative (FP32 dot products of length <= 64); gradient 1e-4 of the
largest gradient component; Adam update 1e-6 absolute at lr = 1e-3."""
import numpy as np
import pytest

from symboliclib_amd.costmodel import CostModel, unpack

pytestmark = pytest.mark.gpu
F, H = 24, 64


def _forward(params, X, mean, dev, ym):
    W1, b1, W2, b2, w3, b3 = [np.asarray(a, np.float64) for a in unpack(params, F, H)]
    Xs = (X - mean) / dev
    h1 = np.maximum(Xs @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    out = h2 @ w3 + b3
    return Xs, h1, h2, out, out * ym[1] + ym[0]


def _grad(params, X, y, mean, dev, ym):
    W1, b1, W2, b2, w3, b3 = [np.asarray(a, np.float64) for a in unpack(params, F, H)]
    Xs, h1, h2, out, _ = _forward(params, X, mean, dev, ym)
    t = (y - ym[0]) / ym[1]
    n = len(y)
    dy = 2 * (out - t) / n
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    dh2 = np.outer(dy, w3) * (h2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2) * (h1 > 0)
    gW1 = dh1.T @ Xs
    gb1 = dh1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])


_FIXED = np.random.default_rng(12345)
_SCALE = _FIXED.uniform(0.5, 3, F).astype(np.float32)
_W = _FIXED.uniform(0.1, 1.0, F)


def _data(n, seed):
    """Samples of one fixed smooth function; `seed` draws only the inputs."""
    X = np.random.default_rng(seed).normal(0, 1, (n, F)).astype(np.float32) * _SCALE
    y = np.log1p((_W * X.astype(np.float64) ** 2).sum(1)).astype(np.float32)
    return X, y


def test_forward_matches_numpy(require_device):
    X, y = _data(300, 1)
    m = CostModel.create(F, H, seed=3)
    mean, dev = X.mean(0), X.std(0) + 0.1
    ym = np.array([0.5, 2.0], np.float32)
    m.set_scaling(mean, dev, ym)
    got = m.predict(X)
    ref = _forward(m.params(), X.astype(np.float64), mean, dev, ym)[4]
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-5), np.abs(got - ref).max()


def test_gradient_and_adam_step_match_numpy(require_device):
    X, y = _data(200, 2)
    m = CostModel.create(F, H, seed=4)
    mean, dev = X.mean(0), X.std(0)
    ym = np.array([y.mean(), y.std()], np.float32)
    m.set_scaling(mean, dev, ym)
    p0 = m.params().astype(np.float64)
    lr = 1e-3
    m.step(X, y, lr)
    g_ref = _grad(p0.astype(np.float32), X.astype(np.float64), y.astype(np.float64), mean, dev, ym)
    g = m.gradient().astype(np.float64)
    scale = np.abs(g_ref).max()
    assert np.abs(g - g_ref).max() <= 1e-4 * scale, (np.abs(g - g_ref).max(), scale)
    # first Adam step: m_hat = g, v_hat = g^2 -> p -= lr g / (|g| + eps)
    p1 = m.params().astype(np.float64)
    expect = p0 - lr * g / (np.abs(g) + 1e-8)
    assert np.abs(p1 - expect).max() <= 1e-6


def test_training_fits_a_smooth_function(require_device):
    X, y = _data(8000, 5)
    Xt, yt = _data(1000, 6)
    m = CostModel.create(F, H, seed=7)
    loss = m.train(X, y, epochs=40, batch=128, lr=2e-3, seed=1)
    pred = m.predict(Xt)
    rms = float(np.sqrt(((pred - yt) ** 2).mean()))
    r = np.corrcoef(pred, yt)[0, 1]
    print(f"\ncost model fit: train loss {loss:.4f} (standardised), held-out corr {r:.4f}, "
          f"held-out rms {rms:.4f} (target std {yt.std():.4f})")
    # the fit explains most of the held-out variance of a function it has not seen these inputs of
    assert loss < 0.05 and r > 0.9 and rms < 0.5 * yt.std()


def test_save_and_load_roundtrip(require_device, tmp_path):
    X, y = _data(500, 8)
    m = CostModel.create(F, H, seed=9)
    m.train(X, y, epochs=5, batch=64, lr=1e-3)
    path = tmp_path / "model.txt"
    m.save(path)
    m2 = CostModel.load(path)
    assert np.allclose(m.predict(X), m2.predict(X), rtol=1e-6, atol=1e-6)
