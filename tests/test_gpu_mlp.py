"""GPU parity of the lambda warm-start network kernels against the FP64 oracle
(oracle/mlp_oracle.py, whose own gradient is pinned by finite differences).

Tolerances: FP32 training kernels within 1e-4 normwise relative of the FP64
gradient and loss; Adam within 1e-5 of the learning rate per parameter; the
BF16 MFMA inference within 2e-3 (relative to max |y|) of an FP64 evaluation
that applies the same BF16 roundings."""
import os
import sys

import numpy as np
import pytest

from gpu_common import record, rel_err

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu

sa = pytest.importorskip("sphpc_amd")


def _dataset(count, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((count, sa.MLP_STRIDE), np.float32)
    scale = np.array([2, 0.05, 0.1, 0.3, 0.02, 4, 1, 1, 1, 0.2, 0.2, 0.2])
    shift = np.array([1, 0.01, 0.02, 0.9, 0.0, 1, 0, 0, 0, 0, 0, 0])
    x[:, :12] = shift + scale * rng.normal(size=(count, 12))
    x[:, 12:] = rng.normal(size=(count, 4))  # ignored columns
    t = (0.03 * np.tanh(x[:, 1] / 0.05) + 0.01 * x[:, 3] + 0.02 * np.sin(x[:, 2] * 10)
         ).astype(np.float32)
    return x, t, shift, scale


def _net(seed, mean, std, tm, ts):
    net = sa.Mlp()
    theta = M.init(seed).astype(np.float32)
    theta[1024:1088] = np.random.default_rng(seed).normal(0, 0.1, 64)  # b1
    net.set_weights(theta)
    net.set_normalization(mean, std, tm, ts)
    return net, theta.astype(np.float64)


def _norm_inputs(x, t, mean, std, tm, ts):
    xn = M.normalize(x, mean, std)
    tn = (np.float32(t) - np.float32(tm)) / np.float32(ts)
    return xn, tn.astype(np.float64)


@pytest.mark.parametrize("count", [1, 1000, 200_003])
def test_gradient_matches_oracle(count):
    x, t, shift, scale = _dataset(count, count)
    mean, std, tm, ts = shift, scale, 0.01, 0.03
    net, theta = _net(1, mean, std, tm, ts)
    xb, tb = sa.DeviceBuffer.from_array(x), sa.DeviceBuffer.from_array(t)
    loss, g = net.gradient(xb, tb, count)
    xn, tn = _norm_inputs(x, t, mean, std, tm, ts)
    l64, g64 = M.loss_and_grad(theta, xn, tn)
    e_g, e_l = rel_err(g, g64), abs(loss - l64) / max(l64, 1e-30)
    record(f"mlp_gradient_{count}", grad_rel=e_g, loss_rel=e_l)
    assert e_g < 1e-4, e_g
    assert e_l < 1e-4, e_l
    assert np.all(g[12 * 64:16 * 64] == 0.0)  # padding inputs


def test_adam_steps_match_oracle():
    count = 4096
    x, t, shift, scale = _dataset(count, 7)
    net, theta = _net(2, shift, scale, 0.0, 0.02)
    xb, tb = sa.DeviceBuffer.from_array(x), sa.DeviceBuffer.from_array(t)
    m = np.zeros(M.N_PARAMS)
    v = np.zeros(M.N_PARAMS)
    lr = 1e-3
    th = net.weights().astype(np.float64)
    for step in (1, 2, 3):
        _, g = net.gradient(xb, tb, count)  # the gradient the step below uses
        net.train_step(xb, tb, count, 0, lr, 0)
        th, m, v = M.adam(th, m, v, g.astype(np.float64), step, lr=lr)
        got = net.weights().astype(np.float64)
        err = float(np.max(np.abs(got - th)))
        record(f"mlp_adam_step{step}", abs_err_over_lr=err / lr)
        assert err < 1e-5 * lr + 4 * np.spacing(np.float32(np.max(np.abs(th)))), (step, err)
        th = got  # continue from the device's FP32 weights


def test_forward_fp32_matches_oracle():
    count = 10_000
    x, t, shift, scale = _dataset(count, 3)
    net, theta = _net(3, shift, scale, 0.02, 0.5)
    xb = sa.DeviceBuffer.from_array(x)
    yb = sa.DeviceBuffer(4 * count)
    net.forward(xb, count, yb)
    y = yb.to_array(np.float32, count)
    y64 = M.forward(theta, M.normalize(x, shift, scale))[0] * 0.5 + 0.02
    e = rel_err(y, y64)
    record("mlp_forward_fp32", rel=e)
    assert e < 1e-5, e


@pytest.mark.parametrize("count", [1, 63, 64, 130, 50_000])
def test_infer_bf16_matches_emulation(count):
    """MFMA lane layouts, the k permutation of layer 2 and the BF16 roundings:
    a wrong layout gives O(1) errors, a missing rounding errors near 1e-2."""
    x, t, shift, scale = _dataset(count, 11 + count)
    net, theta = _net(4, shift, scale, 0.0, 1.0)
    xb = sa.DeviceBuffer.from_array(x)
    yb = sa.DeviceBuffer(4 * count)
    net.infer_bf16(xb, count, yb)
    y = yb.to_array(np.float32, count).astype(np.float64)
    xn = M.normalize(x, shift, scale)
    y_emul = M.forward_bf16(theta, xn)
    y_f64 = M.forward(theta, xn)[0]
    ymax = max(float(np.max(np.abs(y_emul))), 1e-30)
    e_emul = float(np.max(np.abs(y - y_emul))) / ymax
    e_f64 = float(np.max(np.abs(y - y_f64))) / ymax
    record(f"mlp_infer_bf16_{count}", rel_vs_emulation=e_emul, rel_vs_fp64=e_f64)
    assert e_emul < 2e-3, e_emul
    if count >= 1000:
        assert e_emul < 0.5 * e_f64  # closer to the BF16 model than to FP64


def test_fit_normalization_and_save_load(tmp_path):
    count = 100_000
    x, t, _, _ = _dataset(count, 5)
    net = sa.Mlp(seed=9)
    xb, tb = sa.DeviceBuffer.from_array(x), sa.DeviceBuffer.from_array(t)
    net.fit_normalization(xb, tb, count)
    mean, std, tm, ts = net.normalization()
    assert np.allclose(mean, x[:, :12].astype(np.float64).mean(0), rtol=1e-5, atol=1e-6)
    assert np.allclose(std, x[:, :12].astype(np.float64).std(0), rtol=1e-4)
    assert abs(tm - t.astype(np.float64).mean()) < 1e-6
    assert abs(ts - t.astype(np.float64).std()) < 1e-4 * ts
    path = tmp_path / "net.bin"
    net.save(path)
    other = sa.Mlp(seed=1)
    other.load(path)
    assert np.array_equal(other.weights(), net.weights())
    m2, s2, tm2, ts2 = other.normalization()
    assert np.allclose(m2, mean) and np.allclose(s2, std, rtol=1e-6)
    assert tm2 == tm and ts2 == ts


def test_training_on_gpu_reduces_loss():
    count = 65_536
    x, t, shift, scale = _dataset(count, 21)
    net = sa.Mlp(seed=3)
    xb, tb = sa.DeviceBuffer.from_array(x), sa.DeviceBuffer.from_array(t)
    net.fit_normalization(xb, tb, count)
    l0, _ = net.gradient(xb, tb, count)
    for step in range(400):
        net.train_step(xb, tb, count, 4096, 2e-3, 1000 + step)
    l1, _ = net.gradient(xb, tb, count)
    record("mlp_training", loss0=l0, loss400=l1)
    assert l1 < 0.1 * l0, (l0, l1)
