"""Pins the FP64 MLP oracle: analytic gradient against central finite
differences, SiLU derivative, padding inputs inert, Adam's first step size."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import mlp_oracle as M  # noqa: E402


def test_parameter_count():
    assert M.N_PARAMS == 16 * 64 + 64 + 64 * 64 + 64 + 64 + 1 == 5313


def test_silu_derivative_matches_finite_difference():
    z = np.linspace(-6, 6, 101)
    e = 1e-6
    fd = (M.silu(z + e) - M.silu(z - e)) / (2 * e)
    assert np.max(np.abs(fd - M.dsilu(z))) < 1e-8


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    theta = M.init(1)
    x = rng.normal(size=(17, M.N_IN))
    t = rng.normal(size=17)
    _, g = M.loss_and_grad(theta, x, t)
    idx = rng.choice(M.N_PARAMS, 300, replace=False)
    # every block of the parameter vector is probed
    idx = np.unique(np.concatenate([idx, [0, 5, 16 * 64 + 3, 1100, 5000, 5250, 5312]]))
    e = 1e-6
    for i in idx:
        tp = theta.copy(); tp[i] += e
        tm = theta.copy(); tm[i] -= e
        fd = (M.loss_and_grad(tp, x, t)[0] - M.loss_and_grad(tm, x, t)[0]) / (2 * e)
        if M.N_IN * M.N_H <= i < M.N_PAD * M.N_H:  # padding rows of W1
            assert g[i] == 0.0
        else:
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_padding_inputs_do_not_change_output():
    theta = M.init(2)
    x = np.random.default_rng(3).normal(size=(5, M.N_IN))
    xp = M.pad(x)
    xp[:, M.N_IN:] = 7.0
    assert np.allclose(M.forward(theta, x)[0], M.forward(theta, xp)[0])


def test_adam_first_step_is_lr_times_sign():
    theta = np.zeros(4)
    g = np.array([0.5, -2.0, 1e-3, 0.0])
    th, m, v = M.adam(theta, np.zeros(4), np.zeros(4), g, 1, lr=0.01)
    assert np.allclose(th[:3], -0.01 * np.sign(g[:3]), rtol=1e-4)
    assert th[3] == 0.0


def test_training_reduces_loss_on_a_smooth_target():
    rng = np.random.default_rng(4)
    x = rng.normal(size=(512, M.N_IN))
    t = np.tanh(x[:, 0] - 0.5 * x[:, 3] + 0.3 * x[:, 7] * x[:, 1])
    theta = M.init(5)
    m = np.zeros_like(theta); v = np.zeros_like(theta)
    l0 = M.loss_and_grad(theta, x, t)[0]
    for step in range(1, 301):
        l, g = M.loss_and_grad(theta, x, t)
        theta, m, v = M.adam(theta, m, v, g, step, lr=3e-3)
    assert l < 0.1 * l0


def test_bf16_rounding_known_values():
    v = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, 1 + 2 ** -7 + 2 ** -8, -2.5,
                  3.0e-3, 65504.0])
    r = M.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie -> even mantissa
    assert r[2] == 1 + 2 ** -7              # above the tie -> up
    assert r[3] == 1 + 2 ** -6              # tie from an odd mantissa -> up to even
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8
    assert r[6] == 65536.0                  # 65504 is closer to 2^16 at 8 bits
    u = M.bf16(np.random.default_rng(0).normal(size=1000)).astype(np.float32).view(np.uint32)
    assert np.all((u & 0xFFFF) == 0)


def test_bf16_forward_close_to_fp64_forward():
    theta = M.init(3)
    xn = np.random.default_rng(5).normal(size=(256, M.N_IN))
    y64 = M.forward(theta, xn)[0]
    yb = M.forward_bf16(theta, xn)
    assert np.max(np.abs(yb - y64)) < 0.05 * np.max(np.abs(y64))
    assert np.max(np.abs(yb - y64)) > 0  # the rounding is really applied
