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


# ---------------------------------------------------------------------------
# Closed forms of the warm-start features the reference has no code for
# ---------------------------------------------------------------------------
def _lattice64(n, s, origin=0.1):
    """Perfect cell-centred lattice in FP64, x fastest (n^3 particles)."""
    iz, iy, ix = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return origin + (np.stack([ix.ravel(), iy.ravel(), iz.ravel()], 1) + 0.5) * s


def _offsets():
    import itertools
    return [k for k in itertools.product(range(-2, 3), repeat=3)
            if 0 < k[0] ** 2 + k[1] ** 2 + k[2] ** 2 < 4]


def _features(xs, xo, s):
    import orc
    p = orc.make_params(s, (1.0, 1.0, 1.0))
    X = [np.ascontiguousarray(xs[:, a]) for a in range(3)]
    nbr, cnt, dropped = orc.neighbors(p, *X)
    assert dropped == 0
    n = len(xs)
    f = M.warm_features(xs, xo, nbr.reshape(n, orc.MAXN), cnt, np.zeros(n), p.h, p.rest_mass,
                        p.rest_density, p.lambda_eps)
    return f, p


def test_velocity_divergence_feature_uniform_expansion():
    # Feature 4 is (m/rho0) DelW_norm sum_j (h - r)^2/r ((dt v_j - dt v_i) . d),
    # d = x*_i - x*_j.  Under a uniform expansion v = a (x - c) the predicted
    # lattice is x* = c + (1 + a dt)(x - c), so dt (v_j - v_i) = -a dt/(1 + a dt) d
    # and the feature is -a dt/(1 + a dt) (m/rho0) sum_k DelW(r_k) r_k with
    # r_k = (1 + a dt) s |k| over the lattice offsets inside h = 2s.
    s, a, dt = 0.02, 3.0, 1.0 / 300.0
    xo = _lattice64(12, s)
    c = xo.mean(0)
    xs = c + (1.0 + a * dt) * (xo - c)
    f, p = _features(xs, xo, s)
    h, m, rho0 = p.h, p.rest_mass, p.rest_density
    lat = 0.0
    for k in _offsets():
        r = (1.0 + a * dt) * s * np.sqrt(k[0] ** 2 + k[1] ** 2 + k[2] ** 2)
        if r < h:
            lat += -45.0 / (np.pi * h ** 6) * (h - r) ** 2 * r
    expected = -a * dt / (1.0 + a * dt) * m / rho0 * lat
    idx = np.arange(len(xs)).reshape(12, 12, 12)[4:8, 4:8, 4:8].ravel()
    assert np.max(np.abs(f[idx, 4] - expected)) < 1e-12 * abs(expected)
    # the SPH estimate of dt div v = 3 a dt, to the lattice's discretisation error
    assert abs(expected / (3.0 * a * dt) - 1.0) < 0.15
    # a rigid translation has no divergence
    f2, _ = _features(xo + np.array([0.01, -0.02, 0.005]), xo, s)
    assert np.max(np.abs(f2[:, 4])) < 1e-12


def test_density_moment_feature_half_space():
    # Features 9-11 are -(m W_norm/(rho h)) sum_j (h^2 - r^2)^3 (x*_i - x*_j) with
    # rho the particle's own density.  On a perfect lattice (h = 2s) the terms
    # are s^6 (4 - |k|^2)^3 at offsets s k, so the moment is
    # (s/h) sum (4-|k|^2)^3 k / (64 + sum (4-|k|^2)^3): zero in the interior and
    # (0, 0, -63/267)/2 for a particle of the top layer (the k_z = -1 layer
    # sums to 27 + 4*8 + 4*1 = 63, the k_z <= 0 half to 140 + 63, plus 64 for
    # the particle itself).
    s = float(np.float32(0.02))  # h = 2 float(s) exactly, as the parameters derive it
    xs = _lattice64(12, s)
    f, _ = _features(xs, xs, s)
    grid = np.arange(len(xs)).reshape(12, 12, 12)
    inner = grid[4:8, 4:8, 4:8].ravel()
    top = grid[-1, 4:8, 4:8].ravel()
    assert np.max(np.abs(f[inner, 9:12])) < 1e-12
    expected = np.array([0.0, 0.0, -63.0 / 267.0 / 2.0])
    assert np.max(np.abs(f[top, 9:12] - expected)) < 1e-12
    # feature 1: the same particle's C = rho/rho0 - 1 with rho/rho_interior = 267/330
    assert np.allclose((f[top, 1] + 1.0) / (f[inner[0], 1] + 1.0), 267.0 / 330.0, rtol=1e-12)
