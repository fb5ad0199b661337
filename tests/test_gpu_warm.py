"""Learned lambda warm start in the solver: logged features and targets against
the FP64 oracle, persistence of the previous multiplier, the fused MFMA
inference against the standalone one, acceptance of the tolerance solve, and
the gain of a network trained on the GPU from the solver's own samples.

Tolerances (GPU FP32 vs oracle FP64 from the same start-of-substep state):
neighbour counts exact; rho/rho0 - 1 within 1e-5; |grad C|^2 share, SDF
features within 1e-4; divergence and density moments within 1e-4 of their
largest magnitude; targets (accumulated multipliers of 3 Jacobi iterations)
within 1e-4 normwise."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import orc
from gpu_common import Scene, block_scene, d64, record, rel_err

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu

sa = pytest.importorskip("sphpc_amd")
L = orc.oracle()


def _scene(iters=3, tol=0.0, lambda_eps=600.0):
    s = 0.02
    sc = Scene(s, (2.0, 1.0, 1.0), iters=iters, tol=tol, lambda_eps=lambda_eps)
    pos, vel = block_scene((20, 20, 20), s * 0.97, (0.3, 0.0, 0.2), jitter=0.01, seed=4,
                           vel_sigma=0.2, swirl=0.5)
    return sc, pos, vel


def _samples(ctx):
    n = ctx.count()
    xb = sa.DeviceBuffer(4 * sa.MLP_STRIDE * n)
    tb = sa.DeviceBuffer(4 * n)
    assert ctx.warm_samples(xb, tb, 0, n, every=1) == n
    rows = xb.to_array(np.float32, n * sa.MLP_STRIDE).reshape(n, sa.MLP_STRIDE)
    return rows, tb.to_array(np.float32, n), xb


def _predicted(sc, pos, vel):
    """x* of ApplyGravity + PredictPositions (particles.c:516-527, 748-784), FP64."""
    p = sc.o
    x, y, z = [d64(pos[:, i]) for i in range(3)]
    vx, vy, vz = [d64(vel[:, i]) for i in range(3)]
    L.orc_apply_gravity(len(x), vz, C.byref(p))
    xs, ys, zs = np.zeros_like(x), np.zeros_like(x), np.zeros_like(x)
    L.orc_predict_positions(len(x), x, y, z, vx, vy, vz, xs, ys, zs, C.byref(p))
    return xs, ys, zs


def _check_features(rows_by_id, ref, tag):
    n = len(ref)
    assert np.array_equal(rows_by_id[:, 0], ref[:, 0].astype(np.float32))
    e1 = float(np.max(np.abs(rows_by_id[:, 1] - ref[:, 1])))
    e3 = float(np.max(np.abs(rows_by_id[:, 3] - ref[:, 3])))
    e2 = rel_err(rows_by_id[:, 2], ref[:, 2]) if np.any(ref[:, 2]) else \
        float(np.max(np.abs(rows_by_id[:, 2])))
    e4 = rel_err(rows_by_id[:, 4], ref[:, 4])
    e5 = float(np.max(np.abs(rows_by_id[:, 5:9] - ref[:, 5:9])))
    e9 = rel_err(rows_by_id[:, 9:12], ref[:, 9:12])
    e12 = rel_err(rows_by_id[:, 12], ref[:, 12])
    record(tag, rho_abs=e1, cprev_rel=e2, gradc_abs=e3, div_rel=e4, sdf_abs=e5,
           moment_rel=e9, denom_rel=e12, n=n)
    assert e1 < 1e-5, e1
    assert e2 < 1e-4, e2
    assert e3 < 1e-4, e3
    assert e4 < 1e-4, e4
    assert e5 < 1e-4, e5
    assert e9 < 1e-4, e9
    assert e12 < 1e-5, e12


@pytest.mark.parametrize("with_obstacle", [False, True])
def test_logged_features_and_targets_match_oracle(with_obstacle):
    sc, pos, vel = _scene()
    n = len(pos)
    ctx = sc.context(n)
    sdf = None
    if with_obstacle:
        dims, wmin, wmax = (48, 32, 32), (0.2, 0.0, 0.0), (0.92, 0.48, 0.48)
        grid, dx = orc.sphere_sdf(dims, wmin, wmax, (0.55, 0.1, 0.35), 0.12)
        ctx.set_obstacle(grid, dims, dx, wmin, wmax)
        sdf = dict(data=grid, dims=dims, wmin=wmin, wmax=wmax,
                   m2w=float(np.float32(wmax[0] - wmin[0]) / np.float32(dims[0] * dx)))
    ctx.upload(pos, vel)
    ctx.set_sample_logging(True)
    ctx.step(1)
    rows, t, _ = _samples(ctx)
    _, ids = ctx.sorted_pairs()
    nbr, cnt = ctx.neighbors()
    by_id = np.zeros_like(rows, dtype=np.float64)
    by_id[ids] = rows
    t_id = np.zeros(n)
    t_id[ids] = t
    xs, ys, zs = _predicted(sc, pos, vel)
    X = np.stack([xs, ys, zs], 1)
    o = sc.o
    ref = M.warm_features(X, d64(pos), nbr.reshape(n, 60), cnt, np.zeros(n), o.h,
                          o.rest_mass, o.rest_density, o.lambda_eps, sdf)
    _check_features(by_id, ref, f"warm_features_obstacle{int(with_obstacle)}")
    # targets: accumulated multiplier of the plain 3-iteration solve
    ob = orc.obstacle(sdf["data"], sdf["dims"], (sdf["wmax"][0] - sdf["wmin"][0]) / sdf["dims"][0],
                      sdf["wmin"], sdf["wmax"]) if sdf else None
    dens, lam = np.zeros(n), np.zeros(n)
    dpx, dpy, dpz = np.zeros(n), np.zeros(n), np.zeros(n)
    acc = np.zeros(n)
    lam_first = None
    for it in range(sc.iters):
        L.orc_compute_densities(n, xs, ys, zs, nbr, cnt, C.byref(o), dens)
        L.orc_compute_lambda(n, xs, ys, zs, dens, nbr, cnt, C.byref(o), lam)
        if lam_first is None:
            lam_first = lam.copy()
        acc += L.orc_stiffness(C.byref(o), it) * lam
        L.orc_update_dps(n, xs, ys, zs, lam, nbr, cnt, C.byref(o), it, dpx, dpy, dpz)
        L.orc_update_position_stars(n, xs, ys, zs, dpx, dpy, dpz, C.byref(o),
                                    C.byref(ob) if ob else None)
    t_ref = -(acc - lam_first) * ref[:, 12]  # remainder after the first iteration
    assert rel_err(by_id[:, 13], lam_first) < 1e-4
    e = rel_err(t_id, t_ref)
    lam_sum = ctx.field(sa.FIELD_LAMBDA_SUM)
    e_sum = rel_err(lam_sum, acc)
    record(f"warm_targets_obstacle{int(with_obstacle)}", target_rel=e, lambda_sum_rel=e_sum)
    assert np.max(np.abs(t_ref)) > 0
    assert e < 1e-4, e
    assert e_sum < 1e-4, e_sum


def test_previous_multiplier_persists_in_particle_order():
    sc, pos, vel = _scene()
    n = len(pos)
    ctx = sc.context(n)
    ctx.upload(pos, vel)
    ctx.set_sample_logging(True)
    ctx.step(1)
    lam1 = d64(ctx.field(sa.FIELD_LAMBDA_SUM))
    assert np.max(np.abs(lam1)) > 0
    ctx.step(1)
    rows, _, _ = _samples(ctx)
    _, ids = ctx.sorted_pairs()
    f2 = np.zeros(n)
    f2[ids] = rows[:, 2]
    den = np.zeros(n)
    den[ids] = rows[:, 12]
    e = rel_err(f2, -lam1 * den)
    record("warm_prev_multiplier", rel=e)
    assert e < 1e-6, e


def _trained_like_net(seed=5):
    net = sa.Mlp(seed=seed)
    mean = np.array([1.0, 0.0, 0.01, 0.9, 0.0, 4.0, 0, 0, 0, 0, 0, 0], np.float32)
    std = np.array([0.3, 0.01, 0.01, 0.05, 0.01, 1.0, 1, 1, 1, 0.1, 0.1, 0.1], np.float32)
    net.set_normalization(mean, std, 0.005, 0.01)
    return net


def test_fused_inference_matches_standalone():
    sc, pos, vel = _scene(iters=1)
    n = len(pos)
    ctx = sc.context(n)
    ctx.upload(pos, vel)
    ctx.set_sample_logging(True)
    ctx.step(1)
    rows, _, xb = _samples(ctx)
    _, ids = ctx.sorted_pairs()
    net = _trained_like_net()
    yb = sa.DeviceBuffer(4 * n)
    net.infer_bf16(xb, n, yb)
    y = yb.to_array(np.float32, n).astype(np.float64)
    expect = np.zeros(n)
    expect[ids] = rows[:, 13] - np.maximum(y, 0.0) / rows[:, 12]
    # same state again, warm-started with one iteration: the accumulated
    # multiplier is then exactly the warm first step's lambda
    ctx.set_sample_logging(False)
    ctx.set_warm_start(net)
    ctx.upload(pos, vel)
    ctx.step(1)
    got = d64(ctx.field(sa.FIELD_LAMBDA_SUM))
    e = rel_err(got, expect)
    record("warm_fused_vs_standalone", rel=e, frac_positive=float(np.mean(y > 0)))
    assert np.mean(y > 0) > 0.05  # the comparison is not vacuous
    assert e < 1e-5, e


def test_warm_start_keeps_the_tolerance():
    """The prediction only changes the starting point: a converged substep
    still has max density error <= tolerance, and the cap still binds."""
    sc, pos, vel = _scene(iters=10, tol=1e-3)
    n = len(pos)
    ctx = sc.context(n)
    ctx.upload(pos, vel)
    ctx.set_warm_start(_trained_like_net(7))
    for _ in range(5):
        ctx.step(1)
        st = ctx.stats()
        assert 0 <= st.iterations <= 10
        if st.converged:
            assert st.max_density_error <= 1e-3
        else:
            assert st.iterations == 10


def test_trained_network_saves_iterations():
    """Trains on samples the solver logs from six small seeded scenes (the
    production recipe of sphpc_amd/warm_train.py at s = 0.02 m) and compares
    plain and warm-started tolerance solves on a held-out scene."""
    from sphpc_amd import warm_train as W
    s = float(np.float32(0.02))
    box = (1.5, 1.0, 1.0)

    def params(iters, tol):
        return sa.default_params(2.0 * s, time_step=1.0 / 300.0, g=9.81, k=0.5, c=0.01,
                                 rest_density=1000.0, rest_mass=1000.0 * s ** 3,
                                 lambda_epsilon=600.0, vorticity_epsilon=0.01,
                                 vorticity_model=1, solve_iterations=iters,
                                 density_tolerance=tol)

    def block(rng, floor=False):
        sides = rng.uniform(0.25, 0.5, 3)
        counts = [int(a / s) for a in sides]
        origin = [rng.uniform(0.0, box[i] - counts[i] * s) for i in range(3)]
        if floor:
            origin[2] = 0.0
        return sa.Block(*counts, *origin)

    ctx = sa.Context(params(10, 1e-3), box, 25 ** 3)
    cap = 2_000_000
    xb, tb = sa.DeviceBuffer(4 * 16 * cap), sa.DeviceBuffer(4 * cap)
    rows = 0
    for sc in range(6):
        rng = np.random.default_rng(100 + sc)
        ctx.upload(sa.scene_lattice([block(rng)], s, jitter=0.01, seed=sc))
        ctx.set_params(params(10, 1e-3))
        for k in range(1, 41):
            logged = k % 5 in (0, 4)
            if logged:
                ctx.set_params(params(100, 1e-4))
                ctx.set_sample_logging(True)
            ctx.step(1)
            if k % 5 == 0:
                rows += ctx.warm_samples(xb, tb, rows, cap, 1)
            if logged:
                ctx.set_sample_logging(False)
                ctx.set_params(params(10, 1e-3))
    net = sa.Mlp(seed=2)
    net.fit_normalization(xb, tb, rows)
    l0, _ = net.gradient(xb, tb, rows)
    W.train(net, xb, tb, rows, 1500, 8192, 2e-3, 1, log=lambda m: None)
    l1, _ = net.gradient(xb, tb, rows)
    # held-out block collapsing on the floor: compressed from the first substep
    rng = np.random.default_rng(999)
    ctx.upload(sa.scene_lattice([block(rng, floor=True)], s, jitter=0.01, seed=99))
    ctx.set_params(params(10, 1e-3))
    ctx.step(30)
    pos, vel = ctx.download()
    cmp = W.compare(ctx, net, pos, vel, params(10, 1e-3), 10, log=lambda m: None)
    pl, wm = cmp["plain"], cmp["warm"]
    record("warm_gain_small", rows=rows, loss0=l0, loss1=l1,
           plain_iters=pl["iters_per_substep"], warm_iters=wm["iters_per_substep"],
           plain_final=pl["mean_final_error"], warm_final=wm["mean_final_error"],
           plain_max=pl["max_density_error"], warm_max=wm["max_density_error"],
           warm_fallbacks=wm["fallbacks"])
    assert l1 < 0.5 * l0
    assert pl["mean_final_error"] > 1e-3  # the scene is compressed: not vacuous
    # same acceptance test, fewer iterations or a smaller error at the cap
    assert (wm["iters_per_substep"] < pl["iters_per_substep"] or
            wm["mean_final_error"] < 0.9 * pl["mean_final_error"]), cmp


def _face_net(amplitude):
    """A network whose output is ~amplitude * relu(-m_x): large only for the
    particles on the +x face of a block (x density moment < 0), zero inside."""
    net = sa.Mlp(seed=1)
    theta = np.zeros(sa.MLP_PARAMS, np.float32)
    theta[9 * 64 + 0] = -1.0             # W1[9][0]: hidden 1 unit 0 = silu(-m_x / 0.05)
    theta[1088 + 0 * 64 + 0] = 1.0       # W2[0][0]
    theta[5248 + 0] = float(amplitude)   # w3[0]
    net.set_weights(theta)
    mean = np.zeros(12, np.float32)
    std = np.ones(12, np.float32)
    std[9] = 0.05
    net.set_normalization(mean, std, 0.0, 1.0)
    return net


def test_bad_prediction_falls_back_to_plain():
    """A prediction that is large only on one face of a block pushes the
    layer behind that face into the interior: the density error after the
    step exceeds the error before it, so the solver must revert the step,
    report it, and continue from the unmodified positions."""
    sc, pos, vel = _scene(iters=6)
    n = len(pos)
    plain = sc.context(n)
    plain.upload(pos, vel)
    plain.step(1)
    r_plain = plain.residuals()
    warm = sc.context(n)
    warm.upload(pos, vel)
    warm.set_warm_start(_face_net(40.0))
    warm.step(1)
    st = warm.stats()
    r_warm = warm.residuals()
    record("warm_fallback", fallback=st.warm_fallback, plain_final=float(r_plain[-1]),
           warm_final=float(r_warm[-1]), e0=float(r_warm[0]))
    assert st.warm_fallback == 1
    assert len(r_plain) == len(r_warm) == 6
    assert abs(r_warm[0] - r_plain[0]) <= 1e-6   # same state before any update
    assert r_warm[1] == r_warm[0]                # slot of the reverted step: pre-step error
    assert r_warm[-1] <= 2.0 * r_plain[-1] + 1e-4


def test_harmless_prediction_is_kept():
    """A zero remainder makes the warm first step the plain first step: no
    fallback, and the substep equals the plain one."""
    sc, pos, vel = _scene(iters=4)
    n = len(pos)
    plain = sc.context(n)
    plain.upload(pos, vel)
    plain.step(1)
    warm = sc.context(n)
    warm.upload(pos, vel)
    warm.set_warm_start(_face_net(0.0))
    warm.step(1)
    r_plain = plain.residuals()
    assert r_plain[1] <= r_plain[0]  # the plain first step lowers the error here
    assert warm.stats().warm_fallback == 0
    pw, _ = warm.download()
    pp, _ = plain.download()
    e = float(np.max(np.abs(pw - pp)))
    record("warm_zero_remainder", max_abs_dx=e)
    assert e <= 1e-6


def test_residuals_report_every_density_pass():
    sc, pos, vel = _scene(iters=5)
    n = len(pos)
    ctx = sc.context(n)
    ctx.upload(pos, vel)
    ctx.step(1)
    r = ctx.residuals()
    assert len(r) == 5 and np.all(r >= 0)
    assert r[-1] == ctx.stats().max_density_error
    p = sc.gpu_params()
    p.density_tolerance = 1e-6
    p.solve_iterations = 4
    ctx.set_params(p)
    ctx.step(1)
    r = ctx.residuals()
    st = ctx.stats()
    assert len(r) == st.iterations + 1 == 5   # capped: 4 updates + the final pass
    assert r[-1] == st.max_density_error
