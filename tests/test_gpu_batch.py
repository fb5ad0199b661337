"""Batched independent scenes against one context per scene.

With a fixed number of Jacobi iterations every per-particle computation of a
batched context is the single-scene computation with the scene's grid, box
and obstacle, so states, neighbour lists and per-scene density errors must be
bit-identical to separate contexts running the same scenes."""
import numpy as np
import pytest

import sphpc_amd as sa_mod  # noqa: F401  (import check)
from gpu_common import record

pytestmark = pytest.mark.gpu

sa = pytest.importorskip("sphpc_amd")


def _scenes(seed, count):
    rng = np.random.default_rng(seed)
    s = float(np.float32(0.02))
    out = []
    for k in range(count):
        size = (float(rng.uniform(0.4, 0.8)), float(rng.uniform(0.3, 0.6)),
                float(rng.uniform(0.3, 0.5)))
        nb = [int(rng.integers(6, 14)), int(rng.integers(5, 12)), int(rng.integers(5, 12))]
        pos = sa.scene_lattice([sa.Block(*nb, 0.0, 0.0, 0.0)], s, jitter=0.01, seed=seed + k)
        vel = rng.normal(0, 0.1, pos.shape).astype(np.float32)
        sdf = None
        if k % 2 == 0:
            r = 0.06
            c = (0.75 * size[0], 0.5 * size[1], r)
            dims = (24, 24, 24)
            wmin = tuple(ci - 1.5 * r for ci in c)
            wmax = tuple(ci + 1.5 * r for ci in c)
            grid = sa.sdf_sample_shapes([sa.sphere(c, r)], dims, wmin, wmax)
            sdf = dict(sdf=grid, dims=dims, spacing=3 * r / dims[0], wmin=wmin, wmax=wmax)
        out.append(dict(particles=len(pos), size=size, sdf=sdf, pos=pos, vel=vel))
    return s, out


def _params(s, iters=3):
    return sa.default_params(2.0 * s, time_step=1.0 / 300.0, g=9.81, k=0.5, c=0.01,
                             rest_density=1000.0, rest_mass=1000.0 * s ** 3,
                             lambda_epsilon=600.0, vorticity_epsilon=0.01, vorticity_model=1,
                             solve_iterations=iters)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_bit_identical_to_separate_contexts(graph):
    s, scenes = _scenes(7, 6)
    p = _params(s)
    batch = sa.BatchContext(p, scenes)
    batch.upload(np.concatenate([sc["pos"] for sc in scenes]),
                 np.concatenate([sc["vel"] for sc in scenes]))
    batch.set_graph(graph)
    singles = []
    for sc in scenes:
        c = sa.Context(p, sc["size"], sc["particles"])
        if sc["sdf"] is not None:
            o = sc["sdf"]
            c.set_obstacle(o["sdf"], o["dims"], o["spacing"], o["wmin"], o["wmax"])
        c.upload(sc["pos"], sc["vel"])
        singles.append(c)
    for step in range(4):
        batch.step(1)
        for c in singles:
            c.step(1)
        bp, bv = batch.download()
        errs = batch.scene_errors()
        for k, (sc, c) in enumerate(zip(scenes, singles)):
            a, b = batch.first[k], batch.first[k + 1]
            sp, sv = c.download()
            assert np.array_equal(bp[a:b], sp), (step, k)
            assert np.array_equal(bv[a:b], sv), (step, k)
            assert errs[k] == c.stats().max_density_error, (step, k)
    # neighbour lists: batch ids are the single ids shifted by the scene offset
    nb_ids, nb_cnt = batch.neighbors()
    nb_ids = nb_ids.reshape(-1, sa.MAX_NEIGHBORS)
    for k, c in enumerate(singles):
        ids, cnt = c.neighbors()
        ids = ids.reshape(-1, sa.MAX_NEIGHBORS)
        a, b = batch.first[k], batch.first[k + 1]
        assert np.array_equal(nb_cnt[a:b], cnt)
        for i in range(b - a):
            assert np.array_equal(nb_ids[a + i, :cnt[i]], ids[i, :cnt[i]] + a)
    record("batch_vs_single", scenes=len(scenes), particles=int(batch.first[-1]),
           graph=int(graph))


def _stage_substep(ctx, iters):
    """One substep through the stage entry points, in the reference's order."""
    ctx.apply_gravity()
    ctx.predict_positions()
    ctx.find_all_neighbors()
    for it in range(iters):
        ctx.compute_densities()
        ctx.compute_lambda()
        ctx.update_dps(it)
        ctx.update_position_stars()
    ctx.update_velocities()
    ctx.compute_surface_tension()
    ctx.apply_viscosity()
    ctx.compute_vorticity()
    ctx.apply_vorticity_confinement()
    ctx.update_positions()


def test_batch_stage_path_bit_identical_to_separate_contexts():
    """The stage entry points on a batched context (per-scene boxes and
    obstacles in the delta-p stage) give the bits of separate contexts."""
    s, scenes = _scenes(11, 4)
    p = _params(s)
    batch = sa.BatchContext(p, scenes)
    batch.upload(np.concatenate([sc["pos"] for sc in scenes]),
                 np.concatenate([sc["vel"] for sc in scenes]))
    singles = []
    for sc in scenes:
        c = sa.Context(p, sc["size"], sc["particles"])
        if sc["sdf"] is not None:
            o = sc["sdf"]
            c.set_obstacle(o["sdf"], o["dims"], o["spacing"], o["wmin"], o["wmax"])
        c.upload(sc["pos"], sc["vel"])
        singles.append(c)
    for step in range(3):
        _stage_substep(batch, p.solve_iterations)
        for c in singles:
            _stage_substep(c, p.solve_iterations)
        bp, bv = batch.download()
        for k, c in enumerate(singles):
            a, b = batch.first[k], batch.first[k + 1]
            sp, sv = c.download()
            assert np.array_equal(bp[a:b], sp), (step, k)
            assert np.array_equal(bv[a:b], sv), (step, k)


def test_batch_rejects_partial_upload_and_obstacle_change():
    s, scenes = _scenes(3, 2)
    batch = sa.BatchContext(_params(s), scenes)
    with pytest.raises(sa.SphError):
        batch.upload(scenes[0]["pos"])
    with pytest.raises(sa.SphError):
        batch.set_obstacle(np.zeros(8, np.float32), (2, 2, 2), 0.1, (0, 0, 0), (0.2, 0.2, 0.2))


def test_batch_warm_start_falls_back_per_scene():
    """Warm start in a batch with a fixed iteration count: each scene decides
    the plain fallback on its own maximum error, exactly as a context of its own
    does, so every scene is bit-identical to a separate warm-started context and
    the batch reports the number of scenes that fell back.  Compressed blocks
    (the +x-face prediction raises their error) and sparse ones (they stay
    below rest density) alternate, so both outcomes occur."""
    from test_gpu_warm import _face_net
    s = float(np.float32(0.02))
    rng = np.random.default_rng(11)
    scenes = []
    for k in range(6):
        f = 0.97 if k % 2 == 0 else 1.3
        nb = [int(rng.integers(9, 14)), int(rng.integers(8, 12)), int(rng.integers(8, 12))]
        pos = sa.scene_lattice([sa.Block(*nb, 0.05, 0.02, 0.0)], s * f, jitter=0.01, seed=20 + k)
        vel = rng.normal(0, 0.05, pos.shape).astype(np.float32)
        scenes.append(dict(particles=len(pos), size=(0.6, 0.5, 0.5), sdf=None, pos=pos, vel=vel))
    p = _params(s, iters=4)
    net = _face_net(40.0)
    batch = sa.BatchContext(p, scenes)
    batch.upload(np.concatenate([sc["pos"] for sc in scenes]),
                 np.concatenate([sc["vel"] for sc in scenes]))
    batch.set_warm_start(net)
    singles = []
    for sc in scenes:
        c = sa.Context(p, sc["size"], sc["particles"])
        c.upload(sc["pos"], sc["vel"])
        c.set_warm_start(net)
        singles.append(c)
    for step in range(3):
        batch.step(1)
        falls = 0
        for c in singles:
            c.step(1)
            falls += c.stats().warm_fallback
        bp, bv = batch.download()
        for k, (sc, c) in enumerate(zip(scenes, singles)):
            a, b = batch.first[k], batch.first[k + 1]
            sp, sv = c.download()
            assert np.array_equal(bp[a:b], sp), (step, k)
            assert np.array_equal(bv[a:b], sv), (step, k)
        assert batch.stats().warm_fallback == falls, step
        if step == 0:
            first = falls
            assert 0 < falls < len(scenes), falls  # both outcomes occur
    record("batch_warm_fallback", scenes=len(scenes), first_substep_fallbacks=first)
