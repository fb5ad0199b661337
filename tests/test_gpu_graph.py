"""Whole-substep HIP graph replay against eager launches.

Every kernel of the substep is deterministic (integer atomics and max
reductions only), so a replayed graph must give bit-identical states and
statistics.  The graph is replayed several times with eager work and host
synchronisations in between, and every replay is compared."""
import numpy as np
import pytest

from gpu_common import Scene, block_scene

pytestmark = pytest.mark.gpu

sa = pytest.importorskip("sphpc_amd")


def _pair(tol, iters=4):
    s = 0.02
    sc = Scene(s, (2.0, 1.0, 1.0), iters=iters, tol=tol, k=0.5, c=0.01, vort_eps=0.01,
               vort_model=1, lambda_eps=600.0)
    pos, vel = block_scene((24, 20, 22), s, (0.1, 0.05, 0.1), jitter=0.01, seed=3,
                           vel_sigma=0.1, swirl=0.3)
    ctxs = []
    for graph in (False, True):
        c = sc.context(len(pos))
        dims, wmin, wmax = (32, 32, 32), (0.5, 0.0, 0.0), (1.3, 0.8, 0.8)
        ax = [wmin[i] + (np.arange(32) + 0.5) * 0.025 for i in range(3)]
        Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        sdf = (np.sqrt((X - 0.9) ** 2 + (Y - 0.3) ** 2 + (Z - 0.2) ** 2) - 0.15).astype(np.float32)
        c.set_obstacle(sdf.ravel(), dims, 0.025, wmin, wmax)
        c.upload(pos, vel)
        c.set_graph(graph)
        ctxs.append(c)
    return sc, ctxs


def _same(a, b, step):
    pa, va = a.download()
    pb, vb = b.download()
    assert np.array_equal(pa, pb), step
    assert np.array_equal(va, vb), step
    sa_, sb = a.stats(), b.stats()
    assert sa_.iterations == sb.iterations, step
    assert sa_.converged == sb.converged, step
    assert sa_.max_density_error == sb.max_density_error, step
    assert sa_.neighbor_pairs == sb.neighbor_pairs, step
    assert sa_.substeps == sb.substeps, step


@pytest.mark.parametrize("tol", [0.0, 1e-3])
def test_graph_replay_bit_identical(tol):
    sc, (eager, graph) = _pair(tol)
    for step in range(6):
        eager.step(1)
        graph.step(1)
        graph.synchronize()
        _same(eager, graph, step)
        # eager work on the graph context between replays (FP64 block sums
        # are added atomically in any order: equal to rounding)
        ke_g, ke_e = graph.kinetic_energy(), eager.kinetic_energy()
        assert abs(ke_g - ke_e) <= 1e-12 * max(abs(ke_e), 1e-300)
    # a parameter change rebuilds the graph
    p = sc.gpu_params()
    p.solve_iterations = 6
    for c in (eager, graph):
        c.set_params(p)
    eager.step(3)
    graph.step(3)
    _same(eager, graph, "after set_params")


def test_graph_replay_with_warm_start():
    sc, (eager, graph) = _pair(1e-3, iters=10)
    net = sa.Mlp(seed=4)
    mean = np.array([1.0, 0.0, 0.01, 0.9, 0.0, 4.0, 0, 0, 0, 0, 0, 0], np.float32)
    std = np.array([0.3, 0.01, 0.01, 0.05, 0.01, 1.0, 1, 1, 1, 0.1, 0.1, 0.1], np.float32)
    net.set_normalization(mean, std, 0.005, 0.01)
    for c in (eager, graph):
        c.set_warm_start(net)
    for step in range(5):
        eager.step(1)
        graph.step(1)
        _same(eager, graph, step)
        assert np.array_equal(eager.field(sa.FIELD_LAMBDA_SUM), graph.field(sa.FIELD_LAMBDA_SUM))


def test_autotune_keeps_state_and_results():
    """The tuner restores the state it started from, and substeps with the
    tuned launch geometry are bit-identical to the default geometry."""
    sc, (ref, tuned) = _pair(1e-3)
    p0, v0 = tuned.download()
    t = tuned.autotune(2)
    p1, v1 = tuned.download()
    assert np.array_equal(p0, p1) and np.array_equal(v0, v1)
    assert all(b in (64, 128, 256) for b in t["block"].values())
    assert all(d in (8, 16, 32) for d in t["depth"].values())
    assert t["ms_per_substep_tuned"] > 0
    for step in range(3):
        ref.step(1)
        tuned.step(1)
        _same(ref, tuned, step)


def test_list_format_does_not_change_results():
    """Compact and wide neighbour lists give bit-identical substeps (same
    neighbours in the same order), eager and graph-replayed."""
    sc, (compact, wide) = _pair(1e-3)
    wide.set_tuning(wide_lists=True)
    assert wide.tuning()["wide_lists"] and not compact.tuning()["wide_lists"]
    for step in range(4):
        compact.step(1)
        wide.step(1)
        _same(compact, wide, step)
    ia, ca = compact.neighbors()
    ib, cb = wide.neighbors()
    assert np.array_equal(ca, cb) and np.array_equal(ia, ib)


def test_set_tuning_validates_and_round_trips():
    sc, (a, b) = _pair(0.0)
    a.set_tuning(block={"density_lambda": 64, "xsph": 128}, depth={"dp_update": 32},
                 split_collide=False, wide_lists=True)
    t = a.tuning()
    assert t["block"]["density_lambda"] == 64 and t["block"]["xsph"] == 128
    assert t["depth"]["dp_update"] == 32 and t["depth"]["density_lambda"] == 16
    assert t["wide_lists"] and not t["split_collide"]
    with pytest.raises(sa.SphError):
        a.set_tuning(block={"color": 96})
    with pytest.raises(sa.SphError):
        a.set_tuning(depth={"color": 12})
    for step in range(3):
        a.step(1)
        b.step(1)
        _same(a, b, step)


@pytest.mark.parametrize("depth,block,wide", [(8, 64, False), (32, 128, False), (8, 128, True),
                                              (32, 64, True)])
def test_launch_choices_do_not_change_results(depth, block, wide):
    """Every gather kernel at another depth (its own template instance) and
    block size, compact or wide lists: bit-identical to the defaults."""
    sc, (ref, other) = _pair(1e-3)
    names = [k for k in sa.kernel_names()
             if k in ("build_lists", "density_lambda", "dp_update", "color", "surface", "xsph",
                      "vorticity", "confine")]
    other.set_tuning(block={k: block for k in names}, depth={k: depth for k in names},
                     wide_lists=wide, split_collide=False)
    for step in range(3):
        ref.step(1)
        other.step(1)
        _same(ref, other, step)
