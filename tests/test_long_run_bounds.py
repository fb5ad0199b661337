"""The long-run bounds and the golden file they come from (CPU only).

* The golden FP64 run of the 32k dam break, rerun for its first 50 substeps,
  reproduces the stored checkpoint exactly (the oracle and the golden file
  agree; the statistics do not depend on the thread count).
* The same run with the particle mass, or gravity, changed by 1e-4 (a
  mistake planted in the oracle) falls outside the bounds the GPU is held to
  at that checkpoint: the bounds are loose enough for float32 and still tell
  a changed algorithm apart.
* Every case's envelope holds the four kinds of runs at the golden run's
  checkpoints."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import long_run_bounds as lrb
import orc

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_long_run import CASES, scene, statistics  # noqa: E402

NAME = "plain_dam_break_32k"
SUB = 50


def _run(g=9.81, mass_factor=1.0):
    case = CASES[NAME]
    s, pos, box, _ = scene(case)
    p = orc.make_params(s, box, dt=1.0 / 300.0, g=g, mass_factor=mass_factor, **case["physics"])
    n = len(pos)
    a = dict(x=np.ascontiguousarray(pos[:, 0], dtype=np.float64),
             y=np.ascontiguousarray(pos[:, 1], dtype=np.float64),
             z=np.ascontiguousarray(pos[:, 2], dtype=np.float64),
             vx=np.zeros(n), vy=np.zeros(n), vz=np.zeros(n), xs=np.zeros(n), ys=np.zeros(n),
             zs=np.zeros(n), density=np.zeros(n), lambda_=np.zeros(n),
             nbr=np.zeros(n * orc.MAXN, np.uint32), cnt=np.zeros(n, np.int32))
    st = orc.OrcState(**{k: v.ctypes.data for k, v in a.items()})
    L = orc.oracle()
    for _ in range(SUB):
        L.orc_substep(n, C.byref(st), C.byref(p), None, case["iters"], 0.0, None)
    v = np.stack([a["vx"], a["vy"], a["vz"]], 1)
    return statistics(a["density"], v, a["cnt"], s)


@pytest.fixture(scope="module")
def bounds():
    if len(lrb.GOLD[NAME].get("envelope", [])) != lrb.MEMBERS:
        pytest.fail("envelope runs missing from golden/long_run.json")
    return lrb.envelope(NAME)


def test_golden_checkpoint_reproduces(bounds):
    gold, _ = bounds
    got = _run()
    for k in lrb.STATS + ("at_cap",):
        assert got[k] == gold[SUB][k], (k, got[k], gold[SUB][k])


@pytest.mark.parametrize("mistake", [dict(mass_factor=1.0001), dict(g=9.81 * 1.0001)],
                         ids=["mass+1e-4", "gravity+1e-4"])
def test_planted_mistake_falls_outside_bounds(bounds, mistake):
    gold, env = bounds
    bad, frac = lrb.check(NAME, SUB, _run(**mistake), gold, env)
    assert bad, frac


@pytest.mark.parametrize("name", sorted(lrb.GOLD))
def test_envelope_members_cover_the_checkpoints(name):
    """Each case's envelope holds one run of each kind (one ulp up, one ulp
    down, float32 state, float32 storage), at the golden run's checkpoints."""
    g = lrb.GOLD[name]
    env = g.get("envelope", [])
    if not env:
        pytest.skip(f"{name}: envelope runs not in the golden file yet")
    kinds = sorted(("nudge", e["nudge"]) if "nudge" in e else
                   ("fp32_state",) if e.get("fp32_state") else ("f32_storage",) for e in env)
    assert kinds == [("f32_storage",), ("fp32_state",), ("nudge", -1), ("nudge", 1)], kinds
    subs = [c["substep"] for c in g["checkpoints"]]
    for e in env:
        assert [c["substep"] for c in e["checkpoints"]] == subs
