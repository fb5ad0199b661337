"""Long runs against the FP64 oracle, by statistics.

FP32 on the GPU and FP64 in the oracle start from the same state and follow
the same algorithm, but a dam break is chaotic: single trajectories drift
apart, so the comparison is by statistics at checkpoints (the oracle's, in
golden/long_run.json, written by golden/make_long_run.py): kinetic energy,
mean and maximum compression error of the last density pass, mean neighbour
count.  The bounds are written per statistic below; the maximum over particles
is the noisiest of them."""
import json
import os

import numpy as np
import pytest

from gpu_common import Scene, record

pytestmark = pytest.mark.gpu

sa = pytest.importorskip("sphpc_amd")

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "long_run.json")))

# relative bound, absolute floor.  The maximum over particles follows single
# splashing particles and is only held to a factor of two.
# Observed worst deviations: kinetic energy 0.9%, mean error 3%, 99th
# percentile 20%, maximum 83%, mean neighbour count 0.12% (32k plain dam
# break, 500 substeps; the 131k case with all terms stays within half of these).
BOUNDS = {"ke": (0.03, 1e-3), "mean_err": (0.10, 5e-4), "p99_err": (0.25, 2e-3),
          "max_err": (1.0, 0.02), "mean_nb": (0.02, 0.0)}
# The 8M double dam break turns unstable after ~80 substeps (in the oracle as
# on the GPU: tens of thousands of particles at the 60-neighbour cap); by
# substep 110 its kinetic energies differ by 5% while mean error, 99th
# percentile and neighbour count still agree within 0.2%.
# The 1M block falling onto the sphere, run for its full 1000 substeps: the
# splash off the sphere is chaotic.  From substep 300 on the kinetic energies of
# the FP32 and FP64 runs differ by 3-20%; two FP32 runs on the GPU whose initial
# states differ by one float32 ulp per coordinate differ by 3-18% over the same
# checkpoints, and two FP64 oracle runs with that perturbation by up to 50% (mean
# error up to 15%, 99th percentile up to 42%), so that is the flow's own
# sensitivity.  The density statistics stay within the common bounds.
CASE_BOUNDS = {"double_dam_break_8m": {"ke": (0.08, 1e-3)},
               "sphere_block_1m_1000": {"ke": (0.25, 1e-3)}}


def _context(name):
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_long_run import CASES, scene
    case = CASES[name]
    s, pos, box, sc = scene(case)
    ph = dict(case["physics"])
    for key in ("surface", "viscosity", "vorticity"):
        ph.pop(key, None)
    scn = Scene(s, box, iters=case["iters"], **ph)
    ctx = scn.context(len(pos))
    if sc is not None:
        ctx.set_obstacle(sc["sdf"], sc["dims"], sc["dx"], sc["wmin"], sc["wmax"])
    ctx.upload(pos)
    return case, s, ctx


@pytest.mark.parametrize("name", sorted(GOLD))
def test_long_run_statistics_match_oracle(name):
    case, s, ctx = _context(name)
    gold = {c["substep"]: c for c in GOLD[name]["checkpoints"]}
    worst = {k: 0.0 for k in BOUNDS}
    bad = []
    done = 0
    for sub in sorted(gold):
        ctx.step(sub - done)
        done = sub
        rho = ctx.field(sa.FIELD_DENSITY).astype(np.float64)
        e = np.maximum(rho - 1000.0, 0.0) / 1000.0
        cnt = ctx.neighbor_counts()
        got = dict(ke=ctx.kinetic_energy(), mean_err=float(e.mean()), max_err=float(e.max()),
                   p99_err=float(np.quantile(e, 0.99)), mean_nb=float(cnt.mean()))
        for k, (rel, floor) in {**BOUNDS, **CASE_BOUNDS.get(name, {})}.items():
            ref = gold[sub][k]
            worst[k] = max(worst[k], abs(got[k] - ref) / max(abs(ref), 1e-300))
            if abs(got[k] - ref) > rel * abs(ref) + floor:
                bad.append((sub, k, got[k], ref))
        if int(np.sum(cnt >= sa.MAX_NEIGHBORS)) > gold[sub]["at_cap"] + len(cnt) // 1000:
            bad.append((sub, "at_cap", int(np.sum(cnt >= sa.MAX_NEIGHBORS)), gold[sub]["at_cap"]))
    record("long_run_" + name, substeps=done, **{"worst_rel_" + k: v for k, v in worst.items()})
    assert not bad, bad
