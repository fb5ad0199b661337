"""The 8,388,608-particle double dam break (two 256x128x128 blocks at
s = 0.008 m in a 6 x 2 x 2 m box, torus + box obstacle on a 256^3 SDF, all
terms on, 4 Jacobi iterations, vorticity model 1) against the FP64 oracle
from shared states.

The GPU runs the workload with the fused substep.  At each checked substep
the GPU's state at the start of the substep is the shared state:

* the stage entry points run the substep from it, and after every stage the
  oracle runs the same stage on the GPU's own inputs (the whole state,
  promoted to double): cell keys, (key, id) order, cell tables and the
  60-neighbour lists bit-exact (the cap's truncation included, which the late
  substeps exercise: tens of thousands of particles sit at the cap there),
  density per particle within 1e-5 relative, lambda and delta p normwise
  within 1e-5, positions within 8 FP32 ulp of |x| (up to 6 m) plus 1e-5 of the
  displacement, velocity passes as in the small-scene tests;
* the fused substep runs from it, and the oracle runs the rest of the substep
  (lists, the Jacobi iterations, the velocity passes, the commit) from the
  GPU's predicted positions, whose own error is checked to two ulp: positions
  must agree within 8 ulp of each particle's largest coordinate plus 1e-5 of
  the displacement, velocities normwise within 1e-5 plus that position bound over
  dt, for every particle at substeps 20-29.  At 100-101, after the
  instability, particles pressed nearly onto each other take delta p of up to
  metres whose direction a change of an ulp turns: there at most 1% of the
  particles may fall outside the bound (positions, and velocities within that
  bound over dt plus 1e-5 of the largest), and the test records how many the
  oracle's own substep moves outside it when x* is moved by one float32 ulp
  per coordinate (up or down at random, same lists).  (From the oracle's own FP64
  prediction instead, pairs within an ulp of
  r = h fall on different sides of it in the two predictions; list membership
  is fixed for the substep, and where the update then moves such a pair closer
  it interacts in one solve only: at substep 20, 3,652 particles' neighbour
  counts differ between the two predictions, moving particles by up to 6e-5 m,
  while from the GPU's prediction the two solves agree to 1.4e-6 m.)

The substeps are 20-29 (the two columns collapsing and spreading) and 100-101
(after the instability of the configuration, with the 60-neighbour cap in
use)."""
import pytest

import orc
import stage_checks as chk
from gpu_common import Scene

pytestmark = pytest.mark.gpu

sa = pytest.importorskip("sphpc_amd")
from sphpc_amd import workloads  # noqa: E402

CHECKED = list(range(20, 30)) + [100, 101]


@pytest.fixture(scope="module")
def run():
    wl = workloads.double_dam_break(1.0)
    sc = Scene(wl["s"], wl["box"], iters=4, k=0.5, c=0.01, lambda_eps=600.0, vort_eps=0.01,
               vort_model=1)
    n = len(wl["pos"])
    ctx = sc.context(n)
    ctx.set_obstacle(wl["sdf"], wl["dims"], wl["dx"], wl["wmin"], wl["wmax"])
    ctx.upload(wl["pos"])
    obs = orc.obstacle(wl["sdf"], wl["dims"], wl["dx"], wl["wmin"], wl["wmax"])
    st = dict(ctx=ctx, sc=sc, obs=obs, done=0, n=n)
    yield st
    ctx.close()


@pytest.mark.parametrize("substep", CHECKED)
def test_substep_from_shared_state(run, substep):
    chk.substep_from_shared_state(run, sa, substep, f"window8m_{substep}", late=substep >= 100)
