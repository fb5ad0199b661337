"""The 1,048,576-particle block flowing onto a sphere (one 128x128x64 block
at s = 0.01 m in the origin corner of a 4 x 2 x 2 m box, a sphere of radius
0.3 m at the box centre on a 128^3 SDF, all terms on, 4 Jacobi iterations,
vorticity model 1) against the FP64 oracle from shared states, as the 8M
double dam break is checked in test_gpu_window_8m.py (stage_checks.py:
substep_from_shared_state).

Substeps 10-11: the column collapsing.  Substeps 800-801: the flow splashing
over and around the sphere (about 100 particles within 2 cm of its surface
and 20-40 inside it before the push, particles thrown up to the ceiling), so
the obstacle push and the clamp of every wall act in the checked substeps;
there, as after the 8M run's instability, the fused substep is held at the
distribution level (at most 1% of the particles outside the per-particle
bound) and every stage per particle."""
import pytest

import orc
import stage_checks as chk
from gpu_common import Scene

pytestmark = pytest.mark.gpu

sa = pytest.importorskip("sphpc_amd")
from sphpc_amd import workloads  # noqa: E402

CHECKED = [10, 11, 800, 801]


@pytest.fixture(scope="module")
def run():
    wl = workloads.sphere_block(1.0)
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
    chk.substep_from_shared_state(run, sa, substep, f"window1m_{substep}", late=substep >= 100)
