"""The plain dam break of the reference's CPU-runnable case against the FP64
oracle from shared states: 32,768 particles in a 32^3 lattice block (s = 0.02
m, seeded +-0.01 s jitter) in the corner of a 2 x 1 x 1 m box, 3 Jacobi
iterations, relaxation 600, no surface tension, XSPH, vorticity or obstacle;
the same scene as the 500-substep long run.  Checked as the larger workloads
are (stage_checks.py: substep_from_shared_state): substeps 0-1 from rest,
50-51 with the column collapsing, and 300-301, where one-ulp changes of the
initial state already send the oracle's own runs apart (the fused substep is
then held at the distribution level, every stage per particle)."""
import pytest

import orc
import stage_checks as chk
from gpu_common import Scene

pytestmark = pytest.mark.gpu

sa = pytest.importorskip("sphpc_amd")

CHECKED = [0, 1, 50, 51, 300, 301]


@pytest.fixture(scope="module")
def run():
    sc = Scene(0.02, (2.0, 1.0, 1.0), iters=3, lambda_eps=600.0)
    pos = orc.lattice_block(32, 32, 32, 0.02, origin=(0.0, 0.0, 0.0), jitter=0.01, seed=2024)
    ctx = sc.context(len(pos))
    ctx.upload(pos)
    st = dict(ctx=ctx, sc=sc, obs=None, done=0, n=len(pos))
    yield st
    ctx.close()


@pytest.mark.parametrize("substep", CHECKED)
def test_substep_from_shared_state(run, substep):
    chk.substep_from_shared_state(run, sa, substep, f"window32k_{substep}", late=substep >= 100)
