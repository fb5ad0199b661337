"""Bounds of the long-run statistics, from the FP64 oracle's own envelope.

golden/long_run.json holds, per case, the golden FP64 run's statistics at
checkpoints and its envelope: the same run from initial positions nudged by
one float32 ulp (up and down), with its state rounded to float32 after every
substep, and with every array a stage writes rounded to float32 after the
stage.  At a checkpoint the envelope of a statistic is the largest distance of
those runs from the golden run.  A run passes where each statistic lies within
ENV_FACTOR times that distance of the golden value, plus FLOOR (a float32
rounding of the statistic itself, which matters only where every envelope run
lands within rounding of the golden one); the particle count at the cap has
0.1% of the particles as its floor.

ENV_FACTOR: an FP32 solver rounds inside every stage's arithmetic too (lambda
is within a few 1e-6 normwise of FP64, a larger perturbation per substep than
the rounding of stored values), so it may drift somewhat further than the
envelope runs; the recorded MI355X runs stay within 5.3 times the envelope at
every checkpoint (the 32k case; 3.8 for 131k, at most 2.1 for the others).  The factor still separates the
oracle's own algorithm from a changed one: the oracle with the particle mass
or gravity changed by 1e-4 (relative) lands 25-90 times the envelope away in
the maximum density error at the first checkpoint of the 32k dam break
(tests/test_long_run_bounds.py plants exactly that and checks it fails)."""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "long_run.json")))

STATS = ("ke", "mean_err", "p99_err", "max_err", "mean_nb")
ENV_FACTOR = 6.0
# float32 rounding of the statistic itself: relative, absolute
FLOOR = {"ke": (1e-6, 0.0), "mean_err": (1e-6, 1e-7), "p99_err": (1e-6, 1e-7),
         "max_err": (1e-6, 1e-7), "mean_nb": (1e-6, 0.0)}
MEMBERS = 4


def envelope(name):
    """(golden checkpoints by substep, envelope by substep): per statistic
    (the particle count at the 60-neighbour cap included) the largest
    distance of an envelope run from the golden run."""
    g = GOLD[name]
    gold = {c["substep"]: c for c in g["checkpoints"]}
    env = {sub: {k: 0.0 for k in STATS + ("at_cap",)} for sub in gold}
    for run in g["envelope"]:
        for c in run["checkpoints"]:
            e = env[c["substep"]]
            for k in STATS + ("at_cap",):
                e[k] = max(e[k], abs(c[k] - gold[c["substep"]][k]))
    return gold, env


def cap_limit(env_at_cap, particles):
    """Bound on the distance of the particle count at the cap from the golden
    run's: ENV_FACTOR times the envelope's, plus 0.1% of the particles."""
    return ENV_FACTOR * env_at_cap + particles // 1000


def limit(ref, env_k, k):
    rel, floor = FLOOR[k]
    return ENV_FACTOR * env_k + rel * abs(ref) + floor


def check(name, sub, got, gold, env):
    """Statistics `got` at checkpoint `sub` against the bounds: returns
    (violations, distance of each statistic in units of its limit)."""
    bad, frac = [], {}
    for k in STATS:
        ref = gold[sub][k]
        lim = limit(ref, env[sub][k], k)
        d = abs(got[k] - ref)
        frac[k] = d / lim
        if d > lim:
            bad.append((name, sub, k, got[k], ref, env[sub][k]))
    return bad, frac
