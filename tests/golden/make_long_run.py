"""Writes long_run.json: statistics of long FP64 oracle runs (the oracle's
stages are pinned bit-for-bit to the reference's own C by
tests/test_oracle_vs_ref.py; the substep composes them in the order of
Simulation/src/simulation.c:54-116).  Trajectories of a dam break diverge
chaotically between FP32 and FP64, so the GPU is compared with these by
statistics: maximum, 99th percentile and mean compression error of the last
density pass,
kinetic energy, mean neighbour count, particles at the 60-neighbour cap.

The envelope of a case is the same run from initial positions nudged by one
float32 ulp per coordinate (towards +inf, and towards -inf), the same run with
its positions and velocities rounded to float32 after every substep, and the
same run with every array a stage writes rounded to float32 after the stage
(orc_substep_f32_storage: x* after the prediction and after every Jacobi
update, density, lambda, delta p, velocities, colour and vorticity fields --
what an FP32 solver stores): how far the statistics of the FP64 algorithm
itself move under the smallest changes of its input and of its storage,
checkpoint by checkpoint.  The test's bounds are set from it.

Run from the repository root:
  python3 tests/golden/make_long_run.py [case ...]             golden runs
  python3 tests/golden/make_long_run.py --envelope [case ...]  envelope runs
  python3 tests/golden/make_long_run.py --fp32-state case ...   its float32-state run only
  python3 tests/golden/make_long_run.py --f32-storage case ...  its float32-storage run only
  python3 tests/golden/make_long_run.py --out F --envelope c   write to F instead
  python3 tests/golden/make_long_run.py --merge F              copy F's entries in"""
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import orc  # noqa: E402

CASES = {
    # plain PBF dam break: one 32^3 block in the corner of a 2 x 1 x 1 m box
    "plain_dam_break_32k": dict(
        lattice=dict(dims=(32, 32, 32), s=0.02, origin=(0.0, 0.0, 0.0), jitter=0.01, seed=2024),
        box=(2.0, 1.0, 1.0), iters=3, substeps=500, every=50,
        physics=dict(lambda_eps=600.0)),
    # the 8M double dam break at a quarter of its lattice per axis, all terms and the
    # torus + box obstacle
    "double_dam_break_131k": dict(
        workload="double_dam_break", scale=0.25, iters=4, substeps=200, every=20,
        physics=dict(k=0.5, c=0.01, lambda_eps=600.0, vort_eps=0.01, vort_model=1,
                     surface=True, viscosity=True, vorticity=True)),
    # one 128x128x64 block (1M particles) at s = 0.01 m in a 4 x 2 x 2 m box with a
    # sphere of radius 0.3 m at the centre, all terms
    "sphere_block_1m": dict(
        workload="sphere_block", scale=1.0, iters=4, substeps=150, every=15,
        physics=dict(k=0.5, c=0.01, lambda_eps=600.0, vort_eps=0.01, vort_model=1,
                     surface=True, viscosity=True, vorticity=True)),
    # the same block over the full 1000 substeps of its configuration: it falls,
    # splits on the sphere and spreads over the floor (about 35 minutes of CPU on
    # 8 threads)
    "sphere_block_1m_1000": dict(
        workload="sphere_block", scale=1.0, iters=4, substeps=1000, every=100,
        physics=dict(k=0.5, c=0.01, lambda_eps=600.0, vort_eps=0.01, vort_model=1,
                     surface=True, viscosity=True, vorticity=True)),
    # the full 8M double dam break: after about 80 substeps the 1 m water columns
    # compress past what 4 Jacobi iterations hold, particles pile past the
    # 60-neighbour cap and the maximum error grows to ~10 (the statistics of the
    # run, this included, are what the GPU must reproduce); about 30 minutes of
    # CPU on 8 threads
    "double_dam_break_8m": dict(
        workload="double_dam_break", scale=1.0, iters=4, substeps=110, every=10,
        physics=dict(k=0.5, c=0.01, lambda_eps=600.0, vort_eps=0.01, vort_model=1,
                     surface=True, viscosity=True, vorticity=True)),
}


def scene(case):
    """Initial FP32 positions, box, obstacle (or None) of a case."""
    if "lattice" in case:
        lt = case["lattice"]
        pos = orc.lattice_block(*lt["dims"], lt["s"], origin=lt["origin"], jitter=lt["jitter"],
                                seed=lt["seed"])
        return lt["s"], pos, case["box"], None
    from sphpc_amd import workloads
    sc = getattr(workloads, case["workload"])(case["scale"])
    return sc["s"], sc["pos"], sc["box"], sc


def statistics(rho, v, cnt, s, rest=1000.0):
    e = np.maximum(rho - rest, 0.0) / rest
    ke = 0.5 * rest * s ** 3 * float(np.sum(v.astype(np.float64) ** 2))
    return dict(max_err=float(e.max()), p99_err=float(np.quantile(e, 0.99)),
                mean_err=float(e.mean()), ke=ke, mean_nb=float(cnt.mean()),
                at_cap=int(np.sum(cnt >= orc.MAXN)))


def run(name, case, nudge=0, fp32_state=False, f32_storage=False):
    s, pos, box, sc = scene(case)
    if nudge:
        pos = np.nextafter(pos.astype(np.float32), np.float32(np.inf if nudge > 0 else -np.inf))
    L = orc.oracle()
    p = orc.make_params(s, box, dt=1.0 / 300.0, g=9.81, **case["physics"])
    ob = orc.obstacle(sc["sdf"], sc["dims"], sc["dx"], sc["wmin"], sc["wmax"]) if sc else None
    n = len(pos)
    cols = [np.ascontiguousarray(pos[:, i], dtype=np.float64) for i in range(3)]
    a = dict(x=cols[0], y=cols[1], z=cols[2], vx=np.zeros(n), vy=np.zeros(n), vz=np.zeros(n),
             xs=np.zeros(n), ys=np.zeros(n), zs=np.zeros(n), density=np.zeros(n),
             lambda_=np.zeros(n), nbr=np.zeros(n * orc.MAXN, np.uint32), cnt=np.zeros(n, np.int32))
    st = orc.OrcState(**{k: v.ctypes.data for k, v in a.items()})
    out = []
    t0 = time.time()
    substep = L.orc_substep_f32_storage if f32_storage else L.orc_substep
    for i in range(case["substeps"]):
        substep(n, C.byref(st), C.byref(p), C.byref(ob) if ob else None, case["iters"], 0.0, None)
        if fp32_state:
            for k in ("x", "y", "z", "vx", "vy", "vz"):
                a[k][:] = a[k].astype(np.float32)
        if (i + 1) % case["every"] == 0:
            v = np.stack([a["vx"], a["vy"], a["vz"]], 1)
            out.append(dict(substep=i + 1, **statistics(a["density"], v, a["cnt"], s)))
            kind = "fp32_state" if fp32_state else "f32_storage" if f32_storage else nudge
            print(name, kind, out[-1], f"{time.time() - t0:.0f}s", flush=True)
    return dict(particles=n, **{k: v for k, v in case.items()}, checkpoints=out)


def _member(e):
    for kind in ("fp32_state", "f32_storage"):
        if e.get(kind):
            return (kind, True)
    return ("nudge", e.get("nudge"))


def _save(path, update):
    """Read-modify-write of the JSON file: update maps case -> fields.  Envelope
    runs are merged member by member (a new run replaces the one of its kind)."""
    res = json.load(open(path)) if os.path.exists(path) else {}
    for name, fields in update.items():
        entry = res.setdefault(name, {})
        for key, val in fields.items():
            if key == "envelope":
                have = {_member(e): e for e in entry.get("envelope", [])}
                have.update({_member(e): e for e in val})
                entry["envelope"] = list(have.values())
            else:
                entry[key] = val
    with open(path, "w") as f:
        json.dump(res, f, indent=1)


if __name__ == "__main__":
    args = sys.argv[1:]
    path = os.path.join(HERE, "long_run.json")
    if "--out" in args:
        k = args.index("--out")
        path = args[k + 1]
        del args[k:k + 2]
    if "--merge" in args:
        src = json.load(open(args[args.index("--merge") + 1]))
        _save(path, src)
        sys.exit(0)
    envelope = "--envelope" in args
    for flag, kind in (("--fp32-state", "fp32_state"), ("--f32-storage", "f32_storage")):
        if flag in args:
            # only this member of the envelope, merged into it
            for name in [a for a in args if not a.startswith("--")]:
                env = [{kind: True,
                        "checkpoints": run(name, CASES[name], **{kind: True})["checkpoints"]}]
                _save(path, {name: dict(envelope=env)})
            sys.exit(0)
    names = [a for a in args if not a.startswith("--")] or list(CASES)
    for name in names:
        if envelope:
            env = [dict(nudge=d, checkpoints=run(name, CASES[name], d)["checkpoints"])
                   for d in (1, -1)]
            for kind in ("fp32_state", "f32_storage"):
                env.append({kind: True,
                            "checkpoints": run(name, CASES[name], **{kind: True})["checkpoints"]})
            _save(path, {name: dict(envelope=env)})
        else:
            _save(path, {name: run(name, CASES[name])})
