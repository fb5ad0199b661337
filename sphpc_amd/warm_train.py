"""Trains the lambda warm-start network on the GPU from the solver's own samples.

    python -m sphpc_amd.warm_train --source dam_break --substeps 100 \
        --out sphpc_amd/data/warm_net_s0.008.bin      (the committed network)

Scenes (seeded): one fluid block per scene with sides drawn from U(0.2, 1.0) m
on the s = 0.008 m lattice (jitter 0.01 s), at a random position in the
6 x 2 x 2 m box with the torus + box obstacle and the physics of the 8M
double-dam-break workload.  Each scene runs `--substeps` substeps in the
tolerance mode the warm start is used in (density error 1e-3, cap 10).  Every
`--log-every` substeps, two consecutive substeps are solved with logging to
1e-4 (cap 100): the first gives the previous-multiplier feature, the second
is sampled.  Samples are (12 features, C_eff) rows, C_eff = -Lambda (|grad C|^2
+ eps) with Lambda the accumulated multiplier of the converged plain solve.

With `--source dam_break` the scenes are instead the double dam break itself at
the lattice scales `--scales` (fractions of the 8M lattice per axis, same box,
obstacle and physics), each run for `--substeps` substeps with the same
logging; the last scale is held out (validation loss and the step-scale fit).
The default scales 0.5 and 0.375 leave the 0.25 and 1.0 runs out of training
and out of the fit.

Training: Adam on random mini-batches with the project's HIP forward/backward
kernels, cosine learning-rate decay; the last scenes are held out for the
validation loss.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import sphpc_amd as sa  # noqa: E402

SPACING = float(np.float32(0.008))
BOX = (6.0, 2.0, 2.0)


def workload_params(iters=10, tol=1e-3):
    s = SPACING
    return sa.default_params(2.0 * s, time_step=1.0 / 300.0, g=9.81, k=0.5, c=0.01,
                             rest_density=1000.0, rest_mass=1000.0 * s ** 3,
                             lambda_epsilon=600.0, vorticity_epsilon=0.01, vorticity_model=1,
                             solve_iterations=iters, density_tolerance=tol)


def workload_obstacle(dims=(256, 256, 256)):
    """Torus (R 0.4, r 0.12, axis z) united with a box, sampled onto a grid."""
    wmin, wmax = (2.2, 0.2, 0.0), (3.8, 1.8, 1.6)
    shapes = [sa.torus((3.0, 0.75, 0.12), 0.4, 0.12, axis=2),
              sa.box_shape((3.0, 1.45, 0.25), (0.25, 0.2, 0.25))]
    sdf = sa.sdf_sample_shapes(shapes, dims, wmin, wmax)
    return dict(sdf=sdf, dims=dims, wmin=wmin, wmax=wmax, dx=(wmax[0] - wmin[0]) / dims[0])


def random_block(rng, floor=False):
    sides = rng.uniform(0.2, 1.0, 3)
    counts = [max(1, int(a / SPACING)) for a in sides]
    ext = [c * SPACING for c in counts]
    origin = [rng.uniform(0.0, BOX[i] - ext[i]) for i in range(3)]
    if floor:
        origin[2] = 0.0
    return sa.Block(*counts, *origin)


def generate(ctx, xb, tb, capacity, scenes, seed, substeps, log_every, every, log=print):
    """Runs the scenes and appends samples to the device buffers.  Returns the
    number of rows and the row offset at which each scene starts."""
    rows = 0
    starts = []
    plain = workload_params(10, 1e-3)
    logged = workload_params(100, 1e-4)
    for sc in range(scenes):
        rng = np.random.default_rng(seed * 1000 + sc)
        pos = sa.scene_lattice([random_block(rng)], SPACING, jitter=0.01, seed=seed * 1000 + sc)
        ctx.upload(pos)
        starts.append(rows)
        t0 = time.perf_counter()
        rows, iters = _run_logged(ctx, xb, tb, rows, capacity, substeps, log_every, every,
                                  plain, logged)
        log(f"[warm_train] scene {sc}: {len(pos)} particles, rows {rows}, "
            f"logged solve iterations {iters}, {time.perf_counter() - t0:.1f} s")
        if rows >= capacity:
            break
    return rows, starts


def _run_logged(ctx, xb, tb, rows, capacity, substeps, log_every, every, plain, logged):
    """Runs the uploaded scene; every `log_every` substeps the substep before and
    the substep itself are solved with logging, and the latter is sampled."""
    iters = []
    for k in range(1, substeps + 1):
        sample = k % log_every == 0
        pre = (k + 1) % log_every == 0
        if sample or pre:
            ctx.set_params(logged)
            ctx.set_sample_logging(True)
        ctx.step(1)
        if sample:
            iters.append(ctx.stats().iterations)
            rows += ctx.warm_samples(xb, tb, rows, capacity, every)
        if sample or pre:
            ctx.set_sample_logging(False)
            ctx.set_params(plain)
    return rows, iters


def dam_break_context(scale):
    """A context holding the double dam break at `scale` of the 8M lattice per
    axis, with its obstacle and the initial positions uploaded."""
    from sphpc_amd import workloads
    sc = workloads.double_dam_break(scale)
    ctx = sa.Context(workload_params(), sc["box"], len(sc["pos"]))
    ctx.set_obstacle(sc["sdf"], sc["dims"], sc["dx"], sc["wmin"], sc["wmax"])
    ctx.upload(sc["pos"])
    return ctx


def generate_dam_break(xb, tb, capacity, scales, substeps, log_every, every, log=print):
    """Samples of the double dam break at each lattice scale (one run each)."""
    rows, starts = 0, []
    plain = workload_params(10, 1e-3)
    logged = workload_params(100, 1e-4)
    for scale in scales:
        ctx = dam_break_context(scale)
        ctx.set_params(plain)
        starts.append(rows)
        t0 = time.perf_counter()
        rows, iters = _run_logged(ctx, xb, tb, rows, capacity, substeps, log_every, every,
                                  plain, logged)
        log(f"[warm_train] dam break at scale {scale}: {ctx.count()} particles, rows {rows}, "
            f"logged solve iterations {iters}, {time.perf_counter() - t0:.1f} s")
        ctx.close()
        if rows >= capacity:
            break
    return rows, starts


def train(net, xb, tb, count, steps, batch, lr, seed, log=print):
    """Adam with cosine decay from lr to lr/20 on rows [0, count)."""
    net.reset_optimizer()
    t0 = time.perf_counter()
    for k in range(steps):
        a = lr * (0.05 + 0.95 * 0.5 * (1.0 + math.cos(math.pi * k / max(steps - 1, 1))))
        want = (k % 500 == 0) or k == steps - 1
        loss = net.train_step(xb, tb, count, batch, a, seed * 1_000_003 + k, want_loss=want)
        if want:
            log(f"[warm_train] step {k}: batch loss {loss:.4g} (lr {a:.3g}, "
                f"{time.perf_counter() - t0:.1f} s)")
    net.synchronize()
    return time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=os.path.join(ROOT, "sphpc_amd", "data",
                                                  "warm_net_s0.008.bin"))
    ap.add_argument("--scenes", type=int, default=64)
    ap.add_argument("--holdout", type=int, default=8)
    ap.add_argument("--substeps", type=int, default=60)
    ap.add_argument("--log-every", type=int, default=10)
    ap.add_argument("--every", type=int, default=1, help="sample every k-th particle")
    ap.add_argument("--max-rows", type=int, default=60_000_000)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--summary", default=None, help="write a JSON summary here")
    ap.add_argument("--compare-substeps", type=int, default=20)
    ap.add_argument("--clip-quantile", type=float, default=0.995)
    ap.add_argument("--source", choices=("blocks", "dam_break"), default="blocks")
    ap.add_argument("--scales", default="0.5,0.375",
                    help="dam-break lattice scales (the last is held out)")
    args = ap.parse_args()

    xb = sa.DeviceBuffer(4 * sa.MLP_STRIDE * args.max_rows)
    tb = sa.DeviceBuffer(4 * args.max_rows)
    t0 = time.perf_counter()
    if args.source == "dam_break":
        scales = [float(v) for v in args.scales.split(",")]
        rows, starts = generate_dam_break(xb, tb, args.max_rows, scales, args.substeps,
                                          args.log_every, args.every)
        hold = starts[-1] if len(starts) > 1 else rows
        ctx = dam_break_context(scales[-1])
    else:
        ob = workload_obstacle()
        ctx = sa.Context(workload_params(), BOX, 125 ** 3)
        ctx.set_obstacle(ob["sdf"], ob["dims"], ob["dx"], ob["wmin"], ob["wmax"])
        rows, starts = generate(ctx, xb, tb, args.max_rows, args.scenes, args.seed,
                                args.substeps, args.log_every, args.every)
        hold = starts[-args.holdout] if args.holdout and len(starts) > args.holdout else rows
    t_gen = time.perf_counter() - t0
    print(f"[warm_train] {rows} rows ({hold} training) in {t_gen:.1f} s", flush=True)

    # Targets are heavy-tailed (particles pinned at walls accumulate large
    # multipliers over the 1e-4 solve): clip at a high quantile, and cap the
    # network output in the solver at the same bound.
    t_host = tb.to_array(np.float32, rows)
    rng = np.random.default_rng(args.seed)
    sample = t_host[:hold][rng.integers(0, hold, min(hold, 4_000_000))]
    clip = float(np.quantile(sample, args.clip_quantile))
    quant = {str(q): float(np.quantile(sample, q)) for q in (0.5, 0.9, 0.99, 0.999, 1.0)}
    print(f"[warm_train] target quantiles {quant}; clip at {clip:.4g}", flush=True)
    np.minimum(t_host, np.float32(clip), out=t_host)
    sa.check(sa.lib().sph_memcpy_htod(tb.ptr, t_host.ctypes.data, t_host.nbytes))
    net = sa.Mlp(seed=args.seed)
    net.fit_normalization(xb, tb, hold)
    net.set_output_clip(clip)
    t_train = train(net, xb, tb, hold, args.steps, args.batch, args.lr, args.seed)
    train_loss, _ = net.gradient(xb, tb, hold)
    val_loss = None
    if hold < rows:
        # validation rows: the held-out scenes, addressed from their first row
        xv = sa.DeviceBuffer(4 * sa.MLP_STRIDE * (rows - hold))
        sa.check(sa.lib().sph_device_copy(xv.ptr, _offset(xb, 4 * sa.MLP_STRIDE * hold),
                                          4 * sa.MLP_STRIDE * (rows - hold), None))
        tv = sa.DeviceBuffer.from_array(t_host[hold:])
        val_loss, _ = net.gradient(xv, tv, rows - hold)
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    net.save(args.out)
    summary = dict(rows=rows, training_rows=hold, scenes=len(starts), generate_s=t_gen,
                   target_quantiles=quant, clip=clip,
                   train_s=t_train, steps=args.steps, batch=args.batch,
                   train_loss=train_loss, validation_loss=val_loss, out=args.out)
    if args.compare_substeps:
        # Held-out scenes, mid-flight: fit the damping of the predicted step on
        # them (mean max-density error at the iteration cap), then compare
        # plain and warm-started solves at the workload's tolerance.
        states = []
        if args.source == "dam_break":
            # the held-out scale's run, at two points of its course
            ctx.set_params(workload_params(10, 1e-3))
            ctx.set_warm_start(None)
            ctx.step(30)
            states.append(ctx.download())
            ctx.step(30)
            states.append(ctx.download())
        for k in range(2 if args.source == "blocks" else 0):
            rng = np.random.default_rng(args.seed * 1000 + 990 + k)
            pos = sa.scene_lattice([random_block(rng, floor=True)], SPACING, jitter=0.01,
                                   seed=70 + k)
            ctx.upload(pos)
            ctx.set_params(workload_params(10, 1e-3))
            ctx.set_warm_start(None)
            ctx.step(30)
            states.append(ctx.download())
        fit = {}
        for scale in (0.125, 0.25, 0.5, 0.75, 1.0):
            net.set_step_scale(scale)
            errs = [compare(ctx, net, p, v, workload_params(10, 1e-3), 6,
                            log=lambda m: None)["warm"]["mean_final_error"] for p, v in states]
            fit[str(scale)] = float(np.mean(errs))
        best = min(fit, key=fit.get)
        net.set_step_scale(float(best))
        net.save(args.out)
        summary["step_scale_fit"] = fit
        summary["step_scale"] = float(best)
        summary["compare"] = [compare(ctx, net, p, v, workload_params(10, 1e-3),
                                      args.compare_substeps) for p, v in states]
    print(json.dumps(summary), flush=True)
    if args.summary:
        with open(args.summary, "w") as f:
            json.dump(summary, f)


def _offset(buf, nbytes):
    import ctypes
    return ctypes.c_void_p(buf.ptr.value + int(nbytes))


def _set_mode(ctx, mode, net, prev_scale):
    ctx.set_warm_start(net if mode == "warm" else None)
    ctx.set_warm_start_previous(prev_scale if mode == "prev" else 0.0)


def compare(ctx, net, pos, vel, params, substeps, log=print, prev_scale=None):
    """Plain vs warm-started tolerance solve from the same state: each mode
    runs one untimed substep (which gives the warm start its previous
    multiplier), then `substeps` substeps, each synchronised and timed.
    Reports iterations, the max density error at the end of each solve, the
    mean error after each number of iterations, and warm-start fallbacks.
    Modes: plain, warm (the network `net`, if given) and prev (the previous
    substep's multiplier times `prev_scale`, if given)."""
    out = {}
    modes = ["plain"] + (["warm"] if net is not None else []) + \
        (["prev"] if prev_scale else [])
    for mode in modes:
        ctx.upload(pos, vel)
        ctx.set_params(params)
        _set_mode(ctx, mode, net, prev_scale)
        ctx.step(1)
        ctx.synchronize()
        its, errs, conv, curves, falls = [], [], [], [], 0
        elapsed = 0.0
        for _ in range(substeps):
            t0 = time.perf_counter()
            ctx.step(1)
            ctx.synchronize()
            elapsed += time.perf_counter() - t0
            st = ctx.stats()
            its.append(st.iterations)
            errs.append(st.max_density_error)
            conv.append(st.converged)
            falls += st.warm_fallback
            curves.append(ctx.residuals())
        width = max(len(c) for c in curves)
        padded = np.array([np.pad(c, (0, width - len(c)), mode="edge") if len(c) else
                           np.zeros(width) for c in curves])
        out[mode] = dict(iters_per_substep=float(np.mean(its)),
                         max_density_error=float(np.max(errs)),
                         mean_final_error=float(np.mean(errs)),
                         converged_fraction=float(np.mean(conv)),
                         fallbacks=int(falls),
                         error_after_iterations=[float(v) for v in padded.mean(0)],
                         ms_per_substep=1e3 * elapsed / substeps,
                         particle_updates_per_s=ctx.count() * substeps / elapsed)
        log(f"[warm_train] {mode}: {out[mode]}")
    ctx.set_warm_start(None)
    ctx.set_warm_start_previous(0.0)
    return out


def equal_accuracy(ctx, net, pos, vel, params, substeps, cap=10, prev_scale=None):
    """Iterations each mode needs to reach the plain solver's error at the cap.
    From the same state, each mode runs `substeps` substeps (after one untimed
    one) with a fixed iteration count k = 1..cap; the target is the plain
    solver's mean max density error at k = cap, and a mode's answer is the
    smallest k whose mean error is within 0.1% of it, with that run's time."""
    modes = ["plain"] + (["warm"] if net is not None else []) + \
        (["prev"] if prev_scale else [])
    runs = {m: [] for m in modes}
    for mode in modes:
        for k in range(1, cap + 1):
            p = type(params).from_buffer_copy(params)
            p.solve_iterations = k
            p.density_tolerance = 0.0
            ctx.upload(pos, vel)
            ctx.set_params(p)
            _set_mode(ctx, mode, net, prev_scale)
            ctx.step(1)
            ctx.synchronize()
            errs, falls, t = [], 0, 0.0
            for _ in range(substeps):
                t0 = time.perf_counter()
                ctx.step(1)
                ctx.synchronize()
                t += time.perf_counter() - t0
                st = ctx.stats()
                errs.append(st.max_density_error)
                falls += st.warm_fallback
            runs[mode].append(dict(iterations=k, mean_error=float(np.mean(errs)),
                                   ms_per_substep=1e3 * t / substeps, fallbacks=int(falls)))
    ctx.set_warm_start(None)
    ctx.set_warm_start_previous(0.0)
    target = runs["plain"][-1]["mean_error"]
    out = {"target_error": target, "substeps": substeps, "curves": runs}
    for mode in modes:
        hit = [r for r in runs[mode] if r["mean_error"] <= target * 1.001]
        out[mode] = hit[0] if hit else None
    if out["plain"]:
        for mode in modes[1:]:
            out[mode + "_speedup"] = (out["plain"]["ms_per_substep"] /
                                      out[mode]["ms_per_substep"]) if out[mode] else None
    return out


if __name__ == "__main__":
    main()
