"""Stage-by-stage comparison of a GPU context with the FP64 oracle from the
GPU's own inputs: after every stage entry point, the oracle runs that stage
on the GPU's FP32 state promoted to double and the outputs are compared
(integer work bit-exact, floating point within the tolerance written next to
each check).  Used on whole workloads, where a full state is compared."""
import ctypes as C

import numpy as np

import orc
from gpu_common import d64, oracle_lists, record, rel_err

L = orc.oracle()


def soa(a):
    a = d64(a)
    return [np.ascontiguousarray(a[:, i]) for i in range(3)]


def ulp(a):
    """FP32 spacing at |a|, elementwise."""
    return np.spacing(np.abs(np.asarray(a, np.float32))).astype(np.float64)


def wall_slack(ref, box, h):
    """Extra allowance for coordinates at an upper wall of the box.  A
    coordinate within half an FP32 ulp of the wall rounds onto it in FP32 and
    is clamped to 1e-4 h below it (particles.c:841-867), where FP64 keeps it
    just under the wall: the two differ by up to 1e-4 h."""
    ref = np.asarray(ref, np.float64)
    hi = np.asarray(box, np.float64)[None, :]
    near = ref >= hi - 1e-4 * h - 4.0 * ulp(hi)
    return np.where(near, 1e-4 * h + 2.0 * ulp(hi), 0.0)


def increment_err(v_before, v_after_gpu, v_after_64):
    """Normwise check of a velocity increment stored as v + dv in FP32:
    returns (max abs error, bound 1e-5 max|dv| + 2 ulp(max|v|), normwise)."""
    d_gpu = d64(v_after_gpu) - d64(v_before)
    d_64 = d64(v_after_64) - d64(v_before)
    abs_err = float(np.max(np.abs(d_gpu - d_64)))
    bound = 1e-5 * float(np.max(np.abs(d_64))) + \
        2.0 * float(np.spacing(np.float32(np.max(np.abs(d64(v_after_64))))))
    return abs_err, bound, rel_err(d_gpu, d_64)


def check_search(ctx, sa, sc, x0, v0, tag):
    """Gravity, prediction and the neighbour search.  Returns (x*, nbr, cnt)
    of the GPU in particle-id order (x* as FP64 columns)."""
    n = len(x0)
    dropped0 = ctx.stats().neighbor_overflow  # a running total over the context's life
    ctx.apply_gravity()
    ctx.predict_positions()
    ctx.find_all_neighbors()
    x, y, z = soa(x0)
    vx, vy, vz = soa(v0)
    L.orc_apply_gravity(n, vz, C.byref(sc.o))
    xs64 = [np.zeros(n) for _ in range(3)]
    L.orc_predict_positions(n, x, y, z, vx, vy, vz, *xs64, C.byref(sc.o))
    xs_gpu = ctx.field(sa.FIELD_POSITION_STAR)
    # x + v dt in FP32: one rounding each for the product and the sum, so a
    # few ulp of x* and of the step v dt, and before it v + g dt, whose FP32
    # g dt (and sum) carries an ulp of g dt, times dt
    ref = np.stack(xs64, 1)
    e_pred = np.abs(d64(xs_gpu) - ref)
    step = np.abs(np.stack([vx, vy, vz], 1)) * sc.o.dt
    gdt = 2.0 * ulp(sc.o.g * sc.o.dt) * sc.o.dt
    bound = 2.0 * ulp(ref) + 2.0 * ulp(step) + gdt + wall_slack(ref, sc.box, sc.h)
    assert np.all(e_pred <= bound), float(np.max(e_pred - bound))
    pred_ulp = float(np.max(np.where(bound > 2.0 * ulp(ref), 0.0, e_pred / ulp(ref))))

    xs = soa(xs_gpu)
    ol = oracle_lists(sc.o, *xs)
    k, i = ctx.sorted_pairs()
    assert np.array_equal(k, ol["keys"]), "cell keys"
    assert np.array_equal(i, ol["ids"]), "(key, id) order"
    st, en = ctx.cell_bounds()
    assert np.array_equal(st, ol["start"]) and np.array_equal(en, ol["end"]), "cell tables"
    del k, i, st, en
    ids, cnt = ctx.neighbors()
    assert np.array_equal(cnt, ol["cnt"]), "neighbour counts"
    ids = ids.reshape(n, -1)
    nref = ol["nbr"].reshape(n, -1)
    mask = np.arange(orc.MAXN)[None, :] < cnt[:, None]
    assert np.array_equal(np.where(mask, ids, 0), np.where(mask, nref, 0)), "neighbour lists"
    del mask, nref
    dropped = ctx.stats().neighbor_overflow - dropped0
    assert dropped == ol["dropped"], (dropped, ol["dropped"])
    at_cap = int(np.sum(cnt >= orc.MAXN))
    record(tag + "_search", predict_max_ulp_off_walls=pred_ulp,
           predict_at_walls=int(np.sum(np.any(wall_slack(ref, sc.box, sc.h) > 0, 1))),
           at_cap=at_cap, dropped=int(ol["dropped"]), mean_nb=float(cnt.mean()),
           max_abs_x=float(np.max(np.abs(xs_gpu))))
    return xs, np.ascontiguousarray(ids.ravel()), np.ascontiguousarray(cnt), at_cap


def check_jacobi(ctx, sa, sc, obs, nbr, cnt, iters, tag):
    """The Jacobi stages of one substep: density (per particle 1e-5
    relative), lambda (normwise 1e-5; per-particle figures recorded), delta p
    from the GPU's lambda (normwise 1e-5) and the position update with clamp
    and collision (8 FP32 ulp of |x| plus 1e-5 of the largest displacement)."""
    n = len(cnt)
    worst = dict(rho=0.0, lam=0.0, dp=0.0, x=0.0)
    for it in range(iters):
        xs, ys, zs = soa(ctx.field(sa.FIELD_POSITION_STAR))
        ctx.compute_densities()
        rho_gpu = ctx.field(sa.FIELD_DENSITY)
        rho64 = np.zeros(n)
        L.orc_compute_densities(n, xs, ys, zs, nbr, cnt, C.byref(sc.o), rho64)
        e_rho = float(np.max(np.abs(d64(rho_gpu) - rho64) / rho64))
        assert e_rho <= 1e-5, (it, e_rho)

        ctx.compute_lambda()
        lam_gpu = ctx.field(sa.FIELD_LAMBDA)
        lam64 = np.zeros(n)
        L.orc_compute_lambda(n, xs, ys, zs, rho64, nbr, cnt, C.byref(sc.o), lam64)
        e_lam = rel_err(lam_gpu, lam64)
        assert e_lam <= 1e-5, (it, e_lam)
        big = np.abs(lam64) >= 1e-2 * np.max(np.abs(lam64))
        pp = np.abs(d64(lam_gpu) - lam64)[big] / np.abs(lam64[big])

        ctx.update_dps(it)
        dp_gpu = ctx.field(sa.FIELD_DP)
        d_iso = [np.zeros(n) for _ in range(3)]
        L.orc_update_dps(n, xs, ys, zs, d64(lam_gpu), nbr, cnt, C.byref(sc.o), it, *d_iso)
        e_dp = rel_err(dp_gpu, np.stack(d_iso, 1))
        assert e_dp <= 1e-5, (it, e_dp)

        ctx.update_position_stars()
        x2_gpu = d64(ctx.field(sa.FIELD_POSITION_STAR))
        x2 = [xs.copy(), ys.copy(), zs.copy()]
        dd = d64(dp_gpu)
        L.orc_update_position_stars(n, *x2, *[np.ascontiguousarray(dd[:, a]) for a in range(3)],
                                    C.byref(sc.o), C.byref(obs) if obs is not None else None)
        ref = np.stack(x2, 1)
        moved = float(np.max(np.abs(ref - np.stack([xs, ys, zs], 1))))
        e_x = np.abs(x2_gpu - ref)
        bound = 8.0 * ulp(ref) + 1e-5 * moved + wall_slack(ref, sc.box, sc.h)
        assert np.all(e_x <= bound), (it, float(np.max(e_x - bound)))
        record(tag + "_jacobi", iteration=it, rho_rel=e_rho, lambda_normwise=e_lam,
               lambda_per_particle_max=float(pp.max()) if len(pp) else 0.0,
               lambda_per_particle_median=float(np.median(pp)) if len(pp) else 0.0,
               dp_normwise=e_dp, x_max_ulp=float(np.max(e_x / ulp(ref))),
               max_displacement=moved)
        worst["rho"] = max(worst["rho"], e_rho)
        worst["lam"] = max(worst["lam"], e_lam)
        worst["dp"] = max(worst["dp"], e_dp)
    return worst


def check_velocity_passes(ctx, sa, sc, nbr, cnt, tag):
    """Velocity update, surface tension (colour and force), XSPH, vorticity
    and confinement from the GPU's own inputs, then the commit to x.  Returns
    the committed positions."""
    n = len(cnt)
    xs, ys, zs = soa(ctx.field(sa.FIELD_POSITION_STAR))
    xo = soa(ctx.field(sa.FIELD_POSITION_OLD))
    rho = d64(ctx.field(sa.FIELD_DENSITY))
    o = sc.o

    ctx.update_velocities()
    v1 = ctx.field(sa.FIELD_VELOCITY)
    o1 = [np.zeros(n) for _ in range(3)]
    L.orc_update_velocities(n, xs, ys, zs, *xo, *o1, C.byref(o))
    # (x* - x)/dt in FP32: x* - x exact for nearby floats, one rounding of the divide
    e_v = rel_err(v1, np.stack(o1, 1))
    assert e_v <= 1e-6, e_v
    out = dict(v=e_v)
    v_prev = v1
    if o.surface_tension_on:
        ctx.compute_surface_tension()
        col = ctx.field(sa.FIELD_COLOR)
        v2 = ctx.field(sa.FIELD_VELOCITY)
        c64 = [np.zeros(n) for _ in range(3)]
        L.orc_compute_color(n, xs, ys, zs, rho, nbr, cnt, C.byref(o), *c64)
        out["color"] = rel_err(col, np.stack(c64, 1))
        assert out["color"] <= 1e-5, out["color"]
        vv = soa(v_prev)
        L.orc_apply_surface_force(n, xs, ys, zs, rho, *soa(col), nbr, cnt, C.byref(o), *vv)
        a, b, out["surface"] = increment_err(v_prev, v2, np.stack(vv, 1))
        assert a <= b, ("surface", a, b)
        v_prev = v2
    if o.viscosity_on:
        ctx.apply_viscosity()
        v3 = ctx.field(sa.FIELD_VELOCITY)
        ov = [np.zeros(n) for _ in range(3)]
        L.orc_apply_viscosity(n, xs, ys, zs, rho, *soa(v_prev), nbr, cnt, C.byref(o), *ov)
        a, b, out["xsph"] = increment_err(v_prev, v3, np.stack(ov, 1))
        assert a <= b, ("xsph", a, b)
        v_prev = v3
    if o.vorticity_on:
        ctx.compute_vorticity()
        w = ctx.field(sa.FIELD_VORTICITY)
        w64 = [np.zeros(n) for _ in range(3)]
        L.orc_compute_vorticity(n, xs, ys, zs, rho, *soa(v_prev), nbr, cnt, C.byref(o), *w64)
        out["vorticity"] = rel_err(w, np.stack(w64, 1))
        assert out["vorticity"] <= 1e-5, out["vorticity"]
        ctx.apply_vorticity_confinement()
        v4 = ctx.field(sa.FIELD_VELOCITY)
        vc = soa(v_prev)
        L.orc_apply_vorticity_confinement(n, xs, ys, zs, *soa(w), nbr, cnt, C.byref(o), *vc)
        a, b, out["confinement"] = increment_err(v_prev, v4, np.stack(vc, 1))
        assert a <= b, ("confinement", a, b)
        v_prev = v4
    ctx.update_positions()
    xg, vg = ctx.download()
    assert np.array_equal(xg, np.stack([xs, ys, zs], 1).astype(np.float32))
    assert np.array_equal(vg, v_prev)
    record(tag + "_velocity", **{k + "_normwise": v for k, v in out.items()})
    return xg


def oracle_substep_from_xstar(sc, obs, xs, x0, v0, lists=None):
    """The oracle's substep (orc_substep, simulation.c:59-108) from predicted
    positions xs (three FP64 columns): lists from xs (or the given (nbr, cnt)),
    the Jacobi iterations, velocities ((x* - x0)/dt, so v0 does not enter),
    surface tension, XSPH, vorticity and confinement, the commit.  Returns
    (x, v, (nbr, cnt))."""
    o = sc.o
    n = len(x0)
    xs = [np.ascontiguousarray(a.copy()) for a in xs]
    if lists is None:
        ol = oracle_lists(o, *xs)
        lists = (ol["nbr"], ol["cnt"])
    nbr, cnt = lists
    rho, lam = np.zeros(n), np.zeros(n)
    d = [np.zeros(n) for _ in range(3)]
    for it in range(sc.iters):
        L.orc_compute_densities(n, *xs, nbr, cnt, C.byref(o), rho)
        L.orc_compute_lambda(n, *xs, rho, nbr, cnt, C.byref(o), lam)
        L.orc_update_dps(n, *xs, lam, nbr, cnt, C.byref(o), it, *d)
        L.orc_update_position_stars(n, *xs, *d, C.byref(o),
                                    C.byref(obs) if obs is not None else None)
    xo = soa(x0)
    v = [np.zeros(n) for _ in range(3)]
    L.orc_update_velocities(n, *xs, *xo, *v, C.byref(o))
    t = [np.zeros(n) for _ in range(3)]
    if o.surface_tension_on:
        L.orc_compute_color(n, *xs, rho, nbr, cnt, C.byref(o), *t)
        L.orc_apply_surface_force(n, *xs, rho, *t, nbr, cnt, C.byref(o), *v)
    if o.viscosity_on:
        u = [np.zeros(n) for _ in range(3)]
        L.orc_apply_viscosity(n, *xs, rho, *v, nbr, cnt, C.byref(o), *u)
        v = u
    if o.vorticity_on:
        L.orc_compute_vorticity(n, *xs, rho, *v, nbr, cnt, C.byref(o), *t)
        L.orc_apply_vorticity_confinement(n, *xs, *t, nbr, cnt, C.byref(o), *v)
    return np.stack(xs, 1), np.stack(v, 1), (nbr, cnt)


def substep_from_shared_state(run, sa, substep, tag, late):
    """One checked substep of a workload run (`run`: ctx, sc, obs, done): the
    run steps to `substep`; from its state there, the stage entry points are
    checked against the oracle stage by stage, then the fused substep against
    the oracle's rest of the substep from the GPU's predicted positions, per
    particle (or, with `late`, at most 1% of the particles outside the bound).
    The run continues from the fused substep."""
    import time
    ctx, sc, obs = run["ctx"], run["sc"], run["obs"]
    t0 = time.time()
    assert substep >= run["done"]
    if substep > run["done"]:
        ctx.step(substep - run["done"])
    run["done"] = substep
    x0, v0 = ctx.download()

    # stage by stage from the shared state
    ctx.upload(x0, v0)
    xs, nbr, cnt, at_cap = check_search(ctx, sa, sc, x0, v0, tag)
    worst = check_jacobi(ctx, sa, sc, obs, nbr, cnt, sc.iters, tag)
    check_velocity_passes(ctx, sa, sc, nbr, cnt, tag)

    # the fused substep from the same state; the run continues from it.  The
    # oracle takes the rest of the substep from the GPU's predicted positions
    # (their own error is checked above): list membership is decided on them,
    # and a pair within an ulp of r = h in FP32 or FP64 alone would otherwise
    # interact in one solve and not the other once the update moves it closer
    # (at substep 20 of the 8M double dam break, 3,652 particles' neighbour
    # counts differ between the two predictions, moving them by up to 6e-5 m).
    ctx.upload(x0, v0)
    ctx.step(1)
    run["done"] = substep + 1
    xg, vg = ctx.download()
    del nbr
    x64, v64, lists = oracle_substep_from_xstar(sc, obs, xs, x0, v0)
    e_x = np.abs(d64(xg) - x64)
    # x: the roundings of x* + delta p over the iterations plus the compounded
    # stage errors (1e-5 of the displacement).  Over several iterations every
    # axis of a particle carries the rounding of its largest coordinate (the
    # others enter through the pair directions), so the ulp is that of max |x_a|.
    disp = float(np.max(np.abs(x64 - np.stack(xs, 1))))
    big = ulp(np.max(np.abs(x64), 1, keepdims=True))
    bound = 8.0 * big + 1e-5 * disp + wall_slack(x64, sc.box, sc.h)
    beyond = np.any(e_x > bound, 1)
    extra = {}
    if late:
        # the chain's own response, by the oracle: the same substep from x*
        # moved by one float32 ulp per coordinate, up or down at random
        # (seeded), on the same lists (one direction everywhere would be
        # nearly a translation, which leaves pair separations unchanged)
        rng = np.random.default_rng(substep)
        nudged = [np.nextafter(a.astype(np.float32),
                               np.where(rng.random(len(a)) < 0.5, -np.inf, np.inf)
                               .astype(np.float32)).astype(np.float64) for a in xs]
        x64u, _, _ = oracle_substep_from_xstar(sc, obs, nudged, x0, v0, lists=lists)
        sens = np.any(np.abs(x64u - x64) > bound, 1)
        extra = dict(beyond_bound=int(beyond.sum()), oracle_nudge_beyond_bound=int(sens.sum()),
                     x_p99_abs=float(np.quantile(np.max(e_x, 1), 0.99)),
                     x_max_abs_all=float(np.max(e_x)))
        del x64u, nudged, sens
        # after the instability a one-ulp change already moves a share of the
        # particles beyond the bound (recorded): at most 1% may be
        assert beyond.mean() <= 0.01, (int(beyond.sum()), extra)
    else:
        assert not np.any(beyond), (int(np.argmax(np.max(e_x - bound, 1))),
                                    float(np.max(e_x - bound)))
    del lists
    within = ~beyond
    # v: the velocity passes normwise 1e-5, plus the position bound over dt
    # (v = (x* - x) / dt before those passes)
    vmax = float(np.max(np.abs(v64[within])))
    if late:
        # XSPH and vorticity carry the velocities of the particles outside the
        # bound into their neighbours': per particle, at most 1% outside
        vb = np.max(bound, 1, keepdims=True) / sc.o.dt + 1e-5 * vmax
        v_beyond = np.any(np.abs(d64(vg) - v64) > vb, 1)
        extra.update(v_beyond_bound=int(v_beyond.sum()))
        assert v_beyond.mean() <= 0.01, (int(v_beyond.sum()), extra)
        e_v = rel_err(vg[~v_beyond], v64[~v_beyond])
    else:
        e_v = rel_err(vg, v64)
        v_bound = 1e-5 + float(np.max(bound)) / sc.o.dt / vmax
        assert e_v <= v_bound, (e_v, v_bound)
        extra.update(v_bound=v_bound)
    record(tag + "_fused", x_max_ulp=float(np.max(e_x[within] / big[within])),
           x_max_abs=float(np.max(e_x[within])), v_normwise=e_v, **extra)
    record(tag, seconds=time.time() - t0, at_cap=at_cap, **{"worst_" + k: v
                                                            for k, v in worst.items()})
