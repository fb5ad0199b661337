"""Pins the FP64 oracle against the reference's own C code.

oracle/_ref/libsphref.so is Simulation/src/particles.c + neighbors.c compiled
unmodified on the host (oracle/Makefile target `ref`).  Every stage the
reference implements in C is compared bit-for-bit (np.array_equal) on seeded,
jittered inputs that exercise the r > h branches, the r floor and the
60-neighbour cap."""
import ctypes as C

import numpy as np
import pytest

import orc

pytestmark = pytest.mark.skipif(not orc.ref_available(),
                                reason="oracle/_ref not built (needs the reference tree)")


def _scene(seed, jitter=0.45, dims=(11, 9, 8), s=0.02, box=(0.5, 0.4, 0.4), spread=1.0):
    pos = orc.lattice_block(*dims, s * spread, origin=(0.03, 0.02, 0.01), jitter=jitter,
                            seed=seed).astype(np.float64)
    p = orc.make_params(s, box, k=0.7, c=0.05, vort_eps=0.0037, g=9.81)
    xs, ys, zs = orc.soa(pos)
    for a, lim in ((xs, box[0]), (ys, box[1]), (zs, box[2])):
        np.clip(a, 0.0, lim * 0.999, out=a)
    rng = np.random.default_rng(seed + 100)
    v = rng.normal(0.0, 0.3, size=(len(xs), 3))
    vx, vy, vz = orc.soa(v)
    return p, xs, ys, zs, vx, vy, vz


@pytest.mark.parametrize("seed", [0, 1])
def test_hash_matches_reference(seed):
    p, xs, ys, zs, *_ = _scene(seed)
    n = len(xs)
    k1 = np.zeros(n, np.uint32); i1 = np.zeros(n, np.uint32)
    k2 = np.zeros(n, np.uint32); i2 = np.zeros(n, np.uint32)
    orc.oracle().orc_hash_particles(n, xs, ys, zs, C.byref(p), k1, i1)
    rp = orc.refh_params(p)
    orc.reference().refh_hash_particles(n, xs, ys, zs, C.byref(rp), k2, i2)
    assert np.array_equal(k1, k2) and np.array_equal(i1, i2)


@pytest.mark.parametrize("seed,spread", [(0, 1.0), (3, 0.55)])
def test_fill_neighbors_matches_reference(seed, spread):
    # spread 0.55 packs particles so many have more than 60 neighbours
    p, xs, ys, zs, *_ = _scene(seed, spread=spread)
    n = len(xs)
    L = orc.oracle()
    keys = np.zeros(n, np.uint32); ids = np.zeros(n, np.uint32)
    L.orc_hash_particles(n, xs, ys, zs, C.byref(p), keys, ids)
    L.orc_sort_by_key(n, keys, ids)
    cells = p.nx * p.ny * p.nz
    st = np.zeros(cells, np.uint32); en = np.zeros(cells, np.uint32)
    L.orc_find_lower_bounds(keys, n, cells, st)
    L.orc_find_upper_bounds(keys, n, cells, en)
    nb1 = np.zeros(n * orc.MAXN, np.uint32); c1 = np.zeros(n, np.int32)
    nb2 = np.zeros(n * orc.MAXN, np.uint32); c2 = np.zeros(n, np.int32)
    dropped = L.orc_fill_neighbors(n, xs, ys, zs, C.byref(p), st, en, ids, nb1, c1)
    rp = orc.refh_params(p)
    orc.reference().refh_fill_neighbors(n, xs, ys, zs, C.byref(rp), st, en, ids, nb2, c2, orc.MAXN)
    assert np.array_equal(c1, c2)
    assert np.array_equal(nb1, nb2)
    if spread < 1.0:
        assert c1.max() == orc.MAXN and dropped > 0


def _run_both(stage, p, n, arrays, nbr, cnt, substep=0):
    """Run one oracle stage (callable) and the reference stage on copies."""
    ref_arrays = {k: v.copy() for k, v in arrays.items()}
    rp = orc.refh_params(p)
    ra = orc.refh_arrays(**ref_arrays)
    orc.reference().refh_stage(orc.REFH[stage], n, C.byref(ra), C.byref(rp),
                               None if nbr is None else nbr.ctypes.data,
                               None if cnt is None else cnt.ctypes.data, orc.MAXN, substep)
    return ref_arrays


def _fields(n):
    return {k: np.zeros(n) for k in ("xs", "ys", "zs", "x", "y", "z", "vx", "vy", "vz",
                                      "dpx", "dpy", "dpz", "wx", "wy", "wz", "cx", "cy", "cz",
                                      "density", "lambda_")}


@pytest.mark.parametrize("seed,spread", [(5, 1.0), (6, 0.55)])
def test_stages_match_reference(seed, spread):
    p, xs, ys, zs, vx, vy, vz = _scene(seed, spread=spread)
    n = len(xs)
    L = orc.oracle()
    nbr, cnt, _ = orc.neighbors(p, xs, ys, zs)
    a = _fields(n)
    a["xs"][:], a["ys"][:], a["zs"][:] = xs, ys, zs
    rng = np.random.default_rng(seed)
    a["x"][:] = xs - rng.normal(0, 1e-3, n)
    a["y"][:] = ys - rng.normal(0, 1e-3, n)
    a["z"][:] = zs - rng.normal(0, 1e-3, n)
    a["vx"][:], a["vy"][:], a["vz"][:] = vx, vy, vz

    # gravity
    r = _run_both("GRAVITY", p, n, a, None, None)
    o = a["vz"].copy(); L.orc_apply_gravity(n, o, C.byref(p))
    assert np.array_equal(o, r["vz"])

    # predict (+ boundary clamp): push some particles out of the box first
    b = dict(a)
    b["vx"] = a["vx"] * 40.0
    b["vz"] = a["vz"] * 40.0
    r = _run_both("PREDICT", p, n, b, None, None)
    ox, oy, oz = np.zeros(n), np.zeros(n), np.zeros(n)
    L.orc_predict_positions(n, b["x"], b["y"], b["z"], b["vx"], b["vy"], b["vz"], ox, oy, oz,
                            C.byref(p))
    assert np.array_equal(ox, r["xs"]) and np.array_equal(oy, r["ys"]) and np.array_equal(oz, r["zs"])

    # density
    r = _run_both("DENSITY", p, n, a, nbr, cnt)
    rho = np.zeros(n); L.orc_compute_densities(n, xs, ys, zs, nbr, cnt, C.byref(p), rho)
    assert np.array_equal(rho, r["density"])
    a["density"][:] = rho

    # lambda
    r = _run_both("LAMBDA", p, n, a, nbr, cnt)
    lam = np.zeros(n); L.orc_compute_lambda(n, xs, ys, zs, rho, nbr, cnt, C.byref(p), lam)
    assert np.array_equal(lam, r["lambda_"])
    assert np.count_nonzero(lam) > n // 4
    a["lambda_"][:] = lam

    # delta p for several substep indices (stiffness schedule)
    for it in (0, 1, 3):
        r = _run_both("DPS", p, n, a, nbr, cnt, substep=it)
        d = [np.zeros(n) for _ in range(3)]
        L.orc_update_dps(n, xs, ys, zs, lam, nbr, cnt, C.byref(p), it, *d)
        assert np.array_equal(d[0], r["dpx"]) and np.array_equal(d[1], r["dpy"])
        assert np.array_equal(d[2], r["dpz"])

    # velocities
    r = _run_both("VELOCITIES", p, n, a, None, None)
    u = [np.zeros(n) for _ in range(3)]
    L.orc_update_velocities(n, xs, ys, zs, a["x"], a["y"], a["z"], *u, C.byref(p))
    for i, k in enumerate(("vx", "vy", "vz")):
        assert np.array_equal(u[i], r[k])

    # surface tension: colour pass then force pass (v in place)
    r = _run_both("SURFACE", p, n, a, nbr, cnt)
    col = [np.zeros(n) for _ in range(3)]
    L.orc_compute_color(n, xs, ys, zs, rho, nbr, cnt, C.byref(p), *col)
    for i, k in enumerate(("cx", "cy", "cz")):
        assert np.array_equal(col[i], r[k])
    v2 = [a["vx"].copy(), a["vy"].copy(), a["vz"].copy()]
    L.orc_apply_surface_force(n, xs, ys, zs, rho, *col, nbr, cnt, C.byref(p), *v2)
    for i, k in enumerate(("vx", "vy", "vz")):
        assert np.array_equal(v2[i], r[k])

    # viscosity (Jacobi reading of the reference, see ref_harness.c)
    ov = [np.zeros(n) for _ in range(3)]
    L.orc_apply_viscosity(n, xs, ys, zs, rho, a["vx"], a["vy"], a["vz"], nbr, cnt, C.byref(p), *ov)
    rv = [np.zeros(n) for _ in range(3)]
    rp = orc.refh_params(p)
    orc.reference().refh_viscosity_jacobi(n, xs, ys, zs, rho, a["vx"], a["vy"], a["vz"], nbr, cnt,
                                          orc.MAXN, C.byref(rp), *rv)
    for i in range(3):
        assert np.array_equal(ov[i], rv[i])
    assert not np.array_equal(ov[0], a["vx"])

    # vorticity
    r = _run_both("VORTICITY", p, n, a, nbr, cnt)
    w = [np.zeros(n) for _ in range(3)]
    L.orc_compute_vorticity(n, xs, ys, zs, rho, a["vx"], a["vy"], a["vz"], nbr, cnt, C.byref(p), *w)
    for i, k in enumerate(("wx", "wy", "wz")):
        assert np.array_equal(w[i], r[k])
    a["wx"][:], a["wy"][:], a["wz"][:] = w

    # confinement: the reference uses eps = 0.1*h (particles.c:289)
    pc = orc.make_params(0.02, (0.5, 0.4, 0.4), k=0.7, c=0.05)
    pc.vort_eps = 0.1 * pc.h
    r = _run_both("CONFINEMENT", pc, n, a, nbr, cnt)
    v3 = [a["vx"].copy(), a["vy"].copy(), a["vz"].copy()]
    L.orc_apply_vorticity_confinement(n, xs, ys, zs, *w, nbr, cnt, C.byref(pc), *v3)
    for i, k in enumerate(("vx", "vy", "vz")):
        assert np.array_equal(v3[i], r[k])

    # positions
    r = _run_both("POSITIONS", p, n, a, None, None)
    q = [np.zeros(n) for _ in range(3)]
    L.orc_update_positions(n, xs, ys, zs, *q)
    for i, k in enumerate(("x", "y", "z")):
        assert np.array_equal(q[i], r[k])


def test_boundary_clamp_matches_reference():
    p = orc.make_params(0.02, (0.5, 0.4, 0.4))
    rng = np.random.default_rng(9)
    n = 500
    a = _fields(n)
    a["xs"][:] = rng.uniform(-0.1, 0.6, n)
    a["ys"][:] = rng.uniform(-0.1, 0.5, n)
    a["zs"][:] = rng.uniform(-0.1, 0.5, n)
    a["xs"][:5] = [0.5, 0.0, -0.0, 0.4999999, 0.5 - 0.0001 * p.h]
    r = _run_both("BOUNDARY", p, n, a, None, None)
    L = orc.oracle()
    ox, oy, oz = a["xs"].copy(), a["ys"].copy(), a["zs"].copy()
    L.orc_predict_positions(n, a["xs"], a["ys"], a["zs"], np.zeros(n), np.zeros(n), np.zeros(n),
                            ox, oy, oz, C.byref(p))
    assert np.array_equal(ox, r["xs"]) and np.array_equal(oy, r["ys"]) and np.array_equal(oz, r["zs"])
