"""Shared setup for GPU parity tests: one scene description turned into both the
GPU context parameters and the oracle parameters, so both sides see identical
constants."""
import ctypes as C
import json
import os

import numpy as np

import orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# measured GPU-vs-oracle errors are appended to the JSON-lines file this
# variable names (no log when unset)
ERR_LOG = os.environ.get("SPH_PARITY_LOG")


def record(name, **vals):
    """Append measured GPU-vs-oracle errors (stated per quantity)."""
    if not ERR_LOG:
        return
    os.makedirs(os.path.dirname(os.path.abspath(ERR_LOG)), exist_ok=True)
    with open(ERR_LOG, "a") as f:
        f.write(json.dumps({"test": name, **{k: float(v) for k, v in vals.items()}}) + "\n")


class Scene:
    def __init__(self, s, box, *, dt=1.0 / 300.0, g=9.81, k=0.0, c=0.0, mass_factor=1.0,
                 lambda_eps=0.01, vort_eps=None, iters=3, tol=0.0, surface=None,
                 viscosity=None, vorticity=None, vort_model=0):
        self.s = float(np.float32(s))
        self.h = 2.0 * self.s
        self.box = tuple(float(b) for b in box)
        self.o = orc.make_params(s, box, dt=dt, g=g, k=k, c=c, mass_factor=mass_factor,
                                 lambda_eps=lambda_eps, vort_eps=vort_eps,
                                 vort_model=vort_model)
        self.o.surface_tension_on = int(k != 0.0) if surface is None else int(surface)
        self.o.viscosity_on = int(c != 0.0) if viscosity is None else int(viscosity)
        self.o.vorticity_on = int(vort_eps is not None) if vorticity is None else int(vorticity)
        self.iters = iters
        self.tol = tol

    def gpu_params(self):
        import sphpc_amd as sa
        o = self.o
        return sa.default_params(
            self.h, time_step=o.dt, g=o.g, k=o.k, c=o.c, rest_density=o.rest_density,
            rest_mass=o.rest_mass, lambda_epsilon=o.lambda_eps, vorticity_epsilon=o.vort_eps,
            v_max=o.v_max, stiffness_k=o.k_stiff, density_tolerance=self.tol,
            solve_iterations=self.iters, surface_tension=o.surface_tension_on,
            viscosity=o.viscosity_on, vorticity=o.vorticity_on, vorticity_model=o.vort_model)

    def context(self, capacity):
        import sphpc_amd as sa
        return sa.Context(self.gpu_params(), self.box, capacity)


def block_scene(dims, s, origin, jitter=0.01, seed=0, vel_sigma=0.0, swirl=0.0):
    pos = orc.lattice_block(*dims, s, origin=origin, jitter=jitter, seed=seed)
    rng = np.random.default_rng(seed + 1)
    vel = rng.normal(0.0, vel_sigma, size=pos.shape) if vel_sigma else np.zeros_like(pos)
    if swirl:
        c = pos.mean(0)
        d = pos - c
        vel = vel + swirl * np.stack([-d[:, 1], d[:, 0], 0.2 * d[:, 0]], 1)
    return pos.astype(np.float32), vel.astype(np.float32)


def d64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def oracle_lists(o, xs, ys, zs):
    n = len(xs)
    L = orc.oracle()
    keys = np.zeros(n, np.uint32)
    ids = np.zeros(n, np.uint32)
    L.orc_hash_particles(n, xs, ys, zs, C.byref(o), keys, ids)
    L.orc_sort_by_key(n, keys, ids)
    cells = o.nx * o.ny * o.nz
    st = np.zeros(cells, np.uint32)
    en = np.zeros(cells, np.uint32)
    L.orc_find_lower_bounds(keys, n, cells, st)
    L.orc_find_upper_bounds(keys, n, cells, en)
    nbr = np.zeros(max(n, 1) * orc.MAXN, np.uint32)
    cnt = np.zeros(max(n, 1), np.int32)
    dropped = L.orc_fill_neighbors(n, xs, ys, zs, C.byref(o), st, en, ids, nbr, cnt)
    return dict(keys=keys, ids=ids, start=st, end=en, nbr=nbr, cnt=cnt, dropped=dropped)


def rel_err(a, b):
    """max |a-b| / max |b| (normwise relative error)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / den) if den > 0 else float(np.max(np.abs(a - b)))
