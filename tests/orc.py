"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and to the reference's
own C physics built on the host (oracle/_ref/libsphref.so), plus small numpy
scene builders.  Test infrastructure only."""
import ctypes as C
import math
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libsphref.so")

MAXN = 60

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


class OrcParams(C.Structure):
    _fields_ = [
        ("h", C.c_double), ("W_norm", C.c_double), ("DelW_poly_norm", C.c_double),
        ("DelW_norm", C.c_double), ("g", C.c_double), ("dt", C.c_double),
        ("k", C.c_double), ("c", C.c_double), ("rest_density", C.c_double),
        ("rest_mass", C.c_double), ("lambda_eps", C.c_double), ("vort_eps", C.c_double),
        ("v_max", C.c_double), ("k_stiff", C.c_double),
        ("box_min", C.c_double * 3), ("box_max", C.c_double * 3),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("max_neighbors", C.c_int),
        ("vort_model", C.c_int), ("surface_tension_on", C.c_int), ("viscosity_on", C.c_int),
        ("vorticity_on", C.c_int),
    ]


class OrcObstacle(C.Structure):
    _fields_ = [("dim", C.c_int * 3), ("spacing", C.c_double), ("wmin", C.c_double * 3),
                ("wmax", C.c_double * 3), ("sdf", C.c_void_p)]


class OrcState(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in
                ("x", "y", "z", "vx", "vy", "vz", "xs", "ys", "zs", "density", "lambda_",
                 "nbr", "cnt")]


class RefhParams(C.Structure):
    _fields_ = [
        ("h", C.c_double), ("W_norm", C.c_double), ("DelW_poly_norm", C.c_double),
        ("DelW_norm", C.c_double), ("g", C.c_double), ("dt", C.c_double),
        ("k", C.c_double), ("c", C.c_double), ("rest_density", C.c_double),
        ("rest_mass", C.c_double), ("box_min", C.c_double * 3), ("box_max", C.c_double * 3),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
    ]


class RefhArrays(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in
                ("xs", "ys", "zs", "x", "y", "z", "vx", "vy", "vz", "dpx", "dpy", "dpz",
                 "wx", "wy", "wz", "cx", "cy", "cz", "density", "lambda_")]


REFH = dict(GRAVITY=0, PREDICT=1, DENSITY=2, LAMBDA=3, DPS=4, VELOCITIES=5, SURFACE=6,
            VORTICITY=7, CONFINEMENT=8, POSITIONS=9, BOUNDARY=10)

_orc = None
_ref = None


def oracle():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle oracle`")
        L = C.CDLL(ORACLE_SO)
        P = C.POINTER(OrcParams)
        L.orc_params_derive.argtypes = [P]
        for f in ("orc_W", "orc_DelWPoly", "orc_DelW", "orc_C"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [C.c_double] * 3
        L.orc_hash_val.restype = C.c_uint32
        L.orc_hash_val.argtypes = [P, C.c_double, C.c_double, C.c_double]
        L.orc_hash_particles.argtypes = [C.c_int, _dp, _dp, _dp, P, _up, _up]
        L.orc_sort_by_key.argtypes = [C.c_int, _up, _up]
        L.orc_find_lower_bounds.argtypes = [_up, C.c_int, C.c_int, _up]
        L.orc_find_upper_bounds.argtypes = [_up, C.c_int, C.c_int, _up]
        L.orc_fill_neighbors.restype = C.c_long
        L.orc_fill_neighbors.argtypes = [C.c_int, _dp, _dp, _dp, P, _up, _up, _up, _up, _ip]
        L.orc_find_all_neighbors.restype = C.c_long
        L.orc_find_all_neighbors.argtypes = [C.c_int, _dp, _dp, _dp, P, _up, _ip]
        L.orc_apply_gravity.argtypes = [C.c_int, _dp, P]
        L.orc_predict_positions.argtypes = [C.c_int] + [_dp] * 9 + [P]
        L.orc_compute_densities.argtypes = [C.c_int, _dp, _dp, _dp, _up, _ip, P, _dp]
        L.orc_compute_lambda.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _up, _ip, P, _dp]
        L.orc_stiffness.restype = C.c_double
        L.orc_stiffness.argtypes = [P, C.c_int]
        L.orc_update_dps.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _up, _ip, P, C.c_int,
                                     _dp, _dp, _dp]
        L.orc_collide.argtypes = [C.c_int, _dp, _dp, _dp, C.POINTER(OrcObstacle)]
        L.orc_update_position_stars.argtypes = [C.c_int] + [_dp] * 6 + [P, C.POINTER(OrcObstacle)]
        L.orc_update_velocities.argtypes = [C.c_int] + [_dp] * 9 + [P]
        L.orc_compute_color.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _up, _ip, P, _dp, _dp, _dp]
        L.orc_apply_surface_force.argtypes = [C.c_int] + [_dp] * 7 + [_up, _ip, P, _dp, _dp, _dp]
        L.orc_apply_viscosity.argtypes = [C.c_int] + [_dp] * 7 + [_up, _ip, P, _dp, _dp, _dp]
        L.orc_compute_vorticity.argtypes = [C.c_int] + [_dp] * 7 + [_up, _ip, P, _dp, _dp, _dp]
        L.orc_apply_vorticity_confinement.argtypes = [C.c_int] + [_dp] * 6 + [_up, _ip, P,
                                                                              _dp, _dp, _dp]
        L.orc_update_positions.argtypes = [C.c_int] + [_dp] * 6
        L.orc_max_density_error.restype = C.c_double
        L.orc_max_density_error.argtypes = [C.c_int, _dp, C.c_double]
        L.orc_set_threads.restype = None
        L.orc_set_threads.argtypes = [C.c_int]
        for f in ("orc_substep", "orc_substep_f32_storage"):
            getattr(L, f).restype = C.c_int
            getattr(L, f).argtypes = [C.c_int, C.POINTER(OrcState), P, C.POINTER(OrcObstacle),
                                      C.c_int, C.c_double, C.POINTER(C.c_double)]
        _orc = L
    return _orc


def ref_available():
    return os.path.exists(REF_SO)


def reference():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_SO)
        P = C.POINTER(RefhParams)
        L.refh_max_neighbors.restype = C.c_int
        L.refh_hash_particles.argtypes = [C.c_int, _dp, _dp, _dp, P, _up, _up]
        L.refh_fill_neighbors.argtypes = [C.c_int, _dp, _dp, _dp, P, _up, _up, _up, _up, _ip,
                                          C.c_int]
        L.refh_stage.argtypes = [C.c_int, C.c_int, C.POINTER(RefhArrays), P, C.c_void_p,
                                 C.c_void_p, C.c_int, C.c_int]
        L.refh_viscosity_jacobi.argtypes = [C.c_int] + [_dp] * 7 + [_up, _ip, C.c_int, P,
                                                                     _dp, _dp, _dp]
        _ref = L
    return _ref


# --------------------------------------------------------------------------
# Parameters
# --------------------------------------------------------------------------
def make_params(s, box, *, dt=1.0 / 300.0, g=9.81, k=0.0, c=0.0, rest_density=1000.0,
                mass=None, mass_factor=1.0, lambda_eps=0.01, vort_eps=None, v_max=100.0,
                k_stiff=0.95, vort_model=0, surface=None, viscosity=None, vorticity=None):
    """Parameters the way SetParameters derives them: h = 2*float(s)
    (simulation.c:147,152); rest mass rho0*s^3*mass_factor
    (particles.c:1035-1037 with V/N = s^3)."""
    p = OrcParams()
    sf = float(np.float32(s))
    p.h = 2.0 * sf
    p.g = g
    p.dt = dt
    p.k = k
    p.c = c
    p.rest_density = rest_density
    p.rest_mass = mass if mass is not None else rest_density * sf ** 3 * mass_factor
    p.lambda_eps = lambda_eps
    p.vort_eps = 0.1 * p.h if vort_eps is None else vort_eps
    p.v_max = v_max
    p.k_stiff = k_stiff
    for i in range(3):
        p.box_min[i] = 0.0
        p.box_max[i] = float(box[i])
    p.max_neighbors = MAXN
    p.vort_model = vort_model
    p.surface_tension_on = int(k != 0.0) if surface is None else int(surface)
    p.viscosity_on = int(c != 0.0) if viscosity is None else int(viscosity)
    p.vorticity_on = int(vort_eps not in (None, 0.0)) if vorticity is None else int(vorticity)
    oracle().orc_params_derive(C.byref(p))
    return p


def refh_params(p):
    r = RefhParams()
    for f in ("h", "W_norm", "DelW_poly_norm", "DelW_norm", "g", "dt", "k", "c",
              "rest_density", "rest_mass", "nx", "ny", "nz"):
        setattr(r, f, getattr(p, f))
    for i in range(3):
        r.box_min[i] = p.box_min[i]
        r.box_max[i] = p.box_max[i]
    return r


# --------------------------------------------------------------------------
# Scenes
# --------------------------------------------------------------------------
def lattice_block(nx, ny, nz, s, origin=(0.0, 0.0, 0.0), jitter=0.0, seed=0):
    """Lattice like ConstructFluidVolume (geometry.c:42-55): x fastest,
    cell-centred (+s/2).  Optional seeded U(-jitter, jitter)*s per axis.
    Returns float32 (n,3)."""
    sf = float(np.float32(s))
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pos = np.stack([origin[0] + ix.ravel() * sf + sf / 2.0,
                    origin[1] + iy.ravel() * sf + sf / 2.0,
                    origin[2] + iz.ravel() * sf + sf / 2.0], axis=1)
    if jitter:
        rng = np.random.default_rng(seed)
        pos = pos + rng.uniform(-jitter, jitter, size=pos.shape) * sf
    return pos.astype(np.float32)


def soa(a):
    a = np.asarray(a, dtype=np.float64)
    return [np.ascontiguousarray(a[:, i]) for i in range(3)]


def neighbors(p, xs, ys, zs):
    n = len(xs)
    nbr = np.zeros(max(n, 1) * MAXN, dtype=np.uint32)
    cnt = np.zeros(max(n, 1), dtype=np.int32)
    dropped = oracle().orc_find_all_neighbors(n, xs, ys, zs, C.byref(p), nbr, cnt)
    return nbr, cnt, dropped


def neighbor_sets(nbr, cnt):
    return [list(nbr[i * MAXN:i * MAXN + cnt[i]]) for i in range(len(cnt))]


def sphere_sdf(dim, wmin, wmax, center, radius):
    """Analytic sphere SDF sampled at texel centres, x fastest (SDFGen layout)."""
    dx = (wmax[0] - wmin[0]) / dim[0]
    ax = [wmin[i] + (np.arange(dim[i]) + 0.5) * ((wmax[i] - wmin[i]) / dim[i]) for i in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    d = np.sqrt((X - center[0]) ** 2 + (Y - center[1]) ** 2 + (Z - center[2]) ** 2) - radius
    return np.ascontiguousarray(d.astype(np.float32).ravel()), dx


def obstacle(sdf, dim, spacing, wmin, wmax):
    o = OrcObstacle()
    for i in range(3):
        o.dim[i] = dim[i]
        o.wmin[i] = wmin[i]
        o.wmax[i] = wmax[i]
    o.spacing = spacing
    o.sdf = sdf.ctypes.data
    o._keep = sdf
    return o


def ptr(a):
    return None if a is None else a.ctypes.data


def refh_arrays(**kw):
    a = RefhArrays()
    for k, v in kw.items():
        setattr(a, k, ptr(v))
    return a
