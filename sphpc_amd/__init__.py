"""Python view of libsphpc_amd's C ABI (include/sphpc_amd.h).

The product is the HIP library; this module only loads it with ctypes and
wraps the context so tests and bench.py can drive it.  There is no fallback:
if the library is missing or cannot be loaded, importing the wrapper's entry
points raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsphpc_amd.so")
MAX_NEIGHBORS = 60

FIELD_POSITION_STAR = 0
FIELD_DENSITY = 1
FIELD_LAMBDA = 2
FIELD_DP = 3
FIELD_VELOCITY = 4
FIELD_COLOR = 5
FIELD_VORTICITY = 6
FIELD_POSITION_OLD = 7
FIELD_LAMBDA_SUM = 8
_FIELD_COMPONENTS = {0: 3, 1: 1, 2: 1, 3: 3, 4: 3, 5: 3, 6: 3, 7: 3, 8: 1}


class SphError(RuntimeError):
    pass


class Params(C.Structure):
    """struct sph_params."""
    _fields_ = [
        ("smoothing_radius", C.c_double), ("time_step", C.c_double), ("g", C.c_double),
        ("k", C.c_double), ("c", C.c_double), ("rest_density", C.c_double),
        ("rest_mass", C.c_double), ("lambda_epsilon", C.c_double),
        ("vorticity_epsilon", C.c_double), ("v_max", C.c_double), ("stiffness_k", C.c_double),
        ("density_tolerance", C.c_double), ("solve_iterations", C.c_int),
        ("surface_tension", C.c_int), ("viscosity", C.c_int), ("vorticity", C.c_int),
        ("vorticity_model", C.c_int),
    ]


class AABB(C.Structure):
    _fields_ = [("min_x", C.c_double), ("max_x", C.c_double), ("min_y", C.c_double),
                ("max_y", C.c_double), ("min_z", C.c_double), ("max_z", C.c_double)]


class ObstacleDesc(C.Structure):
    _fields_ = [("dim_x", C.c_int), ("dim_y", C.c_int), ("dim_z", C.c_int),
                ("spacing", C.c_double), ("world_bounds", AABB), ("sdf", C.c_void_p)]


class Scene(C.Structure):
    _fields_ = [("particles", C.c_int), ("size_x", C.c_double), ("size_y", C.c_double),
                ("size_z", C.c_double), ("obstacle", ObstacleDesc)]


class Tuning(C.Structure):
    _fields_ = [("split_collide", C.c_int), ("wide_lists", C.c_int), ("block", C.c_int * 32),
                ("depth", C.c_int * 32), ("ms_per_substep_default", C.c_double),
                ("ms_per_substep_tuned", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("particles", C.c_int), ("iterations", C.c_int), ("converged", C.c_int),
                ("max_density_error", C.c_float), ("neighbor_overflow", C.c_ulonglong),
                ("cell_clamps", C.c_ulonglong), ("substeps", C.c_longlong),
                ("neighbor_pairs", C.c_ulonglong), ("warm_fallback", C.c_int)]


_lib = None


def lib():
    """Load libsphpc_amd.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SphError(f"{LIB_PATH} is not built: run __graft_entry__.build() or "
                       "`make -C sphpc_amd`")
    L = C.CDLL(LIB_PATH)
    vp, ip, dp_ = C.c_void_p, C.c_int, C.c_double
    fp = C.POINTER(C.c_float)
    up = C.POINTER(C.c_uint32)
    L.sph_last_error.restype = C.c_char_p
    L.sph_default_params.argtypes = [C.POINTER(Params), dp_]
    L.sph_default_params.restype = None
    L.sph_create.argtypes = [C.POINTER(Params), C.POINTER(AABB), ip, C.POINTER(vp)]
    for name in ("sph_destroy", "sph_apply_gravity", "sph_predict_positions",
                 "sph_find_all_neighbors", "sph_compute_densities", "sph_compute_lambda",
                 "sph_update_position_stars", "sph_update_velocities",
                 "sph_compute_surface_tension", "sph_apply_viscosity", "sph_compute_vorticity",
                 "sph_apply_vorticity_confinement", "sph_update_positions", "sph_synchronize",
                 "sph_profile_reset"):
        getattr(L, name).argtypes = [vp]
    L.sph_set_params.argtypes = [vp, C.POINTER(Params)]
    L.sph_get_params.argtypes = [vp, C.POINTER(Params)]
    L.sph_grid_dims.argtypes = [vp, C.POINTER(ip), C.POINTER(ip), C.POINTER(ip)]
    L.sph_set_obstacle.argtypes = [vp, C.POINTER(ObstacleDesc)]
    L.sph_upload_state.argtypes = [vp, ip] + [fp] * 6
    L.sph_download_state.argtypes = [vp] + [fp] * 6
    L.sph_particle_count.argtypes = [vp, C.POINTER(ip)]
    L.sph_update_dps.argtypes = [vp, ip]
    L.sph_step.argtypes = [vp, ip]
    L.sph_get_stats.argtypes = [vp, C.POINTER(Stats)]
    L.sph_compute_energy.argtypes = [vp, C.POINTER(dp_)]
    L.sph_download_field.argtypes = [vp, ip, fp]
    L.sph_download_neighbors.argtypes = [vp, up, C.POINTER(ip)]
    L.sph_download_sorted.argtypes = [vp, up, up]
    L.sph_download_cell_bounds.argtypes = [vp, up, up]
    L.sph_sort_by_key.argtypes = [vp, vp, ip, vp]
    L.sph_sort_by_key_bits.argtypes = [vp, vp, ip, ip, vp]
    L.sph_sort_by_key_bits_noalloc.argtypes = [vp, vp, ip, ip, vp]
    L.sph_find_lower_bounds.argtypes = [vp, ip, ip, vp, vp]
    L.sph_find_upper_bounds.argtypes = [vp, ip, ip, vp, vp]
    L.sph_sdf_create.argtypes = [C.POINTER(ObstacleDesc), C.POINTER(vp)]
    L.sph_sdf_destroy.argtypes = [vp]
    L.sph_calculate_particle_collisions.argtypes = [vp, vp, vp, ip, vp, vp]
    L.sph_calculate_particle_collisions_f64.argtypes = [vp, vp, vp, ip, vp, vp]
    L.sph_device_malloc.argtypes = [C.POINTER(vp), C.c_size_t]
    L.sph_device_free.argtypes = [vp]
    L.sph_device_copy.argtypes = [vp, vp, C.c_size_t, vp]
    L.sph_memcpy_htod.argtypes = [vp, vp, C.c_size_t]
    L.sph_memcpy_dtoh.argtypes = [vp, vp, C.c_size_t]
    L.sph_get_device_count.argtypes = [C.POINTER(ip)]
    L.sph_profile_enable.argtypes = [vp, ip]
    L.sph_profile_read.argtypes = [vp, ip, C.POINTER(dp_), C.POINTER(C.c_longlong)]
    L.sph_profile_kernel_name.restype = C.c_char_p
    L.sph_profile_kernel_name.argtypes = [ip]
    L.sph_profile_kernel_count.argtypes = []
    sz, ull = C.c_size_t, C.c_ulonglong
    L.sph_mlp_create.argtypes = [C.POINTER(vp)]
    L.sph_mlp_destroy.argtypes = [vp]
    L.sph_mlp_init.argtypes = [vp, ull]
    L.sph_mlp_set_weights.argtypes = [vp, fp]
    L.sph_mlp_get_weights.argtypes = [vp, fp]
    L.sph_mlp_set_normalization.argtypes = [vp, fp, fp, C.c_float, C.c_float]
    L.sph_mlp_get_normalization.argtypes = [vp, fp, fp, fp, fp]
    L.sph_mlp_fit_normalization.argtypes = [vp, vp, vp, sz]
    L.sph_mlp_gradient.argtypes = [vp, vp, vp, sz, fp, fp]
    L.sph_mlp_train_step.argtypes = [vp, vp, vp, sz, sz, C.c_float, ull, fp]
    L.sph_mlp_reset_optimizer.argtypes = [vp]
    L.sph_mlp_forward.argtypes = [vp, vp, sz, vp]
    L.sph_mlp_infer_bf16.argtypes = [vp, vp, sz, vp]
    L.sph_mlp_synchronize.argtypes = [vp]
    L.sph_mlp_save.argtypes = [vp, C.c_char_p]
    L.sph_mlp_load.argtypes = [vp, C.c_char_p]
    L.sph_set_warm_start.argtypes = [vp, vp]
    L.sph_set_warm_start_previous.argtypes = [vp, C.c_float]
    L.sph_set_graph.argtypes = [vp, ip]
    L.sph_autotune.argtypes = [vp, ip, C.POINTER(Tuning)]
    L.sph_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.sph_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.sph_create_batch.argtypes = [C.POINTER(Params), ip, C.POINTER(Scene), C.POINTER(vp)]
    L.sph_batch_scene_count.argtypes = [vp, C.POINTER(ip)]
    L.sph_batch_scene_errors.argtypes = [vp, fp]
    L.sph_get_residuals.argtypes = [vp, fp, ip, C.POINTER(ip)]
    L.sph_mlp_set_output_clip.argtypes = [vp, C.c_float]
    L.sph_mlp_get_output_clip.argtypes = [vp]
    L.sph_mlp_get_output_clip.restype = C.c_float
    L.sph_mlp_set_step_scale.argtypes = [vp, C.c_float]
    L.sph_mlp_get_step_scale.argtypes = [vp]
    L.sph_mlp_get_step_scale.restype = C.c_float
    L.sph_set_sample_logging.argtypes = [vp, ip]
    L.sph_warm_samples.argtypes = [vp, vp, vp, sz, sz, ip, C.POINTER(sz)]
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise SphError(f"status {rc}: {lib().sph_last_error().decode()}")


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _uptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def default_params(h, **overrides):
    p = Params()
    lib().sph_default_params(C.byref(p), float(h))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def device_count():
    n = C.c_int(0)
    check(lib().sph_get_device_count(C.byref(n)))
    return n.value


class Context:
    """One sph_ctx: FP32 device state plus the substep stages."""

    def __init__(self, params, box, capacity):
        self.params = params
        b = AABB(0.0, float(box[0]), 0.0, float(box[1]), 0.0, float(box[2]))
        self._h = C.c_void_p()
        check(lib().sph_create(C.byref(params), C.byref(b), int(capacity), C.byref(self._h)))
        self.capacity = int(capacity)

    def close(self):
        if self._h:
            lib().sph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    # state
    def upload(self, pos, vel=None):
        pos = _f32(pos).reshape(-1, 3)
        n = len(pos)
        vel = np.zeros((n, 3), np.float32) if vel is None else _f32(vel).reshape(-1, 3)
        cols = [np.ascontiguousarray(pos[:, i]) for i in range(3)] + \
               [np.ascontiguousarray(vel[:, i]) for i in range(3)]
        check(lib().sph_upload_state(self._h, n, *[_fptr(c) for c in cols]))
        self.n = n

    def download(self):
        n = self.count()
        cols = [np.zeros(n, np.float32) for _ in range(6)]
        check(lib().sph_download_state(self._h, *[_fptr(c) for c in cols]))
        return np.stack(cols[:3], 1), np.stack(cols[3:], 1)

    def count(self):
        n = C.c_int()
        check(lib().sph_particle_count(self._h, C.byref(n)))
        return n.value

    def grid_dims(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        check(lib().sph_grid_dims(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def set_params(self, params):
        check(lib().sph_set_params(self._h, C.byref(params)))
        self.params = params

    def set_obstacle(self, sdf, dims, spacing, wmin, wmax):
        if sdf is None:
            check(lib().sph_set_obstacle(self._h, None))
            return
        sdf = _f32(sdf).ravel()
        d = ObstacleDesc(int(dims[0]), int(dims[1]), int(dims[2]), float(spacing),
                         AABB(wmin[0], wmax[0], wmin[1], wmax[1], wmin[2], wmax[2]),
                         sdf.ctypes.data)
        check(lib().sph_set_obstacle(self._h, C.byref(d)))

    # stages (reference call order)
    def apply_gravity(self): check(lib().sph_apply_gravity(self._h))
    def predict_positions(self): check(lib().sph_predict_positions(self._h))
    def find_all_neighbors(self): check(lib().sph_find_all_neighbors(self._h))
    def compute_densities(self): check(lib().sph_compute_densities(self._h))
    def compute_lambda(self): check(lib().sph_compute_lambda(self._h))
    def update_dps(self, it): check(lib().sph_update_dps(self._h, int(it)))
    def update_position_stars(self): check(lib().sph_update_position_stars(self._h))
    def update_velocities(self): check(lib().sph_update_velocities(self._h))
    def compute_surface_tension(self): check(lib().sph_compute_surface_tension(self._h))
    def apply_viscosity(self): check(lib().sph_apply_viscosity(self._h))
    def compute_vorticity(self): check(lib().sph_compute_vorticity(self._h))
    def apply_vorticity_confinement(self): check(lib().sph_apply_vorticity_confinement(self._h))
    def update_positions(self): check(lib().sph_update_positions(self._h))

    def step(self, substeps=1):
        check(lib().sph_step(self._h, int(substeps)))

    def synchronize(self):
        check(lib().sph_synchronize(self._h))

    def stats(self):
        s = Stats()
        check(lib().sph_get_stats(self._h, C.byref(s)))
        return s

    def kinetic_energy(self):
        e = C.c_double()
        check(lib().sph_compute_energy(self._h, C.byref(e)))
        return e.value

    @staticmethod
    def _tuning_dict(t):
        names = kernel_names()
        return dict(block={names[k]: t.block[k] for k in range(len(names))},
                    depth={names[k]: t.depth[k] for k in range(len(names))},
                    split_collide=bool(t.split_collide), wide_lists=bool(t.wide_lists),
                    ms_per_substep_default=t.ms_per_substep_default,
                    ms_per_substep_tuned=t.ms_per_substep_tuned)

    def autotune(self, substeps=3):
        """Launch autotuner; returns {kernel name: threads per block}, gathers in
        flight, obstacle-push split, list format and timings."""
        t = Tuning()
        check(lib().sph_autotune(self._h, int(substeps), C.byref(t)))
        return self._tuning_dict(t)

    def tuning(self):
        """Current launch choices, as autotune returns them (timings zero)."""
        t = Tuning()
        check(lib().sph_get_tuning(self._h, C.byref(t)))
        return self._tuning_dict(t)

    def set_tuning(self, block=None, depth=None, split_collide=None, wide_lists=None):
        """Applies launch choices; dicts map kernel names to values, unnamed
        kernels keep their current choice."""
        t = Tuning()
        check(lib().sph_get_tuning(self._h, C.byref(t)))
        names = kernel_names()
        for src, dst in ((block, t.block), (depth, t.depth)):
            for name, v in (src or {}).items():
                dst[names.index(name)] = int(v)
        if split_collide is not None:
            t.split_collide = int(bool(split_collide))
        if wide_lists is not None:
            t.wide_lists = int(bool(wide_lists))
        check(lib().sph_set_tuning(self._h, C.byref(t)))

    def residuals(self):
        """Max density error of each density pass of the last fused substep."""
        buf = np.zeros(160, np.float32)
        k = C.c_int()
        check(lib().sph_get_residuals(self._h, _fptr(buf), len(buf), C.byref(k)))
        return buf[:k.value].copy()

    def set_graph(self, enable):
        """Replay substeps from a captured HIP graph."""
        check(lib().sph_set_graph(self._h, int(bool(enable))))

    def set_warm_start(self, net):
        """Warm-start following fused substeps with `net` (an Mlp), None = off."""
        check(lib().sph_set_warm_start(self._h, net.handle if net is not None else None))

    def set_warm_start_previous(self, scale):
        """Warm-start following fused substeps from scale x the previous
        substep's accumulated multiplier (0 = off)."""
        check(lib().sph_set_warm_start_previous(self._h, float(scale)))

    def set_sample_logging(self, enable):
        check(lib().sph_set_sample_logging(self._h, int(bool(enable))))

    def warm_samples(self, x_buf, t_buf, offset, capacity, every=1):
        """Appends the last logged substep's rows to device buffers; returns rows."""
        w = C.c_size_t()
        check(lib().sph_warm_samples(self._h, x_buf.ptr, t_buf.ptr, int(offset), int(capacity),
                                     int(every), C.byref(w)))
        return w.value

    # inspection
    def field(self, which):
        n = self.count()
        comps = _FIELD_COMPONENTS[which]
        out = np.zeros(n * comps, np.float32)
        check(lib().sph_download_field(self._h, int(which), _fptr(out)))
        return out.reshape(n, 3) if comps == 3 else out

    def neighbors(self):
        n = self.count()
        ids = np.zeros(max(n, 1) * MAX_NEIGHBORS, np.uint32)
        cnt = np.zeros(max(n, 1), np.int32)
        check(lib().sph_download_neighbors(self._h, _uptr(ids),
                                           cnt.ctypes.data_as(C.POINTER(C.c_int))))
        return ids, cnt

    def neighbor_counts(self):
        """Neighbour count per particle id (the lists themselves not downloaded)."""
        cnt = np.zeros(max(self.count(), 1), np.int32)
        check(lib().sph_download_neighbors(self._h, None,
                                           cnt.ctypes.data_as(C.POINTER(C.c_int))))
        return cnt

    def sorted_pairs(self):
        n = self.count()
        k = np.zeros(max(n, 1), np.uint32)
        i = np.zeros(max(n, 1), np.uint32)
        check(lib().sph_download_sorted(self._h, _uptr(k), _uptr(i)))
        return k[:n], i[:n]

    def cell_bounds(self):
        nx, ny, nz = self.grid_dims()
        cells = nx * ny * nz
        s = np.zeros(cells, np.uint32)
        e = np.zeros(cells, np.uint32)
        check(lib().sph_download_cell_bounds(self._h, _uptr(s), _uptr(e)))
        return s, e

    # timing
    def profile(self, enable):
        check(lib().sph_profile_enable(self._h, int(bool(enable))))

    def profile_reset(self):
        check(lib().sph_profile_reset(self._h))

    def profile_read(self, kernel):
        ms, n = C.c_double(), C.c_longlong()
        check(lib().sph_profile_read(self._h, int(kernel), C.byref(ms), C.byref(n)))
        return ms.value, n.value


class BatchContext(Context):
    """Batched independent scenes in one context (sph_create_batch).

    scenes: list of dicts with keys particles, size (3 floats) and optionally
    sdf = dict(sdf=<flat float32 grid>, dims=(nx, ny, nz), spacing, wmin, wmax).
    Positions are uploaded and downloaded concatenated in scene order, each
    scene in its own coordinates (box [0, size))."""

    def __init__(self, params, scenes):
        self.params = params
        arr = (Scene * len(scenes))()
        self._keep = []
        for k, sc in enumerate(scenes):
            arr[k].particles = int(sc["particles"])
            arr[k].size_x, arr[k].size_y, arr[k].size_z = [float(v) for v in sc["size"]]
            o = sc.get("sdf")
            if o is not None:
                g = _f32(o["sdf"]).ravel()
                self._keep.append(g)
                wmin, wmax = o["wmin"], o["wmax"]
                arr[k].obstacle = ObstacleDesc(
                    int(o["dims"][0]), int(o["dims"][1]), int(o["dims"][2]), float(o["spacing"]),
                    AABB(wmin[0], wmax[0], wmin[1], wmax[1], wmin[2], wmax[2]), g.ctypes.data)
        self._h = C.c_void_p()
        check(lib().sph_create_batch(C.byref(params), len(scenes), arr, C.byref(self._h)))
        counts = [int(sc["particles"]) for sc in scenes]
        self.capacity = sum(counts)
        self.nscenes = len(scenes)
        self.first = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)

    def scene_errors(self):
        out = np.zeros(self.nscenes, np.float32)
        check(lib().sph_batch_scene_errors(self._h, _fptr(out)))
        return out


MLP_FEATURES = 12
MLP_STRIDE = 16
MLP_PARAMS = 5313


class Mlp:
    """Lambda warm-start network (sph_mlp_*): FP32 training on the GPU, BF16 MFMA
    inference.  Device arrays are DeviceBuffer objects (rows of MLP_STRIDE floats)."""

    def __init__(self, seed=None):
        self._h = C.c_void_p()
        check(lib().sph_mlp_create(C.byref(self._h)))
        if seed is not None:
            check(lib().sph_mlp_init(self._h, int(seed)))

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().sph_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_weights(self, theta):
        theta = _f32(theta)
        assert theta.size == MLP_PARAMS
        check(lib().sph_mlp_set_weights(self._h, _fptr(theta)))

    def weights(self):
        out = np.zeros(MLP_PARAMS, np.float32)
        check(lib().sph_mlp_get_weights(self._h, _fptr(out)))
        return out

    def set_normalization(self, mean, std, target_mean, target_std):
        m, s = _f32(mean), _f32(std)
        check(lib().sph_mlp_set_normalization(self._h, _fptr(m), _fptr(s), float(target_mean),
                                              float(target_std)))

    def normalization(self):
        m, s = np.zeros(MLP_FEATURES, np.float32), np.zeros(MLP_FEATURES, np.float32)
        tm, ts = C.c_float(), C.c_float()
        check(lib().sph_mlp_get_normalization(self._h, _fptr(m), _fptr(s), C.byref(tm),
                                              C.byref(ts)))
        return m, s, tm.value, ts.value

    def set_output_clip(self, max_target):
        check(lib().sph_mlp_set_output_clip(self._h, float(max_target)))

    def output_clip(self):
        return lib().sph_mlp_get_output_clip(self._h)

    def set_step_scale(self, scale):
        check(lib().sph_mlp_set_step_scale(self._h, float(scale)))

    def step_scale(self):
        return lib().sph_mlp_get_step_scale(self._h)

    def fit_normalization(self, x_buf, t_buf, count):
        check(lib().sph_mlp_fit_normalization(self._h, x_buf.ptr, t_buf.ptr, int(count)))

    def gradient(self, x_buf, t_buf, count):
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_float()
        check(lib().sph_mlp_gradient(self._h, x_buf.ptr, t_buf.ptr, int(count), _fptr(g),
                                     C.byref(loss)))
        return loss.value, g

    def train_step(self, x_buf, t_buf, count, batch, lr, seed, want_loss=False):
        loss = C.c_float()
        check(lib().sph_mlp_train_step(self._h, x_buf.ptr, t_buf.ptr, int(count), int(batch),
                                       float(lr), int(seed),
                                       C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def reset_optimizer(self):
        check(lib().sph_mlp_reset_optimizer(self._h))

    def forward(self, x_buf, count, y_buf):
        check(lib().sph_mlp_forward(self._h, x_buf.ptr, int(count), y_buf.ptr))

    def infer_bf16(self, x_buf, count, y_buf):
        check(lib().sph_mlp_infer_bf16(self._h, x_buf.ptr, int(count), y_buf.ptr))

    def synchronize(self):
        check(lib().sph_mlp_synchronize(self._h))

    def save(self, path):
        check(lib().sph_mlp_save(self._h, str(path).encode()))

    def load(self, path):
        check(lib().sph_mlp_load(self._h, str(path).encode()))


def kernel_names():
    L = lib()
    return [L.sph_profile_kernel_name(i).decode() for i in range(L.sph_profile_kernel_count())]


class DeviceBuffer:
    """Raw device allocation for the native sub-boundary entry points."""

    def __init__(self, nbytes):
        self.ptr = C.c_void_p()
        self.nbytes = int(nbytes)
        check(lib().sph_device_malloc(C.byref(self.ptr), self.nbytes))

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        check(lib().sph_memcpy_htod(b.ptr, a.ctypes.data, a.nbytes))
        return b

    def to_array(self, dtype, count):
        out = np.zeros(count, dtype)
        check(lib().sph_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().sph_device_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def sort_by_key(keys, values, key_bits=None):
    """SortByKey (thrust_c.cu:86-91) on device copies; returns host arrays.
    key_bits None: sph_sort_by_key (passes from the keys' highest set bit)."""
    keys = np.ascontiguousarray(keys, np.uint32)
    values = np.ascontiguousarray(values, np.uint32)
    n = len(keys)
    kb, vb = DeviceBuffer.from_array(keys), DeviceBuffer.from_array(values)
    if key_bits is None:
        check(lib().sph_sort_by_key(kb.ptr, vb.ptr, n, None))
    else:
        check(lib().sph_sort_by_key_bits(kb.ptr, vb.ptr, n, int(key_bits), None))
    return kb.to_array(np.uint32, n), vb.to_array(np.uint32, n)


def cell_bounds(keys, cells):
    """FindLowerBounds/FindUpperBounds (thrust_c.cu:93-120)."""
    keys = np.ascontiguousarray(keys, np.uint32)
    kb = DeviceBuffer.from_array(keys if len(keys) else np.zeros(1, np.uint32))
    lo, hi = DeviceBuffer(4 * max(cells, 1)), DeviceBuffer(4 * max(cells, 1))
    check(lib().sph_find_lower_bounds(kb.ptr, len(keys), int(cells), lo.ptr, None))
    check(lib().sph_find_upper_bounds(kb.ptr, len(keys), int(cells), hi.ptr, None))
    return lo.to_array(np.uint32, cells), hi.to_array(np.uint32, cells)


class Block(C.Structure):
    """struct sph_block: lattice block of nx*ny*nz particles at origin."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("origin_x", C.c_double),
                ("origin_y", C.c_double), ("origin_z", C.c_double)]


SHAPE_SPHERE, SHAPE_TORUS, SHAPE_BOX = 0, 1, 2


class Shape(C.Structure):
    """struct sph_shape: analytic signed-distance primitive."""
    _fields_ = [("kind", C.c_int), ("axis", C.c_int), ("center", C.c_double * 3),
                ("radius", C.c_double), ("radius2", C.c_double), ("half", C.c_double * 3)]


def sphere(center, radius):
    return Shape(SHAPE_SPHERE, 0, (C.c_double * 3)(*center), radius, 0.0, (C.c_double * 3)())


def torus(center, ring_radius, tube_radius, axis=2):
    return Shape(SHAPE_TORUS, axis, (C.c_double * 3)(*center), ring_radius, tube_radius,
                 (C.c_double * 3)())


def box_shape(center, half):
    return Shape(SHAPE_BOX, 0, (C.c_double * 3)(*center), 0.0, 0.0, (C.c_double * 3)(*half))


def scene_lattice(blocks, spacing, jitter=0.0, seed=0):
    """sph_scene_lattice: float32 (n, 3) positions of the given blocks."""
    L = lib()
    L.sph_scene_lattice.argtypes = [C.POINTER(Block), C.c_int, C.c_double, C.c_double,
                                    C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                    C.POINTER(C.c_int)]
    arr = (Block * len(blocks))(*blocks)
    n = C.c_int()
    check(L.sph_scene_lattice(arr, len(blocks), spacing, jitter, seed, None, None, None, 0,
                              C.byref(n)))
    cols = [np.zeros(n.value, np.float32) for _ in range(3)]
    check(L.sph_scene_lattice(arr, len(blocks), spacing, jitter, seed, cols[0].ctypes.data,
                              cols[1].ctypes.data, cols[2].ctypes.data, n.value, C.byref(n)))
    return np.stack(cols, 1)


def sdf_sample_shapes(shapes, dims, wmin, wmax):
    """sph_sdf_sample_shapes: union SDF at texel centres, x fastest (float32)."""
    L = lib()
    L.sph_sdf_sample_shapes.argtypes = [C.POINTER(Shape), C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.POINTER(AABB), C.c_void_p]
    arr = (Shape * len(shapes))(*shapes)
    out = np.zeros(int(dims[0]) * int(dims[1]) * int(dims[2]), np.float32)
    b = AABB(wmin[0], wmax[0], wmin[1], wmax[1], wmin[2], wmax[2])
    check(L.sph_sdf_sample_shapes(arr, len(shapes), int(dims[0]), int(dims[1]), int(dims[2]),
                                  C.byref(b), out.ctypes.data))
    return out


def collide(pos, sdf, dims, spacing, wmin, wmax):
    """CalculateParticleCollisions (obstacle_cuda.cu:118-131) on device SoA
    arrays: float32 positions, or float64 ones (the reference's own arrays,
    sph_calculate_particle_collisions_f64).  Returns host positions of the
    same type."""
    dt = np.float64 if np.asarray(pos).dtype == np.float64 else np.float32
    pos = np.ascontiguousarray(pos, dt).reshape(-1, 3)
    n = len(pos)
    sdf = _f32(sdf).ravel()
    d = ObstacleDesc(int(dims[0]), int(dims[1]), int(dims[2]), float(spacing),
                     AABB(wmin[0], wmax[0], wmin[1], wmax[1], wmin[2], wmax[2]), sdf.ctypes.data)
    h = C.c_void_p()
    check(lib().sph_sdf_create(C.byref(d), C.byref(h)))
    entry = (lib().sph_calculate_particle_collisions_f64 if dt == np.float64 else
             lib().sph_calculate_particle_collisions)
    try:
        bufs = [DeviceBuffer.from_array(np.ascontiguousarray(pos[:, i])) for i in range(3)]
        check(entry(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, n, h, None))
        return np.stack([b.to_array(dt, n) for b in bufs], 1)
    finally:
        lib().sph_sdf_destroy(h)
