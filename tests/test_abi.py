"""The C-ABI library loads without a GPU and exports every function that
include/sphpc_amd.h declares (no compute calls)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sphpc_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sph_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_stage_functions():
    names = declared_functions()
    for stage in ("sph_apply_gravity", "sph_predict_positions", "sph_find_all_neighbors",
                  "sph_compute_densities", "sph_compute_lambda", "sph_update_dps",
                  "sph_update_position_stars", "sph_update_velocities",
                  "sph_compute_surface_tension", "sph_apply_viscosity", "sph_compute_vorticity",
                  "sph_apply_vorticity_confinement", "sph_update_positions", "sph_sort_by_key",
                  "sph_find_lower_bounds", "sph_find_upper_bounds",
                  "sph_calculate_particle_collisions"):
        assert stage in names


def test_library_exports_every_declared_symbol():
    import sphpc_amd
    lib = sphpc_amd.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_defaults_and_argument_errors_without_gpu():
    import sphpc_amd as sa
    p = sa.default_params(0.04)
    assert p.solve_iterations == 10 and abs(p.lambda_epsilon - 0.01) < 1e-15
    assert abs(p.vorticity_epsilon - 0.004) < 1e-15 and p.v_max == 100.0
    lib = sa.lib()
    h = C.c_void_p()
    # non-float h is rejected before any device call
    p.smoothing_radius = 0.1
    box = sa.AABB(0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    assert lib.sph_create(C.byref(p), C.byref(box), 10, C.byref(h)) == -1
    assert b"FP32" in lib.sph_last_error()
    p.smoothing_radius = 0.03999999910593033  # float32(0.04)
    p.rest_mass = 0.008
    box.min_x = 0.5
    assert lib.sph_create(C.byref(p), C.byref(box), 10, C.byref(h)) == -1
    assert b"origin" in lib.sph_last_error()


def test_capacity_limit_without_gpu():
    """Contexts are capped at SPH_MAX_PARTICLES (float4 particle arrays below
    4 GiB, addressed with 32-bit byte offsets); the check precedes any device
    call."""
    import sphpc_amd as sa
    text = open(HEADER).read()
    m = re.search(r"#define SPH_MAX_PARTICLES \(\(1 << 28\) - 64\)", text)
    assert m, "SPH_MAX_PARTICLES definition"
    cap = (1 << 28) - 64
    assert (cap + 64) * 16 <= 1 << 32
    lib = sa.lib()
    p = sa.default_params(0.03999999910593033)
    p.rest_mass = 0.008
    box = sa.AABB(0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    h = C.c_void_p()
    assert lib.sph_create(C.byref(p), C.byref(box), cap + 1, C.byref(h)) == -1
    assert b"2^28" in lib.sph_last_error()
    scenes = (sa.Scene * 2)()
    for s in scenes:
        s.particles = cap // 2 + 1
        s.size_x = s.size_y = s.size_z = 1.0
    assert lib.sph_create_batch(C.byref(p), 2, scenes, C.byref(h)) == -1
    assert b"2^28" in lib.sph_last_error()
