"""Host-side inputs of the library (no GPU): lattice scenes, analytic SDF grids,
the SDFGen format and the sim-config.ini reader."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import sphpc_amd as sa


def test_lattice_matches_construct_fluid_volume():
    # geometry.c:42-55: x fastest, centres at min + i*s + s/2 (computed in double)
    s = float(np.float32(0.02))
    pos = sa.scene_lattice([sa.Block(5, 4, 3, 0.1, 0.2, 0.3)], s)
    iz, iy, ix = np.meshgrid(np.arange(3), np.arange(4), np.arange(5), indexing="ij")
    ref = np.stack([0.1 + ix.ravel() * s + s / 2.0, 0.2 + iy.ravel() * s + s / 2.0,
                    0.3 + iz.ravel() * s + s / 2.0], 1).astype(np.float32)
    assert np.array_equal(pos, ref)


def test_lattice_jitter_bounded_seeded_and_multiblock():
    s = 0.008
    blocks = [sa.Block(16, 8, 8, 0.0, 0.5, 0.0), sa.Block(16, 8, 8, 5.872, 0.5, 0.0)]
    a = sa.scene_lattice(blocks, s, jitter=0.01, seed=7)
    b = sa.scene_lattice(blocks, s, jitter=0.01, seed=7)
    c = sa.scene_lattice(blocks, s, jitter=0.01, seed=8)
    plain = sa.scene_lattice(blocks, s)
    assert len(a) == 2 * 16 * 8 * 8
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    d = a.astype(np.float64) - plain
    assert np.max(np.abs(d)) <= 0.01 * s + 2 * np.spacing(np.float32(6.0))
    assert abs(d.mean()) < 0.001 * s
    assert d.std() > 0.004 * s  # U(-0.01s, 0.01s) has std 0.0058 s
    # second block starts where asked
    assert abs(a[16 * 8 * 8:, 0].min() - (5.872 + s / 2)) < 0.02 * s


def test_uniform_generator_range_and_mean():
    L = sa.lib()
    L.sph_scene_uniform.restype = C.c_double
    L.sph_scene_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_int]
    u = np.array([L.sph_scene_uniform(3, i, i % 3) for i in range(20000)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01
    assert abs(np.var(u) - 1.0 / 12.0) < 0.005


def test_analytic_sdf_shapes():
    dims = (20, 16, 12)
    wmin, wmax = (0.0, 0.0, 0.0), (2.0, 1.6, 1.2)
    dx = 0.1
    ax = [wmin[i] + (np.arange(dims[i]) + 0.5) * dx for i in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    # sphere
    g = sa.sdf_sample_shapes([sa.sphere((1.0, 0.8, 0.6), 0.4)], dims, wmin, wmax)
    ref = np.sqrt((X - 1.0) ** 2 + (Y - 0.8) ** 2 + (Z - 0.6) ** 2) - 0.4
    assert np.allclose(g, ref.ravel(), atol=1e-6)
    # torus around z
    g = sa.sdf_sample_shapes([sa.torus((1.0, 0.8, 0.6), 0.5, 0.1, axis=2)], dims, wmin, wmax)
    q = np.sqrt((X - 1.0) ** 2 + (Y - 0.8) ** 2) - 0.5
    ref = np.sqrt(q ** 2 + (Z - 0.6) ** 2) - 0.1
    assert np.allclose(g, ref.ravel(), atol=1e-6)
    # box: inside distance is minus the distance to the nearest face
    g = sa.sdf_sample_shapes([sa.box_shape((1.0, 0.8, 0.6), (0.3, 0.2, 0.25))], dims, wmin, wmax)
    qx, qy, qz = np.abs(X - 1.0) - 0.3, np.abs(Y - 0.8) - 0.2, np.abs(Z - 0.6) - 0.25
    out = np.sqrt(np.maximum(qx, 0) ** 2 + np.maximum(qy, 0) ** 2 + np.maximum(qz, 0) ** 2)
    ref = out + np.minimum(np.maximum(np.maximum(qx, qy), qz), 0)
    assert np.allclose(g, ref.ravel(), atol=1e-6)
    # union is the minimum
    u = sa.sdf_sample_shapes([sa.sphere((0.5, 0.5, 0.5), 0.2), sa.sphere((1.5, 0.8, 0.6), 0.3)],
                             dims, wmin, wmax)
    a = sa.sdf_sample_shapes([sa.sphere((0.5, 0.5, 0.5), 0.2)], dims, wmin, wmax)
    b = sa.sdf_sample_shapes([sa.sphere((1.5, 0.8, 0.6), 0.3)], dims, wmin, wmax)
    assert np.array_equal(u, np.minimum(a, b))


def test_sdfgen_round_trip(tmp_path):
    L = sa.lib()
    L.sph_sdf_write_sdfgen.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(C.c_double), C.c_double, C.c_void_p]
    L.sph_sdf_read_sdfgen.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                      C.POINTER(C.c_double), C.POINTER(C.POINTER(C.c_float))]
    L.sph_host_free.argtypes = [C.c_void_p]
    vals = np.random.default_rng(1).normal(size=4 * 3 * 2).astype(np.float32)
    path = str(tmp_path / "g.sdf").encode()
    origin = (C.c_double * 3)(0.25, -1.5, 3.0)
    assert L.sph_sdf_write_sdfgen(path, 4, 3, 2, origin, 0.125, vals.ctypes.data) == 0
    text = open(path).read().split("\n")
    assert text[0] == "4 3 2" and text[2] == "0.125"  # obstacle.c:20-46 layout
    dims = (C.c_int * 3)()
    org = (C.c_double * 3)()
    dx = C.c_double()
    out = C.POINTER(C.c_float)()
    assert L.sph_sdf_read_sdfgen(path, dims, org, C.byref(dx), C.byref(out)) == 0
    assert tuple(dims) == (4, 3, 2) and tuple(org) == (0.25, -1.5, 3.0) and dx.value == 0.125
    got = np.ctypeslib.as_array(out, shape=(24,)).copy()
    L.sph_host_free(C.cast(out, C.c_void_p))
    assert np.array_equal(got, vals)
    bad = str(tmp_path / "bad.sdf").encode()
    open(bad, "w").write("4 3 2\n0 0 0\n0.1\n1.0\n2.0\n")
    assert L.sph_sdf_read_sdfgen(bad, dims, org, C.byref(dx), C.byref(out)) != 0


class RunConfig(C.Structure):
    _fields_ = [("number_steps", C.c_int), ("time_step", C.c_double),
                ("number_particles", C.c_int), ("g", C.c_double), ("c", C.c_double),
                ("k", C.c_double), ("boundary", sa.AABB), ("initial_fluid", sa.AABB),
                ("has_obstacle", C.c_int), ("sdf_file", C.c_char * 256),
                ("obstacle_min_x", C.c_double), ("obstacle_max_x", C.c_double),
                ("obstacle_min_y", C.c_double), ("obstacle_min_z", C.c_double),
                ("solve_iterations", C.c_int), ("lambda_epsilon", C.c_double),
                ("vorticity_epsilon", C.c_double), ("mass_factor", C.c_double),
                ("density_tolerance", C.c_double), ("vorticity_model", C.c_int),
                ("surface_tension", C.c_int), ("viscosity", C.c_int), ("vorticity", C.c_int),
                ("jitter", C.c_double), ("seed", C.c_ulonglong), ("output_every", C.c_int),
                ("warm_start_net", C.c_char * 256)]


REFERENCE_INI = """[SimParameters]
number_steps = 2000
time_step = 0.00416666
number_particles = 100000

[PhysicalParameters]
g = 9.8
c = 0.01
k = 3.0

[BoundaryVolume]
min_x = 0.0
max_x = 80.0
min_y = 0.0
max_y = 30.0
min_z = 0.0
max_z = 60.0

[InitialFluidVolume]
min_x = 0.0
max_x = 80.0
min_y = 25.0
max_y = 30.0
min_z = 0.0
max_z = 60.0

[Obstacle]
SDF_file_name = bunny_watertight.sdf
min_x = 45.0
max_x = 70.0
min_y = 0.0
min_z = 0.0
"""


def _load(path):
    L = sa.lib()
    L.sph_config_load.argtypes = [C.c_char_p, C.POINTER(RunConfig)]
    L.sph_config_error.restype = C.c_char_p
    cfg = RunConfig()
    rc = L.sph_config_load(str(path).encode(), C.byref(cfg))
    return rc, cfg, L.sph_config_error().decode()


def test_config_reads_reference_keys(tmp_path):
    p = tmp_path / "sim-config.ini"
    p.write_text(REFERENCE_INI)
    rc, cfg, _ = _load(p)
    assert rc == 0
    assert cfg.number_steps == 2000 and cfg.number_particles == 100000
    assert cfg.time_step == 0.00416666 and cfg.g == 9.8 and cfg.c == 0.01 and cfg.k == 3.0
    assert cfg.boundary.max_x == 80.0 and cfg.initial_fluid.min_y == 25.0
    assert cfg.has_obstacle == 1 and cfg.sdf_file == b"bunny_watertight.sdf"
    assert cfg.obstacle_min_x == 45.0 and cfg.obstacle_max_x == 70.0
    # hard-coded constants of the reference as defaults
    assert cfg.solve_iterations == 10 and cfg.lambda_epsilon == 0.01
    assert cfg.mass_factor == 0.991 and cfg.vorticity_epsilon < 0


def test_config_missing_key_aborts_like_reference(tmp_path):
    p = tmp_path / "sim-config.ini"
    p.write_text(REFERENCE_INI.replace("number_steps = 2000\n", ""))
    rc, _, msg = _load(p)
    assert rc != 0 and msg.startswith("Aborting:") and "SimParameters.number_steps" in msg


def test_config_optional_obstacle_and_solver_keys(tmp_path):
    text = REFERENCE_INI.split("[Obstacle]")[0] + "[Solver]\nsolve_iterations = 4\n" \
        "lambda_epsilon = 600\nmass_factor = 1.0\ndensity_tolerance = 0.001\n" \
        "vorticity_model = 1\njitter = 0.01\nseed = 42\n"
    p = tmp_path / "sim-config.ini"
    p.write_text(text)
    rc, cfg, _ = _load(p)
    assert rc == 0 and cfg.has_obstacle == 0
    assert cfg.solve_iterations == 4 and cfg.lambda_epsilon == 600.0
    assert cfg.mass_factor == 1.0 and cfg.density_tolerance == 0.001
    assert cfg.vorticity_model == 1 and cfg.jitter == 0.01 and cfg.seed == 42
    assert cfg.warm_start_net == b""


def test_config_warm_start_net_key(tmp_path):
    p = tmp_path / "sim-config.ini"
    p.write_text(REFERENCE_INI + "\n[Solver]\nwarm_start_net = nets/warm.bin\n")
    rc, cfg, _ = _load(p)
    assert rc == 0 and cfg.warm_start_net == b"nets/warm.bin"


def test_reference_runtime_ini_particle_count():
    # SetParameters/ConstructFluidVolume rules on Runtime/sim-config.ini's volumes:
    # s = float(cbrt(V/N)), counts floor(L/s) per axis
    V = 80.0 * 5.0 * 60.0
    s = float(np.float32((V / 100000) ** (1.0 / 3.0)))
    counts = [math.floor(80.0 / s), math.floor(5.0 / s), math.floor(60.0 / s)]
    pos = sa.scene_lattice([sa.Block(*counts, 0.0, 25.0, 0.0)], s)
    assert len(pos) == counts[0] * counts[1] * counts[2] == 128 * 8 * 96
