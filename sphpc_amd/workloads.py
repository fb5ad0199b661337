"""Synthetic, seeded scenes of the benchmark workloads (host side).

double_dam_break: two 256x128x128 FP32 lattice blocks (8,388,608 particles at
scale 1) at s = 0.008 m, seeded +-0.01 s jitter, at the two x-ends of a
6 x 2 x 2 m box, with a torus (R 0.4, r 0.12, axis z) united with an
axis-aligned box sampled onto a 256^3 FP32 SDF grid.

sphere_block: one 128x128x64 FP32 lattice block (1,048,576 particles) at
s = 0.01 m, seeded +-0.01 s jitter, in the origin corner of a 4 x 2 x 2 m box,
with a sphere of radius 0.3 m at the box centre sampled onto a 128^3 FP32 SDF
grid over the sphere's bounding box grown by 0.1 m.

batched_scenes: independent scenes, each with a particle count log-uniform in
[n_min, n_max], box aspect ratios x:z and y:z uniform in [1, 4], a fluid block
filling a fraction uniform in [0.1, 0.5] of the box (a dam break in the
origin corner) and a sphere obstacle of radius uniform in [0.1, 0.3] of the
box height on the floor, sampled onto a 64^3 FP32 SDF grid.
"""
import math

import numpy as np

import sphpc_amd as sa


def double_dam_break(scale=1.0):
    s = float(np.float32(0.008))
    nx, ny, nz = (max(1, int(round(256 * scale))), max(1, int(round(128 * scale))),
                  max(1, int(round(128 * scale))))
    y0 = (2.0 - ny * s) / 2.0
    blocks = [sa.Block(nx, ny, nz, 0.0, y0, 0.0),
              sa.Block(nx, ny, nz, 6.0 - nx * s, y0, 0.0)]
    pos = sa.scene_lattice(blocks, s, jitter=0.01, seed=2025)
    dims = (256, 256, 256)
    wmin, wmax = (2.2, 0.2, 0.0), (3.8, 1.8, 1.6)
    shapes = [sa.torus((3.0, 0.75, 0.12), 0.4, 0.12, axis=2),
              sa.box_shape((3.0, 1.45, 0.25), (0.25, 0.2, 0.25))]
    sdf = sa.sdf_sample_shapes(shapes, dims, wmin, wmax)
    return dict(s=s, box=(6.0, 2.0, 2.0), pos=pos, sdf=sdf, dims=dims, wmin=wmin, wmax=wmax,
                dx=(wmax[0] - wmin[0]) / dims[0])


def sphere_block(scale=1.0):
    s = float(np.float32(0.01))
    nx, ny, nz = (max(1, int(round(128 * scale))), max(1, int(round(128 * scale))),
                  max(1, int(round(64 * scale))))
    pos = sa.scene_lattice([sa.Block(nx, ny, nz, 0.0, 0.0, 0.0)], s, jitter=0.01, seed=2024)
    dims = (128, 128, 128)
    c, r = (2.0, 1.0, 1.0), 0.3
    wmin = tuple(ci - r - 0.1 for ci in c)
    wmax = tuple(ci + r + 0.1 for ci in c)
    sdf = sa.sdf_sample_shapes([sa.sphere(c, r)], dims, wmin, wmax)
    return dict(s=s, box=(4.0, 2.0, 2.0), pos=pos, sdf=sdf, dims=dims, wmin=wmin, wmax=wmax,
                dx=(wmax[0] - wmin[0]) / dims[0])


def batched_scenes(count=512, seed=4, spacing=0.01, n_min=2000, n_max=128000, sdf_dim=64):
    """Scene list for sph.BatchContext plus the concatenated initial positions."""
    s = float(np.float32(spacing))
    rng = np.random.default_rng(seed)
    scenes, pos = [], []
    for k in range(count):
        n_target = math.exp(rng.uniform(math.log(n_min), math.log(n_max)))
        ax, ay = rng.uniform(1.0, 4.0, 2)
        fill = rng.uniform(0.1, 0.5)
        volume = n_target * s ** 3 / fill
        lz = (volume / (ax * ay)) ** (1.0 / 3.0)
        size = (ax * lz, ay * lz, lz)
        side = fill ** (1.0 / 3.0)
        counts = [max(2, int(side * L / s)) for L in size]
        p = sa.scene_lattice([sa.Block(*counts, 0.0, 0.0, 0.0)], s, jitter=0.01,
                             seed=seed * 100003 + k)
        r = rng.uniform(0.1, 0.3) * lz
        c = (0.7 * size[0], 0.5 * size[1], r)
        wmin = tuple(ci - 1.25 * r for ci in c)
        wmax = tuple(ci + 1.25 * r for ci in c)
        dims = (sdf_dim,) * 3
        grid = sa.sdf_sample_shapes([sa.sphere(c, r)], dims, wmin, wmax)
        scenes.append(dict(particles=len(p), size=size,
                           sdf=dict(sdf=grid, dims=dims, spacing=2.5 * r / sdf_dim, wmin=wmin,
                                    wmax=wmax)))
        pos.append(p)
    return s, scenes, np.concatenate(pos)
