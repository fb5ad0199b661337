"""Known-answer and invariant checks of the FP64 oracle.

These pin the parts the reference's C code cannot (Thrust sort/bounds, the
texture collision) and cross-check the rest against closed forms:
perfect-lattice density/lambda (integer lattice sums), exact rational hashing
and neighbour sets, textbook stable sort and lower/upper bound, the stiffness
schedule, ballistic motion, an analytic sphere obstacle, antisymmetry of the
position correction and Galilean invariance of XSPH/vorticity."""
import ctypes as C
import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

import orc

L = orc.oracle()


def _lattice_offsets(max_r2):
    return [k for k in itertools.product(range(-2, 3), repeat=3)
            if 0 < k[0] ** 2 + k[1] ** 2 + k[2] ** 2 < max_r2]


def test_lattice_density_closed_form():
    # Interior density of a perfect lattice with h = 2s and mass 0.991 rho0 s^3:
    # rho = m*315/(64 pi h^9) * sum_{|k|^2<4} (h^2 - s^2|k|^2)^3
    #     = 0.991*1000*315*330/(64*pi*512)   (sum of (4-|k|^2)^3 = 330)
    s = 0.02
    p = orc.make_params(s, (1.0, 1.0, 1.0), mass_factor=0.991)
    counts = sum((4 - (k[0] ** 2 + k[1] ** 2 + k[2] ** 2)) ** 3 for k in _lattice_offsets(4))
    assert counts + 64 == 330
    expected = 0.991 * 1000.0 * 315.0 * 330.0 / (64.0 * math.pi * 512.0)
    assert abs(expected - 1000.6872) < 1e-4
    pos = orc.lattice_block(12, 12, 12, s, origin=(0.2, 0.2, 0.2)).astype(np.float64)
    xs, ys, zs = orc.soa(pos)
    nbr, cnt, _ = orc.neighbors(p, xs, ys, zs)
    rho = np.zeros(len(xs))
    L.orc_compute_densities(len(xs), xs, ys, zs, nbr, cnt, C.byref(p), rho)
    idx = np.arange(len(xs)).reshape(12, 12, 12)[4:8, 4:8, 4:8].ravel()
    # float32 lattice positions: tolerance from the ~1e-7 relative rounding of s
    assert np.allclose(rho[idx], expected, rtol=2e-6, atol=0)
    # 26 neighbours with |k|^2 <= 3, plus axial pairs at r = h whose float32
    # rounding lands on r^2 < h^2 (W(h) = 0, so density is unaffected)
    assert np.all((cnt[idx] >= 26) & (cnt[idx] <= 32))


def test_lattice_lambda_closed_form():
    s = 0.02
    p = orc.make_params(s, (1.0, 1.0, 1.0), mass_factor=0.991)
    h = p.h
    m = p.rest_mass
    grad2 = sum((45.0 / (math.pi * h ** 6) * (h - s * math.sqrt(k[0] ** 2 + k[1] ** 2 + k[2] ** 2)) ** 2) ** 2
                for k in _lattice_offsets(4))
    rho = 0.991 * 1000.0 * 315.0 * 330.0 / (64.0 * math.pi * 512.0)
    sum_c = m * m * grad2 / 1000.0 ** 2
    # perfect interior lattice: sum |grad C|^2 = 0.36670 / s^2
    assert abs(sum_c * s * s - 0.36670) < 5e-5
    lam_expected = -(rho / 1000.0 - 1.0) / (sum_c + 0.01)
    pos = orc.lattice_block(12, 12, 12, s, origin=(0.2, 0.2, 0.2)).astype(np.float64)
    xs, ys, zs = orc.soa(pos)
    nbr, cnt, _ = orc.neighbors(p, xs, ys, zs)
    n = len(xs)
    dens = np.zeros(n); lam = np.zeros(n)
    L.orc_compute_densities(n, xs, ys, zs, nbr, cnt, C.byref(p), dens)
    L.orc_compute_lambda(n, xs, ys, zs, dens, nbr, cnt, C.byref(p), lam)
    idx = np.arange(n).reshape(12, 12, 12)[4:8, 4:8, 4:8].ravel()
    assert np.allclose(lam[idx], lam_expected, rtol=5e-3)
    # interior delta p vanishes by symmetry; colour too
    d = [np.zeros(n) for _ in range(3)]
    L.orc_update_dps(n, xs, ys, zs, lam, nbr, cnt, C.byref(p), 0, *d)
    # scale of one pair's term; lambda varies by ~1e-3 relative between
    # particles because float32 lattice rounding moves rho by ~1e-6 and
    # C = rho/rho0 - 1 is only 7e-4, so the interior sum cancels to ~1e-3.
    scale = abs(lam_expected) * m * 45.0 / (math.pi * h ** 6) * h * h / 1000.0
    for a in d:
        assert np.max(np.abs(a[idx])) < 5e-3 * scale
    surface = np.arange(n).reshape(12, 12, 12)[-1, 4:8, 4:8].ravel()
    assert np.min(np.abs(d[2][surface])) > 0.1 * scale
    col = [np.zeros(n) for _ in range(3)]
    L.orc_compute_color(n, xs, ys, zs, dens, nbr, cnt, C.byref(p), *col)
    # surface colour is O(1); the interior cancels to float32-rounding level
    for a in col:
        assert np.max(np.abs(a[idx])) < 1e-4
    assert np.min(np.abs(col[2][surface])) > 0.1


def _exact_floor(x, h):
    return math.floor(Fraction(x) / Fraction(h))


def test_hash_is_exact_floor():
    s = 0.008
    p = orc.make_params(s, (6.0, 2.0, 2.0))
    h = p.h
    rng = np.random.default_rng(1)
    k = rng.integers(0, 370, 3000)
    base = (k * np.float32(h)).astype(np.float32)
    pts = [base, np.nextafter(base, np.float32(10)), np.nextafter(base, np.float32(-1)),
           rng.uniform(0, 5.99, 3000).astype(np.float32)]
    xs = np.concatenate(pts).astype(np.float64)
    xs = np.clip(xs, 0.0, 5.99)
    for x in xs[::7]:
        key = L.orc_hash_val(C.byref(p), float(x), 0.5, 1.25)
        gx = _exact_floor(float(x), h)
        gy = _exact_floor(0.5, h)
        gz = _exact_floor(1.25, h)
        assert key == p.nx * p.ny * gz + gy * p.nx + gx


def test_sort_is_stable_and_sorted():
    rng = np.random.default_rng(2)
    for n, hi in ((1, 5), (1000, 7), (20000, 2 ** 23), (5000, 2 ** 32 - 1)):
        keys = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
        vals = np.arange(n, dtype=np.uint32)
        order = np.argsort(keys, kind="stable")
        k2, v2 = keys.copy(), vals.copy()
        L.orc_sort_by_key(n, k2, v2)
        assert np.array_equal(k2, keys[order]) and np.array_equal(v2, vals[order])


def test_cell_bounds_are_lower_upper_bound():
    rng = np.random.default_rng(3)
    cells = 997
    keys = np.sort(rng.integers(0, cells, 4000).astype(np.uint32))
    keys[-50:] = cells - 1
    lo = np.zeros(cells, np.uint32); hi = np.zeros(cells, np.uint32)
    L.orc_find_lower_bounds(keys, len(keys), cells, lo)
    L.orc_find_upper_bounds(keys, len(keys), cells, hi)
    c = np.arange(cells, dtype=np.uint32)
    assert np.array_equal(lo, np.searchsorted(keys, c, "left"))
    assert np.array_equal(hi, np.searchsorted(keys, c, "right"))
    empty = np.zeros(0, np.uint32)
    L.orc_find_lower_bounds(empty, 0, cells, lo)
    assert not lo.any()


def test_neighbor_sets_brute_force_exact():
    # Brute force over all pairs with exact rational distances; order is the
    # (dz,dy,dx) cell order, then particle id (stable sort of ids).
    s = 0.02
    p = orc.make_params(s, (0.3, 0.3, 0.3))
    pos = orc.lattice_block(6, 6, 5, s, origin=(0.01, 0.02, 0.0), jitter=0.6, seed=4)
    pos = np.clip(pos, 0, 0.29).astype(np.float64)
    xs, ys, zs = orc.soa(pos)
    n = len(xs)
    nbr, cnt, dropped = orc.neighbors(p, xs, ys, zs)
    assert dropped == 0
    h = Fraction(p.h)
    F = [[Fraction(float(v)) for v in a] for a in (xs, ys, zs)]
    cell = [tuple(math.floor(F[a][i] / h) for a in range(3)) for i in range(n)]
    for i in range(n):
        hits = []
        for j in range(n):
            if j == i:
                continue
            r2 = sum((F[a][i] - F[a][j]) ** 2 for a in range(3))
            if r2 < h * h:
                d = tuple(cell[j][a] - cell[i][a] for a in range(3))
                assert max(abs(x) for x in d) <= 1
                hits.append(((d[2], d[1], d[0]), j))
        hits.sort()
        assert [j for _, j in hits] == list(nbr[i * orc.MAXN:i * orc.MAXN + cnt[i]])


def test_stiffness_schedule():
    p = orc.make_params(0.02, (1, 1, 1))
    assert L.orc_stiffness(C.byref(p), 0) == 1.0
    assert abs(L.orc_stiffness(C.byref(p), 1) - 0.95) < 1e-15
    assert abs(L.orc_stiffness(C.byref(p), 2) - (1 - math.sqrt(0.05))) < 1e-15
    assert abs(L.orc_stiffness(C.byref(p), 3) - (1 - 0.05 ** (1 / 3))) < 1e-15


def _state(pos, vel=None):
    n = len(pos)
    x, y, z = orc.soa(pos)
    v = np.zeros((n, 3)) if vel is None else vel
    vx, vy, vz = orc.soa(v)
    arrs = dict(x=x, y=y, z=z, vx=vx, vy=vy, vz=vz, xs=np.zeros(n), ys=np.zeros(n),
                zs=np.zeros(n), density=np.zeros(n), lambda_=np.zeros(n),
                nbr=np.zeros(n * orc.MAXN, np.uint32), cnt=np.zeros(n, np.int32))
    st = orc.OrcState(**{k: v.ctypes.data for k, v in arrs.items()})
    return st, arrs


def test_ballistic_single_particle():
    p = orc.make_params(0.02, (1.0, 1.0, 2.0), g=9.81, dt=1.0 / 300.0)
    st, a = _state(np.array([[0.5, 0.5, 1.5]]))
    for step in range(1, 61):
        L.orc_substep(1, C.byref(st), C.byref(p), None, 3, 0.0, None)
        expected = 1.5 - 9.81 * (1.0 / 300.0) ** 2 * step * (step + 1) / 2.0
        assert abs(a["z"][0] - expected) < 1e-12
        assert abs(a["vz"][0] + 9.81 * (1.0 / 300.0) * step) < 1e-9
    # floor clamp
    st, a = _state(np.array([[0.5, 0.5, 0.001]]), np.array([[0.0, 0.0, -3.0]]))
    L.orc_substep(1, C.byref(st), C.byref(p), None, 3, 0.0, None)
    assert a["z"][0] == 0.0


def test_sphere_obstacle_pushes_to_surface():
    dim = (64, 64, 64)
    wmin, wmax = (0.2, 0.2, 0.2), (1.0, 1.0, 1.0)
    center, R = (0.6, 0.6, 0.6), 0.25
    sdf, dx = orc.sphere_sdf(dim, wmin, wmax, center, R)
    o = orc.obstacle(sdf, dim, dx, wmin, wmax)
    rng = np.random.default_rng(5)
    n = 2000
    d = rng.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1)[:, None]
    r = rng.uniform(0.6 * R, 0.98 * R, n)
    pts = np.array(center) + d * r[:, None]
    outside = np.array(center) + d * (R + 0.03)
    xs, ys, zs = orc.soa(np.vstack([pts, outside]))
    before = np.stack([xs.copy(), ys.copy(), zs.copy()], 1)
    L.orc_collide(2 * n, xs, ys, zs, C.byref(o))
    after = np.stack([xs, ys, zs], 1)
    rr = np.linalg.norm(after[:n] - np.array(center), axis=1)
    # trilinear SDF of a sphere: error O(dx^2/R) plus the 1-texel normal stencil
    assert np.max(np.abs(rr - R)) < 0.2 * dx
    radial = (after[:n] - before[:n])
    cosang = np.sum(radial * d, 1) / np.linalg.norm(radial, axis=1)
    assert np.min(cosang) > 0.999
    assert np.array_equal(after[n:], before[n:])


def test_cohesion_spline_values():
    h, norm = 0.04, 1.0
    for r, e in ((0.25 * h, 2 * (0.75 * h) ** 3 * (0.25 * h) ** 3 - h ** 6 / 64),
                 (0.5 * h, h ** 6 / 64), (0.75 * h, (0.25 * h) ** 3 * (0.75 * h) ** 3),
                 (h, 0.0), (1.01 * h, 0.0)):
        assert abs(L.orc_C(r, h, norm) - e) < 1e-12 * h ** 6
    # continuous at h/2 from both sides
    a = L.orc_C(0.5 * h * (1 - 1e-12), h, norm)
    b = L.orc_C(0.5 * h * (1 + 1e-12), h, norm)
    assert abs(a - b) < 1e-9 * h ** 6


def _jittered(seed, n=(10, 9, 8), jitter=0.3):
    s = 0.02
    p = orc.make_params(s, (0.5, 0.5, 0.5), k=0.5, c=0.1, vort_eps=0.01)
    pos = orc.lattice_block(*n, s, origin=(0.05, 0.05, 0.05), jitter=jitter, seed=seed)
    xs, ys, zs = orc.soa(pos.astype(np.float64))
    nbr, cnt, dropped = orc.neighbors(p, xs, ys, zs)
    assert dropped == 0
    return p, xs, ys, zs, nbr, cnt


def test_position_correction_is_antisymmetric():
    p, xs, ys, zs, nbr, cnt = _jittered(6)
    n = len(xs)
    rho = np.zeros(n); lam = np.zeros(n)
    L.orc_compute_densities(n, xs, ys, zs, nbr, cnt, C.byref(p), rho)
    L.orc_compute_lambda(n, xs, ys, zs, rho, nbr, cnt, C.byref(p), lam)
    d = [np.zeros(n) for _ in range(3)]
    L.orc_update_dps(n, xs, ys, zs, lam, nbr, cnt, C.byref(p), 1, *d)
    for a in d:
        assert abs(a.sum()) < 1e-9 * np.abs(a).sum()
    # lists are symmetric
    sets = orc.neighbor_sets(nbr, cnt)
    for i in range(0, n, 13):
        for j in sets[i]:
            assert i in sets[j]


def test_galilean_invariance_of_velocity_passes():
    p, xs, ys, zs, nbr, cnt = _jittered(7)
    n = len(xs)
    rho = np.zeros(n)
    L.orc_compute_densities(n, xs, ys, zs, nbr, cnt, C.byref(p), rho)
    rng = np.random.default_rng(8)
    v = [rng.normal(0, 0.2, n) for _ in range(3)]
    shift = (0.7, -1.3, 2.1)
    vs = [v[i] + shift[i] for i in range(3)]
    o1 = [np.zeros(n) for _ in range(3)]; o2 = [np.zeros(n) for _ in range(3)]
    L.orc_apply_viscosity(n, xs, ys, zs, rho, *v, nbr, cnt, C.byref(p), *o1)
    L.orc_apply_viscosity(n, xs, ys, zs, rho, *vs, nbr, cnt, C.byref(p), *o2)
    for i in range(3):
        assert np.allclose(o2[i] - shift[i], o1[i], atol=1e-12)
    w1 = [np.zeros(n) for _ in range(3)]; w2 = [np.zeros(n) for _ in range(3)]
    L.orc_compute_vorticity(n, xs, ys, zs, rho, *v, nbr, cnt, C.byref(p), *w1)
    L.orc_compute_vorticity(n, xs, ys, zs, rho, *vs, nbr, cnt, C.byref(p), *w2)
    scale = max(np.abs(w1[i]).max() for i in range(3))
    for i in range(3):
        assert np.allclose(w1[i], w2[i], atol=1e-9 * scale)


def _del_w(r, h):
    """Spiky gradient magnitude DelW(r) = -45/(pi h^6) (h - r)^2 (particles.c:40-46)."""
    return -45.0 / (math.pi * h ** 6) * (h - r) ** 2 if r < h else 0.0


def test_rigid_rotation_vorticity_closed_form():
    # v = Omega x r on a perfect lattice.  With d = x_i - x_j, each term is
    # w_j (v_j - v_i) x DelW(r) d/r = w_j DelW(r)/r (r^2 Omega - d (Omega.d)), and
    # cubic symmetry turns sum f(r) d d^T into (1/3) sum f(r) r^2 I, so
    # omega_i = (2/3) sum_k w DelW(s|k|) s|k| Omega over lattice offsets k with
    # |k|^2 < 4 (h = 2s), w = 1 (vort_model 0, particles.c:207-283) or m/rho
    # with rho = m W_norm s^6 * 330 (model 1).
    p, xs, ys, zs, nbr, cnt = _jittered(9, n=(12, 12, 12), jitter=0.0)
    n = len(xs)
    s, h, m = p.h / 2.0, p.h, p.rest_mass
    rho = np.zeros(n)
    L.orc_compute_densities(n, xs, ys, zs, nbr, cnt, C.byref(p), rho)
    om = np.array([0.3, -0.2, 1.1])
    r = np.stack([xs, ys, zs], 1)
    v = np.cross(om, r)
    vx, vy, vz = [np.ascontiguousarray(v[:, i]) for i in range(3)]
    idx = np.arange(n).reshape(12, 12, 12)[4:8, 4:8, 4:8].ravel()
    lat = sum(_del_w(s * math.sqrt(k[0] ** 2 + k[1] ** 2 + k[2] ** 2), h) *
              s * math.sqrt(k[0] ** 2 + k[1] ** 2 + k[2] ** 2) for k in _lattice_offsets(4))
    rho_lat = m * 315.0 / (64.0 * math.pi * h ** 9) * s ** 6 * 330.0
    for model, w in ((0, 1.0), (1, m / rho_lat)):
        p.vort_model = model
        out = [np.zeros(n) for _ in range(3)]
        L.orc_compute_vorticity(n, xs, ys, zs, rho, vx, vy, vz, nbr, cnt, C.byref(p), *out)
        W = np.stack(out, 1)[idx]
        expected = (2.0 / 3.0) * w * lat * om
        # float32 lattice positions: ~1e-7 relative asymmetry per pair
        assert np.max(np.abs(W - expected)) < 1e-5 * np.linalg.norm(expected), model
    # model 1 approximates the curl of the rotation, 2 Omega, to the lattice's
    # discretisation error (0.947 of it here; antiparallel, because the
    # reference's gradient is taken along x_i - x_j)
    assert abs((2.0 / 3.0) * (m / rho_lat) * lat / -2.0 - 1.0) < 0.1


def test_vorticity_two_particles_closed_form():
    # Two particles along x at r = 0.6 h, v_j - v_i = (0, 0, w): omega_i =
    # (0, -w DelW(r) w_j, 0) and omega_j = (0, -w DelW(r) w_i, 0), with
    # w_q = 1 (model 0) or m / rho_q (model 1).  Different densities at i and
    # j pin that the weight is the neighbour's.
    s = 0.02
    p = orc.make_params(s, (0.5, 0.5, 0.5), vort_eps=0.01)
    h, m = p.h, p.rest_mass
    r = 0.6 * h
    xs = np.array([0.2, 0.2 + r]); ys = np.array([0.2, 0.2]); zs = np.array([0.2, 0.2])
    vx = np.array([0.5, 0.5]); vy = np.array([-0.1, -0.1]); vz = np.array([0.0, 0.7])
    rho = np.array([1200.0, 800.0])
    nbr = np.zeros(2 * orc.MAXN, np.uint32)
    nbr[0], nbr[orc.MAXN] = 1, 0
    cnt = np.array([1, 1], np.int32)
    dw = _del_w(r, h)
    for model in (0, 1):
        p.vort_model = model
        w = [np.zeros(2) for _ in range(3)]
        L.orc_compute_vorticity(2, xs, ys, zs, rho, vx, vy, vz, nbr, cnt, C.byref(p), *w)
        wi = 1.0 if model == 0 else m / rho[1]
        wj = 1.0 if model == 0 else m / rho[0]
        assert abs(w[1][0] - (-0.7 * dw * wi)) < 1e-12 * abs(0.7 * dw * wi)
        assert abs(w[1][1] - (-0.7 * dw * wj)) < 1e-12 * abs(0.7 * dw * wj)
        for a in (0, 2):
            assert w[a][0] == 0.0 and w[a][1] == 0.0


def test_confinement_three_particles_closed_form():
    # i at the origin of an L: j at +r x, k at +r y (r = 0.6 h), with
    # |omega_j| = 3, |omega_k| = 1, omega_i = (0, 0, 2).  eta_i =
    # sum_q |omega_q| DelW(r) (x_i - x_q)/r = |DelW(r)| (3, 1, 0) (DelW < 0),
    # so N_i = (3, 1, 0)/sqrt(10) and N x omega_i = (2, -6, 0)/sqrt(10):
    # dv_i = dt eps (2, -6, 0)/sqrt(10) / m for model 0 (particles.c:375-379)
    # and without the 1/m for model 1.  Using |omega_i| for eta would give
    # N = (1, 1, 0)/sqrt(2), a missing or extra 1/m a factor m.
    s = 0.02
    p = orc.make_params(s, (0.5, 0.5, 0.5), vort_eps=0.01)
    h, m, dt, eps = p.h, p.rest_mass, p.dt, p.vort_eps
    r = 0.6 * h
    xs = np.array([0.2, 0.2 + r, 0.2]); ys = np.array([0.2, 0.2, 0.2 + r])
    zs = np.array([0.2, 0.2, 0.2])
    wx = np.array([0.0, 0.0, 1.0]); wy = np.array([0.0, 0.0, 0.0]); wz = np.array([2.0, -3.0, 0.0])
    nbr = np.zeros(3 * orc.MAXN, np.uint32)
    for i, js in enumerate(([1, 2], [0, 2], [0, 1])):
        nbr[i * orc.MAXN:i * orc.MAXN + 2] = js
    cnt = np.array([2, 2, 2], np.int32)
    v0 = (np.array([0.1, 0.0, 0.0]), np.array([0.2, 0.0, 0.0]), np.array([0.3, 0.0, 0.0]))
    for model, mass_p in ((0, m), (1, 1.0)):
        p.vort_model = model
        v = [a.copy() for a in v0]
        L.orc_apply_vorticity_confinement(3, xs, ys, zs, wx, wy, wz, nbr, cnt, C.byref(p), *v)
        k = dt * eps / math.sqrt(10.0) / mass_p
        expected = (2.0 * k, -6.0 * k, 0.0)
        for a in range(3):
            assert abs((v[a][0] - v0[a][0]) - expected[a]) < 1e-12 * 6.0 * k, (model, a)


def test_max_density_error_compression_only():
    d = np.array([990.0, 1000.0, 1003.0, 1001.0])
    assert abs(L.orc_max_density_error(4, d, 1000.0) - 0.003) < 1e-15
    d = np.array([990.0, 999.0])
    assert L.orc_max_density_error(2, d, 1000.0) == 0.0
