"""The list build's candidate test, emulated in FP32 on the CPU (no GPU).

k_build_lists computes r2 = fma(dz, dz, fma(dy, dy, dx * dx)) in FP32 from
the FP32 differences of the positions and compares it with the guard band
h^2 (1 -+ 1e-5) (h2_lo, h2_hi rounded to FP32 as sph_api.hip derives them):
r2 < h2_lo is a neighbour, r2 >= h2_hi is not, and only the band between goes
to the FP64 test that the reference's list follows (neighbors.c:248-290 with
r^2 in double, inside_exact in the kernel).

Checked here on pairs at distances within a few 1e-5 of h, with positions up
to 6 m from the origin and h from 5 mm to 5 cm:
* the FP32 error of r2 against the exact rational r^2 stays below
  8 2^-24 h^2, far inside the 1e-5 h^2 half-width of the band;
* r2 < h2_lo implies the FP64 test passes and r2 >= h2_hi that it fails, so
  the fast decision never differs from the exact one;
* with the band removed (h2_lo = h2_hi = h^2) the same pairs are misclassified,
  so the check can fail.

FP32 fused multiply-adds are emulated as an exact double product plus an
FP32 addend, rounded once to double and once to FP32; that double rounding can
differ from the hardware's single rounding by one FP32 ulp in rare halfway
cases, 1e-7 h^2 here, far inside the margins tested."""
from fractions import Fraction

import numpy as np
import pytest

F32 = np.float32


def fma32(a, b, c):
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(F32)


def r2_fp32(p, q):
    dx, dy, dz = (p[:, k] - q[:, k] for k in range(3))  # FP32 differences
    return fma32(dz, dz, fma32(dy, dy, dx * dx))


def pairs(h, n, spread, seed):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-6.0, 6.0, size=(n, 3)).astype(F32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = h * (1.0 + rng.uniform(-spread, spread, size=n))
    q = (p.astype(np.float64) + d * r[:, None]).astype(F32)
    return p, q


def exact_r2(p, q):
    return [sum((Fraction(float(a[k])) - Fraction(float(b[k]))) ** 2 for k in range(3))
            for a, b in zip(p, q)]


def fp64_inside(p, q, h):
    d = p.astype(np.float64) - q.astype(np.float64)
    r2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
    return r2 < h * h


@pytest.mark.parametrize("h", [0.005, 0.0125, 0.02, 0.05])
def test_fast_decision_matches_fp64_test(h):
    h2_lo = F32(h * h * (1.0 - 1e-5))
    h2_hi = F32(h * h * (1.0 + 1e-5))
    p, q = pairs(h, 4000, 4e-5, seed=int(h * 1e4))
    r2 = r2_fp32(p, q)
    inside = fp64_inside(p, q, h)
    lo, hi = r2 < h2_lo, r2 < h2_hi
    assert np.all(inside[lo]), "a pair below the band fails the FP64 test"
    assert not np.any(inside[~hi]), "a pair above the band passes the FP64 test"
    band = hi & ~lo
    assert 0 < band.sum() < len(band)  # the band holds only the pairs near h
    ex = exact_r2(p[:500], q[:500])
    err = max(abs(Fraction(float(a)) - e) for a, e in zip(r2[:500], ex))
    assert err <= Fraction(8, 2 ** 24) * Fraction(h) ** 2, float(err / Fraction(h) ** 2)


def test_without_guard_band_pairs_are_misclassified():
    h = 0.02
    h2 = F32(h * h)
    p, q = pairs(h, 20000, 2e-6, seed=7)
    inside = fp64_inside(p, q, h)
    assert np.any((r2_fp32(p, q) < h2) != inside), "no pair flips: the band would not be needed"
