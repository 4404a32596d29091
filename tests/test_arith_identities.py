"""Arithmetic identities the velocity kernels rely on (k_resample_gradient32 in
niftyreg_amd/csrc/f3d_velocity.hip), checked on the CPU.

* The forward-difference Jacobian of reg_trilinearResampleGradient
  (reg-lib/cpu/_reg_resampling.cpp:1547-1766) is a weighted sum over eight (c, b, a) corner pairs
  whose weights are products of {1, 0} bases and {-1, 1} derivatives; on an axis's last index the
  difference steps back and the other axes' differences are read one voxel back.  Summed in the
  reference's order in float32 it equals, value for value, the plain forward differences at the
  corner (x - ax, y - ay, z - az).
* Markstein's correction: with rsp = float(1 / sp) and q = float(a * rsp), the residual
  r = fma(-q, sp, a) is exact and fma(r, rsp, q) is the correctly rounded quotient a / sp.
"""
import itertools
from fractions import Fraction as F

import numpy as np

f32 = np.float32


def _pair_sum_jacobian(d, x, y, z, nx, ny, nz):
    """The reference's eight-pair sums (weights as reg_trilinearResampleGradient forms them)."""
    ax, ay, az = int(x == nx - 1), int(y == ny - 1), int(z == nz - 1)
    J = np.zeros((3, 3), f32)
    for c, b, a in itertools.product((0, 1), (0, 1), (0, 1)):
        bx = (f32(a), f32(-1 if a else 1)) if ax else (f32(1 - a), f32(1 if a else -1))
        by = (f32(b), f32(-1 if b else 1)) if ay else (f32(1 - b), f32(1 if b else -1))
        bz = (f32(c), f32(-1 if c else 1)) if az else (f32(1 - c), f32(1 if c else -1))
        w = (f32(f32(bx[1] * by[0]) * bz[0]), f32(f32(bx[0] * by[1]) * bz[0]), f32(f32(bx[0] * by[0]) * bz[1]))
        X = x + a - (2 if ax and a else 0)
        Y = y + b - (2 if ay and b else 0)
        Z = z + c - (2 if az and c else 0)
        for m in range(3):
            for k in range(3):
                J[m, k] = f32(J[m, k] + f32(w[k] * d[m, Z, Y, X]))
    return J


def _forward_difference_jacobian(d, x, y, z, nx, ny, nz):
    X0, Y0, Z0 = x - (x == nx - 1), y - (y == ny - 1), z - (z == nz - 1)
    J = np.zeros((3, 3), f32)
    for m in range(3):
        J[m, 0] = f32(d[m, Z0, Y0, X0 + 1] - d[m, Z0, Y0, X0])
        J[m, 1] = f32(d[m, Z0, Y0 + 1, X0] - d[m, Z0, Y0, X0])
        J[m, 2] = f32(d[m, Z0 + 1, Y0, X0] - d[m, Z0, Y0, X0])
    return J


def test_jacobian_pair_sums_are_forward_differences():
    rng = np.random.default_rng(0)
    checked = 0
    for trial in range(40):
        nx, ny, nz = (int(v) for v in rng.integers(2, 5, 3))
        d = (rng.standard_normal((3, nz, ny, nx)) * rng.choice([1e-3, 1.0, 100.0])).astype(f32)
        if trial % 3 == 0:
            d[rng.random(d.shape) < 0.3] = 0.0
        if trial % 5 == 0:
            d[rng.random(d.shape) < 0.3] = -0.0
        for z, y, x in itertools.product(range(nz), range(ny), range(nx)):
            want = _pair_sum_jacobian(d, x, y, z, nx, ny, nz)
            got = _forward_difference_jacobian(d, x, y, z, nx, ny, nz)
            assert np.array_equal(got, want), (x, y, z, nx, ny, nz)
            checked += 1
    assert checked > 500


def _rn32(v):
    """The float32 nearest to the rational v (ties to even)."""
    f = f32(float(v))
    best = None
    for c in (np.nextafter(f, f32(-np.inf)), f, np.nextafter(f, f32(np.inf))):
        dist = abs(F(float(c)) - v)
        even = (int(np.array(c, f32).view(np.uint32)) & 1) == 0
        if best is None or dist < best[0] or (dist == best[0] and even):
            best = (dist, f32(c))
    return best[1]


def _fma32(a, b, c):
    return _rn32(F(float(a)) * F(float(b)) + F(float(c)))


def test_markstein_division_is_correctly_rounded():
    rng = np.random.default_rng(1)
    spacings = [1.0, 0.5, 1.2, 0.9, 2.0, 3.0, 0.9765625, 1.1, 0.3] + list(rng.uniform(0.1, 10, 6))
    for s in spacings:
        s = f32(s)
        rsp = _rn32(F(1) / F(float(s)))
        xs = (rng.standard_normal(200) * 10.0 ** rng.integers(-6, 6, 200)).astype(f32)
        for a in xs:
            q = f32(a * rsp)
            r = _fma32(-q, s, a)
            assert F(float(r)) == F(float(a)) - F(float(q)) * F(float(s))  # the residual is exact
            assert _fma32(r, rsp, q) == _rn32(F(float(a)) / F(float(s))), (s, a)
