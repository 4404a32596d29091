"""Deterministic synthetic image pairs for the measured workloads (numpy, CPU).

* blob_phantom: sum of 8 Gaussian blobs (centres uniform in the volume, amplitudes U[0.2, 1],
  sigma U[4, 10] voxels) plus N(0, 0.01) noise.
* shell_phantom: three nested ellipsoidal shells (0.2 / 0.5 / 0.8, radii jittered by 5%), a
  second-order polynomial multiplicative bias of +-15% and N(0, 0.02) noise.
* bspline_warp: the image resampled (trilinear, zero outside) through a random cubic B-spline
  displacement with the given control spacing (voxels) and U[-a, a] voxel displacements.
* affine_warp: the image resampled through a random affine (rotation, translation, scale, shear).

Images are float32 arrays indexed [z, y, x] (x fastest in memory), 1 mm isotropic, identity
orientation.
"""
import numpy as np
from scipy import ndimage


def _rng(seed):
    return np.random.default_rng(np.random.SeedSequence(int(seed)))


def blob_phantom(dims, seed=0, blobs=8, noise=0.01):
    nx, ny, nz = dims
    r = _rng(seed)
    z, y, x = np.meshgrid(np.arange(nz, dtype=np.float32), np.arange(ny, dtype=np.float32),
                          np.arange(nx, dtype=np.float32), indexing="ij")
    img = np.zeros((nz, ny, nx), np.float32)
    for _ in range(blobs):
        c = r.uniform(0, 1, 3) * np.array([nx - 1, ny - 1, nz - 1])
        a = r.uniform(0.2, 1.0)
        s = r.uniform(4.0, 10.0)
        d2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
        img += np.float32(a) * np.exp(-d2 / np.float32(2 * s * s))
    img += r.normal(0.0, noise, img.shape).astype(np.float32)
    return img.astype(np.float32)


def shell_phantom(dims, seed=0, noise=0.02):
    nx, ny, nz = dims
    r = _rng(seed)
    z, y, x = np.meshgrid(np.linspace(-1, 1, nz, dtype=np.float32), np.linspace(-1, 1, ny, dtype=np.float32),
                          np.linspace(-1, 1, nx, dtype=np.float32), indexing="ij")
    img = np.zeros((nz, ny, nx), np.float32)
    for radius, val in ((0.85, 0.2), (0.6, 0.5), (0.3, 0.8)):
        rx, ry, rz = radius * r.uniform(0.95, 1.05, 3)
        inside = (x / rx) ** 2 + (y / ry) ** 2 + (z / rz) ** 2 <= 1.0
        img[inside] = val
    # second-order polynomial bias scaled to +-15%
    coef = r.uniform(-1, 1, 9)
    terms = [x, y, z, x * x, y * y, z * z, x * y, y * z, x * z]
    poly = sum(np.float32(c) * t for c, t in zip(coef, terms))
    poly = poly / max(float(np.abs(poly).max()), 1e-6)
    img *= (1.0 + 0.15 * poly).astype(np.float32)
    img += r.normal(0.0, noise, img.shape).astype(np.float32)
    return img.astype(np.float32)


def _bspline_axis(n, spacing):
    """Dense (n x nodes) matrix of cubic B-spline weights, node q at voxel spacing*(q-1)."""
    nodes = int(np.ceil(n / spacing + 3))
    W = np.zeros((n, nodes), np.float64)
    for v in range(n):
        pre = int(v // spacing)
        t = v / spacing - pre
        b = [(1 - t) ** 3 / 6, (3 * t ** 3 - 6 * t ** 2 + 4) / 6, (-3 * t ** 3 + 3 * t ** 2 + 3 * t + 1) / 6, t ** 3 / 6]
        for a in range(4):
            W[v, pre + a] = b[a]
    return W


def random_bspline_displacement(dims, spacing, amplitude, seed):
    """Dense displacement (3, nz, ny, nx) in voxels of a random cubic B-spline grid."""
    nx, ny, nz = dims
    r = _rng(seed)
    Wx, Wy, Wz = _bspline_axis(nx, spacing), _bspline_axis(ny, spacing), _bspline_axis(nz, spacing)
    disp = np.empty((3, nz, ny, nx), np.float32)
    for c in range(3):
        cp = r.uniform(-amplitude, amplitude, (Wz.shape[1], Wy.shape[1], Wx.shape[1]))
        t = np.einsum("kji,xi->kjx", cp, Wx)
        t = np.einsum("kjx,yj->kyx", t, Wy)
        disp[c] = np.einsum("kyx,zk->zyx", t, Wz).astype(np.float32)
    return disp


def warp(img, disp):
    """Trilinear resampling of img at (x + dx, y + dy, z + dz), zero outside."""
    nz, ny, nx = img.shape
    z, y, x = np.meshgrid(np.arange(nz, dtype=np.float32), np.arange(ny, dtype=np.float32),
                          np.arange(nx, dtype=np.float32), indexing="ij")
    coords = np.stack([z + disp[2], y + disp[1], x + disp[0]])
    return ndimage.map_coordinates(img, coords, order=1, mode="constant", cval=0.0).astype(np.float32)


def bspline_warp(img, spacing, amplitude, seed):
    nz, ny, nx = img.shape
    disp = random_bspline_displacement((nx, ny, nz), spacing, amplitude, seed)
    return warp(img, disp)


def random_affine(seed, rot_deg=15.0, trans_mm=10.0, scale=(0.9, 1.1), shear=0.05):
    """4x4 affine (mm) = T * R * S * Sh with parameters drawn uniformly."""
    r = _rng(seed)
    ax, ay, az = np.deg2rad(r.uniform(-rot_deg, rot_deg, 3))
    Rx = np.array([[1, 0, 0], [0, np.cos(ax), -np.sin(ax)], [0, np.sin(ax), np.cos(ax)]])
    Ry = np.array([[np.cos(ay), 0, np.sin(ay)], [0, 1, 0], [-np.sin(ay), 0, np.cos(ay)]])
    Rz = np.array([[np.cos(az), -np.sin(az), 0], [np.sin(az), np.cos(az), 0], [0, 0, 1]])
    S = np.diag(r.uniform(scale[0], scale[1], 3))
    Sh = np.eye(3)
    Sh[0, 1], Sh[0, 2], Sh[1, 2] = r.uniform(-shear, shear, 3)
    A = np.eye(4)
    A[:3, :3] = Rz @ Ry @ Rx @ S @ Sh
    A[:3, 3] = r.uniform(-trans_mm, trans_mm, 3)
    return A


def affine_warp(img, affine_mm):
    """img resampled at A * p about the image centre (identity orientation, 1 mm)."""
    nz, ny, nx = img.shape
    centre = np.array([(nx - 1) / 2, (ny - 1) / 2, (nz - 1) / 2])
    A = np.asarray(affine_mm, np.float64)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    p = np.stack([x - centre[0], y - centre[1], z - centre[2]]).reshape(3, -1)
    q = A[:3, :3] @ p + A[:3, 3:4]
    q = q + centre[:, None]
    coords = np.stack([q[2], q[1], q[0]]).reshape(3, nz, ny, nx)
    return ndimage.map_coordinates(img, coords, order=1, mode="constant", cval=0.0).astype(np.float32)


def config0_pair(seed=0, dims=(64, 64, 64)):
    """Blob phantom and its B-spline-warped copy (5-voxel spacing, U[-2, 2] voxels)."""
    ref = blob_phantom(dims, seed)
    flo = bspline_warp(ref, 5, 2.0, seed + 1000)
    return ref, flo


def config1_pair(seed=0, dims=(256, 256, 160)):
    """Shell phantom and its B-spline-warped copy (5-voxel spacing, U[-4, 4] voxels)."""
    ref = shell_phantom(dims, seed)
    flo = bspline_warp(ref, 5, 4.0, seed + 1000)
    return ref, flo


def config3_pair(seed=0, dims=(128, 128, 128)):
    """Blob phantom and an affinely warped copy (rotations within 15 degrees per axis,
    translations within 10 mm, scales 0.9-1.1, shears within 0.05).  Returns ref, flo, A."""
    ref = blob_phantom(dims, seed)
    A = random_affine(seed + 2000)
    return ref, affine_warp(ref, A), A


def config2_dims(seed):
    """Axis sizes of a batched-F3D pair: each drawn independently from U{96..224}."""
    r = _rng(seed)
    return tuple(int(v) for v in r.integers(96, 225, 3))


def config2_pair(seed):
    """A batched-F3D pair: its axis sizes from config2_dims, content as config1_pair."""
    return config1_pair(seed=seed, dims=config2_dims(seed))


def config4_request(seed):
    """A serving request: a cubic volume whose voxel count is log-uniform between 64^3 and 256^3,
    a shell phantom, and its copy under a random affine (within 5 degrees, 3 mm, 5% scale) and a
    B-spline warp (5-voxel spacing, U[-4, 4] voxels); the measure is NMI or LNCC with equal
    probability.  Returns (dims, ref, flo, measure) with measure 0 (NMI) or 1 (LNCC)."""
    r = _rng(seed)
    side = int(round(np.exp(r.uniform(np.log(64), np.log(256)))))
    dims = (side, side, side)
    ref = shell_phantom(dims, seed)
    A = random_affine(seed + 3000, rot_deg=5.0, trans_mm=3.0, scale=(0.95, 1.05), shear=0.02)
    flo = bspline_warp(affine_warp(ref, A), 5, 4.0, seed + 4000)
    return dims, ref, flo, int(r.integers(0, 2))
