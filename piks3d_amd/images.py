"""Seeded porous-image generators and the raw image format.

The solver reads a binary voxel image: one uint8 per voxel, 0 = fluid,
1 = solid, x fastest, then y, then z, with no header (reference:
src/physicalGrid.F90:257-280 reads it as a Fortran stream; conf/createSphere.F90:43-51
writes it).  Arrays here are numpy arrays of shape (nz, ny, nx) in C order, so
`arr.tofile(path)` writes exactly that byte stream.

Generators (all seeded with numpy's PCG64):

* ``centered_sphere``  -- one solid sphere, voxel (i, j, k) (1-based) solid when
  (i - c)^2 + (j - c)^2 + (k - c)^2 < r^2 with c = (n + 1) / 2, the convention
  of conf/createSphere.F90:13-41.
* ``quarter_sphere``   -- createSphere's own image: centre ((nx + 1)/2, 1, 1).
* ``sphere_pack``      -- overlapping solid spheres, centres uniform in the
  box, radii uniform in [rmin, rmax], added until the porosity drops to the
  target.
* ``gaussian_field``   -- Gaussian random field with Gaussian covariance
  exp(-r^2 / l^2) (white noise filtered in Fourier space), thresholded at the
  porosity quantile, keeping only pore clusters that connect x = 0 to x = n-1.
* ``box_blocks``       -- random axis-aligned solid boxes (used by the tests to
  produce every wall orientation, edge and corner type).
"""
from __future__ import annotations

import numpy as np


def write_raw(path: str, img: np.ndarray) -> None:
    np.ascontiguousarray(img, dtype=np.uint8).tofile(path)


def read_raw(path: str, nx: int, ny: int, nz: int) -> np.ndarray:
    data = np.fromfile(path, dtype=np.uint8)
    if data.size < nx * ny * nz:
        raise ValueError("image file %s has %d bytes, need %d" % (path, data.size, nx * ny * nz))
    return data[: nx * ny * nz].reshape(nz, ny, nx)


def centered_sphere(n: int, radius: float) -> np.ndarray:
    idx = np.arange(1, n + 1, dtype=np.float64) - (n + 1) / 2.0
    d2 = idx[:, None, None] ** 2 + idx[None, :, None] ** 2 + idx[None, None, :] ** 2
    return (d2 < radius * radius).astype(np.uint8)


def quarter_sphere(nx: int, ny: int, nz: int, radius: float) -> np.ndarray:
    i = np.arange(1, nx + 1, dtype=np.float64) - (nx + 1) / 2.0
    j = np.arange(1, ny + 1, dtype=np.float64) - 1.0
    k = np.arange(1, nz + 1, dtype=np.float64) - 1.0
    d2 = k[:, None, None] ** 2 + j[None, :, None] ** 2 + i[None, None, :] ** 2
    return (d2 < radius * radius).astype(np.uint8)


def sphere_pack(n: int, porosity: float, rmin: float, rmax: float, seed: int,
                shape=None) -> np.ndarray:
    """Solid spheres added one at a time until porosity <= target."""
    nz, ny, nx = shape if shape is not None else (n, n, n)
    rng = np.random.Generator(np.random.PCG64(seed))
    img = np.zeros((nz, ny, nx), dtype=np.uint8)
    total = img.size
    solid = 0
    target_solid = (1.0 - porosity) * total
    while solid < target_solid:
        cx, cy, cz = rng.uniform(0, nx), rng.uniform(0, ny), rng.uniform(0, nz)
        r = rng.uniform(rmin, rmax)
        x0, x1 = max(int(np.floor(cx - r)), 0), min(int(np.ceil(cx + r)) + 1, nx)
        y0, y1 = max(int(np.floor(cy - r)), 0), min(int(np.ceil(cy + r)) + 1, ny)
        z0, z1 = max(int(np.floor(cz - r)), 0), min(int(np.ceil(cz + r)) + 1, nz)
        xs = np.arange(x0, x1) + 0.5 - cx
        ys = np.arange(y0, y1) + 0.5 - cy
        zs = np.arange(z0, z1) + 0.5 - cz
        ball = (zs[:, None, None] ** 2 + ys[None, :, None] ** 2 + xs[None, None, :] ** 2) <= r * r
        sub = img[z0:z1, y0:y1, x0:x1]
        solid += int(np.count_nonzero(ball & (sub == 0)))
        sub[ball] = 1
    return img


def gaussian_field(n: int, porosity: float, corr_len: float, seed: int,
                   percolate_x: bool = True, shape=None) -> np.ndarray:
    nz, ny, nx = shape if shape is not None else (n, n, n)
    rng = np.random.Generator(np.random.PCG64(seed))
    noise = rng.standard_normal((nz, ny, nx))
    kz = np.fft.fftfreq(nz)[:, None, None]
    ky = np.fft.fftfreq(ny)[None, :, None]
    kx = np.fft.rfftfreq(nx)[None, None, :]
    # covariance exp(-r^2/l^2)  <=>  filter exp(-(pi l k)^2 / 2) on the noise
    filt = np.exp(-0.5 * (np.pi * corr_len) ** 2 * (kx ** 2 + ky ** 2 + kz ** 2))
    field = np.fft.irfftn(np.fft.rfftn(noise) * filt, s=(nz, ny, nx), axes=(0, 1, 2))
    thr = np.quantile(field, porosity)
    pore = field < thr
    if percolate_x:
        from scipy import ndimage
        lab, nlab = ndimage.label(pore)
        left = set(np.unique(lab[:, :, 0])) - {0}
        right = set(np.unique(lab[:, :, -1])) - {0}
        keep = np.array(sorted(left & right), dtype=lab.dtype)
        pore = np.isin(lab, keep)
    return (~pore).astype(np.uint8)


def box_blocks(shape, nboxes: int, smin: int, smax: int, seed: int) -> np.ndarray:
    nz, ny, nx = shape
    rng = np.random.Generator(np.random.PCG64(seed))
    img = np.zeros(shape, dtype=np.uint8)
    for _ in range(nboxes):
        sx, sy, sz = rng.integers(smin, smax + 1, size=3)
        x0 = rng.integers(0, max(nx - sx, 0) + 1)
        y0 = rng.integers(0, max(ny - sy, 0) + 1)
        z0 = rng.integers(0, max(nz - sz, 0) + 1)
        img[z0:z0 + sz, y0:y0 + sy, x0:x0 + sx] = 1
    return img


def porosity(img: np.ndarray) -> float:
    return 1.0 - float(np.count_nonzero(img)) / img.size
