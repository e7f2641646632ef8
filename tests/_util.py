"""Shared helpers for the tests: seeded inputs and small grid utilities."""
import numpy as np


def set_fine(cells, R, x, y, z):
    cells[(x >> 1) + R * ((y >> 1) + R * (z >> 1))] |= np.uint8(1 << ((x & 1) | ((y & 1) << 1) | ((z & 1) << 2)))


def fine_grid(cells, R):
    """cell bytes -> dense bool array [z, y, x] of fine voxels."""
    c = cells.reshape(R, R, R)  # [cz, cy, cx]
    N = 2 * R
    out = np.zeros((N, N, N), bool)
    for k in range(8):
        out[(k >> 2) & 1::2, (k >> 1) & 1::2, k & 1::2] = (c >> k) & 1
    return out


def cells_from_fine(f):
    N = f.shape[0]
    R = N // 2
    c = np.zeros((R, R, R), np.uint8)
    for k in range(8):
        c |= (f[(k >> 2) & 1::2, (k >> 1) & 1::2, k & 1::2].astype(np.uint8) << k)
    return c.reshape(-1)


def random_blobs(R, seed, nballs=6, rmin=2.0, rmax=6.0, margin=6):
    """Union of random solid balls (fine voxels) kept `margin` cells inside."""
    rng = np.random.default_rng(seed)
    N = 2 * R
    z, y, x = np.mgrid[0:N, 0:N, 0:N]
    f = np.zeros((N, N, N), bool)
    lo, hi = 2 * margin + rmax, N - 2 * margin - rmax
    for _ in range(nballs):
        c = rng.uniform(lo, hi, 3)
        r = rng.uniform(rmin, rmax)
        f |= (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 < r * r
    return cells_from_fine(f)


def random_sparse(R, seed, density=0.01, margin=2):
    rng = np.random.default_rng(seed)
    N = 2 * R
    f = rng.random((N, N, N)) < density
    m = 2 * margin
    f[:m] = f[-m:] = False
    f[:, :m] = f[:, -m:] = False
    f[:, :, :m] = f[:, :, -m:] = False
    return cells_from_fine(f)
