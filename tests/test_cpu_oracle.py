"""CPU tests: the oracle pinned against brute force, closed forms, invariants
and the reference's fixtures; host-side generators; the C-ABI export list."""
import os
import re
from collections import Counter

import numpy as np
import pytest

import _oracle as O
import broxy_amd as B
from _util import cells_from_fine, fine_grid, random_blobs, random_sparse, set_fine

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def morton(x, y, z):
    m = 0
    for i in range(10):
        m |= ((x >> i) & 1) << (3 * i) | ((y >> i) & 1) << (3 * i + 1) | ((z >> i) & 1) << (3 * i + 2)
    return m


# ---------------------------------------------------------------- dilation
@pytest.mark.parametrize("seed", range(4))
def test_dilate_dfs_equals_brute(seed):
    R = 16
    rng = np.random.default_rng(seed)
    I = random_sparse(R, seed, density=0.01, margin=0)
    r = int(rng.integers(2, 7))
    assert np.array_equal(O.dilate(I, R, r=r), O.dilate(I, R, r=r, brute=1))
    assert np.array_equal(O.dilate(I, R, r=r, se=1), O.dilate(I, R, r=r, se=1, brute=1))
    sc = rng.integers(-4, 9, (2 * R) ** 3).astype(np.int8)
    assert np.array_equal(O.dilate(I, R, scale=sc), O.dilate(I, R, scale=sc, brute=1))


@pytest.mark.parametrize("r", [2, 3, 4, 6])
def test_single_voxel_sphere_is_digital_ball(r):
    # SqDist(p, [q,q+1]^3) < (r+1)^2 per axis distance d(t) = t (t>=0), -t-1 (t<0)
    R = 16
    q = (15, 16, 15)
    I = np.zeros(R ** 3, np.uint8)
    set_fine(I, R, *q)
    D = fine_grid(O.dilate(I, R, r=r), R)
    N = 2 * R
    z, y, x = np.mgrid[0:N, 0:N, 0:N]

    def d(t):
        return np.where(t >= 0, t, -t - 1)

    ball = d(q[0] - x) ** 2 + d(q[1] - y) ** 2 + d(q[2] - z) ** 2 < (r + 1) ** 2
    assert np.array_equal(D, ball)
    # count closed form: 8 * #{a,b,c >= 0 : a^2+b^2+c^2 < (r+1)^2}
    cnt = sum(1 for a in range(r + 1) for b in range(r + 1) for c in range(r + 1) if a * a + b * b + c * c < (r + 1) ** 2)
    assert D.sum() == 8 * cnt


@pytest.mark.parametrize("r", [2, 3, 5])
def test_single_voxel_cube_is_box(r):
    # closed box test: q in [p-r-1, p+r]  <=>  p in [q-r, q+r+1]
    R = 16
    q = (14, 17, 15)
    I = np.zeros(R ** 3, np.uint8)
    set_fine(I, R, *q)
    D = fine_grid(O.dilate(I, R, r=r, se=1), R)
    box = np.zeros_like(D)
    box[q[2] - r:q[2] + r + 2, q[1] - r:q[1] + r + 2, q[0] - r:q[0] + r + 2] = True
    assert np.array_equal(D, box)


def test_dilation_contains_input_and_respects_data_res():
    R = 24
    I = random_blobs(R, seed=3)
    dr = np.array([R - 4, R, R - 8], np.int32)
    D = O.dilate(I, R, r=4, data_res=dr)
    assert np.all((I & ~D) == 0)
    # coarse fills only happen inside the data region: a cell beyond it is
    # either untouched or a tight-contour cell refined voxel by voxel
    Dfull = O.dilate(I, R, r=4)
    assert np.all((D & ~Dfull) == 0)


# ---------------------------------------------------------------- erosion
@pytest.mark.parametrize("seed", range(3))
def test_erode_dfs_equals_brute(seed):
    R = 16
    rng = np.random.default_rng(100 + seed)
    I = random_sparse(R, seed + 50, density=0.02, margin=1)
    r = int(rng.integers(2, 6))
    D = O.dilate(I, R, r=r)
    assert np.array_equal(O.erode(I, D, R, r=r), O.erode(I, D, R, r=r, brute=1))
    sc = rng.integers(-2, 7, (2 * R) ** 3).astype(np.int8)
    D2 = O.dilate(I, R, scale=sc)
    assert np.array_equal(O.erode(I, D2, R, scale=sc), O.erode(I, D2, R, scale=sc, brute=1))


def _hole_ball(p, r, N):
    """Fine voxels v whose centre lies closer than r-1 to the box [p, p+1]^3:
    per axis max(0, |v - p| - 1/2), in doubled units (2|t| - 1 for t != 0)."""
    z, y, x = np.mgrid[0:N, 0:N, 0:N]

    def d2(t):
        t = np.abs(t)
        return np.where(t == 0, 0, (2 * t - 1) ** 2)

    return d2(x - p[0]) + d2(y - p[1]) + d2(z - p[2]) < 4 * (r - 1) ** 2


@pytest.mark.parametrize("r", [2, 3, 4, 6, 9])
def test_erosion_single_contour_voxel_closed_form(r):
    # D full except fine voxel p, so K = {p}; I empty. Reference semantics
    # (GMorpho.cu:6627-6702): the coarse pass clears a cell when its centre
    # 2c+1 is closer than r-1 to p's box (GMorpho.cu:4372, radius r-1 at
    # 4430-4437); the fine pass rewrites every cell of the 6-connected base
    # contour of that result (Grid.cu:1499-1570) with the sub-voxels whose
    # centre is not closer than r-1 (GMorpho.cu:5081-5345). The distances to a
    # single box are closed forms: per axis max(0, |t| - 1/2) from a centre.
    R = 16
    N = 2 * R
    p = (15, 16, 17)
    f = np.ones((N, N, N), bool)
    f[p[2], p[1], p[0]] = False
    D = cells_from_fine(f)
    K = fine_grid(O.contour_K(D, R), R)
    assert K.sum() == 1 and K[p[2], p[1], p[0]]
    I = np.zeros_like(D)
    C = fine_grid(O.erode(I, D, R, r=r), R)
    ball = _hole_ball(p, r, N)  # sub-voxel centres in reach
    # coarse: cell centres (2c+1) in reach, in doubled units |2(2c+1) - (2p+1)|
    cz, cy, cx = np.mgrid[0:R, 0:R, 0:R]

    def cd2(cc, pp):
        t = np.abs(2 * (2 * cc + 1) - (2 * pp + 1))
        return np.maximum(0, t - 1) ** 2

    cleared = cd2(cx, p[0]) + cd2(cy, p[1]) + cd2(cz, p[2]) < 4 * (r - 1) ** 2
    inner = np.zeros_like(cleared)
    inner[1:-1, 1:-1, 1:-1] = True
    byte = np.where(cleared, 0, 1)
    diff = np.zeros_like(cleared)
    for ax in range(3):
        for s in (1, -1):
            diff |= byte != np.roll(byte, s, axis=ax)
    U = diff & inner
    up = lambda a: np.repeat(np.repeat(np.repeat(a, 2, 0), 2, 1), 2, 2)  # noqa: E731
    want = np.where(up(U), ~ball, np.where(up(cleared), False, f))
    assert np.array_equal(C, want)
    # never beyond the ideal ball: every eroded voxel is in reach of p
    assert not np.any(~C & ~ball & f)
    # the ideal ball's closed-form count #{t : sum max(0,|t_a|-1/2)^2 < (r-1)^2}
    rr = r - 1
    cnt = sum(1 for a in range(-rr, rr + 1) for b in range(-rr, rr + 1) for c in range(-rr, rr + 1)
              if sum(0 if t == 0 else (2 * abs(t) - 1) ** 2 for t in (a, b, c)) < 4 * rr * rr)
    assert ball.sum() == cnt
    if r == 2:
        # the 3x3x3 block around p: every sum of up to three 1/4 terms is < 1
        assert cnt == 27
    if r == 3:
        # the 5x5x5 block less the voxels two steps out on at least two axes
        # (9/4 + 9/4 >= 4): 8 corners and 3 * 4 * 3 = 36 with exactly two
        assert cnt == 125 - 8 - 36


def test_erosion_coarse_query_at_cell_centre():
    # Coarse erosion queries the cell centre 2c+1 (GMorpho.cu:4372), not the
    # corner 2c. Border cells (x = 0) are outside the base contour
    # ([1,R-2]^3, Grid.cu:1499-1570), so the coarse result stands there.
    # Contour voxel p = (2, 2cy, 2cz) in cell x = 1; cell (0, cy, cz): centre
    # x = 1 is at distance 1 from [2,3], the corner x = 0 at distance 2.
    R = 16
    N = 2 * R
    cy, cz = 7, 8
    p = (2, 2 * cy, 2 * cz)
    f = np.ones((N, N, N), bool)
    f[p[2], p[1], p[0]] = False
    D = cells_from_fine(f)
    I = np.zeros_like(D)
    idx = 0 + R * (cy + R * cz)
    # r = 3: 1 < (r-1)^2 = 4 clears the border cell; a corner query (4 < 4)
    # would not
    C = O.erode(I, D, R, r=3)
    assert C[idx] == 0
    # r = 2: 1 < 1 fails, the border cell keeps D
    C = O.erode(I, D, R, r=2)
    assert C[idx] == 0xFF
    # a cell full in I is never cleared by the coarse pass (GMorpho.cu:4349ff)
    I2 = I.copy()
    I2[idx] = 0xFF
    assert O.erode(I2, D, R, r=3)[idx] == 0xFF


def test_erosion_fine_result_is_not_anded_with_dilation():
    # The fine pass writes the surviving sub-voxels of a base-contour cell
    # whatever D held there (GMorpho.cu:5347-5360 store the byte, no AND).
    # D: one fine voxel f0 = 2c' (bit 0 of cell c'); I full at c' so the
    # coarse pass leaves it; K = the 6 neighbours of f0; r = 2 erodes the
    # 3x3x3 block of every K voxel. The base contour is c' and its six
    # neighbours. In the +x neighbour the voxels at x = 2c'x + 3 are 1.5 from
    # the nearest contour box, so they survive although D is empty there:
    # bits with x offset 1 (0xAA); likewise +y (0xCC) and +z (0xF0). c', the
    # -x/-y/-z neighbours and every other cell end empty.
    R = 16
    c = (7, 8, 6)
    D = np.zeros(R ** 3, np.uint8)
    ci = lambda x, y, z: x + R * (y + R * z)  # noqa: E731
    D[ci(*c)] = 0x01
    I = np.zeros_like(D)
    I[ci(*c)] = 0xFF
    C = O.erode(I, D, R, r=2)
    want = np.zeros_like(D)
    want[ci(c[0] + 1, c[1], c[2])] = 0xAA
    want[ci(c[0], c[1] + 1, c[2])] = 0xCC
    want[ci(c[0], c[1], c[2] + 1)] = 0xF0
    assert np.array_equal(C, want)


def test_contour_K_hand_bytes():
    # TagVoxelsContourByErosion (Grid.cu:1230-1327) on hand-built 3x3x3
    # neighbourhoods; expected bytes written out by hand.
    R = 8
    ci = lambda x, y, z: x + R * (y + R * z)  # noqa: E731
    c = (3, 4, 3)
    # one full cell: each face neighbour gets the 4 voxels touching the face,
    # diagonal neighbours and the full cell itself get nothing
    D = np.zeros(R ** 3, np.uint8)
    D[ci(*c)] = 0xFF
    K = O.contour_K(D, R)
    want = np.zeros_like(D)
    want[ci(c[0] - 1, c[1], c[2])] = 0xAA
    want[ci(c[0] + 1, c[1], c[2])] = 0x55
    want[ci(c[0], c[1] - 1, c[2])] = 0xCC
    want[ci(c[0], c[1] + 1, c[2])] = 0x33
    want[ci(c[0], c[1], c[2] - 1)] = 0xF0
    want[ci(c[0], c[1], c[2] + 1)] = 0x0F
    assert np.array_equal(K, want)
    # one voxel (bit 0): its neighbours inside the cell (bits 1, 2, 4) and
    # across the -x, -y, -z faces (bit 1, 2, 4 of those cells)
    D = np.zeros(R ** 3, np.uint8)
    D[ci(*c)] = 0x01
    K = O.contour_K(D, R)
    want = np.zeros_like(D)
    want[ci(*c)] = 0x16
    want[ci(c[0] - 1, c[1], c[2])] = 0x02
    want[ci(c[0], c[1] - 1, c[2])] = 0x04
    want[ci(c[0], c[1], c[2] - 1)] = 0x10
    assert np.array_equal(K, want)
    # a full cell less voxel 0: the hole itself, and in the -x neighbour the
    # face voxels next to set voxels (bits 3, 5, 7; bit 1 faces the hole)
    D = np.zeros(R ** 3, np.uint8)
    D[ci(*c)] = 0xFE
    K = O.contour_K(D, R)
    assert K[ci(*c)] == 0x01
    assert K[ci(c[0] - 1, c[1], c[2])] == 0xA8
    assert K[ci(c[0] + 1, c[1], c[2])] == 0x55
    # cells on the grid border ([1,R-2]^3 only) are never tagged
    D = np.zeros(R ** 3, np.uint8)
    D[ci(1, 4, 4)] = 0xFF
    K = O.contour_K(D, R)
    assert K[ci(0, 4, 4)] == 0 and K[ci(2, 4, 4)] == 0x55


def test_contour_K_definition():
    R = 16
    D = random_blobs(R, seed=9, rmin=3, rmax=7, margin=2)
    K = fine_grid(O.contour_K(D, R), R)
    f = fine_grid(D, R)
    N = 2 * R
    nb = np.zeros_like(f)
    nb[1:] |= f[:-1]
    nb[:-1] |= f[1:]
    nb[:, 1:] |= f[:, :-1]
    nb[:, :-1] |= f[:, 1:]
    nb[:, :, 1:] |= f[:, :, :-1]
    nb[:, :, :-1] |= f[:, :, 1:]
    cellfull = np.repeat(np.repeat(np.repeat((D.reshape(R, R, R) == 255), 2, 0), 2, 1), 2, 2)
    inner = np.zeros_like(f)
    inner[2:N - 2, 2:N - 2, 2:N - 2] = True
    assert np.array_equal(K, ~f & nb & ~cellfull & inner)


def test_erosion_of_dilation_keeps_input():
    # closing is extensive away from the grid border: every input voxel stays
    R = 24
    I = random_blobs(R, seed=4, margin=8)
    for r in (2, 4):
        D = O.dilate(I, R, r=r)
        C = O.erode(I, D, R, r=r)
        assert np.all((I & ~C) == 0)


def test_closing_of_ball_bounds_it():
    # The reference's closing dilates with r+1 and erodes with r-1 (the proxy
    # must bound the shape): a ball comes back grown by about two voxels,
    # never smaller, and inside the dilation.
    R = 24
    N = 2 * R
    z, y, x = np.mgrid[0:N, 0:N, 0:N]
    d2 = (x - 23.5) ** 2 + (y - 24.5) ** 2 + (z - 23.5) ** 2
    ball = d2 < 10.3 ** 2
    I = cells_from_fine(ball)
    for r in (2, 3, 4):
        D = O.dilate(I, R, r=r)
        C = fine_grid(O.erode(I, D, R, r=r), R)
        assert np.all(C[ball])
        assert not np.any(C & ~fine_grid(D, R))
        assert not np.any(C & (d2 >= (10.3 + 3.5) ** 2))


# ---------------------------------------------------------------- contours / order
def test_mcm_list_is_morton_sorted_and_complete():
    R = 16
    C = random_blobs(R, seed=2, margin=2)
    lst = O.contour_list(C, R, 2)
    xyz = [(c % R, (c // R) % R, c // (R * R)) for c in lst]
    codes = [morton(*p) for p in xyz]
    assert codes == sorted(codes) and len(set(codes)) == len(codes)
    f = fine_grid(C, R)
    expect = set()
    for cz in range(1, R - 1):
        for cy in range(1, R - 1):
            for cx in range(1, R - 1):
                blk = f[2 * cz:2 * cz + 3, 2 * cy:2 * cy + 3, 2 * cx:2 * cx + 3]
                for j in range(8):
                    cube = blk[(j >> 2) & 1:(j >> 2) + 2, (j >> 1) & 1:((j >> 1) & 1) + 2, j & 1:(j & 1) + 2]
                    if cube.any() and not cube.all():
                        expect.add(cx + R * (cy + R * cz))
                        break
    assert set(lst.tolist()) == expect


# ---------------------------------------------------------------- marching cubes
def test_mc_tables():
    e, n, t = O.mc_tables()
    corners_edges = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
    for c in range(256):
        m = sum(1 << k for k, (a, b) in enumerate(corners_edges) if ((c >> a) & 1) != ((c >> b) & 1))
        assert e[c] == m
        listed = int((t[c] != -1).sum())
        assert listed % 3 == 0
        # every triangle uses crossing edges only
        assert all((e[c] >> int(k)) & 1 for k in t[c][:listed])
        # every crossing edge is used by some triangle
        assert set(int(k) for k in t[c][:listed]) == {k for k in range(12) if (e[c] >> k) & 1}
        if c != 91:
            assert n[c] == listed // 3
    assert n[91] == 4 and (t[91] != -1).sum() == 15
    ntri, ln, tri = B.mc_tables()
    for c in range(256):
        listed = [int(x) for x in t[c] if x != -1]
        packed = [(int(tri[c]) >> (4 * i)) & 0xF for i in range(len(listed))]
        assert packed == listed and ln[c] == len(listed) // 3 and ntri[c] == n[c]


def test_mc_tables_equal_reference_fixture():
    """Case tables equal the reference's tables.h data (fixture written by
    tests/golden/make_golden.py)."""
    g = np.load(os.path.join(GOLDEN, "mc_tables.npz"))
    e, n, t = O.mc_tables()
    assert np.array_equal(g["edgeTable"], np.asarray(e, np.uint16))
    assert np.array_equal(g["numTriTable"], np.asarray(n, np.uint8))
    assert np.array_equal(g["triTable"].reshape(-1), np.asarray(t, np.int8).reshape(-1))


def _mesh_checks(V, T):
    assert T.max() < len(V)
    E = Counter()
    for a, b, c in T:
        for u, v in ((a, b), (b, c), (c, a)):
            E[(min(u, v), max(u, v))] += 1
    return E


def test_mc_closed_manifold_euler():
    R = 24
    N = 2 * R
    z, y, x = np.mgrid[0:N, 0:N, 0:N]
    ball = (x - 23.5) ** 2 + (y - 24.5) ** 2 + (z - 23.5) ** 2 < 13.0 ** 2
    C = cells_from_fine(ball)
    V, Nn, T = O.mc(C, R, np.zeros(3, np.float32), 1.0)
    E = _mesh_checks(V, T)
    assert all(v == 2 for v in E.values())
    assert len(V) - len(E) + len(T) == 2
    # every vertex sits at 0.05 or 0.95 along its fine edge (h = 0.5)
    frac = (V.astype(np.float64) / 0.5) % 1.0
    dist = np.min(np.abs(frac[..., None] - np.array([0.0, 0.05, 0.95, 1.0])), axis=-1)
    assert np.all(dist < 1e-5)
    assert np.all((np.abs(frac - 0.05) < 1e-5).sum(1) + (np.abs(frac - 0.95) < 1e-5).sum(1) == 1)
    # outward normals: the normal points away from the ball centre
    c = np.array([24.0, 25.0, 24.0]) * 0.5
    assert np.mean(np.einsum("ij,ij->i", Nn, V - c) > 0) > 0.99
    assert np.allclose(np.linalg.norm(Nn, axis=1), 1.0, atol=1e-5)


def test_mc_single_voxel():
    # one occupied fine voxel -> an octahedron-like closed surface: 8 cubes x 1 triangle
    R = 16
    I = np.zeros(R ** 3, np.uint8)
    set_fine(I, R, 15, 16, 15)
    V, Nn, T = O.mc(I, R, np.zeros(3, np.float32), 1.0)
    assert len(T) == 8 and len(V) == 6
    E = _mesh_checks(V, T)
    assert all(v == 2 for v in E.values())


def test_mc_vertex_count_closed_form():
    R = 16
    C = random_blobs(R, seed=12, margin=3)
    nv, nt = O.mc_counts(C, R)
    e, n, t = O.mc_tables()
    f = fine_grid(C, R)
    N = 2 * R
    v = t_ = 0
    for cz in range(1, R - 1):
        for cy in range(1, R - 1):
            for cx in range(1, R - 1):
                for j in range(8):
                    fx, fy, fz = 2 * cx + (j & 1), 2 * cy + ((j >> 1) & 1), 2 * cz + ((j >> 2) & 1)
                    cube = [f[fz + dz, fy + dy, fx + dx] for dx, dy, dz in
                            [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]]
                    case = sum(int(b) << k for k, b in enumerate(cube))
                    if case in (0, 255):
                        continue
                    v += bin(int(e[case]) & 0x109).count("1")
                    t_ += int(n[case])
    assert (nv, nt) == (v, t_)


# ---------------------------------------------------------------- voxelizer
def _snap(P, bmin, maxext, N):
    v = (P.astype(np.float64) - bmin.astype(np.float64)) * (N / np.float64(maxext))
    return np.floor(v * 256.0 + 0.5).astype(np.int64)


def test_voxelize_convex_parity_matches_inside_test():
    # For a closed convex mesh the parity fill sets voxel (i,j,k) iff the
    # column through the pixel centre enters below k and leaves at or above
    # it: checked against half-space tests on the snapped triangles.
    P, T = B.icosphere(2)
    P = (P * 0.4 + 0.5).astype(np.float32)
    R = 16
    N = 2 * R
    bmin, bmax, cell, dr = O.grid_attributes(P, R, 0.05)
    maxext = float((bmax - bmin).max())
    I = fine_grid(O.voxelize(P, T, bmin, maxext, R), R)
    X = _snap(P, bmin, maxext, N)
    A, Bv, C = X[T[:, 0]], X[T[:, 1]], X[T[:, 2]]
    n = np.cross(Bv - A, C - A)
    inside = np.zeros((N, N, N), bool)
    for k in range(N):
        for j in range(N):
            for i in range(N):
                p = np.array([256 * i + 128, 256 * j + 128, 256 * k])
                s = np.einsum("ij,ij->i", n, p - A)
                inside[k, j, i] = np.all(s < 0) or (np.all(s <= 0) and np.any((s == 0) & (n[:, 2] > 0)))
    # the conservative shell may add voxels, but parity voxels are exactly `inside`
    assert np.all(I[inside])
    extra = I & ~inside
    # every extra voxel is a shell voxel: within 2 voxels of the surface
    kk, jj, ii = np.nonzero(extra)
    c = (np.array([0.5, 0.5, 0.5]) - bmin) * (N / maxext)
    rad = 0.4 * (N / maxext)
    d = np.sqrt((ii + 0.5 - c[0]) ** 2 + (jj + 0.5 - c[1]) ** 2 + (kk + 0.5 - c[2]) ** 2)
    assert np.all(np.abs(d - rad) < 2.0)


def test_voxelize_solid_soup_is_filled():
    P, T = B.sphere_soup(4, 3, seed=5, rmin=0.08, rmax=0.12, cmin=0.3, cmax=0.7)
    R = 32
    bmin, bmax, cell, dr = O.grid_attributes(P, R, 0.3 + 1.0 / R)
    I = O.voxelize(P, T, bmin, float((bmax - bmin).max()), R)
    assert fine_grid(I, R).sum() > 1000
    # voxelisation is deterministic and independent of triangle order
    perm = np.random.default_rng(0).permutation(len(T))
    I2 = O.voxelize(P, T[perm], bmin, float((bmax - bmin).max()), R)
    assert np.array_equal(I, I2)


def test_grid_attributes_match_reference_formula():
    P, T = B.sh_blob(2, seed=3)
    R = 64
    margin = np.float32(0.3) + np.float32(2.0) / np.float32(2 * R)
    bmin, bmax, cell, dr = O.grid_attributes(P, R, float(margin))
    mn, mx = P.min(0), P.max(0)
    thick = np.float32(np.max(np.abs(mx - mn)) / np.float32(R))
    off = np.float32(np.float32(2) * thick + margin)
    assert np.array_equal(bmin, (mn - off).astype(np.float32))
    assert np.array_equal(bmax, (mx + off).astype(np.float32))
    cs = np.float32(np.max(np.abs(bmax - bmin)) / np.float32(R))
    assert np.float32(cell) == cs
    assert list(dr) == [int(np.ceil(np.float32((bmax[a] - bmin[a]) / cs))) for a in range(3)]


# ---------------------------------------------------------------- ScaleField lattice
def test_lattice_constant_and_ramp():
    L, N = 4, 32
    lat = np.full((L, L, L), 7.75, np.float32)
    assert np.all(O.scale_from_lattice(lat, N) == 7)
    # trilinear interpolation reproduces a linear ramp exactly
    s = N // L
    ramp = np.zeros((L, L, L), np.float32)
    for i in range(L):
        ramp[:, :, i] = 3 + 2 * i
    out = O.scale_from_lattice(ramp, N).reshape(N, N, N)
    for x in range(N):
        xc = min(x, s * (L - 1))  # clamped beyond the last node
        exp = (256 * 3 * s + 512 * xc) // (256 * s) if x < s * (L - 1) else 3 + 2 * (L - 1)
        assert np.all(out[:, :, x] == exp), x


# ---------------------------------------------------------------- pyramid
def test_pyramid_levels():
    R = 16
    I = random_blobs(R, seed=1, rmin=1.5, rmax=3.0, margin=2)
    pyr = O.pyramid(I, R, r=5)
    lv0 = pyr[:R ** 3]
    assert np.all((lv0 >> np.uint64(56)).astype(np.uint8) == I)
    assert np.all((lv0 & np.uint64(0x7F)) == 5)
    top = pyr[-1]
    nonempty = [(I.reshape(R, R, R)[(l >> 2) * 8:(l >> 2) * 8 + 8, ((l >> 1) & 1) * 8:((l >> 1) & 1) * 8 + 8,
                 (l & 1) * 8:(l & 1) * 8 + 8] != 0).any() for l in range(8)]
    assert [(int(top) >> (56 + l)) & 1 for l in range(8)] == [int(b) for b in nonempty]


# ---------------------------------------------------------------- fixtures / generators
def test_icosphere_matches_reference_fixture():
    """The generator's subdivision-3 icosphere is the reference's sphere.off
    (same counts, same vertex set to float precision)."""
    g = np.load(os.path.join(GOLDEN, "sphere_off.npz"))
    Q, F = g["V"], g["F"]
    P, T = B.icosphere(3)
    assert (len(P), len(T)) == (len(Q), len(F))
    d = np.sqrt(((P[:, None, :] - Q[None, :, :]) ** 2).sum(-1)).min(1)
    assert d.max() < 2e-6


def test_oracle_matches_golden_config0():
    """The oracle's config-0 closing and mesh still hash to the committed
    record (regression pin of the oracle; written by make_golden.py)."""
    import hashlib
    import json
    g = json.load(open(os.path.join(GOLDEN, "oracle_config0.json")))
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    P, T = B.sphere_soup(64, 3, seed=2017)
    R = 64
    margin = np.float32(0.3) + np.float32(2.0) / np.float32(2 * R)
    bmin, bmax, cell, dr = O.grid_attributes(P, R, float(margin))
    assert [float(x) for x in bmin] == g["bmin"] and [int(x) for x in dr] == g["data_res"]
    I = O.voxelize(P, T, bmin, float((bmax - bmin).max()), R)
    D = O.dilate(I, R, r=4, data_res=dr)
    C = O.erode(I, D, R, r=4, data_res=dr)
    V, Nn, Tt = O.mc(C, R, bmin, cell)
    assert (sha(I), sha(D), sha(C)) == (g["I"], g["D"], g["C"])
    assert (len(V), len(Tt), sha(Tt), sha(V)) == (g["nV"], g["nT"], g["T"], g["V"])


def test_generators_deterministic_and_closed():
    P1, T1 = B.sh_blob(3, seed=42)
    P2, T2 = B.sh_blob(3, seed=42)
    assert np.array_equal(P1, P2) and np.array_equal(T1, T2)
    assert P1.min() >= 0 and np.isclose(P1.max(0).max(), 1.0)
    E = _mesh_checks(P1, T1)
    assert all(v == 2 for v in E.values())
    lat = B.scale_lattice(8, seed=1)
    assert lat.min() >= 2 and lat.max() <= 24


# ---------------------------------------------------------------- C-ABI
def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(os.path.dirname(B.HERE), "include", "gmorpho_c.h")).read()
    names = set(re.findall(r"\b(gm_[a-z0-9_]+)\s*\(", hdr))
    lib = B.hip_lib()
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing and len(names) >= 25
    mg = open(os.path.join(os.path.dirname(B.HERE), "include", "gmorpho_meshgen.h")).read()
    for n in set(re.findall(r"\b(gmg_[a-z0-9_]+)\s*\(", mg)):
        assert hasattr(B.meshgen_lib(), n)


def test_compiled_chord_tables_equal_runtime_tables():
    """The streaming kernels' compile-time chord tables (one per fixed radius)
    list exactly the chords the host table builder emits (host-only check)."""
    import ctypes
    f = getattr(B.hip_lib(), "_ZN2gm12ct_tables_okEv")
    f.restype = ctypes.c_bool
    assert f()


def test_facade_driver_fails_loudly_without_device():
    """The C++ facade (include/gmorpho.hpp) compiles and links against the
    library; without a device its GMorpho constructor throws the reference's
    exception type, which the driver reports."""
    import subprocess
    exe = os.path.join(B.LIBDIR, "bin", "facade_closing")
    assert os.path.exists(exe), "run broxy_amd/build.py driver"
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"):
        pytest.skip("a HIP device may be present")
    p = subprocess.run([exe, "/tmp/gm_facade_nodev"], capture_output=True, text=True, timeout=120)
    assert p.returncode != 0 and "GMorpho Error:" in p.stderr


def test_context_without_device_fails_loudly():
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"):
        pytest.skip("a HIP device may be present")
    with pytest.raises(B.GMorphoError):
        B.Context(0)


# ---------------------------------------------------------------- bilateral normals
def _pairwise_bilateral(P, N, nbrs, iters):
    """Edge-pair form of BilateralMeshFiltering (GMorpho.cu:6539-6620) for a
    closed mesh where every edge lies in exactly two triangles: each vertex
    sums w(v,u) n_u over its edge neighbours u (twice, which normalization
    removes), w(v,u) = exp(-(|p_v-p_u|^2 + |n_v-n_u|^2 / sigma_n^2)), with
    sigma_n = 1e10 for the first two iterations and 0.2 after
    (GMorpho.cu:6957-6978)."""
    N = N.astype(np.float64)
    for it in range(iters):
        sn2 = (1e10 if it < 2 else 0.2) ** 2
        out = np.zeros_like(N)
        for v, us in enumerate(nbrs):
            for u in us:
                w = np.exp(-(np.sum((P[v] - P[u]) ** 2) + np.sum((N[v] - N[u]) ** 2) / sn2))
                out[v] += w * N[u]
        N = out / np.linalg.norm(out, axis=1, keepdims=True)
    return N


def test_bilateral_tetrahedron_pairwise():
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32) * np.float32(0.7)
    T = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], np.uint32)
    rng = np.random.default_rng(5)
    N = rng.normal(size=(4, 3))
    N = (N / np.linalg.norm(N, axis=1, keepdims=True)).astype(np.float32)
    nbrs = [[u for u in range(4) if u != v] for v in range(4)]
    for iters in (1, 2, 3, 5):
        got = O.bilateral_normals(P, N, T, iters)
        want = _pairwise_bilateral(P.astype(np.float64), N, nbrs, iters)
        # float vs double: with sigma_n = 0.2 the exponent reaches ~100, so a
        # rounding of the exponent moves a weight by ~1e-5 relative
        assert np.allclose(got, want, atol=2e-5), iters
    # a vertex's own normal takes no part: with every other normal equal the
    # result is that normal whatever the vertex held
    N2 = np.tile(np.array([[0, 0, 1]], np.float32), (4, 1))
    N2[0] = [1, 0, 0]
    assert np.allclose(O.bilateral_normals(P, N2, T, 1)[0], [0, 0, 1], atol=1e-7)


def test_bilateral_planar_patch_is_fixed():
    # a flat grid of triangles with equal normals stays unchanged
    n = 6
    P = np.array([[x * 0.1, y * 0.1, 0.3] for y in range(n) for x in range(n)], np.float32)
    T = []
    for y in range(n - 1):
        for x in range(n - 1):
            a = y * n + x
            T += [[a, a + 1, a + n + 1], [a, a + n + 1, a + n]]
    T = np.array(T, np.uint32)
    N = np.tile(np.array([[0, 0, 1]], np.float32), (len(P), 1))
    assert np.array_equal(O.bilateral_normals(P, N, T, 7), N)


# ---------------------------------------------------------------- scale-field brush
def _brush_case(mode, prev, scale=0.5, support=0.8, gscale=0.125):
    N = 32
    cs = 0.0625
    bmin = np.array([-0.25, 0.0, 0.125], np.float32)
    pos = np.array([0.77, 1.03, 1.21], np.float32)
    g0 = np.full(N ** 3, prev, np.int8)
    g = O.scale_brush(g0, N, bmin, cs, pos, scale, support, gscale, mode).reshape(N, N, N)
    z, y, x = np.mgrid[0:N, 0:N, 0:N]
    u = np.stack([x, y, z], -1) * cs + bmin.astype(np.float64)
    dist = np.linalg.norm(u - pos.astype(np.float64), axis=-1)
    h = (support - scale) / np.sqrt(1 - np.sqrt(0.5))
    return g, dist, scale, h, cs


def test_scale_brush_max_closed_forms():
    # ScaleField.cu:81-141, MAX_BRUSH over a uniform field of 4: inside the
    # brush radius the field takes floor(scale / cell) = 8; beyond scale + h
    # it is untouched; in the tukey band (cuda_math.h:696-704) it is
    # floor(prev (1 - t) + t * 8), t = (1 - ((d - v)/h)^2)^2, never below prev
    g, d, v, h, cs = _brush_case(1, 4)
    safe = lambda a, b: np.abs(a - b) > 1e-5  # noqa: E731
    inside = (d < v) & safe(d, v)
    outside = (d > v + h) & safe(d, v + h)
    assert inside.sum() > 100 and outside.sum() > 1000
    assert np.all(g[inside] == 8)
    assert np.all(g[outside] == 4)
    band = (d > v) & (d < v + h) & safe(d, v) & safe(d, v + h)
    t = (1 - ((d - v) / h) ** 2) ** 2
    val = 4 * (1 - t) + t * 8
    ok = band & (np.abs(val - np.round(val)) > 1e-4)
    assert ok.sum() > 100
    assert np.array_equal(g[ok], np.maximum(np.floor(val[ok]), 4).astype(np.int8))
    assert np.all((g >= 4) & (g <= 8))


def test_scale_brush_min_zero_modes():
    # MIN_BRUSH from 12 only lowers (to 8 inside); ZERO_BRUSH pulls toward
    # floor(global_scale / cell) = 2 inside, whatever the field held
    g, d, v, h, cs = _brush_case(2, 12)
    inside = (d < v - 1e-5)
    assert np.all(g[inside] == 8) and np.all((g >= 8) & (g <= 12))
    assert np.all(g[d > v + h + 1e-5] == 12)
    g, d, v, h, cs = _brush_case(3, 12)
    assert np.all(g[d < v - 1e-5] == 2)
    assert np.all(g[d > v + h + 1e-5] == 12)
    # NO_EDITION changes nothing
    g, d, v, h, cs = _brush_case(0, 7)
    assert np.all(g == 7)
    # values are clamped to [0, 127]: a brush of 200 cells is 127 inside
    g, d, v, h, cs = _brush_case(1, 3, scale=0.5, support=0.6, gscale=0.0)
    g2 = O.scale_brush(np.full(32 ** 3, 3, np.int8), 32, np.zeros(3, np.float32), 0.0025, np.zeros(3, np.float32),
                       0.5, 0.6, 0.0, 1)
    assert g2.max() == 127


def test_exp_weight_accuracy():
    # the float exp of the bilateral weights: within 3 ulp of exp over its
    # range, exactly 1 at 0, 0 below -87
    x = -np.concatenate([np.linspace(0, 87, 200001), np.geomspace(1e-8, 1e-2, 1000)]).astype(np.float32)
    got = O.exp_weight(x).astype(np.float64)
    ref = np.exp(x.astype(np.float64))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(got - ref) / ulp) <= 3.0
    assert O.exp_weight(np.array([0.0, -88.0, -1e30], np.float32)).tolist() == [1.0, 0.0, 0.0]


def _box_mesh(lo, hi):
    """Closed axis-aligned box, outward-facing triangles."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    P = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float32)
    T = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                  [3, 7, 6], [3, 6, 2], [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], np.uint32)
    return P, T


def _ms_grid(P, T, R):
    # bmin = 0 and maxext = N: mesh coordinates are fine-voxel units
    N = 2 * R
    return fine_grid(O.voxelize_multislice(P, T, np.zeros(3, np.float32), float(N), R), R)


def test_multislice_box_closed_form():
    """MultiSlicing of a box: voxel (x, y, z) is set iff floor(x0) <= x <
    floor(x1) (the crossing voxel toggles the column from there on) and the
    column centre (y + 1/2, z + 1/2) lies inside the face; no shell."""
    R = 24
    lo, hi = (10.3, 5.2, 7.1), (30.7, 20.9, 15.6)
    f = _ms_grid(*_box_mesh(lo, hi), R)
    want = np.zeros_like(f)
    want[7:16, 5:21, 10:30] = True  # [z][y][x]: k + 1/2 in (7.1, 15.6), j + 1/2 in (5.2, 20.9)
    assert np.array_equal(f, want)
    # triangle order and winding of the input do not matter
    P, T = _box_mesh(lo, hi)
    assert np.array_equal(_ms_grid(P, T[::-1, ::-1].copy(), R), want)


def test_multislice_tie_rule_and_parity():
    """Faces through column centres: the top-left rule keeps exactly one of
    the two boundary columns per axis (15 of the 16 centres on [5.5, 20.5]),
    and two touching boxes cover the union's columns once each; a box inside
    another leaves a hollow (parity)."""
    R = 24
    f = _ms_grid(*_box_mesh((10.3, 5.5, 7.5), (30.7, 20.5, 15.5)), R)
    cols = f.any(axis=2)  # [z][y]
    assert cols.sum() == 15 * 8
    assert cols[:, 5:21].sum() == 15 * 8 and cols[7:16].sum() == 15 * 8
    # two boxes sharing the face y = 12.5: together exactly the columns of one
    Pa, Ta = _box_mesh((10.3, 5.5, 7.5), (30.7, 12.5, 15.5))
    Pb, Tb = _box_mesh((10.3, 12.5, 7.5), (30.7, 20.5, 15.5))
    P2 = np.concatenate([Pa, Pb])
    T2 = np.concatenate([Ta, Tb + 8]).astype(np.uint32)
    assert np.array_equal(_ms_grid(P2, T2, R), f)
    # nested boxes: parity leaves the inner box empty
    Po, To = _box_mesh((4.3, 4.2, 4.1), (40.7, 40.9, 40.6))
    Pi, Ti = _box_mesh((12.3, 12.2, 12.1), (20.7, 20.9, 20.6))
    g = _ms_grid(np.concatenate([Po, Pi]), np.concatenate([To, Ti + 8]).astype(np.uint32), R)
    outer = np.zeros_like(g)
    outer[4:41, 4:41, 4:40] = True
    inner = np.zeros_like(g)
    inner[12:21, 12:21, 12:20] = True
    assert np.array_equal(g, outer & ~inner)


def test_multislice_convex_matches_crossings():
    """For a closed convex mesh, each covered column is set exactly from the
    voxel of its entry crossing up to (not including) the voxel of its exit
    crossing, with the crossings computed here from the snapped triangles in
    exact rational arithmetic."""
    from fractions import Fraction
    P, T = B.icosphere(1)
    R = 12
    N = 2 * R
    P = (P * 9.0 + np.array([12.137, 11.891, 12.073])).astype(np.float32)  # fine-voxel units, inside [2, 22]
    f = _ms_grid(P, T, R)
    X = _snap(P, np.zeros(3, np.float32), float(N), N)
    want = np.zeros_like(f)
    for k in range(N):
        for j in range(N):
            py, pz = 256 * j + 128, 256 * k + 128
            xs = []
            for a, b, c in X[T]:
                d = (b[1] - a[1]) * (c[2] - a[2]) - (b[2] - a[2]) * (c[1] - a[1])
                if d == 0:
                    continue
                # barycentric weights of (py, pz) in the (y, z) projection
                wb = Fraction(int((py - a[1]) * (c[2] - a[2]) - (pz - a[2]) * (c[1] - a[1])), int(d))
                wc = Fraction(int((b[1] - a[1]) * (pz - a[2]) - (b[2] - a[2]) * (py - a[1])), int(d))
                wa = 1 - wb - wc
                if min(wa, wb, wc) < 0:
                    continue
                if min(wa, wb, wc) == 0:
                    raise AssertionError("column centre on an edge: pick another scale")
                xs.append((wa * int(a[0]) + wb * int(b[0]) + wc * int(c[0])) / 256)
            if xs:
                assert len(xs) == 2
                lo, hi = min(xs), max(xs)
                want[k, j, int(np.floor(float(lo))):int(np.floor(float(hi)))] = True
    assert np.array_equal(f, want)
