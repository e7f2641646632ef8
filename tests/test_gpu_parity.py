"""HIP path vs the oracle, stage by stage, through the C-ABI (bit-exact grids,
bit-exact MC index/triangle lists, vertices within 1e-6 relative)."""
import sys

import numpy as np
import pytest

import _oracle as O
import broxy_amd as B
from _util import random_blobs, random_sparse

pytestmark = pytest.mark.gpu

VERT_RTOL = 1e-6
NORMAL_ATOL = 2e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def grid_setup(ctx, R, cells, data_res=None):
    bmin = np.array([-1.0, -0.5, 0.25], np.float32)
    cell = np.float32(2.0 / R)
    bmax = bmin + cell * R
    dr = np.array(data_res if data_res is not None else [R, R, R], np.int32)
    ctx.grid_init(bmin, bmax, R, dr, float(cell))
    ctx.set_cells(B.GRID_INPUT, cells)
    return bmin, float(cell), dr


def test_layout_roundtrip(ctx):
    R = 32
    rng = np.random.default_rng(0)
    cells = rng.integers(0, 256, R ** 3, dtype=np.uint8)
    grid_setup(ctx, R, cells)
    say("layout roundtrip R=32")
    assert np.array_equal(ctx.get_cells(B.GRID_INPUT), cells)


@pytest.mark.parametrize("r", [2, 3, 5, 8])
def test_dilate_erode_uniform(ctx, r):
    R = 32
    I = random_blobs(R, seed=10 + r)
    dr = np.array([R, R - 3, R - 5], np.int32)
    grid_setup(ctx, R, I, dr)
    ctx.scale_uniform(r)
    say("dilate/erode uniform r=%d R=%d" % (r, R))
    ctx.dilate(B.SE_SPHERE)
    D = ctx.get_cells(B.GRID_DILATION)
    Dref = O.dilate(I, R, r=r, data_res=dr)
    assert np.array_equal(D, Dref), "dilation differs in %d cells" % int((D != Dref).sum())
    ctx.erode()
    K = ctx.get_cells(B.GRID_CONTOUR)
    assert np.array_equal(K, O.contour_K(Dref, R))
    C = ctx.get_cells(B.GRID_CLOSING)
    Cref = O.erode(I, Dref, R, r=r, data_res=dr)
    assert np.array_equal(C, Cref), "erosion differs in %d cells" % int((C != Cref).sum())


@pytest.mark.parametrize("R", [16, 80, 144])
def test_dilate_erode_row_widths(ctx, R):
    """Grids of 1, 5 and 9 cell-row words per row. The contour kernels take
    the words of neighbouring rows from the lanes of the wave that hold
    them (64 / W rows per wave) and load the others, so row widths that do
    not divide 64 mix both at every lane position."""
    r = 3
    I = random_blobs(R, seed=40 + R, nballs=8, rmax=4.0 if R < 32 else 8.0, margin=3)
    dr = np.array([R, R - 3, R - 5], np.int32)
    grid_setup(ctx, R, I, dr)
    ctx.scale_uniform(r)
    say("dilate/erode row widths r=%d R=%d" % (r, R))
    ctx.dilate(B.SE_SPHERE)
    D = ctx.get_cells(B.GRID_DILATION)
    Dref = O.dilate(I, R, r=r, data_res=dr)
    assert np.array_equal(D, Dref), "dilation differs in %d cells" % int((D != Dref).sum())
    ctx.erode()
    K = ctx.get_cells(B.GRID_CONTOUR)
    assert np.array_equal(K, O.contour_K(Dref, R))
    C = ctx.get_cells(B.GRID_CLOSING)
    Cref = O.erode(I, Dref, R, r=r, data_res=dr)
    assert np.array_equal(C, Cref), "erosion differs in %d cells" % int((C != Cref).sum())


@pytest.mark.parametrize("r", [1, 11, 15, 16, 17, 23, 31])
def test_dilate_erode_radius_range(ctx, r):
    """Radii across the supported range (both the compiled chord tables and
    the runtime-table path), a grid whose side is not a power of two
    (R = 48: 3 row words), a few isolated seeds and a clipped data region."""
    R = 48
    I = random_sparse(R, seed=100 + r, density=0.00003)
    dr = np.array([R - 2, R, R - 7], np.int32)
    grid_setup(ctx, R, I, dr)
    ctx.scale_uniform(r)
    say("dilate/erode radius range r=%d R=%d" % (r, R))
    ctx.dilate(B.SE_SPHERE)
    Dref = O.dilate(I, R, r=r, data_res=dr)
    D = ctx.get_cells(B.GRID_DILATION)
    assert np.array_equal(D, Dref), "dilation differs in %d cells" % int((D != Dref).sum())
    ctx.erode()
    C = ctx.get_cells(B.GRID_CLOSING)
    Cref = O.erode(I, Dref, R, r=r, data_res=dr)
    assert np.array_equal(C, Cref), "erosion differs in %d cells" % int((C != Cref).sum())


@pytest.mark.parametrize("R,r", [(48, 32), (64, 47), (96, 63), (96, 127)])
def test_large_radius_uniform(ctx, R, r):
    """Radii above the chord tables (32..127) run the distance-transform
    passes: dilation, K, erosion and the mesh equal the oracle's, with a
    clipped data region. The sources are a sparse cluster near one corner,
    so that the ball reaches the grid's faces but not the far corner and
    the closing is neither empty nor the whole grid."""
    from _util import cells_from_fine
    N = 2 * R
    rng = np.random.default_rng(300 + r)
    f = np.zeros((N, N, N), bool)
    f[6:20, 8:22, 5:18] = rng.random((14, 14, 13)) < 0.05
    f[30:34, 10:13, 12:15] = True
    I = cells_from_fine(f)
    # a clipped data region, except at the largest radius, where the clipped
    # faces' contour would erode everything
    dr = np.array([R - 3, R, R - 5] if r < 100 else [R, R, R], np.int32)
    bmin, cell, _ = grid_setup(ctx, R, I, dr)
    ctx.scale_uniform(r)
    say("large radius R=%d r=%d" % (R, r))
    ctx.close_grid(B.SE_SPHERE)
    Dref = O.dilate(I, R, r=r, data_res=dr)
    D = ctx.get_cells(B.GRID_DILATION)
    assert np.array_equal(D, Dref), "dilation differs in %d cells" % int((D != Dref).sum())
    assert np.array_equal(ctx.get_cells(B.GRID_CONTOUR), O.contour_K(Dref, R))
    Cref = O.erode(I, Dref, R, r=r, data_res=dr)
    C = ctx.get_cells(B.GRID_CLOSING)
    assert np.array_equal(C, Cref), "erosion differs in %d cells" % int((C != Cref).sum())
    assert 0 < int((Cref == 255).sum()) < R ** 3
    ctx.extract_mesh()
    V, Nn, T = ctx.mesh()
    Vr, Nr, Tr = O.mc(Cref, R, bmin, cell)
    assert np.array_equal(T, Tr) and np.allclose(V, Vr, rtol=1e-6, atol=0)


def test_large_radius_scale_grid(ctx):
    """A per-voxel radius field reaching above 31 (values 2..80): the
    dilation takes each query's radius, the erosion weighs every source by
    its own; both equal the oracle's. Then the same context closes with a
    small uniform radius again (chord path), and dilates with the cube at
    radius 40 and with the field, all equal to the oracle."""
    R = 48
    N = 2 * R
    I = random_blobs(R, seed=77, nballs=5, rmin=2.0, rmax=6.0, margin=4)
    rng = np.random.default_rng(78)
    sc = rng.integers(2, 81, N ** 3).astype(np.int8)
    sc[rng.random(N ** 3) < 0.6] = 3
    grid_setup(ctx, R, I)
    ctx.scale_grid(sc)
    say("large radius scale grid R=%d" % R)
    ctx.close_grid(B.SE_SPHERE)
    Dref = O.dilate(I, R, scale=sc)
    assert np.array_equal(ctx.get_cells(B.GRID_DILATION), Dref)
    assert np.array_equal(ctx.get_cells(B.GRID_CLOSING), O.erode(I, Dref, R, scale=sc))
    ctx.scale_uniform(6)
    ctx.close_grid(B.SE_SPHERE)
    Dref = O.dilate(I, R, r=6)
    assert np.array_equal(ctx.get_cells(B.GRID_DILATION), Dref)
    assert np.array_equal(ctx.get_cells(B.GRID_CLOSING), O.erode(I, Dref, R, r=6))
    # the cube above radius 31: the per-axis (Chebyshev) transform
    ctx.scale_uniform(40)
    ctx.dilate(B.SE_CUBE)
    assert np.array_equal(ctx.get_cells(B.GRID_DILATION), O.dilate(I, R, r=40, se=1))
    ctx.scale_grid(sc)
    ctx.dilate(B.SE_CUBE)
    assert np.array_equal(ctx.get_cells(B.GRID_DILATION), O.dilate(I, R, scale=sc, se=1))


@pytest.mark.parametrize("r", [2, 4, 7])
def test_sparse_input(ctx, r):
    R = 32
    I = random_sparse(R, seed=r, density=0.004)
    grid_setup(ctx, R, I)
    ctx.scale_uniform(r)
    say("sparse input r=%d" % r)
    ctx.close_grid(B.SE_SPHERE)
    Dref = O.dilate(I, R, r=r)
    assert np.array_equal(ctx.get_cells(B.GRID_DILATION), Dref)
    assert np.array_equal(ctx.get_cells(B.GRID_CLOSING), O.erode(I, Dref, R, r=r))


@pytest.mark.parametrize("R", [32, 48, 64])
def test_cube_dilation(ctx, R):
    # uniform radii run the compiled box kernels (rho = r, 1..31); N = 64
    # has 2 words per row, N = 96 and 128 three and four
    I = random_blobs(R, seed=3 + R)
    grid_setup(ctx, R, I)
    for r in ((1, 2, 3, 5, 8, 13, 21, 31) if R == 32 else (2, 7, 12)):
        ctx.scale_uniform(r)
        say("cube dilation R=%d r=%d" % (R, r))
        ctx.dilate(B.SE_CUBE)
        assert np.array_equal(ctx.get_cells(B.GRID_DILATION), O.dilate(I, R, r=r, se=1)), r


def test_scale_grid_i8(ctx):
    R = 32
    N = 2 * R
    I = random_blobs(R, seed=5)
    grid_setup(ctx, R, I)
    rng = np.random.default_rng(7)
    sc = rng.integers(-3, 8, N ** 3).astype(np.int8)
    ctx.scale_grid(sc)
    say("per-voxel int8 radii")
    ctx.close_grid(B.SE_SPHERE)
    Dref = O.dilate(I, R, scale=sc)
    assert np.array_equal(ctx.get_cells(B.GRID_DILATION), Dref)
    assert np.array_equal(ctx.get_cells(B.GRID_CLOSING), O.erode(I, Dref, R, scale=sc))


def test_scale_lattice(ctx):
    R = 32
    N = 2 * R
    I = random_blobs(R, seed=6)
    grid_setup(ctx, R, I)
    L = 8
    lat = B.scale_lattice(L, seed=11, lo=2.0, hi=7.0, smooth=1)
    ctx.scale_lattice(lat, L)
    sc = O.scale_from_lattice(lat.reshape(L, L, L), N)
    say("lattice radii")
    ctx.close_grid(B.SE_SPHERE)
    Dref = O.dilate(I, R, scale=sc)
    assert np.array_equal(ctx.get_cells(B.GRID_DILATION), Dref)
    assert np.array_equal(ctx.get_cells(B.GRID_CLOSING), O.erode(I, Dref, R, scale=sc))


def test_scale_lattice_wide_clamped(ctx):
    # radii 2..23 over a grid that is not a multiple of the tile sizes, with
    # a data region smaller than the grid on two axes
    R = 80
    N = 2 * R
    I = random_blobs(R, seed=8, nballs=24, rmin=3.0, rmax=14.0)
    dr = [72, R, 61]
    grid_setup(ctx, R, I, data_res=dr)
    L = 10
    lat = B.scale_lattice(L, seed=12, lo=2.0, hi=24.0, smooth=0)
    ctx.scale_lattice(lat, L)
    sc = O.scale_from_lattice(lat.reshape(L, L, L), N)
    assert sc.min() == 2 and sc.max() >= 20
    say("lattice radii 2..23, R=80, data region", dr)
    ctx.close_grid(B.SE_SPHERE)
    Dref = O.dilate(I, R, scale=sc, data_res=dr)
    assert np.array_equal(ctx.get_cells(B.GRID_DILATION), Dref)
    assert np.array_equal(ctx.get_cells(B.GRID_CLOSING), O.erode(I, Dref, R, scale=sc, data_res=dr))


def test_scale_lattice_large_magnitudes(ctx):
    # lattice values far below zero (radius 2 there): the interpolation sums
    # exceed 32 bits and take the 64-bit path
    R = 48
    N = 2 * R
    I = random_blobs(R, seed=14, nballs=10, rmin=3.0, rmax=10.0)
    grid_setup(ctx, R, I)
    L = 6
    rng = np.random.default_rng(15)
    lat = rng.uniform(2.0, 16.0, L ** 3).astype(np.float32)
    lat[rng.random(L ** 3) < 0.3] = -4.0e6
    ctx.scale_lattice(lat, L)
    sc = O.scale_from_lattice(lat.reshape(L, L, L), N)
    say("lattice with large negative values")
    ctx.close_grid(B.SE_SPHERE)
    Dref = O.dilate(I, R, scale=sc)
    assert np.array_equal(ctx.get_cells(B.GRID_DILATION), Dref)
    assert np.array_equal(ctx.get_cells(B.GRID_CLOSING), O.erode(I, Dref, R, scale=sc))


def test_scale_grid_cube(ctx):
    R = 32
    N = 2 * R
    I = random_blobs(R, seed=9)
    grid_setup(ctx, R, I)
    rng = np.random.default_rng(13)
    sc = rng.integers(0, 6, N ** 3).astype(np.int8)
    ctx.scale_grid(sc)
    say("per-voxel int8 radii, cube dilation")
    ctx.close_grid(B.SE_CUBE)
    Dref = O.dilate(I, R, scale=sc, se=1)
    assert np.array_equal(ctx.get_cells(B.GRID_DILATION), Dref)
    assert np.array_equal(ctx.get_cells(B.GRID_CLOSING), O.erode(I, Dref, R, scale=sc))


def test_scale_field_changes(ctx):
    # a new field (or a uniform radius) replaces the radius planes of the last one
    R = 48
    N = 2 * R
    I = random_blobs(R, seed=10)
    grid_setup(ctx, R, I)
    L = 6
    for seed, lo, hi in ((21, 2.0, 9.0), (22, 5.0, 14.0)):
        lat = B.scale_lattice(L, seed=seed, lo=lo, hi=hi, smooth=1)
        ctx.scale_lattice(lat, L)
        sc = O.scale_from_lattice(lat.reshape(L, L, L), N)
        say("lattice seed", seed)
        ctx.dilate(B.SE_SPHERE)
        Dref = O.dilate(I, R, scale=sc)
        assert np.array_equal(ctx.get_cells(B.GRID_DILATION), Dref)
        ctx.erode()
        assert np.array_equal(ctx.get_cells(B.GRID_CLOSING), O.erode(I, Dref, R, scale=sc))
    ctx.scale_uniform(6)
    say("uniform after lattice")
    ctx.close_grid(B.SE_SPHERE)
    Dref = O.dilate(I, R, r=6)
    assert np.array_equal(ctx.get_cells(B.GRID_CLOSING), O.erode(I, Dref, R, r=6))


def test_voxelize_soup(ctx):
    P, T = B.sphere_soup(8, 3, seed=99)
    R = 64
    say("voxelize soup R=64")
    ctx.load(P, T, R, 0.3)
    bmin, bmax, res, dr, cell = ctx.info()
    rb, rB, rc, rd = O.grid_attributes(P, R, 0.3 + 2.0 / (2 * R))
    assert np.array_equal(bmin, rb) and np.array_equal(bmax, rB) and cell == np.float32(rc)
    assert np.array_equal(dr, rd)
    Iref = O.voxelize(P, T, rb, float((rB - rb).max()), R)
    I = ctx.get_cells(B.GRID_INPUT)
    assert np.array_equal(I, Iref), "voxelization differs in %d cells" % int((I != Iref).sum())


@pytest.mark.parametrize("R,kind", [(64, "soup"), (128, "blob"), (48, "soup")])
def test_voxelize_multislice(ctx, R, kind):
    """MultiSlicing (the reference's alternative voxelizer): the same grid
    attributes, the parity grid bit-exact with the oracle's; then closed and
    meshed like any input grid."""
    P, T = B.sphere_soup(12, 3, seed=41) if kind == "soup" else B.sh_blob(5, seed=42)
    say("voxelize multislice R=%d %s" % (R, kind))
    ctx.load_mesh(P, T)
    margin = 0.3 + 2.0 / (2 * R)
    ctx.voxelize(R, margin, multislice=True)
    bmin, bmax, res, dr, cell = ctx.info()
    rb, rB, rc, rd = O.grid_attributes(P, R, margin)
    assert np.array_equal(bmin, rb) and np.array_equal(bmax, rB) and cell == np.float32(rc)
    Iref = O.voxelize_multislice(P, T, rb, float((rB - rb).max()), R)
    I = ctx.get_cells(B.GRID_INPUT)
    assert np.array_equal(I, Iref), "voxelization differs in %d cells" % int((I != Iref).sum())
    assert int((Iref != 0).sum()) > 100
    ctx.scale_uniform(3)
    ctx.close_grid(B.SE_SPHERE)
    Dref = O.dilate(Iref, R, r=3, data_res=rd)
    assert np.array_equal(ctx.get_cells(B.GRID_CLOSING), O.erode(Iref, Dref, R, r=3, data_res=rd))
    # the conservative path afterwards still starts from clear marks
    ctx.load(P, T, R, 0.3)
    assert np.array_equal(ctx.get_cells(B.GRID_INPUT), O.voxelize(P, T, rb, float((rB - rb).max()), R))


def _mesh_compare(ctx, C, R, bmin, cell):
    V, N, T = ctx.mesh()
    Vr, Nr, Tr = O.mc(C, R, bmin, cell)
    assert V.shape == Vr.shape and T.shape == Tr.shape
    assert np.array_equal(T, Tr)
    assert np.allclose(V, Vr, rtol=VERT_RTOL, atol=0)
    assert np.allclose(N, Nr, atol=NORMAL_ATOL)
    return V, T


def test_mesh_blobs(ctx):
    R = 32
    I = random_blobs(R, seed=21, nballs=8)
    bmin, cell, dr = grid_setup(ctx, R, I)
    ctx.scale_uniform(3)
    ctx.close_grid(B.SE_SPHERE)
    ctx.extract_mesh()
    say("mesh of closed blobs R=32")
    C = ctx.get_cells(B.GRID_CLOSING)
    _mesh_compare(ctx, C, R, bmin, cell)


def test_stage_times(ctx):
    """Per-stage timings: every stage is timed, and each stream kernel alone
    takes no longer than its pass (tile selection + stream)."""
    R = 32
    I = random_blobs(R, seed=22, nballs=8)
    grid_setup(ctx, R, I)
    ctx.scale_uniform(4)
    say("stage times R=32")
    ctx.set_timing(True)
    ctx.close_grid(B.SE_SPHERE)
    ctx.extract_mesh()
    ms = dict(zip(B.Context.STAGES, ctx.stage_times()))
    ctx.set_timing(False)
    assert len(ms) == 16 and all(v >= 0 for v in ms.values()), ms
    for p in ("dilate_coarse", "dilate_fine", "erode_coarse", "erode_fine"):
        assert ms[p + "_stream"] <= ms[p], (p, ms)


def test_config0_end_to_end(ctx):
    P, T = B.sphere_soup(64, 3, seed=1234)
    R = 64
    say("config-0 end to end (N=128, r=4)")
    ctx.load(P, T, R, 0.3)
    bmin, bmax, res, dr, cell = ctx.info()
    ctx.scale_uniform(4)
    ctx.close_grid(B.SE_SPHERE)
    ctx.extract_mesh()
    I = ctx.get_cells(B.GRID_INPUT)
    Iref = O.voxelize(P, T, bmin, float((bmax - bmin).max()), R)
    assert np.array_equal(I, Iref)
    Dref = O.dilate(Iref, R, r=4, data_res=dr)
    assert np.array_equal(ctx.get_cells(B.GRID_DILATION), Dref)
    Cref = O.erode(Iref, Dref, R, r=4, data_res=dr)
    assert np.array_equal(ctx.get_cells(B.GRID_CLOSING), Cref)
    V, Tt = _mesh_compare(ctx, Cref, R, bmin, cell)
    assert len(Tt) > 1000


def test_config0_golden(ctx):
    """The GPU path reproduces the committed oracle record of the config-0
    workload (seed 2017) without running the oracle."""
    import hashlib
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_config0.json")))
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    P, T = B.sphere_soup(64, 3, seed=2017)
    say("config-0 golden record (seed 2017)")
    ctx.load(P, T, 64, 0.3)
    bmin, bmax, res, dr, cell = ctx.info()
    assert [float(x) for x in bmin] == g["bmin"] and [int(x) for x in dr] == g["data_res"]
    ctx.scale_uniform(4)
    ctx.close_grid(B.SE_SPHERE)
    ctx.extract_mesh()
    assert sha(ctx.get_cells(B.GRID_INPUT)) == g["I"]
    assert sha(ctx.get_cells(B.GRID_DILATION)) == g["D"]
    assert sha(ctx.get_cells(B.GRID_CLOSING)) == g["C"]
    V, Nn, Tt = ctx.mesh()
    assert (len(V), len(Tt)) == (g["nV"], g["nT"])
    assert sha(np.ascontiguousarray(Tt, np.uint32)) == g["T"]


def test_facade_driver_matches_golden(tmp_path):
    """The C++ facade, used as the reference's viewer uses GMorpho (Load,
    ScaleField global scale, Update -> Mesh), reproduces the oracle's record."""
    import hashlib
    import json
    import os
    import subprocess
    g = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_config0.json")))
    exe = os.path.join(B.LIBDIR, "bin", "facade_closing")
    say("C++ facade driver, config-0 seed 2017")
    pre = str(tmp_path / "facade")
    p = subprocess.run([exe, pre, "2017"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    nv, nt = map(int, p.stdout.split())
    assert (nv, nt) == (g["nV"], g["nT"])
    h = lambda f: hashlib.sha256(open(f, "rb").read()).hexdigest()
    assert h(pre + "_closing.u8") == g["C"]
    assert h(pre + "_tris.u32") == g["T"]


def test_pyramid(ctx):
    R = 32
    I = random_blobs(R, seed=8)
    grid_setup(ctx, R, I)
    ctx.scale_uniform(6)
    say("pyramid uniform")
    assert np.array_equal(ctx.build_pyramid(B.GRID_INPUT), O.pyramid(I, R, r=6))
    sc = np.random.default_rng(1).integers(-5, 40, (2 * R) ** 3).astype(np.int8)
    ctx.scale_grid(sc)
    say("pyramid int8 scale")
    assert np.array_equal(ctx.build_pyramid(B.GRID_INPUT), O.pyramid(I, R, scale=sc))


def test_empty_grid(ctx):
    R = 32
    grid_setup(ctx, R, np.zeros(R ** 3, np.uint8))
    ctx.scale_uniform(5)
    say("empty grid")
    ctx.close_grid(B.SE_SPHERE)
    assert not ctx.get_cells(B.GRID_CLOSING).any()
    ctx.extract_mesh()
    V, N, T = ctx.mesh()
    assert len(V) == 0 and len(T) == 0


def _case91_grid(R):
    """Fine pattern (x+y+z) % 4 != 2 in an inner box: about one cube in eight
    is case 91, whose 5th listed triangle is not counted (tables.h:412)."""
    N = 2 * R
    z, y, x = np.mgrid[0:N, 0:N, 0:N]
    f = (x + y + z) % 4 != 2
    m = np.zeros_like(f)
    m[8:N - 8, 8:N - 8, 8:N - 8] = True
    from _util import cells_from_fine
    return cells_from_fine(f & m)


def test_mesh_overflow_with_late_triangles(ctx):
    """A mesh beyond the initial output (4 N^2 vertices, 8 N^2 triangles)
    with over 20,000 late (case-91) triangles, more than the late list
    holds at first: reading the mesh grows the output and the list and emits
    again from an empty late list. Checked against the oracle twice in a row
    (the second extraction starts from the grown buffers)."""
    R = 32
    N = 2 * R
    C0 = _case91_grid(R)
    bmin, cell, dr = grid_setup(ctx, R, C0)
    ctx.scale_uniform(0)  # radius <= 1: the closing is the input
    for k in range(2):
        say("late-triangle overflow R=%d, extraction %d" % (R, k))
        ctx.close_grid(B.SE_SPHERE)
        ctx.extract_mesh()
        V, T = _mesh_compare(ctx, C0, R, bmin, cell)
        assert len(V) > 4 * N * N and len(T) > 8 * N * N


def test_largest_grid_closed_surface():
    """The largest supported grid (base_res 1024, N = 2048): a sphere closes
    to one closed genus-0 surface (2V - T = 4), every index in range, unit
    normals; the oracle is too slow at this size, so these invariants stand
    in for it."""
    c = B.Context(0)
    P, T = B.icosphere(5)
    say("base_res 1024 sphere")
    c.load(P, T, 1024, 0.3)
    c.scale_uniform(8)
    c.close_grid(B.SE_SPHERE)
    c.extract_mesh()
    V, N, Tm = c.mesh()
    c.close()
    assert len(Tm) > 0 and Tm.max() < len(V)
    assert 2 * len(V) - len(Tm) == 4
    n = np.linalg.norm(N[:: max(1, len(N) // 100000)], axis=1)
    assert np.allclose(n, 1.0, atol=1e-5)


BIL_ATOL = 0.0  # filtered unit normals: the same float operations as the oracle


def test_bilateral_normals(ctx):
    """gm_filter_normals against the oracle's restatement of the reference's
    filter (GMorpho.cu:6539-6620, 6957-6978), from the same unit normals;
    twice, to show that the result is the same every run."""
    P, T = B.sphere_soup(64, 3, seed=1234)
    R = 64
    ctx.load(P, T, R, 0.3)
    ctx.scale_uniform(4)
    ctx.close_grid(B.SE_SPHERE)
    ctx.extract_mesh()
    V, N0, Tt = ctx.mesh()
    for iters in (1, 2, 20):
        say("bilateral normal filtering, %d iterations" % iters)
        ctx.extract_mesh()
        ctx.filter_normals(iters)
        V1, N1, T1 = ctx.mesh()
        assert np.array_equal(V1, V) and np.array_equal(T1, Tt)
        ref = O.bilateral_normals(V, N0, Tt, iters)
        err = float(np.abs(N1 - ref).max())
        say("  max |n - n_oracle| = %.3g" % err)
        assert err <= BIL_ATOL
        ctx.extract_mesh()
        ctx.filter_normals(iters)
        assert np.array_equal(ctx.mesh()[1], N1)


def _run_facade(tmp_path, mode):
    import os
    import subprocess
    exe = os.path.join(B.LIBDIR, "bin", "facade_closing")
    pre = str(tmp_path / ("facade_" + mode))
    p = subprocess.run([exe, pre, "2017", mode], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    nv, nt = map(int, p.stdout.split())
    C = np.fromfile(pre + "_closing.u8", np.uint8)
    Tt = np.fromfile(pre + "_tris.u32", np.uint32).reshape(-1, 3)
    pn = np.fromfile(pre + "_mesh.pn", np.float32).reshape(-1, 6)
    assert (len(pn), len(Tt)) == (nv, nt)
    return C, Tt, pn[:, :3], pn[:, 3:]


def _facade_oracle(se):
    P, T = B.sphere_soup(64, 3, seed=2017)
    R = 64
    bmin, bmax, cell, dr = O.grid_attributes(P, R, 0.3 + 2.0 / (2 * R))
    I = O.voxelize(P, T, bmin, float((bmax - bmin).max()), R)
    D = O.dilate(I, R, r=4, data_res=dr, se=se)
    C = O.erode(I, D, R, r=4, data_res=dr)
    return (C,) + O.mc(C, R, bmin, cell)


def test_facade_viewer_defaults(tmp_path):
    """The viewer's default configuration through the C++ facade: asymmetric
    (cube) closing and bilateral normal filtering, 200 iterations
    (BroxyViewer.cpp:66, 125, 138; GMorpho.cu:7809-7812), checked against the
    oracle; the mesh also goes through the facade's OFF store and load."""
    say("C++ facade driver, viewer defaults (cube closing + bilateral normals)")
    C, Tt, V, Nn = _run_facade(tmp_path, "viewer")
    Cr, Vr, Nr, Tr = _facade_oracle(1)
    assert np.array_equal(C, Cr)
    assert np.array_equal(Tt, Tr)
    assert np.allclose(V, Vr, rtol=VERT_RTOL, atol=0)
    # the filter starts from the GPU's unfiltered unit normals (within 2e-5 of
    # the oracle's, summed in another order): 200 iterations at sigma_n = 0.2
    # amplify such differences, so the oracle filters the same starting normals
    c = B.Context(0)
    try:
        P, T = B.sphere_soup(64, 3, seed=2017)
        c.load(P, T, 64, 0.3)
        c.scale_uniform(4)
        c.close_grid(B.SE_CUBE)
        c.extract_mesh()
        V0, N0, T0 = c.mesh()
    finally:
        c.close()
    assert np.array_equal(V0, V) and np.array_equal(T0, Tt)
    assert np.allclose(N0, Nr, atol=NORMAL_ATOL)
    ref = O.bilateral_normals(V0, N0, T0, 200)
    err = float(np.abs(Nn - ref).max())
    say("  max |n - n_oracle| after 200 iterations = %.3g" % err)
    assert err <= BIL_ATOL


def test_facade_fixed_radius_update(tmp_path):
    """GMorpho::Update(Mesh&, float se_size) (GMorpho.cu:7819-7845) through
    the facade: cube dilation and ball erosion at floor(2 se_size / cell)."""
    say("C++ facade driver, Update(mesh, se_size)")
    C, Tt, V, Nn = _run_facade(tmp_path, "fixed")
    Cr, Vr, Nr, Tr = _facade_oracle(1)
    assert np.array_equal(C, Cr)
    assert np.array_equal(Tt, Tr)
    assert np.allclose(V, Vr, rtol=VERT_RTOL, atol=0)
    assert np.allclose(Nn, Nr, atol=NORMAL_ATOL)


def test_scale_brush_edits_and_closing(ctx):
    """ScaleField brush strokes on the device (UpdateScaleGrid,
    ScaleField.cu:81-141) against the oracle after every stroke, and the
    closings with the edited field; the second closing follows strokes that
    rebuilt the radius planes of the touched bricks only."""
    P, T = B.sphere_soup(64, 3, seed=99)
    R = 64
    N = 2 * R
    ctx.load(P, T, R, 0.3)
    bmin, bmax, res, dr, cell = ctx.info()
    fc = np.float32(0.5 * cell)
    I = ctx.get_cells(B.GRID_INPUT)
    say("scale brush: fill 3, MAX stroke")
    ctx.scale_fill(3)
    g = np.full(N ** 3, 3, np.int8)
    assert np.array_equal(ctx.scale_grid_get(), g)
    strokes = [((0.5, 0.5, 0.5), 0.12, 0.2, 0.0, B.BRUSH_MAX)]
    later = [((0.4, 0.6, 0.5), 0.05, 0.1, 0.0, B.BRUSH_MIN),
             ((0.62, 0.41, 0.55), 0.08, 0.12, 0.03, B.BRUSH_ZERO),
             ((0.3, 0.35, 0.7), 0.2, 0.21, 0.0, B.BRUSH_MAX)]

    def apply(st):
        nonlocal g
        for pos, sc, sup, gs, mode in st:
            ctx.scale_brush(pos, sc, sup, gs, mode)
            g = O.scale_brush(g, N, bmin, fc, np.array(pos, np.float32), sc, sup, gs, mode)
            assert np.array_equal(ctx.scale_grid_get(), g), "brush mode %d differs" % mode

    def closing():
        ctx.close_grid(B.SE_SPHERE)
        Dref = O.dilate(I, R, scale=g, data_res=dr)
        assert np.array_equal(ctx.get_cells(B.GRID_DILATION), Dref)
        assert np.array_equal(ctx.get_cells(B.GRID_CLOSING), O.erode(I, Dref, R, scale=g, data_res=dr))

    apply(strokes)
    assert g.max() > 3  # the stroke raised the field
    closing()
    say("scale brush: MIN, ZERO and MAX strokes, closing again")
    apply(later)
    closing()


def test_facade_brush_strokes(tmp_path):
    """Brush strokes through the facade's ScaleField (AddPoint + UpdateGrid,
    a MAX then a MIN stroke, each followed by Update) against the oracle's
    brush and closing."""
    import os
    say("C++ facade driver, brush strokes")
    C, Tt, V, Nn = _run_facade(tmp_path, "brush")
    P, T = B.sphere_soup(64, 3, seed=2017)
    R = 64
    N = 2 * R
    bmin, bmax, cell, dr = O.grid_attributes(P, R, 0.3 + 2.0 / (2 * R))
    fc = np.float32(0.5 * cell)
    g = np.full(N ** 3, int(np.floor(np.float32(4.25) * fc / fc)), np.int8)
    g = O.scale_brush(g, N, bmin, fc, np.array([0.5, 0.5, 0.5], np.float32), 0.12, 0.2, 4.25 * fc, B.BRUSH_MAX)
    g = O.scale_brush(g, N, bmin, fc, np.array([0.35, 0.6, 0.45], np.float32), 0.05, 0.09, 4.25 * fc, B.BRUSH_MIN)
    got = np.fromfile(str(tmp_path / "facade_brush_scale.i8"), np.int8)
    assert np.array_equal(got, g)
    I = O.voxelize(P, T, bmin, float((bmax - bmin).max()), R)
    D = O.dilate(I, R, scale=g, data_res=dr)
    Cr = O.erode(I, D, R, scale=g, data_res=dr)
    assert np.array_equal(C, Cr)
    Vr, Nr, Tr = O.mc(Cr, R, bmin, cell)
    assert np.array_equal(Tt, Tr)
    assert np.allclose(V, Vr, rtol=VERT_RTOL, atol=0)
    # the viewer's proxy (BroxyViewer.cpp:1060-1125): the facade's Decimator
    # loop on the closing mesh and the grid from GetScaleGridCPU equals the
    # library's on the same input, and is a smaller closed triangle mesh
    d = B.Decimator(V, Tt, 20, 0.001, use_linear_constraints=True, max_edge_length=1.0,
                    bbox_min=np.asarray(bmin, np.float32), res=(N, N, N), cell=float(fc), scale_grid=g, hc_iters=3)
    more = True
    while more and d.state()[0] > 20 and d.state()[2] < 0.001:
        more = d.optimize()
    Pp, Tp = d.mesh()
    d.close()
    tok = open(str(tmp_path / "facade_brush_proxy.off")).read().split()
    nv, nf = int(tok[1]), int(tok[2])
    Pf = np.array(tok[4:4 + 3 * nv], np.float32).reshape(nv, 3)
    Ff = np.array(tok[4 + 3 * nv:], np.int64).reshape(nf, 4)
    assert np.array_equal(Ff[:, 1:], Tp.astype(np.int64)) and np.array_equal(Pf, Pp)
    assert 0 < nf < len(Tt)
    edges = np.sort(np.concatenate([Tp[:, [0, 1]], Tp[:, [1, 2]], Tp[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(edges, axis=0, return_counts=True)
    assert np.all(cnt == 2)  # every edge shared by exactly two faces


def test_graph_replays_match_eager(ctx):
    """gm_close / gm_extract_mesh replayed as captured graphs equal the same
    steps launched eagerly, replay after replay, with eager work (grid
    uploads and reads) between the replays, across a radius change (a new
    capture) and with per-voxel radii."""
    R = 48
    grids = [random_blobs(R, seed=60 + k, nballs=8) for k in range(3)]

    def step(c, I, r=None, sc=None):
        c.set_cells(B.GRID_INPUT, I)
        if sc is not None:
            c.scale_grid(sc)
        else:
            c.scale_uniform(r)
        c.close_grid(B.SE_SPHERE)
        c.extract_mesh()
        V, Nn, T = c.mesh()
        return c.get_cells(B.GRID_CLOSING), V, Nn, T

    eager = B.Context(0)
    try:
        eager.set_graphs(False)
        bmin, cell, dr = grid_setup(eager, R, grids[0])
        grid_setup(ctx, R, grids[0])
        ctx.set_graphs(True)
        sc = np.random.default_rng(3).integers(1, 7, (2 * R) ** 3).astype(np.int8)
        plan = [(0, 4, None), (1, 4, None), (2, 4, None), (0, 4, None), (1, 6, None), (2, 6, None), (0, None, sc),
                (1, None, sc), (2, None, sc)]
        for k, (g, r, s_) in enumerate(plan):
            say("graph replay %d: grid %d radius %s" % (k, g, r if r is not None else "per-voxel"))
            got = step(ctx, grids[g], r, s_)
            ref = step(eager, grids[g], r, s_)
            for a, b in zip(got, ref):
                assert np.array_equal(a, b)
    finally:
        eager.close()


def test_update_matches_close_then_extract(ctx):
    """gm_update (closing and marching cubes as one captured sequence) gives
    the closing grid and mesh of gm_close then gm_extract_mesh on an eager
    context, replay after replay with eager work between the replays, across
    a radius change, per-voxel radii, the cube element, a radius <= 1 (no
    morphology) and a radius above the chord tables; the closing is also
    the oracle's."""
    R = 48
    grids = [random_blobs(R, seed=70 + k, nballs=8) for k in range(2)]
    sc = np.random.default_rng(5).integers(1, 7, (2 * R) ** 3).astype(np.int8)

    def setup(c, I, r, s_):
        c.set_cells(B.GRID_INPUT, I)
        if s_ is not None:
            c.scale_grid(s_)
        else:
            c.scale_uniform(r)

    def read(c):
        V, Nn, T = c.mesh()
        return c.get_cells(B.GRID_CLOSING), V, Nn, T

    eager = B.Context(0)
    try:
        eager.set_graphs(False)
        bmin, cell, dr = grid_setup(eager, R, grids[0])
        grid_setup(ctx, R, grids[0])
        ctx.set_graphs(True)
        plan = [(0, 4, None, B.SE_SPHERE), (1, 4, None, B.SE_SPHERE), (0, 4, None, B.SE_SPHERE),
                (1, 6, None, B.SE_CUBE), (0, None, sc, B.SE_SPHERE), (1, 1, None, B.SE_SPHERE),
                (0, 40, None, B.SE_SPHERE)]
        for k, (g, r, s_, se) in enumerate(plan):
            say("update %d: grid %d radius %s se %d" % (k, g, r if r is not None else "per-voxel", se))
            setup(ctx, grids[g], r, s_)
            ctx.update(se)
            got = read(ctx)
            setup(eager, grids[g], r, s_)
            eager.close_grid(se)
            eager.extract_mesh()
            ref = read(eager)
            for a, b in zip(got, ref):
                assert np.array_equal(a, b)
            if se == B.SE_SPHERE and s_ is None and r > 1:
                Dref = O.dilate(grids[g], R, r=r, data_res=dr)
                assert np.array_equal(got[0], O.erode(grids[g], Dref, R, r=r, data_res=dr))
    finally:
        eager.close()


def test_graph_replay_new_lattice_values(ctx):
    """A new lattice of the same size and radius range keeps the captured
    closing (its launch arguments do not change): the replay must read the
    new values. Each field's closing equals the eager context's and the
    oracle's."""
    R = 32
    N = 2 * R
    I = random_blobs(R, seed=91, nballs=6)
    lat_a = B.scale_lattice(8, seed=5, lo=2.0, hi=9.0, smooth=1)
    rng = np.random.default_rng(6)
    lat_b = rng.permutation(lat_a.reshape(-1)).reshape(lat_a.shape).astype(np.float32)  # same range
    eager = B.Context(0)
    try:
        eager.set_graphs(False)
        grid_setup(eager, R, I)
        grid_setup(ctx, R, I)
        ctx.set_graphs(True)
        for k, lat in enumerate((lat_a, lat_b, lat_a, lat_b)):
            say("graph replay with lattice %d" % k)
            out = []
            for c in (ctx, eager):
                c.scale_lattice(lat, 8)
                c.close_grid(B.SE_SPHERE)
                out.append(c.get_cells(B.GRID_CLOSING))
            assert np.array_equal(out[0], out[1]), "replay %d differs from eager" % k
            sc = O.scale_from_lattice(lat.reshape(8, 8, 8), N)
            Dref = O.dilate(I, R, scale=sc)
            assert np.array_equal(out[0], O.erode(I, Dref, R, scale=sc))
    finally:
        eager.close()


def test_graph_repeat_mode_matches_eager(ctx):
    """Capture-on-repeat: the first closing runs eagerly, the second (same
    launch sequence, new grid contents) is captured, the later ones replay;
    every one equals the eager context's result."""
    R = 48
    grids = [random_blobs(R, seed=70 + k, nballs=8) for k in range(4)]
    eager = B.Context(0)
    try:
        eager.set_graphs(False)
        grid_setup(eager, R, grids[0])
        grid_setup(ctx, R, grids[0])
        ctx.set_graphs("repeat")
        eager.scale_uniform(5)
        ctx.scale_uniform(5)
        for k, I in enumerate(grids):
            say("graph repeat mode: closing %d" % k)
            out = []
            for c in (ctx, eager):
                c.set_cells(B.GRID_INPUT, I)
                c.close_grid(B.SE_SPHERE)
                c.extract_mesh()
                out.append((c.get_cells(B.GRID_CLOSING),) + tuple(c.mesh()))
            for a, b in zip(*out):
                assert np.array_equal(a, b)
    finally:
        ctx.set_graphs(True)
        eager.close()


def test_mc_block_list_grows(ctx):
    """One voxel in every 8^3-cell Morton block of a 128^3-cell grid: all
    4096 blocks are listed, twice the per-block buffers' initial capacity,
    so the first extraction is flagged, the buffers grow and the mesh is
    extracted again; checked against the oracle, then once more (grown)."""
    R = 128
    N = 2 * R
    f = np.zeros((N, N, N), bool)
    f[5::16, 7::16, 9::16] = True
    from _util import cells_from_fine
    C0 = cells_from_fine(f)
    bmin, cell, dr = grid_setup(ctx, R, C0)
    ctx.scale_uniform(0)
    for k in range(2):
        say("block-list growth R=%d, extraction %d" % (R, k))
        ctx.close_grid(B.SE_SPHERE)
        ctx.extract_mesh()
        V, T = _mesh_compare(ctx, C0, R, bmin, cell)
        assert len(T) == 8 * 16 ** 3


def test_autotune_shapes(ctx, tmp_path):
    """gm_autotune runs every launch shape of the streaming passes on the
    grid (each must give the same closing, or it fails), records the fastest
    and writes it; the closing with the recorded shape equals the oracle's,
    and the file reads back."""
    R = 64
    I = random_blobs(R, seed=77, nballs=10)
    bmin, cell, dr = grid_setup(ctx, R, I, [R, R - 2, R - 5])
    ctx.scale_uniform(6)
    path = str(tmp_path / "tune.jsonl")
    say("autotune R=64 r=6")
    z, w = ctx.autotune(B.SE_SPHERE, 2, path)
    assert z in (16, 32) and w in (2, 4, 8)
    lines = open(path).read().strip().splitlines()
    assert any('"base_res": 64, "radius": 6, "se": 0' in ln for ln in lines)
    B.tune_load(path)
    ctx.close_grid(B.SE_SPHERE)
    Dref = O.dilate(I, R, r=6, data_res=dr)
    assert np.array_equal(ctx.get_cells(B.GRID_DILATION), Dref)
    assert np.array_equal(ctx.get_cells(B.GRID_CLOSING), O.erode(I, Dref, R, r=6, data_res=dr))
