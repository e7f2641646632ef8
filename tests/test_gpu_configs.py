"""The benchmark workloads at their full sizes against the oracle, through
the C-ABI: the headline closing (SH icosphere subdivision 7, N = 512, ball
r = 8), the spatially varying closing (subdivision 8, N = 1024, 64^3 FP32
radius lattice), the same headline grid closed with the cube element, and a
served mix of uneven jobs run together in the job engine. Grids after
voxelize, dilate, contour and closing are compared bit for bit, triangle
lists exactly, vertices within 1e-6 relative and unit normals within 2e-5."""
import sys

import numpy as np
import pytest

import _oracle as O
import broxy_amd as B

pytestmark = pytest.mark.gpu

VERT_RTOL = 1e-6
NORMAL_ATOL = 2e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check_mesh(V, Nn, Tt, Cref, R, bmin, cell):
    Vr, Nr, Tr = O.mc(Cref, R, bmin, cell)
    assert V.shape == Vr.shape and Tt.shape == Tr.shape, (V.shape, Vr.shape, Tt.shape, Tr.shape)
    assert np.array_equal(Tt, Tr)
    assert np.allclose(V, Vr, rtol=VERT_RTOL, atol=0)
    assert np.allclose(Nn, Nr, atol=NORMAL_ATOL)
    return len(V), len(Tt)


def _full_pipeline(ctx, P, T, R, r=None, lat=None, se=B.SE_SPHERE):
    """GPU closing + mesh and the oracle's, stage by stage."""
    ctx.load(P, T, R, 0.3)
    bmin, bmax, res, dr, cell = ctx.info()
    sc = None
    if lat is not None:
        L = int(round(lat.size ** (1 / 3.0)))
        ctx.scale_lattice(lat, L)
        sc = O.scale_from_lattice(lat.reshape(L, L, L), 2 * R)
    else:
        ctx.scale_uniform(r)
    ctx.close_grid(se)
    ctx.extract_mesh()
    I = ctx.get_cells(B.GRID_INPUT)
    Iref = O.voxelize(P, T, bmin, float((bmax - bmin).max()), R)
    assert np.array_equal(I, Iref), "voxelized grid differs"
    Dref = O.dilate(Iref, R, r=r or 0, data_res=dr, scale=sc, se=se)
    assert np.array_equal(ctx.get_cells(B.GRID_DILATION), Dref), "dilation differs"
    assert np.array_equal(ctx.get_cells(B.GRID_CONTOUR), O.contour_K(Dref, R)), "contour K differs"
    Cref = O.erode(Iref, Dref, R, r=r or 0, data_res=dr, scale=sc)
    assert np.array_equal(ctx.get_cells(B.GRID_CLOSING), Cref), "closing differs"
    V, Nn, Tt = ctx.mesh()
    return _check_mesh(V, Nn, Tt, Cref, R, bmin, cell)


def test_headline_sh512_ball8(ctx):
    P, T = B.sh_blob(7, seed=2017, lmax=8, amp=0.1)
    say("headline: SH icosphere subdiv 7, N=512, ball r=8")
    nv, nt = _full_pipeline(ctx, P, T, 256, r=8)
    assert nt > 900000 and 2 * nv - nt == 4  # one closed genus-0 surface


def test_headline_grid_cube8(ctx):
    # asymmetric closing: cube dilation, ball erosion (BroxyViewer.cpp:138)
    P, T = B.sh_blob(7, seed=2017, lmax=8, amp=0.1)
    say("SH icosphere subdiv 7, N=512, cube r=8")
    _full_pipeline(ctx, P, T, 256, r=8, se=B.SE_CUBE)


def test_lattice1024(ctx):
    P, T = B.sh_blob(8, seed=2017, lmax=8, amp=0.1)
    lat = B.scale_lattice(64, seed=2017, lo=2.0, hi=24.0, smooth=3)
    say("varying: SH icosphere subdiv 8, N=1024, 64^3 lattice U[2,24] smoothed 3x")
    nv, nt = _full_pipeline(ctx, P, T, 512, lat=lat)
    assert nt > 3000000


# A served mix: grid edges 128..1024, radii 2..24 and lattice fields, run
# together in the job engine (continuous batching, 4 slots), each result
# checked against the oracle.
SERVED = [(128, 2, False), (256, 24, False), (512, 13, True), (1024, 24, False), (128, 9, True),
          (256, 7, False), (512, 2, False), (1024, 0, True), (128, 16, False), (256, 19, True)]


@pytest.fixture(scope="module")
def served_results():
    from broxy_amd import workloads
    pool = workloads._mesh_pool(77, 7)
    jobs = []
    for i, (N, r, field) in enumerate(SERVED):
        P, T = (pool["soups"] if N <= 256 else pool["blobs"])[i % 2]
        lat = pool["lattices"][i % 3] if field else None
        jobs.append({"P": P, "T": T, "R": N // 2, "r": 0 if field else r, "lat": lat})
    say("served mix: %d jobs in the engine" % len(jobs))
    eng = B.Engine(0, slots=4, hbm_budget=200 << 30)
    try:
        ids = [eng.submit(j["P"], j["T"], j["R"], 0.3, radius=j["r"], lattice=j["lat"]) for j in jobs]
        out = [eng.take(i) for i in ids]
    finally:
        eng.close()
    return jobs, out


@pytest.mark.parametrize("k", range(len(SERVED)))
def test_served_mix_against_oracle(served_results, k):
    jobs, out = served_results
    j = jobs[k]
    R = j["R"]
    say("served job %d: N=%d r=%d lattice=%s" % (k, 2 * R, j["r"], j["lat"] is not None))
    bmin, bmax, cell, dr = O.grid_attributes(j["P"], R, 0.3 + 2.0 / (2 * R))
    Iref = O.voxelize(j["P"], j["T"], bmin, float((bmax - bmin).max()), R)
    sc = None
    if j["lat"] is not None:
        sc = O.scale_from_lattice(j["lat"].reshape(64, 64, 64), 2 * R)
    Dref = O.dilate(Iref, R, r=j["r"], data_res=dr, scale=sc)
    Cref = O.erode(Iref, Dref, R, r=j["r"], data_res=dr, scale=sc)
    V, Nn, Tt = out[k]
    _check_mesh(V, Nn, Tt, Cref, R, bmin, cell)
