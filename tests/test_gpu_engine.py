"""Job engine (batched / served use): every job's mesh equals the same job run
alone on one context, bit for bit, whatever else is on the device; contexts
reused across grid sizes stay exact against the oracle."""
import hashlib
import sys

import numpy as np
import pytest

import _oracle as O
import broxy_amd as B
from _util import random_blobs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def h(*arrs):
    m = hashlib.sha256()
    for a in arrs:
        m.update(np.ascontiguousarray(a).tobytes())
    return m.hexdigest()


def jobs(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        P, T = B.sphere_soup(int(rng.integers(8, 33)), 3, seed=int(rng.integers(1 << 30)))
        base_res = int(rng.choice([32, 48, 64, 96]))
        r = int(rng.integers(2, 13))
        lat = None
        if i % 5 == 4:
            lat = B.scale_lattice(8, seed=i, lo=2.0, hi=9.0, smooth=1)
        out.append((P, T, base_res, r, lat))
    return out


def alone(ctx, P, T, base_res, r, lat):
    ctx.load(P, T, base_res, 0.3)
    if lat is not None:
        ctx.scale_lattice(lat, 8)
    else:
        ctx.scale_uniform(r)
    ctx.close_grid(B.SE_SPHERE)
    ctx.extract_mesh()
    return ctx.mesh()


def test_engine_jobs_equal_alone(ctx):
    J = jobs(5, 14)
    say("engine: %d mixed jobs, 4 slots" % len(J))
    eng = B.Engine(0, slots=4, hbm_budget=8 << 30)
    try:
        ids = [eng.submit(P, T, br, 0.3, radius=r, lattice=lat) for P, T, br, r, lat in J]
        mixed = [eng.take(i) for i in ids]
    finally:
        eng.close()
    for k, (P, T, br, r, lat) in enumerate(J):
        V, N, Tt = alone(ctx, P, T, br, r, lat)
        Vm, Nm, Tm = mixed[k]
        assert h(V, N, Tt) == h(Vm, Nm, Tm), "job %d differs from its single-context run" % k


def test_engine_budget_and_errors():
    say("engine: budget admission and job errors")
    eng = B.Engine(0, slots=2, hbm_budget=256 << 20)
    try:
        P, T = B.sphere_soup(8, 3, seed=3)
        with pytest.raises(B.GMorphoError):
            eng.submit(P, T, 512, 0.3, radius=4)  # does not fit 256 MiB
        with pytest.raises(B.GMorphoError):
            eng.submit(P, T, 40, 0.3, radius=4)   # not a multiple of 16
        bad = eng.submit(P, T, 32, 0.3, radius=200)  # radius above the reference's 7-bit range
        ok = eng.submit(P, T, 32, 0.3, radius=3)
        with pytest.raises(B.GMorphoError):
            eng.wait(bad)
        nv, nt = eng.wait(ok)
        assert nt > 0
        reserved, running, queued = eng.stats()
        assert reserved <= 256 << 20 and running == 0 and queued == 0
        # a radius above the chord tables needs distance-transform scratch
        # of its grid (8 N^3 bytes): refused at submit when that cannot fit,
        # never left queued
        with pytest.raises(B.GMorphoError):
            eng.submit(P, T, 192, 0.3, radius=100)
        # two small jobs leave both slots holding buffers; a larger grid that
        # fits only once an idle slot's buffers are released still runs
        ids = [eng.submit(P, T, 32, 0.3, radius=3) for _ in range(2)]
        for i in ids:
            eng.wait(i)
        big = eng.submit(P, T, 240, 0.3, radius=3)
        nv2, nt2 = eng.wait(big)
        assert nt2 > 0
        reserved, running, queued = eng.stats()
        assert reserved <= 256 << 20 and running == 0 and queued == 0
    finally:
        eng.close()


def test_context_reuse_across_sizes(ctx):
    """One context, grids shrinking and growing within its buffers, with
    per-voxel radii: each result equals the oracle."""
    for R in (64, 32, 48, 64, 16):
        N = 2 * R
        I = random_blobs(R, seed=R, nballs=4, rmin=2.0, rmax=min(6.0, R / 6), margin=3)
        bmin = np.array([0.0, 0.0, 0.0], np.float32)
        cell = np.float32(1.0 / R)
        ctx.grid_init(bmin, bmin + cell * R, R, np.array([R, R, R], np.int32), float(cell))
        ctx.set_cells(B.GRID_INPUT, I)
        sc = np.random.default_rng(R).integers(-1, 6, N ** 3).astype(np.int8)
        ctx.scale_grid(sc)
        say("context reuse R=%d" % R)
        ctx.close_grid(B.SE_SPHERE)
        Dref = O.dilate(I, R, scale=sc)
        assert np.array_equal(ctx.get_cells(B.GRID_DILATION), Dref)
        assert np.array_equal(ctx.get_cells(B.GRID_CLOSING), O.erode(I, Dref, R, scale=sc))


def test_engine_borrowed_views_and_repeats(ctx):
    """Borrowed inputs and in-place result views give the same meshes as
    copied inputs and take(); the same job three times on one slot (whose
    context captures launch sequences only when they repeat) gives the same
    mesh each time; phase times are ordered."""
    # lattice jobs need N / 8 to be a power of two
    J = [j for j in jobs(9, 10) if j[4] is None or j[2] in (32, 64)]
    say("engine: borrowed inputs, views, %d jobs, 3 slots" % len(J))
    eng = B.Engine(0, slots=3, hbm_budget=8 << 30)
    try:
        ids = [eng.submit(P, T, br, 0.3, radius=r, lattice=lat, borrow=True) for P, T, br, r, lat in J]
        got = []
        for i in ids:
            V, N, Tt = eng.view(i)
            got.append(h(V, N, Tt))
            t = eng.phases(i)
            assert np.all(np.diff(t) >= 0) and t[4] > t[0]
            eng.release(i)
        with pytest.raises(B.GMorphoError):
            eng.view(ids[0])  # released
        for k, (P, T, br, r, lat) in enumerate(J):
            assert got[k] == h(*alone(ctx, P, T, br, r, lat)), "job %d differs from its single-context run" % k
    finally:
        eng.close()
    P, T, br, r, lat = J[0]
    ref = h(*alone(ctx, P, T, br, r, lat))
    eng = B.Engine(0, slots=1, hbm_budget=8 << 30)
    try:
        for rep in range(3):
            say("engine repeat %d" % rep)
            i = eng.submit(P, T, br, 0.3, radius=r, borrow=True)
            assert h(*eng.view(i)) == ref
            eng.release(i)
    finally:
        eng.close()
