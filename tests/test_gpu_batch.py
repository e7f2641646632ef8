"""Batched closing + mesh (gm_batch_close_mesh): every context's dilation,
closing and mesh equal the same context run alone with gm_close +
gm_extract_mesh, bit for bit, for a mix of grid sizes and radii (uniform
sphere radii run batched, the rest on their own streams)."""
import hashlib
import sys

import numpy as np
import pytest

import _oracle as O
import broxy_amd as B
from _util import cells_from_fine, random_blobs

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def h(*arrs):
    m = hashlib.sha256()
    for a in arrs:
        m.update(np.ascontiguousarray(a).tobytes())
    return m.hexdigest()


def make_jobs(seed, n):
    """(mesh, base_res, radius or lattice) jobs; radii 0..17, two lattice jobs."""
    rng = np.random.default_rng(seed)
    radii = [2, 3, 5, 8, 13, 17, 1, 0, 4, 6, 9, 2, 7, 11, 3, 5]
    out = []
    for i in range(n):
        P, T = B.sphere_soup(int(rng.integers(8, 33)), 3, seed=int(rng.integers(1 << 30)))
        base_res = int(rng.choice([16, 32, 48, 64, 96]))
        lat = None
        if i % 7 == 6:  # the lattice needs a power-of-two ratio to the grid
            lat = B.scale_lattice(8, seed=i, lo=2.0, hi=9.0, smooth=1)
            base_res = 32
        out.append((P, T, base_res, radii[i % len(radii)], lat))
    return out


def load(c, job):
    P, T, base_res, r, lat = job
    c.load(P, T, base_res, 0.3)
    if lat is not None:
        c.scale_lattice(lat, 8)
    else:
        c.scale_uniform(r)


def result(c):
    V, N, T = c.mesh()
    return h(c.get_cells(B.GRID_DILATION), c.get_cells(B.GRID_CLOSING), V, N, T)


def test_batch_equals_alone():
    jobs = make_jobs(31, 16)
    ctxs = [B.Context(0) for _ in jobs]
    for c, j in zip(ctxs, jobs):
        load(c, j)
    say("alone: %d contexts" % len(ctxs))
    alone = []
    for c in ctxs:
        c.close_grid(B.SE_SPHERE)
        c.extract_mesh()
        alone.append(result(c))
    bt = B.Batch(0)
    for rep in range(2):
        say("batch call %d" % rep)
        bt.close_mesh(ctxs)
        got = [result(c) for c in ctxs]
        bad = [i for i, (a, b) in enumerate(zip(alone, got)) if a != b]
        assert not bad, "batched differs from alone for jobs %s" % bad
    bt.close()
    for c in ctxs:
        c.close()


def test_batch_against_oracle_and_regrowth():
    # a smaller batch after a larger one (buffers kept), then larger grids;
    # a few jobs checked against the oracle directly
    bt = B.Batch(0)
    for seed, R in ((5, 32), (6, 80)):
        ctxs = []
        refs = []
        for i, r in enumerate((2, 6, 10)):
            I = random_blobs(R, seed=seed * 10 + i)
            c = B.Context(0)
            bmin = np.array([-1.0, -0.5, 0.25], np.float32)
            cell = np.float32(2.0 / R)
            c.grid_init(bmin, bmin + cell * R, R, np.array([R, R - 3, R - 5], np.int32), float(cell))
            c.set_cells(B.GRID_INPUT, I)
            c.scale_uniform(r)
            ctxs.append(c)
            Dref = O.dilate(I, R, r=r, data_res=[R, R - 3, R - 5])
            refs.append((Dref, O.erode(I, Dref, R, r=r, data_res=[R, R - 3, R - 5]), bmin, float(cell)))
        say("batch of %d at R=%d" % (len(ctxs), R))
        bt.close_mesh(ctxs)
        for c, (Dref, Cref, bmin, cell) in zip(ctxs, refs):
            assert np.array_equal(c.get_cells(B.GRID_DILATION), Dref)
            C = c.get_cells(B.GRID_CLOSING)
            assert np.array_equal(C, Cref)
            V, N, T = c.mesh()
            Vr, Nr, Tr = O.mc(C, R, bmin, cell)
            assert np.array_equal(T, Tr)
            assert V.shape == Vr.shape and np.allclose(V, Vr, rtol=1e-6, atol=0)
            c.close()
    bt.close_mesh([])
    bt.close()


def test_batch_mesh_overflow_emitted_again():
    """A mesh far larger than a fresh context's initial output capacity,
    closed in a batch: reading the mesh grows the output and emits it again
    with the single-job emit, which walks the emit order that the batched
    scan built. Every result equals the job closed alone."""
    R = 48
    N = 2 * R
    f = np.zeros((N, N, N), bool)
    for z in range(6, N - 9, 8):  # 11^3 separate 3^3-voxel cubes
        for y in range(6, N - 9, 8):
            for x in range(6, N - 9, 8):
                f[z:z + 3, y:y + 3, x:x + 3] = True
    I = cells_from_fine(f)

    def make():
        c = B.Context(0)
        bmin = np.zeros(3, np.float32)
        cell = np.float32(1.0 / R)
        c.grid_init(bmin, bmin + cell * R, R, np.array([R, R, R], np.int32), float(cell))
        c.set_cells(B.GRID_INPUT, I)
        c.scale_uniform(2)
        return c

    say("overflow: alone at R=%d" % R)
    alone = make()
    alone.close_grid(B.SE_SPHERE)
    alone.extract_mesh()
    V, _, _ = alone.mesh()
    assert V.shape[0] > 4 * N * N  # beyond the initial output capacity
    ref = result(alone)
    alone.close()
    ctxs = [make(), make()]
    bt = B.Batch(0)
    say("overflow: batch of %d fresh contexts" % len(ctxs))
    bt.close_mesh(ctxs)
    for c in ctxs:
        assert result(c) == ref
        c.close()
    bt.close()


def test_batch_graph_replays():
    """The same batch closed four times: the first call captures the batched
    launches, the others replay them, with eager single-context work and
    grid uploads between the calls; every replay equals the batch run without
    graphs. A changed radius (new capture) is checked the same way."""
    jobs = make_jobs(11, 12)
    ctxs = [B.Context(0) for _ in jobs]
    ref_ctxs = [B.Context(0) for _ in jobs]
    bt, bt_ref = B.Batch(0), B.Batch(0)
    try:
        bt_ref.set_graphs(False)
        for c, c2, j in zip(ctxs, ref_ctxs, jobs):
            load(c, j)
            load(c2, j)
        for rep in range(4):
            say("batch graph replay %d" % rep)
            if rep == 2:  # grids change between replays (same buffers): the replay reads them
                for c, c2, j in zip(ctxs, ref_ctxs, jobs):
                    I = c.get_cells(B.GRID_INPUT)
                    I2 = np.roll(I, 1)
                    c.set_cells(B.GRID_INPUT, I2)
                    c2.set_cells(B.GRID_INPUT, I2)
            if rep == 3:
                for c, c2 in zip(ctxs, ref_ctxs):
                    c.scale_uniform(3)
                    c2.scale_uniform(3)
            bt.close_mesh(ctxs)
            bt_ref.close_mesh(ref_ctxs)
            # eager work on a context between the calls
            ctxs[0].close_grid(B.SE_SPHERE)
            ctxs[0].extract_mesh()
            ctxs[0].mesh_counts()
            ref_ctxs[0].close_grid(B.SE_SPHERE)
            ref_ctxs[0].extract_mesh()
            ref_ctxs[0].mesh_counts()
            for c, c2 in zip(ctxs, ref_ctxs):
                assert result(c) == result(c2)
    finally:
        bt.close()
        bt_ref.close()
        for c in ctxs + ref_ctxs:
            c.close()
