#!/usr/bin/env python3
"""Closing + mesh throughput of the GMorpho path on one MI355X.

Workload (the headline configuration): SH-displaced icosphere (subdivision 7,
327,680 triangles, lmax 8, amplitude 0.1, seeded), voxelized at N = 512
(base_res 256, se_size 0.3), closed with a uniform ball of radius 8 fine
voxels, then marching cubes. A step is one closing + mesh extraction
(dilate -> erode -> MC) of the grid already resident in HBM; voxelization is
timed and reported separately. value = N^3 / step time (Gvoxels/s).

Prints one JSON line (rank 0). Single GPU only: --gpus > 1 exits.
"""
import argparse
import json
import os
import queue
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import broxy_amd as B  # noqa: E402

HBM_PEAK_GBS = 8000.0
STAGES = list(B.Context.STAGES)
# kernel (name in a rocprof trace) timed by each stage
KERNEL_OF_STAGE = {"brick_summary": "k_brick_summary", "dilate_coarse": "k_pass_ct<0,rho>",
                   "tight_contour": "k_tight", "dilate_fine": "k_pass_ct<1,rho>", "contour_K": "k_contour_K",
                   "erode_coarse": "k_pass_ct<2,rho>", "base_contour": "k_base_contour",
                   "erode_fine": "k_pass_ct<3,rho>", "mc_count": "k_mc_blocks_c + k_mc_count",
                   "mc_scan": "k_scan_down", "mc_emit": "k_mc_emit",
                   # no kernel of its own: the emit's last workgroup, only when a
                   # cube lists more triangles than it counts
                   "mc_fixup_normalize": "k_mc_emit (last workgroup)",
                   "dilate_coarse_stream": "k_pass_ct<0,rho>", "dilate_fine_stream": "k_pass_ct<1,rho>",
                   "erode_coarse_stream": "k_pass_ct<2,rho>", "erode_fine_stream": "k_pass_ct<3,rho>"}
# Stages timed over exactly one kernel launch: the dominant kernel is chosen
# among these (a pass stage also holds its tile selection).
SINGLE_KERNEL_STAGES = ("brick_summary", "tight_contour", "contour_K", "base_contour", "dilate_coarse_stream",
                        "dilate_fine_stream", "erode_coarse_stream", "erode_fine_stream", "mc_emit")


# per-voxel radii: one compiled-radius launch per radius value
VAR_KERNEL = {"dilate_coarse_stream": "k_pass_ct<0,v+1> per radius value v",
              "dilate_fine_stream": "k_pass_ct<1,v+1> per radius value v",
              "erode_coarse_stream": "k_pass_ct<2,v-1> per radius value v",
              "erode_fine_stream": "k_pass_ct<3,v-1> per radius value v"}


def edt_bytes(N, nq):
    """One separable transform of a bit grid to nq^3 query points: the bit
    grid read, then each axis stage's 4-byte values written and read back."""
    return N ** 3 / 8.0 + 8.0 * N * N * nq + 8.0 * N * nq * nq + 4.0 * nq ** 3


def algorithmic_bytes(stage, N, nv, nt, edt=False):
    """Compulsory HBM bytes of one launch of a stage's kernel: bit grids N^3/8
    bytes, cell flags N^3/64 bytes, mesh 12 bytes per vertex position, normal
    and triangle. edt: radii above 31, whose four passes are distance
    transforms (plus the threshold pass that writes the flags or bits)."""
    g = N ** 3 / 8.0
    f = N ** 3 / 64.0
    stage = stage.replace("_stream", "")
    if edt and stage in ("dilate_coarse", "dilate_fine", "erode_coarse", "erode_fine"):
        return {"dilate_coarse": edt_bytes(N, N) + 4.0 * (N // 2) ** 3 + 2 * f,
                "dilate_fine": 4.0 * N ** 3 + 2 * g + f,
                "erode_coarse": edt_bytes(N, N // 2) + 4.0 * (N // 2) ** 3 + 3 * f,
                "erode_fine": edt_bytes(N, N) + 4.0 * N ** 3 + 2 * g + f}[stage]
    return {
        "brick_summary": g + f,            # read I; write I-full cell flags
        "dilate_coarse": g + 2 * f,        # read I, I-full flags; write DF
        "tight_contour": 2 * f + 2 * g,    # read DF, I; write TG, D
        "dilate_fine": 2 * g + f,          # read I, TG; set bits of D (upper bound: every word)
        "contour_K": 2 * g + f,            # read D; write K, D-any cell flags
        "erode_coarse": g + 3 * f,         # read K, D-any and I-full flags; write CL
        "base_contour": 2 * g + 2 * f,     # read D, CL; write UG, C
        "erode_fine": 2 * g + f,           # read K, UG; clear bits of C (upper bound: every word)
        "mc_count": 2 * g,                 # read C (block list, then the count)
        "mc_scan": 16.0 * N ** 3 / 4096,   # block counts and bases (2 u32 per 8^3 cells, read + write)
        "mc_emit": g + 24 * nv + 12 * nt,  # read C; write positions, normal sums, triangles
        "mc_fixup_normalize": 24 * nv,     # normals read + write
    }[stage]


def rho_of_stage(stage, r):
    """Chord-table radius of a streaming pass (dilation r+1, erosion r-1)."""
    return r + 1 if stage.startswith("dilate") else r - 1


# rocprofv3 counter summaries per (config, structuring element)
COUNTER_FILES = {("large", "sphere"): "counters_sh512_r8.json"}


def dominant_kernel(stage, r, se):
    """rocprof name of the kernel a stage times (a uniform cube radius runs
    the compiled box kernels, passes 4 and 5; radii above 31 the
    distance-transform passes)."""
    if r and r > 31 and stage.endswith("_stream"):
        return "k_edt_x + k_edt_axis x2 + threshold (distance-transform passes)"
    if se == "cube" and stage in ("dilate_coarse_stream", "dilate_fine_stream"):
        if r:
            return "k_pass_ct<%d,%d>" % (4 if "coarse" in stage else 5, r)
        return "k_dil_coarse" if "coarse" in stage else "k_dil_fine"
    if r:
        return KERNEL_OF_STAGE[stage].replace("rho", str(rho_of_stage(stage, r)))
    return VAR_KERNEL.get(stage, KERNEL_OF_STAGE[stage])


def build_workload(args):
    """(P, T, base_res, uniform radius or None, radius lattice or None, description)."""
    if args.config == "large":
        P, T = B.sh_blob(7, seed=args.seed, lmax=8, amp=0.1)
        r = args.radius if args.radius else 8
        return P, T, 256, r, None, "sh-icosphere subdiv7 (327,680 tris), N=512, ball r=%d" % r
    if args.config == "plain":
        P, T = B.sphere_soup(64, 3, seed=args.seed)
        return P, T, 64, 4, None, "64-icosphere soup subdiv3, N=128, ball r=4"
    if args.config == "varying":
        P, T = B.sh_blob(8, seed=args.seed, lmax=8, amp=0.1)
        lat = B.scale_lattice(64, seed=args.seed, lo=2.0, hi=24.0, smooth=3)
        return P, T, 512, None, lat, ("sh-icosphere subdiv8 (1,310,720 tris), N=1024, radii from a 64^3 FP32 "
                                      "lattice U[2,24] box-smoothed 3x, trilinear")
    raise SystemExit("unknown config")


def cpu_baseline(I, R, r, dr, bmin, cell, N, lat=None, se=0):
    """The oracle (CPU restatement, OpenMP) on the same closing + mesh. Its
    grids and mesh are kept: every bench line compares the GPU's with them."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    sc = None
    if lat is not None:
        L = int(round(lat.size ** (1 / 3.0)))
        sc = O.scale_from_lattice(lat.reshape(L, L, L), N)
    t0 = time.perf_counter()
    D = O.dilate(I, R, r=r or 0, data_res=dr, scale=sc, se=se)
    C = O.erode(I, D, R, r=r or 0, data_res=dr, scale=sc)
    V, Nn, T = O.mc(C, R, bmin, cell)
    dt = time.perf_counter() - t0
    return {"value": round(N ** 3 / dt / 1e9, 4), "unit": "Gvoxels/s", "cores": int(O.num_threads()),
            "kind": "port", "sample": "full closing + mesh of the same grid, once (%.1f s)" % dt}, D, C, (V, Nn, T)


def parity(ctx, P, T, R, bmin, bmax, D, C, mesh):
    """GPU grids and mesh of the last step against the oracle's: voxelized
    grid, dilation and closing bit for bit, triangles exactly, vertices within
    1e-6 relative, unit normals within 2e-5."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    Vr, Nr, Tr = mesh
    V, Nn, Tt = ctx.mesh()
    Iref = O.voxelize(P, T, bmin, float((bmax - bmin).max()), R)
    same = V.shape == Vr.shape and Tt.shape == Tr.shape
    out = {"voxelized_bit_exact": bool(np.array_equal(ctx.get_cells(B.GRID_INPUT), Iref)),
           "dilation_bit_exact": bool(np.array_equal(ctx.get_cells(B.GRID_DILATION), D)),
           "closing_bit_exact": bool(np.array_equal(ctx.get_cells(B.GRID_CLOSING), C)),
           "triangles_exact": bool(same and np.array_equal(Tt, Tr)),
           "vertices_rtol1e-6": bool(same and np.allclose(V, Vr, rtol=1e-6, atol=0)),
           "normals_atol2e-5": bool(same and np.allclose(Nn, Nr, atol=2e-5))}
    out["all"] = all(out.values())
    return out


def progress(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def _mesh_hash(V, N, T):
    import hashlib
    m = hashlib.sha256()
    for a in (V, N, T):
        m.update(np.ascontiguousarray(a).tobytes())
    return m.hexdigest()


def run_alone(ctx, j):
    ctx.load(j["P"], j["T"], j["base_res"], 0.3)
    if j["lattice"] is not None:
        ctx.scale_lattice(j["lattice"], int(round(j["lattice"].size ** (1 / 3.0))))
    else:
        ctx.scale_uniform(j["radius"])
    ctx.update(B.SE_SPHERE)
    return ctx.mesh()


def run_engine(eng, jobs):
    """Every job's mesh hash (inputs borrowed, meshes read in place)."""
    ids = [eng.submit(j["P"], j["T"], j["base_res"], 0.3, radius=j["radius"], lattice=j["lattice"], borrow=True)
           for j in jobs]
    out = []
    for i in ids:
        out.append(_mesh_hash(*eng.view(i)))
        eng.release(i)
    return out


def run_engine_timed(eng, jobs):
    """The same jobs, each mesh released as soon as it is on the host."""
    ids = [eng.submit(j["P"], j["T"], j["base_res"], 0.3, radius=j["radius"], lattice=j["lattice"], borrow=True)
           for j in jobs]
    for i in ids:
        eng.view(i)
        eng.release(i)


def batched(args):
    """256 closing + mesh jobs of uneven size and radius (config 3).

    resident: every job is voxelized into its own context first (untimed),
    then closing + mesh extraction of all of them is timed, one by one (one
    job, wait, next) against gm_batch_close_mesh (all jobs, one launch per
    step). end_to_end: voxelize + close + mesh + copy-back per job from host
    meshes, one by one against the job engine (slot count tuned on the first
    64 jobs). Every job's mesh must be identical between the modes."""
    from broxy_amd import workloads
    jobs = workloads.batched_jobs(seed=args.seed, n=args.batch_jobs)
    vox = sum((2 * j["base_res"]) ** 3 for j in jobs)
    # ---- resident
    ctxs = []
    for j in jobs:
        c = B.Context(0)
        c.load(j["P"], j["T"], j["base_res"], 0.3)
        c.scale_uniform(j["radius"])
        ctxs.append(c)
    bt = B.Batch(0)

    def one_by_one():
        for c in ctxs:
            c.update(B.SE_SPHERE)
            c.mesh_counts()

    def all_at_once():
        bt.close_mesh(ctxs)
        for c in ctxs:
            c.mesh_counts()

    one_by_one()
    all_at_once()  # warm both paths (allocations, code objects)
    t_one, t_all = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        one_by_one()
        t_one.append(time.perf_counter() - t0)
    ref = [_mesh_hash(*c.mesh()) for c in ctxs]
    for _ in range(3):
        t0 = time.perf_counter()
        all_at_once()
        t_all.append(time.perf_counter() - t0)
    same_res = all(_mesh_hash(*c.mesh()) == hsh for c, hsh in zip(ctxs, ref))
    progress("batched resident: one by one %.3f s, batched %.3f s" % (float(np.median(t_one)), float(np.median(t_all))))
    bt.close()
    for c in ctxs:
        c.close()
    del ctxs
    t_one, t_all = float(np.median(t_one)), float(np.median(t_all))
    # ---- end to end from host meshes (the faster one-by-one is the baseline).
    # Each mode first runs the jobs once untimed with every mesh hashed (also
    # warming buffers), then times them without hashing.
    big = max(jobs, key=lambda j: j["base_res"])
    def median_time(fn, reps=3):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    ctx = B.Context(0)
    run_alone(ctx, big)
    ref_e = [_mesh_hash(*run_alone(ctx, j)) for j in jobs]
    t_ctx = median_time(lambda: [run_alone(ctx, j) for j in jobs])
    ctx.close()
    eng = B.Engine(0, slots=1, hbm_budget=args.hbm_budget << 30)
    same_e = run_engine(eng, jobs) == ref_e
    t_eng1 = median_time(lambda: run_engine_timed(eng, jobs))
    eng.close()
    # slot count tuned on the first half of the jobs
    tune = {}
    for k in (16, 32, 64):
        eng = B.Engine(0, slots=k, hbm_budget=args.hbm_budget << 30)
        run_engine(eng, jobs[:128])
        tune[k] = median_time(lambda: run_engine_timed(eng, jobs[:128]))
        eng.close()
    kbest = min(tune, key=tune.get)
    progress("batched end to end: slot tuning %s" % {k: round(v, 3) for k, v in tune.items()})
    eng = B.Engine(0, slots=kbest, hbm_budget=args.hbm_budget << 30)
    same_e = same_e and run_engine(eng, jobs) == ref_e
    t_eng = median_time(lambda: run_engine_timed(eng, jobs))
    eng.close()
    t_seq = min(t_ctx, t_eng1)
    n = len(jobs)
    return {"jobs": n, "jobs_per_s": round(n / t_all, 2), "one_by_one_jobs_per_s": round(n / t_one, 2),
            "speedup": round(t_one / t_all, 3), "Gvoxels_per_s": round(vox / t_all / 1e9, 3),
            "bit_exact_vs_one_by_one": bool(same_res),
            "mode": "resident: grids voxelized into one context per job beforehand; timed = closing + mesh of "
                    "every job with counts on the host (median of 3)",
            "end_to_end": {"jobs_per_s": round(n / t_eng, 2), "one_by_one_jobs_per_s": round(n / t_seq, 2),
                           "speedup": round(t_seq / t_eng, 3),
                           "one_by_one_context_jobs_per_s": round(n / t_ctx, 2),
                           "one_by_one_engine_1slot_jobs_per_s": round(n / t_eng1, 2), "slots": kbest,
                           "slot_tuning_s_per_128_jobs": {str(k): round(v, 4) for k, v in tune.items()},
                           "bit_exact_vs_one_by_one": bool(same_e),
                           "mode": "from host meshes: voxelize + close + mesh + copy-back per job; engine: inputs "
                                   "borrowed, each mesh copied by the device into pinned host memory and released "
                                   "once there; one by one: one context, meshes copied into numpy; each mode is "
                                   "run once hashed (bit_exact), then timed 3 times (median)"},
            "workload": "sphere-union jobs (8-128 icospheres), N log-uniform [64,384] step 32, r U{2..16}"}


def served(args):
    """Poisson arrivals at several rates: continuous admission into the job
    engine vs one request at a time, same request sequence. Latency is
    arrival (submit) to mesh on host."""
    from broxy_amd import workloads
    out = {}
    for lam in args.rates:
        # rate 0: every request arrives at t = 0 (serving capacity)
        reqs = workloads.served_requests(lam if lam > 0 else 1e9, args.requests)
        row = {}
        hashes = {}
        for mode, slots in (("continuous", args.slots), ("one_at_a_time", 1)):
            eng = B.Engine(0, slots=slots, hbm_budget=args.hbm_budget << 30)
            # warm every slot at the largest size of the mix
            big = max(reqs, key=lambda q: q["base_res"])
            run_engine(eng, [big] * slots)
            ids = queue.Queue()
            lat, done, hs, phase = [], [], [], []

            def consume():
                # the client: reads each mesh where the engine left it, then
                # releases the job
                for i in iter(ids.get, None):
                    V, N, T = eng.view(i)
                    hs.append(_mesh_hash(V, N, T))
                    t = eng.phases(i)
                    eng.release(i)
                    lat.append(t[4] - t[0])
                    done.append(t[4])
                    phase.append(np.diff(t))

            th = threading.Thread(target=consume)
            th.start()
            start = eng.now()
            for q in reqs:
                while eng.now() - start < q["t"]:
                    time.sleep(min(0.002, max(0.0, q["t"] - (eng.now() - start))))
                ids.put(eng.submit(q["P"], q["T"], q["base_res"], 0.3, radius=q["radius"], lattice=q["lattice"],
                                   borrow=True))
            ids.put(None)
            th.join()
            eng.close()
            lat = np.array(lat) * 1e3
            span = max(done) - start
            progress("served rate %s %s: %d requests done" % (lam, mode, len(reqs)))
            ph = np.array(phase) * 1e3
            row[mode] = {"p50_ms": round(float(np.percentile(lat, 50)), 2),
                         "p99_ms": round(float(np.percentile(lat, 99)), 2),
                         "jobs_per_s": round(len(reqs) / span, 2), "slots": slots,
                         "p50_phase_ms": {k: round(float(np.percentile(ph[:, n], 50)), 3)
                                          for n, k in enumerate(("queued", "load", "close_mesh", "copy_back"))},
                         "p99_phase_ms": {k: round(float(np.percentile(ph[:, n], 99)), 3)
                                          for n, k in enumerate(("queued", "load", "close_mesh", "copy_back"))}}
            hashes[mode] = hs
        row["bit_exact_mix_vs_alone"] = hashes["continuous"] == hashes["one_at_a_time"]
        row["throughput_ratio"] = round(row["continuous"]["jobs_per_s"] / row["one_at_a_time"]["jobs_per_s"], 3)
        out["lambda_%d" % lam if lam > 0 else "burst"] = row
    return {"requests_per_rate": args.requests, "rates": out,
            "workload": "N in {128,256,512,1024} p {.4,.3,.2,.1}, r U[2,24], 20% with a 64^3 radius lattice"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="large")
    ap.add_argument("--se", choices=["sphere", "cube"], default="sphere",
                    help="dilation structuring element (cube: the asymmetric closing, cube dilation + ball erosion)")
    ap.add_argument("--seed", type=int, default=2017)
    ap.add_argument("--radius", type=int, default=0,
                    help="uniform radius for --config large (default 8; above 31 the distance-transform passes run)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the oracle parity check")
    ap.add_argument("--no-tune", action="store_true", help="default launch shape (no autotuning before timing)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-jobs measurement")
    ap.add_argument("--no-served", action="store_true", help="skip the served-requests measurement")
    ap.add_argument("--batch-jobs", type=int, default=256)
    ap.add_argument("--requests", type=int, default=300, help="served requests per arrival rate")
    ap.add_argument("--rates", type=int, nargs="+", default=[20, 50, 100, 200, 0],
                    help="Poisson arrival rates (req/s); 0 = all requests at once")
    ap.add_argument("--slots", type=int, default=8, help="engine slots for the served measurement")
    ap.add_argument("--hbm-budget", type=int, default=200, help="engine HBM budget in GiB")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 3 copies of MIB MiB by the library's copy kernel (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        B.copy_probe(args.copy_probe << 20, reps=3)
        print(json.dumps({"copy_probe_bytes": args.copy_probe << 20, "kernel": "k_copy_probe"}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the closing path is single-device (--gpus %d refused)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    P, T, R, r, lat, desc = build_workload(args)
    se = B.SE_CUBE if args.se == "cube" else B.SE_SPHERE
    if args.se == "cube":
        desc = desc.replace("ball r=", "cube dilation + ball erosion r=")
    N = 2 * R
    ctx = B.Context(0)
    t0 = time.perf_counter()
    ctx.load(P, T, R, 0.3)
    vox_ms = (time.perf_counter() - t0) * 1e3
    # the same load into the context again: its buffers exist (a serving
    # slot's steady state); mesh upload from host + checks + voxelization
    warm = []
    for _ in range(3):
        t0 = time.perf_counter()
        ctx.load(P, T, R, 0.3)
        warm.append((time.perf_counter() - t0) * 1e3)
    vox_warm_ms = float(np.median(warm))
    bmin, bmax, res, dr, cell = ctx.info()
    if lat is not None:
        ctx.scale_lattice(lat, int(round(lat.size ** (1 / 3.0))))
    else:
        ctx.scale_uniform(r)

    shape = None
    if not args.no_tune and (r is None or r <= 31):
        # launch shape of the streaming passes for this grid and radius,
        # chosen by timing every candidate (outside the timed region); radii
        # above 31 run the distance-transform passes, which have none
        shape = ctx.autotune(se, 5)
    # a step: GMorpho::Update up to the mesh (closing, then marching cubes),
    # enqueued as one captured sequence
    for _ in range(args.warmup):
        ctx.update(se)
        nv, nt = ctx.mesh_counts()

    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.update(se)
    ctx.sync()
    t1 = time.perf_counter()
    nv, nt = ctx.mesh_counts()
    ms = (t1 - t0) * 1e3 / args.steps

    # per-stage device times from HIP events on the context stream
    ctx.set_timing(True)
    acc = np.zeros(len(STAGES))
    reps = max(3, min(args.steps, 10))
    for _ in range(reps):
        ctx.close_grid(se)
        ctx.extract_mesh()
        acc += ctx.stage_times()
    ctx.set_timing(False)
    stage_ms = acc / reps
    planes_ms = field_step_ms = None
    if lat is not None:
        # a new radius field: its first closing also builds the radius bit
        # planes (inside the brick_summary stage); later closings reuse them
        ctx.scale_lattice(lat, int(round(lat.size ** (1 / 3.0))))
        ctx.set_timing(True)
        ctx.close_grid(se)
        ctx.extract_mesh()
        planes_ms = float(ctx.stage_times()[0] - stage_ms[0])
        ctx.set_timing(False)
        # the same steps with a new radius field before each one (host upload
        # of the lattice, radius planes rebuilt on the device, then the step)
        L = int(round(lat.size ** (1 / 3.0)))
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.scale_lattice(lat, L)
            ctx.update(se)
        ctx.sync()
        field_step_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    dom = max((i for i, s in enumerate(STAGES) if s in SINGLE_KERNEL_STAGES), key=lambda i: stage_ms[i])
    dname = STAGES[dom]
    dbytes = algorithmic_bytes(dname, N, nv, nt, edt=r is not None and r > 31)
    achieved = dbytes / (stage_ms[dom] * 1e-3) / 1e9
    whole_bytes = 5 * N ** 3 / 8.0 + 12 * nv + 12 * nt
    whole_gbs = whole_bytes / (ms * 1e-3) / 1e9

    # HBM bytes per launch of that kernel, and its VALU / LDS use, from the
    # rocprofv3 counter passes committed beside the profiles: bytes from the
    # L2-to-memory request counts (128 B per read request, 64 B per write
    # request, both checked on a copy of known size), else FETCH_SIZE /
    # WRITE_SIZE scaled by that copy
    traffic, counters = None, None
    kname = dominant_kernel(dname, r, args.se).replace(" ", "")
    cf = os.path.join(ROOT, "profiles", COUNTER_FILES.get((args.config, args.se), "-"))
    if os.path.exists(cf):
        with open(cf) as fh:
            for k in json.load(fh)["kernels"]:
                if k["kernel"].replace(" ", "") == kname:
                    traffic = k["read_bytes_from_requests"] + k["write_bytes_from_requests"]
                    counters = {"valu_busy_pct": k["VALUBusy_pct"], "valu_utilization_pct": k["VALUUtilization_pct"],
                                "lds_util_pct": k["LdsUtil_pct"], "lds_bank_conflict_pct": k["LDSBankConflict_pct"],
                                "occupancy_pct": k["Occupancy_pct"], "source": "profiles/" + os.path.basename(cf)}
    if traffic is None:
        tf = os.path.join(ROOT, "profiles", "traffic_%s.json" % args.config)
        if os.path.exists(tf):
            with open(tf) as fh:
                tr = json.load(fh)
            ent = tr.get("stages", {}).get(dname.replace("_stream", ""))
            if ent:
                traffic = ent.get("bytes")

    out = {
        "metric": ("Gvoxels/s closing 512^3 r=8 (+%HBM roof); batched jobs/s vs 1-by-1; p99 lat"
                   if args.config == "large" else "Gvoxels/s closing"),
        "value": round(N ** 3 / (ms * 1e-3) / 1e9, 3),
        "unit": "Gvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded mesh generator)",
        "config": {"workload": desc, "N": N, "radius": r if r is not None else "lattice [2,24]", "se": args.se,
                   "mesh_vertices": nv,
                   "mesh_triangles": nt, "parallelism": "single GPU",
                   "launch_shape": ({"tile_planes": shape[0], "waves": shape[1]} if shape else "default")},
        "roofline": {"bound": "hbm",
                     "kernel": dominant_kernel(dname, r, args.se),
                     "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "algorithmic_bytes": int(dbytes),
                     "avg_launch_ms": round(float(stage_ms[dom]), 5), "counters": counters},
        "pipeline_roofline": {"bytes": int(whole_bytes), "achieved_GBs": round(whole_gbs, 1),
                              "frac": round(whole_gbs / HBM_PEAK_GBS, 4)},
        "stage_ms": {s: round(float(v), 5) for s, v in zip(STAGES, stage_ms)},
        "voxelize_ms_incl_h2d": round(vox_ms, 3),
        "voxelize_ms_incl_h2d_warm": round(vox_warm_ms, 3),
    }
    if planes_ms is not None:
        out["radius_planes_ms_once_per_field"] = round(planes_ms, 4)
        out["ms_per_step_new_field_each_step"] = round(field_step_ms, 4)
    if not args.no_cpu:
        I = ctx.get_cells(B.GRID_INPUT)
        cb, Dref, Cref, mesh = cpu_baseline(I, R, r, dr, bmin, cell, N, lat, se)
        out["cpu_baseline"] = cb
        out["parity"] = parity(ctx, P, T, R, bmin, bmax, Dref, Cref, mesh)
    ctx.close()
    progress("headline done: %.4f ms per step" % ms)
    if not args.no_batched:
        out["batched"] = batched(args)
    if not args.no_served:
        out["served"] = served(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
