"""Seeded job mixes for the batched and served closing workloads.

batched_jobs: 256 independent closing jobs, each a sphere-union mesh of
8-128 icospheres (subdivision 3), grid edge N log-uniform in [64, 384]
rounded to a multiple of 32 (base_res = N/2), uniform radius integer in
[2, 16] fine voxels, in seeded order.

served_requests: Poisson arrivals at rate lam; grid edge from
{128, 256, 512, 1024} with probabilities {0.4, 0.3, 0.2, 0.1}, radius uniform
in [2, 24], 20% of requests with a radius field given as a smoothed FP32 64^3
lattice with values in [2, 24]. Meshes come from the sphere-union generator
(N <= 256) and the SH-displaced icosphere (N >= 512), drawn from a small
seeded pool so that generating them stays off the measured path.
"""
import numpy as np

from . import scale_lattice, sh_blob, sphere_soup


def batched_jobs(seed=2017, n=256):
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(n):
        ns = int(rng.integers(8, 129))
        N = float(np.exp(rng.uniform(np.log(64.0), np.log(384.0))))
        N = int(min(384, max(64, 32 * round(N / 32.0))))
        r = int(rng.integers(2, 17))
        P, T = sphere_soup(ns, 3, seed=int(rng.integers(1 << 62)))
        jobs.append({"P": P, "T": T, "base_res": N // 2, "radius": r, "lattice": None, "N": N, "spheres": ns})
    return jobs


def served_requests(lam, n, seed=4242, pool_seed=77, big_subdiv=7):
    """n requests with Poisson arrival times (seconds from 0) at rate lam."""
    rng = np.random.default_rng(seed + int(lam))
    pool = _mesh_pool(pool_seed, big_subdiv)
    t = np.cumsum(rng.exponential(1.0 / lam, n))
    reqs = []
    for i in range(n):
        N = int(rng.choice([128, 256, 512, 1024], p=[0.4, 0.3, 0.2, 0.1]))
        r = int(rng.integers(2, 25))
        lat = None
        if rng.random() < 0.2:
            lat = pool["lattices"][int(rng.integers(len(pool["lattices"])))]
        meshes = pool["soups"] if N <= 256 else pool["blobs"]
        P, T = meshes[int(rng.integers(len(meshes)))]
        reqs.append({"t": float(t[i]), "P": P, "T": T, "base_res": N // 2, "radius": r, "lattice": lat, "N": N})
    return reqs


_POOL = {}


def _mesh_pool(seed, big_subdiv):
    key = (seed, big_subdiv)
    if key not in _POOL:
        rng = np.random.default_rng(seed)
        soups = [sphere_soup(int(rng.integers(8, 129)), 3, seed=int(rng.integers(1 << 62))) for _ in range(8)]
        blobs = [sh_blob(big_subdiv, seed=int(rng.integers(1 << 30)), lmax=8, amp=0.1) for _ in range(2)]
        lattices = [scale_lattice(64, seed=int(rng.integers(1 << 30)), lo=2.0, hi=24.0, smooth=3) for _ in range(3)]
        _POOL[key] = {"soups": soups, "blobs": blobs, "lattices": lattices}
    return _POOL[key]
