#!/usr/bin/env python3
"""Writes the committed fixtures under tests/golden/ (run in the development
container, where the reference tree is readable as text).

* mc_tables.npz   - edgeTable / numTriTable / triTable as data, parsed from the
                    reference's GMorpho/tables.h (Bourke/Bloyd's public-domain
                    marching-cubes tables; tables.h:25, 61, 321).
* sphere_off.npz  - vertices and faces of Bin/Resources/Models/sphere.off, the
                    reference's only mesh file (icosphere, subdivision 3).
* oracle_config0.json - SHA-256 of the oracle's grids and mesh for the seeded
                    64-icosphere soup at N = 128, r = 4 (a regression record of
                    the oracle itself; its independent checks are the brute-force
                    and closed-form tests).
"""
import hashlib
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def tables():
    src = open(os.path.join(REF, "GMorpho/tables.h")).read()

    def arr(name):
        i = src.index(name)
        j = src.index("{", i)
        k = src.index("};", j)
        return [int(x, 0) for x in re.findall(r"-?0x[0-9a-fA-F]+|-?\d+", src[j + 1:k])]

    e = np.array(arr("edgeTable"), np.uint16)
    n = np.array(arr("numTriTable"), np.uint8)
    t = np.array(arr("triTable"), np.int8).reshape(256, 16)
    np.savez_compressed(os.path.join(HERE, "mc_tables.npz"), edgeTable=e, numTriTable=n, triTable=t)


def sphere_off():
    lines = [l for l in open(os.path.join(REF, "Bin/Resources/Models/sphere.off")).read().split("\n") if l.strip()]
    nv, nf = map(int, lines[1].split()[:2])
    V = np.array([[float(x) for x in l.split()[:3]] for l in lines[2:2 + nv]])
    F = np.array([[int(x) for x in l.split()[1:4]] for l in lines[2 + nv:2 + nv + nf]], np.int32)
    np.savez_compressed(os.path.join(HERE, "sphere_off.npz"), V=V, F=F)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def oracle_config0():
    import _oracle as O
    import broxy_amd as B
    P, T = B.sphere_soup(64, 3, seed=2017)
    R, r = 64, 4
    # margin of GMorpho::Load (GMorpho.cu:7737) for se_size 0.3
    margin = np.float32(0.3) + np.float32(2.0) / np.float32(2 * R)
    bmin, bmax, cell, dr = O.grid_attributes(P, R, float(margin))
    I = O.voxelize(P, T, bmin, float((bmax - bmin).max()), R)
    D = O.dilate(I, R, r=r, data_res=dr)
    C = O.erode(I, D, R, r=r, data_res=dr)
    V, Nn, Tt = O.mc(C, R, bmin, cell)
    out = {"workload": "64-icosphere soup subdiv 3, seed 2017, R=64 (N=128), r=4, se_size 0.3",
           "bmin": [float(x) for x in bmin], "cell": float(cell), "data_res": [int(x) for x in dr],
           "I": sha(I), "D": sha(D), "C": sha(C), "nV": int(len(V)), "nT": int(len(Tt)), "T": sha(Tt),
           "V": sha(V)}
    json.dump(out, open(os.path.join(HERE, "oracle_config0.json"), "w"), indent=1)


if __name__ == "__main__":
    tables()
    sphere_off()
    sys.path.insert(0, os.path.join(os.path.dirname(HERE)))
    oracle_config0()
    print("fixtures written")
