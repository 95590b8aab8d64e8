"""Writes tests/golden/c_boundary_expected.bin: what tests/c_api/boundary.c
must write, computed by the CPU oracle from the same inputs (the program's
integer-valued volumes, its 32-bit LCG points and rays, binary32 arithmetic
in the same order). Run from the repository root:

    python tests/golden/make_c_boundary.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

DIM, NPTS, NGRAD, NRAYS = 32, 4096, 256, 16
F = np.float32


class Lcg:
    def __init__(self):
        self.s = 12345

    def next(self):
        self.s = (self.s * 1664525 + 1013904223) & 0xFFFFFFFF
        return self.s

    def uniform(self, lo, extent):
        return F(self.next() >> 8) * F(1.0 / 16777216.0) * F(extent) + F(lo)


def expected():
    x = np.arange(DIM)
    z, y, xx = np.meshgrid(x, x, x, indexing="ij")
    dense = (((xx * 7 + y * 13 + z * 29) % 97).astype(F) / F(8.0)).astype(F)
    g = O.Grid.dense(dense, origin=(-1.0, 0.5, 2.0), spacing=(0.5, 1.0, 1.5))
    rng = Lcg()
    pts = np.empty((NPTS, 3), F)
    for i in range(NPTS):
        pts[i] = [rng.uniform(-1.5, 17.0), rng.uniform(0.0, 33.0), rng.uniform(1.0, 49.0)]
    parts = [g.sample(pts), g.gradient(pts[:NGRAD])]

    origins, leaves = [], []
    l8 = np.arange(8)
    vx, vy, vz = np.meshgrid(l8, l8, l8, indexing="ij")  # leaf order x*64 + y*8 + z
    for lx in range(4):
        for ly in range(4):
            for lz in range(4):
                if (lx * 16 + ly * 4 + lz) % 5 == 4:
                    continue
                v = ((lx * 8 + vx) * 5 + (ly * 8 + vy) * 11 + (lz * 8 + vz) * 3) % 61
                leaves.append((v.astype(F) / F(16.0)).astype(F).reshape(-1))
                origins.append([lx * 8, ly * 8, lz * 8])
    n = len(origins)
    vdb = O.Grid.vdb(np.full(n, 3, np.uint32), np.array(origins, np.int32), np.full(n, O.FORMAT_DENSE_ZYX, np.uint32),
                     leaf_data=[[a] for a in leaves])
    for i in range(NPTS):
        pts[i] = [rng.uniform(-1.0, 34.0), rng.uniform(-1.0, 34.0), rng.uniform(-1.0, 34.0)]
    parts.append(vdb.sample(pts))

    org, dirs = np.empty((NRAYS, 3), F), np.empty((NRAYS, 3), F)
    for r in range(NRAYS):
        org[r] = [rng.uniform(-8.0, 4.0), rng.uniform(0.0, 32.0), rng.uniform(0.0, 32.0)]
        dirs[r] = [1.0, rng.uniform(-0.5, 1.0), rng.uniform(-0.5, 1.0)]
    out = b"".join(p.astype(F).tobytes() for p in parts)
    for r in range(NRAYS):
        iv = vdb.intervals(org[r], dirs[r], max_out=256)
        out += np.uint32(len(iv)).tobytes() + iv.astype(F).tobytes()
    for r in range(NRAYS):
        h = vdb.hits(org[r], dirs[r], [1.0, 2.5], max_out=64)
        out += np.uint32(len(h)).tobytes() + h.astype(F).tobytes()
    return out


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c_boundary_expected.bin")
    data = expected()
    with open(path, "wb") as f:
        f.write(data)
    print(f"wrote {path} ({len(data)} bytes)")
