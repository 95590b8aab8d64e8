"""Seeded synthetic volumes and points shared by the tests and the bench.

Generators follow the reference's procedural test volumes
(testing/volume/procedural_functions.h:38-145: wavelet, xyz, x/y/z ramps;
testing/volume/ProceduralStructuredVolume.h:126-143: voxel (x,y,z) sits at
object position gridOrigin + (x,y,z) * gridSpacing, x fastest).
"""
from __future__ import annotations

import numpy as np

SEED = 20261202


def object_coords(n, origin, spacing):
    """Per-axis object coordinates of voxel indices 0..n-1 (binary32 math)."""
    idx = np.arange(n, dtype=np.float32)
    return [(np.float32(origin[a]) + idx * np.float32(spacing[a])).astype(np.float64) for a in range(3)]


def procedural(dims, fn, origin=(0, 0, 0), spacing=(1, 1, 1), dtype=np.float32):
    """Array [nz, ny, nx] of fn(x, y, z) at the voxel object positions."""
    nx, ny, nz = dims
    cx = object_coords(nx, origin, spacing)[0]
    cy = object_coords(ny, origin, spacing)[1]
    cz = object_coords(nz, origin, spacing)[2]
    z, y, x = np.meshgrid(cz, cy, cx, indexing="ij")
    return fn(x, y, z).astype(dtype)


def wavelet(x, y, z):
    # procedural_functions.h:38-61 (time = 0)
    return np.sin(3 * x) + np.sin(3 * y) + np.cos(3 * z)


def xyz(x, y, z):
    return x * y * z


def zramp(x, y, z):
    return z + 0 * x


def wavelet_volume(n=128, seed=SEED):
    """sin(x) cos(y) sin(z) + U(0, 0.1), origin 0, spacing 1, float32."""
    rng = np.random.default_rng(seed)
    v = procedural((n, n, n), lambda x, y, z: np.sin(x) * np.cos(y) * np.sin(z))
    return (v + rng.uniform(0, 0.1, v.shape).astype(np.float32)).astype(np.float32)


def uniform_points(count, lo, hi, seed=SEED + 1):
    rng = np.random.default_rng(seed)
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    return (lo + rng.random((count, 3), dtype=np.float32) * (hi - lo)).astype(np.float32)


def sparse_shell_leaves(res=256, radius=0.35, width=0.06, seed=SEED + 2, fbm_amp=0.5):
    """Sphere-shell density with value noise in a res^3 index domain, stored as
    8^3 leaves (z fastest inside a leaf); only leaves touching the shell are
    kept. Returns (origins [n,3] int32, leaves [n,512] float32)."""
    rng = np.random.default_rng(seed)
    nl = res // 8
    g = np.arange(nl) * 8 + 4.0
    cz, cy, cx = np.meshgrid(g, g, g, indexing="ij")
    c = res / 2.0
    r = np.sqrt((cx - c) ** 2 + (cy - c) ** 2 + (cz - c) ** 2) / res
    keep = np.abs(r - radius) < width + 8.0 / res
    kz, ky, kx = np.nonzero(keep)
    origins = np.stack([kx * 8, ky * 8, kz * 8], 1).astype(np.int32)
    # voxel centres of every kept leaf, leaf-local order x*64 + y*8 + z
    l = np.arange(8)
    lx, ly, lz = np.meshgrid(l, l, l, indexing="ij")
    px = origins[:, 0:1] + lx.reshape(1, -1) + 0.5
    py = origins[:, 1:2] + ly.reshape(1, -1) + 0.5
    pz = origins[:, 2:3] + lz.reshape(1, -1) + 0.5
    rr = np.sqrt((px - c) ** 2 + (py - c) ** 2 + (pz - c) ** 2) / res
    shell = np.clip(1.0 - np.abs(rr - radius) / width, 0.0, 1.0)
    phase = rng.uniform(0, 2 * np.pi, 6)
    noise = (np.sin(px * 0.11 + phase[0]) * np.sin(py * 0.13 + phase[1]) * np.sin(pz * 0.07 + phase[2])
             + 0.5 * np.sin(px * 0.23 + phase[3]) * np.sin(py * 0.29 + phase[4]) * np.sin(pz * 0.31 + phase[5]))
    vals = shell * (1.0 + fbm_amp * noise)
    return origins, vals.astype(np.float32)


def dense_vdb_leaves(dims, fns, origin=(0, 0, 0), spacing=(1, 1, 1)):
    """Every 8^3 leaf of a dims^3 index domain (dims a multiple of 8), with
    attribute a = fns[a](x, y, z) at the voxel object positions (binary32
    coordinates). Returns (origins [n,3] int32, [leaves [n,512] float32 per
    attribute]), leaf-local order x*64 + y*8 + z, leaves in x-slowest order
    (testing/volume/ProceduralVdbVolume.h:213-259)."""
    nl = dims // 8
    k = np.arange(nl)
    kx, ky, kz = np.meshgrid(k, k, k, indexing="ij")
    origins = np.stack([kx.ravel(), ky.ravel(), kz.ravel()], 1).astype(np.int32) * 8
    l = np.arange(8)
    lx, ly, lz = np.meshgrid(l, l, l, indexing="ij")
    ix = origins[:, 0:1] + lx.reshape(1, -1)
    iy = origins[:, 1:2] + ly.reshape(1, -1)
    iz = origins[:, 2:3] + lz.reshape(1, -1)
    o, sp = np.asarray(origin, np.float32), np.asarray(spacing, np.float32)
    x = (o[0] + ix.astype(np.float32) * sp[0]).astype(np.float64)
    y = (o[1] + iy.astype(np.float32) * sp[1]).astype(np.float64)
    z = (o[2] + iz.astype(np.float32) * sp[2]).astype(np.float64)
    return origins, [fn(x, y, z).astype(np.float32) for fn in fns]


def temporal_volume(dims, fn, kind, dtype=np.float32, origin=(0, 0, 0), spacing=(1, 1, 1), seed=SEED + 5,
                    index64=False):
    """Motion-blur voxel data of fn(x, y, z, t) at the voxel object positions
    (testing/volume/TestingStructuredVolume.h TemporalConfig):
    ("structured", T) - T equally spaced times per voxel, time fastest;
    ("unstructured", times) - the same sorted times for every voxel;
    ("varying", None) - 1..5 sorted random times per voxel (first 0 and last 1
    when there are two or more).
    Returns (data 1-D, kwargs for the volume constructors)."""
    nx, ny, nz = dims
    cx = object_coords(nx, origin, spacing)[0]
    cy = object_coords(ny, origin, spacing)[1]
    cz = object_coords(nz, origin, spacing)[2]
    z, y, x = np.meshgrid(cz, cy, cx, indexing="ij")
    x, y, z = x.reshape(-1), y.reshape(-1), z.reshape(-1)
    nvox = x.size
    if kind[0] == "structured":
        steps = int(kind[1])
        t = np.linspace(0, 1, steps, dtype=np.float32) if steps > 1 else np.zeros(1, np.float32)
        vals = fn(x[:, None], y[:, None], z[:, None], t[None, :].astype(np.float64))
        return vals.reshape(-1).astype(dtype), {"structured_steps": steps}
    if kind[0] == "unstructured":
        t = np.asarray(kind[1], np.float32)
        counts = np.full(nvox, t.size, np.int64)
        times = np.tile(t, nvox)
    else:
        rng = np.random.default_rng(seed)
        counts = rng.integers(1, 6, nvox)
        times = []
        for c in counts:
            if c == 1:
                times.append(rng.random(1, dtype=np.float32))
            else:
                inner = np.sort(rng.choice(np.arange(1, 1000), c - 2, replace=False)).astype(np.float32) / 1000
                times.append(np.concatenate([[0.0], inner, [1.0]]).astype(np.float32))
        times = np.concatenate(times).astype(np.float32)
    idx = np.zeros(nvox + 1, np.uint64 if index64 else np.uint32)
    idx[1:] = np.cumsum(counts)
    rep = np.repeat(np.arange(nvox), counts)
    vals = fn(x[rep], y[rep], z[rep], times.astype(np.float64))
    return vals.astype(dtype), {"tu_indices": idx, "tu_times": times}


def xyz_time(x, y, z, t):
    """(1 - t) x y z (procedural_functions.h getXYZValue, time-linear)."""
    return (1.0 - t) * x * y * z


def wavelet_time(x, y, z, t):
    return (np.sin(3 * x) + np.sin(3 * y) + np.cos(3 * z)) * (1.0 + t)


def temporal_vdb_nodes(fn, seed=SEED + 6, dtype=np.float32, n_leaves=120):
    """vdb nodes with per-node motion blur (node.data form): dense 8^3 leaves
    in [0, 128)^3 and two level-2 tiles at x >= 128, each node temporally
    constant, structured (2..4 steps) or unstructured (per voxel 1..4 sorted
    times, 0 and 1 at the ends when there are two or more). Values are
    fn(x, y, z, t) at the voxel index positions. Returns a dict of the node
    arrays for Device.vdb / Grid.vdb_temporal."""
    rng = np.random.default_rng(seed)
    cells = rng.choice(16 ** 3, n_leaves, replace=False)
    leaf_org = np.stack([cells // 256, (cells // 16) % 16, cells % 16], 1) * 8
    origins = np.concatenate([leaf_org, [[128, 0, 0], [128, 128, 0]]]).astype(np.int32)
    n = len(origins)
    level = np.array([3] * n_leaves + [2, 2], np.uint32)
    fmt = np.array([1] * n_leaves + [0, 0], np.uint32)
    tformat = rng.integers(0, 3, n).astype(np.uint32)
    steps = np.where(tformat == 1, rng.integers(2, 5, n), 0).astype(np.int32)
    leaf_data, tu_indices, tu_times = [], [], []
    l = np.arange(8)
    lx, ly, lz = np.meshgrid(l, l, l, indexing="ij")
    for i in range(n):
        o = origins[i]
        if fmt[i] == 1:
            px, py, pz = (o[0] + lx.reshape(-1)).astype(np.float64), (o[1] + ly.reshape(-1)).astype(np.float64), \
                (o[2] + lz.reshape(-1)).astype(np.float64)
        else:
            px, py, pz = np.array([o[0] + 64.0]), np.array([o[1] + 64.0]), np.array([o[2] + 64.0])
        nv = px.size
        if tformat[i] == 0:
            leaf_data.append([fn(px, py, pz, 0.0).astype(dtype)])
            tu_indices.append(None)
            tu_times.append(None)
        elif tformat[i] == 1:
            t = np.linspace(0, 1, steps[i], dtype=np.float32).astype(np.float64)
            leaf_data.append([fn(px[:, None], py[:, None], pz[:, None], t[None, :]).reshape(-1).astype(dtype)])
            tu_indices.append(None)
            tu_times.append(None)
        else:
            counts = rng.integers(1, 5, nv)
            times = []
            for c in counts:
                if c == 1:
                    times.append(rng.random(1, dtype=np.float32))
                else:
                    inner = np.sort(rng.choice(np.arange(1, 1000), c - 2, replace=False)).astype(np.float32) / 1000
                    times.append(np.concatenate([[0.0], inner, [1.0]]).astype(np.float32))
            times = np.concatenate(times).astype(np.float32)
            idx = np.zeros(nv + 1, np.uint32)
            idx[1:] = np.cumsum(counts)
            rep = np.repeat(np.arange(nv), counts)
            leaf_data.append([fn(px[rep], py[rep], pz[rep], times.astype(np.float64)).astype(dtype)])
            tu_indices.append(idx)
            tu_times.append(times)
    return dict(level=level, origin=origins, fmt=fmt, leaf_data=leaf_data, temporal_format=tformat, steps=steps,
                tu_indices=tu_indices, tu_times=tu_times)


def leaves_from_dense(arr):
    """Split an [nz, ny, nx] array (every edge a multiple of 8) into 8^3 vdb
    leaves. Returns (origins [n,3] int32, leaves [n,512]) with leaf-local
    order x*64 + y*8 + z and leaves in x-slowest order."""
    nz, ny, nx = arr.shape
    a = arr.reshape(nz // 8, 8, ny // 8, 8, nx // 8, 8)        # kz lz ky ly kx lx
    leaves = a.transpose(4, 2, 0, 5, 3, 1).reshape(-1, 512)     # kx ky kz | lx ly lz
    k = np.stack(np.meshgrid(np.arange(nx // 8), np.arange(ny // 8), np.arange(nz // 8), indexing="ij"), -1)
    origins = (k.reshape(-1, 3) * 8).astype(np.int32)
    return origins, np.ascontiguousarray(leaves)


def _gradient_noise(px, py, pz, perm):
    """Improved gradient noise (Perlin, SIGGRAPH 2002: quintic fade, 12 edge
    gradients picked by a hashed permutation) at float64 points; perm is a
    permutation of 0..255 repeated twice."""
    xi, yi, zi = (np.floor(px).astype(np.int64) & 255, np.floor(py).astype(np.int64) & 255,
                  np.floor(pz).astype(np.int64) & 255)
    x, y, z = px - np.floor(px), py - np.floor(py), pz - np.floor(pz)

    def fade(t):
        return t * t * t * (t * (t * 6 - 15) + 10)

    def grad(h, gx, gy, gz):
        h = h & 15
        u = np.where(h < 8, gx, gy)
        v = np.where(h < 4, gy, np.where((h == 12) | (h == 14), gx, gz))
        return np.where(h & 1, -u, u) + np.where(h & 2, -v, v)

    u, v, w = fade(x), fade(y), fade(z)
    A, B = perm[xi] + yi, perm[xi + 1] + yi
    AA, AB, BA, BB = perm[A] + zi, perm[A + 1] + zi, perm[B] + zi, perm[B + 1] + zi

    def lerp(t, a, b):
        return a + t * (b - a)

    return lerp(w, lerp(v, lerp(u, grad(perm[AA], x, y, z), grad(perm[BA], x - 1, y, z)),
                        lerp(u, grad(perm[AB], x, y - 1, z), grad(perm[BB], x - 1, y - 1, z))),
                lerp(v, lerp(u, grad(perm[AA + 1], x, y, z - 1), grad(perm[BA + 1], x - 1, y, z - 1)),
                     lerp(u, grad(perm[AB + 1], x, y - 1, z - 1), grad(perm[BB + 1], x - 1, y - 1, z - 1))))


def torus_vdb_leaves(res=128, octaves=4, seed=SEED + 12):
    """A turbulent torus (major radius 0.75, minor 0.375 of the half-extent,
    space distorted by `octaves` of gradient-noise turbulence) with
    0.5 + 0.5 noise inside, as the non-empty 8^3 leaves of a res^3 domain --
    the shape of the reference's torus test volume
    (testing/volume/TestingVdbTorusVolume.h:104-200), from this module's own
    noise with a seeded permutation. Returns (origins, leaves [n,512])."""
    perm = np.random.default_rng(seed).permutation(256)
    perm = np.concatenate([perm, perm])
    c = (np.arange(res) + 0.5) / res
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    turb = np.zeros_like(x)
    scale = 1.0
    for _ in range(octaves):
        turb += _gradient_noise(scale * x * 12, scale * y * 12, scale * z * 12, perm) / scale
        scale *= 2.0
    k = 1.4 + 0.4 * turb
    X, Y, Z = k * (2 * x - 1), k * (2 * y - 1), k * (2 * z - 1)
    t = np.sqrt(X * X + Z * Z) - 0.75
    inside = t * t + Y * Y - 0.375 * 0.375 < 0
    vals = np.where(inside, 0.5 + 0.5 * _gradient_noise(x * 12, y * 12, z * 12, perm), 0.0).astype(np.float32)
    origins, leaves = leaves_from_dense(vals)
    keep = np.any(leaves != 0, axis=1)
    return origins[keep], np.ascontiguousarray(leaves[keep])
