"""numpy front end of the CPU parity oracle (oracle/vkl_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py. The shipped library never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

VKL_UCHAR, VKL_SHORT, VKL_USHORT, VKL_HALF, VKL_FLOAT, VKL_DOUBLE = 2500, 3000, 3500, 5800, 6000, 7000
NEAREST, LINEAR, CUBIC = 0, 100, 200
FORMAT_TILE, FORMAT_DENSE_ZYX = 0, 1

DTYPE_CODE = {
    np.dtype(np.uint8): VKL_UCHAR,
    np.dtype(np.int16): VKL_SHORT,
    np.dtype(np.uint16): VKL_USHORT,
    np.dtype(np.float16): VKL_HALF,
    np.dtype(np.float32): VKL_FLOAT,
    np.dtype(np.float64): VKL_DOUBLE,
}

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++, -ffp-contract=off)."""
    cmd = ["make", "-C", _HERE] + (["-B"] if force else [])
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def use_native():
    """Load the -O3 -march=native build of the same source (built here, for
    the host that runs it) instead of the default one: the CPU baseline's
    timed copy. Call before the first lib() use."""
    global _LIB_PATH, _lib
    native = os.path.join(_HERE, "_build", "liboracle_native.so")
    if _lib is not None:
        if _LIB_PATH == native:
            return
        raise RuntimeError("the oracle library is already loaded")
    subprocess.run(["make", "-B", "-C", _HERE, "native"], check=True, stdout=subprocess.DEVNULL)
    _LIB_PATH = native


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, u64, u32, i32, f32 = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, C.c_float
        L.oracle_dense_new.restype = vp
        L.oracle_dense_new.argtypes = [vp, vp, vp, i32, u32, vp, vp, vp]
        L.oracle_vdb_new.restype = vp
        L.oracle_vdb_new.argtypes = [u64, vp, vp, vp, u32, vp, i32, vp, vp, vp, vp, vp, vp]
        L.oracle_free.argtypes = [vp]
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_set_sampler.argtypes = [vp, i32, i32, i32]
        L.oracle_set_cubic_double.argtypes = [vp, i32]
        L.oracle_sample.argtypes = [vp, u64, vp, vp, u32]
        L.oracle_sample_mt.argtypes = [vp, u64, vp, vp, u32, i32]
        L.oracle_sample_m.argtypes = [vp, u64, vp, vp, u32, vp]
        L.oracle_gradient.argtypes = [vp, u64, vp, vp, u32]
        L.oracle_bounding_box.argtypes = [vp, vp]
        L.oracle_value_range.argtypes = [vp, u32, vp]
        L.oracle_num_attributes.restype = u32
        L.oracle_num_attributes.argtypes = [vp]
        L.oracle_grid_info.argtypes = [vp, vp, vp, vp, vp]
        L.oracle_level_num_nodes.restype = u64
        L.oracle_level_num_nodes.argtypes = [vp, i32]
        L.oracle_level_copy.argtypes = [vp, i32, vp, vp, vp]
        L.oracle_map_depth.restype = i32
        L.oracle_map_depth.argtypes = [f32]
        L.oracle_iterate_intervals.restype = u32
        L.oracle_iterate_intervals.argtypes = [vp, u32, i32, i32, vp, vp, vp, vp, vp, u32]
        L.oracle_iterate_hits.restype = u32
        L.oracle_iterate_hits.argtypes = [vp, u32, i32, i32, vp, vp, vp, vp, f32, vp, u32]
        L.oracle_iterate_many.restype = None
        L.oracle_iterate_many.argtypes = [vp, u32, i32, i32, i32, vp, C.c_uint64, vp, vp, u32, u32, vp, vp]
        L.oracle_tricubic_matrix.argtypes = [vp]
        L.oracle_set_leaf_access.argtypes = [vp, vp]
        L.oracle_dense_new_temporal.restype = vp
        L.oracle_dense_new_temporal.argtypes = [vp, vp, vp, i32, u32, vp, vp, vp, u32, u32, vp, vp, u64]
        L.oracle_sample_times.argtypes = [vp, u64, vp, vp, vp, u32]
        L.oracle_vdb_new_temporal.restype = vp
        L.oracle_vdb_new_temporal.argtypes = [u64, vp, vp, vp, u32, vp, i32, vp, vp, vp, vp, vp, vp,
                                              vp, vp, vp, vp, vp]
        L.oracle_gradient_times.argtypes = [vp, u64, vp, vp, vp, u32]
        L.oracle_inner_nodes.restype = u64
        L.oracle_inner_nodes.argtypes = [vp, i32, vp, u64]
        L.oracle_woodcock.argtypes = [vp, u64, vp, vp, vp, vp, f32, vp, u32, f32, f32, f32, f32, u32, u32, vp, vp, vp]
        L.oracle_ray_march_intervals.argtypes = [vp, u32, i32, i32, vp, u64, vp, vp, vp, f32, f32, vp, u32, f32, f32,
                                                 u32, vp, vp]
        L.oracle_hit_render.argtypes = [vp, u32, i32, vp, u64, vp, vp, vp, f32, vp, u32, f32, f32, u32, vp, vp]
        L.oracle_tea.argtypes = [vp, vp, u32]
        L.oracle_flight_log.restype = f32
        L.oracle_flight_log.argtypes = [f32]
        L.oracle_half_to_float.restype = f32
        L.oracle_half_to_float.argtypes = [C.c_uint16]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def affine_from_origin_spacing(origin, spacing):
    """Row-major 3x4 index->object transform for gridOrigin / gridSpacing."""
    m = np.zeros(12, np.float32)
    m[0], m[4], m[8] = spacing
    m[9:12] = origin
    return m


class Grid:
    """One oracle volume (dense structuredRegular or sparse vdb)."""

    def __init__(self, handle, keep):
        if not handle:
            raise RuntimeError(lib().oracle_last_error().decode())
        self.h = handle
        self._keep = keep
        self.set_sampler()

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_free(self.h)
            self.h = None

    # -- construction ---------------------------------------------------------
    @classmethod
    def dense(cls, data, origin=(0, 0, 0), spacing=(1, 1, 1), i2o=None, cell_centered=False,
              index_origin=(0, 0, 0), background=None):
        """data: array [nz, ny, nx] (x fastest) or a list of them (attributes)."""
        arrs = data if isinstance(data, (list, tuple)) else [data]
        arrs = [np.ascontiguousarray(a) for a in arrs]
        nz, ny, nx = arrs[0].shape
        dims = np.array([nx, ny, nz], np.int32)
        io = np.array(index_origin, np.int32)
        m = np.asarray(i2o, np.float32) if i2o is not None else affine_from_origin_spacing(origin, spacing)
        dt = np.array([DTYPE_CODE[a.dtype] for a in arrs], np.uint32)
        ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        bg = None if background is None else np.asarray(background, np.float32).reshape(-1)
        h = lib().oracle_dense_new(_p(dims), _p(io), _p(m), int(cell_centered), len(arrs), _p(dt), ptrs, _p(bg))
        return cls(h, (arrs, dims, io, m, dt, ptrs, bg))

    @classmethod
    def dense_temporal(cls, dims, data, structured_steps=None, tu_indices=None, tu_times=None, origin=(0, 0, 0),
                       spacing=(1, 1, 1), i2o=None, cell_centered=False, index_origin=(0, 0, 0), background=None):
        """structuredRegular with motion-blur data: data is a 1-D array (or a
        list of them) holding structured_steps values per voxel (x-fastest
        voxels, time fastest), or tu_indices[voxels] values addressed through
        tu_indices / tu_times."""
        arrs = data if isinstance(data, (list, tuple)) else [data]
        arrs = [np.ascontiguousarray(a).reshape(-1) for a in arrs]
        d = np.asarray(dims, np.int32)
        io = np.array(index_origin, np.int32)
        m = np.asarray(i2o, np.float32) if i2o is not None else affine_from_origin_spacing(origin, spacing)
        dt = np.array([DTYPE_CODE[a.dtype] for a in arrs], np.uint32)
        ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        bg = None if background is None else np.asarray(background, np.float32).reshape(-1)
        if structured_steps:
            fmt, idx, tms, steps = 1, None, None, int(structured_steps)
        else:
            fmt, steps = 2, 0
            idx = np.ascontiguousarray(tu_indices, np.uint64)
            tms = np.ascontiguousarray(tu_times, np.float32)
        h = lib().oracle_dense_new_temporal(_p(d), _p(io), _p(m), int(cell_centered), len(arrs), _p(dt), ptrs, _p(bg),
                                            fmt, steps, _p(idx), _p(tms), 0 if tms is None else len(tms))
        return cls(h, (arrs, d, io, m, dt, ptrs, bg, idx, tms))

    @classmethod
    def vdb_temporal(cls, level, origin, fmt, leaf_data, temporal_format, steps=None, tu_indices=None,
                     tu_times=None, i2o=None, background=None):
        """vdb with per-node motion blur (node.data form). temporal_format: per
        node 0 / 1 / 2; steps: per node time steps (structured nodes);
        tu_indices / tu_times: per node arrays (None for other nodes)."""
        level = np.ascontiguousarray(level, np.uint32)
        origin = np.ascontiguousarray(origin, np.int32).reshape(-1, 3)
        fmt = np.ascontiguousarray(fmt, np.uint32)
        n = level.shape[0]
        m = np.asarray(i2o, np.float32) if i2o is not None else affine_from_origin_spacing((0, 0, 0), (1, 1, 1))
        ld = [[np.ascontiguousarray(a) for a in leaf] for leaf in leaf_data]
        na = len(ld[0])
        dt = np.array([DTYPE_CODE[a.dtype] for a in ld[0]], np.uint32)
        ldp = (C.c_void_p * (n * na))(*[a.ctypes.data for leaf in ld for a in leaf])
        tf = np.ascontiguousarray(temporal_format, np.uint32)
        st = np.ascontiguousarray(steps if steps is not None else np.zeros(n), np.uint32)
        ix = [None if a is None else np.ascontiguousarray(a, np.uint64) for a in (tu_indices or [None] * n)]
        tm = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (tu_times or [None] * n)]
        ixp = (C.c_void_p * n)(*[None if a is None else a.ctypes.data for a in ix])
        tmp = (C.c_void_p * n)(*[None if a is None else a.ctypes.data for a in tm])
        cnt = np.array([0 if a is None else a.size for a in tm], np.uint64)
        bg = None if background is None else np.asarray(background, np.float32).reshape(-1)
        keep = [level, origin, fmt, m, ld, ldp, dt, tf, st, ix, tm, ixp, tmp, cnt, bg]
        h = lib().oracle_vdb_new_temporal(n, _p(level), _p(origin), _p(fmt), na, _p(dt), 0, None, None, ldp, _p(m),
                                          _p(bg), None, _p(tf), _p(st), ixp, tmp, _p(cnt))
        g = cls(h, keep)
        g.num_leaves = n
        return g

    @classmethod
    def vdb(cls, level, origin, fmt, packed_dense=None, packed_tile=None, leaf_data=None,
            i2o=None, background=None, clip=None):
        """Sparse vdb. packed_dense / packed_tile: list (per attribute) of 1-D
        arrays; leaf_data: list over leaves of lists over attributes."""
        level = np.ascontiguousarray(level, np.uint32)
        origin = np.ascontiguousarray(origin, np.int32).reshape(-1, 3)
        fmt = np.ascontiguousarray(fmt, np.uint32)
        n = level.shape[0]
        m = np.asarray(i2o, np.float32) if i2o is not None else affine_from_origin_spacing((0, 0, 0), (1, 1, 1))
        keep = [level, origin, fmt, m]
        if packed_dense is not None:
            pd = [np.ascontiguousarray(a) for a in packed_dense]
            pt = [np.ascontiguousarray(a) for a in packed_tile] if packed_tile is not None else None
            na = len(pd)
            dt = np.array([DTYPE_CODE[a.dtype] for a in pd], np.uint32)
            pdp = (C.c_void_p * na)(*[a.ctypes.data for a in pd])
            ptp = (C.c_void_p * na)(*[a.ctypes.data for a in pt]) if pt is not None else None
            keep += [pd, pt, pdp, ptp]
            ldp = None
            packed = 1
        else:
            ld = [[np.ascontiguousarray(a) for a in leaf] for leaf in leaf_data]
            na = len(ld[0])
            dt = np.array([DTYPE_CODE[a.dtype] for a in ld[0]], np.uint32)
            ldp = (C.c_void_p * (n * na))(*[a.ctypes.data for leaf in ld for a in leaf])
            pdp = ptp = None
            keep += [ld, ldp]
            packed = 0
        bg = None if background is None else np.asarray(background, np.float32).reshape(-1)
        cl = None if clip is None else np.asarray(clip, np.int32)
        keep += [dt, bg, cl]
        h = lib().oracle_vdb_new(n, _p(level), _p(origin), _p(fmt), na, _p(dt), packed, pdp, ptp, ldp,
                                 _p(m), _p(bg), _p(cl))
        g = cls(h, keep)
        g.num_leaves = n
        return g

    # -- observers ----------------------------------------------------------------
    def observe_leaf_access(self, on=True):
        """LeafNodeAccess: returns the uint32 array (one entry per input node)
        that later sampling calls mark with 1; on=False stops marking."""
        if not on:
            lib().oracle_set_leaf_access(self.h, None)
            self._access = None
            return None
        self._access = np.zeros(self.num_leaves, np.uint32)
        lib().oracle_set_leaf_access(self.h, self._access.ctypes.data)
        return self._access

    def inner_nodes(self, max_depth=1):
        """InnerNode observer: array [count, 6 + 2 * numAttributes] (object-
        space box, then one value range per attribute)."""
        count = lib().oracle_inner_nodes(self.h, int(max_depth), None, 0)
        nf = 6 + 2 * lib().oracle_num_attributes(self.h)
        out = np.empty((count, nf), np.float32)
        lib().oracle_inner_nodes(self.h, int(max_depth), out.ctypes.data, count)
        return out

    # -- sampler ------------------------------------------------------------------
    def set_sampler(self, filter=LINEAR, gradient_filter=None, max_sampling_depth=3):
        gf = filter if gradient_filter is None else gradient_filter
        lib().oracle_set_sampler(self.h, filter, gf, max_sampling_depth)

    def set_cubic_double(self, on=True):
        lib().oracle_set_cubic_double(self.h, int(bool(on)))

    def sample(self, xyz, attr=0, threads=1, times=None):
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        out = np.empty(xyz.shape[0], np.float32)
        if times is not None:
            tm = np.ascontiguousarray(np.broadcast_to(np.asarray(times, np.float32), (xyz.shape[0],)))
            lib().oracle_sample_times(self.h, xyz.shape[0], _p(xyz), _p(tm), _p(out), attr)
        elif threads > 1:
            lib().oracle_sample_mt(self.h, xyz.shape[0], _p(xyz), _p(out), attr, threads)
        else:
            lib().oracle_sample(self.h, xyz.shape[0], _p(xyz), _p(out), attr)
        return out

    def woodcock(self, org, dir, t_ranges, rng, sigma_max, tf, tf_range, time=0.0, shutter=0.0,
                 max_collisions=1 << 20, attr=0):
        """Woodcock tracking per ray; rng (n x 2 uint32) is updated in place.
        Returns (eventT, samples, collisions)."""
        org = np.ascontiguousarray(org, np.float32).reshape(-1, 3)
        dir = np.ascontiguousarray(dir, np.float32).reshape(-1, 3)
        tr = np.ascontiguousarray(t_ranges, np.float32).reshape(-1, 2)
        assert rng.dtype == np.uint32 and rng.flags["C_CONTIGUOUS"] and rng.size == 2 * len(org)
        tf = np.ascontiguousarray(tf, np.float32).reshape(-1, 4)
        n = len(org)
        ev, sm, cnt = np.empty(n, np.float32), np.empty(n, np.float32), np.empty(n, np.uint32)
        lib().oracle_woodcock(self.h, n, _p(org), _p(dir), _p(tr), _p(rng), float(sigma_max), _p(tf), len(tf),
                              float(tf_range[0]), float(tf_range[1]), float(time), float(shutter),
                              int(max_collisions), attr, _p(ev), _p(sm), _p(cnt))
        return ev, sm, cnt

    def sample_m(self, xyz, attrs):
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        at = np.ascontiguousarray(attrs, np.uint32)
        out = np.empty((xyz.shape[0], at.shape[0]), np.float32)
        lib().oracle_sample_m(self.h, xyz.shape[0], _p(xyz), _p(out), at.shape[0], _p(at))
        return out

    def gradient(self, xyz, attr=0, times=None):
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1, 3)
        out = np.empty((xyz.shape[0], 3), np.float32)
        if times is not None:
            tm = np.ascontiguousarray(np.broadcast_to(np.asarray(times, np.float32), (xyz.shape[0],)))
            lib().oracle_gradient_times(self.h, xyz.shape[0], _p(xyz), _p(tm), _p(out), attr)
        else:
            lib().oracle_gradient(self.h, xyz.shape[0], _p(xyz), _p(out), attr)
        return out

    # -- metadata -----------------------------------------------------------------
    def bounding_box(self):
        out = np.empty(6, np.float32)
        lib().oracle_bounding_box(self.h, _p(out))
        return out

    def value_range(self, attr=0):
        out = np.empty(2, np.float32)
        lib().oracle_value_range(self.h, attr, _p(out))
        return out

    def info(self):
        ro = np.empty(3, np.int32)
        sz = np.empty(3, np.uint32)
        db = np.empty(6, np.float32)
        o2i = np.empty(12, np.float32)
        lib().oracle_grid_info(self.h, _p(ro), _p(sz), _p(db), _p(o2i))
        return dict(root_origin=ro, active_size=sz, domain_box=db, o2i=o2i)

    def level(self, l):
        n = lib().oracle_level_num_nodes(self.h, l)
        res = [64, 32, 16][l]
        na = lib().oracle_num_attributes(self.h)
        org = np.empty((n, 3), np.uint32)
        vox = np.empty(n * res ** 3, np.uint64)
        rng = np.empty((n * res ** 3, na, 2), np.float32)
        lib().oracle_level_copy(self.h, l, _p(org), _p(vox), _p(rng))
        return org, vox, rng

    # -- iterators ----------------------------------------------------------------
    def intervals(self, org, dir, t_range=(0.0, np.inf), attr=0, depth=None, value_ranges=None, max_out=4096):
        depth = map_depth(0.5) if depth is None else depth
        vr = None if value_ranges is None else np.ascontiguousarray(value_ranges, np.float32).reshape(-1, 2)
        nr = 0 if vr is None else vr.shape[0]
        o = np.asarray(org, np.float32)
        d = np.asarray(dir, np.float32)
        t = np.asarray(t_range, np.float32)
        out = np.empty((max_out, 5), np.float32)
        n = lib().oracle_iterate_intervals(self.h, attr, depth, nr, _p(vr), _p(o), _p(d), _p(t), _p(out), max_out)
        return out[:n].copy()

    def ray_march_intervals(self, org, dir, tf, tf_range, t_ranges=None, time=0.0, sampling_rate=1.0, attr=0,
                            depth=None, value_ranges=None, max_samples=1 << 20):
        """Per ray composited (r, g, b, opacity) and samples taken."""
        depth = map_depth(0.5) if depth is None else depth
        vr = None if value_ranges is None else np.ascontiguousarray(value_ranges, np.float32).reshape(-1, 2)
        nr = 0 if vr is None else vr.shape[0]
        org = np.ascontiguousarray(org, np.float32).reshape(-1, 3)
        dir = np.ascontiguousarray(dir, np.float32).reshape(-1, 3)
        tr = None if t_ranges is None else np.ascontiguousarray(t_ranges, np.float32).reshape(-1, 2)
        tf = np.ascontiguousarray(tf, np.float32).reshape(-1, 4)
        n = len(org)
        out, cnt = np.empty((n, 4), np.float32), np.empty(n, np.uint32)
        lib().oracle_ray_march_intervals(self.h, attr, depth, nr, _p(vr), n, _p(org), _p(dir), _p(tr), float(time),
                                         float(sampling_rate), _p(tf), len(tf), float(tf_range[0]),
                                         float(tf_range[1]), int(max_samples), _p(out), _p(cnt))
        return out, cnt

    def hit_render(self, org, dir, values, tf, tf_range, t_ranges=None, time=0.0, attr=0, max_hits=1 << 16):
        """Hit-iterator renderer per ray: (r, g, b, opacity) and hits shaded."""
        v = np.ascontiguousarray(values, np.float32)
        org = np.ascontiguousarray(org, np.float32).reshape(-1, 3)
        dir = np.ascontiguousarray(dir, np.float32).reshape(-1, 3)
        tr = None if t_ranges is None else np.ascontiguousarray(t_ranges, np.float32).reshape(-1, 2)
        tf = np.ascontiguousarray(tf, np.float32).reshape(-1, 4)
        n = len(org)
        out, cnt = np.empty((n, 4), np.float32), np.empty(n, np.uint32)
        lib().oracle_hit_render(self.h, attr, len(v), _p(v), n, _p(org), _p(dir), _p(tr), float(time), _p(tf),
                                len(tf), float(tf_range[0]), float(tf_range[1]), int(max_hits), _p(out), _p(cnt))
        return out, cnt

    def iterate_many(self, org, dir, max_per_ray, values=None, depth=None, attr=0, threads=0):
        """Interval (values None; depth default map_depth(0.5)) or hit
        iteration (iso values; default depth as hits()) of many rays with t
        range [0, inf) on `threads` host threads (0 = all). Returns (results
        [R, max_per_ray, 5 or 3], counts [R])."""
        org = np.ascontiguousarray(org, np.float32).reshape(-1, 3)
        dir = np.ascontiguousarray(dir, np.float32).reshape(-1, 3)
        hits = values is not None
        v = np.ascontiguousarray(values if hits else [], np.float32)
        if depth is None:
            depth = -1 if hits else map_depth(0.5)
        n = len(org)
        out = np.empty((n, max_per_ray, 3 if hits else 5), np.float32)
        cnt = np.empty(n, np.uint32)
        lib().oracle_iterate_many(self.h, attr, depth, int(hits), v.shape[0], _p(v), n, _p(org), _p(dir),
                                  max_per_ray, threads, _p(cnt), _p(out))
        return out, cnt

    def hits(self, org, dir, values, t_range=(0.0, np.inf), attr=0, depth=-1, time=0.0, max_out=1024):
        v = np.ascontiguousarray(values, np.float32)
        o = np.asarray(org, np.float32)
        d = np.asarray(dir, np.float32)
        t = np.asarray(t_range, np.float32)
        out = np.empty((max_out, 3), np.float32)
        n = lib().oracle_iterate_hits(self.h, attr, depth, v.shape[0], _p(v), _p(o), _p(d), _p(t), time,
                                      _p(out), max_out)
        return out[:n].copy()


def map_depth(hint: float) -> int:
    return int(lib().oracle_map_depth(hint))


def tricubic_matrix() -> np.ndarray:
    out = np.empty((64, 64), np.int32)
    lib().oracle_tricubic_matrix(_p(out))
    return out


def half_to_float(h: int) -> float:
    return float(lib().oracle_half_to_float(h))


def tea(v, key, rounds):
    """TEA over (v0, v1) with key (4 words) and `rounds` rounds."""
    st = np.array(v, np.uint32)
    k = np.array(key, np.uint32)
    lib().oracle_tea(_p(st), _p(k), rounds)
    return int(st[0]), int(st[1])


def flight_log(x):
    return np.float32(lib().oracle_flight_log(float(x)))
