"""Python binding of the HIP device's C ABI (libopenvkl_hip.so).

A ctypes mirror of the Open VKL C API (openvkl/include/openvkl/*.h and
openvkl/devices/cpu/include/openvkl/device/openvkl.h) -- the same calls,
argument meaning and error behaviour -- plus small helpers that build volumes
from numpy arrays and pass torch tensors (device memory) or numpy arrays
(host memory) to the sampling / iteration calls.

The library must be built (openvkl_amd/build.py, or __graft_entry__.build());
importing this module raises if it is missing -- there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libopenvkl_hip.so")

# ---- enums (values fixed by the reference headers) -----------------------------------
VKL_OBJECT, VKL_DATA, VKL_VOLUME = 0x8000000, 0x8000000 + 100, 0x8000000 + 101
VKL_BOOL, VKL_STRING, VKL_VOID_PTR = 250, 1500, 200
VKL_UCHAR, VKL_SHORT, VKL_USHORT, VKL_INT, VKL_UINT = 2500, 3000, 3500, 4000, 4500
VKL_HALF, VKL_FLOAT, VKL_DOUBLE, VKL_ULONG = 5800, 6000, 7000, 5550
VKL_VEC3I, VKL_VEC3F, VKL_BOX1F, VKL_BOX3I, VKL_AFFINE3F = 4002, 6002, 10000, 8002, 12003
VKL_DATA_DEFAULT, VKL_DATA_SHARED_BUFFER = 0, 1
VKL_FILTER_NEAREST, VKL_FILTER_LINEAR, VKL_FILTER_CUBIC = 0, 100, 200
VKL_FORMAT_TILE, VKL_FORMAT_DENSE_ZYX = 0, 1
VKL_NO_ERROR, VKL_UNKNOWN_ERROR, VKL_INVALID_ARGUMENT = 0, 1, 2

NP_TO_VKL = {
    np.dtype(np.uint8): VKL_UCHAR,
    np.dtype(np.int16): VKL_SHORT,
    np.dtype(np.uint16): VKL_USHORT,
    np.dtype(np.int32): VKL_INT,
    np.dtype(np.uint32): VKL_UINT,
    np.dtype(np.float16): VKL_HALF,
    np.dtype(np.float32): VKL_FLOAT,
    np.dtype(np.float64): VKL_DOUBLE,
}


class VKLObject(C.Structure):
    _fields_ = [("host", C.c_void_p), ("device", C.c_void_p)]


class vec3f(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class range1f(C.Structure):
    _fields_ = [("lower", C.c_float), ("upper", C.c_float)]


class box3f(C.Structure):
    _fields_ = [("lower", vec3f), ("upper", vec3f)]


class VKLInterval(C.Structure):
    _fields_ = [("tRange", range1f), ("valueRange", range1f), ("nominalDeltaT", C.c_float)]


class VKLHit(C.Structure):
    _fields_ = [("t", C.c_float), ("sample", C.c_float), ("epsilon", C.c_float)]


def _wide(w):
    class VV(C.Structure):
        _fields_ = [("x", C.c_float * w), ("y", C.c_float * w), ("z", C.c_float * w)]

    class VR(C.Structure):
        _fields_ = [("lower", C.c_float * w), ("upper", C.c_float * w)]

    class VI(C.Structure):
        _fields_ = [("tRange", VR), ("valueRange", VR), ("nominalDeltaT", C.c_float * w)]

    class VH(C.Structure):
        _fields_ = [("t", C.c_float * w), ("sample", C.c_float * w), ("epsilon", C.c_float * w)]

    return VV, VR, VI, VH


WIDE = {w: _wide(w) for w in (4, 8, 16)}

_lib = None
ERROR_CALLBACK = C.CFUNCTYPE(None, C.c_void_p, C.c_uint32, C.c_char_p)


def lib():
    """The loaded C-ABI library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with openvkl_amd/build.py")
    L = C.CDLL(LIB_PATH)
    vp, u32, i32, f32, sz = C.c_void_p, C.c_uint32, C.c_int, C.c_float, C.c_size_t
    O = VKLObject
    sig = {
        "vklInit": (None, []),
        "vklNewDevice": (vp, [C.c_char_p]),
        "vklDeviceSetErrorCallback": (None, [vp, ERROR_CALLBACK, vp]),
        "vklDeviceSetInt": (None, [vp, C.c_char_p, i32]),
        "vklDeviceSetVoidPtr": (None, [vp, C.c_char_p, vp]),
        "vklCommitDevice": (None, [vp]),
        "vklDeviceGetLastErrorCode": (u32, [vp]),
        "vklDeviceGetLastErrorMsg": (C.c_char_p, [vp]),
        "vklGetNativeSIMDWidth": (i32, [vp]),
        "vklReleaseDevice": (None, [vp]),
        "vklCommit": (None, [O]),
        "vklRelease": (None, [O]),
        "vklNewData": (O, [vp, sz, u32, vp, u32, sz]),
        "vklNewVolume": (O, [vp, C.c_char_p]),
        "vklGetBoundingBox": (box3f, [O]),
        "vklGetNumAttributes": (C.c_uint, [O]),
        "vklGetValueRange": (range1f, [O, C.c_uint]),
        "vklNewSampler": (O, [O]),
        "vklNewIntervalIteratorContext": (O, [O]),
        "vklNewVolumeObserver": (O, [O, C.c_char_p]),
        "vklNewSamplerObserver": (O, [O, C.c_char_p]),
        "vklMapObserver": (vp, [O]),
        "vklUnmapObserver": (None, [O]),
        "vklGetObserverElementType": (u32, [O]),
        "vklGetObserverElementSize": (sz, [O]),
        "vklGetObserverNumElements": (sz, [O]),
        "vklNewHitIteratorContext": (O, [O]),
        "vklSetParam": (None, [O, C.c_char_p, u32, vp]),
        "vklSetBool": (None, [O, C.c_char_p, i32]),
        "vklSetFloat": (None, [O, C.c_char_p, f32]),
        "vklSetInt": (None, [O, C.c_char_p, i32]),
        "vklSetVec3f": (None, [O, C.c_char_p, f32, f32, f32]),
        "vklSetVec3i": (None, [O, C.c_char_p, i32, i32, i32]),
        "vklSetData": (None, [O, C.c_char_p, O]),
        "vklComputeSample": (f32, [C.POINTER(O), C.POINTER(vec3f), C.c_uint, f32]),
        "vklComputeSampleN": (None, [C.POINTER(O), C.c_uint, vp, vp, C.c_uint, vp]),
        "vklComputeSampleM": (None, [C.POINTER(O), C.POINTER(vec3f), vp, C.c_uint, vp, f32]),
        "vklComputeSampleMN": (None, [C.POINTER(O), C.c_uint, vp, vp, C.c_uint, vp, vp]),
        "vklComputeGradient": (vec3f, [C.POINTER(O), C.POINTER(vec3f), C.c_uint, f32]),
        "vklComputeGradientN": (None, [C.POINTER(O), C.c_uint, vp, vp, C.c_uint, vp]),
        "vklComputeSampleNAsync": (None, [C.POINTER(O), C.c_uint, vp, vp, C.c_uint, vp]),
        "vklComputeSampleNNoAlloc": (None, [C.POINTER(O), C.c_uint, vp, vp, C.c_uint, vp]),
        "vklComputeGradientNAsync": (None, [C.POINTER(O), C.c_uint, vp, vp, C.c_uint, vp]),
        "vklGetIntervalIteratorSize": (sz, [C.POINTER(O)]),
        "vklInitIntervalIterator": (vp, [C.POINTER(O), C.POINTER(vec3f), C.POINTER(vec3f), C.POINTER(range1f), f32, vp]),
        "vklIterateInterval": (i32, [vp, C.POINTER(VKLInterval)]),
        "vklGetHitIteratorSize": (sz, [C.POINTER(O)]),
        "vklInitHitIterator": (vp, [C.POINTER(O), C.POINTER(vec3f), C.POINTER(vec3f), C.POINTER(range1f), f32, vp]),
        "vklIterateHit": (i32, [vp, C.POINTER(VKLHit)]),
        "vklComputeRayMarchN": (None, [C.POINTER(O), C.c_uint, vp, vp, vp, C.c_uint, vp, C.c_uint]),
        "vklComputeWoodcockN": (None, [C.POINTER(O), C.c_uint, vp, vp, vp, vp, f32, vp, C.c_uint, range1f, f32, f32,
                                       C.c_uint, C.c_uint, vp, vp, vp]),
        "vklComputeRayMarchIntervalsN": (None, [C.POINTER(O), C.c_uint, vp, vp, vp, f32, f32, vp, C.c_uint, range1f,
                                                C.c_uint, vp, vp]),
        "vklComputeHitRenderN": (None, [C.POINTER(O), C.c_uint, vp, vp, vp, f32, vp, C.c_uint, range1f, C.c_uint, vp,
                                        vp]),
        "vklIterateIntervalsBatch": (None, [C.POINTER(O), C.c_uint, vp, vp, vp, vp, vp, C.c_uint, vp]),
        "vklIterateHitsBatch": (None, [C.POINTER(O), C.c_uint, vp, vp, vp, vp, vp, C.c_uint, vp]),
        "vklNewSampleBatchPlan": (vp, [C.c_uint, vp, vp, vp, vp, vp, i32]),
        "vklExecuteSampleBatchPlan": (None, [vp, vp]),
        "vklGetSampleBatchPlanLaunch": (None, [vp, C.POINTER(C.c_uint), C.POINTER(C.c_uint), C.POINTER(C.c_uint)]),
        "vklGetSampleBatchPlanLaunchCount": (C.c_uint, [vp]),
        "vklReleaseSampleBatchPlan": (None, [vp]),
        "vklNewRayBatchPlan": (vp, [C.c_uint, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "vklExecuteRayBatchPlan": (None, [vp, vp]),
        "vklGetRayBatchPlanLaunchCount": (C.c_uint, [vp]),
        "vklReleaseRayBatchPlan": (None, [vp]),
        "vklHipGetVolumeLevel": (i32, [O, C.c_uint, C.POINTER(C.c_uint64), C.POINTER(C.POINTER(C.c_uint64)),
                                       C.POINTER(C.POINTER(range1f))]),
        "vklNewNeuralField": (vp, [C.POINTER(O), C.c_uint, C.c_uint, C.c_uint]),
        "vklReleaseNeuralField": (None, [vp]),
        "vklNeuralFieldTrain": (f32, [vp, C.c_uint, C.c_uint, f32, f32, C.c_uint, vp]),
        "vklNeuralFieldFinalize": (C.c_uint, [vp, f32, C.c_uint, f32, C.c_uint]),
        "vklNeuralFieldSampleN": (None, [vp, C.c_uint, vp, vp, i32, C.POINTER(C.c_uint), vp]),
        "vklNeuralFieldGradients": (f32, [vp, C.c_uint, vp, vp, vp, vp, vp, vp]),
        "vklNeuralFieldGetParams": (None, [vp, vp, vp, vp, vp, vp, vp, C.POINTER(f32)]),
        "vklNeuralFieldNumApprovedLeaves": (C.c_uint, [vp, C.POINTER(C.c_uint)]),
    }
    for w in (4, 8, 16):
        VV, VR, VI, VH = WIDE[w]
        sig[f"vklComputeSample{w}"] = (None, [vp, C.POINTER(O), C.POINTER(VV), vp, C.c_uint, vp])
        sig[f"vklComputeGradient{w}"] = (None, [vp, C.POINTER(O), C.POINTER(VV), C.POINTER(VV), C.c_uint, vp])
        sig[f"vklComputeSampleM{w}"] = (None, [vp, C.POINTER(O), C.POINTER(VV), vp, C.c_uint, vp, vp])
        sig[f"vklGetIntervalIteratorSize{w}"] = (sz, [C.POINTER(O)])
        sig[f"vklInitIntervalIterator{w}"] = (vp, [vp, C.POINTER(O), C.POINTER(VV), C.POINTER(VV), C.POINTER(VR), vp, vp])
        sig[f"vklIterateInterval{w}"] = (None, [vp, vp, C.POINTER(VI), vp])
        sig[f"vklGetHitIteratorSize{w}"] = (sz, [C.POINTER(O)])
        sig[f"vklInitHitIterator{w}"] = (vp, [vp, C.POINTER(O), C.POINTER(VV), C.POINTER(VV), C.POINTER(VR), vp, vp])
        sig[f"vklIterateHit{w}"] = (None, [vp, vp, C.POINTER(VH), vp])
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


# ---- pointer helpers ----------------------------------------------------------------------

def _ptr(a):
    """Address of a numpy array or torch tensor (None -> NULL)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("arrays passed to the library must be C-contiguous")
        return a.ctypes.data
    if not a.is_contiguous():  # torch.Tensor (host or device)
        raise ValueError("tensors passed to the library must be contiguous")
    return a.data_ptr()


def _tf_rows(tf):
    """A transfer function as contiguous float32 rows (r, g, b, opacity):
    numpy or torch, [n, 4] or flat [4n]. Other element types are converted;
    a size that is not a multiple of 4 raises."""
    size = int(tf.size if isinstance(tf, np.ndarray) else tf.numel())
    if size == 0 or size % 4:
        raise ValueError(f"a transfer function holds (r, g, b, opacity) rows; got {size} values")
    if isinstance(tf, np.ndarray):
        return np.ascontiguousarray(tf, np.float32).reshape(-1, 4)
    import torch

    return tf.to(torch.float32).reshape(-1, 4).contiguous()


def _n_points(a):
    return int(a.shape[0])


class VklError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"VKL error {code}: {msg}")
        self.code = code
        self.msg = msg


class Device:
    """vklNewDevice("hip") + vklCommitDevice. Errors are collected, not printed."""

    def __init__(self, name: str = "hip", stream=None):
        L = lib()
        self.h = L.vklNewDevice(name.encode())
        if not self.h:
            raise RuntimeError(f"vklNewDevice({name!r}) failed")
        self.errors = []
        self._cb = ERROR_CALLBACK(lambda u, c, m: self.errors.append((int(c), m.decode())))
        L.vklDeviceSetErrorCallback(self.h, self._cb, None)
        if stream is not None:
            L.vklDeviceSetVoidPtr(self.h, b"stream", stream)
        L.vklCommitDevice(self.h)
        self.check()

    def last_error(self):
        L = lib()
        return int(L.vklDeviceGetLastErrorCode(self.h)), L.vklDeviceGetLastErrorMsg(self.h).decode()

    def check(self):
        if self.errors:
            code, msg = self.errors[-1]
            self.errors.clear()
            raise VklError(code, msg)

    def release(self):
        if self.h:
            lib().vklReleaseDevice(self.h)
            self.h = None

    # -- object creation ------------------------------------------------------------------
    def data(self, array, dtype=None, shared=False, stride_factor=None):
        """vklNewData from a numpy array (copied unless shared) or a torch
        tensor (always shared). stride_factor f: the items are placed f
        elements apart in a padded buffer and passed with byteStride f x the
        element size (f = 0 passes byteStride 0, the natural stride)."""
        if isinstance(array, np.ndarray):
            a = np.ascontiguousarray(array)
            t = dtype if dtype is not None else NP_TO_VKL[a.dtype]
            n = a.size if t not in (VKL_VEC3I, VKL_VEC3F, VKL_BOX1F) else a.size // (3 if t != VKL_BOX1F else 2)
            stride = 0
            if stride_factor is not None:
                f = max(int(stride_factor), 1)
                buf = np.zeros(a.size * f, a.dtype)
                buf[::f] = a.reshape(-1)
                a, stride = buf, int(stride_factor) * a.itemsize
            h = lib().vklNewData(self.h, n, t, a.ctypes.data, VKL_DATA_SHARED_BUFFER if shared else 0, stride)
            obj = Handle(self, h, keep=a if shared else None)
        else:  # torch tensor, always shared (device memory stays in place)
            import torch

            tmap = {torch.float32: VKL_FLOAT, torch.float16: VKL_HALF, torch.uint8: VKL_UCHAR,
                    torch.int16: VKL_SHORT, torch.float64: VKL_DOUBLE, torch.int32: VKL_INT}
            t = dtype if dtype is not None else tmap[array.dtype]
            n, stride = array.numel(), 0
            if stride_factor is not None:
                f = max(int(stride_factor), 1)
                buf = torch.zeros(n * f, dtype=array.dtype, device=array.device)
                buf[::f] = array.reshape(-1)
                array, stride = buf, int(stride_factor) * buf.element_size()
            h = lib().vklNewData(self.h, n, t, array.data_ptr(), VKL_DATA_SHARED_BUFFER, stride)
            obj = Handle(self, h, keep=array)
        self.check()
        return obj

    def data_of_data(self, items):
        """VKL_DATA array of handles (None entries are null objects)."""
        arr = (VKLObject * len(items))(*[VKLObject() if i is None else i.h for i in items])
        h = lib().vklNewData(self.h, len(items), VKL_DATA, C.addressof(arr), 0, 0)
        self.check()
        return Handle(self, h)

    def new_volume(self, type_name):
        h = lib().vklNewVolume(self.h, type_name.encode())
        self.check()
        return Volume(self, h)

    def structured_regular(self, voxels, origin=(0, 0, 0), spacing=(1, 1, 1), filter=None, cell_centered=None,
                           background=None, index_origin=None, index_to_object=None, brick_shift=None,
                           corner_packed=None, dims=None, structured_steps=None, tu_indices=None, tu_times=None,
                           data_shared=False, data_stride_factor=None, cubic_records=None, cubic_stencils=None):
        """voxels: array [nz, ny, nx] (x fastest) or a list of them. Motion
        blur: pass dims (nx, ny, nz) and 1-D voxels with structured_steps
        values per voxel, or tu_indices / tu_times (temporallyUnstructured*;
        uint32 or uint64 indices). data_shared / data_stride_factor: how the
        voxel VKLData are made (see data()). cubic_records: also store the
        per-voxel records of the wavefront-cooperative tricubic kernels
        ("cubicRecords", 16x the voxel memory); cubic_stencils: per cell its
        whole stencil instead ("cubicStencils", 64x)."""
        arrs = voxels if isinstance(voxels, (list, tuple)) else [voxels]
        v = self.new_volume("structuredRegular")
        nx, ny, nz = dims if dims is not None else arrs[0].shape[::-1]
        v.set_vec3i("dimensions", nx, ny, nz)
        if structured_steps is not None:
            v.set_int("temporallyStructuredNumTimesteps", structured_steps)
        if tu_indices is not None:
            ti = np.ascontiguousarray(tu_indices)
            d = self.data(ti, VKL_ULONG if ti.dtype == np.uint64 else VKL_UINT)
            v.set_data("temporallyUnstructuredIndices", d)
            d.release()
        if tu_times is not None:
            d = self.data(np.ascontiguousarray(tu_times, np.float32))
            v.set_data("temporallyUnstructuredTimes", d)
            d.release()
        v.set_vec3f("gridOrigin", *origin)
        v.set_vec3f("gridSpacing", *spacing)
        datas = [self.data(a if isinstance(a, np.ndarray) else a.reshape(-1), shared=data_shared,
                           stride_factor=data_stride_factor) for a in arrs]
        # Device tensors are shared (VKL_DATA_SHARED_BUFFER): the volume reads
        # them in place, so it keeps them alive.
        v._shared = [d._keep for d in datas if d._keep is not None]
        if len(datas) == 1:
            v.set_data("data", datas[0])
        else:
            dd = self.data_of_data(datas)
            v.set_data("data", dd)
            dd.release()
        for d in datas:
            d.release()
        if filter is not None:
            v.set_int("filter", filter)
        if cell_centered is not None:
            v.set_bool("cellCentered", cell_centered)
        if background is not None:
            v.set_float("background", background)
        if index_origin is not None:
            v.set_vec3i("indexOrigin", *index_origin)
        if brick_shift is not None:
            v.set_vec3i("brickShift", *brick_shift)
        if corner_packed is not None:
            v.set_bool("cornerPacked", corner_packed)
        if cubic_records is not None:
            v.set_bool("cubicRecords", cubic_records)
        if cubic_stencils is not None:
            v.set_bool("cubicStencils", cubic_stencils)
        if index_to_object is not None:
            d = self.data(np.asarray(index_to_object, np.float32))
            v.set_data("indexToObject", d)
            d.release()
        v.commit()
        return v

    def vdb(self, level, origin, fmt, packed_dense=None, packed_tile=None, leaf_data=None, background=None,
            index_to_object=None, filter=None, corner_packed=None, max_sampling_depth=None, temporal_format=None,
            steps=None, tu_indices=None, tu_times=None, leaf_index_grid=None, data_shared=False,
            data_stride_factor=None):
        """Sparse vdb volume from node arrays (packed or per-leaf data).
        Motion blur (per-leaf data only): temporal_format per node, steps per
        node (node.temporallyStructuredNumTimesteps), tu_indices / tu_times
        per node (arrays, None for other nodes). data_shared /
        data_stride_factor: how the voxel VKLData are made (see data())."""
        v = self.new_volume("vdb")
        tmp = []

        def d(a, t=None):
            h = self.data(a, t)
            tmp.append(h)
            return h

        def vox(a):
            h = self.data(a if isinstance(a, np.ndarray) else a.reshape(-1), shared=data_shared,
                          stride_factor=data_stride_factor)
            tmp.append(h)
            return h

        v.set_data("node.level", d(np.ascontiguousarray(level, np.uint32)))
        v.set_data("node.origin", d(np.ascontiguousarray(origin, np.int32).reshape(-1), VKL_VEC3I))
        v.set_data("node.format", d(np.ascontiguousarray(fmt, np.uint32)))
        if packed_dense is not None:
            pd = self.data_of_data([vox(a) for a in packed_dense])
            tmp.append(pd)
            v.set_data("nodesPackedDense", pd)
            if packed_tile is not None:
                pt = self.data_of_data([d(a) for a in packed_tile])
                tmp.append(pt)
                v.set_data("nodesPackedTile", pt)
        else:
            per_leaf = []
            for leaf in leaf_data:
                if len(leaf) == 1:
                    per_leaf.append(vox(leaf[0]))
                else:
                    x = self.data_of_data([vox(a) for a in leaf])
                    tmp.append(x)
                    per_leaf.append(x)
            nd = self.data_of_data(per_leaf)
            tmp.append(nd)
            v.set_data("node.data", nd)
        if background is not None:
            v.set_float("background", background)
        if filter is not None:
            v.set_int("filter", filter)
        if index_to_object is not None:
            v.set_data("indexToObject", d(np.asarray(index_to_object, np.float32)))
        if corner_packed is not None:
            v.set_bool("cornerPacked", corner_packed)
        if leaf_index_grid is not None:
            v.set_int("leafIndexGrid", int(leaf_index_grid))
        if max_sampling_depth is not None:
            v.set_int("maxSamplingDepth", max_sampling_depth)
        if temporal_format is not None:
            v.set_data("node.temporalFormat", d(np.ascontiguousarray(temporal_format, np.uint32)))
        if steps is not None:
            v.set_data("node.temporallyStructuredNumTimesteps", d(np.ascontiguousarray(steps, np.int32)))
        for name, arrs in (("node.temporallyUnstructuredIndices", tu_indices),
                           ("node.temporallyUnstructuredTimes", tu_times)):
            if arrs is None:
                continue
            items = []
            for a in arrs:
                if a is None:
                    items.append(None)
                else:
                    a = np.ascontiguousarray(a)
                    items.append(d(a, VKL_ULONG if a.dtype == np.uint64 else None))
            dd = self.data_of_data(items)
            tmp.append(dd)
            v.set_data(name, dd)
        # Shared buffers (device tensors, data_shared host arrays) are read in
        # place by the volume: it keeps them alive.
        v._shared = [h._keep for h in tmp if h._keep is not None]
        v.commit()
        for h in tmp:
            h.release()
        return v


class Handle:
    """A VKLObject handle with typed setters."""

    def __init__(self, device, h, keep=None):
        self.device = device
        self.h = h
        self._keep = keep

    def _set(self, fn, name, *args):
        getattr(lib(), fn)(self.h, name.encode(), *args)
        self.device.check()

    def set_int(self, name, x):
        self._set("vklSetInt", name, int(x))

    def set_float(self, name, x):
        self._set("vklSetFloat", name, float(x))

    def set_bool(self, name, b):
        self._set("vklSetBool", name, int(bool(b)))

    def set_vec3f(self, name, x, y, z):
        self._set("vklSetVec3f", name, float(x), float(y), float(z))

    def set_vec3i(self, name, x, y, z):
        self._set("vklSetVec3i", name, int(x), int(y), int(z))

    def set_data(self, name, data):
        self._set("vklSetData", name, data.h)

    def commit(self):
        lib().vklCommit(self.h)
        self.device.check()

    def release(self):
        if self.h is not None and self.h.host:
            lib().vklRelease(self.h)
            self.h = VKLObject()


class Volume(Handle):
    def bounding_box(self):
        b = lib().vklGetBoundingBox(self.h)
        return np.array([b.lower.x, b.lower.y, b.lower.z, b.upper.x, b.upper.y, b.upper.z], np.float32)

    def value_range(self, attr=0):
        r = lib().vklGetValueRange(self.h, attr)
        return np.array([r.lower, r.upper], np.float32)

    def num_attributes(self):
        return int(lib().vklGetNumAttributes(self.h))

    def level(self, l):
        n = C.c_uint64()
        vox = C.POINTER(C.c_uint64)()
        rng = C.POINTER(range1f)()
        ok = lib().vklHipGetVolumeLevel(self.h, l, C.byref(n), C.byref(vox), C.byref(rng))
        self.device.check()
        assert ok
        res = [64, 32, 16][l]
        count = n.value * res ** 3
        na = self.num_attributes()
        v = np.ctypeslib.as_array(vox, shape=(count,)).copy() if count else np.zeros(0, np.uint64)
        r = (np.ctypeslib.as_array(C.cast(rng, C.POINTER(C.c_float)), shape=(count * na * 2,)).reshape(count, na, 2).copy()
             if count else np.zeros((0, na, 2), np.float32))
        return n.value, v, r

    def observer(self, type_name="InnerNode", max_depth=None):
        """vklNewVolumeObserver ("InnerNode" on vdb volumes; int "maxDepth")."""
        h = lib().vklNewVolumeObserver(self.h, type_name.encode())
        self.device.check()
        o = Observer(self.device, h)
        if max_depth is not None:
            o.set_int("maxDepth", max_depth)
            o.commit()
        return o

    def sampler(self, filter=None, gradient_filter=None, max_sampling_depth=None, binned=None):
        """binned: None = automatic, False / True = \"binnedSampling\" 0 / 1 (spatially
        binned stream sampling of dense volumes, linear filter)."""
        h = lib().vklNewSampler(self.h)
        self.device.check()
        s = Sampler(self.device, h)
        s.volume = self
        if binned is not None:
            s.set_int("binnedSampling", int(bool(binned)))
        if filter is not None:
            s.set_int("filter", filter)
        if gradient_filter is not None:
            s.set_int("gradientFilter", gradient_filter)
        if max_sampling_depth is not None:
            s.set_int("maxSamplingDepth", max_sampling_depth)
        s.commit()
        return s


class Observer(Handle):
    """vklMapObserver / vklUnmapObserver and the element queries."""

    def map(self):
        """Maps, copies the elements out and unmaps: uint32 array for
        LeafNodeAccess, float32 [n, elementSize / 4] for InnerNode."""
        L = lib()
        p = L.vklMapObserver(self.h)
        self.device.check()
        try:
            t, es, n = (L.vklGetObserverElementType(self.h), L.vklGetObserverElementSize(self.h),
                        L.vklGetObserverNumElements(self.h))
            self.device.check()
            if t == VKL_UINT:
                return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint32)), shape=(n,)).copy() if n else \
                    np.zeros(0, np.uint32)
            nf = es // 4
            if not n:
                return np.zeros((0, nf), np.float32)
            return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float)), shape=(n * nf,)).reshape(n, nf).copy()
        finally:
            L.vklUnmapObserver(self.h)

    def element_type(self):
        return int(lib().vklGetObserverElementType(self.h))


class Sampler(Handle):
    def observer(self, type_name="LeafNodeAccess"):
        """vklNewSamplerObserver ("LeafNodeAccess" on samplers of vdb volumes)."""
        h = lib().vklNewSamplerObserver(self.h, type_name.encode())
        self.device.check()
        return Observer(self.device, h)

    def _out(self, like, shape, dtype=np.float32):
        if isinstance(like, np.ndarray):
            return np.empty(shape, dtype)
        import torch

        return torch.empty(shape, dtype=torch.float32, device=like.device)

    def sample(self, xyz, attr=0, out=None, times=None):
        """vklComputeSampleN; xyz: [N,3] float32 numpy (host) or torch (device);
        times: [N] float32 like xyz, or None (time 0)."""
        n = _n_points(xyz)
        out = self._out(xyz, (n,)) if out is None else out
        lib().vklComputeSampleN(C.byref(self.h), n, _ptr(xyz), _ptr(out), attr, _ptr(times))
        self.device.check()
        return out

    def sample_noalloc(self, xyz, attr=0, out=None, times=None):
        """vklComputeSampleNNoAlloc: as sample(), device tensors only; the
        binned kernels' scratch is the library's fixed per-GPU array."""
        n = _n_points(xyz)
        out = self._out(xyz, (n,)) if out is None else out
        lib().vklComputeSampleNNoAlloc(C.byref(self.h), n, _ptr(xyz), _ptr(out), attr, _ptr(times))
        self.device.check()
        return out

    def sample_async(self, xyz, out, attr=0, stream=None):
        lib().vklComputeSampleNAsync(C.byref(self.h), _n_points(xyz), _ptr(xyz), _ptr(out), attr, stream)
        self.device.check()
        return out

    def gradient(self, xyz, attr=0, out=None, times=None):
        n = _n_points(xyz)
        out = self._out(xyz, (n, 3)) if out is None else out
        lib().vklComputeGradientN(C.byref(self.h), n, _ptr(xyz), _ptr(out), attr, _ptr(times))
        self.device.check()
        return out

    def ray_march(self, org, dir, t_ranges, steps, attr=0, out=None):
        """vklComputeRayMarchN: per-ray sum of `steps` midpoint samples."""
        r = _n_points(org)
        out = self._out(org, (r,)) if out is None else out
        lib().vklComputeRayMarchN(C.byref(self.h), r, _ptr(org), _ptr(dir), _ptr(t_ranges), steps, _ptr(out), attr)
        self.device.check()
        return out

    def woodcock(self, org, dir, t_ranges, rng, sigma_max, tf, tf_range, time=0.0, shutter=0.0,
                 max_collisions=1 << 20, attr=0, counts=True):
        """vklComputeWoodcockN: one free-flight (delta tracking) event per ray.
        rng (numRays x 2 uint32, TEA state) is updated in place; tf holds
        (r, g, b, opacity) rows over tf_range. Returns (eventT, samples,
        collisions or None)."""
        r = _n_points(org)
        ev = self._out(org, (r,))
        sm = self._out(org, (r,))
        cnt = None
        if counts:
            if isinstance(org, np.ndarray):
                cnt = np.empty(r, np.uint32)
            else:
                import torch

                cnt = torch.empty(r, dtype=torch.int32, device=org.device)
        tfa = _tf_rows(tf)
        lib().vklComputeWoodcockN(C.byref(self.h), r, _ptr(org), _ptr(dir), _ptr(t_ranges), _ptr(rng),
                                  float(sigma_max), _ptr(tfa), int(tfa.shape[0]), range1f(*[float(x) for x in tf_range]),
                                  float(time), float(shutter), int(max_collisions), attr, _ptr(ev), _ptr(sm),
                                  _ptr(cnt))
        self.device.check()
        return ev, sm, cnt

    def sample_m(self, xyz, attrs, times=None):
        n = _n_points(xyz)
        at = np.ascontiguousarray(attrs, np.uint32)
        out = self._out(xyz, (n, at.size))
        lib().vklComputeSampleMN(C.byref(self.h), n, _ptr(xyz), _ptr(out), at.size, at.ctypes.data, _ptr(times))
        self.device.check()
        return out

    def sample1(self, p, attr=0, time=0.0):
        v = vec3f(*[float(x) for x in p])
        r = lib().vklComputeSample(C.byref(self.h), C.byref(v), attr, float(time))
        self.device.check()
        return r

    def gradient1(self, p, attr=0):
        v = vec3f(*[float(x) for x in p])
        r = lib().vklComputeGradient(C.byref(self.h), C.byref(v), attr, 0.0)
        self.device.check()
        return np.array([r.x, r.y, r.z], np.float32)

    def sample_wide(self, pts, attr=0, width=4, valid=None, times=None):
        VV = WIDE[width][0]
        vv = VV()
        k = len(pts)
        for i in range(k):
            vv.x[i], vv.y[i], vv.z[i] = [float(c) for c in pts[i]]
        val = (C.c_int * width)(*([1] * k + [0] * (width - k))) if valid is None else (C.c_int * width)(*valid)
        out = (C.c_float * width)()
        tm = None
        if times is not None:
            tm = (C.c_float * width)(*([float(x) for x in times] + [0.0] * (width - len(times))))
        getattr(lib(), f"vklComputeSample{width}")(C.addressof(val), C.byref(self.h), C.byref(vv), C.addressof(out),
                                                   attr, None if tm is None else C.addressof(tm))
        self.device.check()
        return np.array(out[:k], np.float32)

    def interval_context(self, attr=0, value_ranges=None, hint=None):
        h = lib().vklNewIntervalIteratorContext(self.h)
        self.device.check()
        c = IteratorContext(self.device, h)
        c.set_int("attributeIndex", attr)
        if value_ranges is not None:
            d = self.device.data(np.asarray(value_ranges, np.float32).reshape(-1), VKL_BOX1F)
            c.set_data("valueRanges", d)
            d.release()
        if hint is not None:
            c.set_float("intervalResolutionHint", hint)
        c.commit()
        c.hit = False
        return c

    def hit_context(self, values, attr=0):
        h = lib().vklNewHitIteratorContext(self.h)
        self.device.check()
        c = IteratorContext(self.device, h)
        c.set_int("attributeIndex", attr)
        d = self.device.data(np.asarray(values, np.float32))
        c.set_data("values", d)
        d.release()
        c.commit()
        c.hit = True
        return c


class IteratorContext(Handle):
    def intervals(self, org, dir, t_range=(0.0, np.inf), time=0.0, max_out=100000):
        """Scalar API: vklInitIntervalIterator + vklIterateInterval loop."""
        L = lib()
        buf = C.create_string_buffer(L.vklGetIntervalIteratorSize(C.byref(self.h)))
        o, d = vec3f(*map(float, org)), vec3f(*map(float, dir))
        t = range1f(*map(float, t_range))
        it = L.vklInitIntervalIterator(C.byref(self.h), C.byref(o), C.byref(d), C.byref(t), time, buf)
        self.device.check()
        out = []
        iv = VKLInterval()
        while len(out) < max_out and L.vklIterateInterval(it, C.byref(iv)):
            out.append([iv.tRange.lower, iv.tRange.upper, iv.valueRange.lower, iv.valueRange.upper, iv.nominalDeltaT])
        self.device.check()
        return np.array(out, np.float32).reshape(-1, 5)

    def hits(self, org, dir, t_range=(0.0, np.inf), time=0.0, max_out=10000):
        L = lib()
        buf = C.create_string_buffer(L.vklGetHitIteratorSize(C.byref(self.h)))
        o, d = vec3f(*map(float, org)), vec3f(*map(float, dir))
        t = range1f(*map(float, t_range))
        it = L.vklInitHitIterator(C.byref(self.h), C.byref(o), C.byref(d), C.byref(t), time, buf)
        self.device.check()
        out = []
        h = VKLHit()
        while len(out) < max_out and L.vklIterateHit(it, C.byref(h)):
            out.append([h.t, h.sample, h.epsilon])
        self.device.check()
        return np.array(out, np.float32).reshape(-1, 3)

    def batch(self, org, dir, t_ranges=None, times=None, max_per_ray=64):
        """vklIterateIntervalsBatch / vklIterateHitsBatch. org/dir: [R,3] numpy
        (host) or torch (device) float32. Returns (results, counts)."""
        r = _n_points(org)
        width = 3 if self.hit else 5
        if isinstance(org, np.ndarray):
            out = np.empty((r, max_per_ray, width), np.float32)
            cnt = np.empty(r, np.uint32)
        else:
            import torch

            out = torch.empty((r, max_per_ray, width), dtype=torch.float32, device=org.device)
            cnt = torch.empty(r, dtype=torch.int32, device=org.device)
        fn = lib().vklIterateHitsBatch if self.hit else lib().vklIterateIntervalsBatch
        fn(C.byref(self.h), r, _ptr(org), _ptr(dir), _ptr(t_ranges), _ptr(times), _ptr(out), max_per_ray, _ptr(cnt))
        self.device.check()
        return out, cnt

    def render_hits(self, org, dir, tf, tf_range, t_ranges=None, time=0.0, max_hits=1 << 16):
        """vklComputeHitRenderN (hit contexts): per-ray shaded (r, g, b,
        opacity) and hits shaded."""
        r = _n_points(org)
        if isinstance(org, np.ndarray):
            out, cnt = np.empty((r, 4), np.float32), np.empty(r, np.uint32)
        else:
            import torch

            out = torch.empty((r, 4), dtype=torch.float32, device=org.device)
            cnt = torch.empty(r, dtype=torch.int32, device=org.device)
        tfa = _tf_rows(tf)
        lib().vklComputeHitRenderN(C.byref(self.h), r, _ptr(org), _ptr(dir), _ptr(t_ranges), float(time), _ptr(tfa),
                                   int(tfa.shape[0]), range1f(*[float(x) for x in tf_range]), int(max_hits), _ptr(out),
                                   _ptr(cnt))
        self.device.check()
        return out, cnt

    def ray_march(self, org, dir, tf, tf_range, t_ranges=None, time=0.0, sampling_rate=1.0,
                  max_samples=1 << 20, counts=True):
        """vklComputeRayMarchIntervalsN (interval contexts): per-ray composited
        (r, g, b, opacity) and samples taken. tf: (r, g, b, opacity) rows."""
        r = _n_points(org)
        if isinstance(org, np.ndarray):
            out = np.empty((r, 4), np.float32)
            cnt = np.empty(r, np.uint32) if counts else None
        else:
            import torch

            out = torch.empty((r, 4), dtype=torch.float32, device=org.device)
            cnt = torch.empty(r, dtype=torch.int32, device=org.device) if counts else None
        tfa = _tf_rows(tf)
        lib().vklComputeRayMarchIntervalsN(C.byref(self.h), r, _ptr(org), _ptr(dir), _ptr(t_ranges), float(time),
                                           float(sampling_rate), _ptr(tfa), int(tfa.shape[0]),
                                           range1f(*[float(x) for x in tf_range]), int(max_samples), _ptr(out),
                                           _ptr(cnt))
        self.device.check()
        return out, cnt


class SampleBatchPlan:
    """vklNewSampleBatchPlan: one graph-captured persistent launch sampling
    many (sampler, device point array) pairs."""

    def __init__(self, device, samplers, points, outs, attrs=None, autotune=True):
        """autotune: True / 1 tunes shape and grouping; False / 0 fixed shape,
        per-type groups; 2 fixed shape, structuredRegular batches of a filter
        merged into one launch."""
        n = len(samplers)
        self.device = device
        self._samplers = (VKLObject * n)(*[s.h for s in samplers])
        self._attrs = (C.c_uint * n)(*(attrs or [0] * n))
        self._counts = (C.c_uint * n)(*[_n_points(p) for p in points])
        self._pts = (C.c_void_p * n)(*[_ptr(p) for p in points])
        self._outs = (C.c_void_p * n)(*[_ptr(o) for o in outs])
        self._keep = (points, outs)
        self.h = lib().vklNewSampleBatchPlan(n, C.addressof(self._samplers), C.addressof(self._attrs),
                                             C.addressof(self._counts), C.addressof(self._pts),
                                             C.addressof(self._outs), int(autotune))
        device.check()
        if not self.h:
            raise RuntimeError("vklNewSampleBatchPlan failed")

    def launch_shape(self):
        b, t, k = C.c_uint(), C.c_uint(), C.c_uint()
        lib().vklGetSampleBatchPlanLaunch(self.h, C.byref(b), C.byref(t), C.byref(k))
        return b.value, t.value, k.value

    def launch_count(self):
        """Kernel launches per execution (one per group of batches)."""
        return int(lib().vklGetSampleBatchPlanLaunchCount(self.h))

    def execute(self, stream=None):
        lib().vklExecuteSampleBatchPlan(self.h, stream)
        self.device.check()

    def release(self):
        if self.h:
            lib().vklReleaseSampleBatchPlan(self.h)
            self.h = None


class RayBatchPlan:
    """vklNewRayBatchPlan: interval / hit iteration of many (context, device
    rays) batches, one graph replay. batches: list of (context, org, dir,
    max_per_ray) or (context, org, dir, max_per_ray, t_ranges, times) with
    torch device tensors. Allocates the outputs: self.results[i] [R, max, 5 or
    3] float32, self.counts[i] [R] int32."""

    def __init__(self, device, batches):
        import torch

        n = len(batches)
        self.device = device
        self.results, self.counts, keep = [], [], []
        ctx = (VKLObject * n)(*[b[0].h for b in batches])
        nr = (C.c_uint * n)(*[_n_points(b[1]) for b in batches])
        mx = (C.c_uint * n)(*[int(b[3]) for b in batches])
        org, dr, tr, tm, res, cnt = ([None] * n for _ in range(6))
        for i, b in enumerate(batches):
            r = _n_points(b[1])
            width = 3 if b[0].hit else 5
            self.results.append(torch.empty((r, int(b[3]), width), dtype=torch.float32, device=b[1].device))
            self.counts.append(torch.empty(r, dtype=torch.int32, device=b[1].device))
            org[i], dr[i] = _ptr(b[1]), _ptr(b[2])
            tr[i] = _ptr(b[4]) if len(b) > 4 else None
            tm[i] = _ptr(b[5]) if len(b) > 5 else None
            res[i], cnt[i] = _ptr(self.results[-1]), _ptr(self.counts[-1])
            keep.append(b)
        arrs = [(C.c_void_p * n)(*a) for a in (org, dr, tr, tm, res, cnt)]
        self._keep = (keep, arrs, ctx, nr, mx)
        self.h = lib().vklNewRayBatchPlan(n, C.addressof(ctx), C.addressof(nr), C.addressof(arrs[0]),
                                          C.addressof(arrs[1]), C.addressof(arrs[2]), C.addressof(arrs[3]),
                                          C.addressof(arrs[4]), C.addressof(mx), C.addressof(arrs[5]))
        device.check()
        if not self.h:
            raise RuntimeError("vklNewRayBatchPlan failed")

    def launch_count(self):
        return int(lib().vklGetRayBatchPlanLaunchCount(self.h))

    def execute(self, stream=None):
        lib().vklExecuteRayBatchPlan(self.h, stream)
        self.device.check()

    def release(self):
        if self.h:
            lib().vklReleaseRayBatchPlan(self.h)
            self.h = None


NF_LEVELS, NF_T = 16, 1 << 19


class NeuralField:
    """vklNewNeuralField: a hash-grid + MLP approximation of a vdb sampler,
    trained on the device from exact samples, with per-leaf error bounds and
    exact fallback for leaves that miss the tolerance."""

    def __init__(self, sampler, attr=0, max_resolution=0, seed=0):
        self.device = sampler.device
        self.sampler = sampler
        self.h = lib().vklNewNeuralField(C.byref(sampler.h), attr, max_resolution, seed)
        self.device.check()
        if not self.h:
            raise RuntimeError("vklNewNeuralField failed")

    def train(self, steps, batch=1 << 16, lr=1e-2, leaf_fraction=0.9, seed=1):
        hist = np.zeros(max(steps, 1), np.float32)
        last = lib().vklNeuralFieldTrain(self.h, steps, batch, lr, leaf_fraction, seed, hist.ctypes.data)
        self.device.check()
        return last, hist[:steps]

    def finalize(self, tolerance, points_per_leaf=64, safety=1.0, seed=7):
        n = lib().vklNeuralFieldFinalize(self.h, tolerance, points_per_leaf, safety, seed)
        self.device.check()
        return n

    def approved(self):
        total = C.c_uint()
        ok = lib().vklNeuralFieldNumApprovedLeaves(self.h, C.byref(total))
        return ok, total.value

    def sample(self, xyz, fallback=True, out=None, workspace=None):
        """xyz: torch device tensor [N, 3]; returns (samples, numFallback)."""
        import torch
        n = _n_points(xyz)
        if out is None:
            out = torch.empty((n,), dtype=torch.float32, device=xyz.device)
        nfb = C.c_uint(0)
        lib().vklNeuralFieldSampleN(self.h, n, _ptr(xyz), _ptr(out), int(fallback), C.byref(nfb), _ptr(workspace))
        self.device.check()
        return out, nfb.value

    def gradients(self, xyz, targets):
        """Summed gradients of sum((y - t)^2) at host points (no optimizer step)."""
        xyz = np.ascontiguousarray(xyz, np.float32)
        targets = np.ascontiguousarray(targets, np.float32)
        g = dict(table=np.zeros((NF_LEVELS, NF_T, 2), np.float32), W1=np.zeros((32, 64), np.float32),
                 W2=np.zeros((64, 64), np.float32), W3=np.zeros(64, np.float32))
        loss = lib().vklNeuralFieldGradients(self.h, xyz.shape[0], xyz.ctypes.data, targets.ctypes.data,
                                             g["table"].ctypes.data, g["W1"].ctypes.data, g["W2"].ctypes.data,
                                             g["W3"].ctypes.data)
        self.device.check()
        return loss, g

    def params(self):
        table = np.zeros((NF_LEVELS, NF_T, 2), np.uint16)
        W1, W2 = np.zeros((32, 64), np.uint16), np.zeros((64, 64), np.uint16)
        W3, res = np.zeros(64, np.float32), np.zeros(NF_LEVELS, np.uint32)
        lo, inv = np.zeros(3, np.float32), C.c_float()
        lib().vklNeuralFieldGetParams(self.h, table.ctypes.data, W1.ctypes.data, W2.ctypes.data, W3.ctypes.data,
                                      res.ctypes.data, lo.ctypes.data, C.byref(inv))
        self.device.check()
        return dict(table=table.view(np.float16), W1=W1.view(np.float16), W2=W2.view(np.float16), W3=W3,
                    res=res, lo=lo, inv_extent=np.float32(inv.value))

    def release(self):
        if self.h:
            lib().vklReleaseNeuralField(self.h)
            self.h = None
