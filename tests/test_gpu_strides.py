"""Strided VKLData (structured_regular_volume_strides.cpp,
vdb_volume.cpp:824-950): voxel arrays passed with byteStride 0 (natural),
1, 2 and 3 x the element size, copied (VKL_DATA_DEFAULT) or shared
(VKL_DATA_SHARED_BUFFER, host and device memory), sample bit for bit as the
compact array does, and the compact volume matches the oracle."""
import numpy as np
import pytest

import oracle as O
import openvkl_amd as vkl
from tests import volumes as V

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

DTYPES = [np.uint8, np.int16, np.uint16, np.float16, np.float32, np.float64]


def same_bits(a, b):
    return bool(np.all((a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))))


def wavelet_voxels(dtype, dims=(21, 18, 13)):
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    w = V.wavelet(x / nx * 2 - 1.0, y / ny * 2 - 1.0, z / nz * 2 - 1.0)
    if np.issubdtype(dtype, np.integer):
        info = np.iinfo(dtype)
        w = (w - w.min()) / (w.max() - w.min()) * (info.max - info.min) + info.min
    return w.astype(dtype)


@pytest.mark.parametrize("dtype", DTYPES, ids=lambda t: np.dtype(t).name)
def test_structured_regular_strides(device, dtype):
    data = wavelet_voxels(dtype)
    nz, ny, nx = data.shape
    rng = np.random.default_rng(3)
    pts = (rng.random((4096, 3)) * [nx + 1, ny + 1, nz + 1] - 1).astype(np.float32)
    compact = device.structured_regular(data)
    want = compact.sampler().sample(pts)
    grid = O.Grid.dense(data)
    assert same_bits(want, grid.sample(pts))
    for shared in (False, True):
        for f in (0, 1, 2, 3):
            v = device.structured_regular(data, data_shared=shared, data_stride_factor=f)
            got = v.sampler().sample(pts)
            assert same_bits(got, want), (shared, f)
            v.release()
    # device memory, shared in place with a stride (torch has no uint16
    # element type the binding maps)
    if dtype is not np.uint16:
        on_device = torch.from_numpy(np.ascontiguousarray(data)).cuda()
        for f in (2, 3):
            dv = device.structured_regular(on_device, data_stride_factor=f)
            assert same_bits(dv.sampler().sample(pts), want), f
            dv.release()


@pytest.mark.parametrize("packed", [True, False], ids=["nodesPackedDense", "node.data"])
def test_vdb_strides(device, packed):
    origins, leaves = V.dense_vdb_leaves(32, [V.wavelet])
    leaves = leaves[0]
    n = len(origins)
    lv, fm = np.full(n, 3, np.uint32), np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32)
    rng = np.random.default_rng(4)
    pts = (rng.random((4096, 3)) * 33 - 0.5).astype(np.float32)

    def build(**kw):
        if packed:
            return device.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)], **kw)
        return device.vdb(lv, origins, fm, leaf_data=[[leaves[i]] for i in range(n)], **kw)

    compact = build()
    want = compact.sampler().sample(pts)
    grid = O.Grid.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    assert same_bits(want, grid.sample(pts))
    for shared in (False, True):
        for f in (0, 1, 2, 3):
            v = build(data_shared=shared, data_stride_factor=f)
            assert same_bits(v.sampler().sample(pts), want), (shared, f)
            v.release()
