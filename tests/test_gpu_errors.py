"""Error contract (openvkl/api/API.cpp:37-89, vdb_volume.cpp:17-79): invalid
input sets the device's last error code / message; no GPU kernel faults."""
import numpy as np
import pytest

import openvkl_amd as vkl

pytestmark = pytest.mark.gpu


def commit_expect(device, vol, code, text):
    vkl.lib().vklCommit(vol.h)
    c, m = device.last_error()
    device.errors.clear()
    assert c == code and text in m, (c, m)


def vdb_volume(device, level, voxels):
    v = device.new_volume("vdb")
    tmp = [device.data(np.array([level], np.uint32)), device.data(np.zeros(3, np.int32), vkl.VKL_VEC3I),
           device.data(np.array([vkl.VKL_FORMAT_DENSE_ZYX], np.uint32))]
    leaf = device.data(voxels)
    nd = device.data_of_data([leaf])
    v.set_data("node.level", tmp[0])
    v.set_data("node.origin", tmp[1])
    v.set_data("node.format", tmp[2])
    v.set_data("node.data", nd)
    for h in tmp + [leaf, nd]:
        h.release()
    return v


def test_invalid_level(device):
    v = vdb_volume(device, 4, np.zeros(512, np.float32))
    commit_expect(device, v, 1, "invalid node level 4 for this vdb configuration")
    v.release()


def test_node_data_too_small(device):
    v = vdb_volume(device, 3, np.zeros(1, np.float32))
    commit_expect(device, v, 1, "Node data too small")
    v.release()


def test_missing_data(device):
    v = device.new_volume("structuredRegular")
    v.set_vec3i("dimensions", 4, 4, 4)
    commit_expect(device, v, 1, "missing required 'data' parameter")
    v.release()


def test_wrong_element_count(device):
    v = device.new_volume("structuredRegular")
    v.set_vec3i("dimensions", 4, 4, 4)
    d = device.data(np.zeros(10, np.float32))
    v.set_data("data", d)
    d.release()
    commit_expect(device, v, 1, "attribute 0 has 10 elements, but expected 64")
    v.release()


def test_illegal_attribute_index(device):
    vol = device.structured_regular(np.zeros((8, 8, 8), np.float32))
    s = vol.sampler()
    h = vkl.lib().vklNewIntervalIteratorContext(s.h)
    ctx = vkl.IteratorContext(device, h)
    device.errors.clear()
    vkl.lib().vklSetInt(ctx.h, b"attributeIndex", 3)
    vkl.lib().vklCommit(ctx.h)
    c, m = device.last_error()
    device.errors.clear()
    assert c == 1 and "illegal attributeIndex" in m
    out = np.zeros(4, np.float32)
    pts = np.zeros((4, 3), np.float32)
    import ctypes
    vkl.lib().vklComputeSampleN(ctypes.byref(s.h), 4, pts.ctypes.data, out.ctypes.data, 7, None)
    c, m = device.last_error()
    device.errors.clear()
    assert c == 1 and "illegal attributeIndex" in m


def test_unknown_volume_type(device):
    h = vkl.lib().vklNewVolume(device.h, b"amr")
    assert not h.host
    c, m = device.last_error()
    device.errors.clear()
    assert c == 1 and "amr" in m
