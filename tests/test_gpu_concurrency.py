"""The reference's threading contract (multi_device.cpp:133-236; README
"sampling and iteration on committed objects are thread safe"): several host
threads sampling, differentiating and iterating one committed sampler at the
same time, with host and device pointers, get the results of serial calls;
iterating while another thread creates and releases LeafNodeAccess observers
on the sampler stays correct."""
import threading

import numpy as np
import pytest

import openvkl_amd as vkl
from tests import volumes as V
from tests.test_gpu_dense import assert_same
from tests.test_gpu_iterators import rays_through

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def run_threads(fns):
    errors = []

    def wrap(f):
        try:
            f()
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(e)

    ts = [threading.Thread(target=wrap, args=(f,)) for f in fns]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=300)
    assert not any(t.is_alive() for t in ts), "a thread did not finish"
    if errors:
        raise errors[0]


def test_concurrent_sampling_and_iteration(device):
    data = V.wavelet_volume(96)
    vol = device.structured_regular(data)
    s = vol.sampler()
    ctx = s.interval_context()
    pts_h = [V.uniform_points(50000 + 1000 * k, (-1, -1, -1), (96, 96, 96), seed=200 + k) for k in range(8)]
    ref_s = [s.sample(p) for p in pts_h]
    ref_g = [s.gradient(p[:5000]) for p in pts_h]
    org, d = rays_through(np.array([0, 0, 0, 95, 95, 95], np.float32), 24, seed=210)
    ref_i = [ctx.intervals(org[r], d[r]) for r in range(len(org))]
    results = {}

    def worker(k):
        def f():
            for rep in range(3):
                if k % 2 == 0:   # host pointers
                    got_s = s.sample(pts_h[k])
                    got_g = s.gradient(pts_h[k][:5000])
                else:            # device pointers
                    got_s = s.sample(torch.from_numpy(pts_h[k]).cuda()).cpu().numpy()
                    got_g = s.gradient(torch.from_numpy(pts_h[k][:5000]).cuda()).cpu().numpy()
                iv = [ctx.intervals(org[r], d[r]) for r in range(k, len(org), 8)]
                results[k, rep] = (got_s, got_g, iv)
        return f

    run_threads([worker(k) for k in range(8)])
    for (k, rep), (got_s, got_g, iv) in results.items():
        assert_same(got_s, ref_s[k])
        assert_same(got_g, ref_g[k])
        for j, r in enumerate(range(k, len(org), 8)):
            assert_same(iv[j], ref_i[r])
    assert len(results) == 24
    ctx.release()
    s.release()
    vol.release()


def test_iteration_while_observers_come_and_go(device):
    origins, leaves = V.sparse_shell_leaves(res=128)
    n = len(origins)
    vol = device.vdb(np.full(n, 3, np.uint32), origins, np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32),
                     packed_dense=[leaves.reshape(-1)])
    s = vol.sampler()
    ictx = s.interval_context()
    hctx = s.hit_context([0.5])
    org, d = rays_through(vol.bounding_box(), 32, seed=220)
    ref_i = [ictx.intervals(org[r], d[r]) for r in range(len(org))]
    ref_h = [hctx.hits(org[r], d[r]) for r in range(len(org))]
    stop = threading.Event()
    got = {}

    def observers():
        while not stop.is_set():
            o = s.observer("LeafNodeAccess")
            o.map()
            o.release()

    def iterate(k):
        def f():
            for r in range(k, len(org), 4):
                got[r] = (ictx.intervals(org[r], d[r]), hctx.hits(org[r], d[r]))
        return f

    obs = threading.Thread(target=observers)
    obs.start()
    try:
        run_threads([iterate(k) for k in range(4)])
    finally:
        stop.set()
        obs.join(timeout=60)
    for r in range(len(org)):
        assert_same(got[r][0], ref_i[r])
        assert_same(got[r][1], ref_h[r])
    hctx.release()
    ictx.release()
    s.release()
    vol.release()
