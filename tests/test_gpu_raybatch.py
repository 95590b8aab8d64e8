"""Ray batch plans (vklNewRayBatchPlan): interval and hit batches over
different volumes and contexts, replayed as one graph, give exactly the
per-context batch calls' results (themselves bit-identical to the oracle),
and a subset is checked against the oracle directly."""
import numpy as np
import pytest

import oracle as O
import openvkl_amd as vkl
from tests import volumes as V
from tests.test_gpu_iterators import rays_through

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))


@pytest.fixture(scope="module")
def volumes(device):
    origins, leaves = V.sparse_shell_leaves(res=128)
    n = len(origins)
    lv, fm = np.full(n, 3, np.uint32), np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32)
    vdb = device.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    vdb_grid = O.Grid.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    data = V.procedural((40, 36, 44), V.wavelet)
    dense = device.structured_regular(data)
    return vdb, vdb_grid, dense, O.Grid.dense(data)


def test_ray_batch_plan_matches_batch_calls_and_oracle(device, volumes):
    vdb, vdb_grid, dense, dense_grid = volumes
    sv, sd = vdb.sampler(), dense.sampler()
    contexts = [sv.interval_context(), sv.interval_context(hint=1.0), sv.hit_context([0.5]),
                sd.interval_context(), sd.hit_context([0.3, 1.1]), sv.hit_context([0.25, 0.75])]
    counts = [700, 300, 500, 0, 257, 64]
    caps = [256, 512, 16, 64, 32, 8]
    batches, rays = [], []
    for k, (c, n, cap) in enumerate(zip(contexts, counts, caps)):
        box = (vdb if k in (0, 1, 2, 5) else dense).bounding_box()
        org, d = rays_through(np.asarray(box, np.float32), max(n, 1), seed=400 + k)
        org, d = org[:n], d[:n]
        rays.append((org, d))
        batches.append((c, torch.from_numpy(np.ascontiguousarray(org)).cuda(),
                        torch.from_numpy(np.ascontiguousarray(d)).cuda(), cap))
    plan = vkl.RayBatchPlan(device, batches)
    assert plan.launch_count() == sum(1 for n in counts if n)
    want = [c.batch(b[1], b[2], max_per_ray=b[3]) if b[1].shape[0] else None for c, b in zip(contexts, batches)]
    for rep in range(2):
        for r, cn in zip(plan.results, plan.counts):
            r.fill_(-1.0)
            cn.fill_(-1)
        torch.cuda.synchronize()
        plan.execute()
        torch.cuda.synchronize()
        for k, w in enumerate(want):
            if w is None:
                continue
            got_cnt, want_cnt = plan.counts[k].cpu().numpy(), w[1].cpu().numpy()
            assert np.array_equal(got_cnt, want_cnt), (rep, k)
            gr, wr = plan.results[k].cpu().numpy(), w[0].cpu().numpy()
            live = np.arange(gr.shape[1])[None, :] < got_cnt[:, None]
            assert same(gr[live], wr[live]), (rep, k)
    # the oracle on a subset of two batches
    org, d = rays[0]
    res, cnt = plan.results[0].cpu().numpy(), plan.counts[0].cpu().numpy()
    for r in range(0, len(org), 23):
        iv = vdb_grid.intervals(org[r], d[r], depth=O.map_depth(0.5), max_out=256)
        assert same(res[r, :cnt[r]], iv), r
    org, d = rays[4]
    res, cnt = plan.results[4].cpu().numpy(), plan.counts[4].cpu().numpy()
    for r in range(0, len(org), 7):
        h = dense_grid.hits(org[r], d[r], [0.3, 1.1], max_out=32)
        assert same(res[r, :cnt[r]], h), r
    plan.release()
    for c in contexts:
        c.release()


def test_ray_batch_plan_more_batches_than_branches(device, volumes):
    """20 batches over 16 graph branches (two batches share a branch in
    order), alternating an interval and a hit context of one volume."""
    vdb = volumes[0]
    s = vdb.sampler()
    ictx, hctx = s.interval_context(), s.hit_context([0.5])
    box = np.asarray(vdb.bounding_box(), np.float32)
    batches = []
    for k in range(20):
        org, d = rays_through(box, 40 + 7 * k, seed=600 + k)
        c = ictx if k % 2 == 0 else hctx
        batches.append((c, torch.from_numpy(org).cuda(), torch.from_numpy(d).cuda(), 64 if k % 2 == 0 else 8))
    plan = vkl.RayBatchPlan(device, batches)
    assert plan.launch_count() == 20
    plan.execute()
    torch.cuda.synchronize()
    for k, b in enumerate(batches):
        w_out, w_cnt = b[0].batch(b[1], b[2], max_per_ray=b[3])
        got_cnt = plan.counts[k].cpu().numpy()
        assert np.array_equal(got_cnt, w_cnt.cpu().numpy()), k
        gr, wr = plan.results[k].cpu().numpy(), w_out.cpu().numpy()
        live = np.arange(gr.shape[1])[None, :] < got_cnt[:, None]
        assert same(gr[live], wr[live]), k
    plan.release()
    ictx.release()
    hctx.release()
