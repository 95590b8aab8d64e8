"""Pins of the oracle's observers (oracle/vkl_oracle.cpp: oracle_inner_nodes,
leaf-access marking in the sparse traversal).

InnerNode: the set of reported boxes follows from the input nodes alone (a
leaf pointer on a level at or above maxDepth reports the node's own box, a
deeper node reports the level-maxDepth voxel that contains it), and an inner
voxel's value range is the union of the ranges of the leaf pointers below it.
LeafNodeAccess: the marked entries are exactly the input nodes that contain a
filter-stencil voxel inside the domain whose node level is within
maxSamplingDepth, found by brute force over the stencil voxels; packed input
marks the per-format packed index instead of the input index."""
import warnings

import numpy as np
import pytest

import oracle as O
from tests import volumes as V

RES = {1: 4096, 2: 128, 3: 8}


def mixed_node_arrays(seed, nattr=1):
    """Dense 8^3 leaves in [0, 256)^3, level-2 tiles at x >= 256, one
    level-1 tile at x = 4096, in shuffled input order. Returns levels,
    origins, formats, per-attribute packed dense / tile arrays and per-node
    data (list over nodes of lists over attributes)."""
    rng = np.random.default_rng(seed)
    cells = rng.choice(32 ** 3, 300, replace=False)
    leaf_org = np.stack([cells // 1024, (cells // 32) % 32, cells % 32], 1) * 8
    tile_org = np.array([[256, 0, 0], [384, 128, 0], [256, 256, 128], [512, 0, 256]])
    levels = np.concatenate([np.full(len(leaf_org), 3), np.full(len(tile_org), 2), [1]]).astype(np.uint32)
    origins = np.concatenate([leaf_org, tile_org, [[4096, 0, 0]]]).astype(np.int32)
    fmt = np.where(levels == 3, O.FORMAT_DENSE_ZYX, O.FORMAT_TILE).astype(np.uint32)
    # interleave tiles among leaves so input and packed indices differ
    order = rng.permutation(len(levels))
    levels, origins, fmt = levels[order], origins[order], fmt[order]
    nd, nt = int((fmt == 1).sum()), int((fmt == 0).sum())
    dense = [(rng.standard_normal(nd * 512) * 3 + a).astype(np.float32) for a in range(nattr)]
    tile = [(rng.standard_normal(nt) * 3 - a).astype(np.float32) for a in range(nattr)]
    per_node = []
    di = ti = 0
    for f in fmt:
        if f == O.FORMAT_DENSE_ZYX:
            per_node.append([d[di * 512:(di + 1) * 512] for d in dense])
            di += 1
        else:
            per_node.append([t[ti:ti + 1] for t in tile])
            ti += 1
    return levels, origins, fmt, dense, tile, per_node


def mixed_nodes(seed, packed, nattr=1):
    levels, origins, fmt, dense, tile, per_node = mixed_node_arrays(seed, nattr)
    if packed:
        g = O.Grid.vdb(levels, origins, fmt, packed_dense=dense, packed_tile=tile)
    else:
        g = O.Grid.vdb(levels, origins, fmt, leaf_data=per_node)
    return g, levels, origins, fmt


def packed_index(fmt):
    out = np.empty(len(fmt), np.int64)
    nd = nt = 0
    for i, f in enumerate(fmt):
        if f == O.FORMAT_DENSE_ZYX:
            out[i] = nd
            nd += 1
        else:
            out[i] = nt
            nt += 1
    return out


def expected_boxes(levels, origins, max_depth):
    boxes = set()
    for L, o in zip(levels, origins):
        if L - 1 <= max_depth:
            size = RES[int(L)]
            lo = o
        else:
            size = RES[max_depth + 1]
            lo = (o // size) * size
        boxes.add(tuple(int(v) for v in lo) + tuple(int(v) + size for v in lo))
    return boxes


@pytest.mark.parametrize("max_depth", [0, 1, 2, 7])
def test_inner_node_boxes_follow_from_input_nodes(max_depth):
    g, levels, origins, _ = mixed_nodes(5, packed=False)
    nodes = g.inner_nodes(max_depth)
    assert nodes.shape[1] == 8
    got = {tuple(int(v) for v in row[:6]) for row in nodes}
    assert len(got) == nodes.shape[0]  # no box twice
    assert got == expected_boxes(levels, origins, min(max_depth, 2))


def test_inner_node_boxes_through_transform():
    """An exactly representable index-to-object transform maps every box."""
    lv = np.array([3, 3, 2], np.uint32)
    org = np.array([[0, 0, 0], [136, 8, 16], [256, 0, 0]], np.int32)
    fm = np.array([1, 1, 0], np.uint32)
    i2o = np.array([0.5, 0, 0, 0, 2.0, 0, 0, 0, 0.25, 4.0, -8.0, 1.0], np.float32)  # 3x3 rows, then t
    dense = [np.arange(512, dtype=np.float32), -np.arange(512, dtype=np.float32)]
    g = O.Grid.vdb(lv, org, fm, packed_dense=[np.concatenate(dense)], packed_tile=[np.array([7.0], np.float32)],
                   i2o=i2o)
    nodes = g.inner_nodes(2)
    want = set()
    for lo in expected_boxes(lv, org, 2):
        a, b = np.array(lo[:3], np.float64), np.array(lo[3:], np.float64)
        sc, t = np.array([0.5, 2.0, 0.25]), np.array([4.0, -8.0, 1.0])
        want.add(tuple(np.concatenate([a * sc + t, b * sc + t]).tolist()))
    assert {tuple(float(v) for v in r[:6]) for r in nodes} == want


def test_inner_node_ranges_are_unions_of_leaf_ranges():
    g, levels, origins, _ = mixed_nodes(6, packed=True)
    leaves = g.inner_nodes(2)
    for d in (0, 1):
        for row in g.inner_nodes(d):
            lo, hi = row[:3], row[3:6]
            inside = np.all((leaves[:, :3] >= lo) & (leaves[:, 3:6] <= hi), axis=1)
            assert inside.any()
            assert row[6] == leaves[inside, 6].min() and row[7] == leaves[inside, 7].max()
    # a leaf's range covers its own values (tiles: exactly one value inside)
    for row in leaves:
        assert row[6] <= row[7]


def stencil(points, filt):
    p = np.asarray(points, np.float32)
    if filt == O.NEAREST:
        return np.floor(p).astype(np.int64)[:, None, :]
    ic = np.floor(p).astype(np.int64)
    r = (0, 1) if filt == O.LINEAR else (-1, 0, 1, 2)
    offs = np.array([(a, b, c) for a in r for b in r for c in r])
    return ic[:, None, :] + offs[None, :, :]


def expected_marks(g, levels, origins, fmt, points, filt, md, packed):
    info = g.info()
    root, size = np.asarray(info["root_origin"]), np.asarray(info["active_size"])
    vox = stencil(points, filt).reshape(-1, 3)
    off = vox - root
    vox = vox[np.all((off >= 0) & (off < size), axis=1)]
    key = {}
    for i, (L, o) in enumerate(zip(levels, origins)):
        key[(int(L),) + tuple(int(v) // RES[int(L)] for v in o)] = i
    idx = packed_index(fmt) if packed else np.arange(len(levels))
    want = np.zeros(len(levels), np.uint32)
    for L in (1, 2, 3):
        if L > md:
            continue
        for c in {tuple(v) for v in (vox // RES[L]).tolist()}:
            i = key.get((L,) + c)
            if i is not None:
                want[idx[i]] = 1
    return want


@pytest.mark.parametrize("packed", [False, True])
@pytest.mark.parametrize("filt", [O.NEAREST, O.LINEAR, O.CUBIC])
def test_leaf_access_marks_stencil_leaves(packed, filt):
    g, levels, origins, fmt = mixed_nodes(7, packed)
    rng = np.random.default_rng(8)
    near = origins[rng.choice(len(origins), 150)] + rng.random((150, 3)) * 10 - 1  # around leaves and tiles
    pts = np.concatenate([near, rng.random((300, 3)) * 260 - 2,  # leaves and the edge of the domain
                          rng.random((40, 3)) * [700, 400, 400],
                          [[4100.5, 3.25, 7.75], [-5, -5, -5]]]).astype(np.float32)
    g.set_sampler(filt)
    acc = g.observe_leaf_access()
    g.sample(pts)
    want = expected_marks(g, levels, origins, fmt, pts, filt, 3, packed)
    np.testing.assert_array_equal(acc, want)
    assert want.sum() > 10 and want.sum() < len(want)
    # gradients traverse the same stencil
    if filt != O.NEAREST:
        acc = g.observe_leaf_access()
        g.gradient(pts)
        np.testing.assert_array_equal(acc, want)


@pytest.mark.parametrize("md", [0, 1, 2])
def test_leaf_access_respects_max_sampling_depth(md):
    """Traversals stopped above a node's level end on an error voxel, which
    marks nothing (VdbSampleInner.ih.in:46-54)."""
    g, levels, origins, fmt = mixed_nodes(9, packed=False)
    rng = np.random.default_rng(10)
    near = origins[rng.choice(len(origins), 150)] + rng.random((150, 3)) * 10 - 1
    pts = np.concatenate([near, rng.random((300, 3)) * 256, rng.random((60, 3)) * [700, 400, 400],
                          [[4100.5, 3.25, 7.75]]]).astype(np.float32)
    g.set_sampler(O.LINEAR, max_sampling_depth=md)
    acc = g.observe_leaf_access()
    g.sample(pts)
    want = expected_marks(g, levels, origins, fmt, pts, O.LINEAR, md, False)
    np.testing.assert_array_equal(acc, want)
    if md == 0:
        assert want.sum() == 0


def test_leaf_access_off_marks_nothing():
    g, *_ = mixed_nodes(11, packed=False)
    acc = g.observe_leaf_access()
    g.observe_leaf_access(False)
    g.sample(np.random.default_rng(12).random((100, 3)).astype(np.float32) * 256)
    assert acc.sum() == 0


# ---- the reference's own known answers -----------------------------------------------

KA_CASES = [((0.0, 0.0, 0.0), (1.0, 1.0, 1.0)), ((100.0, 200.0, 300.0), (1.0, 2.0, 3.0))]
KA_FIELDS = [V.xyz, V.wavelet, V.zramp]


def known_answer_volume(origin, spacing, nattr, dims=256):
    """testing/apps/tests/vdb_volume_inner_node_observer.cpp:138-168: a vdb
    volume filled with the xyz field (and further fields for the
    multi-attribute case) on a dims^3 grid with the given origin / spacing."""
    origins, leaves = V.dense_vdb_leaves(dims, KA_FIELDS[:nattr], origin, spacing)
    n = len(origins)
    i2o = np.array([spacing[0], 0, 0, 0, spacing[1], 0, 0, 0, spacing[2], *origin], np.float32)
    return (np.full(n, 3, np.uint32), origins, np.full(n, O.FORMAT_DENSE_ZYX, np.uint32),
            [a.reshape(-1) for a in leaves], i2o)


def check_inner_nodes_known_answer(inner_nodes, sample_nearest, bbox, dims, nattr, max_points=None):
    """vdb_volume_inner_node_observer.cpp:13-134: for maxDepth 0..2 the node
    count is ceil(dims / node resolution)^3, values sampled (nearest) on a
    step-2 grid over each node (step 16 at depth 0) lie inside the node's
    range for every attribute (NaN background samples are ignored, as
    range1f::extend does), and the union of the node boxes contains the
    volume's bounding box (equals it at depth 2)."""
    for max_depth in range(3):
        res = {0: 4096, 1: 128, 2: 8}[max_depth]
        nodes = inner_nodes(max_depth)
        assert nodes.shape == ((-(-dims // res)) ** 3, 6 + 2 * nattr)
        step = 16 if max_depth == 0 else 2
        k = np.arange(res // step, dtype=np.float32) * step / np.float32(res)
        gx, gy, gz = np.meshgrid(k, k, k, indexing="ij")
        frac = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1)
        if max_points and len(frac) > max_points:
            frac = frac[np.random.default_rng(max_depth).choice(len(frac), max_points, replace=False)]
        lo, hi = nodes[:, None, :3], nodes[:, None, 3:6]
        pts = (lo + frac[None] * (hi - lo)).astype(np.float32)
        for a in range(nattr):
            s = sample_nearest(pts.reshape(-1, 3), a).reshape(len(nodes), -1)
            with warnings.catch_warnings():  # all-NaN rows: nodes outside the domain
                warnings.simplefilter("ignore", RuntimeWarning)
                smin, smax = np.nanmin(s, axis=1), np.nanmax(s, axis=1)
            ok = np.isnan(smin) | ((smin >= nodes[:, 6 + 2 * a]) & (smax <= nodes[:, 7 + 2 * a]))
            assert ok.all(), (max_depth, a)
        ext = np.concatenate([nodes[:, :3].min(0), nodes[:, 3:6].max(0)])
        if max_depth == 2:
            np.testing.assert_array_equal(ext, bbox)
        else:
            assert np.all(bbox[:3] >= ext[:3]) and np.all(bbox[3:] <= ext[3:])


@pytest.mark.parametrize("origin,spacing", KA_CASES)
@pytest.mark.parametrize("nattr", [1, 3])
def test_inner_nodes_reference_known_answer(origin, spacing, nattr):
    level, origins, fmt, dense, i2o = known_answer_volume(origin, spacing, nattr, dims=128)
    g = O.Grid.vdb(level, origins, fmt, packed_dense=dense, i2o=i2o)
    g.set_sampler(O.NEAREST)
    check_inner_nodes_known_answer(g.inner_nodes, lambda p, a: g.sample(p, attr=a, threads=8),
                                   g.bounding_box(), 128, nattr, max_points=4096)
