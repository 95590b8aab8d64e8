"""Pins of the oracle's interval ray marching with transfer-function
compositing, the ray-march iterator renderer (examples/interactive/renderer/
ray_march_iterator/RayMarchIterator.ispc:39-137; oracle/vkl_oracle.cpp:
oracle_ray_march_intervals).

Closed forms on a constant volume with a constant transfer function
(r, g, b, a): the transmittance 1 - alpha is the product of (1 - a dt) over the
sub-intervals, so -ln(1 - alpha) lies between a L and a L (1 + a dt), L the
ray's length in the volume's box (slab test); the colour is (r, g, b) alpha; the sub-intervals
of nominalDeltaT / samplingRate make ceil(length / step) samples per interval
of the (separately pinned) interval iterator; an opaque a dt >= 1 stops after
one sample with the transfer function's colour, which np.interp gives per
channel; rays that miss the box, or value ranges that exclude the volume,
give zero."""
import math

import numpy as np
import pytest

import oracle as O


def constant(value=0.5, n=16):
    return O.Grid.dense(np.full((n, n, n), value, np.float32))


def rays(count, seed, n=16):
    rng = np.random.default_rng(seed)
    org = (rng.random((count, 3)) * n - np.array([0, 0, 2.0 * n])).astype(np.float32)
    org[:, 2] = -5.0
    tgt = rng.random((count, 3)) * (n - 1)
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return org.astype(np.float32), d.astype(np.float32)


def box_length(org, d, lo, hi):
    o, dd = org.astype(np.float64), d.astype(np.float64)
    with np.errstate(divide="ignore"):
        inv = 1.0 / dd
    t0 = np.minimum((lo - o) * inv, (hi - o) * inv).max(1)
    t1 = np.maximum((lo - o) * inv, (hi - o) * inv).min(1)
    return np.maximum(t1 - np.maximum(t0, 0.0), 0.0)


def flat_tf(rgba, n=3):
    return np.tile(np.asarray(rgba, np.float32), (n, 1))


def test_transmittance_colour_and_sample_count():
    g = constant()
    org, d = rays(300, 1)
    a, rgb, rate = 0.02, (0.2, 0.5, 0.9), 2.0
    out, cnt = g.ray_march_intervals(org, d, flat_tf(rgb + (a,)), (0.0, 1.0), sampling_rate=rate)
    box = g.bounding_box().astype(np.float64)
    L = box_length(org, d, box[:3], box[3:])
    # -ln T = sum -ln(1 - a dt_i), and x <= -ln(1 - x) <= x + x^2 for x <= 1/2,
    # so 0 <= -ln T - a L <= a L (a dt_max), dt_max = nominalDeltaT / rate.
    ndt = np.array([g.intervals(org[r], d[r])[0, 4] for r in range(300)], np.float64)
    excess = -np.log(1.0 - out[:, 3].astype(np.float64)) - a * L
    assert excess.min() >= -1e-5
    assert np.all(excess <= a * L * (a * ndt / rate) * 1.01 + 1e-5)
    np.testing.assert_allclose(out[:, :3], np.outer(out[:, 3], rgb), rtol=1e-5, atol=1e-7)
    for r in range(0, 300, 13):
        iv = g.intervals(org[r], d[r])
        step = iv[:, 4].astype(np.float64) / rate
        want = int(np.ceil((iv[:, 1].astype(np.float64) - iv[:, 0]) / step - 1e-6).sum())
        assert abs(int(cnt[r]) - want) <= len(iv), (r, cnt[r], want)  # per interval, the last step may round
        assert cnt[r] > 0


def test_opaque_stops_after_one_sample_with_tf_colour():
    rng = np.random.default_rng(2)
    tf = rng.random((6, 4)).astype(np.float32)
    tf[:, 3] = 50.0  # a dt >= 1: opaque at the first sample
    lo, hi = -1.0, 2.0
    for value in (-2.0, 0.1, 0.77, 1.5, 3.0):
        g = constant(value)
        org, d = rays(40, 3)
        out, cnt = g.ray_march_intervals(org, d, tf, (lo, hi))
        assert np.all(cnt == 1) and np.all(out[:, 3] == 1.0)
        xs = np.linspace(lo, hi, 6)
        want = [np.interp(value, xs, tf[:, k]) for k in range(3)]
        np.testing.assert_allclose(out[:, :3], np.tile(want, (40, 1)), rtol=1e-6, atol=1e-7)


def test_early_termination_at_099():
    g = constant()
    org, d = rays(200, 4)
    out, cnt = g.ray_march_intervals(org, d, flat_tf((1, 1, 1, 0.4)), (0.0, 1.0), sampling_rate=4.0)
    box = g.bounding_box().astype(np.float64)
    L = box_length(org, d, box[:3], box[3:])
    long_rays = np.exp(-0.4 * L) < 0.005
    assert long_rays.any()
    assert np.all(out[long_rays, 3] >= 0.99)
    # it stops at the first sample that reaches 0.99: one sample fewer stays below
    for r in np.nonzero(long_rays)[0][:20]:
        o1, c1 = g.ray_march_intervals(org[r:r + 1], d[r:r + 1], flat_tf((1, 1, 1, 0.4)), (0.0, 1.0),
                                       sampling_rate=4.0, max_samples=int(cnt[r]) - 1)
        assert c1[0] == cnt[r] - 1 and o1[0, 3] < 0.99


def test_misses_and_culled_ranges_give_zero():
    g = constant(0.5)
    org, d = rays(50, 5)
    away = -d
    out, cnt = g.ray_march_intervals(org, away, flat_tf((1, 1, 1, 0.5)), (0.0, 1.0))
    assert np.all(out == 0) and np.all(cnt == 0)
    out, cnt = g.ray_march_intervals(org, d, flat_tf((1, 1, 1, 0.5)), (0.0, 1.0), value_ranges=[[2.0, 3.0]])
    assert np.all(out == 0) and np.all(cnt == 0)
    out, cnt = g.ray_march_intervals(org, d, flat_tf((1, 1, 1, 0.5)), (0.0, 1.0), max_samples=3)
    assert np.all(cnt <= 3)
