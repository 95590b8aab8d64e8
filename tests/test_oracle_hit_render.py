"""Pins of the oracle's hit-iterator renderer (examples/interactive/renderer/
hit_iterator_renderer/HitIteratorRenderer.ispc:51-150; oracle/vkl_oracle.cpp:
oracle_hit_render).

Closed forms on the ramp field f = x, whose gradient is (1, 0, 0) and whose
iso surfaces are the planes x = iso, hit by rays along -x and +x:

* a ray along +x sees each plane from behind the normal: both lights are on
  the other side, so each hit gets the ambient term 0.1 alone;
* a ray along -x faces the normal, so each light counts unless a shadow query
  towards +x meets another iso plane inside the box: the far plane is lit,
  0.1 + (1 + 0.8) / sqrt(3), and the near plane is shadowed by the far one
  for each light whose shadow ray reaches it before leaving the box;
* the colour is the sum of (1 - alpha) illum albedo, albedo = np.interp of
  the transfer function at the hit, and alpha = 1 - 0.4^hits;
* compositing stops at the sixth hit (1 - 0.4^6 >= 0.99 > 1 - 0.4^5);
* a nearest gradient filter gives no normal, hence illum 1;
* misses give zero."""
import math

import numpy as np
import pytest

import oracle as O

LIT = 0.1 + (1.0 + 0.8) / math.sqrt(3.0)


def ramp(n=32):
    x = np.arange(n, dtype=np.float32)
    return O.Grid.dense(np.broadcast_to(x, (n, n, n)).copy())


def tf_table():
    tf = np.zeros((4, 4), np.float32)
    tf[:, 0] = [0.1, 0.9, 0.3, 0.6]
    tf[:, 1] = [0.5, 0.2, 0.8, 0.4]
    tf[:, 2] = [0.7, 0.3, 0.1, 0.9]
    tf[:, 3] = 1.0
    return tf


def albedo(v, lo=0.0, hi=31.0):
    xs = np.linspace(lo, hi, 4)
    return np.array([np.interp(v, xs, tf_table()[:, k]) for k in range(3)])


def x_rays(count, sign, seed=1):
    rng = np.random.default_rng(seed)
    org = np.empty((count, 3), np.float32)
    org[:, 0] = -3.0 if sign > 0 else 34.0
    org[:, 1:] = 2 + rng.random((count, 2)) * 27
    d = np.zeros((count, 3), np.float32)
    d[:, 0] = sign
    return org, d


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_two_planes_shading_and_shadows(sign):
    g = ramp()
    iso = [7.25, 20.5]
    org, d = x_rays(50, sign)
    out, cnt = g.hit_render(org, d, iso, tf_table(), (0.0, 31.0))
    assert np.all(cnt == 2)
    first, second = (iso[0], iso[1]) if sign > 0 else (iso[1], iso[0])
    if sign > 0:
        il1 = il2 = np.full(50, 0.1)
    else:
        # the shadow ray from the near plane (x = 7.25) towards light (1, 1, +-1)
        # meets the far plane (x = 20.5) only if it stays in the box [0, 31]
        # in y and z over dx = 13.25
        dx = iso[1] - iso[0]
        y, z = org[:, 1].astype(np.float64), org[:, 2].astype(np.float64)
        blocked0 = (y + dx <= 31) & (z + dx <= 31)
        blocked1 = (y + dx <= 31) & (z - dx >= 0)
        il1 = np.full(50, LIT)
        il2 = 0.1 + np.where(blocked0, 0, 1) / math.sqrt(3) + np.where(blocked1, 0, 0.8) / math.sqrt(3)
        assert blocked0.any() and (~blocked0).any()
    want = il1[:, None] * albedo(first)[None] + 0.4 * il2[:, None] * albedo(second)[None]
    np.testing.assert_allclose(out[:, :3], want, rtol=2e-4, atol=1e-5)
    np.testing.assert_allclose(out[:, 3], 0.84, rtol=1e-6)


def test_stops_at_the_sixth_hit():
    g = ramp()
    iso = [2.5 + 3 * k for k in range(9)]
    org, d = x_rays(20, 1.0, seed=2)
    out, cnt = g.hit_render(org, d, iso, tf_table(), (0.0, 31.0))
    assert np.all(cnt == 6)
    np.testing.assert_allclose(out[:, 3], 1 - 0.4 ** 6, rtol=1e-6)
    out, cnt = g.hit_render(org, d, iso, tf_table(), (0.0, 31.0), max_hits=2)
    assert np.all(cnt == 2)


def test_no_normal_is_unshaded_and_misses_are_zero():
    g = ramp()
    g.set_sampler(O.LINEAR, gradient_filter=O.NEAREST)
    org, d = x_rays(20, -1.0, seed=3)
    out, cnt = g.hit_render(org, d, [11.0], tf_table(), (0.0, 31.0))
    assert np.all(cnt == 1)
    np.testing.assert_allclose(out[:, :3], np.tile(albedo(11.0), (20, 1)), rtol=2e-4, atol=1e-5)
    g.set_sampler(O.LINEAR)
    out, cnt = g.hit_render(org, -d, [11.0], tf_table(), (0.0, 31.0))
    assert np.all(out == 0) and np.all(cnt == 0)
