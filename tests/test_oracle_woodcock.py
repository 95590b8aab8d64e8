"""Pins of the oracle's Woodcock (delta) tracking, the free-flight sampler of
the reference density path tracer (examples/interactive/renderer/
density_path_tracer/DensityPathTracer.ispc:61-110; oracle/vkl_oracle.cpp:
oracle_woodcock, oracle_tea, oracle_flight_log).

Known answers: the published 32-round TEA vector (key 0, block 0 ->
41ea3a0a 94baa940) pins the cipher round the renderer's 8-round generator
uses. Library routines: the logarithm against numpy's to 1 ulp; the transfer
function's opacity against np.interp. Closed forms: with opacity 1 every ray
stops at its first tentative collision, t0 - ln(1 - u0) / sigmaMax with u0
from the first TEA draw; with constant opacity a the escape probability over a
length L is exp(-sigmaMax a L) and event distances follow the truncated
exponential (Kolmogorov-Smirnov); with opacity 0 the number of tentative
collisions is Poisson(sigmaMax L); the motion-blur time of a collision is
clamp(time + (u2 - 0.5) shutter, 0, 1) with u2 from the second TEA draw."""
import math

import numpy as np
import pytest

import oracle as O

KEY = (0xa341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)
TO_FLOAT = np.float32(2.0 ** -32)


def test_tea_published_vector():
    assert O.tea((0, 0), (0, 0, 0, 0), 32) == (0x41EA3A0A, 0x94BAA940)


def test_flight_log_against_library_log():
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.random(20000, dtype=np.float32),
                         np.float32(2.0) ** -np.arange(0, 149, dtype=np.float32),
                         np.nextafter(np.float32(1), np.float32(0)) - np.arange(64, dtype=np.float32) * 2 ** -24,
                         np.array([0.70710677, 0.70710683, 1.4142135, 1e-40, 3e-45], np.float32)])
    xs = xs[xs > 0]
    got = np.array([O.flight_log(x) for x in xs], np.float32)
    want = np.log(xs.astype(np.float64))
    ulp = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
    err = np.abs(got.astype(np.float64) - want) / np.maximum(ulp, np.spacing(np.float32(0)))
    assert err.max() <= 1.0, (xs[err.argmax()], got[err.argmax()], want[err.argmax()])
    assert O.flight_log(1.0) == 0.0
    assert O.flight_log(0.0) == -np.inf
    assert np.isnan(O.flight_log(-1.0)) and np.isnan(O.flight_log(float("nan")))


def constant_grid(value=1.0, n=8):
    return O.Grid.dense(np.full((n, n, n), value, np.float32))


def x_rays(count, y=3.5, z=3.5, t1=7.0):
    org = np.tile(np.array([[0.0, y, z]], np.float32), (count, 1))
    d = np.tile(np.array([[1.0, 0.0, 0.0]], np.float32), (count, 1))
    tr = np.tile(np.array([[0.0, t1]], np.float32), (count, 1))
    return org, d, tr


def seeds(count, frame=7):
    """RandomTEA__Constructor(frame, pixel) (Random.ih:32-38)."""
    s = np.empty((count, 2), np.uint32)
    s[:, 0] = frame
    s[:, 1] = np.arange(count, dtype=np.uint32)
    return s


def flat_tf(opacity, n=2):
    tf = np.zeros((n, 4), np.float32)
    tf[:, :3] = 0.5
    tf[:, 3] = opacity
    return tf


def test_opacity_one_stops_at_first_draw():
    g = constant_grid()
    n = 3000
    org, d, tr = x_rays(n)
    st = seeds(n)
    st0 = st.copy()
    sigma = np.float32(0.8)
    ev, sm, cnt = g.woodcock(org, d, tr, st, sigma, flat_tf(1.0), (0.0, 2.0))
    for r in range(0, n, 7):
        a, b = O.tea(st0[r], KEY, 8)
        u0 = np.float32(np.float32(a) * TO_FLOAT)
        t = -math.log1p(-float(u0)) / float(sigma)
        if t > 7.0:
            assert ev[r] == np.inf and cnt[r] == 0
        else:
            assert ev[r] == pytest.approx(t, rel=3e-6, abs=1e-6)
            assert cnt[r] == 1 and sm[r] == 1.0
        assert tuple(st[r]) == O.tea((a, b), KEY, 8)  # two draws per collision


@pytest.mark.parametrize("value", [0.3, 1.7])
def test_escape_fraction_and_event_distribution(value):
    """Constant volume, 5-entry transfer function: opacity a = interp(value);
    P(escape) = exp(-sigma a L); event t | event ~ truncated Exp(sigma a)."""
    tf = np.zeros((5, 4), np.float32)
    tf[:, 3] = [0.1, 0.6, 0.25, 0.9, 0.4]
    lo, hi = -0.5, 2.5
    a = float(np.interp(value, np.linspace(lo, hi, 5), tf[:, 3]))
    g = constant_grid(value)
    n, sigma, L = 40000, 0.6, 7.0
    org, d, tr = x_rays(n, t1=L)
    ev, sm, cnt = g.woodcock(org, d, tr, seeds(n, frame=3), sigma, tf, (lo, hi))
    p = math.exp(-sigma * a * L)
    esc = np.isinf(ev).mean()
    assert abs(esc - p) < 4.5 * math.sqrt(p * (1 - p) / n)
    hit = ev[np.isfinite(ev)].astype(np.float64)
    np.testing.assert_allclose(sm[np.isfinite(ev)], value, rtol=1e-6)  # trilinear of a constant
    hit.sort()
    lam = sigma * a
    cdf = (1 - np.exp(-lam * hit)) / (1 - math.exp(-lam * L))
    emp = np.arange(1, len(hit) + 1) / len(hit)
    ks = max(np.abs(emp - cdf).max(), np.abs(emp - 1 / len(hit) - cdf).max())
    assert ks < 1.63 / math.sqrt(len(hit))  # KS at the 1 % level


def test_zero_opacity_counts_are_poisson():
    g = constant_grid()
    n, sigma, L = 20000, 1.5, 7.0
    org, d, tr = x_rays(n, t1=L)
    ev, sm, cnt = g.woodcock(org, d, tr, seeds(n), sigma, flat_tf(0.0), (0.0, 2.0))
    assert np.all(np.isinf(ev)) and np.all(sm == 0)
    lam = sigma * L
    assert abs(cnt.mean() - lam) < 4.5 * math.sqrt(lam / n)
    assert abs(cnt.var() - lam) < 0.1 * lam


def test_no_majorant_and_collision_cap():
    g = constant_grid()
    n = 64
    org, d, tr = x_rays(n)
    st = seeds(n)
    ev, sm, cnt = g.woodcock(org, d, tr, st, 0.0, flat_tf(1.0), (0.0, 2.0))
    assert np.all(ev == np.inf) and np.all(cnt == 0) and np.array_equal(st, seeds(n))
    tr[:, 1] = 1e30
    ev, sm, cnt = g.woodcock(org, d, tr, st, 2.0, flat_tf(0.0), (0.0, 2.0), max_collisions=3)
    assert np.all(np.isnan(ev)) and np.all(cnt == 3) and np.all(sm == 0)


@pytest.mark.parametrize("time,shutter", [(0.5, 0.4), (0.1, 0.6), (0.95, 0.3), (0.3, 0.0)])
def test_motion_blur_time_of_each_collision(time, shutter):
    """Two structured time steps holding 0 and 1: a nearest sample equals the
    collision's time; opacity 1 accepts the first collision."""
    n = 6
    data = np.zeros(n ** 3 * 2, np.float32)
    data[1::2] = 1.0
    g = O.Grid.dense_temporal((n, n, n), data, structured_steps=2)
    g.set_sampler(O.NEAREST)
    r = 2000
    org, d, tr = x_rays(r, y=2.5, z=2.5, t1=5.0)
    st = seeds(r, frame=11)
    st0 = st.copy()
    ev, sm, cnt = g.woodcock(org, d, tr, st, 1e4, flat_tf(1.0), (0.0, 1.0), time=time, shutter=shutter)
    assert np.all(cnt == 1)
    for k in range(0, r, 5):
        s1 = O.tea(st0[k], KEY, 8)
        s2 = O.tea(s1, KEY, 8)
        u2 = np.float32(np.float32(s2[0]) * TO_FLOAT)
        tm = np.float32(time)
        if shutter != 0:
            tm = np.float32(tm + np.float32(np.float32(u2 - np.float32(0.5)) * np.float32(shutter)))
        tm = min(max(tm, np.float32(0)), np.float32(1))
        assert sm[k] == tm, (k, sm[k], tm)
