"""Pins the travel-time oracle (oracle/traveltime.py) against closed forms,
hand-computed samples and finite differences.  CPU only."""
import math

import numpy as np

from oracle import traveltime as ott

N = (12, 10, 14)
O = (100.0, -50.0, 0.0)
H = (10.0, 10.0, 10.0)


def _picks(rng, k, margin=0.0):
    lo = np.array(O) - margin
    hi = np.array(O) + np.array(N) * np.array(H) + margin
    return rng.uniform(lo, hi, (k, 3)), rng.uniform(lo, hi, (k, 3))


def test_homogeneous_model_gives_distance_over_velocity():
    rng = np.random.default_rng(1)
    src, sta = _picks(rng, 50)
    vp = np.full(int(np.prod(N)), 2500.0)
    t = ott.forward(src, sta, N, O, H, vp)
    d = np.linalg.norm(src - sta, axis=1)
    np.testing.assert_allclose(t, d / 2500.0, rtol=1e-14)


def test_axis_aligned_ray_samples_by_hand():
    # source in cell (3, 4, 0), station in cell (3, 4, 9): dist = 90 m -> n_s = max(int(90/50), 2) = 2,
    # samples at frac 0.25 and 0.75 -> iz = int(2.25) = 2 and int(6.75) = 6, ds = 45 m
    src = np.array([[O[0] + 35.0, O[1] + 45.0, O[2] + 5.0]])
    sta = np.array([[O[0] + 35.0, O[1] + 45.0, O[2] + 95.0]])
    vp = np.full(int(np.prod(N)), 2000.0)
    idx = lambda x, y, z: (x * N[1] + y) * N[2] + z
    vp[idx(3, 4, 2)] = 1500.0
    vp[idx(3, 4, 6)] = 3000.0
    t = ott.forward(src, sta, N, O, H, vp)
    assert t[0] == 45.0 / 1500.0 + 45.0 / 3000.0
    g = ott.gradient(src, sta, N, O, H, vp, np.array([2.0]), np.array([0.5]))
    expect = np.zeros_like(vp)
    expect[idx(3, 4, 2)] = 2.0 * (-45.0 / (1500.0 * 1500.0)) * 0.5
    expect[idx(3, 4, 6)] = 2.0 * (-45.0 / (3000.0 * 3000.0)) * 0.5
    np.testing.assert_array_equal(g, expect)


def test_points_outside_the_model_clamp_and_invalid_velocities():
    # both ends beyond the model: cells clamp to the faces; an invalid velocity is 1500 in the forward model
    # and skipped in the gradient
    src = np.array([[O[0] - 500.0, O[1] + 45.0, O[2] + 45.0]])
    sta = np.array([[O[0] + 900.0, O[1] + 45.0, O[2] + 45.0]])
    vp = np.full(int(np.prod(N)), 2000.0)
    idx = lambda x, y, z: (x * N[1] + y) * N[2] + z
    vp[idx(0, 4, 4)] = -1.0
    vp[idx(11, 4, 4)] = np.nan
    d = 1400.0
    ns = int(d / 50.0)
    ds = d / ns
    cells = [min(int(0 + (s + 0.5) / ns * 11), 11) for s in range(ns)]
    t = 0.0
    for c in cells:
        t += ds / (1500.0 if c in (0, 11) else 2000.0)
    assert ott.forward(src, sta, N, O, H, vp)[0] == t
    g = ott.gradient(src, sta, N, O, H, vp, np.array([1.0]))
    assert g[idx(0, 4, 4)] == 0 and g[idx(11, 4, 4)] == 0
    assert g[idx(5, 4, 4)] == sum(-ds / 2000.0 ** 2 for c in cells if c == 5)
    # coincident source and station: zero length, two samples of zero weight
    assert ott.forward(src, src, N, O, H, vp)[0] == 0.0


def test_gradient_is_minus_the_derivative_of_the_doubly_weighted_misfit():
    """The reference passes weighted residuals w (obs - t) and multiplies by w
    again, so its gradient is -dJ/dv for J = 1/2 sum w^2 (obs - t)^2."""
    rng = np.random.default_rng(4)
    src, sta = _picks(rng, 40, margin=20.0)
    V = int(np.prod(N))
    vp = rng.uniform(1800, 3500, V)
    w = rng.uniform(0.5, 2.0, 40)
    obs = ott.forward(src, sta, N, O, H, rng.uniform(1800, 3500, V))
    t = ott.forward(src, sta, N, O, H, vp)
    g = ott.gradient(src, sta, N, O, H, vp, w * (obs - t), w)

    def J(v):
        tt = ott.forward(src, sta, N, O, H, v)
        return 0.5 * float((w ** 2 * (obs - tt) ** 2).sum())
    for k in range(4):
        d = rng.standard_normal(V)
        eps = 1e-3
        fd = (J(vp + eps * d) - J(vp - eps * d)) / (2 * eps)
        assert math.isclose(-fd, float(g @ d), rel_tol=1e-6)
