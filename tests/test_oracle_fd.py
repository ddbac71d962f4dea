"""Pins the FD oracle (oracle/fd_oracle.c) against closed forms, an independent
Kronecker-product operator, invariants and hand-computed values.  CPU only."""
from fractions import Fraction as F

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import fd as ofd


def test_weights_are_the_central_8th_order_ones():
    # symbol a0 + 2 sum a_k cos(k t) = -t^2 + O(t^10): moment conditions
    a = [F(-205, 72), F(8, 5), F(-1, 5), F(8, 315), F(-1, 560)]
    assert a[0] + 2 * sum(a[1:]) == 0
    assert sum(a[k] * k**2 for k in range(1, 5)) == 1
    for m in (2, 3, 4):
        assert sum(a[k] * k ** (2 * m) for k in range(1, 5)) == 0
    np.testing.assert_allclose(ofd.weights(8), [float(x) for x in a], rtol=0, atol=1e-15)
    np.testing.assert_array_equal(ofd.weights(2), [-2.0, 1.0])


def test_stability_limits():
    # 3-D leapfrog: v dt / h <= sqrt(4 / (3 |symbol(pi)|))
    assert ofd.stability_limit(2) == pytest.approx(1 / np.sqrt(3), rel=1e-14)
    s8 = -205 / 72 + 2 * (-8 / 5 - 1 / 5 - 8 / 315 - 1 / 560)
    assert ofd.stability_limit(8) == pytest.approx(np.sqrt(4 / (3 * abs(s8))), rel=1e-14)
    assert ofd.stability_limit(8) == pytest.approx(0.45286, abs=1e-5)


def _kron_laplacian(n, hw, order, h):
    a = ofd.weights(order)
    R = len(a) - 1
    mats = []
    for ax in range(3):
        m = n[ax] + 2 * hw
        diags = [np.full(m - abs(k), a[abs(k)] / h[ax] ** 2) for k in range(-R, R + 1)]
        mats.append(sp.diags(diags, list(range(-R, R + 1)), shape=(m, m), format="csr"))
    I = [sp.identity(n[ax] + 2 * hw, format="csr") for ax in range(3)]
    return (sp.kron(sp.kron(mats[0], I[1]), I[2]) + sp.kron(sp.kron(I[0], mats[1]), I[2])
            + sp.kron(sp.kron(I[0], I[1]), mats[2])).tocsr()


@pytest.mark.parametrize("order,hw", [(2, 1), (8, 4), (2, 4)])
def test_one_step_matches_independent_kron_operator(order, hw):
    n = (7, 9, 6)
    h = (10.0, 12.5, 8.0)
    c = ofd.cfg(n, hw=hw, order=order, h=h, dt=5e-4, sponge_width=0)
    nh = ofd.nh(c)
    rng = np.random.default_rng(1)
    interior = np.zeros(nh, bool)
    interior[hw:-hw, hw:-hw, hw:-hw] = True
    u = np.where(interior, rng.standard_normal(nh), 0.0)
    p = np.where(interior, rng.standard_normal(nh), 0.0)
    v = 1500 + 1000 * rng.random(nh)
    s, nxt = ofd.stencil(c, u, p, v)
    assert s == 0
    L = _kron_laplacian(n, hw, order, h)
    expect = 2 * u - p + (v**2 * c.dt**2) * (L @ u.ravel()).reshape(nh)
    expect[~interior] = 0.0
    np.testing.assert_allclose(nxt, expect, rtol=1e-12, atol=1e-12 * np.abs(expect).max())


def test_plane_wave_symbol_closed_form():
    # far from the boundary L8 sin(k x) = symbol(k h)/h^2 sin(k x) exactly
    n = (40, 12, 12)
    c = ofd.cfg(n, hw=4, order=8, dt=1e-3, sponge_width=0)
    nh = ofd.nh(c)
    kx = 0.7 / c.dx
    x = np.arange(nh[0]) * c.dx
    u = np.zeros(nh)
    u[4:-4, 4:-4, 4:-4] = np.sin(kx * x)[4:-4, None, None]
    v = np.full(nh, 2000.0)
    _, nxt = ofd.stencil(c, u, np.zeros(nh), v)
    a = ofd.weights(8)
    sym = (a[0] + 2 * sum(a[k] * np.cos(k * kx * c.dx) for k in range(1, 5))) / c.dx**2
    # y/z taps of a y,z-constant field vanish only at cells >= 4 from the y,z faces
    sl = (slice(8, -8), slice(8, -8), slice(8, -8))
    expect = 2 * u + (2000.0 * c.dt) ** 2 * sym * u
    np.testing.assert_allclose(nxt[sl], expect[sl], rtol=1e-11, atol=1e-13)


def test_sponge_hand_values():
    c = ofd.cfg((50, 50, 50), hw=4, sponge_width=20, sponge_damping=0.015)
    f = ofd.lib().orc_fd_sponge_factor
    import ctypes as C
    f.argtypes = [C.POINTER(ofd.OrcCfg), C.c_int64, C.c_int64, C.c_int64]
    assert f(C.byref(c), 4, 29, 29) == pytest.approx(0.985, abs=1e-15)           # on the x face
    assert f(C.byref(c), 9, 34, 29) == pytest.approx(1 - 0.015 * 225 / 400, abs=1e-15)  # d = 5
    assert f(C.byref(c), 29, 29, 29) == 1.0                                      # centre
    assert f(C.byref(c), 2, 29, 29) == 1.0                                       # halo untouched
    assert f(C.byref(c), 53, 53, 53) == pytest.approx(0.985, abs=1e-15)          # far corner interior


def test_ricker_closed_form():
    f, dt = 15.0, 1e-3
    w = ofd.ricker(f, dt, 1 / f, 200)
    t0_idx = round((1 / f) / dt)
    assert abs(w[t0_idx] - 1.0) < 1e-3
    tau = np.arange(200) * dt - 1 / f
    np.testing.assert_allclose(w, (1 - 2 * (np.pi * f * tau) ** 2) * np.exp(-(np.pi * f * tau) ** 2), rtol=1e-13)
    from seismic_engine import synthetic
    np.testing.assert_allclose(synthetic.ricker(f, dt, 200), w, rtol=1e-13, atol=1e-15)


def test_time_reversibility_fp64():
    n = (20, 18, 16)
    c = ofd.cfg(n, hw=4, order=8, dt=1e-3, sponge_width=0)
    nh = ofd.nh(c)
    rng = np.random.default_rng(3)
    u0 = np.zeros(nh)
    u0[4:-4, 4:-4, 4:-4] = rng.standard_normal(n)
    u1 = u0.copy()
    v = np.full(nh, 3000.0)
    cur, prev = u1.copy(), u0.copy()
    ofd.run(c, v, cur, prev, 60)
    cur, prev = prev.copy(), cur.copy()
    ofd.run(c, v, cur, prev, 60)
    np.testing.assert_allclose(cur, u0, atol=1e-9 * np.abs(u0).max())
    np.testing.assert_allclose(prev, u1, atol=1e-9 * np.abs(u0).max())


@pytest.mark.parametrize("factor,stable", [(0.95, True), (1.05, False)])
def test_stability_threshold(factor, stable):
    n = (24, 24, 24)
    lim = ofd.stability_limit(8)
    dt = factor * lim * 10.0 / 3000.0
    c = ofd.cfg(n, hw=4, order=8, dt=dt, sponge_width=0)
    nh = ofd.nh(c)
    rng = np.random.default_rng(5)
    cur = np.zeros(nh)
    cur[4:-4, 4:-4, 4:-4] = rng.standard_normal(n)
    prev = cur.copy()
    ofd.run(c, np.full(nh, 3000.0), cur, prev, 400)
    growth = np.abs(cur).max() / 4.0
    assert (growth < 10) == stable


def test_linearity_in_source_amplitude():
    n = (16, 16, 16)
    c = ofd.cfg(n, hw=4, order=8, sponge_width=4)
    v = np.full(ofd.nh(c), 2000.0)
    w = ofd.ricker(15.0, 1e-3, 1 / 15, 60)
    a, _, _, _ = ofd.forward(c, v, w, 60)
    b, _, _, _ = ofd.forward(c, v, 2.5 * w, 60)
    np.testing.assert_allclose(b, 2.5 * a, rtol=1e-12, atol=1e-12 * np.abs(a).max())


def test_mirror_symmetry_homogeneous():
    n = (17, 17, 17)
    c = ofd.cfg(n, hw=4, order=8, sponge_width=5)
    v = np.full(ofd.nh(c), 2000.0)
    w = ofd.ricker(15.0, 1e-3, 1 / 15, 50)
    cur, _, _, res = ofd.forward(c, v, w, 50)
    assert res.completed_successfully == 1
    for ax in range(3):
        np.testing.assert_allclose(cur, np.flip(cur, ax), rtol=0, atol=1e-12 * np.abs(cur).max())
    np.testing.assert_allclose(cur, cur.transpose(1, 0, 2), rtol=0, atol=1e-12 * np.abs(cur).max())


def test_cfl_violation_stops_before_stepping():
    n = (10, 10, 10)
    c = ofd.cfg(n, hw=4, order=8, dt=1e-3)
    v = np.full(ofd.nh(c), 5000.0)   # 0.5 > 0.49
    _, _, _, res = ofd.forward(c, v, np.ones(5), 5)
    assert res.status == 2 and res.timesteps_completed == 0 and res.completed_successfully == 0
    v[:] = 4500.0                    # 0.45 passes both bounds (0.49 and 0.45286)
    assert ofd.verify_cfl(c, v)
    v[:] = 4550.0                    # 0.455: under 0.49 but over the order-8 bound
    assert not ofd.verify_cfl(c, v)


def test_nan_guards_follow_reference_rules():
    n = (8, 8, 8)
    c = ofd.cfg(n, hw=4, order=8, sponge_width=0)
    nh = ofd.nh(c)
    u = np.zeros(nh)
    u[4:-4, 4:-4, 4:-4] = 1.0
    u[8, 8, 8] = np.nan
    p = u.copy()
    p[8, 8, 8] = 0.0
    s, nxt = ofd.stencil(c, u, p, np.full(nh, 2000.0))
    assert s == 1
    assert nxt[8, 8, 8] == 0.0                 # non-finite centre -> 0
    assert nxt[8, 8, 10] == 1.0                # non-finite neighbour -> keep centre
    assert nxt[8, 8, 11] == 1.0
    assert np.isfinite(nxt).all()


def test_repair_field_nan_hand_case():
    # 3x3x3 with halo 1 -> only the centre is interior
    d = np.arange(27, dtype=float)
    d[13] = np.nan
    d[0] = np.inf
    out = ofd.repair_field_nan(d, 3, 3, 3, 1, -7.0)
    # centre has no interior neighbours (all are halo) -> replace value
    assert out[13] == -7.0 and out[0] == -7.0
    # 4x3x3, halo 1 -> interior cells (1,1,1),(2,1,1): idx 13 and 22
    d = np.arange(36, dtype=float)
    d[13] = np.nan
    out = ofd.repair_field_nan(d, 4, 3, 3, 1, 0.0)
    assert out[13] == 22.0   # only interior neighbour is (2,1,1)


def _sponge_run(sponge, n=32, sw=12, nt=800):
    from seismic_engine import synthetic
    c = ofd.cfg((n, n, n), hw=4, order=8, sponge_width=sw, sponge=sponge)
    v = np.full(ofd.nh(c), 3000.0)
    w = ofd.ricker(25.0, 1e-3, 1 / 25, nt)
    cur = np.zeros(ofd.nh(c))
    prev = np.zeros_like(cur)
    src = synthetic.cell_index((n, n, n), 4, n // 2, n // 2, n // 2)
    peaks = []
    for k in range(0, nt, 100):
        ofd.run(c, v, cur, prev, 100, src_idx=[src], src_w=w[None, k:k + 100])
        peaks.append(np.abs(cur).max())
    return peaks


def test_reference_sponge_grows_and_cerjan_sponge_decays():
    """The reference sponge damps the current and previous fields before the
    swap, leaving u^{t+1} = 2u^t - D u^{t-1} + C L u^t: for smooth modes
    inside the sponge the recurrence has the root 1 + sqrt(1 - D) > 1 and the
    field grows long after the wave has left.  The Cerjan form (next and
    current damped) has the root D < 1 there and decays."""
    ref = _sponge_run("reference")
    cer = _sponge_run("cerjan")
    assert ref[-1] > 1e3 * ref[0]
    assert cer[-1] < 0.1 * cer[0]
    assert max(cer) < 2 * cer[0]


def test_cerjan_step_damps_next_and_current():
    """One step from the same state: the new previous field is D u in both
    forms, the new current field of the Cerjan form is D times the reference
    one, and without a sponge the two forms are the same scheme."""
    n = (18, 16, 20)
    out = {}
    for mode in ("reference", "cerjan"):
        c = ofd.cfg(n, hw=4, order=8, sponge_width=6, sponge=mode)
        cur = np.zeros(ofd.nh(c))
        cur[4:-4, 4:-4, 4:-4] = np.random.default_rng(8).standard_normal(n)
        prev = np.zeros_like(cur)
        prev[4:-4, 4:-4, 4:-4] = np.random.default_rng(9).standard_normal(n)
        u0 = cur.copy()
        ofd.run(c, np.full(ofd.nh(c), 2500.0), cur, prev, 1)
        out[mode] = (cur, prev, u0)
    d = ofd.sponge_factors(ofd.cfg(n, hw=4, order=8, sponge_width=6))
    np.testing.assert_allclose(out["cerjan"][1], d * out["cerjan"][2], rtol=1e-15, atol=0)
    np.testing.assert_array_equal(out["cerjan"][1], out["reference"][1])
    np.testing.assert_allclose(out["cerjan"][0], d * out["reference"][0], rtol=1e-15, atol=0)
    assert not np.array_equal(out["cerjan"][0], out["reference"][0])
    runs = []
    for mode in ("reference", "cerjan"):
        c = ofd.cfg(n, hw=4, order=8, sponge_width=0, sponge=mode)
        w = ofd.ricker(25.0, 1e-3, 1 / 25, 60)
        runs.append(ofd.forward(c, np.full(ofd.nh(c), 2500.0), w, 60)[0])
    np.testing.assert_array_equal(runs[0], runs[1])
