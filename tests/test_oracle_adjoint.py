"""Pins the reverse-time oracle (oracle/adjoint_oracle.c): its gradient equals
central finite differences of the forward oracle's misfit, and its image is
linear in the injected traces.  CPU only."""
import numpy as np
import pytest

from oracle import adjoint as oad
from oracle import fd as ofd
from seismic_engine import synthetic

N = (14, 12, 10)
HW = 4
NT = 70


def _case(sponge="reference"):
    c = ofd.cfg(N, hw=HW, order=8, sponge_width=4, sponge=sponge)
    v_true = synthetic.pad_halo(synthetic.smooth_random_velocity(*N, seed=1, sigma=2.5, vmin=1800, vmax=2600), HW)
    v0 = synthetic.pad_halo(synthetic.smooth_random_velocity(*N, seed=2, sigma=3.0, vmin=1900, vmax=2500), HW)
    w = synthetic.ricker(25.0, 1e-3, NT)
    src = synthetic.cell_index(N, HW, 7, 6, 2)
    rcv = synthetic.surface_receivers(*N, HW, iz=1)
    return c, v_true, v0, w, src, rcv


@pytest.mark.parametrize("sponge", ["reference", "cerjan"])
def test_gradient_equals_finite_differences(sponge):
    c, v_true, v0, w, src, rcv = _case(sponge)
    _, _, obs = oad.adjoint(c, v_true, w, NT, src, rcv)
    _, _, tr = oad.adjoint(c, v0, w, NT, src, rcv)
    acc, _, _ = oad.adjoint(c, v0, w, NT, src, rcv, inj=tr - obs, scale_c=True, snap_lap=True)
    g = 2 * acc / v0[HW:-HW, HW:-HW, HW:-HW]
    assert oad.misfit(c, v0, w, NT, src, rcv, obs) == np.float64(0.5 * ((tr - obs) ** 2).sum()) or \
        abs(oad.misfit(c, v0, w, NT, src, rcv, obs) - 0.5 * ((tr - obs) ** 2).sum()) < 1e-12
    rng = np.random.default_rng(0)
    for _ in range(3):
        d = np.zeros_like(v0)
        d[HW:-HW, HW:-HW, HW:-HW] = rng.standard_normal(N)
        eps = 1e-2
        fdv = (oad.misfit(c, v0 + eps * d, w, NT, src, rcv, obs)
               - oad.misfit(c, v0 - eps * d, w, NT, src, rcv, obs)) / (2 * eps)
        an = float((g * d[HW:-HW, HW:-HW, HW:-HW]).sum())
        assert abs(fdv - an) < 1e-6 * abs(fdv)


def test_image_is_linear_in_injection_and_zero_without_it():
    c, v_true, v0, w, src, rcv = _case()
    _, _, data = oad.adjoint(c, v_true, w, NT, src, rcv)
    a1, ill, _ = oad.adjoint(c, v0, w, NT, src, rcv, inj=data, scale_c=False, snap_lap=False, ill=True)
    a2, _, _ = oad.adjoint(c, v0, w, NT, src, rcv, inj=-3.0 * data, scale_c=False, snap_lap=False)
    np.testing.assert_allclose(a2, -3.0 * a1, rtol=1e-12, atol=1e-12 * np.abs(a1).max())
    a0, _, _ = oad.adjoint(c, v0, w, NT, src, rcv, inj=None)
    assert not a0.any()
    assert (ill >= 0).all() and ill.max() > 0


def test_laplacian_matches_stencil_step():
    c = ofd.cfg(N, hw=HW, order=8, sponge_width=0)
    rng = np.random.default_rng(3)
    u = np.zeros(ofd.nh(c))
    u[HW:-HW, HW:-HW, HW:-HW] = rng.standard_normal(N)
    v = np.full(ofd.nh(c), 2000.0)
    _, nxt = ofd.stencil(c, u, np.zeros_like(u), v)
    lap = oad.laplacian(c, u)
    inner = (slice(HW, -HW),) * 3
    np.testing.assert_allclose(nxt[inner], (2 * u + (2000.0 * c.dt) ** 2 * lap)[inner], rtol=1e-12, atol=1e-12)
