"""GPU travel-time operator (stp_tt_forward / stp_tt_gradient through the C
ABI) against the oracle restatement of the reference's Python loops:
travel times bit-exact (FP64, same operation order), gradient within 1e-12
relative (FP64 atomics change the order of the additions per cell)."""
from dataclasses import dataclass, field

import numpy as np
import pytest
import torch

from oracle import traveltime as ott
from seismic_engine.traveltime import Geometry, TravelTimeOperator

pytestmark = pytest.mark.gpu

G = Geometry(24, 18, 30, dx=12.5, dy=10.0, dz=7.5, origin_x=-100.0, origin_y=250.0, origin_z=3.0)
N = (G.nx, G.ny, G.nz)
O = (G.origin_x, G.origin_y, G.origin_z)
H = (G.dx, G.dy, G.dz)


def _case(k, seed):
    rng = np.random.default_rng(seed)
    lo = np.array(O) - 40.0
    hi = np.array(O) + np.array(N) * np.array(H) + 40.0
    src = rng.uniform(lo, hi, (k, 3))
    sta = rng.uniform(lo, hi, (k, 3))
    V = int(np.prod(N))
    vp = rng.uniform(1500, 4500, V)
    vp[rng.integers(0, V, 30)] = -5.0
    vp[rng.integers(0, V, 30)] = np.nan
    vp[rng.integers(0, V, 10)] = np.inf
    return src, sta, vp, rng


def test_forward_is_bit_exact(cuda):
    src, sta, vp, _ = _case(3000, 1)
    op = TravelTimeOperator(G, src, sta, device=cuda)
    t = op.forward(vp).cpu().numpy()
    t_o = ott.forward(src, sta, N, O, H, vp)
    assert np.array_equal(t, t_o)


def test_gradient_matches_oracle(cuda):
    src, sta, vp, rng = _case(3000, 2)
    w = rng.uniform(0.2, 3.0, len(src))
    r = rng.standard_normal(len(src)) * 1e-2
    r[::97] = np.nan
    op = TravelTimeOperator(G, src, sta, weights=w, device=cuda)
    g = op.gradient(vp, r).cpu().numpy()
    g_o = ott.gradient(src, sta, N, O, H, vp, r, w)
    assert np.linalg.norm(g - g_o) <= 1e-12 * np.linalg.norm(g_o)
    assert np.all(g[~np.isfinite(vp) | (vp <= 0)] == 0)


def test_misfit_residuals_and_missing_picks(cuda):
    @dataclass
    class P:
        source_id: str
        station_id: str
        weight: float = 1.0

    @dataclass
    class Pt:
        x: float
        y: float
        z: float

    @dataclass
    class Obs:
        picks: list = field(default_factory=list)
        sources: dict = field(default_factory=dict)
        stations: dict = field(default_factory=dict)

    src, sta, vp, rng = _case(6, 3)
    obs = Obs()
    for i in range(6):
        obs.sources[f"s{i}"] = Pt(*src[i])
        obs.stations[f"r{i}"] = Pt(*sta[i])
        obs.picks.append(P(f"s{i}", f"r{i}", weight=0.5 + i))
    obs.picks.append(P("missing", "r0", weight=2.0))
    op = TravelTimeOperator.from_obs(G, obs, device=cuda)
    t = op.forward(vp).cpu().numpy()
    assert t[6] == 0.0
    assert np.array_equal(t[:6], ott.forward(src, sta, N, O, H, vp))
    observed = t + rng.standard_normal(7) * 1e-3
    w = np.array([p.weight for p in obs.picks])
    assert abs(op.misfit(torch.as_tensor(t, device=cuda), observed) - 0.5 * float((w * (observed - t) ** 2).sum())) \
        < 1e-15
    wr = op.weighted_residuals(torch.as_tensor(t, device=cuda), observed)
    g = op.gradient(vp, wr).cpu().numpy()
    g_o = ott.gradient(src, sta, N, O, H, vp, (w * (observed - t))[:6], w[:6])
    assert np.linalg.norm(g - g_o) <= 1e-12 * np.linalg.norm(g_o)
    empty = TravelTimeOperator(G, np.zeros((0, 3)), np.zeros((0, 3)), device=cuda)
    assert empty.forward(vp).numel() == 0
    assert not empty.gradient(vp, np.zeros(0)).any()
