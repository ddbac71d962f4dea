"""Gradient post-processing kernels and the travel-time inversion loop on the
GPU against the CPU oracle (oracle/gradient_post.py): the elementwise
kernels bit-exact, the loop's misfit history and model within 1e-9."""
import numpy as np
import pytest
import torch

from oracle import gradient_post as ogp
from oracle import sparse as osp
from oracle import traveltime as ott
from seismic_engine.gradient import GradientComputer
from seismic_engine.inversion import TravelTimeInversion
from seismic_engine.traveltime import Geometry, TravelTimeOperator

pytestmark = pytest.mark.gpu


def test_post_processing_kernels_match_oracle(cuda):
    nx, ny, nz = 17, 12, 23
    rng = np.random.default_rng(0)
    v = rng.uniform(1500, 4500, nx * ny * nz)
    g = rng.standard_normal(nx * ny * nz)
    gc = GradientComputer(nx, ny, nz, smoothing_sigma=1.5, regularization_lambda=3e-3)
    r = gc.apply_regularization(v, g).cpu().numpy()
    assert np.array_equal(r, ogp.tikhonov(v, g, nx, ny, nz, 3e-3))
    # z columns across wavefront (64) and workgroup (256) edges, y rows across
    # the 16-row chunks, and single-cell axes
    for sx, sy, sz in ((4, 37, 300), (3, 1, 130), (1, 5, 64), (2, 18, 1)):
        vv = rng.uniform(1500, 4500, sx * sy * sz)
        gg = rng.standard_normal(sx * sy * sz)
        rr = GradientComputer(sx, sy, sz, regularization_lambda=0.7).apply_regularization(vv, gg).cpu().numpy()
        assert np.array_equal(rr, ogp.tikhonov(vv, gg, sx, sy, sz, 0.7)), (sx, sy, sz)
    assert np.array_equal(gc.normalize_gradient(r).cpu().numpy(), ogp.normalize(r))
    tiny = np.full(50, 1e-31)
    assert np.array_equal(gc.normalize_gradient(tiny).cpu().numpy(), tiny)
    nan = g.copy()
    nan[7] = np.nan
    out = gc.normalize_gradient(nan).cpu().numpy()
    assert np.array_equal(out[np.isfinite(out)], g[np.isfinite(nan)]) and np.isnan(out[7])
    p = gc.build_depth_preconditioner(5, dz=300.0).cpu().numpy()
    assert np.array_equal(p, ogp.depth_weights(nx, ny, nz, 5, dz=300.0))
    assert np.array_equal(gc.apply_preconditioning(g).cpu().numpy(), g * p)
    s = gc.apply_smoothing(g).cpu().numpy()
    np.testing.assert_allclose(s, osp.smooth_gradient(g, nx, ny, nz, 1.5), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("descent", [False, True])
def test_inversion_loop_matches_oracle(descent, cuda):
    rng = np.random.default_rng(5 + descent)
    geo = Geometry(10, 9, 8, dx=10.0, dy=10.0, dz=10.0)
    n, o, h = (10, 9, 8), (0.0, 0.0, 0.0), (10.0, 10.0, 10.0)
    src = rng.uniform(0, 80, (300, 3))
    sta = rng.uniform(0, 80, (300, 3))
    w = rng.uniform(0.5, 1.5, 300)
    true = rng.uniform(2800, 3600, 720)
    obs = ott.forward(src, sta, n, o, h, true)
    vp0 = np.full(720, 3200.0)
    kw = dict(max_iterations=4, step_length=40.0, smoothing_sigma=1.0, regularization_lambda=1e-4,
              misfit_type="l2", descent=descent)
    op = TravelTimeOperator(geo, src, sta, weights=w, device=cuda)
    inv = TravelTimeInversion(op, obs, vp0, **kw)
    res = inv.run()
    vp_o, hist_o, conv_o = ogp.invert(src, sta, n, o, h, w, obs, vp0, max_iterations=4, step_length=40.0, sigma=1.0,
                                      lam=1e-4, descent=descent)
    np.testing.assert_allclose(res["misfit_history"], hist_o, rtol=1e-9)
    np.testing.assert_allclose(inv.vp.cpu().numpy(), vp_o, rtol=1e-9)
    assert res["converged"] == conv_o
    if descent:
        assert hist_o[-1] < hist_o[0]
