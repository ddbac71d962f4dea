"""Pins the gradient post-processing oracle (oracle/gradient_post.py) by
closed forms and hand values.  CPU only."""
import numpy as np

from oracle import gradient_post as ogp


def test_tikhonov_of_a_quadratic_field():
    nx, ny, nz = 6, 5, 7
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    v = 3.0 * x ** 2 + 0.5 * y ** 2 - z ** 2 + 7.0
    g = np.ones(v.size)
    out = ogp.tikhonov(v, g, nx, ny, nz, 0.25).reshape(nx, ny, nz)
    # second differences: 6 along x, 1 along y, -2 along z, each only away from that axis' faces
    lap = (6.0 * ((x > 0) & (x < nx - 1)) + 1.0 * ((y > 0) & (y < ny - 1)) - 2.0 * ((z > 0) & (z < nz - 1)))
    np.testing.assert_allclose(out, 1.0 - 0.25 * lap, rtol=0, atol=1e-12)


def test_normalize_and_depth_weights():
    g = np.array([0.5, -4.0, 2.0])
    np.testing.assert_array_equal(ogp.normalize(g), [0.125, -1.0, 0.5])
    tiny = np.array([1e-31, -2e-31])
    np.testing.assert_array_equal(ogp.normalize(tiny), tiny)
    p = ogp.depth_weights(2, 1, 5, 2, dz=500.0).reshape(2, 1, 5)
    np.testing.assert_allclose(p[1, 0], [0, 0, 1.0, 1 / 1.05, 1 / 1.1])


def test_descent_inversion_lowers_the_misfit():
    rng = np.random.default_rng(3)
    n, o, h = (8, 8, 8), (0.0, 0.0, 0.0), (10.0, 10.0, 10.0)
    src = rng.uniform(0, 80, (120, 3))
    sta = rng.uniform(0, 80, (120, 3))
    true = np.full(512, 3000.0)
    true[200:260] = 3400.0
    from oracle import traveltime as ott
    obs = ott.forward(src, sta, n, o, h, true)
    _, hist, _ = ogp.invert(src, sta, n, o, h, np.ones(120), obs, np.full(512, 3000.0), max_iterations=5,
                            step_length=50.0, sigma=1.0, descent=True)
    assert hist[-1] < 0.5 * hist[0]
