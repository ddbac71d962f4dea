"""Pins the CPU oracle of the learned preconditioner (oracle/learned.py)
against what the mathematics fixes: a loop convolution, adjointness of the
stencils, the Jacobi limit of an untrained network, the exact Kaporin number,
finite differences of the loss, and the optimality of the FSAI factor."""
import numpy as np
import pytest
import torch

from oracle import learned as ol
from oracle import sparse as osp
from seismic_engine import synthetic


def tiny_system(grid=(4, 4, 4), nrays=120, seed=5, lam=1e-2):
    pts = synthetic.ray_endpoints(nrays, grid, 1.0, seed=seed)
    rp, ci, val = osp.ray_matrix(grid, 1.0, pts)
    return ol.normal_matrix(rp, ci, val, int(np.prod(grid)), lam)


def test_conv_matches_loop_definition():
    rng = np.random.default_rng(0)
    grid = (4, 3, 5)
    x = rng.standard_normal((60, 16))
    w = rng.standard_normal((16, 27, 16))
    ref = ol.conv_bruteforce(x, w, grid)
    got = ol.conv(torch.tensor(x), torch.tensor(w), grid).numpy()
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-10)


def test_pool_and_upcat():
    rng = np.random.default_rng(1)
    grid = (4, 4, 2)
    x = torch.tensor(rng.standard_normal((32, 3)))
    p = ol.pool(x, grid)
    x5 = x.numpy().reshape(4, 4, 2, 3)
    assert np.allclose(p.numpy().reshape(2, 2, 1, 3), x5.reshape(2, 2, 2, 2, 1, 2, 3).mean(axis=(1, 3, 5)))
    fine = torch.tensor(rng.standard_normal((32, 2)))
    cat = ol.upcat(p, fine, grid).numpy()
    for q in range(32):
        X, Y, Z = q // 8, (q // 2) % 4, q % 2
        assert np.allclose(cat[q, :3], p.numpy()[(X // 2 * 2 + Y // 2) * 1 + Z // 2])
        assert np.allclose(cat[q, 3:], fine.numpy()[q])


def test_stencil_transpose_and_triangularity():
    grid = (4, 3, 5)
    V = 60
    rng = np.random.default_rng(2)
    g = torch.tensor(rng.standard_normal((ol.NG, V)))
    Gd = ol.dense_g(g, grid).numpy()
    # upper triangular in the x-major cell order, diagonal = g[0]
    assert np.allclose(np.tril(Gd, -1), 0)
    assert np.allclose(np.diag(Gd), g[0].numpy())
    y = torch.tensor(rng.standard_normal(V))
    assert np.allclose(ol.apply_gt(g, y, grid).numpy(), Gd.T @ y.numpy())


def test_untrained_network_is_jacobi():
    grid = (4, 4, 4)
    N = tiny_system(grid)
    feat, dsq = ol.features(N, grid, 1e-2)
    out = ol.network(torch.tensor(ol.init_params(0).astype(np.float64)), feat, grid)
    g = ol.coefficients(out)
    assert torch.all(g[0] == 1) and torch.all(g[1:] == 0)
    s = torch.tensor(np.random.default_rng(3).standard_normal(64))
    assert np.allclose(ol.apply_m(g, dsq, s, grid).numpy(), s.numpy() / N.diagonal())


def test_features_definition():
    grid = (4, 4, 4)
    N = tiny_system(grid)
    feat, dsq = ol.features(N, grid, 1e-2, round_bf16=False)
    Nd = N.toarray()
    i = (1 * 4 + 2) * 4 + 3          # cell (1, 2, 3)
    j = (2 * 4 + 1) * 4 + 3          # its neighbour (+1, -1, 0): tap (2, 0, 1) = 19 -> channel 18
    assert feat[i, 18] == pytest.approx(Nd[i, j] / np.sqrt(Nd[i, i] * Nd[j, j]))
    assert feat[i, 26] == 1.0 and torch.all(feat[:, 27:] == 0)
    assert np.allclose(dsq.numpy(), 1 / np.sqrt(np.diag(Nd)))
    # cell (0, *, *) has no x-1 neighbours: channels 0..8 vanish
    assert torch.all(feat[:16, :9] == 0)


def test_loss_with_basis_probes_is_exact_kaporin():
    grid = (4, 4, 4)
    V = 64
    N = tiny_system(grid)
    _, dsq = ol.features(N, grid, 1e-2)
    Jd = torch.tensor(N.toarray()) * torch.outer(dsq, dsq)
    rng = np.random.default_rng(4)
    g = torch.tensor(rng.standard_normal((ol.NG, V)) * 0.1)
    g[0] = torch.tensor(np.exp(rng.standard_normal(V) * 0.2))
    probes = torch.eye(V, dtype=torch.float64) * np.sqrt(V)
    loss = float(ol.kaporin_loss(g, ol.j_operator(N, dsq), probes, grid))
    exact = ol.kaporin_exact(g, Jd, grid)
    assert loss - float(torch.logdet(Jd)) / V == pytest.approx(exact, rel=1e-10, abs=1e-12)
    assert exact >= 0
    # scale invariance: K(cG) = K(G)
    assert ol.kaporin_exact(g * 3.0, Jd, grid) == pytest.approx(exact, rel=1e-10)


def test_gradient_matches_finite_differences():
    grid = (4, 4, 4)
    N = tiny_system(grid)
    p = ol.init_params(7, zero_output=False) * 0.3
    probes = np.random.default_rng(8).choice([-1.0, 1.0], size=(3, 64))
    loss, grad, _, _ = ol.loss_and_grad(p, N, grid, 1e-2, probes, round_bf16=False)
    rng = np.random.default_rng(9)
    for _ in range(4):
        d = rng.standard_normal(p.size)
        h = 1e-6   # the loss is strongly curved (log, exp): larger steps bias the difference
        lp = ol.loss_and_grad(p.astype(np.float64) + h * d, N, grid, 1e-2, probes, round_bf16=False)[0]
        lm = ol.loss_and_grad(p.astype(np.float64) - h * d, N, grid, 1e-2, probes, round_bf16=False)[0]
        fd = (lp - lm) / (2 * h)
        assert grad @ d == pytest.approx(fd, rel=1e-6, abs=1e-9)


def test_fsai_factor_minimises_exact_kaporin():
    """Row-wise FSAI on the upper 63-tap pattern is the exact minimiser of K:
    random perturbations of it only increase K."""
    grid = (4, 4, 4)
    V = 64
    N = tiny_system(grid)
    d = N.diagonal()
    Jd = N.toarray() / np.sqrt(np.outer(d, d))
    offs = ol.g_offsets()
    g = np.zeros((ol.NG, V))
    for i in range(V):
        x, y, z = i // 16, (i // 4) % 4, i % 4
        ks, cols = [], []
        for k, (a, b, c) in enumerate(offs):
            if 0 <= x + a < 4 and 0 <= y + b < 4 and 0 <= z + c < 4:
                ks.append(k)
                cols.append(((x + a) * 4 + y + b) * 4 + z + c)
        e = np.zeros(len(cols))
        e[0] = 1.0                       # k = 0 is the diagonal
        sol = np.linalg.solve(Jd[np.ix_(cols, cols)], e)
        sol /= np.sqrt(sol[0])
        g[ks, i] = sol
    Jt = torch.tensor(Jd)
    k_fsai = ol.kaporin_exact(torch.tensor(g), Jt, grid)
    k_jac = ol.kaporin_exact(torch.tensor(np.vstack([np.ones((1, V)), np.zeros((ol.NG - 1, V))])), Jt, grid)
    assert k_fsai < k_jac
    rng = np.random.default_rng(10)
    mask = np.abs(g) > 0
    for _ in range(5):
        gp = g + 1e-2 * rng.standard_normal(g.shape) * mask
        assert ol.kaporin_exact(torch.tensor(gp), Jt, grid) > k_fsai
