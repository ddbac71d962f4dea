"""Pins the CPU oracle of the learned preconditioner (oracle/learned.py)
against what the mathematics fixes: a loop convolution, an edge-by-edge
matrix of M, the Kronecker-product grid Laplacian of an untrained network,
symmetry and positive definiteness, the energy-norm identity of CG, and
finite differences of the output gradient and of the unrolled-CG loss."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import learned as ol
from oracle import sparse as osp
from seismic_engine import synthetic


def tiny_system(grid=(4, 4, 4), nrays=120, seed=5, lam=1e-2):
    pts = synthetic.ray_endpoints(nrays, grid, 1.0, seed=seed)
    rp, ci, val = osp.ray_matrix(grid, 1.0, pts)
    return ol.normal_matrix(rp, ci, val, int(np.prod(grid)), lam)


def grid_laplacian(grid):
    """Neumann (graph) Laplacian of the 7-point grid graph as a Kronecker sum."""
    def path(n):
        T = sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tolil()
        T[0, 0] = T[n - 1, n - 1] = 1
        return T.tocsr()
    n0, n1, n2 = grid
    I0, I1, I2 = sp.identity(n0), sp.identity(n1), sp.identity(n2)
    return (sp.kron(sp.kron(path(n0), I1), I2) + sp.kron(sp.kron(I0, path(n1)), I2)
            + sp.kron(sp.kron(I0, I1), path(n2))).toarray()


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


def test_apply_matches_edge_by_edge_matrix():
    grid = (4, 3, 5)
    V = 60
    rng = np.random.default_rng(2)
    out = torch.tensor(rng.standard_normal((16, V)) * 0.5)
    dsq = torch.tensor(rng.uniform(0.5, 2.0, V))
    coef = ol.coefficients(out, dsq, grid)
    # far-face edges carry no weight
    x, y, z = np.meshgrid(np.arange(4), np.arange(3), np.arange(5), indexing="ij")
    assert torch.all(coef[1][torch.tensor((x == 3).ravel())] == 0)
    assert torch.all(coef[3][torch.tensor((z == 4).ravel())] == 0)
    M = ol.dense_m(coef, grid)
    r = torch.tensor(rng.standard_normal(V))
    assert np.allclose(ol.apply_m(coef, r, grid).numpy(), M @ r.numpy(), rtol=1e-12, atol=1e-12)


def test_m_is_symmetric_positive_definite_for_any_output():
    grid = (4, 4, 4)
    rng = np.random.default_rng(3)
    dsq = torch.tensor(rng.uniform(0.1, 3.0, 64))
    for scale in (0.1, 1.0, 3.0):
        coef = ol.coefficients(torch.tensor(rng.standard_normal((16, 64)) * scale), dsq, grid)
        M = ol.dense_m(coef, grid)
        assert np.allclose(M, M.T, rtol=0, atol=1e-12 * np.abs(M).max())
        # I + L_w >= I, so M >= S^2
        ev = np.linalg.eigvalsh(M / np.outer(coef[0].numpy(), coef[0].numpy()))
        assert ev.min() >= 1 - 1e-9


def test_untrained_network_is_jacobi_scaled_grid_laplacian():
    grid = (4, 4, 4)
    N = tiny_system(grid)
    feat, dsq = ol.features(N, grid, 1e-2)
    out = ol.network(torch.tensor(ol.init_params(0).astype(np.float64)), feat, grid)
    assert torch.all(out == 0)
    coef = ol.coefficients(out, dsq, grid)
    d = N.diagonal()
    Dm = np.diag(d ** -0.5)
    expect = Dm @ (np.eye(64) + ol.W0 * grid_laplacian(grid)) @ Dm
    assert np.allclose(ol.dense_m(coef, grid), expect, rtol=1e-12, atol=1e-14)


def test_output_gradient_matches_finite_differences():
    grid = (4, 4, 4)
    rng = np.random.default_rng(4)
    out = torch.tensor(rng.standard_normal((16, 64)) * 0.3)
    dsq = torch.tensor(rng.uniform(0.5, 2.0, 64))
    r = rng.standard_normal(64)
    g = rng.standard_normal(64)
    grad = ol.pair_output_gradient(out, dsq, grid, r, g).numpy()
    assert np.all(grad[4:] == 0)

    def f(o):
        return float(torch.tensor(g) @ ol.apply_m(ol.coefficients(o, dsq, grid), torch.tensor(r), grid))
    for _ in range(4):
        d = torch.tensor(rng.standard_normal((16, 64)))
        h = 1e-6
        fd = (f(out + h * d) - f(out - h * d)) / (2 * h)
        assert float((grad * d.numpy()).sum()) == pytest.approx(fd, rel=1e-7, abs=1e-10)


def exact_system(grid=(4, 4, 4), lam=1e-2):
    N = tiny_system(grid, lam=lam)
    Nd = N.toarray()
    rng = np.random.default_rng(6)
    rhs = Nd @ rng.uniform(0.5, 1.5, Nd.shape[0])
    return N, rhs, np.linalg.solve(Nd, rhs)


def test_energy_identity_of_the_loss():
    """e_k . r_k = |e_k|_N^2 along preconditioned CG (the loss needs no
    extra product with N), and the loss at k1 = k2 = 1 is the closed form of
    one preconditioned steepest-descent step."""
    grid = (4, 4, 4)
    N, rhs, xs = exact_system(grid)
    Nd = torch.tensor(N.toarray())
    rng = np.random.default_rng(7)
    coef = ol.coefficients(torch.tensor(rng.standard_normal((16, 64)) * 0.3), torch.tensor(N.diagonal() ** -0.5), grid)
    M = torch.tensor(ol.dense_m(coef, grid))
    b = torch.tensor(rhs)
    x_star = torch.tensor(xs)
    loss1 = float(ol.unrolled_cg_loss(coef, Nd, b, x_star, grid, 1, 1))
    z = M @ b
    a = (b @ z) / (z @ Nd @ z)
    e1 = x_star - a * z
    assert loss1 == pytest.approx(0.5 * np.log(float(e1 @ Nd @ e1) / float(x_star @ Nd @ x_star)), rel=1e-10)
    # the energy error of CG iterates decreases monotonically
    vals = [float(ol.unrolled_cg_loss(coef, Nd, b, x_star, grid, k, k)) for k in range(1, 12)]
    assert all(v2 < v1 for v1, v2 in zip(vals, vals[1:]))


def test_training_gradient_matches_finite_differences():
    grid = (4, 4, 4)
    N, rhs, xs = exact_system(grid)
    p = ol.init_params(7, zero_output=False) * 0.3
    loss, grad, _, _ = ol.loss_and_grad(p, N, grid, 1e-2, rhs, xs, 2, 6, round_bf16=False)
    rng = np.random.default_rng(9)
    for _ in range(3):
        d = rng.standard_normal(p.size)
        h = 1e-6
        lp = ol.loss_and_grad(p.astype(np.float64) + h * d, N, grid, 1e-2, rhs, xs, 2, 6, round_bf16=False)[0]
        lm = ol.loss_and_grad(p.astype(np.float64) - h * d, N, grid, 1e-2, rhs, xs, 2, 6, round_bf16=False)[0]
        assert grad @ d == pytest.approx((lp - lm) / (2 * h), rel=1e-5, abs=1e-9)
