"""Learned preconditioner on the GPU against the CPU oracle (oracle/learned.py):
BF16 MFMA convolutions forward and backward, coupling features, the network's
loss and parameter gradient, the factor stencils, and CGLS with the trained
network against the Jacobi fallback."""
import numpy as np
import pytest
import torch

from oracle import learned as ol
from oracle import sparse as osp
from seismic_engine import synthetic
from seismic_engine.learned import (LearnedPreconditioner, conv3d_bf16, conv3d_bf16_bwd, rademacher,
                                    to_bf16_bits, tomography_system, train)
from seismic_engine.sparse import DeviceCSR

pytestmark = pytest.mark.gpu
LAM = 1e-2


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def device_system(grid, nrays, seed):
    """Siddon matrix from the CPU oracle, uploaded as FP32 / int32."""
    V = int(np.prod(grid))
    pts = synthetic.ray_endpoints(nrays, grid, 1.0, seed=seed)
    rp, ci, val = osp.ray_matrix(grid, 1.0, pts)
    val32 = val.astype(np.float32)
    A = DeviceCSR.from_device(len(rp) - 1, V, torch.as_tensor(rp, dtype=torch.int32, device="cuda"),
                              torch.as_tensor(ci, dtype=torch.int32, device="cuda"),
                              torch.as_tensor(val32, device="cuda"))
    return A, ol.normal_matrix(rp, ci, val32, V, LAM)


@pytest.mark.parametrize("cin,cout", [(32, 16), (16, 32), (32, 64), (96, 32), (48, 16), (16, 64)])
def test_conv3d_forward_matches_oracle(cin, cout, cuda):
    grid = (8, 4, 12)
    V = int(np.prod(grid))
    g = torch.Generator().manual_seed(cin * 100 + cout)
    x = torch.randn(V, cin, generator=g).to(torch.bfloat16)
    w = torch.randn(cout, 27, cin, generator=g) * 0.2
    for relu in (False, True):
        out = conv3d_bf16(x.view(torch.int16).to(cuda), w.to(cuda), grid, relu=relu).cpu()
        ref = ol.conv(x.double(), ol.bf16(w).double(), grid)
        if relu:
            ref = torch.relu(ref)
        assert rel(out, ref) < 1e-5, (cin, cout, relu)


@pytest.mark.parametrize("cin,cout", [(32, 16), (96, 32), (16, 64)])
def test_conv3d_backward_matches_oracle(cin, cout, cuda):
    grid = (8, 8, 4)
    V = int(np.prod(grid))
    g = torch.Generator().manual_seed(7 + cin + cout)
    x = torch.randn(V, cin, generator=g).to(torch.bfloat16)
    w = ol.bf16(torch.randn(cout, 27, cin, generator=g) * 0.2)
    dout = torch.randn(V, cout, generator=g).to(torch.bfloat16)
    din, dw = conv3d_bf16_bwd(x.view(torch.int16).to(cuda), w.to(cuda), dout.view(torch.int16).to(cuda), grid)
    xd = x.double().requires_grad_(True)
    wd = w.double().requires_grad_(True)
    (ol.conv(xd, wd, grid) * dout.double()).sum().backward()
    assert rel(din.cpu(), xd.grad) < 1e-5
    assert rel(dw.cpu(), wd.grad) < 1e-5


def test_features_match_oracle(cuda):
    grid = (8, 8, 8)
    A, N = device_system(grid, 700, seed=11)
    net = LearnedPreconditioner(seed=0)
    f, d = net.features(A, grid, LAM)
    fo, do = ol.features(N, grid, LAM)
    # BF16 features: equal up to one BF16 rounding step of the FP64 couplings
    diff = (f.cpu().double() - fo).abs()
    assert float(diff.max()) <= 2.0 ** -8 * max(1.0, float(fo.abs().max()))
    assert float((diff > 0).double().mean()) < 0.01
    assert rel(d.cpu(), do) < 1e-6


def test_loss_and_gradient_match_oracle(cuda):
    grid = (8, 8, 8)
    V = 512
    A, N = device_system(grid, 700, seed=12)
    p = ol.init_params(3, zero_output=False) * 0.5
    net = LearnedPreconditioner(seed=0, params=p)
    y = rademacher(4, V, seed=5)
    loss, grad = net.loss_grad(A, grid, LAM, y.to(cuda))
    loss_o, grad_o, g_o, dsq_o = ol.loss_and_grad(p, N, grid, LAM, y.double().numpy())
    assert abs(loss - loss_o) < 2e-3, (loss, loss_o)
    gg = grad.cpu().double().numpy()
    cos = gg @ grad_o / (np.linalg.norm(gg) * np.linalg.norm(grad_o))
    assert cos > 0.999 and rel(gg, grad_o) < 3e-2, (cos, rel(gg, grad_o))
    # the prepared factor and M s
    net.prepare(A, grid, LAM)
    gc = net.coefficients()
    assert rel(gc.cpu(), g_o) < 1e-2
    s = torch.randn(V, generator=torch.Generator().manual_seed(3))
    z = net.apply(s.to(cuda)).cpu()
    _, d = net.features(A, grid, LAM)
    zo = ol.apply_m(gc.cpu().double(), d.cpu().double(), s.double(), grid)
    assert rel(z, zo) < 1e-5


def test_untrained_network_is_jacobi(cuda):
    grid = (16, 16, 16)
    A = tomography_system(grid, 12000, seed=21)
    net = LearnedPreconditioner(seed=1)
    net.prepare(A, grid, LAM)
    g = net.coefficients()
    assert torch.all(g[0] == 1) and torch.all(g[1:] == 0)
    s_true = torch.as_tensor(1.0 / synthetic.smooth_random_velocity(16, 16, 16, seed=3, sigma=2.5),
                             dtype=torch.float64, device=cuda).reshape(-1)
    b = A.matvec(s_true)
    xj, ij, _ = A.cgls(b, LAM, 1e-6, 2000, "jacobi")
    xl, il, _ = A.cgls(b, LAM, 1e-6, 2000, "learned", network=net)
    assert ij.converged and il.converged
    assert abs(il.iterations - ij.iterations) <= 1
    assert rel(xl.cpu(), xj.cpu()) < 1e-5


def test_training_reduces_condition_number_and_iterations(cuda):
    """Train on 32^3 systems of the benchmark recipe (same ray density per
    cell as the 200k-ray 64^3 system), then solve a held-out system.  The
    gain the training objective measures: a lower Kaporin number of the
    preconditioned operator than the untrained network (which is Jacobi).
    CGLS with the trained network converges to the Jacobi solution; its
    iteration count stays within 25 % of Jacobi's (the Kaporin optimum is
    not the iteration-count optimum on these systems, so no fewer
    iterations are promised)."""
    grid = (32, 32, 32)
    nrays = 50000
    net = LearnedPreconditioner(seed=2)
    A_test = tomography_system(grid, nrays, seed=10_000)
    y = rademacher(8, 32 ** 3, seed=99).to(cuda)
    k0, _ = net.loss_grad(A_test, grid, LAM, y)
    log = train(net, ((grid, tomography_system(grid, nrays, seed=s)) for s in range(100)), n_systems=100,
                steps_per_system=2, nprobe=4, lr=2e-3, lr_final=2e-4)
    k1, _ = net.loss_grad(A_test, grid, LAM, y)
    assert all(np.isfinite(log.losses))
    assert k1 < k0 - 0.1, (k0, k1)
    s_true = torch.as_tensor(1.0 / synthetic.smooth_random_velocity(32, 32, 32, seed=4, sigma=5.0),
                             dtype=torch.float64, device=cuda).reshape(-1)
    b = A_test.matvec(s_true)
    xj, ij, _ = A_test.cgls(b, LAM, 1e-6, 4000, "jacobi")
    net.prepare(A_test, grid, LAM)
    xl, il, _ = A_test.cgls(b, LAM, 1e-6, 4000, "learned", network=net)
    print(f"held-out 32^3: log K {k0:.4f} -> {k1:.4f}; CGLS iterations Jacobi {ij.iterations}, "
          f"learned {il.iterations}", flush=True)
    assert ij.converged and il.converged
    assert il.iterations <= 1.25 * ij.iterations
    assert rel(xl.cpu(), xj.cpu()) < 1e-3   # both meet the 1e-6 normal-equation residual
