"""GPU reverse-time pass (adjoint-state gradient and RTM image) against the
FP64 oracle, finite differences and invariances."""
import numpy as np
import pytest

from oracle import adjoint as oad
from oracle import fd as ofd
from seismic_engine import FDPropagator, synthetic

pytestmark = pytest.mark.gpu

N = (24, 20, 28)
HW = 4
NT = 110


def _setup(cuda, nshots=1, srcs=((12, 10, 3),), sponge="reference"):
    v_true = synthetic.smooth_random_velocity(*N, seed=1, sigma=3.0, vmin=1800, vmax=2600)
    v0 = synthetic.smooth_random_velocity(*N, seed=2, sigma=4.0, vmin=1900, vmax=2500)
    w = synthetic.ricker(25.0, 1e-3, NT)
    rcv = synthetic.surface_receivers(*N, HW, iz=1)
    fd = FDPropagator(*N, order=8, sponge_width=6, nshots=nshots, device=cuda, sponge=sponge)
    fd.set_sources([fd.index(*s) for s in srcs])
    fd.set_receivers(rcv)
    return fd, v_true, v0, w, rcv


def _obs(fd, v_true, w):
    fd.set_velocity(v_true)
    fd.reset()
    return fd.propagate(NT, w)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("sponge", ["reference", "cerjan"])
def test_gradient_matches_fp64_oracle(sponge, cuda):
    fd, v_true, v0, w, rcv = _setup(cuda, sponge=sponge)
    obs = _obs(fd, v_true, w)
    fd.set_velocity(v0)
    J, g, tr = fd.misfit_gradient(NT, w, obs)
    c = ofd.cfg(N, hw=HW, order=8, sponge_width=6, sponge=sponge)
    src = fd.index(12, 10, 3)
    v0h = synthetic.pad_halo(v0, HW)
    obs_np = obs[0].cpu().numpy().astype(np.float64)
    Jo = oad.misfit(c, v0h, w, NT, src, rcv, obs_np)
    _, _, tro = oad.adjoint(c, v0h, w, NT, src, rcv)
    acc, _, _ = oad.adjoint(c, v0h, w, NT, src, rcv, inj=tro - obs_np, scale_c=True, snap_lap=True)
    go = 2 * acc / v0
    assert abs(J - Jo) / Jo < 1e-5
    assert rel(tr[0].cpu().numpy(), tro) < 1e-5
    assert rel(g.cpu().numpy(), go) < 1e-4


def test_checkpoint_interval_does_not_change_results(cuda):
    import torch
    fd, v_true, v0, w, rcv = _setup(cuda)
    obs = _obs(fd, v_true, w)
    fd.set_velocity(v0)
    outs = [fd.reverse_time(NT, w, obs=obs, checkpoint_interval=k) for k in (1, 7, 0, NT)]
    for o in outs[1:]:
        assert torch.equal(o["acc"], outs[0]["acc"])
        assert torch.equal(o["traces"], outs[0]["traces"])
        assert o["misfit"] == outs[0]["misfit"]


def test_gradient_finite_differences_8_directions(cuda):
    import torch
    fd, v_true, v0, w, rcv = _setup(cuda)
    obs = _obs(fd, v_true, w)
    fd.set_velocity(v0)
    J0, g, _ = fd.misfit_gradient(NT, w, obs)
    g = g.double().cpu().numpy()
    rng = np.random.default_rng(2026)
    errs = []
    for k in range(8):
        d = synthetic.smooth_random_velocity(*N, seed=100 + k, sigma=2.0, vmin=-1.0, vmax=1.0)
        eps = 2.0    # m/s on ~2000 m/s
        Js = []
        for sgn in (1, -1):
            fd.set_velocity(v0 + sgn * eps * d)
            Js.append(fd.reverse_time(NT, w, obs=obs, record=False)["misfit"])
        fdv = (Js[0] - Js[1]) / (2 * eps)
        an = float((g * d).sum())
        errs.append(abs(fdv - an) / abs(fdv))
    assert max(errs) < 2e-2, errs


@pytest.mark.parametrize("sponge", ["reference", "cerjan"])
def test_rtm_image_matches_oracle_and_finalize(sponge, cuda):
    import torch
    fd, v_true, v0, w, rcv = _setup(cuda, sponge=sponge)
    data = _obs(fd, v_true, w)
    fd.set_velocity(v0)
    out = fd.reverse_time(NT, w, inject=data, snapshot="field", scale_by_c=False, illumination=True)
    c = ofd.cfg(N, hw=HW, order=8, sponge_width=6, sponge=sponge)
    img_o, ill_o, _ = oad.adjoint(c, synthetic.pad_halo(v0, HW), w, NT, fd.index(12, 10, 3), rcv,
                                  inj=data[0].cpu().numpy().astype(np.float64), scale_c=False, snap_lap=False,
                                  ill=True)
    assert rel(out["acc"][0].cpu().numpy(), img_o) < 1e-4
    assert rel(out["ill"][0].cpu().numpy(), ill_o) < 1e-5


def test_batched_shots_gradient(cuda):
    import torch
    srcs = ((6, 5, 3), (18, 14, 3))
    fd, v_true, v0, w, rcv = _setup(cuda, nshots=2, srcs=srcs)
    obs = _obs(fd, v_true, w)
    fd.set_velocity(v0)
    out = fd.reverse_time(NT, w, obs=obs)
    for s, src in enumerate(srcs):
        f1, _, _, _, _ = _setup(cuda, 1, (src,))
        f1.set_velocity(v0)
        o1 = f1.reverse_time(NT, w, obs=obs[s:s + 1])
        assert rel(out["acc"][s].cpu().numpy(), o1["acc"][0].cpu().numpy().astype(np.float64)) < 1e-6


def test_autograd_function(cuda):
    import torch
    from seismic_engine.autograd import waveform_misfit
    fd, v_true, v0, w, rcv = _setup(cuda)
    obs = _obs(fd, v_true, w)
    v = torch.tensor(v0, dtype=torch.float32, device=cuda, requires_grad=True)
    J, tr = waveform_misfit(v, fd, w, obs, NT)
    J.backward()
    fd.set_velocity(v0)
    J2, g2, _ = fd.misfit_gradient(NT, w, obs)
    assert abs(J.item() - J2) <= 1e-12 * J2
    assert torch.equal(v.grad, g2)


def test_misfit_without_gradient_is_one_forward_pass(cuda):
    """waveform_misfit without a gradient requested runs the forward pass only
    (stp_fd_adjoint with no accumulator): the same misfit and traces as the
    gradient call, and the same misfit as the FP64 oracle."""
    import torch
    from seismic_engine.autograd import waveform_misfit
    fd, v_true, v0, w, rcv = _setup(cuda)
    obs = _obs(fd, v_true, w)
    v = torch.tensor(v0, dtype=torch.float32, device=cuda)
    with torch.no_grad():
        J0, tr0 = waveform_misfit(v.clone().requires_grad_(True), fd, w, obs, NT)
    J1, tr1 = waveform_misfit(v, fd, w, obs, NT)
    vg = v.clone().requires_grad_(True)
    J2, tr2 = waveform_misfit(vg, fd, w, obs, NT)
    J2.backward()
    assert J0.item() == J1.item() == J2.item()
    assert torch.equal(tr0, tr2) and torch.equal(tr1, tr2)
    assert vg.grad is not None and float(vg.grad.abs().max()) > 0
    c = ofd.cfg(N, hw=HW, order=8, sponge_width=6)
    Jo = oad.misfit(c, synthetic.pad_halo(v0, HW), w, NT, fd.index(12, 10, 3), rcv, obs[0].cpu().numpy().astype(np.float64))
    assert abs(J1.item() - Jo) / Jo < 1e-5
