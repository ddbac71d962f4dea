"""GPU parity of the FD propagator (libseismic_hip.so through its C ABI)
against the FP64 oracle.  Tolerance: 1e-5 relative L2 for FP32 wavefields and
traces; receiver indexing and step statuses exact."""
import numpy as np
import pytest

from oracle import fd as ofd
from seismic_engine import FDPropagator, synthetic

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _run_pair(n, order, hw, nt, sponge_width, vel, rcv, cuda, src=None, sponge="reference"):
    c = ofd.cfg(n, hw=hw, order=order, sponge_width=sponge_width, sponge=sponge)
    v_h = synthetic.pad_halo(vel, hw)
    w = synthetic.ricker(15.0, 1e-3, nt)
    if src is None:
        src = (n[0] // 2, n[1] // 2, n[2] // 2)
    cur_o, prev_o, tr_o, res = ofd.forward(c, v_h, w, nt, src=src, rcv=rcv)
    assert res.completed_successfully == 1
    fd = FDPropagator(*n, order=order, halo_width=hw, sponge_width=sponge_width, device=cuda, sponge=sponge)
    fd.set_velocity(v_h)
    assert fd.verify_cfl()
    fd.set_sources([fd.index(*src)])
    fd.set_receivers(rcv)
    tr = fd.propagate(nt, w)
    import torch
    torch.cuda.synchronize()
    return fd, (cur_o, prev_o, tr_o), tr


@pytest.mark.parametrize("n,order,hw", [
    ((40, 36, 44), 8, 4),    # float4 path, partial tiles on every axis
    ((3, 30, 28), 8, 4),     # an x chunk shorter than the halo width and the prefetch distance
    ((13, 18, 24), 8, 4),    # one full unrolled pass of 12 planes and a one-plane tail
    ((70, 33, 45), 8, 4),    # scalar path (nz % 4 != 0), two x chunks
    ((30, 20, 25), 2, 1),    # reference 7-point stencil, halo 1
    ((24, 20, 16), 2, 4),    # 7-point stencil on the float4 path
])
def test_forward_matches_oracle(n, order, hw, cuda):
    nt = 120
    vel = synthetic.two_layer_velocity(*n)
    rcv = synthetic.surface_receivers(*n, hw)
    fd, (cur_o, prev_o, tr_o), tr = _run_pair(n, order, hw, nt, 10, vel, rcv, cuda)
    cur = fd.cur[0].cpu().numpy()
    prev = fd.previous_damped()[0].cpu().numpy()
    assert rel_l2(cur, cur_o) < TOL
    assert rel_l2(prev, prev_o) < TOL
    assert rel_l2(tr[0].cpu().numpy(), tr_o) < TOL
    # halo stays zero (Dirichlet), exactly
    hwsl = fd.cur[0].cpu().numpy().copy()
    hwsl[hw:-hw, hw:-hw, hw:-hw] = 0
    assert not hwsl.any()


@pytest.mark.parametrize("n,order,hw", [
    ((48, 40, 36), 8, 4),    # float4 path
    ((38, 30, 25), 8, 4),    # scalar path
    ((30, 20, 25), 2, 1),    # reference 7-point stencil
])
def test_cerjan_sponge_matches_oracle(n, order, hw, cuda):
    """Cerjan sponge on a smooth random 1500-4500 m/s model, long enough for
    the reference form to grow: the FP32 fields and traces follow the FP64
    oracle of the same form and stay bounded."""
    nt = 400
    vel = synthetic.smooth_random_velocity(*n, seed=4, sigma=6.0)
    rcv = synthetic.surface_receivers(*n, hw)
    fd, (cur_o, prev_o, tr_o), tr = _run_pair(n, order, hw, nt, 12, vel, rcv, cuda, sponge="cerjan")
    assert rel_l2(fd.cur[0].cpu().numpy(), cur_o) < TOL
    assert rel_l2(fd.previous_damped()[0].cpu().numpy(), prev_o) < TOL
    assert rel_l2(tr[0].cpu().numpy(), tr_o) < TOL
    early = float(np.abs(tr_o[:, :150]).max())
    assert float(np.abs(tr_o).max()) < 10 * early


def test_receiver_indexing_is_exact(cuda):
    n = (32, 28, 24)
    hw = 4
    rng = np.random.default_rng(7)
    nh = tuple(k + 2 * hw for k in n)
    ix = rng.integers(0, n[0], 300)
    iy = rng.integers(0, n[1], 300)
    iz = rng.integers(0, n[2], 300)
    rcv = (((ix + hw) * nh[1] + iy + hw) * nh[2] + iz + hw).astype(np.int64)
    fd = FDPropagator(*n, order=8, device=cuda)
    fd.set_velocity(synthetic.two_layer_velocity(*n))
    fd.set_sources([fd.index(16, 14, 12)])
    fd.set_receivers(rcv)
    w = synthetic.ricker(15.0, 1e-3, 40)
    tr = fd.propagate(40, w)
    flat = fd.cur[0].reshape(-1)
    import torch
    assert torch.equal(tr[0, :, -1], flat[torch.as_tensor(rcv, device=cuda)])


def test_split_calls_equal_one_call(cuda):
    n = (36, 30, 28)
    vel = synthetic.two_layer_velocity(*n)
    rcv = synthetic.surface_receivers(*n, 4)
    w = synthetic.ricker(15.0, 1e-3, 90)
    outs = []
    for chunks in ((90,), (31, 59), (1, 1, 88)):
        fd = FDPropagator(*n, order=8, sponge_width=8, device=cuda)
        fd.set_velocity(vel)
        fd.set_sources([fd.index(18, 15, 14)])
        fd.set_receivers(rcv)
        trs = [fd.propagate(k, w) for k in chunks]
        import torch
        outs.append((fd.cur.clone(), torch.cat(trs, dim=2)))
    for cur, tr in outs[1:]:
        assert torch.equal(cur, outs[0][0])
        assert torch.equal(tr, outs[0][1])


def test_batched_shots_match_single_shot_oracle(cuda):
    n = (34, 30, 32)
    hw = 4
    nt = 80
    vel = synthetic.two_layer_velocity(*n)
    v_h = synthetic.pad_halo(vel, hw)
    rcv = synthetic.surface_receivers(*n, hw)
    w = synthetic.ricker(15.0, 1e-3, nt)
    srcs = [(10, 10, 10), (20, 15, 16), (5, 25, 28)]
    fd = FDPropagator(*n, order=8, sponge_width=8, nshots=3, device=cuda)
    fd.set_velocity(v_h)
    fd.set_sources([fd.index(*s) for s in srcs])
    fd.set_receivers(rcv)
    wb = np.stack([w, 0.5 * w, -w])
    tr = fd.propagate(nt, wb).cpu().numpy()
    c = ofd.cfg(n, hw=hw, order=8, sponge_width=8)
    for s, (src, scale) in enumerate(zip(srcs, (1.0, 0.5, -1.0))):
        cur_o, _, tr_o, _ = ofd.forward(c, v_h, scale * w, nt, src=src, rcv=rcv)
        assert rel_l2(fd.cur[s].cpu().numpy(), cur_o) < TOL
        assert rel_l2(tr[s], tr_o) < TOL


def test_smooth_random_velocity_96(cuda):
    n = (96, 96, 96)
    nt = 300
    vel = synthetic.smooth_random_velocity(*n, seed=2026, sigma=20.0 * 96 / 512)
    rcv = synthetic.line_receivers(*n, 4, iy=48)
    fd, (cur_o, _, tr_o), tr = _run_pair(n, 8, 4, nt, 20, vel, rcv, cuda)
    assert rel_l2(fd.cur[0].cpu().numpy(), cur_o) < TOL
    assert rel_l2(tr[0].cpu().numpy(), tr_o) < TOL


def test_nan_guards_and_status(cuda):
    import torch
    n = (20, 20, 20)
    hw = 4
    c = ofd.cfg(n, hw=hw, order=8, sponge_width=0)
    nh = ofd.nh(c)
    rng = np.random.default_rng(11)
    u = np.zeros(nh)
    u[hw:-hw, hw:-hw, hw:-hw] = rng.standard_normal(n)
    p = u * 0.9
    u[10, 10, 10] = np.nan
    p[14, 12, 9] = np.inf
    v_h = np.full(nh, 2000.0)
    v_h[8, 8, 8] = np.nan       # velocity substitution -> 1500
    v_h[9, 9, 9] = -3.0
    cur_o, prev_o = u.copy(), p.copy()
    s_o, _ = ofd.run(c, v_h, cur_o, prev_o, 3)
    fd = FDPropagator(*n, order=8, sponge_width=0, device=cuda)
    fd.set_velocity(v_h)
    fd.cur[0] = torch.as_tensor(u, dtype=torch.float32)
    fd.prev[0] = torch.as_tensor(p, dtype=torch.float32)
    _, st = fd.propagate(3, status=True)
    st = st.cpu().numpy()
    assert s_o == 1
    assert st[0] == 1 and st[1] == 1          # NaN centre, then NaN as previous
    got = fd.cur[0].cpu().numpy()
    assert np.isfinite(got).all()
    assert rel_l2(got, cur_o) < TOL


def test_invalid_arguments_raise(cuda):
    from seismic_engine import SeismicError
    fd = FDPropagator(16, 16, 16, order=8, device=cuda)
    with pytest.raises(SeismicError):
        fd.set_receivers([0])                 # halo cell, not interior
    with pytest.raises(SeismicError):
        fd.set_sources([fd.index(1, 1, 1), fd.index(2, 2, 2)])   # one per shot
    with pytest.raises(SeismicError):
        FDPropagator(16, 16, 16, order=4, device=cuda)


# Entries of STP_FD_VARIANTS in csrc/fd_kernels.hip (tile shape, prefetch
# distance, early or late halos, non-temporal flags); STP_FD_TILE selects one
# at plan creation.
FD_VARIANTS = range(4)


def _plan_with_variant(variant, *args, **kw):
    import os
    old = os.environ.pop("STP_FD_TILE", None)
    if variant is not None:
        os.environ["STP_FD_TILE"] = str(variant)
    try:
        return FDPropagator(*args, **kw)
    finally:
        os.environ.pop("STP_FD_TILE", None)
        if old is not None:
            os.environ["STP_FD_TILE"] = old


def _run_fd(cuda, n, nt, chunks, variant, nshots=1, seed=5):
    import torch
    fd = _plan_with_variant(variant, *n, order=8, sponge_width=6, nshots=nshots, device=cuda, sponge="cerjan")
    vel = synthetic.smooth_random_velocity(*n, seed=seed, sigma=3.0, vmin=1500.0, vmax=3500.0)
    fd.set_velocity(vel)
    srcs = [fd.index(n[0] // 2 + s, n[1] // 3, n[2] // 2 - s) for s in range(nshots)]
    fd.set_sources(srcs)
    rcv = synthetic.line_receivers(*n, 4, iy=n[1] // 2, step=3)
    rcv = np.concatenate([rcv, [srcs[0]]])          # one receiver on a source cell
    fd.set_receivers(rcv)
    w = synthetic.ricker(20.0, 1e-3, nt)
    trs = [fd.propagate(k, w) for k in chunks]
    return fd.cur.clone(), fd.prev.clone(), torch.cat(trs, dim=2)


@pytest.mark.parametrize("n,chunks,nshots", [
    ((70, 45, 68), (40,), 1),          # one x chunk of 70 planes: the ring's guarded tail
    ((70, 45, 68), (6, 11, 23), 1),    # the same steps split over three calls
    ((36, 30, 28), (18,), 2),          # two shots (batched plans keep C cached)
    ((300, 36, 136), (12,), 1),        # x chunks of 256 and 44 planes, both with tails
    ((3, 30, 28), (9,), 1),            # a chunk shorter than the halo width and every ring
])
def test_tile_variants_agree_bit_for_bit(n, chunks, nshots, cuda):
    """Every tile variant (tile shape, loads one to three planes ahead, halos
    with the queue front or with the centre plane, non-temporal accesses)
    does the same arithmetic per cell: fields and traces are bit-identical,
    on grids whose x chunks end in partial ring passes."""
    import torch
    nt = sum(chunks)
    ref = _run_fd(cuda, n, nt, chunks, None, nshots=nshots)
    for v in FD_VARIANTS:
        got = _run_fd(cuda, n, nt, chunks, v, nshots=nshots)
        for a, b in zip(ref, got):
            assert torch.equal(a, b), f"variant {v}"


@pytest.mark.parametrize("variant", FD_VARIANTS)
def test_tile_variants_guards_and_status(variant, cuda):
    """Non-finite values: every variant applies the reference's guards and
    reports the same step statuses as the oracle's step-by-step run."""
    import torch
    n = (20, 20, 20)
    hw = 4
    c = ofd.cfg(n, hw=hw, order=8, sponge_width=0)
    nh = ofd.nh(c)
    rng = np.random.default_rng(12)
    u = np.zeros(nh)
    u[hw:-hw, hw:-hw, hw:-hw] = rng.standard_normal(n)
    p = u * 0.9
    u[10, 10, 10] = np.nan
    p[14, 12, 9] = np.inf
    v_h = np.full(nh, 2000.0)
    cur_o, prev_o = u.copy(), p.copy()
    s_o, _ = ofd.run(c, v_h, cur_o, prev_o, 4)
    fd = _plan_with_variant(variant, *n, order=8, sponge_width=0, device=cuda)
    fd.set_velocity(v_h)
    fd.cur[0] = torch.as_tensor(u, dtype=torch.float32)
    fd.prev[0] = torch.as_tensor(p, dtype=torch.float32)
    _, st = fd.propagate(4, status=True)
    st = st.cpu().numpy()
    assert s_o == 1
    assert st[0] == 1 and st[1] == 1
    got = fd.cur[0].cpu().numpy()
    assert np.isfinite(got).all()
    assert rel_l2(got, cur_o) < TOL


def _with_sweep(sweep, fn):
    import os
    old = os.environ.pop("STP_FD_SWEEP", None)
    if sweep is not None:
        os.environ["STP_FD_SWEEP"] = sweep
    try:
        return fn()
    finally:
        os.environ.pop("STP_FD_SWEEP", None)
        if old is not None:
            os.environ["STP_FD_SWEEP"] = old


@pytest.mark.parametrize("n,chunks,nshots", [
    ((300, 36, 136), (7,), 1),         # two x chunks with tails, an odd step count
    ((70, 45, 68), (6, 11, 23), 1),    # calls that start on odd and even steps
    ((36, 30, 28), (9,), 2),           # batched shots
    ((3, 30, 28), (5,), 1),            # a chunk shorter than the halo width
])
def test_alternating_sweeps_equal_one_way_sweeps(n, chunks, nshots, cuda):
    """Steps march x upward and downward by turns (so each starts on the
    planes the previous one left in the caches); the x taps commute, so the
    fields and traces equal those of steps that always march upward, bit for
    bit."""
    import torch
    nt = sum(chunks)
    alt = _with_sweep(None, lambda: _run_fd(cuda, n, nt, chunks, None, nshots=nshots))
    up = _with_sweep("up", lambda: _run_fd(cuda, n, nt, chunks, None, nshots=nshots))
    for a, b in zip(alt, up):
        assert torch.equal(a, b)


def test_alternating_sweeps_in_the_adjoint_passes(cuda):
    """The checkpointed forward pass, the snapshot recompute and the reverse
    step alternate their sweeps as well: image, illumination and traces
    equal those of one-way sweeps, bit for bit."""
    import torch
    n, nt = (40, 28, 32), 37
    vel = synthetic.smooth_random_velocity(*n, seed=3, sigma=3.0, vmin=1800.0, vmax=2600.0)
    w = synthetic.ricker(25.0, 1e-3, nt)

    def run():
        fd = FDPropagator(*n, order=8, sponge_width=6, device=cuda, sponge="cerjan")
        fd.set_velocity(vel)
        fd.set_sources([fd.index(n[0] // 2, n[1] // 2, 3)])
        fd.set_receivers(synthetic.surface_receivers(*n, 4, iz=1))
        obs = fd.propagate(nt, w).clone()
        fd.reset()
        rtm = fd.reverse_time(nt, w, inject=obs, snapshot="field", illumination=True)   # image + illumination
        grad = fd.reverse_time(nt, w, obs=obs * 0.5)                                       # residual injection
        tensors = [o[k].clone() for o in (rtm, grad) for k in ("acc", "ill", "traces") if o[k] is not None]
        return tensors, grad["misfit"]

    (alt, m_alt), (up, m_up) = _with_sweep(None, run), _with_sweep("up", run)
    assert len(alt) == len(up) == 5
    for a, b in zip(alt, up):
        assert torch.equal(a, b)
    assert m_alt == m_up
