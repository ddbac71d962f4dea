"""stp_cgls_noalloc (the CGLS solve on the library's static device workspace,
for callers that cannot allocate) against stp_cgls on the same systems:
identical iterates, so x, the residual history and the solve record must be
equal bit for bit.  Two no-allocation calls run back to back, the second on a
smaller system, so the second sees the workspace the first left behind."""
import ctypes as C

import numpy as np
import pytest

from seismic_engine import synthetic
from seismic_engine._lib import SeismicError, SolveInfo, call, ptr, stream_handle
from seismic_engine.learned import DEFAULT_WEIGHTS, LearnedPreconditioner
from seismic_engine.sparse import DeviceCSR

pytestmark = pytest.mark.gpu

MAX_ITER = 5000


def _system(n, nrays, seed, cuda):
    import torch
    pts = synthetic.ray_endpoints(nrays, n, seed=seed)
    A = DeviceCSR.from_rays(pts, n)
    s = torch.as_tensor(1.0 / synthetic.smooth_random_velocity(*n, seed=seed + 1, sigma=2.0), dtype=torch.float64,
                        device=cuda).reshape(-1)
    return A, A.matvec(s)


def _solve(name, A, b, lam, tol, kind, net=None):
    import torch
    x = torch.full((A.cols,), float("nan"), dtype=torch.float64, device=b.device)
    hist = torch.zeros(MAX_ITER, dtype=torch.float64, device=b.device)
    info = SolveInfo()
    call(name, A._h, ptr(b), C.c_double(lam), C.c_double(tol), MAX_ITER, kind, None if net is None else net.handle,
         ptr(x), ptr(hist), C.byref(info), stream_handle())
    torch.cuda.synchronize()
    return x, info, hist[: info.iterations].clone()


def _same(a, b):
    import torch
    xa, ia, ha = a
    xb, ib, hb = b
    assert torch.equal(xa, xb)
    assert torch.equal(ha, hb)
    assert (ia.iterations, ia.converged, ia.replacements) == (ib.iterations, ib.converged, ib.replacements)
    assert ia.rel_residual == ib.rel_residual and ia.true_rel_residual == ib.true_rel_residual


@pytest.mark.parametrize("layout", [True, False])
def test_noalloc_cgls_matches_cgls(layout, cuda):
    big = _system((24, 20, 28), 9000, 71, cuda)
    small = _system((12, 12, 12), 3000, 73, cuda)
    cases = []
    for A, b in (big, small):
        if layout:
            A.prepare_layout()
        else:
            A.set_grid(None)
        for lam, tol, kind in ((1e-2, 1e-6, 1), (0.0, 1e-10, 1), (1e-2, 1e-6, 0)):
            cases.append((A, b, lam, tol, kind, _solve("stp_cgls", A, b, lam, tol, kind)))
    # no-allocation calls back to back: the big system's three solves, then the small one's
    for A, b, lam, tol, kind, ref in cases:
        got = _solve("stp_cgls_noalloc", A, b, lam, tol, kind)
        assert got[1].converged
        _same(got, ref)


def test_noalloc_cgls_learned_preconditioner(cuda):
    net = LearnedPreconditioner.load(DEFAULT_WEIGHTS)
    for n, nrays, seed in (((24, 20, 28), 9000, 81), ((12, 12, 12), 3000, 83)):
        A, b = _system(n, nrays, seed, cuda)
        A.prepare_layout()
        net.prepare(A, n, 1e-2)
        ref = _solve("stp_cgls", A, b, 1e-2, 1e-6, 2, net)
        got = _solve("stp_cgls_noalloc", A, b, 1e-2, 1e-6, 2, net)
        assert got[1].converged
        _same(got, ref)


def test_noalloc_cgls_rejects_what_needs_allocation(cuda):
    import torch
    # a cell grid without its solver layout: building the layout allocates
    A, b = _system((12, 12, 12), 3000, 91, cuda)
    x = torch.empty(A.cols, dtype=torch.float64, device=cuda)
    info = SolveInfo()
    with pytest.raises(SeismicError, match="stp_csr_prepare_layout"):
        call("stp_cgls_noalloc", A._h, ptr(b), C.c_double(1e-2), C.c_double(1e-6), 100, 1, None, ptr(x), None,
             C.byref(info), stream_handle())
    # more columns than the static workspace holds (2^20)
    n = (128, 128, 65)
    pts = np.array([[[0.0, 0.5, 0.5], [128.0, 0.5, 0.5]]])
    Abig = DeviceCSR.from_rays(pts, n)
    Abig.set_grid(None)
    bb = torch.ones(Abig.rows, dtype=torch.float64, device=cuda)
    xb = torch.empty(Abig.cols, dtype=torch.float64, device=cuda)
    with pytest.raises(SeismicError, match="at most 1048576"):
        call("stp_cgls_noalloc", Abig._h, ptr(bb), C.c_double(1e-2), C.c_double(1e-6), 100, 1, None, ptr(xb), None,
             C.byref(info), stream_handle())
