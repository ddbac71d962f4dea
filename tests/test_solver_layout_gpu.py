"""CGLS in the Morton solver layout (stp_csr_set_grid / stp_csr_prepare_layout)
against the same solve on the matrix as given: renumbering cells and rays
changes only the FP32 summation order, so both solves converge in about the
same number of iterations to solutions that agree to the solver tolerance,
and both agree with the FP64 oracle's tightly converged solution."""
import numpy as np
import pytest
import torch

from oracle import sparse as osp
from seismic_engine import synthetic
from seismic_engine.learned import LearnedPreconditioner, make_system
from seismic_engine.sparse import DeviceCSR

pytestmark = pytest.mark.gpu
LAM = 1e-2


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _pair(A, b, precond, tol=1e-6, network=None):
    A.prepare_layout()
    xl, il, _ = A.cgls(b, LAM, tol, 5000, precond, network=network)
    A.set_grid(None)
    xp, ip, _ = A.cgls(b, LAM, tol, 5000, precond, network=network)
    print(f"{precond}: layout {il.iterations} its, plain {ip.iterations} its", flush=True)
    assert il.converged and ip.converged
    assert il.true_rel_residual < tol and ip.true_rel_residual < tol
    assert abs(il.iterations - ip.iterations) <= max(3, 0.05 * ip.iterations)
    return xl.cpu().numpy(), xp.cpu().numpy()


@pytest.mark.parametrize("n", [(24, 24, 24), (20, 12, 7)])
@pytest.mark.parametrize("precond,kind", [("none", 0), ("jacobi", 1)])
def test_layout_solve_matches_plain_solve(n, precond, kind, cuda):
    pts = synthetic.ray_endpoints(6 * int(np.prod(n)) // 5, n, seed=41)
    A = DeviceCSR.from_rays(pts, n)
    rp, ci, v = (t.cpu().numpy() for t in A.export())
    rp, ci, v = rp.astype(np.int64), ci.astype(np.int64), v.astype(np.float64)
    s_true = 1.0 / synthetic.smooth_random_velocity(*n, seed=42, sigma=3.0)
    b = osp.spmv(rp, ci, v, s_true.ravel())
    xl, xp = _pair(A, b, precond)
    xs, _, _, _ = osp.cgls(rp, ci, v, b, A.cols, LAM, 1e-13, 20000, 0)
    xo, _, _, _ = osp.cgls(rp, ci, v, b, A.cols, LAM, 1e-6, 5000, kind)
    e_orc = np.linalg.norm(xo - xs) + 1e-9 * np.linalg.norm(xs)
    assert np.linalg.norm(xl - xs) <= 3.0 * e_orc
    assert np.linalg.norm(xp - xs) <= 3.0 * e_orc


def test_layout_with_learned_preconditioner(cuda):
    sysm = make_system((24, 24, 24), 20_000, seed=43)
    net = LearnedPreconditioner(seed=2)
    net.prepare(sysm.A, sysm.grid, LAM)
    xl, xp = _pair(sysm.A, sysm.b, "learned", network=net)
    assert rel(xl, xp) < 1e-3 and rel(xl, sysm.x_star.cpu()) < 1e-3


def test_layout_with_empty_rows_and_grid_checks(cuda):
    rng = np.random.default_rng(44)
    n = (6, 5, 4)
    V, rows = int(np.prod(n)), 300
    lens = rng.integers(0, 9, rows)
    lens[[0, 7, 8, rows - 1]] = 0                      # empty rows, first and last included
    cols = [np.sort(rng.choice(V, k, replace=False)) for k in lens]
    cols[5] = np.append(cols[5][cols[5] < V - 1], V - 1)
    lens[5] = cols[5].size
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate(cols).astype(np.int32)
    val = rng.uniform(0.1, 1.0, ci.size).astype(np.float32)
    A = DeviceCSR.from_arrays(rp, ci, val)
    assert A.cols == V
    with pytest.raises(RuntimeError):
        A.set_grid((5, 5, 5))
    A.set_grid(n)
    b = rng.standard_normal(rows)
    xl, xp = _pair(A, b, "jacobi", tol=1e-8)
    dense = np.zeros((rows, V))
    for r in range(rows):
        dense[r, ci[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    xd = np.linalg.solve(dense.T @ dense + LAM * np.eye(V), dense.T @ b)
    assert rel(xl, xd) < 1e-5 and rel(xp, xd) < 1e-5
    # a second solve reuses the layout, a new grid rebuilds it
    A.set_grid(n)
    x2, i2, _ = A.cgls(torch.as_tensor(b, device="cuda"), LAM, 1e-8, 5000, "jacobi")
    assert i2.converged and rel(x2.cpu(), xd) < 1e-5
    x3, i3, _ = A.cgls(torch.as_tensor(b, device="cuda"), LAM, 1e-8, 5000, "jacobi")
    assert torch.equal(x2, x3)
    A.set_grid((4, 5, 6))
    x4, i4, _ = A.cgls(torch.as_tensor(b, device="cuda"), LAM, 1e-8, 5000, "jacobi")
    assert i4.converged and rel(x4.cpu(), xd) < 1e-5


@pytest.mark.parametrize("n", [(24, 24, 24), (20, 12, 7)])
def test_products_through_the_layout_match_the_given_order(n, cuda):
    pts = synthetic.ray_endpoints(5 * int(np.prod(n)) // 4, n, seed=45)
    A = DeviceCSR.from_rays(pts, n)
    rng = np.random.default_rng(46)
    x = torch.as_tensor(rng.standard_normal(A.cols), device="cuda")
    u = torch.as_tensor(rng.standard_normal(A.rows), device="cuda")
    plain = [A.matvec(x), A.matvec(u, True), A.matvec(x.float()), A.matvec(u.float(), True)]
    A.prepare_layout()
    lay = [A.matvec(x), A.matvec(u, True), A.matvec(x.float()), A.matvec(u.float(), True)]
    for k, (p, q) in enumerate(zip(plain, lay)):
        p, q = p.double().cpu().numpy(), q.double().cpu().numpy()
        tol = 1e-13 if k < 2 else 1e-6          # FP64 / FP32 accumulation, different summation order
        assert rel(q, p) < tol, (k, rel(q, p))


@pytest.mark.parametrize("cells_per_row,rows,tile", [(0, 0, 64), (5, 20_000, 32), (8, 30_000, 0)])
def test_staged_transpose_tile_heights(cells_per_row, rows, tile, cuda):
    """The staged A^T takes 64-cell tiles; 32-cell tiles when a 64-cell tile
    would read more than 8192 distinct rows; none (global gathers) when even a
    32-cell tile would.  Random rows over an 8^3 grid reach the last two
    cases; every case solves to the dense FP64 normal-equation solution."""
    rng = np.random.default_rng(71)
    if cells_per_row == 0:
        n = (12, 12, 12)
        A = DeviceCSR.from_rays(synthetic.ray_endpoints(3_000, n, seed=72), n)
        rp_, ci_, v_ = (t.cpu().numpy() for t in A.export())
    else:
        n = (8, 8, 8)
        V = int(np.prod(n))
        ci_ = np.concatenate([np.sort(rng.choice(V, cells_per_row, replace=False)) for _ in range(rows)])
        rp_ = np.arange(0, rows * cells_per_row + 1, cells_per_row)
        v_ = rng.uniform(0.1, 1.0, ci_.size).astype(np.float32)
        A = DeviceCSR.from_arrays(rp_.astype(np.int32), ci_.astype(np.int32), v_)
        assert A.cols == V
        A.set_grid(n)
    A.prepare_layout()
    t, nlist = A.staging()
    assert t == tile
    assert (nlist > 0) == (tile > 0)
    V = A.cols
    dense = np.zeros((A.rows, V))
    rp_ = rp_.astype(np.int64)
    for r in range(A.rows):
        dense[r, ci_[rp_[r]:rp_[r + 1]]] = v_[rp_[r]:rp_[r + 1]]
    b = rng.standard_normal(A.rows)
    x, info, _ = A.cgls(torch.as_tensor(b, device="cuda"), LAM, 1e-6, 5000, "jacobi")
    assert info.converged and info.true_rel_residual < 1e-6
    xd = np.linalg.solve(dense.T @ dense + LAM * np.eye(V), dense.T @ b)
    assert rel(x.cpu(), xd) < 1e-4
