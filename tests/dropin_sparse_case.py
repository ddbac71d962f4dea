"""Inputs and checks of the C++ CSRMatrix / SparseOps case that
tests/cpp/dropin_driver.cpp runs as its `sparse` command: a 40 x 30 matrix
from triplets with out-of-range rows and columns, both products, the
directional derivative, the 6^3 7-point FD matrix scaled by -1 and shifted
by 6 I (CG and Jacobi-PCG), and Jacobi CGLS on the rectangular matrix.
Shared by the GPU test and the CPU stand-in test of the C++ layer."""
import numpy as np

from oracle import sparse as osp

ROWS, COLS, M, FDN = 40, 30, 900, 6


def write_inputs(d):
    rng = np.random.default_rng(3)
    r = rng.integers(-2, ROWS + 2, M)
    c = rng.integers(-2, COLS + 2, M)
    v = rng.standard_normal(M)
    r.astype(np.int64).tofile(d / "r.i64")
    c.astype(np.int64).tofile(d / "c.i64")
    v.tofile(d / "v.f64")
    x = rng.standard_normal(COLS)
    u = rng.standard_normal(ROWS)
    x.tofile(d / "x.f64")
    u.tofile(d / "u.f64")
    fdvel = rng.uniform(1500, 2500, FDN ** 3)
    fdvel.tofile(d / "fdvel.f64")
    (d / "params.txt").write_text(f"rows {ROWS}\ncols {COLS}\nfdn {FDN}\n")
    return dict(r=r, c=c, v=v, x=x, u=u, fdvel=fdvel)


def _f64(d, name):
    return np.fromfile(d / name, dtype=np.float64)


def _rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def check_outputs(d, text, inp):
    """The driver's output files against the oracle's assembly and dense
    solves.  Products carry FP32 matrix values (rel 1e-6); CG and Jacobi-PCG
    to 1e-10 against a dense solve (rel 1e-6); CGLS to 1e-8 against the damped
    normal equations (rel 1e-5)."""
    rp_o, ci_o, val_o = osp.csr_from_triplets(ROWS, COLS, inp["r"], inp["c"], inp["v"])
    assert np.array_equal(np.fromfile(d / "rp.i64", np.int64), rp_o)
    assert np.array_equal(np.fromfile(d / "ci.i64", np.int64), ci_o)
    assert np.array_equal(_f64(d, "val.f64"), val_o)
    dense = np.zeros((ROWS, COLS))
    for i in range(ROWS):
        dense[i, ci_o[rp_o[i]:rp_o[i + 1]]] = val_o[rp_o[i]:rp_o[i + 1]]
    assert f"nnz {ci_o.size} " in text
    assert _rel(_f64(d, "y.f64"), dense @ inp["x"]) < 1e-6
    assert _rel(_f64(d, "yt.f64"), dense.T @ inp["u"]) < 1e-6
    assert _rel(_f64(d, "dd.f64"), (dense @ inp["x"])[: ROWS - 5]) < 1e-6
    dot = float(text.split("dot ")[1].split()[0])
    assert abs(dot - inp["x"] @ inp["x"]) <= 1e-13 * (inp["x"] @ inp["x"])
    frp = np.fromfile(d / "fd_rp.i64", np.int64)
    fci = np.fromfile(d / "fd_ci.i64", np.int64)
    fval = _f64(d, "fd_val.f64")
    N = FDN ** 3
    F = np.zeros((N, N))
    for i in range(N):
        F[i, fci[frp[i]:frp[i + 1]]] = fval[frp[i]:frp[i + 1]]
    assert np.allclose(np.diag(F), 2 - 2 * inp["fdvel"] ** 2 * 1e-6 * 3 / 100, rtol=1e-12)
    assert frp[-1] == 7 * N - 6 * FDN ** 2     # 7 slots per cell less the missing neighbours
    H = 6 * np.eye(N) - F
    assert np.allclose(_f64(d, "h_val.f64"), -fval + 6 * (fci == np.repeat(np.arange(N), np.diff(frp))))
    xs = np.linalg.solve(H, np.ones(N))
    assert "cg iters" in text and "converged 1 pcg" in text
    assert _rel(_f64(d, "cg.f64"), xs) < 1e-6
    assert _rel(_f64(d, "pcg.f64"), xs) < 1e-6
    xl = np.linalg.solve(dense.T @ dense + 1e-2 * np.eye(COLS), dense.T @ inp["u"])
    assert "cgls iters" in text and "converged 1" in text.split("cgls iters")[1]
    assert _rel(_f64(d, "cgls.f64"), xl) < 1e-5
