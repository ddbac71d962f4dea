"""CPU oracle of the learned preconditioner (test infrastructure only).

A float64 torch restatement of what csrc/learned_precond.hip and
csrc/conv3d.hip compute: the coupling features of N = A^T A + lam I, the
3-level convolutional network (with the same BF16 roundings of weights and
activations, or none), the upper-triangular 63-tap factor G, the
preconditioner M = D^-1/2 G G^T D^-1/2 and the Kaporin loss
    log K~ = log(q / (P n)) - (2/n) sum_i log G_ii,  q = sum_y (G y)^T J (G y),
with its gradient by autograd.  The reference has no learned preconditioner
(its only one is Jacobi, src/cpp/sparse/sparse_ops.cpp:52-113), so this is
pinned by the checks in tests/test_oracle_learned.py: a brute-force loop
convolution, the Jacobi limit of an untrained network, finite differences of
the loss, and the exact Kaporin number on small systems.

Layouts follow the library: activations channels-last [cell][channel] with
cell = (x*n1 + y)*n2 + z; weights [cout][27][cin] with tap t = (dx+1)*9 +
(dy+1)*3 + (dz+1); coefficients [63][cell].
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch
import torch.nn.functional as F

LAYERS = ((32, 16), (16, 32), (32, 64), (96, 32), (48, 16), (16, 64))
NG = 63
NFEAT = 32


def bf16(x: torch.Tensor, on: bool = True) -> torch.Tensor:
    """Round to BF16 (nearest even) and back, in the tensor's dtype."""
    return x.to(torch.bfloat16).to(x.dtype) if on else x


def bf16_ste(x: torch.Tensor, on: bool = True) -> torch.Tensor:
    """BF16 rounding with an identity gradient (the library trains FP32 master
    weights through BF16 copies)."""
    return x + (bf16(x.detach(), on) - x.detach()) if on else x


def offsets27() -> np.ndarray:
    return np.array([(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)], dtype=np.int64)


def g_offsets() -> np.ndarray:
    o = np.array([(a, b, c) for a in range(-2, 3) for b in range(-2, 3) for c in range(-2, 3)], dtype=np.int64)
    return o[62:]


def layer_slices():
    out, o = [], 0
    for cin, cout in LAYERS:
        n = cin * cout * 27
        out.append((o, o + n, cin, cout))
        o += n
    return out


def num_params() -> int:
    return layer_slices()[-1][1]


def init_params(seed: int = 0, zero_output: bool = True) -> np.ndarray:
    """He-uniform weights over the real fan-in (numpy generator, not the
    library's own initialisation); output layer zero unless asked otherwise."""
    rng = np.random.default_rng(seed)
    w = np.zeros(num_params(), np.float32)
    for l, (a, b, cin, cout) in enumerate(layer_slices()):
        cin_real = 27 if l == 0 else cin
        bound = np.sqrt(6.0 / (27.0 * cin_real))
        v = rng.uniform(-bound, bound, size=(cout, 27, cin)).astype(np.float32)
        v[:, :, cin_real:] = 0
        if l == len(LAYERS) - 1 and zero_output:
            v[:] = 0
        w[a:b] = v.reshape(-1)
    return w


# ------------------------------------------------------------------ system
def normal_matrix(rp, ci, val, ncols: int, lam: float) -> sp.csr_matrix:
    """N = A^T A + lam I in FP64 from FP32 CSR values (small systems only)."""
    A = sp.csr_matrix((np.asarray(val, np.float32).astype(np.float64), np.asarray(ci), np.asarray(rp)),
                      shape=(len(rp) - 1, ncols))
    return (A.T @ A + lam * sp.eye(ncols)).tocsr()


def features(N: sp.csr_matrix, grid, lam: float, round_bf16: bool = True):
    """[cell, 32] inputs: for each of the 26 neighbours (3x3x3 order, centre
    skipped) N[i][j] / sqrt(N_ii N_jj) (0 outside the grid), then a constant 1
    channel and 5 zero channels; and dsq = 1 / sqrt(N_ii)."""
    n0, n1, n2 = grid
    V = n0 * n1 * n2
    d = N.diagonal()
    Nd = N.toarray() if V <= 8192 else None
    feat = np.zeros((V, NFEAT))
    idx = np.arange(V).reshape(n0, n1, n2)
    x, y, z = np.meshgrid(np.arange(n0), np.arange(n1), np.arange(n2), indexing="ij")
    ch = 0
    for s, (a, b, c) in enumerate(offsets27()):
        if s == 13:
            continue
        xx, yy, zz = x + a, y + b, z + c
        ok = (xx >= 0) & (xx < n0) & (yy >= 0) & (yy < n1) & (zz >= 0) & (zz < n2)
        i = idx[ok]
        j = idx[xx[ok], yy[ok], zz[ok]]
        nij = Nd[i, j] if Nd is not None else np.asarray(N[i, j]).ravel()
        col = np.zeros(V)
        col[i] = nij / np.sqrt(d[i] * d[j])
        feat[:, ch] = col
        ch += 1
    feat[:, 26] = 1.0
    f = torch.tensor(feat)
    if round_bf16:
        f = bf16(f.float()).double()
    return f, torch.tensor(1.0 / np.sqrt(d))


# ------------------------------------------------------------------ network
def _cl_to_t(x: torch.Tensor, grid) -> torch.Tensor:
    n0, n1, n2 = grid
    return x.reshape(n0, n1, n2, -1).permute(3, 0, 1, 2).unsqueeze(0)


def _t_to_cl(x: torch.Tensor) -> torch.Tensor:
    return x[0].permute(1, 2, 3, 0).reshape(-1, x.shape[1])


def conv(x_cl: torch.Tensor, w: torch.Tensor, grid) -> torch.Tensor:
    """3x3x3 zero-padded convolution, channels-last in and out; w [cout][27][cin]."""
    cout, _, cin = w.shape
    wt = w.reshape(cout, 3, 3, 3, cin).permute(0, 4, 1, 2, 3)
    return _t_to_cl(F.conv3d(_cl_to_t(x_cl, grid), wt, padding=1))


def conv_bruteforce(x_cl: np.ndarray, w: np.ndarray, grid) -> np.ndarray:
    """Loop form of conv() with the tap/offset convention spelled out."""
    n0, n1, n2 = grid
    cout = w.shape[0]
    out = np.zeros((n0 * n1 * n2, cout))
    offs = offsets27()
    for p in range(n0 * n1 * n2):
        x, y, z = p // (n1 * n2), (p // n2) % n1, p % n2
        for t, (a, b, c) in enumerate(offs):
            xx, yy, zz = x + a, y + b, z + c
            if 0 <= xx < n0 and 0 <= yy < n1 and 0 <= zz < n2:
                out[p] += w[:, t, :] @ x_cl[(xx * n1 + yy) * n2 + zz]
    return out


def pool(x_cl, grid):
    return _t_to_cl(F.avg_pool3d(_cl_to_t(x_cl, grid), 2))


def upcat(coarse_cl, fine_cl, grid):
    cgrid = tuple(g // 2 for g in grid)
    up = _cl_to_t(coarse_cl, cgrid).repeat_interleave(2, 2).repeat_interleave(2, 3).repeat_interleave(2, 4)
    return torch.cat([_t_to_cl(up), fine_cl], dim=1)


def network(params: torch.Tensor, feat: torch.Tensor, grid, round_bf16: bool = True) -> torch.Tensor:
    """Network output [64][cell] (before the exponential of channel 0)."""
    r = round_bf16
    W = [bf16_ste(params[a:b], r).reshape(cout, 27, cin) for a, b, cin, cout in layer_slices()]
    g1 = tuple(grid)
    g2 = tuple(g // 2 for g in g1)
    g3 = tuple(g // 4 for g in g1)
    a = bf16_ste(torch.relu(conv(feat, W[0], g1)), r)
    pa = bf16_ste(pool(a, g1), r)
    b = bf16_ste(torch.relu(conv(pa, W[1], g2)), r)
    pb = bf16_ste(pool(b, g2), r)
    c = bf16_ste(torch.relu(conv(pb, W[2], g3)), r)
    d = bf16_ste(torch.relu(conv(upcat(c, b, g2), W[3], g2)), r)
    e = bf16_ste(torch.relu(conv(upcat(d, a, g1), W[4], g1)), r)
    return conv(e, W[5], g1).T.contiguous()


def coefficients(out: torch.Tensor) -> torch.Tensor:
    """[63][cell]: diagonal exp(clamp(out0, -30, 30)), the rest as output."""
    return torch.cat([torch.exp(out[:1].clamp(-30, 30)), out[1:NG]], dim=0)


def apply_g(g: torch.Tensor, y: torch.Tensor, grid) -> torch.Tensor:
    """(G y)_i = sum_k g[k][i] y[i + d_k] (zero outside the grid)."""
    n0, n1, n2 = grid
    y3 = F.pad(y.reshape(1, 1, n0, n1, n2), (2, 2, 2, 2, 2, 2))[0, 0]
    out = torch.zeros_like(y)
    for k, (a, b, c) in enumerate(g_offsets()):
        out = out + g[k] * y3[2 + a:2 + a + n0, 2 + b:2 + b + n1, 2 + c:2 + c + n2].reshape(-1)
    return out


def apply_gt(g: torch.Tensor, v: torch.Tensor, grid) -> torch.Tensor:
    """(G^T v)_j = sum_k g[k][j - d_k] v[j - d_k]."""
    n0, n1, n2 = grid
    out = torch.zeros_like(v)
    for k, (a, b, c) in enumerate(g_offsets()):
        w = F.pad((g[k] * v).reshape(1, 1, n0, n1, n2), (2, 2, 2, 2, 2, 2))[0, 0]
        out = out + w[2 - a:2 - a + n0, 2 - b:2 - b + n1, 2 - c:2 - c + n2].reshape(-1)
    return out


def dense_g(g: torch.Tensor, grid) -> torch.Tensor:
    """G as a dense matrix (tiny grids)."""
    V = int(np.prod(grid))
    return torch.stack([apply_g(g, e, grid) for e in torch.eye(V, dtype=g.dtype)], dim=1)


def apply_m(g, dsq, s, grid):
    """z = D^-1/2 G G^T D^-1/2 s."""
    return dsq * apply_g(g, apply_gt(g, dsq * s, grid), grid)


def kaporin_loss(g: torch.Tensor, J, probes: torch.Tensor, grid, diag_exponent: torch.Tensor | None = None):
    """log(q / (P n)) - (2/n) sum log G_ii with q = sum_y (G y)^T J (G y);
    J is a callable x -> J x.  diag_exponent (the network output of channel
    0, clamped) replaces log G_ii when given."""
    n = g.shape[1]
    q = 0.0
    for y in probes:
        u = apply_g(g, y, grid)
        q = q + u @ J(u)
    logd = diag_exponent.sum() if diag_exponent is not None else torch.log(g[0]).sum()
    return torch.log(q / (probes.shape[0] * n)) - 2.0 * logd / n


def kaporin_exact(g: torch.Tensor, Jd: torch.Tensor, grid) -> float:
    """Exact log K(G^T J G) = log(tr(B)/n) - (1/n) log det B (dense, tiny grids)."""
    Gd = dense_g(g, grid)
    B = Gd.T @ Jd @ Gd
    n = B.shape[0]
    return float(torch.log(torch.trace(B) / n) - torch.logdet(B) / n)


def j_operator(N: sp.csr_matrix, dsq: torch.Tensor):
    Nt = torch.tensor(N.toarray())

    def J(x):
        return dsq * (Nt @ (dsq * x))
    return J


def loss_and_grad(params_np: np.ndarray, N: sp.csr_matrix, grid, lam: float, probes: np.ndarray,
                  round_bf16: bool = True):
    """Oracle of stp_lpc_loss_grad: (loss, gradient w.r.t. the FP32 master weights)."""
    feat, dsq = features(N, grid, lam, round_bf16)
    p = torch.tensor(np.asarray(params_np, np.float64), requires_grad=True)
    out = network(p, feat, grid, round_bf16)
    g = coefficients(out)
    loss = kaporin_loss(g, j_operator(N, dsq), torch.tensor(np.asarray(probes, np.float64)), grid,
                        diag_exponent=out[0].clamp(-30, 30))
    loss.backward()
    return float(loss.detach()), p.grad.numpy().copy(), g.detach(), dsq
