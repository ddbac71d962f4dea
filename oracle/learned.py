"""CPU oracle of the learned preconditioner (test infrastructure only).

A float64 torch restatement of what csrc/learned_precond.hip and
csrc/conv3d.hip compute: the coupling features of N = A^T A + lam I, the
3-level convolutional network (with the same BF16 roundings of weights and
activations, or none), the coefficients s_i = exp(o0_i)/sqrt(N_ii),
w_e = exp(o_e) of the preconditioner M = S (I + L_w) S (L_w the weighted
Laplacian of the 7-point grid graph), its application, and the unrolled-CG
training loss with its gradient by autograd.  The reference has no learned
preconditioner (its only one is Jacobi, src/cpp/sparse/sparse_ops.cpp:52-113),
so this is pinned by the checks in tests/test_oracle_learned.py: a
brute-force loop convolution, an explicit edge-by-edge matrix of M and the
Kronecker-product grid Laplacian of an untrained network, symmetry and
positive definiteness, the energy-norm identity of CG, and finite
differences of the loss.

Layouts follow the library: activations channels-last [cell][channel] with
cell = (x*n1 + y)*n2 + z; weights [cout][27][cin] with tap t = (dx+1)*9 +
(dy+1)*3 + (dz+1); coefficients [4][cell] (s, w_x, w_y, w_z; the weight of
cell i belongs to the edge (i, i + e_axis) and is 0 at the far face).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import torch
import torch.nn.functional as F

LAYERS = ((32, 16), (16, 32), (32, 64), (96, 32), (48, 16), (16, 16))
NCOEF = 4
NFEAT = 32
W0 = 1.0          # edge weight of an untrained network
OCLAMP = 30.0


def bf16(x: torch.Tensor, on: bool = True) -> torch.Tensor:
    """Round to BF16 (nearest even) and back, in the tensor's dtype."""
    return x.to(torch.bfloat16).to(x.dtype) if on else x


def bf16_ste(x: torch.Tensor, on: bool = True) -> torch.Tensor:
    """BF16 rounding with an identity gradient (the library trains FP32 master
    weights through BF16 copies)."""
    return x + (bf16(x.detach(), on) - x.detach()) if on else x


def offsets27() -> np.ndarray:
    return np.array([(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)], dtype=np.int64)


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
    """Network output [16][cell] (channels 0..3 used)."""
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


# ------------------------------------------------------------------ M = S (I + L_w) S
def coefficients(out: torch.Tensor, dsq: torch.Tensor, grid) -> torch.Tensor:
    """[4][cell]: s = dsq exp(o0), w_d = W0 exp(o_{1+d}) where cell + e_d is
    inside the grid, else 0 (outputs clamped to +-30 first)."""
    n0, n1, n2 = grid
    o = out[:NCOEF].clamp(-OCLAMP, OCLAMP)
    x, y, z = np.meshgrid(np.arange(n0), np.arange(n1), np.arange(n2), indexing="ij")
    inside = [torch.tensor((x + 1 < n0).ravel()), torch.tensor((y + 1 < n1).ravel()), torch.tensor((z + 1 < n2).ravel())]
    rows = [dsq * torch.exp(o[0])]
    for d in range(3):
        rows.append(torch.where(inside[d], W0 * torch.exp(o[1 + d]), torch.zeros_like(o[1 + d])))
    return torch.stack(rows)


def apply_m(coef: torch.Tensor, r: torch.Tensor, grid) -> torch.Tensor:
    """z = S (I + L_w) S r: (L_w y)_i = sum over the edges (i, j) of
    w_ij (y_i - y_j)."""
    n0, n1, n2 = grid
    y = (coef[0] * r).reshape(n0, n1, n2)
    u = y.clone()
    for d in range(3):
        w = coef[1 + d].reshape(n0, n1, n2)
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[d] = slice(0, (n0, n1, n2)[d] - 1)
        hi[d] = slice(1, None)
        f = w[tuple(lo)] * (y[tuple(lo)] - y[tuple(hi)])
        pad_hi = [0] * 6
        pad_hi[2 * (2 - d) + 1] = 1
        pad_lo = [0] * 6
        pad_lo[2 * (2 - d)] = 1
        u = u + F.pad(f, pad_hi) - F.pad(f, pad_lo)
    return coef[0] * u.reshape(-1)


def dense_m(coef: torch.Tensor, grid) -> np.ndarray:
    """M as a dense matrix, assembled edge by edge (tiny grids)."""
    n0, n1, n2 = grid
    V = n0 * n1 * n2
    c = coef.detach().numpy()
    P = np.eye(V)
    st = (n1 * n2, n2, 1)
    for i in range(V):
        for d in range(3):
            w = c[1 + d, i]
            if w == 0.0:
                continue
            j = i + st[d]
            P[i, i] += w
            P[j, j] += w
            P[i, j] -= w
            P[j, i] -= w
    return c[0][:, None] * P * c[0][None, :]


def pair_output_gradient(out: torch.Tensor, dsq: torch.Tensor, grid, r, g) -> torch.Tensor:
    """d<g, M r>/d(network output) [16][cell] (what stp_lpc_backward_accumulate
    adds up, through the exponentials)."""
    o = out.detach().clone().requires_grad_(True)
    val = torch.as_tensor(g, dtype=torch.float64) @ apply_m(coefficients(o, dsq, grid),
                                                           torch.as_tensor(r, dtype=torch.float64), grid)
    val.backward()
    return o.grad


def unrolled_cg_loss(coef: torch.Tensor, N: torch.Tensor, rhs: torch.Tensor, x_star: torch.Tensor, grid,
                     k1: int, k2: int) -> torch.Tensor:
    """Preconditioned CG from x = 0 for k2 iterations; the mean over k in
    [k1, k2] of 0.5 log(e_k . r_k / x_star . rhs), e_k = x_star - x_k."""
    x = torch.zeros_like(rhs)
    r = rhs.clone()
    e0 = x_star @ rhs
    z = apply_m(coef, r, grid)
    p = z
    rz = r @ z
    terms = []
    for k in range(k2):
        q = N @ p
        a = rz / (p @ q)
        x = x + a * p
        r = r - a * q
        if k + 1 >= k1:
            terms.append(0.5 * torch.log(((x_star - x) @ r) / e0))
        if k + 1 == k2:
            break
        z = apply_m(coef, r, grid)
        rz2 = r @ z
        p = z + (rz2 / rz) * p
        rz = rz2
    return torch.stack(terms).mean()


def loss_and_grad(params_np: np.ndarray, N: sp.csr_matrix, grid, lam: float, rhs, x_star, k1: int, k2: int,
                  round_bf16: bool = True):
    """Oracle of one training step: (unrolled-CG loss, gradient with respect
    to the FP32 master weights, coefficients, dsq)."""
    feat, dsq = features(N, grid, lam, round_bf16)
    p = torch.tensor(np.asarray(params_np, np.float64), requires_grad=True)
    out = network(p, feat, grid, round_bf16)
    coef = coefficients(out, dsq, grid)
    loss = unrolled_cg_loss(coef, torch.tensor(N.toarray()), torch.as_tensor(rhs, dtype=torch.float64),
                            torch.as_tensor(x_star, dtype=torch.float64), grid, k1, k2)
    loss.backward()
    return float(loss.detach()), p.grad.numpy().copy(), coef.detach(), dsq
