"""Learned preconditioner for the tomography normal equations (stp_lpc_* in
libseismic_hip.so).

The reference preconditions CG with Jacobi only
(src/cpp/sparse/sparse_ops.cpp:52-113); that stays the fallback and the judge.
Here a 3-level 3-D convolutional network (16/32/64 channels, 3x3x3, ReLU,
BF16 on MFMA with FP32 accumulation) maps the Jacobi-scaled local couplings of
N = A^T A + lam I to an upper-triangular 63-tap factor G, and CGLS runs with
M = D^-1/2 G G^T D^-1/2.  Untrained, G = I and M is Jacobi.  Training runs on
the GPU through the library's own forward and backward kernels; it minimises
Kaporin's condition number of G^T J G, estimated from Rademacher probes.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from ._lib import call, lib, ptr, stream_handle

NG = 63            # taps of G
NFEAT = 32         # network input channels
LAYERS = ((32, 16), (16, 32), (32, 64), (96, 32), (48, 16), (16, 64))   # (cin, cout): e1 e2 e3 d2 d1 o


def _torch():
    import torch
    return torch


def g_offsets() -> np.ndarray:
    """(63, 3) offsets d_k of the taps of G: the (62 + k)-th offset of
    [-2,2]^3 in lexicographic (x, y, z) order; row 0 is the diagonal."""
    o = np.array([(a, b, c) for a in range(-2, 3) for b in range(-2, 3) for c in range(-2, 3)], dtype=np.int64)
    return o[62:]


def rademacher(nprobe: int, ncell: int, seed: int, device=None):
    """Probe vectors with entries +-1 (E[y y^T] = I), [nprobe, ncell] FP32."""
    torch = _torch()
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    y = torch.randint(0, 2, (nprobe, ncell), generator=g, dtype=torch.int8).to(torch.float32) * 2 - 1
    return y.to(device) if device is not None else y


class LearnedPreconditioner:
    """Handle of one network (stp_lpc).  `grid` is the cell grid (n0, n1, n2)
    of the matrices it is used with (A's columns, x-major)."""

    def __init__(self, seed: int = 0, params: np.ndarray | None = None):
        h = C.c_void_p()
        call("stp_lpc_create", C.c_uint64(seed), C.byref(h))
        self.handle = h
        n = C.c_int64()
        call("stp_lpc_num_params", C.byref(n))
        self.num_params = n.value
        self.steps = 0
        if params is not None:
            self.set_params(params)

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib().stp_lpc_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    # ---------------------------------------------------------------- weights
    def get_params(self) -> np.ndarray:
        w = np.empty(self.num_params, dtype=np.float32)
        call("stp_lpc_get_params", self.handle, w.ctypes.data, C.c_int64(w.size))
        return w

    def set_params(self, w) -> None:
        w = np.ascontiguousarray(np.asarray(w, dtype=np.float32).reshape(-1))
        call("stp_lpc_set_params", self.handle, w.ctypes.data, C.c_int64(w.size))
        self.steps = 0

    def save(self, path) -> None:
        np.savez(path, params=self.get_params(), layers=np.asarray(LAYERS), steps=self.steps)

    @classmethod
    def load(cls, path) -> "LearnedPreconditioner":
        with np.load(path, allow_pickle=False) as z:
            net = cls(params=z["params"])
            net.steps = int(z["steps"])
        return net

    # ---------------------------------------------------------------- one system
    @staticmethod
    def _grid(grid) -> np.ndarray:
        return np.ascontiguousarray(np.asarray(grid, dtype=np.int64).reshape(3))

    def features(self, A, grid, lam: float):
        """Network input [cell, 32] (FP32 copy of the BF16 features) and
        1/sqrt(N_ii) [cell]."""
        torch = _torch()
        n = self._grid(grid)
        V = int(np.prod(n))
        f = torch.empty(V, NFEAT, dtype=torch.float32, device="cuda")
        d = torch.empty(V, dtype=torch.float32, device="cuda")
        call("stp_lpc_features", self.handle, A._h, n.ctypes.data, C.c_double(lam), ptr(f), ptr(d), stream_handle())
        return f, d

    def prepare(self, A, grid, lam: float) -> None:
        n = self._grid(grid)
        call("stp_lpc_prepare", self.handle, A._h, n.ctypes.data, C.c_double(lam), stream_handle())
        self._prepared_cells = int(np.prod(n))

    def coefficients(self):
        """[63, cell] coefficients of G for the prepared system."""
        torch = _torch()
        g = torch.empty(NG, self._prepared_cells, dtype=torch.float32, device="cuda")
        call("stp_lpc_coefficients", self.handle, ptr(g), stream_handle())
        return g

    def apply(self, s):
        """z = M s for the prepared system."""
        torch = _torch()
        s = s.to(device="cuda", dtype=torch.float32).contiguous()
        z = torch.empty_like(s)
        call("stp_lpc_apply", self.handle, ptr(s), ptr(z), stream_handle())
        return z

    def loss_grad(self, A, grid, lam: float, probes):
        torch = _torch()
        n = self._grid(grid)
        probes = probes.to(device="cuda", dtype=torch.float32).contiguous()
        loss = C.c_double()
        g = torch.empty(self.num_params, dtype=torch.float32, device="cuda")
        call("stp_lpc_loss_grad", self.handle, A._h, n.ctypes.data, C.c_double(lam), ptr(probes),
             int(probes.shape[0]), C.byref(loss), ptr(g), stream_handle())
        return loss.value, g

    def train_step(self, A, grid, lam: float, probes, lr: float) -> float:
        torch = _torch()
        n = self._grid(grid)
        probes = probes.to(device="cuda", dtype=torch.float32).contiguous()
        loss = C.c_double()
        call("stp_lpc_train_step", self.handle, A._h, n.ctypes.data, C.c_double(lam), ptr(probes),
             int(probes.shape[0]), C.c_double(lr), C.byref(loss), stream_handle())
        self.steps += 1
        return loss.value


# -------------------------------------------------------------------- single layers (tests, tuning)
def conv3d_bf16(x_bits, w, grid, relu: bool = False):
    """One 3x3x3 convolution: x_bits [cell, cin] int16 BF16 bit patterns,
    w [cout, 27, cin] FP32 (rounded to BF16 inside) -> [cell, cout] FP32."""
    torch = _torch()
    n0, n1, n2 = (int(v) for v in grid)
    cout, _, cin = w.shape
    out = torch.empty(n0 * n1 * n2, cout, dtype=torch.float32, device="cuda")
    call("stp_conv3d_bf16", C.c_int64(n0), C.c_int64(n1), C.c_int64(n2), int(cin), int(cout),
         ptr(x_bits.contiguous()), ptr(w.float().contiguous()), int(relu), ptr(out), stream_handle())
    return out


def conv3d_bf16_bwd(x_bits, w, dout_bits, grid, want_din: bool = True, want_dw: bool = True):
    torch = _torch()
    n0, n1, n2 = (int(v) for v in grid)
    cout, _, cin = w.shape
    din = torch.empty(n0 * n1 * n2, cin, dtype=torch.float32, device="cuda") if want_din else None
    dw = torch.empty(cout, 27, cin, dtype=torch.float32, device="cuda") if want_dw else None
    call("stp_conv3d_bf16_bwd", C.c_int64(n0), C.c_int64(n1), C.c_int64(n2), int(cin), int(cout),
         ptr(x_bits.contiguous()), ptr(w.float().contiguous()), ptr(dout_bits.contiguous()), ptr(din), ptr(dw),
         stream_handle())
    return din, dw


def to_bf16_bits(x):
    """FP32 tensor -> int16 BF16 bit patterns (round to nearest even)."""
    torch = _torch()
    return x.to(torch.bfloat16).view(torch.int16)


# -------------------------------------------------------------------- training
@dataclass
class TrainLog:
    losses: list = field(default_factory=list)
    seconds: float = 0.0
    systems: int = 0
    steps: int = 0


def tomography_system(grid, nrays: int, seed: int, h: float = 1.0):
    """Straight-ray system of the tomography benchmark recipe: rays between
    seeded uniform points on opposite faces (face pair round-robin), exact
    Siddon lengths.  Returns the device matrix."""
    from . import synthetic
    from .sparse import DeviceCSR
    pts = synthetic.ray_endpoints(nrays, tuple(grid), h=h, seed=seed)
    return DeviceCSR.from_rays(pts, tuple(grid), h)


def train(net: LearnedPreconditioner, systems, n_systems: int, lam: float = 1e-2, steps_per_system: int = 2,
          nprobe: int = 4, lr: float = 2e-3, lr_final: float = 2e-4, seed: int = 0, log_every: int = 0) -> TrainLog:
    """Adam on log K over an iterable of n_systems (grid, matrix) systems
    (consumed lazily, so each matrix can be freed after use),
    `steps_per_system` steps each with fresh Rademacher probes; the learning
    rate decays geometrically from lr to lr_final."""
    torch = _torch()
    total = max(1, n_systems * steps_per_system)
    log = TrainLog()
    t0 = time.perf_counter()
    k = 0
    for si, (grid, A) in enumerate(systems):
        V = int(np.prod(grid))
        for _ in range(steps_per_system):
            rate = lr * (lr_final / lr) ** (k / max(1, total - 1))
            y = rademacher(nprobe, V, seed * 1_000_003 + k, device="cuda")
            loss = net.train_step(A, grid, lam, y, rate)
            log.losses.append(loss)
            k += 1
            if log_every and k % log_every == 0:
                print(f"[train] step {k}/{total} system {si} loss {loss:.4f}", flush=True)
        log.systems += 1
    torch.cuda.synchronize()
    log.seconds = time.perf_counter() - t0
    log.steps = k
    return log


# -------------------------------------------------------------------- default model
DEFAULT_WEIGHTS = __import__("pathlib").Path(__file__).resolve().parent / "data" / "lpc_tomography.npz"


def solve_iterations(net, grid, nrays: int, seed: int, lam: float = 1e-2, tol: float = 1e-6) -> tuple[int, int]:
    """(Jacobi, learned) CGLS iterations to tol on one system of the recipe."""
    torch = _torch()
    from . import synthetic
    A = tomography_system(grid, nrays, seed)
    n = grid[0]
    s = torch.as_tensor(1.0 / synthetic.smooth_random_velocity(*grid, seed=seed + 1, sigma=20.0 * n / 512 * 4),
                        dtype=torch.float64, device="cuda").reshape(-1)
    b = A.matvec(s)
    _, ij, _ = A.cgls(b, lam, tol, 20000, "jacobi")
    net.prepare(A, grid, lam)
    _, il, _ = A.cgls(b, lam, tol, 20000, "learned", network=net)
    return ij.iterations, il.iterations


def train_recipe(seed: int = 0, n_systems: int = 400, grid=(32, 32, 32), nrays: int = 50_000,
                 val_every: int = 50, val=((64, 64, 64), 200_000, 90_000), log_every: int = 0):
    """The training run behind the default weights: Adam on log K over
    `n_systems` seeded systems of the tomography recipe at `grid` (the
    200k-ray 64^3 density per cell), 2 steps per system with 4 Rademacher
    probes; every `val_every` systems the validation system `val` (grid,
    rays, seed; seeds disjoint from training and from the benchmark's) is
    solved and the weights with the fewest learned-CGLS iterations are kept."""
    net = LearnedPreconditioner(seed=seed)
    best = (solve_iterations(net, *val)[1], net.get_params(), 0)
    history = [(0, best[0])]
    done = 0
    while done < n_systems:
        k = min(val_every, n_systems - done)
        train(net, ((grid, tomography_system(grid, nrays, seed=s)) for s in range(done, done + k)), n_systems=k,
              steps_per_system=2, nprobe=4, lr=2e-3 * (0.3 ** (done / max(1, n_systems))),
              lr_final=2e-3 * (0.3 ** ((done + k) / max(1, n_systems))), seed=seed + done, log_every=log_every)
        done += k
        it = solve_iterations(net, *val)[1]
        history.append((done, it))
        if it < best[0]:
            best = (it, net.get_params(), done)
    net.set_params(best[1])
    return net, {"validation": history, "kept_after_systems": best[2], "validation_iterations": best[0]}


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(description="train the default learned preconditioner on the GPU")
    ap.add_argument("--out", default=str(DEFAULT_WEIGHTS))
    ap.add_argument("--systems", type=int, default=400)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    net, info = train_recipe(seed=args.seed, n_systems=args.systems, log_every=100)
    __import__("pathlib").Path(args.out).parent.mkdir(parents=True, exist_ok=True)
    net.save(args.out)
    print(json.dumps(info), flush=True)
