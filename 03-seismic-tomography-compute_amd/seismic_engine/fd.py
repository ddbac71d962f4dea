"""Device-resident acoustic FD propagator over the C ABI (stp_fd_*).

Wavefields, coefficients and traces are torch tensors on the GPU; every
compute call goes to libseismic_hip.so.  Layout and semantics follow the
reference's WaveEquationSolver (src/cpp/solver/wave_equation.h:50-144):
with-halo x-major arrays (z contiguous), additive source injection before
each step, sponge on all six faces, finite guards inside the step.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import FdConfig, call, ptr, stream_handle


def _torch():
    import torch
    return torch


SPONGE_MODES = {"reference": 0, "cerjan": 1}


class FDPropagator:
    """Leapfrog propagator for `nshots` wavefields sharing one velocity model.

    order 2 is the reference's 7-point stencil (wave_equation.cpp:370-373);
    order 8 is the radius-4 star stencil.  Both advance
        next = 2*cur - D*prev_raw + v^2 dt^2 * Lap(cur)
    where D is the reference sponge factor (wave_equation.cpp:262-297).
    sponge="cerjan" multiplies `next` by D as well (not in the reference): the
    reference form grows inside the sponge on long runs, this one decays.
    """

    def __init__(self, nx: int, ny: int, nz: int, dx: float = 10.0, dy: float = 10.0, dz: float = 10.0,
                 dt: float = 1e-3, order: int = 8, halo_width: int | None = None, sponge_width: int = 20,
                 sponge_damping: float = 0.015, amplitude_limit: float = 1e10, cfl_safety_factor: float = 0.49,
                 nshots: int = 1, device: str = "cuda", sponge: str = "reference"):
        torch = _torch()
        if halo_width is None:
            halo_width = order // 2
        self.n = (int(nx), int(ny), int(nz))
        self.hw = int(halo_width)
        self.order = int(order)
        self.nshots = int(nshots)
        self.dt = float(dt)
        self.spacing = (float(dx), float(dy), float(dz))
        self.nh = tuple(k + 2 * self.hw for k in self.n)
        self.ncell = int(np.prod(self.nh))
        self.device = torch.device(device)
        self.cfg = FdConfig(nx, ny, nz, self.hw, self.order, dx, dy, dz, dt, sponge_width, sponge_damping,
                            amplitude_limit, cfl_safety_factor, self.nshots)
        h = C.c_void_p()
        call("stp_fd_plan_create", C.byref(self.cfg), C.byref(h))
        self._plan = h
        if sponge not in SPONGE_MODES:
            raise ValueError(f"sponge must be one of {sorted(SPONGE_MODES)}")
        self.sponge = sponge
        call("stp_fd_set_sponge_mode", self._plan, SPONGE_MODES[sponge])
        shape = (self.nshots,) + self.nh
        self.cur = torch.zeros(shape, dtype=torch.float32, device=self.device)
        self.prev = torch.zeros(shape, dtype=torch.float32, device=self.device)
        self.coef = torch.zeros(self.nh, dtype=torch.float32, device=self.device)
        self.velocity = torch.zeros(self.nh, dtype=torch.float32, device=self.device)
        self.prev_damped = False
        self.vmax = 0.0
        self.nrcv = 0
        self.t = 0          # time index of the next step (wavelet sample to inject)
        self._have_src = False

    def __del__(self):
        try:
            if getattr(self, "_plan", None):
                _lib.lib().stp_fd_plan_destroy(self._plan)
                self._plan = None
        except Exception:
            pass

    # ------------------------------------------------------------------ setup
    def index(self, ix: int, iy: int, iz: int) -> int:
        """With-halo flat index of interior cell (ix, iy, iz)."""
        nyh, nzh = self.nh[1], self.nh[2]
        return ((ix + self.hw) * nyh + (iy + self.hw)) * nzh + (iz + self.hw)

    def set_velocity(self, vel) -> float:
        """Velocity as a with-halo array (nxh,nyh,nzh) or interior (nx,ny,nz)
        array (halo filled by edge replication).  Returns vmax."""
        torch = _torch()
        v = vel
        if isinstance(v, np.ndarray):
            v = torch.from_numpy(np.ascontiguousarray(v))
        v = v.to(device=self.device, dtype=torch.float32)
        if tuple(v.shape) == self.n:
            v = torch.nn.functional.pad(v[None, None], (self.hw,) * 6, mode="replicate")[0, 0]
        if tuple(v.shape) != self.nh and v.numel() != self.ncell:
            raise ValueError(f"velocity shape {tuple(v.shape)} is neither {self.n} nor {self.nh}")
        self.velocity.copy_(v.reshape(self.nh))
        vmax = C.c_double()
        call("stp_fd_coefficients", self._plan, ptr(self.velocity), ptr(self.coef), C.byref(vmax),
             stream_handle(self.device))
        self.vmax = vmax.value
        return self.vmax

    def verify_cfl(self) -> bool:
        ok = C.c_int32()
        call("stp_fd_verify_cfl", self._plan, C.c_double(self.vmax), C.byref(ok))
        return bool(ok.value)

    def set_sources(self, idx) -> None:
        """One with-halo flat index per shot."""
        a = np.ascontiguousarray(np.asarray(idx, dtype=np.int64).reshape(-1))
        call("stp_fd_set_sources", self._plan, a.ctypes.data, int(a.size))
        self._have_src = True

    def set_receivers(self, idx) -> None:
        """With-halo flat indices of receiver cells (shared by all shots)."""
        a = np.ascontiguousarray(np.asarray(idx, dtype=np.int64).reshape(-1))
        call("stp_fd_set_receivers", self._plan, a.ctypes.data if a.size else None, int(a.size))
        self.nrcv = int(a.size)

    def reset(self) -> None:
        """reset_field (wave_equation.cpp:501-506)."""
        self.cur.zero_()
        self.prev.zero_()
        self.prev_damped = False
        self.t = 0

    def sponge_factors(self):
        torch = _torch()
        d = torch.empty(self.nh, dtype=torch.float32, device=self.device)
        call("stp_fd_sponge_factors", self._plan, ptr(d), stream_handle(self.device))
        return d

    # ------------------------------------------------------------------ stepping
    def propagate(self, nsteps: int, wavelet=None, record: bool = True, status: bool = False):
        """Advance nsteps steps starting at time index self.t.

        wavelet: (nw,) or (nshots, nw) samples; sample t is injected at each
        shot's source before step t (t < nw).  Returns traces of shape
        (nshots, nrcv, nsteps) when record and receivers are set (else None),
        and the per-step status words if status=True.
        """
        torch = _torch()
        w = None
        nw = 0
        if wavelet is not None:
            if not self._have_src:
                raise RuntimeError("set_sources() before propagating a wavelet")
            w = torch.as_tensor(np.asarray(wavelet) if not torch.is_tensor(wavelet) else wavelet,
                                dtype=torch.float32).to(self.device)
            if w.dim() == 1:
                w = w.expand(self.nshots, -1)
            w = w.contiguous()
            nw = int(w.shape[1])
        traces = None
        if record and self.nrcv > 0:
            traces = torch.zeros((self.nshots, self.nrcv, nsteps), dtype=torch.float32, device=self.device)
        st = torch.zeros(max(nsteps, 1), dtype=torch.int32, device=self.device) if status else None
        call("stp_fd_propagate", self._plan, ptr(self.cur), ptr(self.prev), ptr(self.coef), int(nsteps),
             int(self.prev_damped), ptr(w), nw, int(self.t), ptr(traces), int(nsteps), ptr(st),
             stream_handle(self.device))
        if nsteps % 2 == 1:
            self.cur, self.prev = self.prev, self.cur
        if nsteps > 0:
            self.prev_damped = False
        self.t += nsteps
        if status:
            return traces, st[:nsteps]
        return traces

    def previous_damped(self):
        """The reference's `previous` buffer (sponge already applied)."""
        if self.prev_damped:
            return self.prev
        return self.prev * self.sponge_factors()

    # ------------------------------------------------------------------ reverse time
    def reverse_time(self, nt: int, wavelet, *, obs=None, inject=None, snapshot: str = "laplacian",
                     scale_by_c: bool = True, illumination: bool = False, checkpoint_interval: int = 0,
                     record: bool = True):
        """One forward pass from rest (nt steps, checkpointed) and the reverse
        pass of the adjoint recursion (stp_fd_adjoint).

        obs      : observed traces (nshots, nrcv, nt) -> residual injection and misfit
        inject   : traces to inject instead (e.g. recorded data for RTM)
        snapshot : "laplacian" (gradient accumulator) or "field" (RTM image)
        Returns dict(acc=(nshots,nx,ny,nz), ill=..., traces=(nshots,nrcv,nt), misfit=float|None).
        The field state (cur/prev) is not modified.
        """
        torch = _torch()
        if not self._have_src:
            raise RuntimeError("set_sources() before reverse_time()")
        w = torch.as_tensor(np.asarray(wavelet) if not torch.is_tensor(wavelet) else wavelet,
                            dtype=torch.float32).to(self.device)
        if w.dim() == 1:
            w = w.expand(self.nshots, -1)
        w = w.contiguous()
        shp = (self.nshots,) + self.n
        acc = torch.empty(shp, dtype=torch.float32, device=self.device)
        ill = torch.empty(shp, dtype=torch.float32, device=self.device) if illumination else None
        need_tr = record or obs is not None
        tr = torch.empty((self.nshots, self.nrcv, nt), dtype=torch.float32, device=self.device) \
            if (need_tr and self.nrcv) else None

        def dev3(a):
            if a is None:
                return None
            t = torch.as_tensor(np.asarray(a) if not torch.is_tensor(a) else a, dtype=torch.float32).to(self.device)
            return t.reshape(self.nshots, self.nrcv, nt).contiguous()

        o = dev3(obs)
        inj = dev3(inject)
        mis = C.c_double(float("nan"))
        call("stp_fd_adjoint", self._plan, ptr(self.coef), ptr(w), int(w.shape[1]), int(nt), ptr(o), ptr(inj),
             int(scale_by_c), int(snapshot == "laplacian"), int(checkpoint_interval), ptr(acc), ptr(ill), ptr(tr),
             C.byref(mis), stream_handle(self.device))
        return {"acc": acc, "ill": ill, "traces": tr, "misfit": mis.value if obs is not None else None}

    def misfit(self, nt: int, wavelet, obs):
        """J = 1/2 sum (d - obs)^2 and the modelled traces from one forward
        pass from rest (no checkpoints, no reverse pass); the field state is
        not modified."""
        torch = _torch()
        if not self._have_src:
            raise RuntimeError("set_sources() before misfit()")
        w = torch.as_tensor(np.asarray(wavelet) if not torch.is_tensor(wavelet) else wavelet,
                            dtype=torch.float32).to(self.device)
        if w.dim() == 1:
            w = w.expand(self.nshots, -1)
        w = w.contiguous()
        o = torch.as_tensor(np.asarray(obs) if not torch.is_tensor(obs) else obs, dtype=torch.float32)
        o = o.to(self.device).reshape(self.nshots, self.nrcv, nt).contiguous()
        tr = torch.empty((self.nshots, self.nrcv, nt), dtype=torch.float32, device=self.device)
        mis = C.c_double(float("nan"))
        call("stp_fd_adjoint", self._plan, ptr(self.coef), ptr(w), int(w.shape[1]), int(nt), ptr(o), None, 0, 0, 0,
             None, None, ptr(tr), C.byref(mis), stream_handle(self.device))
        return mis.value, tr

    def misfit_gradient(self, nt: int, wavelet, obs, checkpoint_interval: int = 0):
        """J = 1/2 sum (d - obs)^2 and dJ/dv (nshots summed), FP32 on the GPU.
        dJ/dv = 2 acc / v with acc = sum_k phi^k L u~^{k-1} (adjoint recursion)."""
        out = self.reverse_time(nt, wavelet, obs=obs, snapshot="laplacian", scale_by_c=True,
                                checkpoint_interval=checkpoint_interval, record=True)
        hw = self.hw
        v = self.velocity[hw:-hw, hw:-hw, hw:-hw]
        grad = 2.0 * out["acc"].sum(0) / v
        return out["misfit"], grad, out["traces"]
