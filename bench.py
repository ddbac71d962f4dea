"""Benchmark: 8th-order acoustic FD propagation on one MI355X.

Workload (one bench step): a full forward propagation on a 512^3 grid
(FP32, dx = 10 m, dt = 1 ms, 2000 time steps, 20-cell sponge), velocity =
seeded random field smoothed by a Gaussian of sigma = 20 cells and scaled to
1500-4500 m/s, Ricker 15 Hz source at the centre, 256 receivers on the
surface line iy = 256 (ix = 0, 2, ..., 510) sampled every step.  Inputs are
resident in HBM before timing starts.

Prints ONE JSON line: Gcell-updates/s, the FD step kernel against the HBM
roofline (16 algorithmic bytes per cell-update), and the FP64 CPU oracle
timed on this host as a reported baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "03-seismic-tomography-compute_amd"))

METRIC = "FD Gcell-updates/s + %HBM roofline; time-to-1e-6 CG: learned-PCG vs Jacobi"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = 16.0   # read u^t, u^{t-1}, c; write u^{t+1}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pmc_traffic(n: int) -> tuple[float | None, str | None]:
    """HBM bytes per FD launch from the committed rocprofv3 PMC summary."""
    p = ROOT / "profiles" / f"fd_step_kernel_{n}cube_pmc.md"
    if not p.exists():
        return None, None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", p.read_text())
    return (float(m.group(1)), str(p.relative_to(ROOT))) if m else (None, None)


def omp_threads(k: int) -> None:
    """Thread count of the oracle's OpenMP loops from here on (the runtime
    reads OMP_NUM_THREADS only once, when it starts)."""
    import ctypes
    from oracle import fd as ofd
    ofd.lib()                                        # loads libgomp with the oracle
    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(k))


def host_cpu() -> dict:
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "usable": len(os.sched_getaffinity(0))}


def stream_copy_rate(dev, nbytes: int = 1 << 30, reps: int = 10) -> float:
    """Device-to-device copy of nbytes (torch copy_ on the current stream),
    bytes read + written per second: the measured STREAM-copy rate."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    return 2.0 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9


def stream_3r1w_rate(dev, n: int, reps: int = 10) -> dict:
    """The FD step's access pattern without the stencil (stp_stream_3r1w_f32:
    out = a + b * c over n^3 FP32 cells, three reads and one write per cell),
    plain and with non-temporal loads and stores: algorithmic GB/s of each,
    timed with HIP events on the current stream.  The measured streaming
    ceiling of the step's 16 bytes per cell."""
    import ctypes as C
    import torch
    from seismic_engine._lib import call, ptr, stream_handle
    m = n ** 3
    a, b, c = (torch.full((m,), v, dtype=torch.float32, device=dev) for v in (1.0, 2.0, 0.5))
    o = torch.empty_like(a)
    out = {}
    for nt in (0, 1):
        def run():
            call("stp_stream_3r1w_f32", C.c_int64(m), ptr(a), ptr(b), ptr(c), ptr(o), nt, stream_handle())
        for _ in range(3):
            run()
        torch.cuda.synchronize()
        st = torch.cuda.current_stream(dev)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            run()
        e1.record(st)
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps
        out["nontemporal" if nt else "plain"] = {"us": round(us, 2), "gb_s": round(16.0 * m / (us * 1e-6) / 1e9, 1)}
    assert float(o[0]) == 2.0
    del a, b, c, o
    torch.cuda.empty_cache()
    return out


def cpu_baseline(threads: int, vel_h, n: int, steps: int, sponge: str, order: int = 8, rcv=None):
    """The FP64 CPU restatement of the reference's forward propagation (its
    4-pass step in C, OpenMP over x), built with the reference's release flags
    (oracle liboracle_timing.so), on an n^3 with-halo model: `steps` time
    steps from rest, centre Ricker source.  Returns (Gcell-updates/s, seconds)."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    omp_threads(threads)
    from oracle import fd as ofd
    from seismic_engine import synthetic
    ofd.select_build("timing")
    try:
        hw = 4 if order == 8 else 1
        c = ofd.cfg((n, n, n), hw=hw, order=order, sponge_width=20, sponge_damping=0.015, sponge=sponge)
        w = synthetic.ricker(15.0, 1e-3, steps)
        t0 = time.perf_counter()
        _, _, _, res = ofd.forward(c, vel_h, w, steps, rcv=rcv)
        dt = time.perf_counter() - t0
        assert res.completed_successfully == 1
    finally:
        ofd.select_build("check")
    return float(n) ** 3 * steps / dt / 1e9, dt


def cpu_fd_reference_order(threads: int, steps1: int = 50):
    """The reference's own stencil (2nd order, 7 points, halo 1) on its
    two-layer 128^3 case: 1500 / 2500 m/s split at iz = 64, sponge 20 / 0.015,
    Ricker 15 Hz at the centre, receivers at every surface cell, 500 steps in
    full on `threads` threads, and the first steps1 steps on one thread."""
    from seismic_engine import synthetic
    n = 128
    vel_h = synthetic.pad_halo(synthetic.two_layer_velocity(n, n, n, split_iz=64), 1)
    rcv = synthetic.surface_receivers(n, n, n, 1)
    g, sec = cpu_baseline(threads, vel_h, n, 500, "reference", order=2, rcv=rcv)
    g1, sec1 = cpu_baseline(1, vel_h, n, steps1, "reference", order=2, rcv=rcv)
    omp_threads(threads)
    return {"value": round(g, 5), "unit": "Gcell-updates/s", "cores": threads, "seconds": round(sec, 3),
            "single_core_value": round(g1, 5),
            "sample": f"two-layer 128^3, order 2 (the reference's 7-point stencil, halo 1), reference sponge, "
                      f"500 steps in full on {threads} threads ({sec:.1f} s); one thread: first {steps1} steps "
                      f"({sec1:.1f} s)"}


def cpu_cg_baseline(dev, threads: int, nrays: int = 200_000, n: int = 64, lam: float = 1e-2, tol: float = 1e-6):
    """The tomography solve on the CPU: the 200k-ray 64^3 system of the GPU
    benchmark (FP32 path lengths as FP64, int64 indices), the reference's
    2-accumulator SpMV (OpenMP over rows) and sequential scatter SpMV^T, and
    Jacobi CGLS to 1e-6 (FP64 throughout), timed with the reference's release
    flags.  SpMV bytes per call: 16 nnz + 8 (rows + 1) + 8 rows + 8 cols (x
    counted once)."""
    import numpy as np
    import torch
    from oracle import fd as ofd
    from oracle import sparse as osp
    from seismic_engine import synthetic
    from seismic_engine.sparse import DeviceCSR
    pts = synthetic.ray_endpoints(nrays, (n, n, n), h=1.0, seed=2026)
    A = DeviceCSR.from_rays(pts, (n, n, n), 1.0)
    s_true = torch.as_tensor(1.0 / synthetic.smooth_random_velocity(n, n, n, seed=2027, sigma=20.0 * n / 512 * 4),
                             dtype=torch.float64, device=dev).reshape(-1)
    b = A.matvec(s_true).cpu().numpy()
    rp, ci, v = (t.cpu().numpy() for t in A.export())
    rp, ci, v = rp.astype(np.int64), ci.astype(np.int64), v.astype(np.float64)
    rows, cols, nnz = A.rows, A.cols, A.nnz
    del A
    x = np.random.default_rng(0).standard_normal(cols)
    nbytes = 16 * nnz + 8 * (rows + 1) + 8 * rows + 8 * cols
    out = {"rows": rows, "cols": cols, "nnz": nnz, "lambda": lam, "tol": tol}
    ofd.select_build("timing")
    try:
        def timed(fn, reps):
            fn()
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            return (time.perf_counter() - t0) / reps
        for k in (threads, 1):
            omp_threads(k)
            sec = timed(lambda: osp.spmv(rp, ci, v, x), 10 if k > 1 else 3)
            out[f"spmv_{k}_threads"] = {"ms": round(sec * 1e3, 3), "gb_s": round(nbytes / sec / 1e9, 2)}
        omp_threads(threads)
        r = np.random.default_rng(1).standard_normal(rows)
        sec = timed(lambda: osp.spmv_t(rp, ci, v, r, cols), 3)
        out["spmv_t_sequential"] = {"ms": round(sec * 1e3, 3), "gb_s": round(nbytes / sec / 1e9, 2)}
        t0 = time.perf_counter()
        _, it, rel, _ = osp.cgls(rp, ci, v, b, cols, lam, tol, 20000, precond=1)
        sec = time.perf_counter() - t0
        out["jacobi_cgls"] = {"seconds": round(sec, 3), "iterations": int(it), "rel_residual": rel,
                              "threads": threads}
    finally:
        ofd.select_build("check")
    return out


def accuracy_check(dev, threads: int, vel_int, sponge: str, case: str, steps: int = 500):
    """FP32 GPU propagation against the FP64 oracle on a 128^3 model (full
    `steps` steps, centre Ricker source, receivers at every surface cell):
    relative L2 error of the final field and of the traces, and the largest
    trace amplitude (a growing sponge would show there)."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    import numpy as np
    import torch
    from oracle import fd as ofd
    from seismic_engine import FDPropagator, synthetic
    n = (128, 128, 128)
    c = ofd.cfg(n, hw=4, order=8, sponge_width=20, sponge_damping=0.015, sponge=sponge)
    v = synthetic.pad_halo(vel_int, 4)
    w = synthetic.ricker(15.0, 1e-3, steps)
    rcv = synthetic.surface_receivers(*n, 4)
    cur_o, _, tr_o, res = ofd.forward(c, v, w, steps, rcv=rcv)
    g = FDPropagator(*n, order=8, sponge_width=20, sponge_damping=0.015, device=dev, sponge=sponge)
    g.set_velocity(v)
    g.set_sources([g.index(64, 64, 64)])
    g.set_receivers(rcv)
    trg = g.propagate(steps, w)
    torch.cuda.synchronize()
    e = float(np.linalg.norm(g.cur[0].cpu().numpy() - cur_o) / np.linalg.norm(cur_o))
    et = float(np.linalg.norm(trg[0].cpu().numpy() - tr_o) / np.linalg.norm(tr_o))
    return {"case": f"{case}, 128^3, {steps} steps, {sponge} sponge", "field_rel_l2_vs_fp64": e,
            "traces_rel_l2_vs_fp64": et, "trace_max_abs_fp64": float(np.abs(tr_o).max())}


def adjoint_accuracy(dev, threads: int, n: int = 64, nt: int = 300, sponge: str = "cerjan"):
    """RTM image and adjoint-state gradient of the GPU (FP32) against the FP64
    oracle at n^3: smooth random true and starting models (1500-4500 m/s),
    Ricker 15 Hz source at the centre of the surface (iz = 2), receivers at
    every surface cell, nt steps."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    import numpy as np
    import torch
    from oracle import adjoint as oad
    from oracle import fd as ofd
    from seismic_engine import FDPropagator, synthetic
    hw = 4
    v_true = synthetic.smooth_random_velocity(n, n, n, seed=11, sigma=n / 16)
    v0 = synthetic.smooth_random_velocity(n, n, n, seed=12, sigma=n / 12)
    w = synthetic.ricker(15.0, 1e-3, nt)
    rcv = synthetic.surface_receivers(n, n, n, hw)
    fd = FDPropagator(n, n, n, order=8, sponge_width=20, sponge_damping=0.015, device=dev, sponge=sponge)
    src = fd.index(n // 2, n // 2, 2)
    fd.set_sources([src])
    fd.set_receivers(rcv)
    fd.set_velocity(v_true)
    obs = fd.propagate(nt, w).clone()
    fd.reset()
    fd.set_velocity(v0)
    img = fd.reverse_time(nt, w, inject=obs, snapshot="field", scale_by_c=False, illumination=False)["acc"][0]
    _, g, tr = fd.misfit_gradient(nt, w, obs)
    torch.cuda.synchronize()
    c = ofd.cfg((n, n, n), hw=hw, order=8, sponge_width=20, sponge_damping=0.015, sponge=sponge)
    v0h = synthetic.pad_halo(v0, hw)
    obs_np = obs[0].cpu().numpy().astype(np.float64)
    img_o, _, tro = oad.adjoint(c, v0h, w, nt, src, rcv, inj=obs_np, scale_c=False, snap_lap=False)
    acc_o, _, _ = oad.adjoint(c, v0h, w, nt, src, rcv, inj=tro - obs_np, scale_c=True, snap_lap=True)
    g_o = 2 * acc_o / v0

    def rel(a, b):
        return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))
    return {"case": f"{n}^3, {nt} steps, {sponge} sponge, smooth random models, surface source and receivers",
            "rtm_image_rel_l2_vs_fp64": rel(img.cpu().numpy(), img_o),
            "gradient_rel_l2_vs_fp64": rel(g.cpu().numpy(), g_o)}


def cg_workload(dev, nrays: int = 200_000, n: int = 64, lam: float = 1e-2, tol: float = 1e-6, reps: int = 3,
                kinds=("jacobi", "none", "learned"), given_order: bool = True):
    """Tomography solve: exact straight-ray CSR (seeded source/receiver pairs
    on opposite faces of a 64^3 model), b = A s_true, CGLS on
    (A^T A + lam I) x = A^T b to a 1e-6 relative normal-equation residual."""
    import torch
    from seismic_engine import synthetic
    from seismic_engine.sparse import DeviceCSR
    pts = synthetic.ray_endpoints(nrays, (n, n, n), h=1.0, seed=2026)
    A = DeviceCSR.from_rays(pts, (n, n, n), 1.0)
    s_true = torch.as_tensor(1.0 / synthetic.smooth_random_velocity(n, n, n, seed=2027, sigma=20.0 * n / 512 * 4),
                             dtype=torch.float64, device=dev).reshape(-1)
    b = A.matvec(s_true)
    out = {"rows": A.rows, "cols": A.cols, "nnz": A.nnz, "lambda": lam, "tol": tol}
    # Morton solver layout of the matrix: built once per matrix, kept with it
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    A.prepare_layout()
    torch.cuda.synchronize()
    out["layout_seconds"] = round(time.perf_counter() - t0, 5)
    # SpMV alone (through the solver layout built above, the two vector
    # permutations included): 8 B/nnz + row offsets + y + gathered x (counted once)
    x32 = s_true.float()
    y32 = torch.empty(A.rows, dtype=torch.float32, device=dev)
    for _ in range(3):
        y32 = A.matvec(x32)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        y32 = A.matvec(x32)
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / 20
    spmv_bytes = 8 * A.nnz + 4 * (A.rows + 1) + 4 * A.rows + 4 * A.cols
    out["spmv"] = {"us": round(us, 2), "gb_s": round(spmv_bytes / (us * 1e-6) / 1e9, 1),
                   "frac_hbm": round(spmv_bytes / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)}
    # gather cost of the two CGLS products in the solver layout: the same
    # kernels without the value stream (index stream + gathers), i.e. the
    # ceiling of each product's gather pattern
    from seismic_engine._lib import call, ptr, stream_handle
    gp = {}
    for name, transpose, fp64 in (("A_p_fp32", 0, 0), ("At_r_fp64", 1, 1)):
        xin = torch.rand(A.rows if transpose else A.cols, dtype=torch.float64 if fp64 else torch.float32, device=dev)
        yo = torch.empty(A.cols if transpose else A.rows, dtype=xin.dtype, device=dev)

        def run():
            call("stp_csr_gather_probe", A._h, transpose, fp64, ptr(xin), ptr(yo), stream_handle())
        for _ in range(3):
            run()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(20):
            run()
        e1.record()
        torch.cuda.synchronize()
        gus = e0.elapsed_time(e1) * 1e3 / 20
        idx_bytes = 4 * A.nnz + 4 * ((A.cols if transpose else A.rows) + 1)
        gp[name] = {"us": round(gus, 2), "gathers_per_s": round(A.nnz / (gus * 1e-6) / 1e9, 2),
                    "unit": "G gathers/s", "index_stream_gb_s": round(idx_bytes / (gus * 1e-6) / 1e9, 1)}
    out["gather_probe"] = gp
    from seismic_engine.learned import DEFAULT_WEIGHTS, LearnedPreconditioner, train_recipe
    if "learned" not in kinds:
        net = None
    elif DEFAULT_WEIGHTS.exists():
        net = LearnedPreconditioner.load(DEFAULT_WEIGHTS)
        out["learned_weights"] = str(DEFAULT_WEIGHTS.relative_to(ROOT))
    else:
        log("no trained weights found: training the learned preconditioner now")
        t0 = time.perf_counter()
        net, tinfo = train_recipe()
        out["learned_weights"] = f"trained in this run ({time.perf_counter() - t0:.1f} s)"
    grid = (n, n, n)
    for kind in kinds:
        best = None
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if kind == "learned":
                net.prepare(A, grid, lam)     # coupling features + network: part of the solve time
                torch.cuda.synchronize()
                tprep = time.perf_counter() - t0
            x, info, _ = A.cgls(b, lam, tol, 20000, kind, network=net if kind == "learned" else None)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            if best is None or dt < best[0]:
                best = (dt, info, tprep if kind == "learned" else 0.0)
        dt, info, tp = best
        out[kind] = {"seconds": round(dt, 5), "iterations": info.iterations, "converged": bool(info.converged),
                     "true_rel_residual": info.true_rel_residual, "replacements": info.replacements}
        if kind == "learned":
            out[kind]["prepare_seconds"] = round(tp, 5)
    if "learned" in out and "jacobi" in out:
        out["learned_over_jacobi_time"] = round(out["learned"]["seconds"] / out["jacobi"]["seconds"], 4)
    if "learned" in out and "none" in out:
        out["learned_over_none_time"] = round(out["learned"]["seconds"] / out["none"]["seconds"], 4)
    if net is not None:
        # the untrained network (zero output layer): M = D^-1/2 (I + L) D^-1/2
        raw = LearnedPreconditioner(seed=0)
        best = None
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            raw.prepare(A, grid, lam)
            _, info, _ = A.cgls(b, lam, tol, 20000, "learned", network=raw)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = (dt, info) if best is None or dt < best[0] else best
        out["untrained_network"] = {"seconds": round(best[0], 5), "iterations": best[1].iterations,
                                    "converged": bool(best[1].converged)}
        del raw
    if "jacobi" in kinds and given_order:
        # the same Jacobi solve on the matrix in its given (x-major, ray) order
        A.set_grid(None)
        best = None
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, info, _ = A.cgls(b, lam, tol, 20000, "jacobi")
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        out["jacobi_given_order"] = {"seconds": round(best, 5), "iterations": info.iterations,
                                     "converged": bool(info.converged)}
        A.set_grid((n, n, n))
    if net is not None:
        log("learned preconditioner on the held-out 96^3 system")
        out["heldout_96"] = learned_heldout(dev, net=net)
    return out


def rtm_workload(fd, nt: int, w, traces):
    """RTM imaging on the benchmark model: the source wavefield is propagated
    forward with sqrt(nt) checkpoints, the recorded traces are injected in
    reverse time and the two wavefields are cross-correlated at zero lag
    (image and illumination) inside the reverse step kernel."""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fd.reverse_time(nt, w, inject=traces, snapshot="field", scale_by_c=False, illumination=True, record=False)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    img, ill = out["acc"], out["ill"]
    ok = bool(torch.isfinite(img).all()) and bool((ill >= 0).all()) and float(ill.max()) > 0
    n3 = img[0].numel()
    del out, img, ill
    torch.cuda.empty_cache()
    return {"seconds": round(sec, 4), "image_gcell_steps_per_s": round(n3 * nt / sec / 1e9, 2), "finite": ok,
            "passes": "forward with checkpoints, segment recompute with snapshots, reverse step with fused "
                      "image/illumination accumulation"}


def gradient_workload(dev, n: int = 256, nshots: int = 16, nt: int = 1500, sponge: str = "cerjan"):
    """Adjoint-state gradient: n^3 grid, 16 shots on a 4 x 4 surface grid
    (iz = 2), receivers every 4th surface cell, smooth random true and
    starting velocities (seeds 1 and 2, sigma 20 cells, 1500-4500 m/s), L2
    waveform misfit; observed data from the true model, then one misfit +
    gradient evaluation (forward, checkpointed, and adjoint) is timed."""
    import numpy as np
    import torch
    from seismic_engine import FDPropagator, synthetic
    hw = 4
    w = torch.as_tensor(synthetic.ricker(15.0, 1e-3, nt), dtype=torch.float32, device=dev)
    pos = [n // 8 * (1 + 2 * k) for k in range(4)]
    rcv = synthetic.surface_receivers(n, n, n, hw)[::1].reshape(n, n)[::4, ::4].reshape(-1)
    fd = FDPropagator(n, n, n, order=8, sponge_width=20, sponge_damping=0.015, nshots=nshots, device=dev,
                      sponge=sponge)
    fd.set_sources([fd.index(ix, iy, 2) for ix in pos for iy in pos])
    fd.set_receivers(rcv)
    fd.set_velocity(synthetic.smooth_random_velocity(n, n, n, seed=1, sigma=20.0, device=dev, as_tensor=True))
    obs = fd.propagate(nt, w).clone()
    obs_max = float(obs.abs().max())
    fd.reset()
    fd.set_velocity(synthetic.smooth_random_velocity(n, n, n, seed=2, sigma=20.0, device=dev, as_tensor=True))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J, g, _ = fd.misfit_gradient(nt, w, obs)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    ok = bool(torch.isfinite(g).all()) and np.isfinite(J) and float(g.abs().max()) > 0
    cells = float(n) ** 3 * nshots * nt
    del fd, obs, g
    torch.cuda.empty_cache()
    return {"grid": [n, n, n], "shots": nshots, "time_steps": nt, "receivers": int(rcv.size), "sponge": sponge,
            "seconds": round(sec, 3), "shot_gcell_steps_per_s": round(cells / sec / 1e9, 2), "misfit": J,
            "obs_max_abs": obs_max, "finite": ok}


def learned_heldout(dev, n: int = 96, lam: float = 1e-2, tol: float = 1e-6, net=None):
    """Held-out larger system of the learned-preconditioner recipe: n^3 cells,
    rays at the 200k-ray 64^3 density per face cell (seed 10_096), Jacobi
    against learned CGLS to 1e-6 (learned time includes its preparation)."""
    import torch
    from seismic_engine.learned import make_system
    nrays = int(200_000 * (n / 64) ** 2)
    sysm = make_system((n, n, n), nrays, seed=10_096, lam=lam, tol=1e-8)
    out = {"grid": [n, n, n], "rays": nrays, "nnz": sysm.A.nnz}
    sysm.A.prepare_layout()
    for kind in ("jacobi", "learned"):
        best = None
        for _ in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if kind == "learned":
                net.prepare(sysm.A, sysm.grid, lam)
            _, info, _ = sysm.A.cgls(sysm.b, lam, tol, 20000, kind, network=net if kind == "learned" else None)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            if best is None or dt < best[0]:
                best = (dt, info)
        out[kind] = {"seconds": round(best[0], 5), "iterations": best[1].iterations,
                     "converged": bool(best[1].converged)}
    out["learned_over_jacobi_time"] = round(out["learned"]["seconds"] / out["jacobi"]["seconds"], 4)
    return out


def network_mfma(dev, n: int = 64, reps: int = 20):
    """MFMA utilisation of the learned preconditioner's network at an n^3
    system: each layer's 3x3x3 BF16 convolution (FP32 accumulation) timed
    alone with HIP events on the launching stream; 2 cin cout 27 FLOP per cell
    of the layer's level (full, half, quarter, half, full, full grid) against
    the 2.5 PFLOP/s dense BF16 peak."""
    import torch
    from seismic_engine.learned import LAYERS, conv3d_bf16
    levels = (0, 1, 2, 1, 0, 0)
    gen = torch.Generator(device=dev).manual_seed(11)
    layers, tot_f, tot_t = [], 0.0, 0.0
    for (cin, cout), lv in zip(LAYERS, levels):
        m = n >> lv
        x = torch.randn(m ** 3, cin, device=dev, generator=gen).to(torch.bfloat16).view(torch.int16)
        w = 0.1 * torch.randn(cout, 27, cin, device=dev, generator=gen)
        conv3d_bf16(x, w, (m, m, m))
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            conv3d_bf16(x, w, (m, m, m))
        e1.record()
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) * 1e-3 / reps
        f = 2.0 * cin * cout * 27 * m ** 3
        tot_f += f
        tot_t += t
        layers.append({"cin": cin, "cout": cout, "grid": m, "us": round(t * 1e6, 1),
                       "tflop_s": round(f / t / 1e12, 2)})
    return {"grid": n, "dtype": "bf16 in, fp32 accumulate", "layers": layers,
            "forward_us": round(tot_t * 1e6, 1), "tflop_s": round(tot_f / tot_t / 1e12, 2),
            "frac_bf16_dense_peak": round(tot_f / tot_t / 1e12 / 2500.0, 5)}


def aux_workload(dev, n: int = 512, cpu_n: int = 128, hw: int = 4, reps: int = 10):
    """Per-call rate of the inversion loop's gradient post-processing, the
    terrain boundary and RTM slice extraction at n^3, FP64: HIP events on the
    launching stream around `reps` calls; algorithmic bytes per call (each
    array read or written once) over that time, against 8 TB/s.  CPU: the
    numpy restatements of the post-processing (oracle/gradient_post.py) on a
    cpu_n^3 sample, scaled per cell."""
    import ctypes as C
    from types import SimpleNamespace

    import numpy as np
    import torch
    from oracle import gradient_post as ogp
    from seismic_engine._lib import call, ptr, stream_handle
    from seismic_engine.sparse import SparseOps
    from seismic_engine.terrain import TerrainBoundary
    V = n ** 3
    n3 = (C.c_int64(n), C.c_int64(n), C.c_int64(n))
    gen = torch.Generator(device=dev).manual_seed(7)
    g = torch.randn(V, dtype=torch.float64, device=dev, generator=gen)
    v = 1500.0 + 3000.0 * torch.rand(V, dtype=torch.float64, device=dev, generator=gen)
    w = torch.empty_like(g)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3

    out = {"grid": [n, n, n], "dtype": "f64"}

    def rec(name, sec, nbytes, what):
        out[name] = {"us": round(sec * 1e6, 1), "bytes_per_call": int(nbytes), "gb_s": round(nbytes / sec / 1e9, 1),
                     "frac_hbm": round(nbytes / sec / 1e9 / HBM_PEAK_GBS, 4), "bytes": what}

    st = stream_handle
    rec("tikhonov", timed(lambda: call("stp_gradient_tikhonov_f64", *n3, C.c_double(1e-4), ptr(v), ptr(g), st())),
        24 * V, "v, g read, g written")
    rec("normalize", timed(lambda: call("stp_gradient_normalize_f64", C.c_int64(V), ptr(g), None, st())),
        24 * V, "max pass reads g; scale pass reads and writes g")
    rec("depth_preconditioner", timed(lambda: call("stp_depth_preconditioner_f64", *n3, C.c_int64(16),
                                                   C.c_double(200.0), ptr(w), st())), 8 * V, "weights written")
    rec("precondition", timed(lambda: call("stp_gradient_precondition_f64", C.c_int64(V), ptr(w), ptr(g), st())),
        24 * V, "w, g read, g written")
    rec("smoothing_sigma2", timed(lambda: SparseOps.smooth_gradient(g, n, n, n, 2.0)), 48 * V,
        "three 1-D passes, each reads and writes the volume")
    grid = SimpleNamespace(nx_global=n, ny_global=n, nz_global=n, dx=10.0, dy=10.0, dz=10.0, origin_x=0.0,
                           origin_y=0.0, origin_z=0.0, halo_width=hw)
    tb = TerrainBoundary(grid, 1500.0)
    tb.generate_synthetic_bathymetry(0.05, 20.0)
    del v, w
    torch.cuda.empty_cache()
    nh = n + 2 * hw
    vh = torch.zeros(nh ** 3, dtype=torch.float64, device=dev)
    rec("terrain_velocity_mask", timed(lambda: tb.update_velocity_mask(vh, 1500.0, 1800.0)), 8 * V,
        "interior velocity written")
    rec("terrain_boundary_mask", timed(lambda: tb.build_boundary_mask()), nh ** 3, "one byte per cell written")
    f = torch.randn(nh ** 3, dtype=torch.float64, device=dev, generator=gen)
    ic = g
    o = torch.empty(3, n * n, dtype=torch.float64, device=dev)
    for axis, name in ((0, "slice_x"), (2, "slice_z")):
        rec(name, timed(lambda: call("stp_rtm_extract_slice_f64", axis, C.c_int64(n // 2), *n3, hw, C.c_int64(nh),
                                     C.c_int64(nh), ptr(f), ptr(vh), ptr(ic), ptr(o[0]), ptr(o[1]), ptr(o[2]), st())),
            48 * n * n, "three planes read and written")
    del f, vh, g, o, tb
    torch.cuda.empty_cache()
    rng = np.random.default_rng(7)
    gc = rng.standard_normal(cpu_n ** 3)
    vc = 1500.0 + 3000.0 * rng.random(cpu_n ** 3)
    scale = V / cpu_n ** 3
    cpu = {}
    for name, fn in (("tikhonov", lambda: ogp.tikhonov(vc, gc, cpu_n, cpu_n, cpu_n, 1e-4)),
                     ("normalize", lambda: ogp.normalize(gc)),
                     ("depth_preconditioner", lambda: ogp.depth_weights(cpu_n, cpu_n, cpu_n, 16))):
        t0 = time.perf_counter()
        fn()
        cpu[name + "_ms_scaled"] = round((time.perf_counter() - t0) * scale * 1e3, 1)
    out["cpu_baseline"] = dict(cpu, sample=f"numpy restatements (oracle/gradient_post.py) at {cpu_n}^3, 1 process, "
                                           f"scaled by {scale:.0f} to {n}^3")
    return out


def traveltime_workload(dev, n: int = 64, npicks: int = 200_000, h: float = 10.0, cpu_sample: int = 2000):
    """Travel-time tomography (the reference's straight-ray sampled operator):
    npicks source/station pairs on opposite faces of an n^3 model of h-metre
    cells (the 200k-ray tomography geometry), smooth random velocity; one
    forward model, one gradient, and one iteration of the inversion loop
    (forward, gradient, smoothing, Tikhonov term, normalisation, line search).
    The reference's own pure-Python loops (restated in oracle/traveltime.py)
    are timed on a sample of the picks as the CPU baseline."""
    import numpy as np
    import torch
    from oracle import traveltime as ott
    from seismic_engine import synthetic
    from seismic_engine.inversion import TravelTimeInversion
    from seismic_engine.traveltime import Geometry, TravelTimeOperator
    pts = synthetic.ray_endpoints(npicks, (n, n, n), h=h, seed=2026)
    src, sta = pts[:, 0].copy(), pts[:, 1].copy()
    geo = Geometry(n, n, n, dx=h, dy=h, dz=h)
    vp = synthetic.smooth_random_velocity(n, n, n, seed=2027, sigma=5.0).reshape(-1)
    true = synthetic.smooth_random_velocity(n, n, n, seed=2028, sigma=5.0).reshape(-1)
    op = TravelTimeOperator(geo, src, sta, device=dev)
    obs = op.forward(true)
    t = op.forward(vp)
    r = op.weighted_residuals(t, obs)
    g = op.gradient(vp, r)
    torch.cuda.synchronize()
    reps = 20

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / reps
    t_fwd = timed(lambda: op.forward(vp))
    t_grad = timed(lambda: op.gradient(vp, r))
    def inversion(iters):
        return TravelTimeInversion(op, obs, vp, max_iterations=iters, step_length=50.0, smoothing_sigma=2.0,
                                   misfit_type="l2", descent=True)
    inversion(1).run()                                   # warm-up: first calls of every kernel
    torch.cuda.synchronize()
    calls = {"n": 0}
    forward = op.forward

    def counted(v):
        calls["n"] += 1
        return forward(v)
    op.forward = counted
    t0 = time.perf_counter()
    info = inversion(5).run()
    torch.cuda.synchronize()
    t_it = (time.perf_counter() - t0) / info["iterations"]
    op.forward = forward
    fwd_per_it = calls["n"] / info["iterations"]
    samples = int(sum(max(int(np.linalg.norm(a - b) / (h * 5)), 2) for a, b in zip(src[:2000], sta[:2000])))
    t0 = time.perf_counter()
    ott.forward(src[:cpu_sample], sta[:cpu_sample], (n, n, n), (0.0, 0.0, 0.0), (h, h, h), vp)
    t_cpu = time.perf_counter() - t0
    return {"picks": npicks, "grid": [n, n, n], "samples_per_pick": round(samples / 2000, 2),
            "forward_ms": round(t_fwd * 1e3, 3), "gradient_ms": round(t_grad * 1e3, 3), "timing": "HIP events, 20 calls",
            "inversion_iteration_ms": round(t_it * 1e3, 2),
            "inversion_timing": f"wall time per iteration over {info['iterations']} warm iterations (forward, "
                                f"misfit, residuals, gradient, smoothing, Tikhonov, normalisation, line search; "
                                f"{fwd_per_it:.1f} forward models per iteration)",
            "cpu_reference_loop_forward_ms_per_200k_picks": round(t_cpu * npicks / cpu_sample * 1e3, 1),
            "cpu_sample": f"pure-Python restatement of _forward_model on {cpu_sample} picks, 1 core, scaled"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--nt", type=int, default=2000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=64, help="time steps of the CPU baseline sample")
    ap.add_argument("--no-cg", action="store_true", help="skip the tomography-solve timing")
    ap.add_argument("--no-rtm", action="store_true", help="skip the RTM imaging timing")
    ap.add_argument("--sponge", default="cerjan", choices=("cerjan", "reference"),
                    help="sponge form of the FD, RTM and gradient workloads (the reference form grows inside the "
                         "sponge on these 1500-4500 m/s models; the kernel's memory traffic is the same)")
    ap.add_argument("--no-grad", action="store_true", help="skip the 256^3 x 16-shot adjoint-gradient timing")
    ap.add_argument("--no-tt", action="store_true", help="skip the travel-time tomography timing")
    ap.add_argument("--no-aux", action="store_true",
                    help="skip the gradient post-processing, terrain and slice-extraction timing")
    ap.add_argument("--aux-only", action="store_true", help="only the gradient post-processing, terrain and "
                                                            "slice-extraction timing (one JSON object)")
    ap.add_argument("--cg-only", action="store_true",
                    help="only the tomography-solve timing (one JSON object, for profiling the CGLS kernels)")
    ap.add_argument("--cg-kinds", default="jacobi,none,learned",
                    help="comma-separated CGLS preconditioners to time (jacobi, none, learned)")
    ap.add_argument("--no-cg-given-order", action="store_true",
                    help="skip the Jacobi solve on the matrix in its given order (no solver layout)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (reference for PMC byte counters)")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
                    or min(16, os.cpu_count() or 1))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from seismic_engine import FDPropagator, synthetic

    dev = torch.device("cuda:0")
    if args.copy_calibration:
        a = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 4 << 28, "bytes_written": 4 << 28, "copies": 3}))
        return 0
    if args.aux_only:
        print(json.dumps({"aux": aux_workload(dev)}), flush=True)
        return 0
    if args.cg_only:
        cg = cg_workload(dev, kinds=tuple(args.cg_kinds.split(",")), given_order=not args.no_cg_given_order)
        if "learned" in args.cg_kinds:
            cg["network_mfma"] = network_mfma(dev)
        print(json.dumps({"cg": cg}), flush=True)
        return 0
    n, nt, hw = args.n, args.nt, 4
    log(f"building {n}^3 velocity (seed 2026, sigma 20) on the GPU")
    vel = synthetic.smooth_random_velocity(n, n, n, seed=2026, sigma=20.0, device=dev, as_tensor=True)
    fd = FDPropagator(n, n, n, order=8, sponge_width=20, sponge_damping=0.015, device=dev, sponge=args.sponge)
    fd.set_velocity(vel)
    del vel
    if not fd.verify_cfl():
        raise SystemExit("CFL check failed")
    fd.set_sources([fd.index(n // 2, n // 2, n // 2)])
    rcv = synthetic.line_receivers(n, n, n, hw, iy=n // 2, step=max(1, n // 256))
    fd.set_receivers(rcv)
    w = torch.as_tensor(synthetic.ricker(15.0, 1e-3, nt), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    def one():
        fd.reset()
        return fd.propagate(nt, w)

    for i in range(args.warmup):
        one()
        torch.cuda.synchronize()
        log(f"warmup {i + 1}/{args.warmup} done")

    stream = torch.cuda.current_stream(dev)
    ev_prop = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        fd.reset()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        tr = fd.propagate(nt, w)
        e1.record(stream)
        ev_prop.append((e0, e1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prop_ms = [a.elapsed_time(b) for a, b in ev_prop]
    log(f"timed {args.steps} propagations: wall {wall:.3f} s, propagate {prop_ms} ms")
    assert torch.isfinite(tr).all()
    tr_max = float(tr.abs().max())
    field_max = float(fd.cur.abs().max())
    cells = float(n) ** 3
    value = cells * nt * args.steps / wall / 1e9
    launch_s = sum(prop_ms) / 1e3 / (args.steps * nt)
    achieved = BYTES_PER_CELL * cells / launch_s / 1e9
    traffic, traffic_src = pmc_traffic(n)

    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"fd8_forward_{n}cube_{nt}steps_{len(rcv)}rcv", "grid": [n, n, n],
                   "order": 8, "time_steps": nt, "receivers": int(len(rcv)), "sponge": [20, 0.015],
                   "velocity": "smooth random seed 2026 sigma 20, 1500-4500 m/s", "source": "Ricker 15 Hz centre",
                   "sponge_form": args.sponge, "step": "one full forward propagation"},
        "wavefield": {"trace_max_abs": tr_max, "final_field_max_abs": field_max,
                      "amplitude_limit_reached": bool(field_max >= 0.99e10)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "fd_step_kernel<4,4,128,16,1,256,4,2,6>", "bytes_per_cell_update": BYTES_PER_CELL,
                     "avg_launch_us": round(launch_s * 1e6, 2), "traffic_source": traffic_src},
    }
    copy_gbs = stream_copy_rate(dev)
    result["roofline"]["stream_copy_gbs"] = round(copy_gbs, 1)
    result["roofline"]["frac_of_stream_copy"] = round(achieved / copy_gbs, 4)
    s3 = stream_3r1w_rate(dev, n)
    best3 = max(v["gb_s"] for v in s3.values())
    result["roofline"]["stream_3r1w"] = s3
    result["roofline"]["frac_of_stream_3r1w"] = round(achieved / best3, 4)

    if not args.no_rtm:
        log(f"RTM imaging: {n}^3, {nt} steps, recorded traces back-propagated")
        result["rtm"] = rtm_workload(fd, nt, w, tr)
        result["rtm"]["sponge"] = args.sponge
    del tr
    if not args.no_grad:
        vel_keep = fd.velocity.double().cpu().numpy() if not args.no_cpu_baseline else None
        del fd
        torch.cuda.empty_cache()
        log("adjoint gradient: 256^3, 16 shots, 1500 steps")
        result["gradient"] = gradient_workload(dev, sponge=args.sponge)
    if not args.no_tt:
        log("travel-time tomography: 200k picks, 64^3")
        result["traveltime"] = traveltime_workload(dev)
    if not args.no_aux:
        log("gradient post-processing, terrain boundary and slice extraction at 512^3 (FP64)")
        result["aux"] = aux_workload(dev)
    if not args.no_cg:
        log("tomography solve: 200k rays, 64^3, CGLS to 1e-6")
        result["cg"] = cg_workload(dev)
        log("learned preconditioner network: per-layer BF16 MFMA convolutions at 64^3")
        result["cg"]["network_mfma"] = network_mfma(dev)

    if not args.no_cpu_baseline:
        if args.no_grad:
            vel_h = fd.velocity.double().cpu().numpy()
            del fd
            torch.cuda.empty_cache()
        else:
            vel_h = vel_keep
        steps = args.cpu_steps
        log(f"cpu baseline: FP64 restatement on the {n}^3 workload, first {steps} steps, {args.cpu_threads} threads")
        gcs, sec = cpu_baseline(args.cpu_threads, vel_h, n, steps, args.sponge, rcv=rcv)
        steps1 = 4
        log(f"cpu baseline: the same on 1 core, first {steps1} steps")
        gcs1, sec1 = cpu_baseline(1, vel_h, n, steps1, args.sponge, rcv=rcv)
        omp_threads(args.cpu_threads)
        del vel_h
        result["cpu_baseline"] = {
            "value": round(gcs, 5), "unit": "Gcell-updates/s", "cores": args.cpu_threads, "kind": "port",
            "sample": f"FP64 C restatement of the reference's 4-pass step (OpenMP over x), built with the "
                      f"reference's release flags (-O3 -ffast-math -mavx2 -mfma; x86-64-v3 for -march=native), on "
                      f"the same {n}^3 smooth random model, order 8, sponge 20, {len(rcv)} receivers: first {steps} "
                      f"of {nt} steps, {sec:.1f} s",
            "single_core_value": round(gcs1, 5),
            "single_core_sample": f"the same on 1 thread, first {steps1} steps, {sec1:.1f} s",
            "host": host_cpu()}
        log("cpu baseline: the reference's 7-point stencil on the two-layer 128^3 case, 500 steps")
        result["cpu_baseline"]["fd_order2_c1"] = cpu_fd_reference_order(args.cpu_threads)
        if not args.no_cg:
            log("cpu baseline: SpMV, SpMV^T and Jacobi CGLS on the 200k-ray 64^3 system")
            result["cpu_baseline"]["cg"] = cpu_cg_baseline(dev, args.cpu_threads)
        log("accuracy: 128^3, 500 steps, FP32 GPU vs FP64 oracle")
        result["accuracy"] = [
            accuracy_check(dev, args.cpu_threads, synthetic.two_layer_velocity(128, 128, 128, split_iz=64),
                           "reference", "two-layer 1500/2500 m/s (split iz = 64)"),
            accuracy_check(dev, args.cpu_threads, synthetic.smooth_random_velocity(128, 128, 128, seed=2026, sigma=5.0),
                           args.sponge, "smooth random 1500-4500 m/s, seed 2026, sigma 5 cells")]
        log("accuracy: RTM image and gradient at 64^3 vs the FP64 oracle")
        result["adjoint_accuracy"] = adjoint_accuracy(dev, args.cpu_threads, sponge=args.sponge)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
