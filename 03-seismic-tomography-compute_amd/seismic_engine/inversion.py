"""The reference's runnable inversion loop (AdjointStateInversion.run,
src/python/seismic_tomography/inversion.py:124-272, with FaultRecoveryPolicy
:20-88) for the straight-ray travel-time problem, on the GPU: forward model
and Frechet gradient (traveltime.py), smoothing, Tikhonov term and
normalisation (gradient.py), health checks and repairs (health.py) are
library kernels; the loop, the line search and the fault bookkeeping are host
logic as in the reference.  The HDF5 writer and logging are not part of it.

One divergence is opt-in: the reference updates vp - step * g with g from
_compute_gradient, which is -dJ/dv for J = 1/2 sum w^2 (obs - t)^2 (the
residuals are observed minus calculated), so its steps go uphill and the line
search can only shrink them.  descent=True uses -g instead (a descent
direction); the default keeps the reference's sign."""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from .gradient import GradientComputer
from .health import check_field_health, repair_field_nan
from .traveltime import TravelTimeOperator


def _torch():
    import torch
    return torch


_FAULT_COUNTERS = ("total_faults", "consecutive_faults", "recovered_iterations", "skipped_iterations")


@dataclass
class FaultRecoveryPolicy:
    """Fault bookkeeping of the inversion loop (inversion.py:20-88): when to
    give up, how far to shrink the step after a failed update, and the
    velocity bounds a model is clamped into.  The constructor arguments,
    the method names and the keys of `stats` are the reference's, so code
    that configures or queries the policy keeps working; the rules are:
    stop once `max_consecutive_faults` faults follow one another or when
    more than `nan_ratio_threshold` of the model is non-finite; a reduced
    step is the step times `step_reduction_factor`, never below
    `min_step_length`; non-finite velocities become the lower bound."""

    max_consecutive_faults: int = 5
    nan_ratio_threshold: float = 0.1
    step_reduction_factor: float = 0.5
    min_step_length: float = 1e-8
    velocity_clamp_range: tuple = (1450.0, 8500.0)
    _count: dict = field(default_factory=lambda: dict.fromkeys(_FAULT_COUNTERS, 0), init=False, repr=False)

    def _bump(self, *names: str) -> None:
        for n in names:
            self._count[n] += 1

    def should_continue(self, nan_ratio: float) -> bool:
        too_many = self._count["consecutive_faults"] >= self.max_consecutive_faults
        return not too_many and not nan_ratio > self.nan_ratio_threshold

    def record_fault(self) -> None:
        self._bump("consecutive_faults", "total_faults")

    def record_recovery(self) -> None:
        self._bump("recovered_iterations")
        self._count["consecutive_faults"] = 0

    def record_skip(self) -> None:
        self._bump("skipped_iterations")

    def reduced_step_length(self, current_step: float) -> float:
        return max(current_step * self.step_reduction_factor, self.min_step_length)

    def clamp_velocity(self, vp):
        """Non-finite entries to the lower bound, then every entry into the
        bounds (a torch tensor, on its device)."""
        torch = _torch()
        lo, hi = self.velocity_clamp_range
        return torch.nan_to_num(vp, nan=lo, posinf=lo, neginf=lo).clamp(lo, hi)

    @property
    def stats(self) -> dict:
        return {n: self._count[n] for n in _FAULT_COUNTERS}


class TravelTimeInversion:
    def __init__(self, operator: TravelTimeOperator, observed_times, vp0, max_iterations: int = 50,
                 step_length: float = 0.01, smoothing_sigma: float = 2.0, regularization_lambda: float = 1e-4,
                 misfit_tolerance: float = 1e-6, misfit_type: str = "cross_correlation", line_search_max: int = 10,
                 line_search_factor: float = 0.5, enable_fault_tolerance: bool = True, descent: bool = False):
        torch = _torch()
        self.op = operator
        g = operator.geometry
        self.shape = (g.nx, g.ny, g.nz)
        self.observed = torch.as_tensor(np.asarray(observed_times, np.float64) if not torch.is_tensor(observed_times)
                                        else observed_times, dtype=torch.float64, device=operator.device)
        self.vp = torch.as_tensor(vp0, dtype=torch.float64).to(operator.device).reshape(-1).clone()
        self.max_iterations = max_iterations
        self.step_length = step_length
        self.misfit_tolerance = misfit_tolerance
        self.misfit_type = misfit_type
        self.line_search_max = line_search_max
        self.line_search_factor = line_search_factor
        self.enable_fault_tolerance = enable_fault_tolerance
        self.descent = descent
        self.gradient_computer = GradientComputer(*self.shape, smoothing_sigma=smoothing_sigma,
                                                  regularization_lambda=regularization_lambda)
        self.fault_policy = FaultRecoveryPolicy()
        self._current_iteration = 0
        self._misfit_history: list = []
        self._converged = False
        self._fault_events: list = []

    def _misfit(self, vp) -> float:
        return self.op.misfit(self.op.forward(vp), self.observed, self.misfit_type)

    def _record_fault_event(self, iteration: int, fault_type: str, detail):
        self._fault_events.append({"iteration": iteration, "type": fault_type, "detail": str(detail),
                                   "timestamp": time.time()})

    def _verify_model_health(self):
        torch = _torch()
        vp = self.vp
        self.vp = torch.where(torch.isfinite(vp) & (vp > 0), vp, torch.full_like(vp, 1500.0))

    def _line_search(self, gradient, current_misfit: float) -> float:
        """inversion.py:435-462."""
        step = self.step_length
        for _ in range(self.line_search_max):
            trial = self.vp - step * gradient
            if self.enable_fault_tolerance:
                trial = self.fault_policy.clamp_velocity(trial)
            trial_misfit = self._misfit(trial)
            if not math.isfinite(trial_misfit):
                step *= self.fault_policy.step_reduction_factor
                continue
            if trial_misfit < current_misfit:
                return step
            step *= self.line_search_factor
        return step

    def run(self, callback=None) -> dict:
        torch = _torch()
        nx, ny, nz = self.shape
        if self.enable_fault_tolerance:
            self._verify_model_health()
        for iteration in range(self.max_iterations):
            self._current_iteration = iteration
            if self.enable_fault_tolerance:
                nan_count = check_field_health(self.vp)
                if nan_count > 0:
                    if not self.fault_policy.should_continue(nan_count / self.vp.numel()):
                        break
                    self.fault_policy.record_fault()
                    self.vp = repair_field_nan(self.vp, nx, ny, nz, halo_width=1, replace_value=1500.0)
            times = self.op.forward(self.vp)
            misfit = self.op.misfit(times, self.observed, self.misfit_type)
            if not math.isfinite(misfit):
                self.fault_policy.record_fault()
                self.step_length = self.fault_policy.reduced_step_length(self.step_length)
                self.vp = self.fault_policy.clamp_velocity(self.vp)
                self._record_fault_event(iteration, "non_finite_misfit", misfit)
                self.fault_policy.record_skip()
                continue
            residuals = self.op.weighted_residuals(times, self.observed)
            self._misfit_history.append(misfit)
            if iteration > 0:
                prev = self._misfit_history[-2]
                if abs(prev - misfit) / max(abs(prev), 1e-30) < self.misfit_tolerance:
                    self._converged = True
                    break
            gradient = self.op.gradient(self.vp, residuals)
            if self.enable_fault_tolerance:
                bad = int((~torch.isfinite(gradient)).sum())
                if bad > 0:
                    gradient = torch.where(torch.isfinite(gradient), gradient, torch.zeros_like(gradient))
                    self.fault_policy.record_fault()
                    self._record_fault_event(iteration, "gradient_nan", bad)
            gc = self.gradient_computer
            gradient = gc.apply_smoothing(gradient)
            gradient = gc.apply_regularization(self.vp, gradient)
            gradient = gc.normalize_gradient(gradient)
            if self.descent:
                gradient = -gradient
            step = self._line_search(gradient, misfit)
            vp = self.vp - step * gradient
            if self.enable_fault_tolerance:
                vp = self.fault_policy.clamp_velocity(vp)
            self.vp = vp
            if self.enable_fault_tolerance:
                self.fault_policy.record_recovery()
            if callback is not None:
                callback(iteration, misfit, gradient, step)
        return {"converged": self._converged, "iterations": self._current_iteration + 1,
                "final_misfit": self._misfit_history[-1] if self._misfit_history else float("inf"),
                "misfit_history": self._misfit_history, "fault_events": self._fault_events,
                "fault_stats": self.fault_policy.stats if self.enable_fault_tolerance else None}
