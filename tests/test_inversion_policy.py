"""The inversion loop's fault policy against the reference's rules
(src/python/seismic_tomography/inversion.py:20-88), on the CPU."""
import math

import torch

from seismic_engine.inversion import FaultRecoveryPolicy


def test_defaults_are_the_reference_ones():
    p = FaultRecoveryPolicy()
    assert (p.max_consecutive_faults, p.nan_ratio_threshold, p.step_reduction_factor, p.min_step_length,
            p.velocity_clamp_range) == (5, 0.1, 0.5, 1e-8, (1450.0, 8500.0))
    assert p.stats == {"total_faults": 0, "consecutive_faults": 0, "recovered_iterations": 0,
                       "skipped_iterations": 0}
    # positional arguments in the reference's order
    q = FaultRecoveryPolicy(3, 0.2, 0.25, 1e-3, (1000.0, 2000.0))
    assert (q.max_consecutive_faults, q.nan_ratio_threshold, q.step_reduction_factor, q.min_step_length,
            q.velocity_clamp_range) == (3, 0.2, 0.25, 1e-3, (1000.0, 2000.0))


def test_abort_rules_and_counters():
    p = FaultRecoveryPolicy(max_consecutive_faults=3, nan_ratio_threshold=0.1)
    assert p.should_continue(0.0)
    assert p.should_continue(0.1)            # the threshold itself is allowed
    assert not p.should_continue(0.1000001)
    assert p.should_continue(math.nan)       # a NaN ratio is not above the threshold (as in the reference)
    for _ in range(2):
        p.record_fault()
    assert p.should_continue(0.0)
    p.record_fault()
    assert not p.should_continue(0.0)        # three in a row
    p.record_recovery()
    assert p.should_continue(0.0)
    p.record_fault()
    p.record_skip()
    p.record_skip()
    assert p.stats == {"total_faults": 4, "consecutive_faults": 1, "recovered_iterations": 1,
                       "skipped_iterations": 2}
    # counters are per policy
    assert FaultRecoveryPolicy().stats["total_faults"] == 0


def test_step_reduction_and_floor():
    p = FaultRecoveryPolicy(step_reduction_factor=0.5, min_step_length=1e-3)
    assert p.reduced_step_length(0.01) == 0.005
    assert p.reduced_step_length(0.0015) == 1e-3
    assert p.reduced_step_length(0.0) == 1e-3


def test_velocity_clamp():
    p = FaultRecoveryPolicy(velocity_clamp_range=(1450.0, 8500.0))
    v = torch.tensor([math.nan, math.inf, -math.inf, 100.0, 1450.0, 3000.0, 8500.0, 9e3], dtype=torch.float64)
    got = p.clamp_velocity(v)
    assert got.tolist() == [1450.0, 1450.0, 1450.0, 1450.0, 1450.0, 3000.0, 8500.0, 8500.0]
    assert got.dtype == torch.float64
