// Field health checks and repair (drop-in for the reference's
// src/cpp/core/fault_tolerance.h:13-97).  The counting, clamping and
// neighbour-mean repair run on the GPU (stp_count_nonfinite_f64,
// stp_clamp_field_f64, stp_repair_field_nan_f64); host arrays are copied in
// and out.  The cross-rank parts (global checks, rank exclusion, synchronised
// abort) reduce to their single-process meaning.
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace seismic {

enum class HealthStatus : int {   // fault_tolerance.h:13-20
    HEALTHY = 0,
    NAN_DETECTED = 1,
    INF_DETECTED = 2,
    CFL_VIOLATION = 3,
    DIVERGENCE = 4,
    COMM_FAILURE = 5
};

struct FaultRecord {              // fault_tolerance.h:22-29
    int rank;
    HealthStatus status;
    int64_t timestep;
    int64_t cell_index;
    double offending_value;
    double velocity_at_fault;
};

class FaultTolerance {
public:
    FaultTolerance(int rank = 0, int size = 1);
    // the reference's (MPI_Comm, rank, size) constructor; the communicator is unused
    template <class Comm>
    FaultTolerance(Comm, int rank, int size) : FaultTolerance(rank, size) {}

    int global_health_check(HealthStatus local_status) const { return static_cast<int>(local_status); }
    int global_health_check_field(const double* data, int64_t count) const;
    int broadcast_error(int local_error) const { return local_error; }
    [[noreturn]] void synchronized_abort(int error_code, const std::string& message) const;

    // per-axis vmax * dt / h <= cfl_limit over the finite velocities
    // (fault_tolerance.cpp:49-70)
    bool verify_cfl_condition(const double* velocity, int64_t count, double dx, double dy, double dz, double dt,
                              double cfl_limit = 0.5) const;

    static int64_t count_nan_inf(const double* data, int64_t count);
    static bool is_valid_double(double val) { return std::isfinite(val); }
    static double safe_clamp(double val, double lo, double hi);
    static void clamp_field(double* data, int64_t count, double lo, double hi);
    void repair_field_nan(double* data, int64_t count, int64_t nx, int64_t ny, int64_t nz, int halo_width,
                          double replace_value = 0.0) const;

    void identify_corrupted_ranks(const std::vector<int>& rank_statuses, std::vector<int>& corrupted_ranks) const;
    void build_exclusion_mask(const std::vector<int>& corrupted_ranks, std::vector<uint8_t>& mask) const;

    struct FaultStats {           // fault_tolerance.h:61-67
        int64_t total_nan_count;
        int64_t total_inf_count;
        int64_t cells_repaired;
        int64_t timesteps_recovered;
        int global_abort_count;
    };
    const FaultStats& stats() const { return stats_; }
    void reset_stats();
    // a fused FD step found and replaced non-finite values at `timestep`
    // (its guards repair in place, so no cell count is known)
    void record_step_repair(int64_t timestep);

    // setters clamp as the reference's do: attempts >= 1, limit >= 1, factor
    // in [0.01, 0.99] (fault_tolerance.cpp:186-196)
    void set_max_repair_attempts(int max_attempts);
    int max_repair_attempts() const { return max_repair_attempts_; }
    void set_amplitude_limit(double limit);
    double amplitude_limit() const { return amplitude_limit_; }
    void set_cfl_safety_factor(double factor);
    double cfl_safety_factor() const { return cfl_safety_factor_; }

    // up to max_repair_attempts rounds of neighbour-mean repair of both
    // fields, then clamp to +-amplitude_limit; true when no non-finite value
    // is left (fault_tolerance.cpp:198-257)
    bool attempt_recovery(double* field_current, double* field_previous, int64_t count, int64_t nx, int64_t ny,
                          int64_t nz, int halo_width, int64_t timestep);

    void log_fault(const FaultRecord& record) { fault_log_.push_back(record); }
    const std::vector<FaultRecord>& fault_log() const { return fault_log_; }

private:
    int rank_;
    int size_;
    FaultStats stats_;
    int max_repair_attempts_;
    double amplitude_limit_;
    double cfl_safety_factor_;
    std::vector<FaultRecord> fault_log_;
};

}  // namespace seismic
