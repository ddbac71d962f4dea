// FaultTolerance (reference interface: src/cpp/core/fault_tolerance.h:31-97).
// Field passes run on the GPU through the C ABI; host arrays are staged.
#include "core/fault_tolerance.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "device.h"

namespace seismic {

FaultTolerance::FaultTolerance(int rank, int size)
    : rank_(rank), size_(size), max_repair_attempts_(3), amplitude_limit_(1e10), cfl_safety_factor_(0.49) {
    reset_stats();
}

int FaultTolerance::global_health_check_field(const double* data, int64_t count) const {
    return count_nan_inf(data, count) > 0 ? 1 : 0;
}

void FaultTolerance::synchronized_abort(int error_code, const std::string& message) const {
    std::fprintf(stderr, "[rank %d] fatal: %s (error code %d)\n", rank_, message.c_str(), error_code);
    std::fflush(stderr);
    std::exit(error_code);
}

bool FaultTolerance::verify_cfl_condition(const double* velocity, int64_t count, double dx, double dy, double dz,
                                          double dt, double cfl_limit) const {
    double vmax = 0.0;
    for (int64_t i = 0; i < count; ++i)
        if (std::isfinite(velocity[i])) vmax = std::max(vmax, velocity[i]);
    const double hmin = std::min(dx, std::min(dy, dz));
    return vmax * dt / hmin <= cfl_limit;
}

int64_t FaultTolerance::count_nan_inf(const double* data, int64_t count) {
    if (count <= 0) return 0;
    dev::Buffer<double> d(count);
    dev::upload(d.get(), data, count);
    int64_t n = 0;
    dev::check(stp_count_nonfinite_f64(d.get(), count, &n, nullptr), "count_nan_inf");
    return n;
}

double FaultTolerance::safe_clamp(double val, double lo, double hi) {
    return std::isfinite(val) ? std::min(std::max(val, lo), hi) : 0.0;
}

void FaultTolerance::clamp_field(double* data, int64_t count, double lo, double hi) {
    if (count <= 0) return;
    dev::Buffer<double> d(count);
    dev::upload(d.get(), data, count);
    dev::check(stp_clamp_field_f64(d.get(), count, lo, hi, nullptr), "clamp_field");
    dev::download(data, d.get(), count);
}

void FaultTolerance::repair_field_nan(double* data, int64_t count, int64_t nx, int64_t ny, int64_t nz,
                                      int halo_width, double replace_value) const {
    if (count <= 0) return;
    dev::Buffer<double> d(count);
    dev::upload(d.get(), data, count);
    dev::check(stp_repair_field_nan_f64(d.get(), count, nx, ny, nz, halo_width, replace_value, nullptr),
               "repair_field_nan");
    dev::download(data, d.get(), count);
}

void FaultTolerance::identify_corrupted_ranks(const std::vector<int>& rank_statuses,
                                              std::vector<int>& corrupted_ranks) const {
    corrupted_ranks.clear();
    for (size_t r = 0; r < rank_statuses.size(); ++r)
        if (rank_statuses[r] != 0) corrupted_ranks.push_back(static_cast<int>(r));
}

void FaultTolerance::build_exclusion_mask(const std::vector<int>& corrupted_ranks, std::vector<uint8_t>& mask) const {
    mask.assign(static_cast<size_t>(size_), 1);
    for (int r : corrupted_ranks)
        if (r >= 0 && r < size_) mask[static_cast<size_t>(r)] = 0;
}

void FaultTolerance::reset_stats() {
    std::memset(&stats_, 0, sizeof stats_);
    fault_log_.clear();
}

void FaultTolerance::record_step_repair(int64_t timestep) {
    stats_.timesteps_recovered += 1;
    log_fault(FaultRecord{rank_, HealthStatus::NAN_DETECTED, timestep, -1, 0.0, 0.0});
}

void FaultTolerance::set_max_repair_attempts(int max_attempts) { max_repair_attempts_ = std::max(1, max_attempts); }
void FaultTolerance::set_amplitude_limit(double limit) { amplitude_limit_ = std::max(1.0, limit); }
void FaultTolerance::set_cfl_safety_factor(double factor) {
    cfl_safety_factor_ = std::min(0.99, std::max(0.01, factor));
}

bool FaultTolerance::attempt_recovery(double* cur, double* prev, int64_t count, int64_t nx, int64_t ny, int64_t nz,
                                      int halo_width, int64_t timestep) {
    if (count <= 0) return true;
    dev::Buffer<double> dc(count), dp(count);
    dev::upload(dc.get(), cur, count);
    dev::upload(dp.get(), prev, count);
    int64_t bad = 0;
    dev::check(stp_count_nonfinite_f64(dc.get(), count, &bad, nullptr), "attempt_recovery");
    if (bad == 0) return true;
    stats_.total_nan_count += bad;
    FaultRecord rec{rank_, HealthStatus::NAN_DETECTED, timestep, -1, 0.0, 0.0};
    for (int64_t i = 0; i < count; ++i)
        if (!std::isfinite(cur[i])) {
            rec.cell_index = i;
            rec.offending_value = cur[i];
            break;
        }
    log_fault(rec);
    bool ok = false;
    for (int a = 0; a < max_repair_attempts_ && !ok; ++a) {
        for (double* f : {dc.get(), dp.get()}) {
            dev::check(stp_repair_field_nan_f64(f, count, nx, ny, nz, halo_width, 0.0, nullptr), "attempt_recovery");
            dev::check(stp_clamp_field_f64(f, count, -amplitude_limit_, amplitude_limit_, nullptr),
                       "attempt_recovery");
        }
        int64_t left = 0;
        dev::check(stp_count_nonfinite_f64(dc.get(), count, &left, nullptr), "attempt_recovery");
        stats_.cells_repaired += bad - left;
        bad = left;
        ok = left == 0;
    }
    if (ok) stats_.timesteps_recovered += 1;
    dev::download(cur, dc.get(), count);
    dev::download(prev, dp.get(), count);
    return ok;
}

}  // namespace seismic
