// 3-D acoustic leapfrog solver (drop-in for the reference's
// src/cpp/solver/wave_equation.h:16-144) running on one MI355X.
//
// The wavefields live on the GPU (FP32); every step is one launch of the
// fused stencil / sponge / guard kernel (stp_fd_propagate).  field() returns a
// host mirror in FP64 that is refreshed from the device when it is read after
// a step; mutable_field() marks the mirror as edited, and the next step
// uploads it.  The halo exchange is a no-op (one GPU holds the grid).
//
// Order: SolverConfig::fd_order (2 = the reference's 7-point stencil, 8 =
// radius-4 star stencil; 0 = 8 when the halo width is >= 4, else 2).
//
// run_rtm with receivers set (set_receivers) is a true reverse-time
// migration: a checkpointed forward pass of the source wavelet, then the
// receiver residuals (one trace of nt samples per receiver) propagated
// backwards with zero-lag cross-correlation into rtm_engine()->result()
// (image += S R, illumination += S S), then finalize_image.  With no
// receivers and a with-halo residual field (n_h^3 values, the reference's
// input) it follows the reference's flow: the forward field correlated with
// itself after every step, then once with the residual field.  Inputs that
// fit neither form return completed_successfully = false (no exception).
// Snapshot ring buffers and the mmap drain thread are not part of this build.
#pragma once

#include <atomic>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "core/config.h"
#include "core/fault_tolerance.h"
#include "core/grid.h"
#include "core/mpi_context.h"
#include "rtm/rtm_engine.h"

namespace seismic {

struct WaveField {            // wave_equation.h:16-22
    std::vector<double> current;
    std::vector<double> previous;   // as the reference stores it: sponge applied
    std::vector<double> next;
    std::vector<double> velocity;
    std::vector<double> snapshots;  // interior of `current`, appended by save_snapshot
};

enum class StepResult {       // wave_equation.h:24-29
    OK = 0,
    NAN_DETECTED_AND_REPAIRED = 1,
    CFL_VIOLATION = 2,
    FATAL_DIVERGENCE = 3
};

struct PropagationResult {    // wave_equation.h:31-37
    StepResult status;
    int64_t total_nan_encountered;
    int64_t total_cells_repaired;
    int timesteps_completed;
    bool completed_successfully;
};

struct RTMConfig {            // wave_equation.h:39-48 (ring buffer / mmap fields kept, unused)
    bool enable_rtm = false;
    bool enable_snapshots = false;
    int snapshot_interval = 50;
    bool enable_mmap_writer = false;
    std::string mmap_filename = "rtm_snapshots.dat";
    int64_t mmap_max_bytes = 2LL * 1024LL * 1024LL * 1024LL;
    size_t ring_buffer_capacity = 1024;
    std::vector<SliceSpec> slice_specs;
    int checkpoint_interval = 0;   // new: forward checkpoints every k steps (0: ceil(sqrt(nt)))
};

class WaveEquationSolver {
public:
    // mpi and grid are held by reference and must outlive the solver
    WaveEquationSolver(const MpiContext& mpi, const CartesianGrid& grid, const SolverConfig& config);
    ~WaveEquationSolver();
    WaveEquationSolver(const WaveEquationSolver&) = delete;
    WaveEquationSolver& operator=(const WaveEquationSolver&) = delete;

    void initialize_velocity(double vp0);
    // copies min(count, cells with halo) values in the with-halo layout
    void set_velocity_field(const double* vp_data, int64_t count);

    // adds wavelet[0] to the current field at interior cell (ix, iy, iz);
    // non-finite samples and out-of-range cells are ignored
    void inject_source(int64_t ix, int64_t iy, int64_t iz, const std::vector<double>& wavelet);

    StepResult step_forward(int timestep);
    StepResult step_adjoint(int timestep);

    // reset, CFL check, then config.nt steps injecting wavelet[t] at the
    // interior centre before step t; snapshot every snapshot_interval steps
    PropagationResult forward_propagate(const std::vector<double>& wavelet);
    // no reset; t = nt-1 .. 0 injecting residual_wavelet[t] at the centre
    PropagationResult adjoint_propagate(const std::vector<double>& residual_wavelet);

    void exchange_halo(WaveField&) {}
    // damps current and previous in the sponge (wave_equation.cpp:262-297);
    // the steps apply the sponge themselves
    void apply_sponge_layer();

    const WaveField& field() const;
    WaveField& mutable_field();
    const CartesianGrid& grid() const { return grid_; }

    void save_snapshot(int timestep);
    void clear_snapshots() { field_.snapshots.clear(); }
    double compute_energy() const;
    void reset_field();

    FaultTolerance& fault_tolerance() { return ft_; }
    const FaultTolerance& fault_tolerance() const { return ft_; }
    bool verify_cfl() const;
    void set_amplitude_limit(double limit);
    double amplitude_limit() const { return amplitude_limit_; }
    int64_t nan_count_current() const;
    const FaultTolerance::FaultStats& fault_stats() const { return ft_.stats(); }

    void configure_rtm(const RTMConfig& rtm_cfg);
    const RTMConfig& rtm_config() const { return rtm_cfg_; }
    RTMEngine* rtm_engine() { return rtm_engine_.get(); }

    PropagationResult run_rtm(const std::vector<double>& source_wavelet,
                              const std::vector<double>& receiver_residuals);

    int64_t snapshots_captured() const { return snapshots_captured_; }
    int64_t snapshots_dropped() const { return 0; }

    // ---- new
    int order() const { return order_; }
    // receivers: interior cells (ix, iy, iz) as with-halo flat indices
    void set_receivers(const std::vector<int64_t>& with_halo_indices);
    const std::vector<int64_t>& receivers() const { return receivers_; }
    // traces recorded by the last forward_propagate / run_rtm, [receiver][t]
    const std::vector<double>& receiver_traces() const { return traces_; }
    // device pointers of the current / previous (undamped) fields and of C = v^2 dt^2
    float* current_device();
    float* previous_device();
    const float* coefficients_device() const { return coef_; }

private:
    PropagationResult run_rtm_reference_flow(const std::vector<double>& source_wavelet,
                                             const std::vector<double>& residual_field);
    const MpiContext& mpi_;
    const CartesianGrid& grid_;
    SolverConfig config_;
    FaultTolerance ft_;
    double amplitude_limit_ = 1e10;
    int order_ = 2;
    int64_t ncell_ = 0;

    void* plan_ = nullptr;          // stp_fd_plan*
    float *cur_ = nullptr, *prev_ = nullptr, *coef_ = nullptr, *vel_ = nullptr, *sponge_ = nullptr;
    float *wave_dev_ = nullptr, *traces_dev_ = nullptr;
    int32_t* status_dev_ = nullptr;
    int64_t wave_cap_ = 0, traces_cap_ = 0, status_cap_ = 0;
    bool prev_damped_ = false;      // device `previous` already carries the sponge factor
    double vmax_ = 0.0;
    bool coef_stale_ = true;

    mutable WaveField field_;
    mutable bool host_stale_ = false;   // device newer than the mirror
    bool host_dirty_ = false;           // mirror edited through mutable_field()
    std::vector<float> sponge_host_;

    std::vector<int64_t> receivers_;
    std::vector<double> traces_;

    RTMConfig rtm_cfg_;
    std::unique_ptr<RTMEngine> rtm_engine_;
    int64_t snapshots_captured_ = 0;

    void sync_to_device();
    void sync_to_host() const;
    void refresh_coefficients();
    void ensure_wave(int64_t n);
    void ensure_traces(int64_t n);
    void ensure_status(int64_t n);
    // advance n steps from time index t0, injecting wave_dev_[t] at the
    // source when `inject`; OR of the step status words
    int advance(int n, int t0, bool inject, int nw, bool record, int trace_stride, int trace_offset);
    void set_centre_source();
};

}  // namespace seismic
