// WaveEquationSolver on one MI355X (reference interface:
// src/cpp/solver/wave_equation.h:50-144).  Each step is one launch of the
// fused FD kernel through stp_fd_propagate; the host mirror of the fields is
// refreshed only when read.
#include "solver/wave_equation.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <utility>

#include "device.h"

namespace seismic {

namespace {

stp_fd_plan* make_plan(const CartesianGrid& g, const SolverConfig& c, int order, double amp, double cfl) {
    const auto& l = g.local();
    const auto& p = g.params();
    stp_fd_config fc{};
    fc.nx = l.nx_local;
    fc.ny = l.ny_local;
    fc.nz = l.nz_local;
    fc.halo_width = p.halo_width;
    fc.order = order;
    fc.dx = p.dx;
    fc.dy = p.dy;
    fc.dz = p.dz;
    fc.dt = c.dt;
    fc.sponge_width = c.sponge_width;
    fc.sponge_damping = c.sponge_damping;
    fc.amplitude_limit = amp;
    fc.cfl_safety_factor = cfl;
    fc.nshots = 1;
    stp_fd_plan* plan = nullptr;
    dev::check(stp_fd_plan_create(&fc, &plan), "WaveEquationSolver");
    return plan;
}

}  // namespace

WaveEquationSolver::WaveEquationSolver(const MpiContext& mpi, const CartesianGrid& grid, const SolverConfig& config)
    : mpi_(mpi), grid_(grid), config_(config), ft_(mpi.config().rank, mpi.config().size) {
    const int hw = grid_.params().halo_width;
    order_ = config_.fd_order > 0 ? config_.fd_order : (hw >= 4 ? 8 : 2);
    if (order_ != 2 && order_ != 8) throw std::invalid_argument("WaveEquationSolver: fd_order must be 2 or 8");
    if (hw < order_ / 2) throw std::invalid_argument("WaveEquationSolver: halo width smaller than order / 2");
    ncell_ = grid_.local().total_cells_with_halo;
    const size_t n = static_cast<size_t>(ncell_);
    field_.current.assign(n, 0.0);
    field_.previous.assign(n, 0.0);
    field_.next.assign(n, 0.0);
    field_.velocity.assign(n, 0.0);
    plan_ = make_plan(grid_, config_, order_, amplitude_limit_, ft_.cfl_safety_factor());
    cur_ = dev::alloc<float>(ncell_);
    prev_ = dev::alloc<float>(ncell_);
    coef_ = dev::alloc<float>(ncell_);
    vel_ = dev::alloc<float>(ncell_);
    sponge_ = dev::alloc<float>(ncell_);
    const int64_t bytes = static_cast<int64_t>(sizeof(float)) * ncell_;
    dev::check(stp_memset_d(cur_, 0, bytes, nullptr), "WaveEquationSolver");
    dev::check(stp_memset_d(prev_, 0, bytes, nullptr), "WaveEquationSolver");
    dev::check(stp_memset_d(vel_, 0, bytes, nullptr), "WaveEquationSolver");
    dev::check(stp_fd_sponge_factors(static_cast<stp_fd_plan*>(plan_), sponge_, nullptr), "WaveEquationSolver");
    sponge_host_.resize(n);
    dev::download(sponge_host_.data(), sponge_, ncell_);
    set_centre_source();
}

WaveEquationSolver::~WaveEquationSolver() {
    rtm_engine_.reset();
    if (plan_) stp_fd_plan_destroy(static_cast<stp_fd_plan*>(plan_));
    for (float** p : {&cur_, &prev_, &coef_, &vel_, &sponge_, &wave_dev_, &traces_dev_}) dev::release(*p);
    dev::release(status_dev_);
}

// ------------------------------------------------------------------ host mirror
void WaveEquationSolver::sync_to_host() const {
    if (!host_stale_) return;
    std::vector<float> buf(static_cast<size_t>(ncell_));
    dev::download(buf.data(), cur_, ncell_);
    std::copy(buf.begin(), buf.end(), field_.current.begin());
    dev::download(buf.data(), prev_, ncell_);
    for (size_t i = 0; i < buf.size(); ++i)
        field_.previous[i] = prev_damped_ ? buf[i] : static_cast<double>(buf[i]) * sponge_host_[i];
    host_stale_ = false;
}

void WaveEquationSolver::sync_to_device() {
    if (!host_dirty_) return;
    std::vector<float> buf(static_cast<size_t>(ncell_));
    std::copy(field_.current.begin(), field_.current.end(), buf.begin());
    dev::upload(cur_, buf.data(), ncell_);
    std::copy(field_.previous.begin(), field_.previous.end(), buf.begin());
    dev::upload(prev_, buf.data(), ncell_);
    prev_damped_ = true;   // the mirror holds `previous` as the reference does, sponge applied
    std::copy(field_.velocity.begin(), field_.velocity.end(), buf.begin());
    dev::upload(vel_, buf.data(), ncell_);
    coef_stale_ = true;
    host_dirty_ = false;
}

const WaveField& WaveEquationSolver::field() const {
    sync_to_host();
    return field_;
}

WaveField& WaveEquationSolver::mutable_field() {
    sync_to_host();
    host_dirty_ = true;
    return field_;
}

float* WaveEquationSolver::current_device() {
    sync_to_device();
    host_stale_ = true;
    return cur_;
}

float* WaveEquationSolver::previous_device() {
    sync_to_device();
    host_stale_ = true;
    return prev_;
}

// ------------------------------------------------------------------ model
void WaveEquationSolver::refresh_coefficients() {
    if (!coef_stale_) return;
    dev::check(stp_fd_coefficients(static_cast<stp_fd_plan*>(plan_), vel_, coef_, &vmax_, nullptr),
               "WaveEquationSolver coefficients");
    coef_stale_ = false;
}

void WaveEquationSolver::initialize_velocity(double vp0) {
    sync_to_host();
    std::fill(field_.velocity.begin(), field_.velocity.end(), vp0);
    std::vector<float> buf(static_cast<size_t>(ncell_), static_cast<float>(vp0));
    dev::upload(vel_, buf.data(), ncell_);
    coef_stale_ = true;
}

void WaveEquationSolver::set_velocity_field(const double* vp, int64_t count) {
    sync_to_host();
    const int64_t n = std::min<int64_t>(std::max<int64_t>(count, 0), ncell_);
    std::copy(vp, vp + n, field_.velocity.begin());
    std::vector<float> buf(field_.velocity.begin(), field_.velocity.end());
    dev::upload(vel_, buf.data(), ncell_);
    coef_stale_ = true;
}

bool WaveEquationSolver::verify_cfl() const {
    const_cast<WaveEquationSolver*>(this)->sync_to_device();
    const_cast<WaveEquationSolver*>(this)->refresh_coefficients();
    int32_t ok = 0;
    dev::check(stp_fd_verify_cfl(static_cast<const stp_fd_plan*>(plan_), vmax_, &ok), "verify_cfl");
    return ok != 0;
}

void WaveEquationSolver::set_amplitude_limit(double limit) {
    amplitude_limit_ = std::max(1.0, limit);
    ft_.set_amplitude_limit(limit);
    // the limit is part of the plan: rebuild it with the same receivers
    stp_fd_plan_destroy(static_cast<stp_fd_plan*>(plan_));
    plan_ = make_plan(grid_, config_, order_, amplitude_limit_, ft_.cfl_safety_factor());
    set_centre_source();
    if (!receivers_.empty())
        dev::check(stp_fd_set_receivers(static_cast<stp_fd_plan*>(plan_), receivers_.data(),
                                        static_cast<int32_t>(receivers_.size())),
                   "set_amplitude_limit");
    coef_stale_ = true;
}

// ------------------------------------------------------------------ sources, receivers
void WaveEquationSolver::set_centre_source() {
    const auto& l = grid_.local();
    const int hw = grid_.params().halo_width;
    const int64_t idx = grid_.local_index_with_halo(l.nx_local / 2 + hw, l.ny_local / 2 + hw, l.nz_local / 2 + hw);
    dev::check(stp_fd_set_sources(static_cast<stp_fd_plan*>(plan_), &idx, 1), "WaveEquationSolver source");
}

void WaveEquationSolver::set_receivers(const std::vector<int64_t>& idx) {
    dev::check(stp_fd_set_receivers(static_cast<stp_fd_plan*>(plan_), idx.empty() ? nullptr : idx.data(),
                                    static_cast<int32_t>(idx.size())),
               "set_receivers");
    receivers_ = idx;
}

void WaveEquationSolver::inject_source(int64_t ix, int64_t iy, int64_t iz, const std::vector<double>& wavelet) {
    const int hw = grid_.params().halo_width;
    const int64_t idx = grid_.local_index_with_halo(ix + hw, iy + hw, iz + hw);
    if (idx < 0 || idx >= ncell_ || wavelet.empty() || !std::isfinite(wavelet[0])) return;
    sync_to_device();
    float v = 0.f;
    dev::download(&v, cur_ + idx, 1);
    v = static_cast<float>(static_cast<double>(v) + wavelet[0]);
    dev::upload(cur_ + idx, &v, 1);
    host_stale_ = true;
}

// ------------------------------------------------------------------ stepping
void WaveEquationSolver::ensure_wave(int64_t n) {
    if (n <= wave_cap_) return;
    dev::release(wave_dev_);
    wave_dev_ = dev::alloc<float>(n);
    wave_cap_ = n;
}

void WaveEquationSolver::ensure_traces(int64_t n) {
    if (n <= traces_cap_) return;
    dev::release(traces_dev_);
    traces_dev_ = dev::alloc<float>(n);
    traces_cap_ = n;
}

void WaveEquationSolver::ensure_status(int64_t n) {
    if (n <= status_cap_) return;
    dev::release(status_dev_);
    status_dev_ = dev::alloc<int32_t>(n);
    status_cap_ = n;
}

int WaveEquationSolver::advance(int n, int t0, bool inject, int nw, bool record, int trace_stride, int trace_offset) {
    if (n <= 0) return 0;
    sync_to_device();
    refresh_coefficients();
    ensure_status(n);
    dev::check(stp_memset_d(status_dev_, 0, static_cast<int64_t>(sizeof(int32_t)) * n, nullptr), "step");
    float* tr = record && !receivers_.empty() ? traces_dev_ + trace_offset : nullptr;
    dev::check(stp_fd_propagate(static_cast<stp_fd_plan*>(plan_), cur_, prev_, coef_, n, prev_damped_ ? 1 : 0,
                                inject ? wave_dev_ : nullptr, inject ? nw : 0, t0, tr, trace_stride, status_dev_,
                                nullptr),
               "step");
    if (n % 2) std::swap(cur_, prev_);
    prev_damped_ = false;
    host_stale_ = true;
    std::vector<int32_t> st(static_cast<size_t>(n));
    dev::download(st.data(), status_dev_, n);
    int any = 0;
    for (int s = 0; s < n; ++s)
        if (st[static_cast<size_t>(s)]) {
            any |= st[static_cast<size_t>(s)];
            ft_.record_step_repair(t0 + s);
        }
    return any;
}

static StepResult status_of(int bits) {
    return (bits & STP_STEP_NAN_DETECTED_AND_REPAIRED) ? StepResult::NAN_DETECTED_AND_REPAIRED : StepResult::OK;
}

StepResult WaveEquationSolver::step_forward(int timestep) { return status_of(advance(1, timestep, false, 0, false, 1, 0)); }
StepResult WaveEquationSolver::step_adjoint(int timestep) { return status_of(advance(1, timestep, false, 0, false, 1, 0)); }

PropagationResult WaveEquationSolver::forward_propagate(const std::vector<double>& wavelet) {
    reset_field();
    ft_.reset_stats();
    PropagationResult r{StepResult::OK, 0, 0, 0, false};
    if (!verify_cfl()) {
        r.status = StepResult::CFL_VIOLATION;
        return r;
    }
    const int nt = config_.nt;
    const int nw = static_cast<int>(std::min<size_t>(wavelet.size(), static_cast<size_t>(std::max(nt, 0))));
    set_centre_source();
    ensure_wave(std::max(nw, 1));
    std::vector<float> w(wavelet.begin(), wavelet.begin() + nw);
    dev::upload(wave_dev_, w.data(), nw);
    const int64_t nrcv = static_cast<int64_t>(receivers_.size());
    if (nrcv) ensure_traces(nrcv * std::max(nt, 1));
    const int si = config_.snapshot_interval;
    int t = 0, bits = 0;
    while (t < nt) {
        // run to the next snapshot step (inclusive), or to the end
        int stop = nt - 1;
        if (si > 0) stop = std::min(nt - 1, (t % si == 0) ? t : (t / si + 1) * si);
        const int n = stop - t + 1;
        bits |= advance(n, t, nw > 0, nw, nrcv > 0, nt, t);
        if (si > 0 && stop % si == 0) save_snapshot(stop);
        t = stop + 1;
    }
    if (nrcv) {
        std::vector<float> h(static_cast<size_t>(nrcv * nt));
        dev::download(h.data(), traces_dev_, nrcv * nt);
        traces_.assign(h.begin(), h.end());
    }
    r.status = status_of(bits);
    r.timesteps_completed = nt;
    r.total_cells_repaired = ft_.stats().cells_repaired;
    r.completed_successfully = true;
    return r;
}

PropagationResult WaveEquationSolver::adjoint_propagate(const std::vector<double>& residual) {
    PropagationResult r{StepResult::OK, 0, 0, 0, false};
    const int nt = config_.nt;
    if (nt <= 0) {
        r.completed_successfully = true;
        return r;
    }
    // step s runs t = nt-1-s and injects residual[t] when t < |residual|;
    // NaN marks "no sample" (the kernel skips non-finite samples)
    std::vector<float> w(static_cast<size_t>(nt), std::numeric_limits<float>::quiet_NaN());
    for (int s = 0; s < nt; ++s) {
        const size_t t = static_cast<size_t>(nt - 1 - s);
        if (t < residual.size()) w[static_cast<size_t>(s)] = static_cast<float>(residual[t]);
    }
    set_centre_source();
    ensure_wave(nt);
    dev::upload(wave_dev_, w.data(), nt);
    r.status = status_of(advance(nt, 0, true, nt, false, 1, 0));
    r.timesteps_completed = nt;
    r.total_cells_repaired = ft_.stats().cells_repaired;
    r.completed_successfully = true;
    return r;
}

void WaveEquationSolver::apply_sponge_layer() {
    sync_to_host();
    for (size_t i = 0; i < field_.current.size(); ++i) {
        field_.current[i] *= sponge_host_[i];
        field_.previous[i] *= sponge_host_[i];
    }
    host_dirty_ = true;
}

void WaveEquationSolver::save_snapshot(int) {
    const auto& l = grid_.local();
    const int hw = grid_.params().halo_width;
    const int64_t n = l.total_cells;
    const size_t base = field_.snapshots.size();
    field_.snapshots.resize(base + static_cast<size_t>(n));
    if (!host_stale_) {
        // the host mirror is current (set through mutable_field, or no step since)
        size_t k = base;
        for (int64_t ix = 0; ix < l.nx_local; ++ix)
            for (int64_t iy = 0; iy < l.ny_local; ++iy)
                for (int64_t iz = 0; iz < l.nz_local; ++iz) {
                    const double v = field_.current[static_cast<size_t>(
                        grid_.local_index_with_halo(ix + hw, iy + hw, iz + hw))];
                    field_.snapshots[k++] = std::isfinite(v) ? v : 0.0;
                }
    } else {
        // the device is ahead: gather the interior there, download only it
        dev::Buffer<double> d(n);
        const int64_t nxh = l.nx_local + 2 * hw, nyh = l.ny_local + 2 * hw, nzh = l.nz_local + 2 * hw;
        dev::check(stp_field_interior_f64(cur_, nxh, nyh, nzh, hw, d.get(), nullptr), "save_snapshot");
        dev::download(field_.snapshots.data() + base, d.get(), n);
    }
    ++snapshots_captured_;
}

double WaveEquationSolver::compute_energy() const {
    double e = 0.0;
    if (!host_stale_) {
        for (double v : field_.current)
            if (std::isfinite(v)) e += v * v;
    } else {
        dev::check(stp_field_energy_f32(cur_, ncell_, &e, nullptr), "compute_energy");
    }
    return mpi_.allreduce_sum(e);
}

int64_t WaveEquationSolver::nan_count_current() const {
    sync_to_host();
    return FaultTolerance::count_nan_inf(field_.current.data(), ncell_);
}

void WaveEquationSolver::reset_field() {
    const int64_t bytes = static_cast<int64_t>(sizeof(float)) * ncell_;
    if (host_dirty_) sync_to_device();   // keep a velocity edited through mutable_field()
    dev::check(stp_memset_d(cur_, 0, bytes, nullptr), "reset_field");
    dev::check(stp_memset_d(prev_, 0, bytes, nullptr), "reset_field");
    std::fill(field_.current.begin(), field_.current.end(), 0.0);
    std::fill(field_.previous.begin(), field_.previous.end(), 0.0);
    std::fill(field_.next.begin(), field_.next.end(), 0.0);
    field_.snapshots.clear();
    prev_damped_ = false;
    host_stale_ = false;
}

// ------------------------------------------------------------------ RTM
void WaveEquationSolver::configure_rtm(const RTMConfig& cfg) {
    rtm_engine_.reset();
    rtm_cfg_ = cfg;
    if (rtm_cfg_.enable_rtm || rtm_cfg_.enable_snapshots) {
        const auto& l = grid_.local();
        const auto& p = grid_.params();
        rtm_engine_ = std::make_unique<RTMEngine>(l.nx_local, l.ny_local, l.nz_local, p.dx, p.dy, p.dz,
                                                  p.halo_width);
        rtm_engine_->set_slice_specs(rtm_cfg_.slice_specs);
    }
}

PropagationResult WaveEquationSolver::run_rtm(const std::vector<double>& source_wavelet,
                                              const std::vector<double>& receiver_residuals) {
    PropagationResult r{StepResult::OK, 0, 0, 0, false};
    if (!rtm_engine_ || !rtm_cfg_.enable_rtm) {
        r.completed_successfully = true;
        return r;
    }
    const int nt = config_.nt;
    const int64_t nrcv = static_cast<int64_t>(receivers_.size());
    if (nrcv == 0) {
        // the reference's own input: a with-halo residual field
        if (static_cast<int64_t>(receiver_residuals.size()) >= ncell_)
            return run_rtm_reference_flow(source_wavelet, receiver_residuals);
        return r;   // neither form: not completed
    }
    if (static_cast<int64_t>(receiver_residuals.size()) < nrcv * nt) return r;
    rtm_engine_->reset_image();
    sync_to_device();
    refresh_coefficients();
    set_centre_source();
    const int nw = static_cast<int>(std::min<size_t>(source_wavelet.size(), static_cast<size_t>(nt)));
    ensure_wave(std::max(nw, 1));
    std::vector<float> w(source_wavelet.begin(), source_wavelet.begin() + nw);
    dev::upload(wave_dev_, w.data(), nw);
    const int64_t nint = rtm_engine_->total_cells();
    dev::Buffer<float> inj(nrcv * nt), img(nint), ill(nint);
    ensure_traces(nrcv * nt);
    std::vector<float> res(receiver_residuals.begin(), receiver_residuals.begin() + nrcv * nt);
    dev::upload(inj.get(), res.data(), nrcv * nt);
    dev::check(stp_fd_adjoint(static_cast<stp_fd_plan*>(plan_), coef_, wave_dev_, nw, nt, nullptr, inj.get(), 1, 0,
                              rtm_cfg_.checkpoint_interval, img.get(), ill.get(), traces_dev_, nullptr, nullptr),
               "run_rtm");
    rtm_engine_->accumulate_device_f32(img.get(), ill.get());
    rtm_engine_->finalize_image();
    std::vector<float> h(static_cast<size_t>(nrcv * nt));
    dev::download(h.data(), traces_dev_, nrcv * nt);
    traces_.assign(h.begin(), h.end());
    r.timesteps_completed = nt;
    r.completed_successfully = true;
    return r;
}

// The reference's run_rtm as written (wave_equation.cpp:621-688): the image
// condition of the forward field with itself after every step (slices
// captured every RTMConfig::snapshot_interval steps), then one term of the
// final forward field with the residual field, then finalize_image.
PropagationResult WaveEquationSolver::run_rtm_reference_flow(const std::vector<double>& source_wavelet,
                                                             const std::vector<double>& residual_field) {
    PropagationResult r{StepResult::OK, 0, 0, 0, false};
    const auto& l = grid_.local();
    const int hw = grid_.params().halo_width;
    const int64_t nyh = l.ny_local + 2 * hw, nzh = l.nz_local + 2 * hw;
    rtm_engine_->reset_image();
    reset_field();
    dev::Buffer<double> f64(ncell_), res(ncell_);
    const int64_t bytes = static_cast<int64_t>(sizeof(double)) * ncell_;
    auto current_f64 = [&]() {
        dev::check(stp_memset_d(f64.get(), 0, bytes, nullptr), "run_rtm");
        dev::check(stp_accumulate_f32_f64(ncell_, cur_, f64.get(), nullptr), "run_rtm");
    };
    const int si = std::max(1, rtm_cfg_.snapshot_interval);
    for (int t = 0; t < config_.nt; ++t) {
        if (t < static_cast<int>(source_wavelet.size()))
            inject_source(l.nx_local / 2, l.ny_local / 2, l.nz_local / 2, {source_wavelet[static_cast<size_t>(t)]});
        const StepResult sr = step_forward(t);
        if (sr == StepResult::FATAL_DIVERGENCE) {
            r.status = sr;
            r.timesteps_completed = t;
            return r;
        }
        if (sr == StepResult::NAN_DETECTED_AND_REPAIRED) r.status = sr;
        current_f64();
        rtm_engine_->apply_image_condition_device(f64.get(), f64.get(), nyh, nzh, t);
        if (t % si == 0 && rtm_engine_->capturing()) {
            const std::vector<double>& cur = field().current;
            rtm_engine_->extract_slices(cur.data(), cur.data(), nullptr, nyh, nzh, t);
        }
        ++r.timesteps_completed;
    }
    dev::upload(res.get(), residual_field.data(), ncell_);
    current_f64();
    rtm_engine_->apply_image_condition_device(f64.get(), res.get(), nyh, nzh, config_.nt);
    rtm_engine_->finalize_image();
    r.total_cells_repaired = ft_.stats().cells_repaired;
    r.completed_successfully = true;
    return r;
}

}  // namespace seismic
