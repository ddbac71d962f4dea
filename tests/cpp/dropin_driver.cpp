// Test program for the C++ drop-in layer: calls the reference-named classes
// (WaveEquationSolver, CSRMatrix, SparseOps, RTMEngine, FaultTolerance,
// load_config, CartesianGrid) the way code written against the reference
// does, reading inputs from and writing outputs to raw binary files in a
// directory, so that tests/test_dropin_cpp*.py can compare them with the
// oracle.
//
//   dropin_driver config <yaml>            print the parsed configuration
//   dropin_driver grid                     print grid index checks
//   dropin_driver fd <dir>                 forward / step / adjoint propagation
//   dropin_driver sparse <dir>             assembly, SpMV, CG, Jacobi-PCG, CGLS
//   dropin_driver rtm <dir>                imaging condition, finalize, run_rtm
//   dropin_driver smooth <dir>             smooth_gradient, gaussian_filter_3d
//   dropin_driver health <dir>             count / clamp / repair of a field
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "core/config.h"
#include "core/fault_tolerance.h"
#include "core/grid.h"
#include "core/mpi_context.h"
#include "rtm/rtm_engine.h"
#include "solver/boundary.h"
#include "rtm/mmap_writer.h"
#include "solver/wave_equation.h"
#include "sparse/csr_matrix.h"
#include "sparse/sparse_ops.h"

using namespace seismic;

namespace {

template <typename T>
std::vector<T> read(const std::string& path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("cannot read " + path);
    const std::streamsize n = f.tellg();
    f.seekg(0);
    std::vector<T> v(static_cast<size_t>(n) / sizeof(T));
    f.read(reinterpret_cast<char*>(v.data()), n);
    return v;
}

template <typename T>
void write(const std::string& path, const T* p, size_t n) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(p), static_cast<std::streamsize>(n * sizeof(T)));
}

template <typename T>
void write(const std::string& path, const std::vector<T>& v) { write(path, v.data(), v.size()); }

// key = value lines
double param(const std::string& dir, const std::string& key) {
    std::ifstream f(dir + "/params.txt");
    std::string k;
    double v;
    while (f >> k >> v)
        if (k == key) return v;
    throw std::runtime_error("missing parameter " + key);
}

const char* step_name(StepResult s) {
    switch (s) {
        case StepResult::OK: return "OK";
        case StepResult::NAN_DETECTED_AND_REPAIRED: return "NAN_DETECTED_AND_REPAIRED";
        case StepResult::CFL_VIOLATION: return "CFL_VIOLATION";
        default: return "FATAL_DIVERGENCE";
    }
}

void print_result(const char* tag, const PropagationResult& r) {
    std::printf("%s status=%s steps=%d ok=%d\n", tag, step_name(r.status), r.timesteps_completed,
                r.completed_successfully ? 1 : 0);
}

int cmd_config(const std::string& path) {
    const TomographyConfig c = load_config(path);
    std::printf("solver.frequency %.17g\nsolver.dt %.17g\nsolver.nt %d\nsolver.water_depth %.17g\n",
                c.solver.frequency, c.solver.dt, c.solver.nt, c.solver.water_depth);
    std::printf("solver.seafloor_gradient %.17g\nsolver.sponge_width %d\nsolver.sponge_damping %.17g\n",
                c.solver.seafloor_gradient, c.solver.sponge_width, c.solver.sponge_damping);
    std::printf("solver.snapshot_interval %d\nsolver.output_dir %s\nsolver.fd_order %d\n", c.solver.snapshot_interval,
                c.solver.output_dir.c_str(), c.solver.fd_order);
    std::printf("model.vp_min %.17g\nmodel.vp_max %.17g\nmodel.vs_min %.17g\nmodel.vs_max %.17g\n", c.model.vp_min,
                c.model.vp_max, c.model.vs_min, c.model.vs_max);
    std::printf("model.rho_min %.17g\nmodel.rho_max %.17g\nmodel.initial_vp %.17g\nmodel.initial_vs %.17g\n",
                c.model.rho_min, c.model.rho_max, c.model.initial_vp, c.model.initial_vs);
    std::printf("model.initial_rho %.17g\n", c.model.initial_rho);
    std::printf("inversion.max_iterations %d\ninversion.step_length %.17g\n", c.inversion.max_iterations,
                c.inversion.step_length);
    std::printf("inversion.gradient_smoothing_sigma %.17g\ninversion.regularization_lambda %.17g\n",
                c.inversion.gradient_smoothing_sigma, c.inversion.regularization_lambda);
    std::printf("inversion.misfit_tolerance %.17g\ninversion.misfit_type %s\n", c.inversion.misfit_tolerance,
                c.inversion.misfit_type.c_str());
    std::printf("inversion.gradient_precondition %d\ninversion.line_search_max %d\n",
                c.inversion.gradient_precondition, c.inversion.line_search_max);
    std::printf("inversion.line_search_factor %.17g\n", c.inversion.line_search_factor);
    std::printf("source_positions %zu\n", c.source_positions.size());
    return 0;
}

int cmd_grid() {
    GridParams p{10, 7, 5, 10.0, 20.0, 5.0, 1.0, 2.0, 3.0, 2};
    const int coords[3] = {1, 0, 0}, dims[3] = {3, 1, 1};
    CartesianGrid g(p, coords, dims);
    const LocalGrid& l = g.local();
    std::printf("nx_local %lld offset_x %lld nx_with_halo %lld total_with_halo %lld\n", (long long)l.nx_local,
                (long long)l.offset_x, (long long)l.nx_with_halo, (long long)l.total_cells_with_halo);
    std::printf("index_with_halo(2,3,4) %lld\n", (long long)g.local_index_with_halo(2, 3, 4));
    std::printf("is_halo(0,0,0) %d is_halo(2,2,2) %d x_coord(1) %.17g\n", g.is_halo(0, 0, 0) ? 1 : 0,
                g.is_halo(2, 2, 2) ? 1 : 0, g.x_coord(1));
    MpiContext mpi;
    std::printf("rank %d size %d neighbor0 %d allreduce %.1f\n", mpi.config().rank, mpi.config().size,
                mpi.config().neighbors[0], mpi.allreduce_sum(2.5));
    return 0;
}

int cmd_fd(const std::string& dir) {
    const int64_t nx = (int64_t)param(dir, "nx"), ny = (int64_t)param(dir, "ny"), nz = (int64_t)param(dir, "nz");
    const int hw = (int)param(dir, "hw");
    GridParams gp{nx, ny, nz, 10.0, 10.0, 10.0, 0.0, 0.0, 0.0, hw};
    CartesianGrid grid(gp);
    MpiContext mpi;
    SolverConfig cfg = load_config(dir + "/none.yaml").solver;
    cfg.dt = 1e-3;
    cfg.nt = (int)param(dir, "nt");
    cfg.sponge_width = (int)param(dir, "sponge_width");
    cfg.sponge_damping = 0.015;
    cfg.snapshot_interval = (int)param(dir, "snapshot_interval");
    cfg.fd_order = (int)param(dir, "order");
    const std::vector<double> vel = read<double>(dir + "/vel.f64");
    const std::vector<double> w = read<double>(dir + "/wavelet.f64");
    const std::vector<int64_t> rcv = read<int64_t>(dir + "/rcv.i64");

    WaveEquationSolver s(mpi, grid, cfg);
    s.set_velocity_field(vel.data(), (int64_t)vel.size());
    s.set_receivers(rcv);
    std::printf("order %d cfl %d\n", s.order(), s.verify_cfl() ? 1 : 0);
    const PropagationResult r = s.forward_propagate(w);
    const double energy = s.compute_energy();   // before field(): reduced on the device
    print_result("forward", r);
    write(dir + "/cur.f64", s.field().current);
    write(dir + "/prev.f64", s.field().previous);
    write(dir + "/traces.f64", s.receiver_traces());
    write(dir + "/snapshots.f64", s.field().snapshots);
    std::printf("snapshots %lld energy %.17g host_energy %.17g\n", (long long)s.snapshots_captured(), energy,
                s.compute_energy());

    // the same run by hand: inject at the centre, then step (wave_equation.cpp:118-125)
    WaveEquationSolver m(mpi, grid, cfg);
    m.set_velocity_field(vel.data(), (int64_t)vel.size());
    for (int t = 0; t < cfg.nt; ++t) {
        if (t < (int)w.size()) m.inject_source(nx / 2, ny / 2, nz / 2, {w[(size_t)t]});
        m.step_forward(t);
    }
    write(dir + "/cur_steps.f64", m.field().current);
    write(dir + "/prev_steps.f64", m.field().previous);

    // adjoint_propagate continues from the forward state (no reset)
    const PropagationResult ra = s.adjoint_propagate(w);
    print_result("adjoint", ra);
    write(dir + "/cur_adj.f64", s.field().current);

    // CFL violation: 10x the velocity fails verify_cfl before any step
    std::vector<double> fast(vel);
    for (double& v : fast) v *= 10.0;
    WaveEquationSolver f(mpi, grid, cfg);
    f.set_velocity_field(fast.data(), (int64_t)fast.size());
    print_result("fast", f.forward_propagate(w));
    return 0;
}

int cmd_sparse(const std::string& dir) {
    const int64_t rows = (int64_t)param(dir, "rows"), cols = (int64_t)param(dir, "cols");
    const std::vector<int64_t> r = read<int64_t>(dir + "/r.i64"), c = read<int64_t>(dir + "/c.i64");
    const std::vector<double> v = read<double>(dir + "/v.f64");
    std::vector<SparseEntry> e(r.size());
    for (size_t k = 0; k < r.size(); ++k) e[k] = SparseEntry{r[k], c[k], v[k]};
    CSRMatrix A;
    A.build_from_triplets(rows, cols, e);
    write(dir + "/rp.i64", A.row_ptr().data(), A.row_ptr().size());
    write(dir + "/ci.i64", A.col_ind().data(), A.col_ind().size());
    write(dir + "/val.f64", A.values().data(), A.values().size());
    std::printf("nnz %lld frobenius %.17g\n", (long long)A.nnz(), A.frobenius_norm());

    const std::vector<double> xv = read<double>(dir + "/x.f64");
    AlignedVector<double> x(xv.size());
    for (size_t i = 0; i < xv.size(); ++i) x[i] = xv[i];
    const AlignedVector<double> y = A.spmv(x);
    write(dir + "/y.f64", y.data(), y.size());
    const std::vector<double> uv = read<double>(dir + "/u.f64");
    AlignedVector<double> u(uv.size());
    for (size_t i = 0; i < uv.size(); ++i) u[i] = uv[i];
    const AlignedVector<double> yt = A.spmv_transpose(u);
    write(dir + "/yt.f64", yt.data(), yt.size());
    // (A x)[i] for the first rows - 5 entries of a gradient-length vector
    AlignedVector<double> glen(static_cast<size_t>(rows - 5), 0.0);
    write(dir + "/dd.f64", SparseOps::compute_directional_derivative(A, x, glen));
    std::printf("dot %.17g\n", SparseOps::dot_product(x, x));

    // square SPD system: the 7-point FD matrix shifted to be diagonally dominant
    const std::vector<double> vel = read<double>(dir + "/fdvel.f64");
    const int64_t n = (int64_t)param(dir, "fdn");
    CSRMatrix F;
    F.build_fd_stencil_7point(n, n, n, vel.data(), 10.0, 10.0, 10.0, 1e-3);
    write(dir + "/fd_rp.i64", F.row_ptr().data(), F.row_ptr().size());
    write(dir + "/fd_ci.i64", F.col_ind().data(), F.col_ind().size());
    write(dir + "/fd_val.f64", F.values().data(), F.values().size());
    F.scale(-1.0);
    CSRMatrix H = SparseOps::build_hessian_approx(F, 6.0);   // 6 I - (FD matrix): SPD
    AlignedVector<double> b(static_cast<size_t>(H.rows()), 1.0);
    SolveInfo i1, i2;
    const AlignedVector<double> x1 = SparseOps::conjugate_gradient(H, b, 1000, 1e-10, &i1);
    const AlignedVector<double> x2 = SparseOps::jacobi_preconditioned_cg(H, b, 1000, 1e-10, &i2);
    write(dir + "/cg.f64", x1.data(), x1.size());
    write(dir + "/pcg.f64", x2.data(), x2.size());
    write(dir + "/h_val.f64", H.values().data(), H.values().size());
    std::printf("cg iters %d converged %d pcg iters %d converged %d\n", i1.iterations, i1.converged ? 1 : 0,
                i2.iterations, i2.converged ? 1 : 0);

    // CGLS on the rectangular triplet matrix
    SolveInfo i3;
    const AlignedVector<double> x3 = SparseOps::cgls(A, u, 1e-2, 1e-8, 5000, CglsPreconditioner::JACOBI, &i3);
    write(dir + "/cgls.f64", x3.data(), x3.size());
    std::printf("cgls iters %d converged %d true_rel %.3e\n", i3.iterations, i3.converged ? 1 : 0,
                i3.true_rel_residual);
    return 0;
}

// SparseOps::smooth_gradient and gaussian_filter_3d on the grid of
// params.txt (no CSR matrix involved)
int cmd_smooth(const std::string& dir) {
    const int64_t nx = (int64_t)param(dir, "nx"), ny = (int64_t)param(dir, "ny"), nz = (int64_t)param(dir, "nz");
    const double sigma = param(dir, "sigma");
    const std::vector<double> g = read<double>(dir + "/g.f64");
    write(dir + "/smooth.f64", SparseOps::smooth_gradient(g, nx, ny, nz, sigma));
    write(dir + "/gauss.f64", SparseOps::gaussian_filter_3d(nx, ny, nz, sigma));
    const std::vector<double> shortg(g.begin(), g.begin() + (std::ptrdiff_t)(g.size() / 2));
    const std::vector<double> z = SparseOps::smooth_gradient(shortg, nx, ny, nz, sigma);
    double zmax = 0.0;
    for (double v : z) zmax = std::max(zmax, std::fabs(v));
    std::printf("short_gradient size %zu max %.3g\n", z.size(), zmax);
    return 0;
}

int cmd_rtm(const std::string& dir) {
    const int64_t nx = (int64_t)param(dir, "nx"), ny = (int64_t)param(dir, "ny"), nz = (int64_t)param(dir, "nz");
    const int hw = (int)param(dir, "hw");
    const std::vector<double> f = read<double>(dir + "/f.f64"), a = read<double>(dir + "/a.f64");
    RTMEngine eng(nx, ny, nz, 10.0, 10.0, 10.0, hw);
    eng.apply_image_condition(f.data(), a.data(), ny + 2 * hw, nz + 2 * hw, 0);
    eng.apply_image_condition(a.data(), f.data(), ny + 2 * hw, nz + 2 * hw, 1);
    eng.finalize_image();
    const RTMImageCondition& res = eng.result();
    write(dir + "/image.f64", res.image);
    write(dir + "/illum.f64", res.illumination);
    write(dir + "/refl.f64", res.reflectivity);
    // slice capture: X at 2, Y and Z at the middle, VOLUME skipped; ic = the image
    eng.set_slice_specs({{SliceAxis::X, 2, 1}, {SliceAxis::Y, -1, 1}, {SliceAxis::VOLUME, 0, 1},
                         {SliceAxis::Z, -1, 1}});
    std::vector<double> slices;
    eng.set_slice_callback([&](const RTMEngine::CapturedSlice& sl) {
        slices.push_back((double)sl.index);
        for (const auto* v : {&sl.forward_slice, &sl.adjoint_slice, &sl.ic_slice})
            slices.insert(slices.end(), v->begin(), v->end());
    });
    eng.extract_slices(f.data(), a.data(), nullptr, ny + 2 * hw, nz + 2 * hw, 7);
    write(dir + "/slices.f64", slices);
    std::printf("slices_captured=%lld\n", (long long)eng.total_captured_slices());

    // run_rtm through the solver: source wavelet and receiver residuals
    GridParams gp{nx, ny, nz, 10.0, 10.0, 10.0, 0.0, 0.0, 0.0, hw};
    CartesianGrid grid(gp);
    MpiContext mpi;
    SolverConfig cfg = load_config(dir + "/none.yaml").solver;
    cfg.dt = 1e-3;
    cfg.nt = (int)param(dir, "nt");
    cfg.sponge_width = 6;
    cfg.fd_order = 8;
    const std::vector<double> vel = read<double>(dir + "/vel.f64");
    const std::vector<double> w = read<double>(dir + "/wavelet.f64");
    const std::vector<int64_t> rcv = read<int64_t>(dir + "/rcv.i64");
    const std::vector<double> resid = read<double>(dir + "/resid.f64");
    WaveEquationSolver s(mpi, grid, cfg);
    s.set_velocity_field(vel.data(), (int64_t)vel.size());
    s.set_receivers(rcv);
    print_result("rtm_disabled", s.run_rtm(w, resid));
    RTMConfig rc;
    rc.enable_rtm = true;
    s.configure_rtm(rc);
    print_result("rtm", s.run_rtm(w, resid));
    write(dir + "/rtm_image.f64", s.rtm_engine()->result().image);
    write(dir + "/rtm_illum.f64", s.rtm_engine()->result().illumination);
    write(dir + "/rtm_traces.f64", s.receiver_traces());

    // the reference's own input: no receivers, a with-halo residual field
    WaveEquationSolver s2(mpi, grid, cfg);
    s2.set_velocity_field(vel.data(), (int64_t)vel.size());
    s2.configure_rtm(rc);
    const std::vector<double> rfield = read<double>(dir + "/resid_field.f64");
    print_result("rtm_field", s2.run_rtm(w, rfield));
    write(dir + "/rtmf_image.f64", s2.rtm_engine()->result().image);
    write(dir + "/rtmf_illum.f64", s2.rtm_engine()->result().illumination);
    write(dir + "/rtmf_refl.f64", s2.rtm_engine()->result().reflectivity);
    // neither form: not completed, and no exception
    print_result("rtm_short", s2.run_rtm(w, std::vector<double>(10, 0.0)));
    return 0;
}

// slice frames and the snapshot file, host only: nslices synthetic slices of
// growing size written to snap.dat (max_bytes small enough to restart the
// ring); every serialized frame also goes to frames.bin as [i64 n][frame]
int cmd_mmap(const std::string& dir) {
    const int n = (int)param(dir, "nslices");
    const int64_t max_bytes = (int64_t)param(dir, "max_bytes");
    MmapSnapshotWriter w;
    if (!w.open(dir + "/snap.dat", max_bytes)) return 3;
    std::vector<uint8_t> all, frame;
    int ok = 0, roundtrip = 0;
    for (int k = 0; k < n; ++k) {
        RTMEngine::CapturedSlice s;
        s.timestep = 10 * k;
        s.axis = static_cast<SliceAxis>(k % 3);
        s.index = k;
        s.dim0 = 3 + k;
        s.dim1 = 4 + (k % 2);
        for (int64_t i = 0; i < s.dim0 * s.dim1; ++i) {
            s.forward_slice.push_back(0.5 * k + i);
            s.adjoint_slice.push_back(-1.25 * i + k);
            s.ic_slice.push_back(1.0 / (1 + i + k));
        }
        s.capture_timestamp = 1000.0 + k;
        serialize_slice(s, (uint64_t)k, frame);
        const int64_t fs = (int64_t)frame.size();
        all.insert(all.end(), reinterpret_cast<const uint8_t*>(&fs), reinterpret_cast<const uint8_t*>(&fs) + 8);
        all.insert(all.end(), frame.begin(), frame.end());
        ok += w.write_slice(s) ? 1 : 0;
        RTMEngine::CapturedSlice back;
        uint64_t seq = 0;
        if (deserialize_slice(frame.data(), frame.size(), back, &seq) && seq == (uint64_t)k &&
            back.forward_slice == s.forward_slice && back.adjoint_slice == s.adjoint_slice &&
            back.ic_slice == s.ic_slice && back.dim0 == s.dim0 && back.axis == s.axis)
            ++roundtrip;
    }
    frame[200] ^= 1;
    RTMEngine::CapturedSlice bad;
    const bool corrupt_rejected = !deserialize_slice(frame.data(), frame.size(), bad, nullptr);
    const std::vector<uint8_t> huge((size_t)max_bytes, 0);
    const bool huge_rejected = !w.write_frame(huge.data(), max_bytes);
    const int64_t count = w.frame_count(), ring = w.header()->ring_start;
    const MmapSnapshotWriter::Stats st = w.stats();
    w.close();
    std::vector<uint8_t> first;
    const bool read0 = MmapSnapshotWriter::read_frame(dir + "/snap.dat", 0, first);
    write(dir + "/frames.bin", all);
    std::printf("written=%d roundtrip=%d corrupt_rejected=%d huge_rejected=%d frame_count=%lld ring_start=%lld "
                "count_frames=%lld stats_written=%lld stats_dropped=%lld read0=%d read0_size=%zu\n",
                ok, roundtrip, corrupt_rejected ? 1 : 0, huge_rejected ? 1 : 0, (long long)count, (long long)ring,
                (long long)MmapSnapshotWriter::count_frames(dir + "/snap.dat"), (long long)st.frames_written,
                (long long)st.frames_dropped_buffer_full, read0 ? 1 : 0, first.size());
    return 0;
}

// CSRMatrix::set_cell_grid argument checks, host only (no device copy is made)
int cmd_csrgrid() {
    CSRMatrix A;
    AlignedVector<int64_t> rp(3), ci(2);
    AlignedVector<double> v(2);
    rp[0] = 0; rp[1] = 1; rp[2] = 2; ci[0] = 0; ci[1] = 23; v[0] = 1.0; v[1] = 2.0;
    A.assign(2, 24, std::move(rp), std::move(ci), std::move(v));
    int bad = 0;
    try { A.set_cell_grid(2, 3, 5); } catch (const std::invalid_argument&) { ++bad; }
    try { A.set_cell_grid(0, 4, 6); } catch (const std::invalid_argument&) { ++bad; }
    A.set_cell_grid(2, 3, 4);
    A.set_cell_grid(0, 0, 0);
    A.set_cell_grid(4, 3, 2);
    CSRMatrix B = A.clone();
    std::printf("rejected=%d accepted=1 clone_ok=%d\n", bad, B.cols() == 24 ? 1 : 0);
    return 0;
}

int cmd_health(const std::string& dir) {
    std::vector<double> d = read<double>(dir + "/field.f64");
    const int64_t nx = (int64_t)param(dir, "nx"), ny = (int64_t)param(dir, "ny"), nz = (int64_t)param(dir, "nz");
    std::printf("nonfinite %lld\n", (long long)FaultTolerance::count_nan_inf(d.data(), (int64_t)d.size()));
    std::vector<double> c(d);
    FaultTolerance::clamp_field(c.data(), (int64_t)c.size(), -1.0, 1.0);
    write(dir + "/clamped.f64", c);
    FaultTolerance ft;
    std::vector<double> rep(d);
    ft.repair_field_nan(rep.data(), (int64_t)rep.size(), nx, ny, nz, 1, -7.0);
    write(dir + "/repaired.f64", rep);
    std::vector<double> cur(d), prev(d);
    const bool ok = ft.attempt_recovery(cur.data(), prev.data(), (int64_t)cur.size(), nx, ny, nz, 1, 3);
    std::printf("recovered %d cells_repaired %lld log %zu\n", ok ? 1 : 0, (long long)ft.stats().cells_repaired,
                ft.fault_log().size());
    write(dir + "/recovered.f64", cur);
    return 0;
}

// TerrainBoundary on the grid of params.txt (synthetic bathymetry plus the
// points of pts.f64), then the velocity mask, the boundary mask and the three
// boundary conditions on copies of field.f64
int cmd_terrain(const std::string& dir) {
    const int64_t nx = (int64_t)param(dir, "nx"), ny = (int64_t)param(dir, "ny"), nz = (int64_t)param(dir, "nz");
    const int hw = (int)param(dir, "hw");
    GridParams gp{nx, ny, nz, param(dir, "dx"), param(dir, "dy"), param(dir, "dz"), 0.0, 0.0, 0.0, hw};
    CartesianGrid grid(gp);
    TerrainBoundary t(grid, param(dir, "water"));
    t.generate_synthetic_bathymetry(param(dir, "gradient"), param(dir, "roughness"));
    const std::vector<double> pts = read<double>(dir + "/pts.f64");
    std::vector<BathymetryPoint> bp;
    for (size_t k = 0; k + 2 < pts.size(); k += 3) bp.push_back({pts[k], pts[k + 1], pts[k + 2]});
    t.set_bathymetry_data(bp);
    write(dir + "/bathy.f64", t.bathymetry());
    const std::vector<double> field = read<double>(dir + "/field.f64");
    std::vector<double> v(field);
    t.update_velocity_mask(v, 1500.0, 1800.0);
    write(dir + "/vel.f64", v);
    t.build_boundary_mask(grid.local(), hw);
    write(dir + "/mask.u8", t.mask());
    std::vector<double> f(field);
    t.apply_free_surface_bc(f, grid.local(), hw);
    write(dir + "/free.f64", f);
    f = field;
    t.apply_seafloor_bc(f, grid.local(), hw);
    write(dir + "/sea.f64", f);
    f = field;
    t.apply_absorbing_bc(f, grid.local(), hw, 5, 0.3);
    write(dir + "/abs.f64", f);
    std::printf("seafloor_index %lld water %d\n", (long long)t.seafloor_index(3, 4), t.is_water(3, 4, 0) ? 1 : 0);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: dropin_driver config|grid|fd|sparse|rtm|health|terrain [arg]\n");
        return 2;
    }
    const std::string cmd = argv[1], arg = argc > 2 ? argv[2] : "";
    try {
        if (cmd == "config") return cmd_config(arg);
        if (cmd == "grid") return cmd_grid();
        if (cmd == "fd") return cmd_fd(arg);
        if (cmd == "sparse") return cmd_sparse(arg);
        if (cmd == "rtm") return cmd_rtm(arg);
        if (cmd == "smooth") return cmd_smooth(arg);
        if (cmd == "health") return cmd_health(arg);
        if (cmd == "terrain") return cmd_terrain(arg);
        if (cmd == "mmap") return cmd_mmap(arg);
        if (cmd == "csrgrid") return cmd_csrgrid();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    std::fprintf(stderr, "unknown command %s\n", cmd.c_str());
    return 2;
}
