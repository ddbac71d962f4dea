// Test-only driver: extern "C" entry points over the reference's own
// translation units (compiled from /root/reference by oracle/build.py into
// oracle/_ref/libref.so).  Only functions whose reference code is safe to run
// are exposed:
//   * CSRMatrix::build_from_triplets / spmv / add_diagonal / scale /
//     frobenius_norm (src/cpp/sparse/csr_matrix.cpp:25-193), only for inputs
//     where every row keeps >= 8 entries after merging, because
//     pad_to_alignment (csr_matrix.cpp:223-264) writes past its buffers
//     otherwise; spmv_transpose is not exposed (it accumulates into
//     uninitialised memory, aligned_alloc.h:61-77);
//   * RTMEngine::apply_image_condition / finalize_image / reset_image
//     (src/cpp/rtm/rtm_engine.cpp:24-57, 170-190);
//   * load_config (src/cpp/core/config.cpp:9-96);
//   * TerrainBoundary (src/cpp/solver/boundary.cpp:8-204) on a single-block
//     CartesianGrid (src/cpp/core/grid.cpp:13-49).
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <vector>

#include "core/config.h"
#include "core/grid.h"
#include "solver/boundary.h"
#include "rtm/rtm_engine.h"
#include "sparse/csr_matrix.h"

using namespace seismic;

extern "C" {

// returns nullptr when some row would keep fewer than 8 entries
void* ref_csr_build(int64_t rows, int64_t cols, int64_t n, const int64_t* r, const int64_t* c, const double* v) {
    std::vector<std::set<int64_t>> uniq(rows > 0 ? rows : 0);
    for (int64_t i = 0; i < n; ++i)
        if (r[i] >= 0 && r[i] < rows && c[i] >= 0 && c[i] < cols) uniq[r[i]].insert(c[i]);
    for (const auto& s : uniq)
        if (s.size() < 8) return nullptr;
    std::vector<SparseEntry> e(n);
    for (int64_t i = 0; i < n; ++i) e[i] = {r[i], c[i], v[i]};
    auto* m = new CSRMatrix();
    m->build_from_triplets(rows, cols, e);
    return m;
}

int64_t ref_csr_nnz(void* h) { return static_cast<CSRMatrix*>(h)->nnz(); }

void ref_csr_arrays(void* h, int64_t* rp, int64_t* ci, double* val) {
    auto* m = static_cast<CSRMatrix*>(h);
    std::memcpy(rp, m->row_ptr().data(), sizeof(int64_t) * (m->rows() + 1));
    std::memcpy(ci, m->col_ind().data(), sizeof(int64_t) * m->nnz());
    std::memcpy(val, m->values().data(), sizeof(double) * m->nnz());
}

void ref_csr_spmv(void* h, const double* x, int64_t nx, double* y) {
    auto* m = static_cast<CSRMatrix*>(h);
    AlignedVector<double> xv(nx);
    for (int64_t i = 0; i < nx; ++i) xv[i] = x[i];
    AlignedVector<double> yv = m->spmv(xv);
    std::memcpy(y, yv.data(), sizeof(double) * yv.size());
}

void ref_csr_add_diagonal(void* h, double a) { static_cast<CSRMatrix*>(h)->add_diagonal(a); }
void ref_csr_scale(void* h, double a) { static_cast<CSRMatrix*>(h)->scale(a); }
double ref_csr_frobenius(void* h) { return static_cast<CSRMatrix*>(h)->frobenius_norm(); }
void ref_csr_free(void* h) { delete static_cast<CSRMatrix*>(h); }

void* ref_rtm_create(int64_t nx, int64_t ny, int64_t nz, int hw) {
    return new RTMEngine(nx, ny, nz, 10.0, 10.0, 10.0, hw);
}
void ref_rtm_apply(void* h, const double* f, const double* a, int64_t nyh, int64_t nzh, int64_t step) {
    static_cast<RTMEngine*>(h)->apply_image_condition(f, a, nyh, nzh, step);
}
void ref_rtm_finalize(void* h) { static_cast<RTMEngine*>(h)->finalize_image(); }
void ref_rtm_reset(void* h) { static_cast<RTMEngine*>(h)->reset_image(); }
void ref_rtm_result(void* h, double* img, double* ill, double* refl) {
    auto* e = static_cast<RTMEngine*>(h);
    const auto& r = e->result();
    std::memcpy(img, r.image.data(), sizeof(double) * r.image.size());
    std::memcpy(ill, r.illumination.data(), sizeof(double) * r.illumination.size());
    std::memcpy(refl, r.reflectivity.data(), sizeof(double) * r.reflectivity.size());
}
void ref_rtm_free(void* h) { delete static_cast<RTMEngine*>(h); }

// RTMEngine::extract_slices with the given specs (axis, index pairs); the
// captured slices are written one after another (forward, adjoint, ic per
// slice) into out, their dims into dims (2 per slice).  Returns the count.
int ref_rtm_extract(void* h, int nspec, const int* axes, const int64_t* idx, const double* fwd, const double* adj,
                    const double* ic, int64_t nyh, int64_t nzh, int64_t step, double* out, int64_t* dims) {
    auto* e = static_cast<RTMEngine*>(h);
    std::vector<SliceSpec> specs;
    for (int k = 0; k < nspec; ++k) specs.push_back({static_cast<SliceAxis>(axes[k]), idx[k], 1});
    e->set_slice_specs(specs);
    int count = 0;
    int64_t off = 0;
    e->set_slice_callback([&](const RTMEngine::CapturedSlice& s) {
        const int64_t n = s.dim0 * s.dim1;
        std::memcpy(out + off, s.forward_slice.data(), sizeof(double) * n);
        std::memcpy(out + off + n, s.adjoint_slice.data(), sizeof(double) * n);
        std::memcpy(out + off + 2 * n, s.ic_slice.data(), sizeof(double) * n);
        off += 3 * n;
        dims[2 * count] = s.dim0;
        dims[2 * count + 1] = s.dim1;
        ++count;
    });
    e->extract_slices(fwd, adj, ic, nyh, nzh, step);
    return count;
}

// load_config: 21 numeric fields in a fixed order, plus the two strings.
// Returns 0, or 1 when the reference threw (bad number).
int ref_load_config(const char* path, double* out, char* misfit_type, char* output_dir, int cap) {
    try {
        TomographyConfig c = load_config(path);
        const double v[] = {c.solver.frequency, c.solver.dt, (double)c.solver.nt, c.solver.water_depth,
                            c.solver.seafloor_gradient, (double)c.solver.sponge_width, c.solver.sponge_damping,
                            (double)c.solver.snapshot_interval, c.model.vp_min, c.model.vp_max, c.model.vs_min,
                            c.model.vs_max, c.model.rho_min, c.model.rho_max, c.model.initial_vp,
                            c.model.initial_vs, c.model.initial_rho, (double)c.inversion.max_iterations,
                            c.inversion.step_length, c.inversion.gradient_smoothing_sigma,
                            c.inversion.regularization_lambda, c.inversion.misfit_tolerance,
                            (double)c.inversion.gradient_precondition, (double)c.inversion.line_search_max,
                            c.inversion.line_search_factor};
        std::memcpy(out, v, sizeof v);
        std::strncpy(misfit_type, c.inversion.misfit_type.c_str(), cap - 1);
        misfit_type[cap - 1] = 0;
        std::strncpy(output_dir, c.solver.output_dir.c_str(), cap - 1);
        output_dir[cap - 1] = 0;
        return 0;
    } catch (...) {
        return 1;
    }
}

// TerrainBoundary on an nx x ny x nz grid (halo hw, spacing h, origin o):
// bathymetry (synthetic when gen != 0), then the velocity mask, the boundary
// mask and the three boundary conditions applied to copies of `field`.
void ref_terrain(int64_t nx, int64_t ny, int64_t nz, int hw, const double* h, const double* o, double water_depth,
                 int gen, double gradient, double roughness, int64_t npts, const double* pts, double water_vp,
                 double sediment_vp, int sponge_width, double damping, const double* field, double* bathy_out,
                 double* vel_out, uint8_t* mask_out, double* free_out, double* sea_out, double* abs_out) {
    GridParams p{nx, ny, nz, h[0], h[1], h[2], o[0], o[1], o[2], hw};
    const int coords[3] = {0, 0, 0}, dims[3] = {1, 1, 1};
    CartesianGrid g(p, coords, dims);
    TerrainBoundary t(g, water_depth);
    if (gen) t.generate_synthetic_bathymetry(gradient, roughness);
    std::vector<BathymetryPoint> bp;
    for (int64_t k = 0; k < npts; ++k) bp.push_back({pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]});
    if (npts) t.set_bathymetry_data(bp);
    std::memcpy(bathy_out, t.bathymetry().data(), sizeof(double) * nx * ny);
    const LocalGrid& loc = g.local();
    const int64_t n = loc.nx_with_halo * loc.ny_with_halo * loc.nz_with_halo;
    std::vector<double> v(field, field + n);
    t.update_velocity_mask(v, water_vp, sediment_vp);
    std::memcpy(vel_out, v.data(), sizeof(double) * n);
    t.build_boundary_mask(loc, hw);
    std::memcpy(mask_out, t.mask().data(), n);
    std::vector<double> f(field, field + n);
    t.apply_free_surface_bc(f, loc, hw);
    std::memcpy(free_out, f.data(), sizeof(double) * n);
    f.assign(field, field + n);
    t.apply_seafloor_bc(f, loc, hw);
    std::memcpy(sea_out, f.data(), sizeof(double) * n);
    f.assign(field, field + n);
    t.apply_absorbing_bc(f, loc, hw, sponge_width, damping);
    std::memcpy(abs_out, f.data(), sizeof(double) * n);
}

}  // extern "C"
