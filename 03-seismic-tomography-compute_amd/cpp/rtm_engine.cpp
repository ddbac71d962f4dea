// RTMEngine (reference interface: src/cpp/rtm/rtm_engine.h:31-96): FP64
// image and illumination on the GPU.
#include "rtm/rtm_engine.h"

#include <chrono>
#include <cmath>
#include <stdexcept>

#include "device.h"

namespace seismic {

RTMEngine::RTMEngine(int64_t nx, int64_t ny, int64_t nz, double dx, double dy, double dz, int halo_width)
    : nx_(nx), ny_(ny), nz_(nz), dx_(dx), dy_(dy), dz_(dz), halo_width_(halo_width) {
    if (nx < 0 || ny < 0 || nz < 0 || halo_width < 0) throw std::invalid_argument("RTMEngine: bad shape");
    const int64_t n = total_cells();
    image_dev_ = dev::alloc<double>(n);
    illum_dev_ = dev::alloc<double>(n);
    refl_dev_ = dev::alloc<double>(n);
    reset_image();
}

RTMEngine::~RTMEngine() {
    dev::release(image_dev_);
    dev::release(illum_dev_);
    dev::release(refl_dev_);
    dev::release(fwd_stage_);
    dev::release(adj_stage_);
}

void RTMEngine::apply_image_condition(const double* fwd, const double* adj, int64_t ny, int64_t nz, int64_t step) {
    const int64_t nxh = nx_ + 2 * halo_width_;
    const int64_t cells = nxh * ny * nz;   // with-halo arrays of strides ny, nz
    if (cells > stage_cells_) {
        dev::release(fwd_stage_);
        dev::release(adj_stage_);
        fwd_stage_ = dev::alloc<double>(cells);
        adj_stage_ = dev::alloc<double>(cells);
        stage_cells_ = cells;
    }
    dev::upload(fwd_stage_, fwd, cells);
    dev::upload(adj_stage_, adj, cells);
    apply_image_condition_device(fwd_stage_, adj_stage_, ny, nz, step);
}

void RTMEngine::apply_image_condition_device(const double* fwd, const double* adj, int64_t ny, int64_t nz,
                                             int64_t) {
    dev::check(stp_rtm_image_condition_f64(nx_, ny_, nz_, halo_width_, ny, nz, fwd, adj, image_dev_, illum_dev_,
                                           nullptr),
               "RTMEngine::apply_image_condition");
    host_valid_ = false;
}

void RTMEngine::accumulate_device_f32(const float* image, const float* illum) {
    if (image) dev::check(stp_accumulate_f32_f64(total_cells(), image, image_dev_, nullptr), "RTMEngine accumulate");
    if (illum) dev::check(stp_accumulate_f32_f64(total_cells(), illum, illum_dev_, nullptr), "RTMEngine accumulate");
    host_valid_ = false;
}

void RTMEngine::extract_slices(const double* fwd, const double* adj, const double* image_condition, int64_t ny,
                               int64_t nz, int64_t timestep) {
    if (!slice_cb_ || slice_specs_.empty()) return;
    const double* ic = image_condition ? image_condition : result().image.data();
    const int64_t hw = halo_width_, nynz = ny * nz;
    auto fin = [](double v) { return std::isfinite(v) ? v : 0.0; };
    for (const SliceSpec& spec : slice_specs_) {
        if (spec.axis == SliceAxis::VOLUME) continue;
        const int a = static_cast<int>(spec.axis);
        const int64_t idx = spec.index >= 0 ? spec.index : (a == 0 ? nx_ : (a == 1 ? ny_ : nz_)) / 2;
        CapturedSlice sl;
        sl.timestep = timestep;
        sl.axis = spec.axis;
        sl.index = idx;
        sl.dim0 = a == 0 ? ny_ : nx_;
        sl.dim1 = a == 2 ? ny_ : nz_;
        const int64_t n = sl.dim0 * sl.dim1;
        sl.forward_slice.resize(static_cast<size_t>(n));
        sl.adjoint_slice.resize(static_cast<size_t>(n));
        sl.ic_slice.resize(static_cast<size_t>(n));
        for (int64_t p = 0; p < sl.dim0; ++p)
            for (int64_t q = 0; q < sl.dim1; ++q) {
                int64_t src, ici;
                if (a == 0) { src = (idx + hw) * nynz + (p + hw) * nz + (q + hw); ici = p * nz_ + q; }
                else if (a == 1) { src = (p + hw) * nynz + (idx + hw) * nz + (q + hw); ici = p * nz_ + q; }
                else { src = (p + hw) * nynz + (q + hw) * nz + (idx + hw); ici = p * ny_ + q; }
                const size_t d = static_cast<size_t>(p * sl.dim1 + q);
                sl.forward_slice[d] = fin(fwd[src]);
                sl.adjoint_slice[d] = fin(adj[src]);
                sl.ic_slice[d] = fin(ic[ici]);
            }
        sl.capture_timestamp =
            std::chrono::duration<double>(std::chrono::high_resolution_clock::now().time_since_epoch()).count();
        slice_cb_(sl);
        slices_captured_.fetch_add(1, std::memory_order_release);
    }
}

void RTMEngine::finalize_image() {
    dev::check(stp_rtm_finalize_f64(total_cells(), image_dev_, illum_dev_, refl_dev_, nullptr),
               "RTMEngine::finalize_image");
    host_valid_ = false;
}

const RTMImageCondition& RTMEngine::result() const {
    if (!host_valid_) {
        const int64_t n = total_cells();
        host_.image.resize(static_cast<size_t>(n));
        host_.illumination.resize(static_cast<size_t>(n));
        host_.reflectivity.resize(static_cast<size_t>(n));
        dev::download(host_.image.data(), image_dev_, n);
        dev::download(host_.illumination.data(), illum_dev_, n);
        dev::download(host_.reflectivity.data(), refl_dev_, n);
        host_valid_ = true;
    }
    return host_;
}

void RTMEngine::reset_image() {
    const int64_t bytes = static_cast<int64_t>(sizeof(double)) * total_cells();
    dev::check(stp_memset_d(image_dev_, 0, bytes, nullptr), "RTMEngine::reset_image");
    dev::check(stp_memset_d(illum_dev_, 0, bytes, nullptr), "RTMEngine::reset_image");
    dev::check(stp_memset_d(refl_dev_, 0, bytes, nullptr), "RTMEngine::reset_image");
    dev::check(stp_stream_sync(nullptr), "RTMEngine::reset_image");
    host_valid_ = false;
}

}  // namespace seismic
