// RTM imaging condition (drop-in for the reference's src/cpp/rtm/rtm_engine.h:12-96).
// Image and illumination stay on the GPU in FP64 and are copied to the host
// when result() is read.  extract_slices gathers the slice planes of the
// caller's host arrays on the host (a plane is too small to be worth a copy of
// the volume to the device) and passes them to the slice callback; device
// arrays go through stp_rtm_extract_slice_f64.
#pragma once

#include <atomic>
#include <cstdint>
#include <functional>
#include <vector>

namespace seismic {

enum class SliceAxis { X = 0, Y = 1, Z = 2, VOLUME = 3 };

struct SliceSpec {
    SliceAxis axis;
    int64_t index;
    int64_t stride;
};

struct RTMImageCondition {    // rtm_engine.h:25-29
    std::vector<double> image;
    std::vector<double> illumination;
    std::vector<double> reflectivity;
};

class RTMEngine {
public:
    RTMEngine(int64_t nx, int64_t ny, int64_t nz, double dx, double dy, double dz, int halo_width);
    ~RTMEngine();
    RTMEngine(const RTMEngine&) = delete;
    RTMEngine& operator=(const RTMEngine&) = delete;

    void set_slice_specs(const std::vector<SliceSpec>& specs) { slice_specs_ = specs; }

    // over the interior: f, a := value if finite else 0; image += f a;
    // illumination += f f (rtm_engine.cpp:24-57).  Host with-halo arrays of
    // y / z strides ny, nz (ny_with_halo, nz_with_halo).
    void apply_image_condition(const double* forward_wavefield, const double* adjoint_wavefield, int64_t ny,
                               int64_t nz, int64_t step_idx);
    // the same on device arrays
    void apply_image_condition_device(const double* forward_dev, const double* adjoint_dev, int64_t ny, int64_t nz,
                                      int64_t step_idx);
    // add device FP32 interior-indexed image / illumination contributions (the
    // fused reverse-time pass accumulates these)
    void accumulate_device_f32(const float* image_dev, const float* illum_dev);

    // rtm_engine.cpp:59-165: per slice spec (negative index = middle of the
    // axis, VOLUME skipped), the plane of the with-halo forward / adjoint host
    // arrays (y / z strides ny, nz) and of the image condition (the engine's
    // image when null; indexed by the in-plane coordinates, as the reference
    // does), non-finite values as 0, handed to the slice callback
    void extract_slices(const double* forward_wavefield, const double* adjoint_wavefield,
                        const double* image_condition, int64_t ny, int64_t nz, int64_t timestep);

    // reflectivity = image / illumination where illumination > 1e-10, else 0
    void finalize_image();
    const RTMImageCondition& result() const;

    int64_t nx() const { return nx_; }
    int64_t ny() const { return ny_; }
    int64_t nz() const { return nz_; }
    int64_t total_cells() const { return nx_ * ny_ * nz_; }

    void reset_image();

    struct CapturedSlice {
        int64_t timestep;
        SliceAxis axis;
        int64_t index;
        int64_t dim0;
        int64_t dim1;
        std::vector<double> forward_slice;
        std::vector<double> adjoint_slice;
        std::vector<double> ic_slice;
        double capture_timestamp;
    };
    using SliceCaptureCallback = std::function<void(const CapturedSlice&)>;
    void set_slice_callback(SliceCaptureCallback cb) { slice_cb_ = std::move(cb); }
    // New: a callback and slice specs are set (extract_slices has work to do)
    bool capturing() const { return static_cast<bool>(slice_cb_) && !slice_specs_.empty(); }
    int64_t total_captured_slices() const { return slices_captured_.load(std::memory_order_acquire); }

    // device arrays (interior-indexed FP64)
    const double* image_device() const { return image_dev_; }
    const double* illumination_device() const { return illum_dev_; }

private:
    int64_t nx_, ny_, nz_;
    double dx_, dy_, dz_;
    int halo_width_;
    double *image_dev_ = nullptr, *illum_dev_ = nullptr, *refl_dev_ = nullptr;
    double *fwd_stage_ = nullptr, *adj_stage_ = nullptr;
    int64_t stage_cells_ = 0;
    mutable RTMImageCondition host_;
    mutable bool host_valid_ = false;
    std::vector<SliceSpec> slice_specs_;
    SliceCaptureCallback slice_cb_;
    std::atomic<int64_t> slices_captured_{0};
};

}  // namespace seismic
