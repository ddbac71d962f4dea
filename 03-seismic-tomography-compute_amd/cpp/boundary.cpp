// TerrainBoundary (reference interface: src/cpp/solver/boundary.h:15-52) over
// the C ABI: field operations on the GPU (stp_terrain_*), bathymetry on the
// host with a device mirror.
#include "solver/boundary.h"

#include <stdexcept>

#include "device.h"

namespace seismic {

namespace {

void ext_of(const CartesianGrid& g, const LocalGrid& loc, int hw, int64_t* ext) {
    ext[0] = loc.nx_with_halo;
    ext[1] = loc.ny_with_halo;
    ext[2] = loc.nz_with_halo;
    ext[3] = hw;
    ext[4] = loc.offset_x;
    ext[5] = loc.offset_y;
    ext[6] = loc.offset_z;
    ext[7] = g.params().nx_global;
    ext[8] = g.params().ny_global;
}

}  // namespace

TerrainBoundary::TerrainBoundary(const CartesianGrid& grid, double water_depth)
    : grid_(grid), water_depth_(water_depth) {
    bathymetry_grid_.assign(static_cast<size_t>(grid.params().nx_global * grid.params().ny_global), water_depth);
}

TerrainBoundary::~TerrainBoundary() { dev::release(d_bathy_); }

const double* TerrainBoundary::device_bathymetry() const {
    const int64_t n = static_cast<int64_t>(bathymetry_grid_.size());
    if (!d_bathy_) d_bathy_ = dev::alloc<double>(n);
    if (dirty_) {
        dev::upload(d_bathy_, bathymetry_grid_.data(), n);
        dirty_ = false;
    }
    return d_bathy_;
}

void TerrainBoundary::set_bathymetry_data(const std::vector<BathymetryPoint>& data) {
    const GridParams& p = grid_.params();
    for (const auto& pt : data) {
        const int64_t ix = static_cast<int64_t>((pt.x - p.origin_x) / p.dx);
        const int64_t iy = static_cast<int64_t>((pt.y - p.origin_y) / p.dy);
        if (ix >= 0 && ix < p.nx_global && iy >= 0 && iy < p.ny_global)
            bathymetry_grid_[static_cast<size_t>(ix * p.ny_global + iy)] = pt.depth;
    }
    dirty_ = true;
}

void TerrainBoundary::generate_synthetic_bathymetry(double gradient, double roughness) {
    const GridParams& p = grid_.params();
    const int64_t n = static_cast<int64_t>(bathymetry_grid_.size());
    if (!d_bathy_) d_bathy_ = dev::alloc<double>(n);
    dev::check(stp_terrain_bathymetry_synthetic_f64(p.nx_global, p.ny_global, p.dx, p.dy, water_depth_, gradient,
                                                    roughness, d_bathy_, nullptr),
               "TerrainBoundary::generate_synthetic_bathymetry");
    dev::download(bathymetry_grid_.data(), d_bathy_, n);
    dirty_ = false;
}

double TerrainBoundary::seafloor_depth(int64_t gx, int64_t gy) const {
    const GridParams& p = grid_.params();
    if (gx < 0 || gx >= p.nx_global || gy < 0 || gy >= p.ny_global) return water_depth_;
    return bathymetry_grid_[static_cast<size_t>(gx * p.ny_global + gy)];
}

int64_t TerrainBoundary::seafloor_index(int64_t gx, int64_t gy) const {
    return static_cast<int64_t>(seafloor_depth(gx, gy) / grid_.params().dz);
}

bool TerrainBoundary::is_subsurface(int64_t gx, int64_t gy, int64_t gz) const {
    return grid_.params().origin_z + gz * grid_.params().dz >= seafloor_depth(gx, gy);
}

bool TerrainBoundary::is_water(int64_t gx, int64_t gy, int64_t gz) const { return !is_subsurface(gx, gy, gz); }

void TerrainBoundary::apply_free_surface_bc(std::vector<double>& field, const LocalGrid& loc, int halo_width) {
    const int64_t n = loc.nx_with_halo * loc.ny_with_halo * loc.nz_with_halo;
    if (static_cast<int64_t>(field.size()) < n) throw std::invalid_argument("apply_free_surface_bc: field too short");
    dev::Buffer<double> f(n);
    dev::upload(f.get(), field.data(), n);
    dev::check(stp_terrain_free_surface_f64(loc.nx_with_halo, loc.ny_with_halo, loc.nz_with_halo, halo_width, f.get(),
                                            nullptr),
               "TerrainBoundary::apply_free_surface_bc");
    dev::download(field.data(), f.get(), n);
}

void TerrainBoundary::apply_seafloor_bc(std::vector<double>& field, const LocalGrid& loc, int halo_width) {
    const int64_t n = static_cast<int64_t>(field.size());
    if (n == 0) return;
    int64_t ext[9];
    ext_of(grid_, loc, halo_width, ext);
    const double geo[3] = {water_depth_, grid_.params().dz, grid_.params().origin_z};
    dev::Buffer<double> f(n);
    dev::upload(f.get(), field.data(), n);
    dev::check(stp_terrain_seafloor_f64(ext, geo, device_bathymetry(), f.get(), n, nullptr),
               "TerrainBoundary::apply_seafloor_bc");
    dev::download(field.data(), f.get(), n);
}

void TerrainBoundary::apply_absorbing_bc(std::vector<double>& field, const LocalGrid& loc, int halo_width,
                                         int sponge_width, double damping) {
    const int64_t n = loc.nx_with_halo * loc.ny_with_halo * loc.nz_with_halo;
    if (static_cast<int64_t>(field.size()) < n) throw std::invalid_argument("apply_absorbing_bc: field too short");
    dev::Buffer<double> f(n);
    dev::upload(f.get(), field.data(), n);
    dev::check(stp_terrain_absorbing_f64(loc.nx_with_halo, loc.ny_with_halo, loc.nz_with_halo, halo_width,
                                         sponge_width, damping, f.get(), nullptr),
               "TerrainBoundary::apply_absorbing_bc");
    dev::download(field.data(), f.get(), n);
}

void TerrainBoundary::update_velocity_mask(std::vector<double>& velocity, double water_vp, double sediment_vp) const {
    const LocalGrid& loc = grid_.local();
    const int64_t n = loc.nx_with_halo * loc.ny_with_halo * loc.nz_with_halo;
    if (static_cast<int64_t>(velocity.size()) < n) throw std::invalid_argument("update_velocity_mask: too short");
    int64_t ext[9];
    ext_of(grid_, loc, grid_.params().halo_width, ext);
    const double geo[3] = {water_depth_, grid_.params().dz, grid_.params().origin_z};
    dev::Buffer<double> v(n);
    dev::upload(v.get(), velocity.data(), n);
    dev::check(stp_terrain_velocity_mask_f64(ext, geo, device_bathymetry(), water_vp, sediment_vp, v.get(), nullptr),
               "TerrainBoundary::update_velocity_mask");
    dev::download(velocity.data(), v.get(), n);
}

void TerrainBoundary::build_boundary_mask(const LocalGrid& loc, int halo_width) {
    const int64_t n = loc.nx_with_halo * loc.ny_with_halo * loc.nz_with_halo;
    int64_t ext[9];
    ext_of(grid_, loc, halo_width, ext);
    const double geo[3] = {water_depth_, grid_.params().dz, grid_.params().origin_z};
    dev::Buffer<uint8_t> m(n);
    dev::check(stp_terrain_boundary_mask(ext, geo, device_bathymetry(), m.get(), nullptr),
               "TerrainBoundary::build_boundary_mask");
    boundary_mask_.assign(static_cast<size_t>(n), 0);
    dev::download(boundary_mask_.data(), m.get(), n);
}

}  // namespace seismic
