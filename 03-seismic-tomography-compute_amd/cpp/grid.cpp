// CartesianGrid (reference interface: src/cpp/core/grid.h:36-68).
#include "core/grid.h"

#include <algorithm>

namespace seismic {

CartesianGrid::CartesianGrid(const GridParams& params) : params_(params) {
    const int coords[3] = {0, 0, 0}, dims[3] = {1, 1, 1};
    decompose(coords, dims);
}

CartesianGrid::CartesianGrid(const GridParams& params, const int mpi_coords[3], const int mpi_dims[3])
    : params_(params) {
    decompose(mpi_coords, mpi_dims);
}

// Block decomposition per axis: the first (n mod p) ranks get one extra
// cell (grid.cpp:13-39).
void CartesianGrid::decompose(const int coords[3], const int dims[3]) {
    const int64_t n[3] = {params_.nx_global, params_.ny_global, params_.nz_global};
    int64_t len[3], off[3];
    for (int a = 0; a < 3; ++a) {
        const int64_t p = std::max(1, dims[a]);
        const int64_t c = coords[a];
        const int64_t q = n[a] / p, r = n[a] % p;
        len[a] = q + (c < r ? 1 : 0);
        off[a] = c * q + std::min(c, r);
    }
    const int64_t hw = params_.halo_width;
    local_.nx_local = len[0];
    local_.ny_local = len[1];
    local_.nz_local = len[2];
    local_.offset_x = off[0];
    local_.offset_y = off[1];
    local_.offset_z = off[2];
    local_.nx_with_halo = len[0] + 2 * hw;
    local_.ny_with_halo = len[1] + 2 * hw;
    local_.nz_with_halo = len[2] + 2 * hw;
    local_.total_cells = len[0] * len[1] * len[2];
    local_.total_cells_with_halo = local_.nx_with_halo * local_.ny_with_halo * local_.nz_with_halo;
}

int64_t CartesianGrid::local_index(int64_t ix, int64_t iy, int64_t iz) const {
    return (ix * local_.ny_local + iy) * local_.nz_local + iz;
}

int64_t CartesianGrid::local_index_with_halo(int64_t ix, int64_t iy, int64_t iz) const {
    return (ix * local_.ny_with_halo + iy) * local_.nz_with_halo + iz;
}

void CartesianGrid::global_to_local(int64_t gx, int64_t gy, int64_t gz, int64_t& lx, int64_t& ly,
                                    int64_t& lz) const {
    lx = gx - local_.offset_x;
    ly = gy - local_.offset_y;
    lz = gz - local_.offset_z;
}

void CartesianGrid::local_to_global(int64_t lx, int64_t ly, int64_t lz, int64_t& gx, int64_t& gy,
                                    int64_t& gz) const {
    gx = lx + local_.offset_x;
    gy = ly + local_.offset_y;
    gz = lz + local_.offset_z;
}

bool CartesianGrid::is_interior(int64_t lx, int64_t ly, int64_t lz) const {
    return lx >= 0 && ly >= 0 && lz >= 0 && lx < local_.nx_local && ly < local_.ny_local && lz < local_.nz_local;
}

bool CartesianGrid::is_halo(int64_t lx, int64_t ly, int64_t lz) const {
    const bool inside = lx >= 0 && ly >= 0 && lz >= 0 && lx < local_.nx_with_halo && ly < local_.ny_with_halo &&
                        lz < local_.nz_with_halo;
    const int64_t hw = params_.halo_width;
    return inside && !is_interior(lx - hw, ly - hw, lz - hw);
}

std::vector<double> CartesianGrid::compute_coordinates() const {
    std::vector<double> xyz;
    xyz.reserve(static_cast<size_t>(3 * local_.total_cells));
    for (int64_t ix = 0; ix < local_.nx_local; ++ix)
        for (int64_t iy = 0; iy < local_.ny_local; ++iy)
            for (int64_t iz = 0; iz < local_.nz_local; ++iz) {
                xyz.push_back(x_coord(ix));
                xyz.push_back(y_coord(iy));
                xyz.push_back(z_coord(iz));
            }
    return xyz;
}

double CartesianGrid::x_coord(int64_t ix) const { return params_.origin_x + params_.dx * (local_.offset_x + ix); }
double CartesianGrid::y_coord(int64_t iy) const { return params_.origin_y + params_.dy * (local_.offset_y + iy); }
double CartesianGrid::z_coord(int64_t iz) const { return params_.origin_z + params_.dz * (local_.offset_z + iz); }

}  // namespace seismic
