// 3-D launch over an nx x ny x nz box of cells stored z fastest: 256 threads
// along z, one block per chunk of `rows` consecutive y rows (grid y) of one x
// plane (grid z).  Workgroup b runs on XCD b % 8, so a block that walks its
// own rows keeps its y neighbours in its XCD's L2 (one block per row would
// put rows iy - 1 and iy + 1 on other XCDs and fetch them again).  Kernels
// loop y chunks and x planes with the grid's stride, so every box size is
// covered, and no index is recovered by division.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace stp {

constexpr int ROWS = 16;   // y rows per block

inline dim3 grid3(int64_t nx, int64_t ny, int64_t nz, int rows = ROWS) {
    const int64_t bz = (nz + 255) / 256, by = (ny + rows - 1) / rows;
    return dim3((unsigned)(bz < 1 ? 1 : bz), (unsigned)(by < 1 ? 1 : (by > 65535 ? 65535 : by)),
                (unsigned)(nx < 1 ? 1 : (nx > 65535 ? 65535 : nx)));
}

}  // namespace stp

// for (ix, iy) over the x planes and y rows of this block (see above)
#define STP_FOR_PLANES_ROWS(ix, iy, nx, ny, rows)                                                  \
    for (int64_t ix = blockIdx.z; ix < (nx); ix += gridDim.z)                                      \
        for (int64_t y0_ = (int64_t)blockIdx.y * (rows); y0_ < (ny); y0_ += (int64_t)gridDim.y * (rows)) \
            for (int64_t iy = y0_, y1_ = (y0_ + (rows) < (ny) ? y0_ + (rows) : (ny)); iy < y1_; ++iy)
