// FD plan (the opaque stp_fd_plan of the C ABI), shared by the C ABI files.
#pragma once
#include <stdint.h>

#include "fd_kernels.h"
#include "seismic_hip.h"

struct stp_fd_plan {
    stp_fd_config cfg;
    int R = 0, VW = 1, variant = 0, tz = 0, ty = 0, lx = 0, ntz = 0, nty = 0, nxc = 0;
    int alt_sweep = 1;   // alternate the x-march direction step by step (STP_FD_SWEEP=up: always up)
    int nxh = 0, nyh = 0, nzh = 0;
    int64_t ncell = 0;
    float cx[5] = {0}, cy[5] = {0}, cz[5] = {0};
    int sponge_mode = STP_SPONGE_REFERENCE;   // stp_fd_set_sponge_mode
    // receivers
    int nrcv = 0, nrcv_pts = 0;
    int* d_blk_ptr = nullptr;
    int64_t* d_rcv_cell = nullptr;
    int* d_rcv_id = nullptr;
    // sources
    int64_t* d_src = nullptr;
    // scratch
    unsigned int* d_word = nullptr;
};

namespace stp {
void fd_fill_args(const stp_fd_plan* p, FdStepArgs* a);
}  // namespace stp
