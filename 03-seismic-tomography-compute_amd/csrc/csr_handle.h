// Device CSR handle shared by the tomography solve and the learned
// preconditioner (internal; the C ABI sees an opaque pointer).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "sparse_kernels.h"

struct stp_csr {
    int rows = 0, cols = 0;
    int64_t nnz = 0;
    int *rp = nullptr, *ci = nullptr, *trp = nullptr, *tci = nullptr;
    float *val = nullptr, *tval = nullptr;
    double* val64 = nullptr;   // optional FP64 values of A (square CG in FP64)
    int W = 64, WT = 64;   // lanes per row for A and A^T
    // workspace
    double *r = nullptr, *part0 = nullptr, *part1 = nullptr, *part2 = nullptr, *part3 = nullptr;
    double *part4 = nullptr, *part5 = nullptr;
    float *p = nullptr, *q = nullptr, *s = nullptr, *z = nullptr, *sp = nullptr, *dinv = nullptr, *r32 = nullptr;
    double *ry = nullptr, *rg = nullptr;   // CGLS: change of r (rows) and A^T r of the pass start (cols)
    double* pnorm = nullptr;               // CGLS: |r|^2 and |y|^2 partials (2 x 256)
    double *dp = nullptr, *dap = nullptr, *dz = nullptr, *dr = nullptr, *ddinv = nullptr;
    stp::SolverScalars* sc = nullptr;
    size_t c_r = 0, c_p0 = 0, c_p1 = 0, c_p2 = 0, c_p3 = 0, c_p4 = 0, c_p5 = 0, c_p = 0, c_q = 0, c_s = 0, c_z = 0, c_sp = 0, c_dinv = 0, c_r32 = 0, c_ry = 0, c_rg = 0, c_pnorm = 0;
    // solver layout (solver_layout.hip): columns are the cells of a
    // gn[0]*gn[1]*gn[2] grid when gn[0] > 0; the renumbered copy is built on
    // the first CGLS solve or by stp_csr_prepare_layout
    int64_t gn[3] = {0, 0, 0};
    bool lay_ready = false;
    int *cperm = nullptr, *rsrc = nullptr;   // cell -> layout column; layout row -> row
    int *lrp = nullptr, *lci = nullptr, *ltrp = nullptr, *ltci = nullptr;
    float *lval = nullptr, *ltval = nullptr;
    double *lx = nullptr, *lb = nullptr;
    int lW = 16, lWT = 8;   // lanes per row in the layout
    // staged copy of the layout A^T: its rows cut into tiles of st_rt rows;
    // per tile the sorted distinct columns its entries read
    // (st_uniq[st_tu[t] .. st_tu[t + 1])) and per entry the position of its
    // column in that list (st_l16); st_rt = 0 when no tile size fits the
    // staging buffer
    int st_rt = 0, st_ntiles = 0, st_maxu = 0;
    int64_t st_nlist = 0;   // total length of the tiles' lists
    int *st_tu = nullptr, *st_uniq = nullptr;
    uint16_t* st_l16 = nullptr;
    size_t c_dp = 0, c_dap = 0, c_dz = 0, c_dr = 0, c_ddinv = 0, c_sc = 0;
};
