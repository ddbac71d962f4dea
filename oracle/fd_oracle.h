/* Shared declarations of the FD oracle (test infrastructure only). */
#ifndef ORC_FD_ORACLE_H
#define ORC_FD_ORACLE_H
#include <stdint.h>

typedef struct {
    int64_t nx, ny, nz;      /* interior cells per axis */
    int32_t hw;              /* halo width */
    int32_t order;           /* 2 or 8 */
    double dx, dy, dz, dt;
    int32_t sponge_width;
    double sponge_damping;
    double amplitude_limit;  /* reference default 1e10 */
    double cfl_safety_factor;/* reference default 0.49 */
    int32_t sponge_mode;     /* 0: reference (damps current and previous before the swap);
                                1: Cerjan (damps next and current: the stored fields both decay) */
} orc_fd_cfg;

typedef struct {
    int64_t total_nan_count;
    int64_t cells_repaired;
    int64_t timesteps_recovered;
} orc_fault_stats;

int orc_fd_weights(int order, double* out);
double orc_fd_sponge_factor(const orc_fd_cfg* c, int64_t ix, int64_t iy, int64_t iz);
#endif
