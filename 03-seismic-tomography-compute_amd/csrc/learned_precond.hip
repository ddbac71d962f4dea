// Learned preconditioner network.  The network kernels are added with the
// learned-preconditioner component; until a network is loaded the entry point
// reports an error instead of running anything.
#include <hip/hip_runtime.h>

#include "capi_common.h"
#include "learned_precond.h"

namespace stp {
int lpc_apply(void* handle, const float* s, float* z, const int* done, hipStream_t st) {
    (void)handle; (void)s; (void)z; (void)done; (void)st;
    return fail(STP_ERR_STATE, "learned preconditioner: no network loaded");
}
}  // namespace stp
