// Single-process MpiContext (reference interface: src/cpp/core/mpi_context.h:20-57).
#include "core/mpi_context.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace seismic {

MpiContext::MpiContext() {
    std::memset(&cfg_, 0, sizeof cfg_);
    cfg_.size = 1;
    for (int a = 0; a < 3; ++a) cfg_.dims[a] = 1;
    for (int k = 0; k < 6; ++k) cfg_.neighbors[k] = PROC_NULL;
}

MpiContext::MpiContext(int&, char**&) : MpiContext() {}

void MpiContext::abort(int error_code) const { std::exit(error_code); }

static void only_proc_null(int peer, const char* what) {
    if (peer != PROC_NULL) {
        std::fprintf(stderr, "MpiContext::%s: rank %d does not exist in a single-process run\n", what, peer);
        std::abort();
    }
}

void MpiContext::send_double(const double*, int, int dest, int) const { only_proc_null(dest, "send_double"); }
void MpiContext::recv_double(double*, int, int src, int) const { only_proc_null(src, "recv_double"); }

void MpiContext::allgather_int(const int* sendbuf, int sendcount, int* recvbuf, int recvcount) const {
    const int n = sendcount < recvcount ? sendcount : recvcount;
    if (n > 0) std::memcpy(recvbuf, sendbuf, sizeof(int) * static_cast<size_t>(n));
}

double MpiContext::time() const {
    using clk = std::chrono::steady_clock;
    return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

void MpiContext::print_rank_info() const {
    std::printf("rank 0 of 1 (single process, no MPI), dims 1x1x1\n");
}

}  // namespace seismic
