// Single-process stand-in for the reference's MpiContext
// (src/cpp/core/mpi_context.h:10-57).  One MI355X holds the whole grid, so
// there is no MPI: rank 0 of 1, every neighbour is PROC_NULL (-1), collectives
// return their argument, point-to-point calls to PROC_NULL do nothing.  The
// type exists so that WaveEquationSolver(const MpiContext&, ...) keeps its
// signature.
#pragma once

#include <cstdint>

namespace seismic {

constexpr int PROC_NULL = -1;

struct MpiConfig {            // mpi_context.h:10-18 (cart_comm is an int handle here)
    int rank;
    int size;
    int local_rank;
    int cart_comm;
    int dims[3];
    int coords[3];
    int neighbors[6];
};

class MpiContext {
public:
    MpiContext();
    MpiContext(int& argc, char**& argv);
    MpiContext(const MpiContext&) = delete;
    MpiContext& operator=(const MpiContext&) = delete;

    const MpiConfig& config() const { return cfg_; }

    void barrier() const {}
    [[noreturn]] void abort(int error_code) const;
    double allreduce_max(double v) const { return v; }
    double allreduce_sum(double v) const { return v; }
    int allreduce_max_int(int v) const { return v; }
    int64_t allreduce_sum_int64(int64_t v) const { return v; }
    // only PROC_NULL peers exist on one process: these are no-ops for it and
    // abort for any other rank
    void send_double(const double* buf, int count, int dest, int tag) const;
    void recv_double(double* buf, int count, int src, int tag) const;
    void allgather_int(const int* sendbuf, int sendcount, int* recvbuf, int recvcount) const;
    void bcast_double(double*, int, int) const {}
    void bcast_int(int*, int, int) const {}
    double time() const;
    void print_rank_info() const;

private:
    MpiConfig cfg_;
};

}  // namespace seismic
