// api_internal.hpp -- definitions of the opaque ABI objects and helpers
// shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "dnx.h"
#include "host.hpp"
#include "kernels.hpp"

namespace dnx {

// Owning device allocation.
struct DevBuf {
    void *ptr = nullptr;
    size_t size = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf();
    int alloc(size_t bytes);
};

int hip_fail(hipError_t e, const char *what);
size_t prec_size(int prec);
void pack_host(int prec, const double *src, size_t count, std::vector<char> &out);
void unpack_host(int prec, const std::vector<char> &in, size_t count, double *dst);
DevParams dev_params(const dnx_params *p);
// looks up the no-allocation scratch array on `device` (current), once
int noalloc_init(int device);

}  // namespace dnx

struct dnx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
};

struct dnx_problem {
    dnx_ctx *ctx = nullptr;
    dnx::HostProblem hp;
    dnx::DevBuf b_cp, b_lits, b_eidx, b_row, b_unit;
    int32_t *d_cp = nullptr, *d_lits = nullptr, *d_eidx = nullptr, *d_row_ptr = nullptr;
    int8_t *d_unit = nullptr;
    // grouped-exchange tables per precision (index 0 FP64, 1 FP32), built on
    // the first state that uses them: 0 not yet, 1 built, -1 not applicable
    dnx::DevBuf b_gw[2], b_gbase[2], b_geidx[2];
    int gx_state[2] = {0, 0};
    // sliced-exchange tables per precision, the same way
    dnx::DevBuf b_sw[2], b_units[2], b_sbrun[2], b_runs[2], b_sq[2], b_pidx[2];
    int sx_state[2] = {0, 0};
    int sx_nunits[2] = {0, 0}, sx_region[2] = {0, 0};
    int exchange = DNX_EXCHANGE_SLICED;  // for states created from now on
};

struct dnx_state {
    dnx_problem *p = nullptr;
    int prec = DNX_FP64;
    bool grouped = false;  // steps use the grouped exchange
    bool sliced = false;   // steps use the sliced exchange (b_vnext, b_cst)
    dnx::DevBuf b_vnext, b_cst;
    dnx::DevParams prm{};
    dnx::DevBuf b_vd, b_xs, b_xl, b_e, b_ctl;
    dnx::Ctl *ctl = nullptr;     // device control block
    dnx::Ctl *h_ctl = nullptr;   // two pinned host snapshots for polling
    long long enq = 0;           // steps enqueued since creation
    hipGraphExec_t graph = nullptr;
    ~dnx_state()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (h_ctl) (void)hipHostFree(h_ctl);
    }
};

namespace dnx {
DevProb dev_prob(const dnx_problem *p);
// the problem as a state's steps see it (with its exchange tables)
DevProb dev_prob_state(const dnx_state *s);
// dnx_portfolio_solve in DNX_FP32: the copies run as slots of one batch.
int portfolio_batched(dnx_problem *p, const dnx_params *prm, int seed, int threads, int64_t step_cap,
                      double walltime_s, int8_t *assign_out, int64_t *steps_out, double *t_out, int *status,
                      int *winner);
}
