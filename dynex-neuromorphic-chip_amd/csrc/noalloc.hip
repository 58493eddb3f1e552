// noalloc.hip -- the circuit right-hand side on caller-owned device buffers
// with no device allocation (dnx_rhs_device_noalloc).
//
// The evaluation needs device memory beyond its input and output: the
// per-literal contributions (read back by the per-variable fold) and the
// per-workgroup partial sums of loc and energy.  dnx_rhs allocates them per
// call; this form keeps them in one file-scope device array of fixed size,
// so a caller that cannot allocate (code under HIP graph capture, a serving
// loop that must not call hipMalloc) can still evaluate the RHS.
//
// Each GPU holds its own copy of the array at its own address, looked up
// once per device with hipGetSymbolAddress when a context is opened on it;
// host code only ever uses that looked-up address.  Every call on one GPU
// shares the array: calls that overlap on one GPU (two streams, two host
// threads) race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "api_internal.hpp"

static constexpr size_t SCRATCH_BYTES = (size_t)DNX_NOALLOC_SCRATCH_BYTES;

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;

std::mutex g_mu;
void *g_ptr[MAX_DEVICES] = {};        // the array's address on each device
hipError_t g_err[MAX_DEVICES] = {};   // lookup result on each device
bool g_done[MAX_DEVICES] = {};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// loc and energy from the per-workgroup partials, energy summed in
// workgroup order (the order dnx_rhs sums them on the host, so both give
// the same double).
__global__ __launch_bounds__(64) void k_rhs_totals(const int *__restrict__ lpart,
                                                   const double *__restrict__ epart, int nb,
                                                   int *__restrict__ loc, double *__restrict__ energy)
{
    if (loc) {
        int s = 0;
        for (int i = threadIdx.x; i < nb; i += 64) s += lpart[i];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (threadIdx.x == 0) *loc = s;
    }
    if (energy && threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < nb; i++) s += epart[i];
        *energy = s;
    }
}

}  // namespace

namespace dnx {

// Looks the array up on `device` (which must be current); once per device.
int noalloc_init(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return DNX_OK;  // entry point reports it
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_done[device]) return DNX_OK;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    g_err[device] = e;
    g_ptr[device] = e == hipSuccess ? ptr : nullptr;
    g_done[device] = true;
    return DNX_OK;
}

}  // namespace dnx

using namespace dnx;

extern "C" int dnx_rhs_device_noalloc(dnx_problem *p, int prec, const dnx_params *prm, const void *x, void *dxdt,
                                      int *loc, double *energy, void *stream)
{
    if (!p || !prm || (prec != DNX_FP64 && prec != DNX_FP32)) return DNX_E_ARG;
    if (!p->ctx) {
        set_error("problem was created without a context and has no device copy");
        return DNX_E_STATE;
    }
    const HostProblem &h = p->hp;
    const size_t N = (size_t)h.n + 2 * (size_t)h.m, es = prec_size(prec);
    if (N > 0 && (!x || !dxdt)) return DNX_E_ARG;
    const int dev = p->ctx->device;
    if (dev < 0 || dev >= MAX_DEVICES) {
        set_error("device index beyond the no-allocation table");
        return DNX_E_ARG;
    }
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_done[dev]) {
            set_error("no context was opened on this device");
            return DNX_E_STATE;
        }
        if (g_err[dev] != hipSuccess) return hip_fail(g_err[dev], "hipGetSymbolAddress(scratch)");
        base = g_ptr[dev];
    }
    // scratch layout: [contributions (nnz+1 values)][loc partials][energy partials]
    const int nb = rhs_clause_blocks(h.m);
    const size_t e_bytes = align256((size_t)h.nnz() * es + es);
    const size_t l_bytes = align256(sizeof(int) * (size_t)nb);
    const size_t need = e_bytes + l_bytes + sizeof(double) * (size_t)nb;
    if (need > SCRATCH_BYTES) {
        set_error("problem needs " + std::to_string(need) + " bytes of scratch; the no-allocation entry point holds " +
                  std::to_string(SCRATCH_BYTES));
        return DNX_E_BUDGET;
    }
    unsigned char *sp = (unsigned char *)base;
    void *e_buf = sp;
    int *lpart = (int *)(sp + e_bytes);
    double *epart = (double *)(sp + e_bytes + l_bytes);
    hipError_t e;
    if ((e = hipSetDevice(dev)) != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipStream_t st = stream ? (hipStream_t)stream : p->ctx->stream;
    // every partial is written by its workgroup: nothing needs clearing
    launch_rhs(prec, dev_prob(p), dev_params(prm), x, dxdt, e_buf, lpart, epart, st);
    if (loc || energy) k_rhs_totals<<<1, 64, 0, st>>>(lpart, epart, nb, loc, energy);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "rhs launch");
    return DNX_OK;
}
