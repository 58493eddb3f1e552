// api.cpp -- the C ABI of libdnx (include/dnx.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "api_internal.hpp"

namespace dnx {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

int hip_fail(hipError_t e, const char *what)
{
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return DNX_E_HIP;
}

DevProb dev_prob(const dnx_problem *p)
{
    DevProb d;
    d.n = p->hp.n;
    d.m = p->hp.m;
    d.max_deg = p->hp.max_deg;
    d.cp = p->hp.uniform3 ? nullptr : p->d_cp;
    d.lits = p->d_lits;
    d.eidx = p->d_eidx;
    d.row_ptr = p->d_row_ptr;
    d.unit = p->hp.has_unit ? p->d_unit : nullptr;
    d.gw = nullptr;
    d.gbase = nullptr;
    d.geidx = nullptr;
    d.sw = nullptr;
    d.units = d.sbrun = d.runs = nullptr;
    d.sq = nullptr;
    d.pidx = nullptr;
    d.nunits = d.sx_region = 0;
    return d;
}

template <typename V>
static int upload_vec(DevBuf &b, const std::vector<V> &v, hipStream_t st)
{
    int rc = b.alloc(sizeof(V) * (v.empty() ? 1 : v.size()));
    if (rc) return rc;
    hipError_t e;
    if (!v.empty() && (e = hipMemcpyAsync(b.ptr, v.data(), sizeof(V) * v.size(), hipMemcpyHostToDevice, st)) !=
                          hipSuccess)
        return hip_fail(e, "table upload");
    return DNX_OK;
}

// The sliced-exchange tables of a problem for one precision, planned for
// the device's compute units and uploaded once.  Returns whether they exist.
static int ensure_sliced(dnx_problem *p, int prec, bool *have)
{
    const int i = prec == DNX_FP64 ? 0 : 1;
    *have = false;
    if (p->sx_state[i] < 0) return DNX_OK;
    if (p->sx_state[i] == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->ctx->device) != hipSuccess ||
            cus <= 0)
            cus = 256;
        SlicedPlan pl;
        if (!sliced_tables(p->hp, sx_slice(prec), sx_sub(prec), SX_F, cus, sx_max_rounds(prec), sx_region_cap(prec),
                           pl)) {
            p->sx_state[i] = -1;
            return DNX_OK;
        }
        for (size_t k = 0; k + 1 < pl.sb_run.size(); k++)
            if (pl.sb_run[k + 1] - pl.sb_run[k] > SX_MAXRUNS) {
                p->sx_state[i] = -1;
                return DNX_OK;
            }
        hipStream_t st = p->ctx->stream;
        int rc;
        if ((rc = upload_vec(p->b_sw[i], pl.gw, st)) || (rc = upload_vec(p->b_units[i], pl.units, st)) ||
            (rc = upload_vec(p->b_sbrun[i], pl.sb_run, st)) || (rc = upload_vec(p->b_runs[i], pl.runs, st)) ||
            (rc = upload_vec(p->b_sq[i], pl.q, st)) || (rc = upload_vec(p->b_pidx[i], pl.pidx, st)))
            return rc;
        hipError_t e;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "exchange tables sync");
        p->sx_nunits[i] = (int)pl.units.size() / 8;
        p->sx_region[i] = pl.stride;
        p->sx_state[i] = 1;
    }
    *have = true;
    return DNX_OK;
}

DevProb dev_prob_state(const dnx_state *s)
{
    DevProb d = dev_prob(s->p);
    if (s->sliced) {
        const int i = s->prec == DNX_FP64 ? 0 : 1;
        const dnx_problem *p = s->p;
        d.sw = (const uint32_t *)p->b_sw[i].ptr;
        d.units = (const int32_t *)p->b_units[i].ptr;
        d.sbrun = (const int32_t *)p->b_sbrun[i].ptr;
        d.runs = (const int32_t *)p->b_runs[i].ptr;
        d.sq = (const uint16_t *)p->b_sq[i].ptr;
        d.pidx = (const int32_t *)p->b_pidx[i].ptr;
        d.nunits = p->sx_nunits[i];
        d.sx_region = p->sx_region[i];
    }
    if (s->grouped) {
        const int i = s->prec == DNX_FP64 ? 0 : 1;
        d.gw = (const uint32_t *)s->p->b_gw[i].ptr;
        d.gbase = (const int32_t *)s->p->b_gbase[i].ptr;
        d.geidx = (const int32_t *)s->p->b_geidx[i].ptr;
    }
    return d;
}

// The grouped-exchange tables of a problem for one precision (its
// variable-pass workgroup size), built and uploaded once.  Returns whether
// they exist.
static int ensure_grouped(dnx_problem *p, int prec, bool *have)
{
    const int i = prec == DNX_FP64 ? 0 : 1;
    *have = false;
    if (p->gx_state[i] < 0) return DNX_OK;
    if (p->gx_state[i] == 0) {
        std::vector<uint32_t> gw;
        std::vector<int32_t> gb, ge;
        if (!grouped_tables(p->hp, gx_var_block(prec), gw, gb, ge)) {
            p->gx_state[i] = -1;
            return DNX_OK;
        }
        int rc;
        if ((rc = p->b_gw[i].alloc(sizeof(uint32_t) * gw.size())) ||
            (rc = p->b_gbase[i].alloc(sizeof(int32_t) * gb.size())) ||
            (rc = p->b_geidx[i].alloc(sizeof(int32_t) * ge.size())))
            return rc;
        hipStream_t st = p->ctx->stream;
        hipError_t e;
        if ((!gw.empty() && (e = hipMemcpyAsync(p->b_gw[i].ptr, gw.data(), p->b_gw[i].size, hipMemcpyHostToDevice,
                                                st)) != hipSuccess) ||
            (!gb.empty() && (e = hipMemcpyAsync(p->b_gbase[i].ptr, gb.data(), p->b_gbase[i].size,
                                                hipMemcpyHostToDevice, st)) != hipSuccess) ||
            (!ge.empty() && (e = hipMemcpyAsync(p->b_geidx[i].ptr, ge.data(), p->b_geidx[i].size,
                                                hipMemcpyHostToDevice, st)) != hipSuccess))
            return hip_fail(e, "exchange tables upload");
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "exchange tables sync");
        p->gx_state[i] = 1;
    }
    *have = true;
    return DNX_OK;
}

DevParams dev_params(const dnx_params *p)
{
    DevParams d;
    d.alpha = p->alpha;
    d.beta = p->beta;
    d.gamma = p->gamma;
    d.delta = p->delta;
    d.epsilon = p->epsilon;
    d.zeta = p->zeta;
    d.xl_max = (double)p->xl_max;
    return d;
}

// Device buffer that frees itself.
DevBuf::~DevBuf()
{
    if (ptr) (void)hipFree(ptr);
}
int DevBuf::alloc(size_t bytes)
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    size = bytes;
    if (bytes == 0) return DNX_OK;
    hipError_t e = hipMalloc(&ptr, bytes);
    if (e != hipSuccess) {
        ptr = nullptr;
        set_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
        return DNX_E_NOMEM;
    }
    return DNX_OK;
}

int upload_problem(dnx_problem *p)
{
    const HostProblem &h = p->hp;
    int rc;
    if ((rc = p->b_cp.alloc(sizeof(int32_t) * h.cp.size())) ||
        (rc = p->b_lits.alloc(sizeof(int32_t) * h.lits.size())) ||
        (rc = p->b_eidx.alloc(sizeof(int32_t) * h.eidx.size())) ||
        (rc = p->b_row.alloc(sizeof(int32_t) * h.row_ptr.size())) ||
        (rc = p->b_unit.alloc(h.unit.size())))
        return rc;
    hipStream_t st = p->ctx->stream;
    hipError_t e;
#define UP(buf, vec)                                                                                  \
    if (!vec.empty() &&                                                                               \
        (e = hipMemcpyAsync(buf.ptr, vec.data(), buf.size, hipMemcpyHostToDevice, st)) != hipSuccess) \
        return hip_fail(e, "upload");
    UP(p->b_cp, h.cp)
    UP(p->b_lits, h.lits)
    UP(p->b_eidx, h.eidx)
    UP(p->b_row, h.row_ptr)
    UP(p->b_unit, h.unit)
#undef UP
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "upload sync");
    p->d_cp = (int32_t *)p->b_cp.ptr;
    p->d_lits = (int32_t *)p->b_lits.ptr;
    p->d_eidx = (int32_t *)p->b_eidx.ptr;
    p->d_row_ptr = (int32_t *)p->b_row.ptr;
    p->d_unit = (int8_t *)p->b_unit.ptr;
    return DNX_OK;
}

size_t prec_size(int prec) { return prec == DNX_FP64 ? 8 : 4; }

// host double <-> device precision staging
void pack_host(int prec, const double *src, size_t count, std::vector<char> &out)
{
    out.resize(count * prec_size(prec));
    if (prec == DNX_FP64) {
        memcpy(out.data(), src, count * 8);
    } else {
        float *f = (float *)out.data();
        for (size_t i = 0; i < count; i++) f[i] = (float)src[i];
    }
}
void unpack_host(int prec, const std::vector<char> &in, size_t count, double *dst)
{
    if (prec == DNX_FP64) {
        memcpy(dst, in.data(), count * 8);
    } else {
        const float *f = (const float *)in.data();
        for (size_t i = 0; i < count; i++) dst[i] = (double)f[i];
    }
}

static const int kChunk = 30;  // steps per captured graph (multiple of 3)

static void init_ctl(Ctl &c)
{
    memset(&c, 0, sizeof(c));
    c.status = DNX_RUNNING;
    c.stop_on_solve = 1;
    c.cap = LLONG_MAX;
}

static void one_step(dnx_state *s, const DevProb &P, int slot, hipStream_t st)
{
    if (s->sliced)
        launch_step_sliced(s->prec, P, s->prm, s->b_vd.ptr, s->b_xs.ptr, s->b_xl.ptr, s->b_e.ptr, s->b_vnext.ptr,
                           s->b_cst.ptr, s->ctl, slot, st);
    else
        launch_step(s->prec, P, s->prm, s->b_vd.ptr, s->b_xs.ptr, s->b_xl.ptr, s->b_e.ptr, s->ctl, slot, st);
}

static int state_enqueue_steps(dnx_state *s, long long count)
{
    hipStream_t st = s->p->ctx->stream;
    DevProb P = dev_prob_state(s);
    // align to the graph's slot phase, then replay whole chunks
    while (count > 0 && (s->enq % 3) != 0) {
        one_step(s, P, (int)(s->enq % 3), st);
        s->enq++;
        count--;
    }
    while (count >= kChunk) {
        if (!s->graph) {
            hipGraph_t g;
            hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
            for (int i = 0; i < kChunk; i++)
                one_step(s, P, i % 3, st);
            if ((e = hipStreamEndCapture(st, &g)) != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
            e = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                s->graph = nullptr;
                return hip_fail(e, "hipGraphInstantiate");
            }
        }
        hipError_t e = hipGraphLaunch(s->graph, st);
        if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
        s->enq += kChunk;
        count -= kChunk;
    }
    while (count > 0) {
        one_step(s, P, (int)(s->enq % 3), st);
        s->enq++;
        count--;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "step launch");
    return DNX_OK;
}

// the grouped exchange applies to 3-SAT problems of up to GX_MAX_VARS variables
static bool m_ok_grouped(const HostProblem &h) { return h.uniform3 && h.n <= GX_MAX_VARS && h.m > 0; }

// clauses of the state x (reference layout) with C >= 1/2, i.e. whose
// literals all have q*clamp(V) <= 0 (dynex.cc:384-387, 499)
static int count_loc(const HostProblem &h, const double *x)
{
    int loc = 0;
    for (int c = 0; c < h.m; c++) {
        bool unsat = true;
        if (h.cp[(size_t)c + 1] - h.cp[(size_t)c] == 1) unsat = false;  // C = 0 (dynex.cc:517)
        for (int q = h.cp[(size_t)c]; unsat && q < h.cp[(size_t)c + 1]; q++) {
            int32_t lit = h.lits[(size_t)q];
            double v = x[std::abs(lit) - 1];
            v = v < -1.0 ? -1.0 : (v > 1.0 ? 1.0 : v);
            if ((lit > 0 ? v : -v) > 0.0) unsat = false;
        }
        loc += unsat;
    }
    return loc;
}

// status of a control-block snapshot: the stored status, or the `done`
// verdict of the last evaluated state when no later pass has run
static int ctl_status(const Ctl &c)
{
    if (c.status != DNX_RUNNING) return c.status;
    return c.done[c.steps % 3];
}

}  // namespace dnx

using namespace dnx;

static int host_only(const dnx_problem *)
{
    set_error("problem was created without a context and has no device copy");
    return DNX_E_STATE;
}

extern "C" {

const char *dnx_last_error(void) { return g_err.c_str(); }

int dnx_params_default(int m, int xl_max_flag, dnx_params *p)
{
    if (!p || m < 0) return DNX_E_ARG;
    p->alpha = 2.5;                 // dynex.cc:132
    p->beta = 20.0;                 // dynex.cc:134
    p->gamma = 25.0 / 100.0;        // dynex.cc:135
    p->delta = 5.0 / 100.0;         // dynex.cc:136
    p->epsilon = 1.0 / 10.0;        // dynex.cc:137
    p->zeta = 1.0 / 10.0;           // dynex.cc:138
    // dynex.cc:2358-2359: int product (wrapping), INT_MAX if not positive
    int32_t w = (int32_t)((uint32_t)xl_max_flag * (uint32_t)m);
    if (w <= 0) w = INT_MAX;
    p->xl_max = w;
    return DNX_OK;
}

int dnx_open(int device, dnx_ctx **out)
{
    if (!out) return DNX_E_ARG;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    if (device < 0 || device >= count) {
        set_error("device " + std::to_string(device) + " not present (" + std::to_string(count) + " visible)");
        return DNX_E_ARG;
    }
    if ((e = hipSetDevice(device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
    (void)noalloc_init(device);  // a failed lookup is reported by the entry point that needs it
    dnx_ctx *c = new dnx_ctx();
    c->device = device;
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) {
        delete c;
        return hip_fail(e, "hipStreamCreate");
    }
    *out = c;
    return DNX_OK;
}

int dnx_close(dnx_ctx *ctx)
{
    if (!ctx) return DNX_E_ARG;
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return DNX_OK;
}

static int finish_problem(dnx_ctx *ctx, dnx_problem *p, dnx_problem **out)
{
    if (!ctx) {  // host-only problem: parsing, inspection, checks
        *out = p;
        return DNX_OK;
    }
    (void)hipSetDevice(ctx->device);
    int rc = upload_problem(p);
    if (rc != DNX_OK) {
        delete p;
        return rc;
    }
    *out = p;
    return DNX_OK;
}

int dnx_problem_from_dimacs(dnx_ctx *ctx, const char *path, dnx_problem **out)
{
    if (!path || !out) return DNX_E_ARG;
    dnx_problem *p = new dnx_problem();
    p->ctx = ctx;
    std::string err;
    int rc = parse_dimacs(path, p->hp, err);
    if (rc != DNX_OK) {
        set_error(err);
        delete p;
        return rc;
    }
    return finish_problem(ctx, p, out);
}

int dnx_problem_from_lits(dnx_ctx *ctx, int n, int m, const int32_t *clause_ptr, const int32_t *lits,
                          dnx_problem **out)
{
    if (!out || n < 0 || m < 0 || (m > 0 && !clause_ptr)) return DNX_E_ARG;
    dnx_problem *p = new dnx_problem();
    p->ctx = ctx;
    p->hp.n = n;
    p->hp.m = m;
    if (clause_ptr) {
        p->hp.cp.assign(clause_ptr, clause_ptr + m + 1);
    } else {
        p->hp.cp.assign(1, 0);
    }
    int64_t nnz = p->hp.cp.back();
    if (nnz < 0 || (nnz > 0 && !lits)) {
        delete p;
        return DNX_E_ARG;
    }
    for (int c = 0; c < m; c++)
        if (p->hp.cp[(size_t)c + 1] < p->hp.cp[(size_t)c]) {
            delete p;
            set_error("clause_ptr is not non-decreasing");
            return DNX_E_ARG;
        }
    if (nnz > 0) p->hp.lits.assign(lits, lits + nnz);
    int rc = finalize_problem(p->hp);
    if (rc != DNX_OK) {
        set_error(rc == DNX_E_TOOLONG ? "clause longer than DNX_MAX_LITS" : "literal 0 or out of range");
        delete p;
        return rc;
    }
    return finish_problem(ctx, p, out);
}

int dnx_problem_set_exchange(dnx_problem *p, int exchange)
{
    if (!p || exchange < DNX_EXCHANGE_LITERAL || exchange > DNX_EXCHANGE_SLICED) return DNX_E_ARG;
    p->exchange = exchange;
    return DNX_OK;
}

int dnx_problem_info(const dnx_problem *p, int *n, int *m, int64_t *nnz)
{
    if (!p) return DNX_E_ARG;
    if (n) *n = p->hp.n;
    if (m) *m = p->hp.m;
    if (nnz) *nnz = p->hp.nnz();
    return DNX_OK;
}

int dnx_problem_clauses(const dnx_problem *p, int32_t *clause_ptr, int32_t *lits)
{
    if (!p) return DNX_E_ARG;
    if (clause_ptr) memcpy(clause_ptr, p->hp.cp.data(), sizeof(int32_t) * p->hp.cp.size());
    if (lits && !p->hp.lits.empty()) memcpy(lits, p->hp.lits.data(), sizeof(int32_t) * p->hp.lits.size());
    return DNX_OK;
}

int dnx_problem_free(dnx_problem *p)
{
    if (!p) return DNX_E_ARG;
    if (p->ctx) {
        (void)hipSetDevice(p->ctx->device);
        (void)hipStreamSynchronize(p->ctx->stream);
    }
    delete p;
    return DNX_OK;
}

int dnx_rhs(dnx_problem *p, int prec, const dnx_params *prm, const double *x, double *dxdt, int *loc,
            double *energy)
{
    if (!p || !prm || !x || !dxdt || (prec != DNX_FP64 && prec != DNX_FP32)) return DNX_E_ARG;
    if (!p->ctx) return host_only(p);
    (void)hipSetDevice(p->ctx->device);
    const HostProblem &h = p->hp;
    const size_t N = (size_t)h.n + 2 * (size_t)h.m, es = prec_size(prec);
    const int nb = rhs_clause_blocks(h.m);
    DevBuf bx, bd, be, bloc, bep;
    int rc;
    if ((rc = bx.alloc(N * es)) || (rc = bd.alloc(N * es)) || (rc = be.alloc((size_t)h.nnz() * es + es)) ||
        (rc = bloc.alloc(sizeof(int) * (size_t)nb)) || (rc = bep.alloc(sizeof(double) * (size_t)nb)))
        return rc;
    hipStream_t st = p->ctx->stream;
    std::vector<char> stage;
    pack_host(prec, x, N, stage);
    hipError_t e;
    if (N && (e = hipMemcpyAsync(bx.ptr, stage.data(), N * es, hipMemcpyHostToDevice, st)) != hipSuccess)
        return hip_fail(e, "rhs upload");
    if ((e = hipMemsetAsync(bloc.ptr, 0, sizeof(int) * (size_t)nb, st)) != hipSuccess) return hip_fail(e, "memset");
    if ((e = hipMemsetAsync(bep.ptr, 0, sizeof(double) * (size_t)nb, st)) != hipSuccess)
        return hip_fail(e, "memset");
    launch_rhs(prec, dev_prob(p), dev_params(prm), bx.ptr, bd.ptr, be.ptr, (int *)bloc.ptr, (double *)bep.ptr, st);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "rhs launch");
    std::vector<double> ep((size_t)nb);
    std::vector<int> lp((size_t)nb);
    if (N && (e = hipMemcpyAsync(stage.data(), bd.ptr, N * es, hipMemcpyDeviceToHost, st)) != hipSuccess)
        return hip_fail(e, "rhs download");
    if ((e = hipMemcpyAsync(lp.data(), bloc.ptr, sizeof(int) * (size_t)nb, hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (e = hipMemcpyAsync(ep.data(), bep.ptr, sizeof(double) * (size_t)nb, hipMemcpyDeviceToHost, st)) !=
            hipSuccess)
        return hip_fail(e, "rhs download");
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "rhs sync");
    unpack_host(prec, stage, N, dxdt);
    if (loc) {
        int tot = 0;
        for (int v : lp) tot += v;
        *loc = tot;
    }
    if (energy) {
        double s = 0.0;
        for (double v : ep) s += v;
        *energy = s;
    }
    return DNX_OK;
}

int dnx_euler(dnx_problem *p, int prec, const dnx_params *prm, const double *x, const double *dxdt, double h,
              double *x_out)
{
    if (!p || !prm || !x || !dxdt || !x_out || (prec != DNX_FP64 && prec != DNX_FP32)) return DNX_E_ARG;
    if (!p->ctx) return host_only(p);
    (void)hipSetDevice(p->ctx->device);
    const size_t N = (size_t)p->hp.n + 2 * (size_t)p->hp.m, es = prec_size(prec);
    if (N == 0) return DNX_OK;
    DevBuf bx, bd, bo;
    int rc;
    if ((rc = bx.alloc(N * es)) || (rc = bd.alloc(N * es)) || (rc = bo.alloc(N * es))) return rc;
    hipStream_t st = p->ctx->stream;
    std::vector<char> s1, s2;
    pack_host(prec, x, N, s1);
    pack_host(prec, dxdt, N, s2);
    hipError_t e;
    if ((e = hipMemcpyAsync(bx.ptr, s1.data(), N * es, hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(bd.ptr, s2.data(), N * es, hipMemcpyHostToDevice, st)) != hipSuccess)
        return hip_fail(e, "euler upload");
    launch_euler(prec, dev_prob(p), dev_params(prm), bx.ptr, bd.ptr, h, bo.ptr, st);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "euler launch");
    if ((e = hipMemcpyAsync(s1.data(), bo.ptr, N * es, hipMemcpyDeviceToHost, st)) != hipSuccess)
        return hip_fail(e, "euler download");
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "euler sync");
    unpack_host(prec, s1, N, x_out);
    return DNX_OK;
}

int dnx_state_create(dnx_problem *p, int prec, const dnx_params *prm, const double *x0, dnx_state **out)
{
    if (!p || !prm || !x0 || !out || (prec != DNX_FP64 && prec != DNX_FP32)) return DNX_E_ARG;
    if (!p->ctx) return host_only(p);
    (void)hipSetDevice(p->ctx->device);
    const HostProblem &h = p->hp;
    const size_t es = prec_size(prec), N = (size_t)h.n + 2 * (size_t)h.m;
    dnx_state *s = new dnx_state();
    s->p = p;
    s->prec = prec;
    s->prm = dev_params(prm);
    int rc;
    if (p->exchange == DNX_EXCHANGE_SLICED && h.uniform3 && h.m > 0 && (rc = ensure_sliced(p, prec, &s->sliced))) {
        delete s;
        return rc;
    }
    if (!s->sliced && p->exchange != DNX_EXCHANGE_LITERAL && m_ok_grouped(h) &&
        (rc = ensure_grouped(p, prec, &s->grouped))) {
        delete s;
        return rc;
    }
    if (s->sliced && ((rc = s->b_vnext.alloc(4 * es * (size_t)sx_vstride(h.n > 0 ? h.n : 1))) ||
                      (rc = s->b_cst.alloc(es * ((size_t)h.m + 1))))) {
        delete s;
        return rc;
    }
    DevBuf bx;
    // one entry past the end of Xs, Xl, e (and the stored C): the sliced
    // clause pass sends the stores of lanes without a clause there; the
    // sliced exchange's e holds every fold block's padded region
    size_t elen = (size_t)h.nnz();
    if (s->sliced) elen = std::max(elen, (size_t)((h.n + SX_F - 1) / SX_F) * (size_t)p->sx_region[prec == DNX_FP64 ? 0 : 1]);
    if ((rc = s->b_vd.alloc(2 * es * (size_t)(h.n > 0 ? h.n : 1))) ||
        (rc = s->b_xs.alloc(es * ((size_t)h.m + 1))) || (rc = s->b_xl.alloc(es * ((size_t)h.m + 1))) ||
        (rc = s->b_e.alloc(es * (elen + 1))) || (rc = s->b_ctl.alloc(sizeof(Ctl))) ||
        (rc = bx.alloc(es * (N ? N : 1)))) {
        delete s;
        return rc;
    }
    s->ctl = (Ctl *)s->b_ctl.ptr;
    hipError_t e;
    if ((e = hipHostMalloc((void **)&s->h_ctl, 2 * sizeof(Ctl), hipHostMallocDefault)) != hipSuccess) {
        s->h_ctl = nullptr;
        delete s;
        return hip_fail(e, "hipHostMalloc");
    }
    Ctl c;
    init_ctl(c);
    c.loc_min = c.loc_last = h.m;
    std::vector<char> stage;
    pack_host(prec, x0, N, stage);
    hipStream_t st = p->ctx->stream;
    if ((N && (e = hipMemcpyAsync(bx.ptr, stage.data(), N * es, hipMemcpyHostToDevice, st)) != hipSuccess) ||
        (e = hipMemcpyAsync(s->ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice, st)) != hipSuccess) {
        delete s;
        return hip_fail(e, "state upload");
    }
    launch_load_state(prec, dev_prob_state(s), bx.ptr, s->b_vd.ptr, s->b_xs.ptr, s->b_xl.ptr, st);
    if ((e = hipStreamSynchronize(st)) != hipSuccess) {
        delete s;
        return hip_fail(e, "state load");
    }
    *out = s;
    return DNX_OK;
}

int dnx_state_free(dnx_state *s)
{
    if (!s) return DNX_E_ARG;
    (void)hipSetDevice(s->p->ctx->device);
    (void)hipStreamSynchronize(s->p->ctx->stream);
    delete s;
    return DNX_OK;
}

static int set_mode(dnx_state *s, int stop, long long cap)
{
    // the two fields sit next to each other only in part: write them apart
    hipStream_t st = s->p->ctx->stream;
    hipError_t e;
    int v = stop;
    if ((e = hipMemcpyAsync(&s->ctl->stop_on_solve, &v, sizeof(int), hipMemcpyHostToDevice, st)) != hipSuccess ||
        (e = hipMemcpyAsync(&s->ctl->cap, &cap, sizeof(long long), hipMemcpyHostToDevice, st)) != hipSuccess)
        return hip_fail(e, "mode upload");
    // the source values live on this stack frame: wait before returning
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "mode sync");
    return DNX_OK;
}

int dnx_state_track_loc(dnx_state *s, int enable)
{
    if (!s) return DNX_E_ARG;
    (void)hipSetDevice(s->p->ctx->device);
    hipStream_t st = s->p->ctx->stream;
    // track_loc, loc_min and loc_last are consecutive ints of the block
    const int v[3] = {enable ? 1 : 0, s->p->hp.m, s->p->hp.m};
    hipError_t e;
    if ((e = hipMemcpyAsync(&s->ctl->track_loc, v, sizeof v, hipMemcpyHostToDevice, st)) != hipSuccess)
        return hip_fail(e, "track_loc upload");
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "track_loc sync");
    return DNX_OK;
}

int dnx_state_best_loc(dnx_state *s, int *best_loc, int *last_loc)
{
    if (!s) return DNX_E_ARG;
    (void)hipSetDevice(s->p->ctx->device);
    hipStream_t st = s->p->ctx->stream;
    int v[3] = {0, 0, 0};
    hipError_t e;
    if ((e = hipMemcpyAsync(v, &s->ctl->track_loc, sizeof v, hipMemcpyDeviceToHost, st)) != hipSuccess)
        return hip_fail(e, "loc download");
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "loc sync");
    if (best_loc) *best_loc = v[1];
    if (last_loc) *last_loc = v[2];
    return DNX_OK;
}

int dnx_state_step(dnx_state *s, int nsteps)
{
    if (!s || nsteps < 0) return DNX_E_ARG;
    (void)hipSetDevice(s->p->ctx->device);
    int rc = set_mode(s, 0, LLONG_MAX);
    if (rc) return rc;
    return state_enqueue_steps(s, nsteps);
}

int dnx_state_time(dnx_state *s, int nsteps, double *ms)
{
    if (!s || nsteps <= 0 || !ms) return DNX_E_ARG;
    (void)hipSetDevice(s->p->ctx->device);
    int rc = set_mode(s, 0, LLONG_MAX);
    if (rc) return rc;
    hipStream_t st = s->p->ctx->stream;
    hipEvent_t a, b;
    hipError_t e;
    if ((e = hipEventCreate(&a)) != hipSuccess) return hip_fail(e, "event");
    if ((e = hipEventCreate(&b)) != hipSuccess) {
        (void)hipEventDestroy(a);
        return hip_fail(e, "event");
    }
    e = hipEventRecord(a, st);
    if (e == hipSuccess) rc = state_enqueue_steps(s, nsteps);
    if (e == hipSuccess && rc == DNX_OK) e = hipEventRecord(b, st);
    if (e == hipSuccess && rc == DNX_OK) e = hipEventSynchronize(b);
    float f = 0.0f;
    if (e == hipSuccess && rc == DNX_OK) e = hipEventElapsedTime(&f, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "state timing");
    *ms = (double)f;
    return DNX_OK;
}

int dnx_state_read(dnx_state *s, double *x, double *t, int64_t *steps, int *last_level, int *loc)
{
    if (!s) return DNX_E_ARG;
    (void)hipSetDevice(s->p->ctx->device);
    const HostProblem &h = s->p->hp;
    const size_t es = prec_size(s->prec), N = (size_t)h.n + 2 * (size_t)h.m;
    hipStream_t st = s->p->ctx->stream;
    hipError_t e;
    DevBuf bo;
    std::vector<char> stage;
    if (x && N) {
        int rc = bo.alloc(N * es);
        if (rc) return rc;
        launch_materialize(s->prec, dev_prob_state(s), s->prm, s->b_vd.ptr, s->b_xs.ptr, s->b_xl.ptr, s->ctl, bo.ptr,
                           st);
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "materialize");
        stage.resize(N * es);
        if ((e = hipMemcpyAsync(stage.data(), bo.ptr, N * es, hipMemcpyDeviceToHost, st)) != hipSuccess)
            return hip_fail(e, "state download");
    }
    Ctl c;
    if ((e = hipMemcpyAsync(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st)) != hipSuccess)
        return hip_fail(e, "ctl download");
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "state sync");
    if (x && N) unpack_host(s->prec, stage, N, x);
    const int sl = (int)(c.steps % 3);
    const int plev = (c.status == DNX_RUNNING && !c.done[sl] && c.stepped[sl]) ? ctl_level(c, sl) : -1;
    static const double dts[5] = {0.125, 0.0625, 0.03125, 0.015625, 0.0078125};
    if (t) *t = c.t + (plev >= 0 ? dts[plev] : 0.0);
    if (steps) *steps = c.steps + (plev >= 0 ? 1 : 0);
    if (last_level) *last_level = plev;
    if (loc) *loc = (x && N) ? count_loc(h, x) : -1;
    return DNX_OK;
}

int dnx_state_solve(dnx_state *s, int64_t step_cap, double walltime_s, int8_t *assign_out, int64_t *steps_out,
                    double *t_out, int *status)
{
    if (!s || step_cap < 0 || !status) return DNX_E_ARG;
    (void)hipSetDevice(s->p->ctx->device);
    const HostProblem &h = s->p->hp;
    hipStream_t st = s->p->ctx->stream;
    hipError_t e;
    // the device step index equals the number of enqueued steps while running;
    // a cap that would pass LLONG_MAX is no cap
    const bool no_cap = step_cap >= LLONG_MAX - 1 - s->enq;
    int rc = set_mode(s, 1, no_cap ? LLONG_MAX : (long long)s->enq + step_cap);
    if (rc) return rc;
    auto t0 = std::chrono::steady_clock::now();
    hipEvent_t ev[2];
    if ((e = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming)) != hipSuccess)
        return hip_fail(e, "hipEventCreate");
    auto issue = [&](int b) -> int {
        int r = state_enqueue_steps(s, kChunk);
        if (r) return r;
        hipError_t ee;
        if ((ee = hipMemcpyAsync(&s->h_ctl[b], s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st)) != hipSuccess)
            return hip_fail(ee, "ctl poll");
        if ((ee = hipEventRecord(ev[b], st)) != hipSuccess) return hip_fail(ee, "event");
        return DNX_OK;
    };
    int st_final = DNX_RUNNING;
    // states 0..step_cap are evaluated: step_cap+1 step pairs reach the cap
    const long long need = no_cap ? LLONG_MAX : (long long)step_cap + 1;
    long long issued = 0;
    int inflight[2], nin = 0, nextb = 0;
    while (rc == DNX_OK) {
        // keep up to two chunks queued so the device never waits on a poll
        while (nin < 2 && issued < need) {
            rc = issue(nextb);
            if (rc) break;
            inflight[nin++] = nextb;
            nextb ^= 1;
            issued += kChunk;
        }
        if (rc || nin == 0) break;
        const int b = inflight[0];
        inflight[0] = inflight[1];
        nin--;
        if ((e = hipEventSynchronize(ev[b])) != hipSuccess) {
            rc = hip_fail(e, "poll sync");
            break;
        }
        int sv = ctl_status(s->h_ctl[b]);
        if (sv != DNX_RUNNING) {
            st_final = sv;
            break;
        }
        if (walltime_s > 0) {
            double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (el > walltime_s) {
                st_final = DNX_UNKNOWN;
                break;
            }
        }
    }
    // drain the queued chunk before its event is destroyed
    e = hipStreamSynchronize(st);
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (e != hipSuccess) return hip_fail(e, "solve drain");
    if (rc) return rc;
    // final snapshot after the queue drained
    Ctl c;
    if ((e = hipMemcpy(&c, s->ctl, sizeof(Ctl), hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(e, "ctl read");
    int sv = ctl_status(c);
    if (sv != DNX_RUNNING) st_final = sv;
    if (steps_out) *steps_out = c.steps;
    if (t_out) *t_out = c.t;
    if (st_final == DNX_SAT) {
        // exact integer check of the verified state's signs on the host
        std::vector<double> x((size_t)h.n + 2 * (size_t)h.m);
        rc = dnx_state_read(s, x.data(), nullptr, nullptr, nullptr, nullptr);
        if (rc) return rc;
        std::vector<int8_t> a((size_t)h.n);
        for (int v = 0; v < h.n; v++) a[(size_t)v] = x[(size_t)v] >= 0.0 ? 1 : -1;
        if (count_unsat(h, a.data()) != 0) {
            set_error("device reported a solution that fails the clause check");
            *status = DNX_FAILED;
            return DNX_E_STATE;
        }
        if (assign_out) memcpy(assign_out, a.data(), (size_t)h.n);
    }
    *status = st_final == DNX_RUNNING ? DNX_UNKNOWN : st_final;
    return DNX_OK;
}

int dnx_solve(dnx_problem *p, int prec, const dnx_params *prm, const double *x0, int64_t step_cap,
              double walltime_s, int8_t *assign_out, int64_t *steps_out, double *t_out, int *status)
{
    dnx_state *s = nullptr;
    int rc = dnx_state_create(p, prec, prm, x0, &s);
    if (rc) return rc;
    rc = dnx_state_solve(s, step_cap, walltime_s, assign_out, steps_out, t_out, status);
    dnx_state_free(s);
    return rc;
}

int dnx_device_copy(dnx_ctx *ctx, int64_t bytes, int repeats, double *ms)
{
    if (!ctx || bytes < 16 || repeats <= 0 || !ms) return DNX_E_ARG;
    (void)hipSetDevice(ctx->device);
    bytes &= ~(int64_t)15;
    DevBuf a, b;
    int rc;
    if ((rc = a.alloc((size_t)bytes)) || (rc = b.alloc((size_t)bytes))) return rc;
    hipStream_t st = ctx->stream;
    hipError_t e;
    if ((e = hipMemsetAsync(a.ptr, 1, (size_t)bytes, st)) != hipSuccess) return hip_fail(e, "memset");
    hipEvent_t e0, e1;
    if ((e = hipEventCreate(&e0)) != hipSuccess) return hip_fail(e, "event");
    if ((e = hipEventCreate(&e1)) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(e, "event");
    }
    e = hipEventRecord(e0, st);
    for (int i = 0; i < repeats && e == hipSuccess; i++) {
        launch_copy(a.ptr, b.ptr, bytes, st);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float f = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return hip_fail(e, "copy");
    *ms = (double)f / repeats;
    return DNX_OK;
}

int dnx_init_reference(const dnx_problem *p, int seed, double *x0)
{
    if (!p || !x0) return DNX_E_ARG;
    init_reference(p->hp, seed, x0);
    return DNX_OK;
}

int dnx_init_reference_thread(const dnx_problem *p, int seed, int tid, double *x0)
{
    if (!p || !x0 || tid < 0) return DNX_E_ARG;
    init_reference_thread(p->hp, seed, tid, x0);
    return DNX_OK;
}

// Portfolio (the reference's -w, dynex.cc:2731-2801): `threads` copies of
// the problem with the reference's per-thread initial states and alpha/i
// (dynex.cc:2657) advance in rounds of equal step counts; the first round in
// which any copy verifies ends the run, and the winner is the verified copy
// with the fewest steps (lowest thread index on a tie).  The reference takes
// whichever thread happens to finish first in wall-clock time; this choice
// is deterministic.  FP32 runs the copies as slots of one batch; FP64 (which
// the batched engine does not run) as one state per copy, in rounds of one
// graph chunk each.
int dnx_portfolio_solve(dnx_problem *p, int prec, const dnx_params *prm, int seed, int threads, int64_t step_cap,
                        double walltime_s, int8_t *assign_out, int64_t *steps_out, double *t_out, int *status,
                        int *winner)
{
    if (!p || !prm || threads <= 0 || step_cap < 0 || !status || (prec != DNX_FP64 && prec != DNX_FP32))
        return DNX_E_ARG;
    if (!p->ctx) return host_only(p);
    (void)hipSetDevice(p->ctx->device);
    if (prec == DNX_FP32)
        return portfolio_batched(p, prm, seed, threads, step_cap, walltime_s, assign_out, steps_out, t_out, status,
                                 winner);
    const HostProblem &h = p->hp;
    const size_t N = (size_t)h.n + 2 * (size_t)h.m;
    std::vector<dnx_state *> st((size_t)threads, nullptr);
    struct Free {
        std::vector<dnx_state *> &v;
        ~Free()
        {
            for (auto *s : v)
                if (s) dnx_state_free(s);
        }
    } guard{st};
    std::vector<double> x(N ? N : 1);
    for (int i = 0; i < threads; i++) {
        init_reference_thread(h, seed, i, x.data());
        dnx_params pi = *prm;
        if (i > 0) pi.alpha = prm->alpha / i;  // dynex.cc:2657
        int rc = dnx_state_create(p, prec, &pi, x.data(), &st[(size_t)i]);
        if (rc) return rc;
        if ((rc = set_mode(st[(size_t)i], 1, step_cap))) return rc;
    }
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = p->ctx->stream;
    std::vector<Ctl> c((size_t)threads);
    *status = DNX_UNKNOWN;
    if (winner) *winner = -1;
    for (;;) {
        for (auto *sp : st) {
            int rc = state_enqueue_steps(sp, kChunk);
            if (rc) return rc;
        }
        hipError_t e = hipSuccess;
        for (int i = 0; i < threads && e == hipSuccess; i++)
            e = hipMemcpyAsync(&c[(size_t)i], st[(size_t)i]->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(e, "portfolio poll");
        int best = -1, running = 0;
        for (int i = 0; i < threads; i++) {
            const int sv = ctl_status(c[(size_t)i]);
            if (sv == DNX_RUNNING) running++;
            if (sv == DNX_SAT && (best < 0 || c[(size_t)i].steps < c[(size_t)best].steps)) best = i;
        }
        if (best >= 0) {
            int stt = DNX_UNKNOWN;
            int rc = dnx_state_solve(st[(size_t)best], 0, 0.0, assign_out, steps_out, t_out, &stt);
            if (rc) return rc;
            *status = stt;
            if (winner) *winner = best;
            return DNX_OK;
        }
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (!running || (walltime_s > 0 && el > walltime_s)) return DNX_OK;
    }
}

int dnx_generate_planted(uint64_t seed, int n, int m, double p1, double p2, int32_t *lits_out, int8_t *planted_out)
{
    if (n < 3 || m < 0 || !lits_out || p1 < 0 || p2 < 0 || p1 + p2 > 1.0) return DNX_E_ARG;
    generate_planted(seed, n, m, p1, p2, lits_out, planted_out);
    return DNX_OK;
}

int dnx_init_uniform(uint64_t seed, int n, int m, double xs0, double xl0, double *x0)
{
    if (n < 0 || m < 0 || !x0) return DNX_E_ARG;
    init_uniform(seed, n, m, xs0, xl0, x0);
    return DNX_OK;
}

int dnx_write_dimacs(const char *path, int n, int m, const int32_t *clause_ptr, const int32_t *lits,
                     const char *comment)
{
    if (!path || n < 0 || m < 0 || (m > 0 && (!clause_ptr || !lits))) return DNX_E_ARG;
    int rc = write_dimacs(path, n, m, clause_ptr, lits, comment);
    if (rc) set_error(std::string("cannot write ") + path);
    return rc;
}

int dnx_check_assignment(const dnx_problem *p, const int8_t *assign, int *unsat)
{
    if (!p || !assign || !unsat) return DNX_E_ARG;
    *unsat = count_unsat(p->hp, assign);
    return DNX_OK;
}

}  // extern "C"
