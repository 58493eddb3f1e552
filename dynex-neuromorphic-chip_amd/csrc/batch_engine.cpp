// batch_engine.cpp -- the batched engine behind dnx_batch_* (include/dnx.h).
//
// Instances live in slots.  Each admitted instance gets one contiguous range
// of a device arena (first-fit, coalescing free list) holding its clause
// table, transpose, state and contribution buffer, and a descriptor in the
// slot table.  An advance launches the fused step kernel once per execution
// shape over the slots that are still running, largest first; a poll copies
// the per-slot status table back through pinned memory.  Solved instances
// stay in their slot until retired, which frees the arena range at once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <limits>
#include <map>
#include <set>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "api_internal.hpp"
#include "solo.hpp"
#include "batch.hpp"

namespace dnx {

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }
// bytes per half of the admission staging buffers
static constexpr size_t ADMIT_CHUNK = (size_t)16 << 20;

// First-fit allocator over [0, cap) with coalescing on release.
struct Arena {
    size_t cap = 0;
    std::map<size_t, size_t> free_;  // offset -> size
    void reset(size_t c)
    {
        cap = c;
        free_.clear();
        if (c) free_[0] = c;
    }
    bool alloc(size_t sz, size_t &off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= sz) {
                off = it->first;
                size_t rest = it->second - sz;
                free_.erase(it);
                if (rest) free_[off + sz] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t sz)
    {
        auto it = free_.emplace(off, sz).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t free_bytes() const
    {
        size_t s = 0;
        for (auto &kv : free_) s += kv.second;
        return s;
    }
};

struct SlotHost {
    bool used = false;
    int n = 0, m = 0;
    int64_t nnz = 0;
    int mode = 0;
    int nchunk = 1;  // mode 0: LDS chunks of its contributions
    size_t lds = 0;  // dynamic LDS bytes the fused kernel needs for it
    size_t off = 0, size = 0;
    bool general = false;  // clause_ptr in the arena (else 3 literals per clause)
};

// Compulsory bytes of one FP32 step of an instance: literals read once
// (4 per literal), Xs and Xl read and written (16 per clause), the (V, dV)
// pair read and written (8 per variable).
static double step_bytes(const SlotHost &sl) { return 4.0 * (double)sl.nnz + 16.0 * sl.m + 8.0 * sl.n; }

// timing buckets: the three fused-kernel modes, wide mode, gang mode, the
// fused kernel's tail launch beside wide mode, and epoch launches
constexpr int BK_WIDE = BM_COUNT, BK_GANG = BM_COUNT + 1, BK_TAIL = BM_COUNT + 2, BK_EPOCH = BM_COUNT + 3,
              NBUCKET = BM_COUNT + 4;

struct Shape {
    int bs;
    int lds_cap;  // bytes of LDS the shape may use per workgroup
};

}  // namespace dnx

using namespace dnx;

struct dnx_batch {
    dnx_ctx *ctx = nullptr;
    int nslots = 0;
    Arena arena;
    void *arena_base = nullptr;
    DevBuf b_inst, b_dyn, b_work, b_head, b_stamp;
    std::vector<BInst> h_inst;
    std::vector<SlotHost> slots;
    std::vector<int> free_slots;
    BDyn *h_dyn = nullptr;    // pinned: [2][nslots] status snapshots
    BDyn *h_stage = nullptr;  // pinned: [nslots] source of status uploads
    int *h_work = nullptr;  // pinned: [2][BM_COUNT * nslots]
    Shape shape[BM_COUNT] = {{1024, 160 * 1024}, {256, 40 * 1024}, {1024, 0}};
    int cus = 0, lds_max = 0;
    double clock_hz = 1e8;
    int buf = 0;            // alternate pinned buffer
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool ev_pending[2] = {false, false};
    hipStream_t mstream[BM_COUNT] = {nullptr, nullptr, nullptr};
    hipEvent_t fork = nullptr, join[BM_COUNT] = {nullptr, nullptr, nullptr};
    hipEvent_t side_done = nullptr;  // end of a tail launch of the fused kernel on its own stream
    // admission uploads run on their own stream and complete before admit
    // returns, so they never wait behind queued kernels
    hipStream_t ustream = nullptr;
    // optional per-shape launch timing (start/stop events around each launch)
    bool timing = false;
    // timing buckets: the three fused-kernel modes, then wide mode
    hipEvent_t t0[NBUCKET] = {};
    hipEvent_t t1[NBUCKET] = {};
    bool t_pending[NBUCKET] = {};
    double t_ms[NBUCKET] = {};
    long long t_launches[NBUCKET] = {};
    // compulsory bytes of the steps taken by the timed launches of
    // dnx_batch_run, per bucket (step_bytes per instance step)
    double t_alg[NBUCKET] = {};
    // the tail (gang mode, else wide mode): entered when the running
    // instances hold at most max(wide_clauses, wide_frac x the clauses
    // running when dnx_batch_run starts) clauses in total; wide_chunk steps
    // between status syncs in wide mode
    long long wide_clauses = 0;
    double wide_frac = 0.0125;
    int wide_chunk = 300;
    // serving: steps per chunk of the wide lane (status syncs and active-set
    // choices happen between chunks)
    int serve_chunk = 30;
    // serving: requests with at least this many clauses run in the wide lane
    long long serve_big = 60000;
    // serving: clause budget of the wide lane's active set per chunk
    long long serve_active_clauses = 6000000;
    // serving: wide-lane members below this many steps are scheduled by
    // least attained service, the others by least expected work (0: all by
    // least attained service)
    long long serve_las_steps = 1024;
    // serving: 3-SAT requests with at least this many clauses run in the
    // sliced lane, at most serve_sliced_active resident at a time (0: no
    // sliced lane)
    long long serve_sliced = 60000;
    int serve_sliced_active = 1024;
    // requests the last serving call retired in each lane: narrow, wide,
    // sliced, and those that could never run (larger than the arena)
    long long served_by[4] = {0, 0, 0, 0};
    // portfolio: dnx_batch_run returns at the first status sync at which
    // any slot is satisfied
    bool group_stop = false;
    // tail: single-chunk mode-0 instances run in the fused kernel beside
    // wide mode (batch_split_tail)
    bool tail_split = true;
    // work-list order of the fused kernel: instances of at least order_big
    // clauses first, largest first, then the others smallest first (0: all
    // largest first)
    long long order_big = 0;
    // dnx_batch_run: the fused launch keeps each instance on its workgroup
    // until it ends or the running clauses fall to the wide-mode threshold
    // (BFlow), instead of relaunching every steps_per_launch steps
    bool flow = true;
    DevBuf b_flow;
    DevBuf b_wctl, b_wlist, b_bmap, b_vmap;
    // gang mode: taken for the tail when the running instances need at
    // most gang_blocks workgroups in total (0 disables it)
    int gang_blocks = 256;
    int gang_cpw = 384;       // target clauses per gang workgroup
    int gang_steps = 512;     // steps per cooperative launch
    int gang_cap = 0;         // co-resident gang workgroups (0: not known yet)
    unsigned long long gang_epoch = 0;
    int gang_arrive = 0;      // arrivals counted on the device so far
    DevBuf b_gmap, b_gwords, b_gerr;
    // elastic tail: once the fused kernel has CUs to spare (fewer running
    // instances than CUs), epoch launches give them to the largest running
    // instances as gangs while the others keep one workgroup each
    // (batch_epoch_run).  epoch_min_clauses: the smallest instance given a
    // gang (0 disables epochs); epoch_steps: gang steps per epoch launch.
    // (4,096 mixed instances: 273-278 ms per batch, against 372 ms with the
    // threshold tail; 12,000 clauses with m / 1536 workgroups: 278-290 ms.)
    long long epoch_min_clauses = 6000;
    int epoch_steps = 1024;
    // a gang forms only with at most this many clauses per workgroup (a
    // regrouped gang's step at m = 20,000: 8.5-10 / 7.3-7.8 / 6.1 us in 10 /
    // 16 / 32 workgroups, against ~17 us in one; at m = 6,547: 5.4 / 5.0 us
    // in 8 / 16, against 7.7 in one)
    int epoch_cpw = 768;
    // epoch gangs of 3-SAT instances regroup their contributions by the
    // workgroup that folds them (gang_rx in batch.hip)
    bool gang_regroup = true;
    // an epoch's upload (batch_epoch_run): pinned on the host, and its device copy
    unsigned char *h_ep = nullptr;
    DevBuf b_ep;
    DevBuf b_gscr;   // regrouped gangs' scratch
    // admission staging: two halves of adm_chunk bytes, pinned on the host
    // (h_adm) and on the device (b_adm); an admission waits for its uploads
    // before it returns
    unsigned char *h_adm = nullptr;
    size_t adm_chunk = 0;
    DevBuf b_adm;
    hipEvent_t adm_ev[2] = {nullptr, nullptr};
    bool adm_pending[2] = {false, false};
    ~dnx_batch()
    {
        if (ustream) (void)hipStreamSynchronize(ustream);
        if (arena_base) (void)hipFree(arena_base);
        if (h_adm) (void)hipHostFree(h_adm);
        if (h_ep) (void)hipHostFree(h_ep);
        for (auto &e : adm_ev)
            if (e) (void)hipEventDestroy(e);
        if (h_dyn) (void)hipHostFree(h_dyn);
        if (h_stage) (void)hipHostFree(h_stage);
        if (h_work) (void)hipHostFree(h_work);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        for (int i = 0; i < BM_COUNT; i++) {
            if (mstream[i]) (void)hipStreamDestroy(mstream[i]);
            if (join[i]) (void)hipEventDestroy(join[i]);
        }
        for (int i = 0; i < NBUCKET; i++) {
            if (t0[i]) (void)hipEventDestroy(t0[i]);
            if (t1[i]) (void)hipEventDestroy(t1[i]);
        }
        if (fork) (void)hipEventDestroy(fork);
        if (side_done) (void)hipEventDestroy(side_done);
        if (ustream) (void)hipStreamDestroy(ustream);
    }
};

namespace dnx {

int dnx_batch_admit_host(dnx_batch *b, const HostProblem &h, const dnx_params *prm, const double *x0,
                         int64_t step_cap, int stop_on_solve, int *slot_out);
int admit_many_host(dnx_batch *b, int count, const HostProblem *const *hs, const dnx_params *prms,
                    const double *const *x0s, int64_t step_cap, int stop_on_solve, int *slots_out);
// Mode 0 holds V, one more float per variable (lds_var_bytes) and the
// contributions in LDS.  When the contributions do not fit beside them, the
// clauses are cut into at most KC_MAX chunks of consecutive clauses whose
// contributions fit; a step then evaluates and folds one chunk at a time,
// keeping each variable's partial fold in that float.
struct Mode0Plan {
    int nchunk = 0;
    size_t lds = 0;                 // dynamic LDS bytes the instance needs
    std::vector<int32_t> bound;     // nchunk + 1 clause indices
};

static bool plan_mode0(const dnx_batch *b, const HostProblem &h, Mode0Plan &P)
{
    const size_t cap = (size_t)b->shape[BM_LDS_ALL].lds_cap;
    const size_t pv = lds_var_bytes(h.n);
    const int64_t nnz = h.nnz();
    P.bound.clear();
    // packed literal words hold a 15-bit variable and a 16-bit LDS slot
    if (h.n > (1 << 15) || cap / 4 > (1 << 16)) return false;
    if (pv + (size_t)nnz * 4 <= cap) {
        P.nchunk = 1;
        P.lds = pv + (size_t)nnz * 4;
        P.bound = {0, h.m};
        return true;
    }
    // with several chunks each variable's partial fold lives in the float
    // beside its V (lds_var_bytes)
    if (pv >= cap) return false;
    const int64_t slots = (int64_t)((cap - pv) / 4);
    if (slots < DNX_MAX_LITS || slots * KC_MAX < nnz) return false;
    // 3-SAT chunks start at multiples of four clauses, so the steady clause
    // pass's groups of four never straddle two chunks (a variable's fold runs
    // on across chunks in slot order, so the cuts do not change any sum)
    P.bound.push_back(0);
    int64_t used = 0, most = 0;
    for (int c = 0; c < h.m; c++) {
        const int k = h.cp[(size_t)c + 1] - h.cp[(size_t)c];
        if (used + k > slots) {
            int cut = c;
            if (h.uniform3 && cut % 4 != 0 && cut - cut % 4 > P.bound.back()) cut -= cut % 4;
            const int64_t moved = h.cp[(size_t)c] - h.cp[(size_t)cut];  // slots of the clauses moved into the new chunk
            most = std::max(most, used - moved);
            P.bound.push_back(cut);
            used = moved;
        }
        used += k;
    }
    P.bound.push_back(h.m);
    most = std::max(most, used);
    P.nchunk = (int)P.bound.size() - 1;
    if (P.nchunk > KC_MAX) return false;
    P.lds = pv + (size_t)most * 4;
    return true;
}

static int choose_mode(const dnx_batch *b, const HostProblem &h, Mode0Plan &P)
{
    if (plan_mode0(b, h, P)) return BM_LDS_ALL;
    P.nchunk = 1;
    P.bound = {0, h.m};
    const size_t pv = lds_var_bytes(h.n);
    P.lds = pv;
    if (pv <= (size_t)b->shape[BM_LDS_V].lds_cap) return BM_LDS_V;
    P.lds = 0;
    return BM_GLOBAL;
}

// Byte layout of one instance inside its arena range.
struct Layout {
    size_t lits, cp, lpos, pk, gpos, eidx, row, unit, xs, xl, vd, e, cbound, cbase, cst, total;
    size_t upload;  // bytes [0, upload) are built on the host
};
static Layout layout_of(int n, int m, int64_t nnz, bool general, bool has_unit, int nchunk)
{
    Layout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t at = o;
        o += al256(bytes);
        return at;
    };
    // tables the host builds (uploaded at admission) ...
    L.lits = take((size_t)nnz * 4);
    L.cp = general ? take(((size_t)m + 1) * 4) : 0;
    L.lpos = take((size_t)nnz * 4);
    // transpose slots of the literals; with one chunk (or outside mode 0)
    // lpos holds them already
    L.gpos = nchunk > 1 ? take((size_t)nnz * 4) : 0;
    L.row = take(((size_t)n + 1) * 4);
    L.unit = has_unit ? take((size_t)n) : 0;
    L.xs = take((size_t)m * 4);
    L.xl = take((size_t)m * 4);
    L.vd = take((size_t)n * sizeof(BF2));
    L.cbound = take(((size_t)nchunk + 1) * 4);
    L.cbase = nchunk > 1 ? take((size_t)nchunk * ((size_t)n + 1) * 4) : 0;
    L.upload = o;
    // ... and the ones the device derives or clears at admission.  Both
    // transposes and the contribution array exist for every instance, so
    // any instance can also run in wide mode.
    L.pk = general ? 0 : take((size_t)nnz * 4);  // filled for mode-0 3-SAT instances
    L.eidx = take((size_t)nnz * 4);
    L.e = take((size_t)nnz * 4);
    L.cst = take((size_t)m * 4);
    L.total = o ? o : 256;
    return L;
}

// Adds the elapsed time of the last timed launch of each shape (the caller
// has waited for the work).
static void batch_collect_timing(dnx_batch *b)
{
    for (int md = 0; md < NBUCKET; md++) {
        if (!b->t_pending[md]) continue;
        // a launch on another stream (the tail's fused-kernel launch) may
        // still run: its timing stays pending until it has ended
        if (hipEventQuery(b->t1[md]) == hipErrorNotReady) continue;
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, b->t0[md], b->t1[md]) == hipSuccess) {
            b->t_ms[md] += ms;
            b->t_launches[md] += 1;
        }
        b->t_pending[md] = false;
    }
}

static int batch_sync(dnx_batch *b)
{
    hipError_t e = hipStreamSynchronize(b->ctx->stream);
    return e == hipSuccess ? DNX_OK : hip_fail(e, "batch sync");
}

// Enqueue one advance of every running slot by up to ksteps steps.  The work
// lists come from `known` (the last polled status); slots that ended since
// are skipped by the kernel itself.
static int batch_enqueue_slots(dnx_batch *b, int ksteps, const std::vector<int> &slots, BFlow *flow = nullptr);

static int batch_enqueue(dnx_batch *b, int ksteps, const BDyn *known, BFlow *flow = nullptr)
{
    std::vector<int> run;
    for (int s = 0; s < b->nslots; s++)
        if (b->slots[(size_t)s].used && known[s].status == DNX_RUNNING) run.push_back(s);
    return batch_enqueue_slots(b, ksteps, run, flow);
}

// Enqueue one advance of the listed slots (fused kernel, one launch per
// execution shape, each on its own stream, joined back into the context
// stream).  With a flow (mode 0 only), ksteps is ignored and the launch runs
// until the flow stops.
static int batch_enqueue_slots(dnx_batch *b, int ksteps, const std::vector<int> &slots, BFlow *flow)
{
    hipStream_t st = b->ctx->stream;
    const int S = b->nslots;
    int *hw = b->h_work + (size_t)b->buf * BM_COUNT * S;
    int cnt[BM_COUNT] = {0, 0, 0};
    std::vector<std::pair<int, int>> order;
    order.reserve(slots.size());
    for (int s : slots) order.push_back({b->slots[(size_t)s].m, s});
    const long long big = b->order_big;
    std::sort(order.begin(), order.end(), [big](const std::pair<int, int> &x, const std::pair<int, int> &y) {
        const bool bx = x.first >= big, by = y.first >= big;
        if (bx != by) return bx;
        if (x.first != y.first) return bx ? x.first > y.first : x.first < y.first;
        return x.second < y.second;
    });
    for (auto &o : order) {
        int md = b->slots[(size_t)o.second].mode;
        hw[(size_t)md * S + cnt[md]++] = o.second;
    }
    if (order.empty()) return DNX_OK;
    // a previous advance that used this pinned work buffer must have consumed it
    int *dw = (int *)b->b_work.ptr + (size_t)b->buf * BM_COUNT * S;
    int *dh = (int *)b->b_head.ptr + b->buf * BM_COUNT;
    hipError_t e;
    for (int md = 0; md < BM_COUNT; md++) {
        if (!cnt[md]) continue;
        if ((e = hipMemcpyAsync(dw + (size_t)md * S, hw + (size_t)md * S, sizeof(int) * cnt[md],
                                hipMemcpyHostToDevice, st)) != hipSuccess)
            return hip_fail(e, "work list upload");
    }
    launch_batch_zero(dh, BM_COUNT, st);
    if ((e = hipEventRecord(b->fork, st)) != hipSuccess) return hip_fail(e, "fork");
    for (int md = 0; md < BM_COUNT; md++) {
        if (!cnt[md]) continue;
        hipStream_t ms = b->mstream[md];
        if ((e = hipStreamWaitEvent(ms, b->fork, 0)) != hipSuccess) return hip_fail(e, "fork wait");
        const Shape &sh = b->shape[md];
        int lds = 0;
        if (md != BM_GLOBAL) {
            // the largest LDS need among this shape's instances
            for (int i = 0; i < cnt[md]; i++) lds = std::max(lds, (int)b->slots[(size_t)hw[(size_t)md * S + i]].lds);
        }
        // When few instances remain, each gets a larger workgroup: the aim is
        // about four waves per SIMD across the chip, so lone long-running
        // instances (the tail of a batch) are not left latency-bound on a
        // small workgroup.
        int bs = sh.bs;
        const long long target_waves = 16LL * b->cus;
        while (bs < 1024 && (long long)order.size() * (bs / 64) < target_waves) bs *= 2;
        int per_cu = batch_occupancy(md, bs, lds);
        if (per_cu < 1) per_cu = 1;
        int grid = std::min(cnt[md], b->cus * per_cu);
        const bool timed = b->timing && !b->t_pending[md];
        if (timed && (e = hipEventRecord(b->t0[md], ms)) != hipSuccess) return hip_fail(e, "timing");
        BFlow *f = md == BM_LDS_ALL ? flow : nullptr;
        launch_batch(md, bs, lds, grid, (const BInst *)b->b_inst.ptr, (BDyn *)b->b_dyn.ptr, dw + (size_t)md * S,
                     cnt[md], dh + md, f ? INT_MAX : ksteps, f, ms);
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "batch launch");
        if (timed) {
            if ((e = hipEventRecord(b->t1[md], ms)) != hipSuccess) return hip_fail(e, "timing");
            b->t_pending[md] = true;
        }
        if ((e = hipEventRecord(b->join[md], ms)) != hipSuccess) return hip_fail(e, "join");
        if ((e = hipStreamWaitEvent(st, b->join[md], 0)) != hipSuccess) return hip_fail(e, "join wait");
    }
    return DNX_OK;
}

}  // namespace dnx

extern "C" {

int dnx_batch_open(dnx_ctx *ctx, int prec, int slots, int64_t arena_bytes, dnx_batch **out)
{
    if (!ctx || !out || slots <= 0) return DNX_E_ARG;
    if (prec != DNX_FP32) {
        set_error("the batched engine runs in DNX_FP32");
        return DNX_E_ARG;
    }
    (void)hipSetDevice(ctx->device);
    hipError_t e;
    size_t freeb = 0, totalb = 0;
    if ((e = hipMemGetInfo(&freeb, &totalb)) != hipSuccess) return hip_fail(e, "hipMemGetInfo");
    if (arena_bytes <= 0) arena_bytes = (int64_t)(0.8 * (double)totalb);
    if ((size_t)arena_bytes > freeb) arena_bytes = (int64_t)(freeb - (freeb >> 4));
    dnx_batch *b = new dnx_batch();
    b->ctx = ctx;
    b->nslots = slots;
    int rc;
    if ((rc = b->b_inst.alloc(sizeof(BInst) * (size_t)slots)) || (rc = b->b_dyn.alloc(sizeof(BDyn) * (size_t)slots)) ||
        (rc = b->b_work.alloc(sizeof(int) * 2 * BM_COUNT * (size_t)slots)) ||
        (rc = b->b_head.alloc(sizeof(int) * 2 * BM_COUNT)) || (rc = b->b_stamp.alloc(sizeof(unsigned long long)))) {
        delete b;
        return rc;
    }
    if ((e = hipMalloc(&b->arena_base, (size_t)arena_bytes)) != hipSuccess) {
        b->arena_base = nullptr;
        delete b;
        set_error("arena hipMalloc(" + std::to_string(arena_bytes) + ") failed");
        return DNX_E_NOMEM;
    }
    b->arena.reset((size_t)arena_bytes);
    if ((e = hipHostMalloc((void **)&b->h_dyn, sizeof(BDyn) * 2 * (size_t)slots, hipHostMallocDefault)) !=
            hipSuccess ||
        (e = hipHostMalloc((void **)&b->h_stage, sizeof(BDyn) * (size_t)slots, hipHostMallocDefault)) != hipSuccess ||
        (e = hipHostMalloc((void **)&b->h_work, sizeof(int) * 2 * BM_COUNT * (size_t)slots, hipHostMallocDefault)) !=
            hipSuccess) {
        delete b;
        return hip_fail(e, "hipHostMalloc");
    }
    for (int i = 0; i < 2; i++)
        if ((e = hipEventCreateWithFlags(&b->ev[i], hipEventDisableTiming)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&b->adm_ev[i], hipEventDisableTiming)) != hipSuccess) {
            delete b;
            return hip_fail(e, "event");
        }
    // admission staging (grown by an admission that needs more)
    b->adm_chunk = ADMIT_CHUNK;
    if ((e = hipHostMalloc((void **)&b->h_adm, 2 * b->adm_chunk, hipHostMallocDefault)) != hipSuccess) {
        b->h_adm = nullptr;
        delete b;
        return hip_fail(e, "admission staging");
    }
    if ((rc = b->b_adm.alloc(2 * b->adm_chunk))) {
        delete b;
        return rc;
    }
    if ((e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->side_done, hipEventDisableTiming)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&b->ustream, hipStreamNonBlocking)) != hipSuccess) {
        delete b;
        return hip_fail(e, "event");
    }
    for (int i = 0; i < BM_COUNT; i++)
        if ((e = hipStreamCreateWithFlags(&b->mstream[i], hipStreamNonBlocking)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming)) != hipSuccess) {
            delete b;
            return hip_fail(e, "stream");
        }
    for (int i = 0; i < NBUCKET; i++)
        if ((e = hipEventCreate(&b->t0[i])) != hipSuccess || (e = hipEventCreate(&b->t1[i])) != hipSuccess) {
            delete b;
            return hip_fail(e, "event");
        }
    if ((rc = b->b_wctl.alloc(sizeof(Ctl) * (size_t)slots))) {
        delete b;
        return rc;
    }
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, ctx->device)) != hipSuccess) {
        delete b;
        return hip_fail(e, "hipGetDeviceProperties");
    }
    b->cus = prop.multiProcessorCount;
    int lds_max = 0;
    (void)hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device);
    // the fused kernel also holds a few hundred bytes of static LDS (work
    // ticket, level words, reduction scratch) beside the dynamic part
    b->lds_max = (lds_max > 0 ? lds_max : 65536) - 1024;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device) == hipSuccess && khz > 0)
        b->clock_hz = (double)khz * 1e3;
    for (int md = 0; md < BM_COUNT; md++)
        if (b->shape[md].lds_cap > b->lds_max) b->shape[md].lds_cap = b->lds_max;
    b->h_inst.assign((size_t)slots, BInst{});
    b->slots.assign((size_t)slots, SlotHost{});
    for (int s = slots - 1; s >= 0; s--) b->free_slots.push_back(s);
    std::vector<BDyn> init((size_t)slots);
    for (auto &d : init) {
        memset(&d, 0, sizeof(d));
        d.status = -1;
        d.plevel = -1;
    }
    memcpy(b->h_dyn, init.data(), sizeof(BDyn) * (size_t)slots);
    memcpy(b->h_dyn + slots, init.data(), sizeof(BDyn) * (size_t)slots);
    if ((e = hipMemcpy(b->b_dyn.ptr, init.data(), sizeof(BDyn) * (size_t)slots, hipMemcpyHostToDevice)) !=
        hipSuccess) {
        delete b;
        return hip_fail(e, "dyn init");
    }
    *out = b;
    return DNX_OK;
}

int dnx_batch_close(dnx_batch *b)
{
    if (!b) return DNX_E_ARG;
    (void)hipSetDevice(b->ctx->device);
    (void)hipDeviceSynchronize();
    delete b;
    return DNX_OK;
}

int dnx_batch_set_shape(dnx_batch *b, int mode, int block, int lds_cap)
{
    if (!b || mode < 0 || mode >= BM_COUNT || (block != 128 && block != 256 && block != 512 && block != 1024))
        return DNX_E_ARG;
    for (auto &sl : b->slots)
        if (sl.used) {
            set_error("shapes can only change while the batch is empty");
            return DNX_E_STATE;
        }
    b->shape[mode].bs = block;
    if (mode != BM_GLOBAL) b->shape[mode].lds_cap = std::max(0, std::min(lds_cap, b->lds_max));
    return DNX_OK;
}

int dnx_batch_admit(dnx_batch *b, const dnx_problem *p, const dnx_params *prm, const double *x0, int64_t step_cap,
                    int stop_on_solve, int *slot_out)
{
    if (!b || !p || !prm || !x0 || !slot_out || step_cap < 0) return DNX_E_ARG;
    return dnx_batch_admit_host(b, p->hp, prm, x0, step_cap, stop_on_solve, slot_out);
}

int dnx_batch_admit_many(dnx_batch *b, int count, const dnx_problem *const *probs, const dnx_params *prms,
                         const double *const *x0s, int64_t step_cap, int stop_on_solve, int *slots_out)
{
    if (!b || count < 0 || (count && (!probs || !x0s || !slots_out)) || step_cap < 0) return DNX_E_ARG;
    std::vector<const HostProblem *> hs((size_t)count);
    for (int i = 0; i < count; i++) {
        if (!probs[i] || !x0s[i]) return DNX_E_ARG;
        hs[(size_t)i] = &probs[i]->hp;
    }
    std::vector<dnx_params> pv;
    if (!prms) {
        pv.resize((size_t)count);
        for (int i = 0; i < count; i++) dnx_params_default(hs[(size_t)i]->m, 10000, &pv[(size_t)i]);
    }
    return admit_many_host(b, count, hs.data(), prms ? prms : pv.data(), x0s, step_cap, stop_on_solve, slots_out);
}

}  // extern "C"

namespace {

// Runs fn(i) for i in [0, n) on the calling thread and a fixed set of helper
// threads; run() returns when every index is done.
class ForkJoin {
  public:
    explicit ForkJoin(int helpers)
    {
        for (int i = 0; i < helpers; i++) th_.emplace_back([this] { loop(); });
    }
    ~ForkJoin()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            quit_ = true;
            gen_++;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    void run(int n, const std::function<void(int)> &fn)
    {
        if (th_.empty() || n <= 1) {
            for (int i = 0; i < n; i++) fn(i);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            fn_ = &fn;
            n_ = n;
            next_.store(0);
            active_ = (int)th_.size();
            gen_++;
        }
        cv_.notify_all();
        work();
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return active_ == 0; });
        fn_ = nullptr;
    }

  private:
    void work()
    {
        for (int i; (i = next_.fetch_add(1)) < n_;) (*fn_)(i);
    }
    void loop()
    {
        unsigned long long seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                if (quit_) return;
            }
            work();
            std::lock_guard<std::mutex> lk(mu_);
            if (--active_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int)> *fn_ = nullptr;
    int n_ = 0, active_ = 0;
    std::atomic<int> next_{0};
    unsigned long long gen_ = 0;
    bool quit_ = false;
};

struct AdmitItem {
    const HostProblem *h = nullptr;
    const double *x0 = nullptr;
    dnx_params prm{};
    int mode = 0, slot = -1;
    Mode0Plan plan;
    Layout L{};
    BInst inst{};
    size_t off = 0;    // arena offset
    size_t soff = 0;   // offset of the staged prefix in its staging buffer
    int chunk = 0;
};

// Writes the part of an instance's arena range that only the host builds
// (everything below L.upload) to `stage`, with every region's alignment
// padding zeroed.
void stage_prefix(const AdmitItem &a, unsigned char *stage)
{
    const HostProblem &h = *a.h;
    const Layout &L = a.L;
    const int n = h.n, m = h.m, nchunk = a.plan.nchunk;
    const int64_t nnz = h.nnz();
    auto tail = [&](size_t at, size_t bytes) {
        const size_t end = al256(bytes);
        if (end > bytes) memset(stage + at + bytes, 0, end - bytes);
    };
    auto put = [&](size_t at, const void *src, size_t bytes) {
        if (bytes) memcpy(stage + at, src, bytes);
        tail(at, bytes);
    };
    put(L.lits, h.lits.data(), (size_t)nnz * 4);
    if (!h.uniform3) put(L.cp, h.cp.data(), ((size_t)m + 1) * 4);
    if (nchunk == 1) {
        put(L.lpos, h.lpos.data(), (size_t)nnz * 4);
    } else {
        // chunk-local LDS slots: within chunk k, variable-major, each
        // variable's slots in clause order
        int32_t *lp = (int32_t *)(stage + L.lpos);
        int32_t *cb = (int32_t *)(stage + L.cbase);
        std::vector<int32_t> run((size_t)n + 1);
        for (int k = 0; k < nchunk; k++) {
            int32_t *base = cb + (size_t)k * ((size_t)n + 1);
            std::fill(base, base + n + 1, 0);
            const int q0 = h.cp[(size_t)a.plan.bound[(size_t)k]], q1 = h.cp[(size_t)a.plan.bound[(size_t)k + 1]];
            for (int q = q0; q < q1; q++) base[std::abs(h.lits[(size_t)q]) - 1 + 1]++;
            for (int v = 0; v < n; v++) base[v + 1] += base[v];
            std::copy(base, base + n, run.begin());
            for (int q = q0; q < q1; q++) lp[q] = run[(size_t)std::abs(h.lits[(size_t)q]) - 1]++;
        }
        tail(L.lpos, (size_t)nnz * 4);
        tail(L.cbase, (size_t)nchunk * ((size_t)n + 1) * 4);
        put(L.gpos, h.lpos.data(), (size_t)nnz * 4);
    }
    put(L.row, h.row_ptr.data(), ((size_t)n + 1) * 4);
    if (h.has_unit) put(L.unit, h.unit.data(), (size_t)n);
    float *xs = (float *)(stage + L.xs), *xl = (float *)(stage + L.xl);
    const double *x0 = a.x0;
    for (int c = 0; c < m; c++) {
        xs[c] = (float)x0[(size_t)n + c];
        xl[c] = (float)x0[(size_t)n + m + c];
    }
    tail(L.xs, (size_t)m * 4);
    tail(L.xl, (size_t)m * 4);
    BF2 *vd = (BF2 *)(stage + L.vd);
    for (int v = 0; v < n; v++) {  // unit pins at admission (dynex.cc:2694-2699)
        vd[v].v = h.unit[(size_t)v] != 0 ? (float)h.unit[(size_t)v] : (float)x0[v];
        vd[v].d = 0.0f;
    }
    tail(L.vd, (size_t)n * sizeof(BF2));
    put(L.cbound, a.plan.bound.data(), ((size_t)nchunk + 1) * 4);
}

}  // namespace

namespace dnx {

int dnx_batch_admit_host(dnx_batch *b, const HostProblem &h, const dnx_params *prm, const double *x0,
                         int64_t step_cap, int stop_on_solve, int *slot_out)
{
    const HostProblem *hp = &h;
    return admit_many_host(b, 1, &hp, prm, &x0, step_cap, stop_on_solve, slot_out);
}

// Admission of `count` instances: plans and arena ranges first (all or
// none), then the host-built tables are staged by worker threads into
// pinned chunks, each chunk uploaded with one copy and placed by one
// kernel (which also derives the tables the device can build), chunks
// alternating between two buffers so staging overlaps the uploads; one wait
// at the end.
int admit_many_host(dnx_batch *b, int count, const HostProblem *const *hs, const dnx_params *prms,
                    const double *const *x0s, int64_t step_cap, int stop_on_solve, int *slots_out)
{
    if (count < 0) return DNX_E_ARG;
    if (count == 0) return DNX_OK;
    if ((size_t)count > b->free_slots.size()) {
        set_error(count == 1 ? "no free slot" : "fewer free slots than instances");
        return DNX_E_BUDGET;
    }
    (void)hipSetDevice(b->ctx->device);
    std::vector<AdmitItem> items((size_t)count);
    const size_t fs = b->free_slots.size();
    auto release_upto = [&](int k) {
        for (int i = 0; i < k; i++) b->arena.release(items[(size_t)i].off, items[(size_t)i].L.total);
    };
    const size_t rec_bytes = sizeof(AdmitRec);
    size_t chunk_need = 0;
    int helpers = 0;
    if (count >= 8) {
        const unsigned hc = std::thread::hardware_concurrency();
        helpers = (int)std::min<unsigned>(hc ? hc : 1, 16) - 1;
    }
    ForkJoin pool(helpers);
    pool.run(count, [&](int i) {
        AdmitItem &a = items[(size_t)i];
        a.h = hs[i];
        a.x0 = x0s[i];
        a.prm = prms[i];
        a.mode = choose_mode(b, *a.h, a.plan);
        a.L = layout_of(a.h->n, a.h->m, a.h->nnz(), !a.h->uniform3, a.h->has_unit, a.plan.nchunk);
    });
    for (int i = 0; i < count; i++) {
        AdmitItem &a = items[(size_t)i];
        if (!b->arena.alloc(a.L.total, a.off)) {
            release_upto(i);
            set_error("instance needs " + std::to_string(a.L.total) + " bytes; arena has " +
                      std::to_string(b->arena.free_bytes()) + " free in fragments");
            return DNX_E_BUDGET;
        }
        a.slot = b->free_slots[fs - 1 - (size_t)i];
        chunk_need = std::max(chunk_need, al256(rec_bytes) + a.L.upload);
    }
    // staging buffers: two halves of `ch` bytes, pinned on the host and on
    // the device; kept by the batch for later admissions
    const size_t need = std::max(chunk_need, ADMIT_CHUNK);
    hipError_t e;
    if (b->adm_chunk < need) {
        const size_t ch = need;
        if (b->h_adm) (void)hipHostFree(b->h_adm);
        b->h_adm = nullptr;
        b->adm_chunk = 0;
        int rc;
        if ((e = hipHostMalloc((void **)&b->h_adm, 2 * ch, hipHostMallocDefault)) != hipSuccess) {
            release_upto(count);
            return hip_fail(e, "admission staging");
        }
        if ((rc = b->b_adm.alloc(2 * ch))) {
            release_upto(count);
            return rc;
        }
        b->adm_chunk = ch;
    }
    const size_t ch = b->adm_chunk;
    // cut the list into chunks: records first, then each prefix
    std::vector<int> cstart{0};
    {
        int k = 0;
        size_t used = 0;
        for (int i = 0; i < count; i++) {
            const size_t up = items[(size_t)i].L.upload;
            if (i > cstart.back() && al256((size_t)(i - cstart.back() + 1) * rec_bytes) + used + up > ch) {
                cstart.push_back(i);
                k++;
                used = 0;
            }
            items[(size_t)i].chunk = k;
            items[(size_t)i].soff = used;  // relative to the end of the records
            used += up;
        }
        cstart.push_back(count);
    }
    unsigned char *arena = (unsigned char *)b->arena_base;
    const int nchunks = (int)cstart.size() - 1;
    int rc = DNX_OK;
    for (int k = 0; k < nchunks && rc == DNX_OK; k++) {
        const int hf = k & 1, i0 = cstart[(size_t)k], i1 = cstart[(size_t)k + 1];
        if (b->adm_pending[hf]) {
            if ((e = hipEventSynchronize(b->adm_ev[hf])) != hipSuccess) {
                rc = hip_fail(e, "admission wait");
                break;
            }
            b->adm_pending[hf] = false;
        }
        unsigned char *hbuf = b->h_adm + (size_t)hf * ch;
        unsigned char *dbuf = (unsigned char *)b->b_adm.ptr + (size_t)hf * ch;
        const size_t hdr = al256((size_t)(i1 - i0) * rec_bytes);
        AdmitRec *recs = (AdmitRec *)hbuf;
        pool.run(i1 - i0, [&](int j) {
            const AdmitItem &a = items[(size_t)(i0 + j)];
            stage_prefix(a, hbuf + hdr + a.soff);
        });
        size_t bytes = hdr;
        for (int i = i0; i < i1; i++) {
            AdmitItem &a = items[(size_t)i];
            const HostProblem &h = *a.h;
            const Layout &L = a.L;
            const bool general = !h.uniform3;
            const int64_t nnz = h.nnz();
            unsigned char *base = arena + a.off;
            AdmitRec R;
            memset(&R, 0, sizeof(R));
            R.src = hdr + a.soff;
            R.dst = a.off;
            R.prefix = L.upload;
            R.nnz = nnz;
            R.slot = a.slot;
            R.fill_pk = a.mode == BM_LDS_ALL && !general;
            R.lits = L.lits;
            R.pos = a.plan.nchunk > 1 ? L.gpos : L.lpos;
            R.lpos = L.lpos;
            R.pk = general ? 0 : L.pk;
            R.eidx = L.eidx;
            R.e = L.e;
            R.cst = L.cst;
            R.pk_words = general ? 0 : al256((size_t)nnz * 4) / 4;
            R.eidx_words = al256((size_t)nnz * 4) / 4;
            R.e_words = al256((size_t)nnz * 4) / 4;
            R.cst_words = al256((size_t)h.m * 4) / 4;
            BInst &I = R.inst;
            I.n = h.n;
            I.m = h.m;
            I.mode = a.mode;
            I.flags = (general ? BF_GENERAL : 0) | (h.has_unit ? BF_UNIT : 0);
            I.lits = (const int32_t *)(base + L.lits);
            I.cp = general ? (const int32_t *)(base + L.cp) : nullptr;
            I.lpos = (const int32_t *)(base + L.lpos);
            I.pk = R.fill_pk ? (const uint32_t *)(base + L.pk) : nullptr;
            I.gpos = a.plan.nchunk > 1 ? (const int32_t *)(base + L.gpos) : I.lpos;
            I.eidx = (const int32_t *)(base + L.eidx);
            I.row_ptr = (const int32_t *)(base + L.row);
            I.unit = h.has_unit ? (const int8_t *)(base + L.unit) : nullptr;
            I.xs = (float *)(base + L.xs);
            I.xl = (float *)(base + L.xl);
            I.vd = (BF2 *)(base + L.vd);
            I.e = (float *)(base + L.e);
            I.cbound = (const int32_t *)(base + L.cbound);
            I.cbase = a.plan.nchunk > 1 ? (const int32_t *)(base + L.cbase) : I.row_ptr;
            I.cst = (float *)(base + L.cst);
            I.nchunk = a.plan.nchunk;
            I.alpha = (float)a.prm.alpha;
            I.beta = (float)a.prm.beta;
            I.gamma = (float)a.prm.gamma;
            I.delta = (float)a.prm.delta;
            I.eps = (float)a.prm.epsilon;
            I.zeta = (float)a.prm.zeta;
            I.xlmax = (float)(double)a.prm.xl_max;
            BDyn &D = R.dyn;
            D.status = DNX_RUNNING;
            D.plevel = -1;
            D.stop_on_solve = stop_on_solve ? 1 : 0;
            D.cap = step_cap;
            recs[i - i0] = R;
            a.inst = I;
            bytes = std::max(bytes, (size_t)R.src + L.upload);
        }
        if ((e = hipMemcpyAsync(dbuf, hbuf, bytes, hipMemcpyHostToDevice, b->ustream)) != hipSuccess) {
            rc = hip_fail(e, "admit upload");
            break;
        }
        launch_admit_place((const AdmitRec *)dbuf, i1 - i0, dbuf, arena, (BInst *)b->b_inst.ptr,
                           (BDyn *)b->b_dyn.ptr, b->ustream);
        if ((e = hipGetLastError()) != hipSuccess || (e = hipEventRecord(b->adm_ev[hf], b->ustream)) != hipSuccess) {
            rc = hip_fail(e, "admit placement");
            break;
        }
        b->adm_pending[hf] = true;
    }
    // the staging buffers are reused by the next admission, and the
    // instances must be in place before anything launches: wait here
    if ((e = hipStreamSynchronize(b->ustream)) != hipSuccess && rc == DNX_OK) rc = hip_fail(e, "admit sync");
    b->adm_pending[0] = b->adm_pending[1] = false;
    if (rc != DNX_OK) {
        release_upto(count);
        return rc;
    }
    for (int i = 0; i < count; i++) {
        const AdmitItem &a = items[(size_t)i];
        const int slot = a.slot;
        b->h_inst[(size_t)slot] = a.inst;
        BDyn D;
        memset(&D, 0, sizeof(D));
        D.status = DNX_RUNNING;
        D.plevel = -1;
        D.stop_on_solve = stop_on_solve ? 1 : 0;
        D.cap = step_cap;
        b->h_dyn[slot] = D;
        b->h_dyn[(size_t)b->nslots + slot] = D;
        SlotHost &sl = b->slots[(size_t)slot];
        sl.used = true;
        sl.n = a.h->n;
        sl.m = a.h->m;
        sl.nnz = a.h->nnz();
        sl.mode = a.mode;
        sl.nchunk = a.plan.nchunk;
        sl.lds = a.plan.lds;
        sl.off = a.off;
        sl.size = a.L.total;
        sl.general = !a.h->uniform3;
        slots_out[i] = slot;
    }
    b->free_slots.resize(fs - (size_t)count);
    return DNX_OK;
}

}  // namespace dnx

extern "C" {

int dnx_batch_advance(dnx_batch *b, int steps)
{
    if (!b || steps <= 0) return DNX_E_ARG;
    (void)hipSetDevice(b->ctx->device);
    // the host-side view of which slots run: the newest completed snapshot
    int rc = batch_enqueue(b, steps, b->h_dyn + (size_t)b->buf * b->nslots);
    if (rc) return rc;
    return DNX_OK;
}

// Copies the status table into pinned buffer `buf` after the queued work.
static int batch_snapshot(dnx_batch *b, int which)
{
    hipError_t e;
    if ((e = hipMemcpyAsync(b->h_dyn + (size_t)which * b->nslots, b->b_dyn.ptr, sizeof(BDyn) * (size_t)b->nslots,
                            hipMemcpyDeviceToHost, b->ctx->stream)) != hipSuccess)
        return hip_fail(e, "status copy");
    if ((e = hipEventRecord(b->ev[which], b->ctx->stream)) != hipSuccess) return hip_fail(e, "event");
    b->ev_pending[which] = true;
    return DNX_OK;
}

static int batch_wait(dnx_batch *b, int which)
{
    if (!b->ev_pending[which]) return DNX_OK;
    hipError_t e = hipEventSynchronize(b->ev[which]);
    b->ev_pending[which] = false;
    return e == hipSuccess ? DNX_OK : hip_fail(e, "status wait");
}

int dnx_batch_poll(dnx_batch *b, int *status, int64_t *steps, double *t, uint64_t *t_done_ticks)
{
    if (!b) return DNX_E_ARG;
    (void)hipSetDevice(b->ctx->device);
    int rc = batch_snapshot(b, b->buf);
    if (!rc) rc = batch_wait(b, b->buf);
    if (rc) return rc;
    const BDyn *d = b->h_dyn + (size_t)b->buf * b->nslots;
    for (int s = 0; s < b->nslots; s++) {
        if (status) status[s] = b->slots[(size_t)s].used ? d[s].status : -1;
        if (steps) steps[s] = d[s].steps;
        if (t) t[s] = d[s].t;
        if (t_done_ticks) t_done_ticks[s] = d[s].t_done;
    }
    return DNX_OK;
}

// True when group_stop is set and a slot in the status snapshot h_dyn
// (buffer 0) is satisfied.
static bool group_done(const dnx_batch *b)
{
    if (!b->group_stop) return false;
    for (int s = 0; s < b->nslots; s++)
        if (b->slots[(size_t)s].used && b->h_dyn[s].status == DNX_SAT) return true;
    return false;
}

// Runs every running slot in wide mode until none runs or the wall-clock
// limit passes.  The status table in h_dyn (buffer 0) is current on return.
static int batch_wide_run(dnx_batch *b, double walltime_s, std::chrono::steady_clock::time_point t0,
                          const std::vector<char> *skip = nullptr)
{
    hipStream_t st = b->ctx->stream;
    std::vector<int> list;
    for (int s = 0; s < b->nslots; s++)
        if (b->slots[(size_t)s].used && b->h_dyn[s].status == DNX_RUNNING && !(skip && (*skip)[(size_t)s]))
            list.push_back(s);
    if (list.empty()) return DNX_OK;
    std::sort(list.begin(), list.end(), [&](int x, int y) {
        return b->slots[(size_t)x].m != b->slots[(size_t)y].m ? b->slots[(size_t)x].m > b->slots[(size_t)y].m
                                                                : x < y;
    });
    std::vector<int> bmap, vmap;
    for (int s : list) {
        const SlotHost &sl = b->slots[(size_t)s];
        for (int j = 0; j * WIDE_BLOCK < sl.m; j++) {
            bmap.push_back(s);
            bmap.push_back(j);
        }
        for (int j = 0; j * WIDE_BLOCK < sl.n; j++) {
            vmap.push_back(s);
            vmap.push_back(j);
        }
    }
    const int count = (int)list.size(), nbc = (int)bmap.size() / 2, nbv = (int)vmap.size() / 2;
    int rc;
    auto grow = [&](DevBuf &buf, size_t bytes) { return buf.size >= bytes ? DNX_OK : buf.alloc(bytes); };
    if ((rc = grow(b->b_wlist, sizeof(int) * list.size())) || (rc = grow(b->b_bmap, sizeof(int) * (bmap.size() + 2))) ||
        (rc = grow(b->b_vmap, sizeof(int) * (vmap.size() + 2))))
        return rc;
    hipError_t e;
    if ((e = hipMemcpy(b->b_wlist.ptr, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (nbc && (e = hipMemcpy(b->b_bmap.ptr, bmap.data(), sizeof(int) * bmap.size(), hipMemcpyHostToDevice)) !=
                    hipSuccess) ||
        (nbv && (e = hipMemcpy(b->b_vmap.ptr, vmap.data(), sizeof(int) * vmap.size(), hipMemcpyHostToDevice)) !=
                    hipSuccess))
        return hip_fail(e, "wide map upload");
    Ctl *wctl = (Ctl *)b->b_wctl.ptr;
    BDyn *dyn = (BDyn *)b->b_dyn.ptr;
    launch_wide_enter(dyn, wctl, (const int *)b->b_wlist.ptr, count, 0, st);
    // chunks of wide_chunk steps (a multiple of 3, so the rotation returns
    // to 0), launched directly: the host enqueues a chunk far faster than
    // the device runs it.  (The maps change at every entry, so a captured
    // graph would be rebuilt each time; and under rocprofv3 kernel tracing
    // hipGraphLaunch of this graph ended the process with SIGSEGV inside the
    // profiler.)
    const int chunk = (b->wide_chunk + 2) / 3 * 3;
    const int r = 0;
    std::vector<long long> before(list.size());
    for (;;) {
        for (size_t i = 0; i < list.size(); i++) before[i] = b->h_dyn[list[i]].steps;
        const bool timed = b->timing && !b->t_pending[BK_WIDE];
        if (timed && (e = hipEventRecord(b->t0[BK_WIDE], st)) != hipSuccess) return hip_fail(e, "timing");
        for (int k = 0; k < chunk; k++)
            launch_wide_step((const BInst *)b->b_inst.ptr, wctl, (const int *)b->b_bmap.ptr, nbc,
                             (const int *)b->b_vmap.ptr, nbv, k % 3, st);
        if (timed) {
            if ((e = hipEventRecord(b->t1[BK_WIDE], st)) != hipSuccess) return hip_fail(e, "timing");
            b->t_pending[BK_WIDE] = true;
        }
        launch_wide_sync(dyn, wctl, (const int *)b->b_wlist.ptr, count, (r + 2) % 3, st);
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "wide launch");
        if ((rc = batch_snapshot(b, 0)) || (rc = batch_wait(b, 0))) return rc;
        batch_collect_timing(b);
        if (b->timing)
            for (size_t i = 0; i < list.size(); i++)
                b->t_alg[BK_WIDE] +=
                    (double)(b->h_dyn[list[i]].steps - before[i]) * step_bytes(b->slots[(size_t)list[i]]);
        int running = 0;
        long long live = 0;
        for (int s : list)
            if (b->h_dyn[s].status == DNX_RUNNING) {
                running++;
                live += b->slots[(size_t)s].m;
            }
        if (!running || group_done(b)) break;
        // finished instances still cost their blocks; rebuild the maps once
        // most of the clause blocks belong to finished instances
        if (2 * (live / WIDE_BLOCK + running) < nbc) break;
        if (walltime_s > 0 &&
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > walltime_s)
            break;
    }
    return DNX_OK;
}

// Gang mode for the tail (see batch.hip).  *used stays false when the
// running set does not suit it -- an instance too large for one gang, or
// more workgroups than gang_blocks or than can be resident at once even at
// the fewest workgroups each instance's clauses allow -- and the caller
// then runs wide mode.  Returns (for the caller to rebuild the map) after
// the running set has lost half its workgroups or, while some gang is cut
// below its instance's wish, after any instance has ended; also when
// nothing runs, or at the wall-clock limit.  The status table in h_dyn
// (buffer 0) is then current.
static int batch_gang_run(dnx_batch *b, double walltime_s, std::chrono::steady_clock::time_point t0, bool *used)
{
    *used = false;
    if (b->gang_blocks <= 0) return DNX_OK;
    if (!b->gang_cap) b->gang_cap = std::max(0, gang_occupancy()) * b->cus;
    // one gang per running instance, each gang on one XCD: its workgroups at
    // block indices x, x + 8, x + 16, ... (the dispatcher deals workgroups to
    // the XCDs in turn).  An instance asks for m / gang_cpw workgroups; when
    // the running set asks for more than the budget, every gang is cut to a
    // common level L (never below what its clauses need in LDS), the highest
    // level whose gangs fit the budget and pack onto the XCDs, largest first
    // to the least-loaded XCD.
    struct Want {
        int lo, hi, slot;
    };
    std::vector<Want> want;
    for (int s = 0; s < b->nslots; s++) {
        if (!b->slots[(size_t)s].used || b->h_dyn[s].status != DNX_RUNNING) continue;
        const int m = b->slots[(size_t)s].m;
        const int lo = std::max(1, (m + GANG_CPW - 1) / GANG_CPW);
        if (lo > GANG_MAX) return DNX_OK;
        want.push_back({lo, std::max(lo, std::min(GANG_MAX, (m + b->gang_cpw - 1) / b->gang_cpw)), s});
    }
    if (want.empty()) return DNX_OK;
    const int budget = std::min(b->gang_blocks, b->gang_cap);
    const int per_xcd = b->gang_cap / GANG_XCD;
    std::vector<std::pair<int, int>> gs;  // (G, slot)
    std::vector<std::vector<std::array<int, 4>>> on(GANG_XCD);
    std::vector<int> list, glist;
    std::vector<char> gcut;  // the gang is smaller than its instance asked for
    bool fit = false;
    for (int L = GANG_MAX; L >= 1 && !fit; L--) {
        gs.clear();
        int total = 0;
        for (const Want &w : want) {
            const int G = std::max(w.lo, std::min(w.hi, L));
            gs.push_back({G, w.slot});
            total += G;
        }
        if (total > budget) continue;
        std::sort(gs.begin(), gs.end(), [](const std::pair<int, int> &x, const std::pair<int, int> &y) {
            return x.first != y.first ? x.first > y.first : x.second < y.second;
        });
        for (auto &v : on) v.clear();
        list.clear();
        glist.clear();
        gcut.clear();
        fit = true;
        int gangs = 0;
        for (auto &g : gs) {
            int x = 0;
            for (int y = 1; y < GANG_XCD; y++)
                if (on[(size_t)y].size() < on[(size_t)x].size()) x = y;
            if ((int)on[(size_t)x].size() + g.first > per_xcd) {
                fit = false;
                break;
            }
            for (int r = 0; r < g.first; r++) on[(size_t)x].push_back({g.second, r, g.first, gangs});
            gangs++;
            list.push_back(g.second);
            glist.push_back(g.first);
            const int m = b->slots[(size_t)g.second].m;
            const int hi = std::max((m + GANG_CPW - 1) / GANG_CPW, std::min(GANG_MAX, (m + b->gang_cpw - 1) / b->gang_cpw));
            gcut.push_back(g.first < hi);
        }
    }
    if (!fit) return DNX_OK;
    const int gangs = (int)list.size();
    size_t W = 0;
    for (auto &v : on) W = std::max(W, v.size());
    std::vector<int> map((size_t)GANG_XCD * W * 4);
    for (size_t j = 0; j < W; j++)
        for (int x = 0; x < GANG_XCD; x++) {
            const size_t at = (j * GANG_XCD + (size_t)x) * 4;
            if (j < on[(size_t)x].size())
                for (int q = 0; q < 4; q++) map[at + (size_t)q] = on[(size_t)x][j][(size_t)q];
            else {
                map[at] = -1;
                map[at + 1] = 0;
                map[at + 2] = 1;
                map[at + 3] = 0;
            }
        }
    *used = true;
    const int nblk = (int)map.size() / 4;
    hipStream_t st = b->ctx->stream;
    int rc;
    hipError_t e;
    const size_t wbytes = (size_t)gangs * 2 * GANG_MAX * GANG_LINE * sizeof(unsigned long long);
    if (b->b_gmap.size < sizeof(int) * map.size() && (rc = b->b_gmap.alloc(sizeof(int) * map.size()))) return rc;
    if (b->b_gwords.size < wbytes) {
        // fresh lines are zero; epochs only grow, so lines never need clearing again
        if ((rc = b->b_gwords.alloc(wbytes))) return rc;
        if ((e = hipMemset(b->b_gwords.ptr, 0, wbytes)) != hipSuccess) return hip_fail(e, "gang words");
    }
    if (!b->b_gerr.ptr) {
        // [0] error flag, [1] arrival count
        if ((rc = b->b_gerr.alloc(2 * sizeof(int)))) return rc;
        if ((e = hipMemset(b->b_gerr.ptr, 0, 2 * sizeof(int))) != hipSuccess) return hip_fail(e, "gang flag");
        b->gang_arrive = 0;
    }
    if ((e = hipMemcpy(b->b_gmap.ptr, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice)) != hipSuccess)
        return hip_fail(e, "gang map upload");
    const int K = std::max(1, b->gang_steps);
    std::vector<long long> before(list.size());
    for (;;) {
        for (size_t i = 0; i < list.size(); i++) before[i] = b->h_dyn[list[i]].steps;
        const bool timed = b->timing && !b->t_pending[BK_GANG];
        if (timed && (e = hipEventRecord(b->t0[BK_GANG], st)) != hipSuccess) return hip_fail(e, "timing");
        const int le = launch_gang((const BInst *)b->b_inst.ptr, (BDyn *)b->b_dyn.ptr, (const int *)b->b_gmap.ptr, nblk,
                                   (unsigned long long *)b->b_gwords.ptr, b->gang_epoch, K, (int *)b->b_gerr.ptr,
                                   (int *)b->b_gerr.ptr + 1, b->gang_arrive, st);
        b->gang_epoch += 2 * (unsigned long long)K + 2;
        b->gang_arrive += nblk;
        if (le != 0) return hip_fail((hipError_t)le, "gang launch");
        if (timed) {
            if ((e = hipEventRecord(b->t1[BK_GANG], st)) != hipSuccess) return hip_fail(e, "timing");
            b->t_pending[BK_GANG] = true;
        }
        if ((rc = batch_snapshot(b, 0)) || (rc = batch_wait(b, 0))) return rc;
        batch_collect_timing(b);
        int err = 0;
        if ((e = hipMemcpy(&err, b->b_gerr.ptr, sizeof(int), hipMemcpyDeviceToHost)) != hipSuccess)
            return hip_fail(e, "gang flag read");
        if (err == 2) {
            // a workgroup ran on another XCD than its gang: the launch changed
            // nothing; gang mode is switched off for this batch
            b->gang_blocks = 0;
            if ((e = hipMemset(b->b_gerr.ptr, 0, sizeof(int))) != hipSuccess) return hip_fail(e, "gang flag");
            break;
        }
        if (err) {
            set_error("gang mode: a workgroup barrier wait timed out; the gang's instances are in an undefined state");
            return DNX_E_STATE;
        }
        if (b->timing)
            for (size_t i = 0; i < list.size(); i++)
                b->t_alg[BK_GANG] +=
                    (double)(b->h_dyn[list[i]].steps - before[i]) * step_bytes(b->slots[(size_t)list[i]]);
        int live = 0;
        bool any = false, ended = false, grow = false;
        for (size_t i = 0; i < list.size(); i++)
            if (b->h_dyn[list[i]].status == DNX_RUNNING) {
                any = true;
                live += glist[i];
                grow |= gcut[i] != 0;
            } else {
                ended = true;
            }
        if (!any || group_done(b)) break;
        // rebuild once half the workgroups idle, or as soon as an instance
        // has ended while a running gang is smaller than it asked for (the
        // freed workgroups can enlarge it)
        if (2 * live < nblk || (ended && grow)) break;
        if (walltime_s > 0 &&
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > walltime_s)
            break;
    }
    return DNX_OK;
}

// The tail split: single-chunk mode-0 instances (all their state in LDS) step
// faster in one workgroup than spread over the chip, the others faster in
// wide mode.  This launches the fused kernel for the first kind on its own
// stream, running each of them until it ends; wide mode then advances the
// rest on the context stream while it runs (the fused kernel holds one CU per
// instance and leaves the others to the wide kernels).  side_done marks its
// end; until then the status snapshots hold no valid entry for its slots.
static int batch_side_narrow(dnx_batch *b, const std::vector<int> &slots, int ksteps)
{
    const int S = b->nslots;
    // work-list region 1 is free: the fused kernel is not otherwise launched
    // while the tail runs
    int *hw = b->h_work + (size_t)BM_COUNT * S;
    int *dw = (int *)b->b_work.ptr + (size_t)BM_COUNT * S;
    int *dh = (int *)b->b_head.ptr + BM_COUNT;
    std::vector<int> order(slots);
    std::sort(order.begin(), order.end(), [&](int x, int y) {
        return b->slots[(size_t)x].m != b->slots[(size_t)y].m ? b->slots[(size_t)x].m > b->slots[(size_t)y].m
                                                                : x < y;
    });
    int lds = 0;
    for (size_t i = 0; i < order.size(); i++) {
        hw[i] = order[i];
        lds = std::max(lds, (int)b->slots[(size_t)order[i]].lds);
    }
    const int cnt = (int)order.size();
    hipStream_t ms = b->mstream[BM_LDS_ALL];
    hipError_t e;
    if ((e = hipMemcpyAsync(dw, hw, sizeof(int) * cnt, hipMemcpyHostToDevice, ms)) != hipSuccess)
        return hip_fail(e, "tail work list upload");
    launch_batch_zero(dh, 1, ms);
    const int bs = batch_max_threads();
    int per_cu = batch_occupancy(BM_LDS_ALL, bs, lds);
    if (per_cu < 1) per_cu = 1;
    // a quarter of the CUs stays free for the wide kernels
    const int grid = std::min(cnt, std::max(1, b->cus * per_cu - b->cus / 4));
    const bool timed = b->timing && !b->t_pending[BK_TAIL];
    if (timed && (e = hipEventRecord(b->t0[BK_TAIL], ms)) != hipSuccess) return hip_fail(e, "timing");
    launch_batch(BM_LDS_ALL, bs, lds, grid, (const BInst *)b->b_inst.ptr, (BDyn *)b->b_dyn.ptr, dw, cnt, dh, ksteps,
                 nullptr, ms);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "tail launch");
    if (timed) {
        if ((e = hipEventRecord(b->t1[BK_TAIL], ms)) != hipSuccess) return hip_fail(e, "timing");
        b->t_pending[BK_TAIL] = true;
    }
    if ((e = hipEventRecord(b->side_done, ms)) != hipSuccess) return hip_fail(e, "tail event");
    return DNX_OK;
}

// The tail with the split above: returns when every running slot has ended
// (the status table in h_dyn, buffer 0, is then current).  *used is false
// when there is no single-chunk mode-0 instance to split off.
static int batch_split_tail(dnx_batch *b, std::chrono::steady_clock::time_point t0, bool *used)
{
    *used = false;
    std::vector<int> nset;
    std::vector<char> skip((size_t)b->nslots, 0);
    for (int s = 0; s < b->nslots; s++) {
        const SlotHost &sl = b->slots[(size_t)s];
        if (sl.used && b->h_dyn[s].status == DNX_RUNNING && sl.mode == BM_LDS_ALL && sl.nchunk == 1) {
            nset.push_back(s);
            skip[(size_t)s] = 1;
        }
    }
    if (nset.empty()) return DNX_OK;
    *used = true;
    std::vector<long long> before(nset.size());
    for (size_t i = 0; i < nset.size(); i++) before[i] = b->h_dyn[nset[i]].steps;
    int rc = batch_side_narrow(b, nset, INT_MAX);
    if (rc) return rc;
    bool side_over = false;
    for (;;) {
        bool wide_running = false;
        for (int s = 0; s < b->nslots; s++)
            if (!skip[(size_t)s] && b->slots[(size_t)s].used && b->h_dyn[s].status == DNX_RUNNING)
                wide_running = true;
        if (!wide_running) break;
        // once the side launch has ended (nothing else holds CUs, and a
        // fresh snapshot has its slots' final states), the instances left
        // may fit gang mode
        if (!side_over && b->gang_blocks > 0 && hipEventQuery(b->side_done) == hipSuccess) {
            side_over = true;
            if ((rc = batch_snapshot(b, 0)) || (rc = batch_wait(b, 0))) return rc;
        }
        if (side_over) {
            bool g = false;
            if ((rc = batch_gang_run(b, 0.0, t0, &g))) return rc;
            if (g) continue;
        }
        if ((rc = batch_wide_run(b, 0.0, t0, &skip))) return rc;
    }
    hipError_t e;
    if ((e = hipStreamWaitEvent(b->ctx->stream, b->side_done, 0)) != hipSuccess) return hip_fail(e, "tail join");
    if ((rc = batch_snapshot(b, 0)) || (rc = batch_wait(b, 0))) return rc;
    batch_collect_timing(b);
    if (b->timing)
        for (size_t i = 0; i < nset.size(); i++)
            b->t_alg[BK_TAIL] += (double)(b->h_dyn[nset[i]].steps - before[i]) * step_bytes(b->slots[(size_t)nset[i]]);
    return DNX_OK;
}

// ---------------------------------------------------------------------------
// Elastic tail.  While more instances run than the chip has CUs, the flow
// launch keeps every CU busy with one instance each.  Once fewer run, CUs
// fall idle while the longest instances still need thousands of steps at
// the pace of one workgroup.  An epoch launch (launch_epoch) gives the spare
// CUs to the largest running instances as gangs and keeps the others on
// single workgroups, in one cooperative launch; each epoch ends when its
// gangs have taken epoch_steps steps, and the next epoch is planned from the
// statuses (more spare CUs as instances end: more or larger gangs).
// ---------------------------------------------------------------------------
static size_t ep_align(size_t x, size_t a) { return (x + a - 1) / a * a; }

struct EpochPlan {
    std::vector<int> gmap;    // GANG_XCD x GANG_MAX entries of (slot, rank, G | regrouped << 16, gang)
    int gcount[GANG_XCD] = {};
    int members = 0;          // gang workgroups in total
    int gangs = 0;
    std::vector<unsigned long long> goff;  // each gang's scratch offset (regrouped gangs)
    unsigned long long scratch = 0;        // scratch bytes in total
    std::vector<int> flow;    // the other running instances, largest first
};

// Instances of at least epoch_min_clauses clauses, largest first, each get
// the workgroups their clauses need in LDS (2048 per workgroup) on the XCD
// with the most room, as long as one CU stays for each instance left to the
// flow workers; then the gangs grow one workgroup at a time, largest
// instance first, up to what they ask for (m / gang_cpw, at most GANG_MAX)
// while CUs are to spare on their XCDs.  False when no gang forms.
static bool epoch_plan(const dnx_batch *b, const std::vector<int> &run, EpochPlan &P)
{
    const int per_xcd = b->cus / GANG_XCD;
    std::vector<int> order(run);
    std::sort(order.begin(), order.end(), [&](int x, int y) {
        const int mx = b->slots[(size_t)x].m, my = b->slots[(size_t)y].m;
        return mx != my ? mx > my : x < y;
    });
    int spare = b->cus - (int)run.size();  // CUs left once every instance has one
    int room[GANG_XCD];
    for (int x = 0; x < GANG_XCD; x++) room[x] = per_xcd;
    struct G {
        int slot, x, size, hi;
    };
    std::vector<G> gs;
    std::vector<char> in((size_t)b->nslots, 0);
    int members = 0;
    for (int s : order) {
        const int m = b->slots[(size_t)s].m;
        if (m < b->epoch_min_clauses) break;
        const int lo = std::max({1, (m + GANG_CPW - 1) / GANG_CPW, (m + b->epoch_cpw - 1) / b->epoch_cpw});
        if (lo > GANG_MAX || lo - 1 > spare) continue;
        int x = 0;
        for (int y = 1; y < GANG_XCD; y++)
            if (room[y] > room[x]) x = y;
        if (room[x] < lo) continue;
        const int hi = std::max(lo, std::min(GANG_MAX, (m + b->gang_cpw - 1) / b->gang_cpw));
        if (members + lo > b->gang_blocks) continue;
        gs.push_back({s, x, lo, hi});
        room[x] -= lo;
        spare -= lo - 1;
        members += lo;
        in[(size_t)s] = 1;
    }
    if (gs.empty()) return false;
    for (bool grew = true; grew && spare > 0;) {
        grew = false;
        for (G &g : gs)
            if (spare > 0 && g.size < g.hi && room[g.x] > 0 && members < b->gang_blocks) {
                g.size++;
                room[g.x]--;
                spare--;
                members++;
                grew = true;
            }
    }
    P.gmap.assign((size_t)GANG_XCD * GANG_MAX * 4, -1);
    for (int x = 0; x < GANG_XCD; x++) P.gcount[x] = 0;
    P.members = 0;
    P.gangs = (int)gs.size();
    P.goff.assign(gs.size(), 0);
    P.scratch = 0;
    for (size_t gi = 0; gi < gs.size(); gi++) {
        const G &g = gs[gi];
        const SlotHost &sl = b->slots[(size_t)g.slot];
        // regrouped form: 3-SAT, and the tables fit (the kernel also checks
        // each workgroup's slot count and runs gang_run if one does not)
        const bool rx = b->gang_regroup && !sl.general && sl.n <= RX_NCAP && rx_vblock(sl.n, g.size) <= RX_VCAP &&
                        (sl.m + g.size - 1) / g.size <= GANG_CPW;
        if (rx) {
            P.goff[gi] = P.scratch;
            P.scratch += rx_layout(sl.n, sl.m, g.size).total;
        }
        for (int r = 0; r < g.size; r++) {
            int *e = &P.gmap[((size_t)g.x * GANG_MAX + (size_t)P.gcount[g.x]) * 4];
            e[0] = g.slot;
            e[1] = r;
            e[2] = g.size | (rx ? 1 << 16 : 0);
            e[3] = (int)gi;
            P.gcount[g.x]++;
        }
        P.members += g.size;
    }
    P.flow.clear();
    for (int s : order)
        if (!in[(size_t)s]) P.flow.push_back(s);
    return true;
}

// One epoch (see above).  *used is false when it does not apply (epochs
// off, an instance outside mode 0, or no gang forms); the status table in
// h_dyn (buffer 0) is current on return when *used.
static int batch_epoch_run(dnx_batch *b, bool *used)
{
    *used = false;
    if (b->epoch_min_clauses <= 0 || b->gang_blocks <= 0 || b->cus <= 0 || b->cus % GANG_XCD != 0) return DNX_OK;
    std::vector<int> run;
    for (int s = 0; s < b->nslots; s++) {
        if (!b->slots[(size_t)s].used || b->h_dyn[s].status != DNX_RUNNING) continue;
        if (b->slots[(size_t)s].mode != BM_LDS_ALL) return DNX_OK;
        run.push_back(s);
    }
    if (run.empty() || (int)run.size() >= b->cus) return DNX_OK;
    EpochPlan P;
    if (!epoch_plan(b, run, P)) return DNX_OK;
    int lds = 0;
    for (int s : P.flow) lds = std::max(lds, (int)b->slots[(size_t)s].lds);
    if (epoch_occupancy(lds) < 1) return DNX_OK;
    *used = true;
    const int S = b->nslots;
    hipStream_t st = b->ctx->stream;
    hipError_t e;
    int rc;
    // everything the launch reads goes over in one copy from a pinned
    // buffer: the gang table and counts, the gangs' scratch offsets, the
    // flow control block, the work list's head and the work list.  The host
    // rewrites the buffer only for the next epoch, after this one's statuses
    // came back, so the copy has long run by then.
    const size_t mapw = P.gmap.size() + GANG_XCD;  // ints
    const size_t o_goff = ep_align(sizeof(int) * mapw, 8), o_flow = ep_align(o_goff + 8 * GANG_XCD * GANG_MAX, 128);
    const size_t o_head = o_flow + sizeof(BFlow), o_work = o_head + 16;
    const size_t ebytes = o_work + sizeof(int) * P.flow.size();
    // the launch's error word comes back into the same pinned buffer, past
    // the largest work list, with the statuses
    const size_t o_err = ep_align(o_work + sizeof(int) * (size_t)S, 64);
    if (!b->h_ep) {
        const size_t cap = o_err + 64;
        if ((e = hipHostMalloc((void **)&b->h_ep, cap, hipHostMallocDefault)) != hipSuccess)
            return hip_fail(e, "hipHostMalloc");
        if ((rc = b->b_ep.alloc(cap))) return rc;
    }
    unsigned char *he = b->h_ep;
    std::memcpy(he, P.gmap.data(), sizeof(int) * P.gmap.size());
    std::memcpy(he + sizeof(int) * P.gmap.size(), P.gcount, sizeof(int) * GANG_XCD);
    std::memcpy(he + o_goff, P.goff.data(), sizeof(unsigned long long) * (size_t)P.gangs);
    BFlow f{};
    f.live = LLONG_MAX / 2;
    f.threshold = -1;
    f.gangs = P.members;
    std::memcpy(he + o_flow, &f, sizeof f);
    std::memset(he + o_head, 0, 16);
    std::memcpy(he + o_work, P.flow.data(), sizeof(int) * P.flow.size());
    if ((e = hipMemcpyAsync(b->b_ep.ptr, he, ebytes, hipMemcpyHostToDevice, st)) != hipSuccess)
        return hip_fail(e, "epoch upload");
    unsigned char *de = (unsigned char *)b->b_ep.ptr;
    // buffers grow at least twofold (a reallocation waits for the device)
    if (P.scratch > 0 && b->b_gscr.size < P.scratch &&
        (rc = b->b_gscr.alloc(std::max((size_t)P.scratch, 2 * b->b_gscr.size))))
        return rc;
    // barrier lines for as many gangs as an epoch can hold (one per gang workgroup)
    const size_t wbytes = (size_t)GANG_XCD * GANG_MAX * 2 * GANG_MAX * GANG_LINE * sizeof(unsigned long long);
    if (b->b_gwords.size < wbytes) {
        // fresh lines are zero; epochs only grow, so lines never need clearing again
        if ((rc = b->b_gwords.alloc(wbytes))) return rc;
        if ((e = hipMemsetAsync(b->b_gwords.ptr, 0, wbytes, st)) != hipSuccess) return hip_fail(e, "gang words");
    }
    if (!b->b_gerr.ptr) {
        if ((rc = b->b_gerr.alloc(2 * sizeof(int)))) return rc;
        if ((e = hipMemsetAsync(b->b_gerr.ptr, 0, 2 * sizeof(int), st)) != hipSuccess) return hip_fail(e, "gang flag");
        b->gang_arrive = 0;
    }
    std::vector<std::pair<int, long long>> before;
    if (b->timing)
        for (int s : run) before.push_back({s, b->h_dyn[s].steps});
    const int K = std::max(1, b->epoch_steps);
    const bool timed = b->timing && !b->t_pending[BK_EPOCH];
    if (timed && (e = hipEventRecord(b->t0[BK_EPOCH], st)) != hipSuccess) return hip_fail(e, "timing");
    const int le = launch_epoch((const BInst *)b->b_inst.ptr, (BDyn *)b->b_dyn.ptr, (const int *)de,
                                (const int *)de + P.gmap.size(), (unsigned long long *)b->b_gwords.ptr, b->gang_epoch,
                                K, (int *)b->b_gerr.ptr, (const int *)(de + o_work), (int)P.flow.size(),
                                (int *)(de + o_head), (BFlow *)(de + o_flow), (unsigned char *)b->b_gscr.ptr,
                                (const unsigned long long *)(de + o_goff), lds, b->cus, st);
    b->gang_epoch += 2 * (unsigned long long)K + 2;
    if (le != 0) return hip_fail((hipError_t)le, "epoch launch");
    if (timed) {
        if ((e = hipEventRecord(b->t1[BK_EPOCH], st)) != hipSuccess) return hip_fail(e, "timing");
        b->t_pending[BK_EPOCH] = true;
    }
    if ((e = hipMemcpyAsync(he + o_err, b->b_gerr.ptr, sizeof(int), hipMemcpyDeviceToHost, st)) != hipSuccess)
        return hip_fail(e, "epoch flag read");
    if ((rc = batch_snapshot(b, 0)) || (rc = batch_wait(b, 0))) return rc;
    batch_collect_timing(b);
    for (auto &q : before)
        b->t_alg[BK_EPOCH] += (double)(b->h_dyn[q.first].steps - q.second) * step_bytes(b->slots[(size_t)q.first]);
    int err = 0;
    std::memcpy(&err, he + o_err, sizeof(int));
    if (err == 3) {
        // not every gang member started: the gangs changed nothing (the flow
        // workers' steps stand); epochs are switched off for this batch
        b->epoch_min_clauses = 0;
        if ((e = hipMemset(b->b_gerr.ptr, 0, sizeof(int))) != hipSuccess) return hip_fail(e, "gang flag");
    } else if (err == 4) {
        set_error("epoch launch: a regrouped gang's exchange tables disagreed; the gang's instances are in an "
                  "undefined state");
        return DNX_E_STATE;
    } else if (err) {
        set_error("epoch launch: a gang barrier wait timed out; the gang's instances are in an undefined state");
        return DNX_E_STATE;
    }
    return DNX_OK;
}

// Idle workgroups at which a flow launch over `grid` workgroups should stop
// so that the largest running instance can get a gang (0: none could).
static int epoch_idle_trigger(const dnx_batch *b, const std::vector<int> &run, int grid)
{
    if (b->epoch_min_clauses <= 0 || b->gang_blocks <= 0 || b->cus % GANG_XCD != 0) return 0;
    // the largest running instance that a gang can take
    int lo = 0;
    for (int s : run) {
        const int m = b->slots[(size_t)s].m;
        if (m < b->epoch_min_clauses) continue;
        const int l = std::max({1, (m + GANG_CPW - 1) / GANG_CPW, (m + b->epoch_cpw - 1) / b->epoch_cpw});
        if (l <= GANG_MAX) lo = std::max(lo, l);
    }
    if (lo == 0) return 0;
    // CUs spare once `need` more workgroups have no instance: a gang of lo
    // takes lo - 1 of them besides its instance's own
    const int have = b->cus - grid;
    return std::max(1, lo - 1 - have);
}

int dnx_batch_run(dnx_batch *b, int steps_per_launch, double walltime_s, int *remaining)
{
    if (!b || steps_per_launch <= 0) return DNX_E_ARG;
    (void)hipSetDevice(b->ctx->device);
    auto t0 = std::chrono::steady_clock::now();
    int rc = batch_snapshot(b, 0);
    if (!rc) rc = batch_wait(b, 0);
    if (rc) return rc;
    b->buf = 0;
    int running = 0;
    long long start_clauses = 0;
    for (int s = 0; s < b->nslots; s++)
        if (b->slots[(size_t)s].used && b->h_dyn[s].status == DNX_RUNNING) start_clauses += b->slots[(size_t)s].m;
    // the fraction of the run's clauses, but no more than 75% of what gang
    // mode can hold (its workgroup budget times GANG_CPW clauses; the rest is
    // slack for each instance's gang rounding up), so that the tail starts
    // where gang mode takes it over whole instead of in wide mode
    long long frac_clauses = (long long)(b->wide_frac * (double)start_clauses);
    if (b->gang_blocks > 0) {
        if (!b->gang_cap) b->gang_cap = std::max(0, gang_occupancy()) * b->cus;
        const long long fit = (long long)std::min(b->gang_blocks, b->gang_cap) * GANG_CPW * 3 / 4;
        if (fit > 0) frac_clauses = std::min(frac_clauses, fit);
    }
    const long long threshold = std::max(b->wide_clauses, frac_clauses);
    for (;;) {
        const BDyn *known = b->h_dyn;
        running = 0;
        for (int s = 0; s < b->nslots; s++) running += b->slots[(size_t)s].used && known[s].status == DNX_RUNNING;
        if (!running || group_done(b)) break;
        if (walltime_s > 0 &&
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > walltime_s)
            break;
        long long clauses = 0;
        bool all0 = true;
        std::vector<int> run;
        for (int s = 0; s < b->nslots; s++)
            if (b->slots[(size_t)s].used && known[s].status == DNX_RUNNING) {
                clauses += b->slots[(size_t)s].m;
                all0 &= b->slots[(size_t)s].mode == BM_LDS_ALL;
                run.push_back(s);
            }
        // the elastic tail (epochs) replaces the clause threshold while every
        // running instance is a mode-0 instance and the run uses flow launches
        const bool elastic = b->epoch_min_clauses > 0 && b->gang_blocks > 0 && b->flow && walltime_s <= 0 &&
                             !b->group_stop && all0 && b->cus % GANG_XCD == 0;
        if (elastic) {
            bool used = false;
            if ((rc = batch_epoch_run(b, &used))) return rc;
            if (used) continue;
        }
        if (!elastic && clauses <= threshold) {
            bool used = false;
            if ((rc = batch_gang_run(b, walltime_s, t0, &used))) return rc;
            if (!used && b->tail_split && !b->group_stop && walltime_s <= 0 &&
                (rc = batch_split_tail(b, t0, &used)))
                return rc;
            if (!used && (rc = batch_wide_run(b, walltime_s, t0))) return rc;
            continue;
        }
        std::vector<std::pair<int, long long>> before;
        if (b->timing)
            for (int s = 0; s < b->nslots; s++)
                if (b->slots[(size_t)s].used && known[s].status == DNX_RUNNING) before.push_back({s, known[s].steps});
        // flow launch: when every running instance is a mode-0 instance and
        // the run has no wall-clock limit or group stop (those need the host
        // to look at the statuses every steps_per_launch steps)
        BFlow *flow = nullptr;
        if (b->flow && walltime_s <= 0 && !b->group_stop) {
            if (all0) {
                if (!b->b_flow.ptr && (rc = b->b_flow.alloc(sizeof(BFlow)))) return rc;
                BFlow f{};
                f.live = clauses;
                f.threshold = elastic ? -1 : threshold;
                f.stop = 0;
                // elastic: stop once enough workgroups idle for the largest
                // running instance to get a gang
                f.idle_trigger = elastic ? epoch_idle_trigger(b, run, std::min(running, b->cus)) : 0;
                hipError_t e;
                if ((e = hipMemcpyAsync(b->b_flow.ptr, &f, sizeof f, hipMemcpyHostToDevice, b->ctx->stream)) !=
                        hipSuccess ||
                    (e = hipStreamSynchronize(b->ctx->stream)) != hipSuccess)
                    return hip_fail(e, "flow upload");
                flow = (BFlow *)b->b_flow.ptr;
            }
        }
        if ((rc = batch_enqueue(b, steps_per_launch, known, flow))) return rc;
        if ((rc = batch_snapshot(b, 0))) return rc;
        if ((rc = batch_wait(b, 0))) return rc;
        batch_collect_timing(b);
        for (auto &q : before) {
            const SlotHost &sl = b->slots[(size_t)q.first];
            b->t_alg[sl.mode] += (double)(b->h_dyn[q.first].steps - q.second) * step_bytes(sl);
        }
    }
    if (remaining) *remaining = running;
    return DNX_OK;
}

int dnx_batch_stamp(dnx_batch *b, uint64_t *ticks, double *ticks_per_s)
{
    if (!b) return DNX_E_ARG;
    (void)hipSetDevice(b->ctx->device);
    launch_batch_stamp((unsigned long long *)b->b_stamp.ptr, b->ctx->stream);
    unsigned long long v = 0;
    hipError_t e = hipMemcpyAsync(&v, b->b_stamp.ptr, sizeof(v), hipMemcpyDeviceToHost, b->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "stamp");
    if (ticks) *ticks = v;
    if (ticks_per_s) *ticks_per_s = b->clock_hz;
    return DNX_OK;
}

int dnx_batch_read_state(dnx_batch *b, int slot, double *x, int *status, int64_t *steps, double *t)
{
    if (!b || slot < 0 || slot >= b->nslots || !b->slots[(size_t)slot].used) return DNX_E_ARG;
    (void)hipSetDevice(b->ctx->device);
    const SlotHost &sl = b->slots[(size_t)slot];
    const size_t N = (size_t)sl.n + 2 * (size_t)sl.m;
    hipStream_t st = b->ctx->stream;
    DevBuf bo;
    int rc = bo.alloc(sizeof(float) * (N ? N : 1));
    if (rc) return rc;
    launch_batch_materialize((const BInst *)b->b_inst.ptr + slot, (const BDyn *)b->b_dyn.ptr + slot, sl.n, sl.m,
                             (float *)bo.ptr, st);
    std::vector<float> xf(N ? N : 1);
    BDyn d;
    hipError_t e;
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "materialize");
    if ((N && (e = hipMemcpyAsync(xf.data(), bo.ptr, sizeof(float) * N, hipMemcpyDeviceToHost, st)) != hipSuccess) ||
        (e = hipMemcpyAsync(&d, (BDyn *)b->b_dyn.ptr + slot, sizeof(BDyn), hipMemcpyDeviceToHost, st)) !=
            hipSuccess)
        return hip_fail(e, "state copy");
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "state sync");
    const bool pend = d.status == DNX_RUNNING && d.plevel >= 0;
    static const double dts[5] = {0.125, 0.0625, 0.03125, 0.015625, 0.0078125};
    if (x)
        for (size_t i = 0; i < N; i++) x[i] = (double)xf[i];
    if (status) *status = d.status;
    if (steps) *steps = d.steps + (pend ? 1 : 0);
    if (t) *t = d.t + (pend ? dts[d.plevel] : 0.0);
    return DNX_OK;
}

int dnx_batch_assignment(dnx_batch *b, int slot, int8_t *assign, int *unsat)
{
    if (!b || slot < 0 || slot >= b->nslots || !b->slots[(size_t)slot].used || !assign) return DNX_E_ARG;
    (void)hipSetDevice(b->ctx->device);
    const SlotHost &sl = b->slots[(size_t)slot];
    std::vector<BF2> vd((size_t)(sl.n ? sl.n : 1));
    hipError_t e;
    if (sl.n &&
        (e = hipMemcpyAsync(vd.data(), b->h_inst[(size_t)slot].vd, sizeof(BF2) * (size_t)sl.n, hipMemcpyDeviceToHost,
                            b->ctx->stream)) != hipSuccess)
        return hip_fail(e, "assignment copy");
    if ((e = hipStreamSynchronize(b->ctx->stream)) != hipSuccess) return hip_fail(e, "assignment sync");
    for (int v = 0; v < sl.n; v++) assign[v] = vd[(size_t)v].v >= 0.0f ? 1 : -1;
    if (unsat) {
        // the clause table the slot holds (uploaded verbatim at admission)
        std::vector<int32_t> lits((size_t)(sl.nnz ? sl.nnz : 1)), cp;
        const BInst &I = b->h_inst[(size_t)slot];
        if (sl.nnz && (e = hipMemcpy(lits.data(), I.lits, sizeof(int32_t) * (size_t)sl.nnz,
                                     hipMemcpyDeviceToHost)) != hipSuccess)
            return hip_fail(e, "clause table copy");
        if (sl.general) {
            cp.resize((size_t)sl.m + 1);
            if ((e = hipMemcpy(cp.data(), I.cp, sizeof(int32_t) * cp.size(), hipMemcpyDeviceToHost)) != hipSuccess)
                return hip_fail(e, "clause table copy");
        }
        int bad = 0;
        for (int c = 0; c < sl.m; c++) {
            bool ok = false;
            const int q0 = sl.general ? cp[(size_t)c] : 3 * c, q1 = sl.general ? cp[(size_t)c + 1] : 3 * c + 3;
            for (int q = q0; q < q1 && !ok; q++) {
                int32_t lit = lits[(size_t)q];
                int8_t a = assign[std::abs(lit) - 1];
                ok = (lit > 0 && a > 0) || (lit < 0 && a < 0);
            }
            bad += !ok;
        }
        *unsat = bad;
    }
    return DNX_OK;
}

int dnx_batch_retire(dnx_batch *b, int slot)
{
    if (!b || slot < 0 || slot >= b->nslots || !b->slots[(size_t)slot].used) return DNX_E_ARG;
    (void)hipSetDevice(b->ctx->device);
    // kernels still queued may read this slot: retire after they finish
    int rc = batch_sync(b);
    if (rc) return rc;
    SlotHost &sl = b->slots[(size_t)slot];
    b->arena.release(sl.off, sl.size);
    BDyn D;
    memset(&D, 0, sizeof(D));
    D.status = -1;
    D.plevel = -1;
    b->h_dyn[slot] = D;
    b->h_dyn[(size_t)b->nslots + slot] = D;
    b->h_stage[slot] = D;
    hipError_t e = hipMemcpy((BDyn *)b->b_dyn.ptr + slot, &b->h_stage[slot], sizeof(BDyn), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "retire");
    sl = SlotHost();
    b->free_slots.push_back(slot);
    return DNX_OK;
}

int dnx_batch_info(dnx_batch *b, int slot, int *n, int *m, int *mode)
{
    if (!b || slot < 0 || slot >= b->nslots) return DNX_E_ARG;
    const SlotHost &sl = b->slots[(size_t)slot];
    if (n) *n = sl.n;
    if (m) *m = sl.m;
    if (mode) *mode = sl.used ? sl.mode : -1;
    return DNX_OK;
}

int dnx_batch_set_wide(dnx_batch *b, int64_t max_clauses, int steps_per_sync)
{
    if (!b || max_clauses < 0 || steps_per_sync <= 0) return DNX_E_ARG;
    b->wide_clauses = max_clauses;
    b->wide_frac = 0.0;
    b->wide_chunk = steps_per_sync;
    return DNX_OK;
}

int dnx_batch_set_gang(dnx_batch *b, int max_blocks, int clauses_per_workgroup, int steps_per_launch)
{
    if (!b || max_blocks < 0 || clauses_per_workgroup <= 0 || clauses_per_workgroup > GANG_CPW ||
        steps_per_launch <= 0)
        return DNX_E_ARG;
    b->gang_blocks = max_blocks;
    b->gang_cpw = clauses_per_workgroup;
    b->gang_steps = steps_per_launch;
    return DNX_OK;
}

int dnx_batch_set_order(dnx_batch *b, int64_t large_clauses)
{
    if (!b || large_clauses < 0) return DNX_E_ARG;
    b->order_big = large_clauses;
    return DNX_OK;
}

int dnx_batch_set_tail_split(dnx_batch *b, int enable)
{
    if (!b) return DNX_E_ARG;
    b->tail_split = enable != 0;
    return DNX_OK;
}

int dnx_batch_set_epoch(dnx_batch *b, int64_t min_clauses, int steps, int clauses_per_workgroup)
{
    if (!b || min_clauses < 0 || steps <= 0 || clauses_per_workgroup < 64) return DNX_E_ARG;
    b->epoch_min_clauses = min_clauses;
    b->epoch_steps = steps;
    b->epoch_cpw = clauses_per_workgroup;
    return DNX_OK;
}

int dnx_batch_set_gang_regroup(dnx_batch *b, int enable)
{
    if (!b) return DNX_E_ARG;
    b->gang_regroup = enable != 0;
    return DNX_OK;
}

int dnx_batch_set_flow(dnx_batch *b, int enable)
{
    if (!b) return DNX_E_ARG;
    b->flow = enable != 0;
    return DNX_OK;
}

int dnx_batch_set_wide_fraction(dnx_batch *b, double fraction, int steps_per_sync)
{
    if (!b || !(fraction >= 0.0 && fraction <= 1.0) || steps_per_sync <= 0) return DNX_E_ARG;
    b->wide_clauses = 0;
    b->wide_frac = fraction;
    b->wide_chunk = steps_per_sync;
    return DNX_OK;
}

int dnx_batch_set_serve_big(dnx_batch *b, int64_t clauses, int64_t active_clauses)
{
    if (!b || clauses < 0 || active_clauses <= 0) return DNX_E_ARG;
    b->serve_big = clauses;
    b->serve_active_clauses = active_clauses;
    return DNX_OK;
}

int dnx_batch_set_serve_sliced(dnx_batch *b, int64_t min_clauses, int max_active)
{
    if (!b || min_clauses < 0 || max_active < 1 || max_active > 4096) return DNX_E_ARG;
    b->serve_sliced = min_clauses;
    b->serve_sliced_active = max_active;
    return DNX_OK;
}

int dnx_batch_served_by(dnx_batch *b, int64_t *counts)
{
    if (!b || !counts) return DNX_E_ARG;
    for (int i = 0; i < 4; i++) counts[i] = b->served_by[i];
    return DNX_OK;
}

int dnx_batch_set_serve_order(dnx_batch *b, int64_t las_steps)
{
    if (!b || las_steps < 0) return DNX_E_ARG;
    b->serve_las_steps = las_steps;
    return DNX_OK;
}

int dnx_batch_timing(dnx_batch *b, int enable, double *ms, int64_t *launches, double *alg_bytes)
{
    if (!b) return DNX_E_ARG;
    (void)hipSetDevice(b->ctx->device);
    int rc = batch_sync(b);
    if (rc) return rc;
    batch_collect_timing(b);
    for (int md = 0; md < NBUCKET; md++) {
        if (ms) ms[md] = b->t_ms[md];
        if (launches) launches[md] = b->t_launches[md];
        if (alg_bytes) alg_bytes[md] = b->t_alg[md];
    }
    if (enable >= 0) {
        b->timing = enable != 0;
        for (int md = 0; md < NBUCKET; md++) {
            b->t_ms[md] = 0.0;
            b->t_launches[md] = 0;
            b->t_alg[md] = 0.0;
        }
    }
    return DNX_OK;
}

int dnx_batch_memory(dnx_batch *b, int64_t *arena_bytes, int64_t *free_bytes)
{
    if (!b) return DNX_E_ARG;
    if (arena_bytes) *arena_bytes = (int64_t)b->arena.cap;
    if (free_bytes) *free_bytes = (int64_t)b->arena.free_bytes();
    return DNX_OK;
}

}  // extern "C"

namespace dnx {

// One request of a served stream.
struct ServeReq {
    const HostProblem *hp = nullptr;
    const double *x0 = nullptr;
    dnx_params prm{};
};

struct ServeOut {
    int *status;
    int64_t *steps;
    double *tts, *wait;
    // optional (may be null, as may their entries): simulated time, sign
    // assignment and final state (reference layout) of each request
    double *t = nullptr;
    int8_t *const *assign = nullptr;
    double *const *x = nullptr;
    int *where = nullptr;  // optional: DNX_SERVE_* per request
};

// The serving loop's sliced lane: 3-SAT requests of at least serve_sliced
// clauses run with the single-instance path's sliced exchange (solo.hpp),
// their tables and state in arena ranges.  Every 30-step chunk advances an
// active set of resident requests together -- one launch pair per step whose
// blocks a map deals to (request, unit) and (request, fold block) -- chosen
// in the wide lane's order up to the lane's clause budget; after the chunk a
// kernel gathers the members' control-block heads for one copy and an event.
// Two chunks are in flight.  Plans are built by planner threads from arrival
// on (units of ~3,000 clauses, so many requests share the device); a request
// whose plan does not apply goes back to the other lanes.
struct SlicedLane {
    static constexpr int CHUNK = 30;  // steps per chunk (a multiple of 3)
    static constexpr int HEADW = 32;  // ints gathered per control block (covers its head)
    struct Run {
        int r = -1;
        size_t off = 0, size = 0;
        std::unique_ptr<SlicedPlan> pl;
        SoloLayout L{};
        DevProb P{};
        DevParams prm{};
        SXDesc desc{};
        unsigned char *base = nullptr;
        long long attained = 0;
        int inflight = 0;   // chunks in flight that include it
        bool ended = false;
        Ctl last{};         // head of its last polled control block
    };
    struct Chunk {
        std::vector<int> members;  // run indices
        hipEvent_t ev = nullptr;
        DevBuf desc, mapc, mapv, heads;
        unsigned char *h_stage = nullptr;  // pinned: desc + maps going up, heads coming down
        size_t h_size = 0;
        size_t heads_off = 0;  // where the heads land in h_stage
        bool busy = false;
    };
    static_assert(offsetof(Ctl, track_loc) <= HEADW * sizeof(int), "the gathered head covers the poll fields");
    dnx_batch *b;
    const std::function<ServeReq(int)> &get;
    bool on = false;
    hipStream_t st = nullptr;
    std::vector<Run> runs;
    Chunk chunks[2];
    int next_chunk = 0;
    std::deque<int> queue;  // routed requests, arrival order
    // planner threads
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int> jobs;
    std::map<int, std::unique_ptr<SlicedPlan>> plans;  // done
    std::set<int> failed;                              // the exchange does not apply
    bool stop = false;
    std::vector<std::thread> planners;
    // plans held at once (ready or being built): a plan of an n=100,000
    // request holds tens of MB of host memory
    static constexpr int AHEAD = 16;
    int building = 0;
    // pinned staging of one admission, free again once `up_ev` has passed
    unsigned char *stage = nullptr;
    size_t stage_size = 0;
    hipEvent_t up_ev = nullptr;
    bool up_pending = false;

    SlicedLane(dnx_batch *bb, const std::function<ServeReq(int)> &g) : b(bb), get(g) {}
    int init()
    {
        on = b->serve_sliced > 0;
        if (!on) return DNX_OK;
        hipError_t e;
        if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&up_ev, hipEventDisableTiming)) != hipSuccess)
            return hip_fail(e, "sliced lane setup");
        for (auto &c : chunks)
            if ((e = hipEventCreateWithFlags(&c.ev, hipEventDisableTiming)) != hipSuccess)
                return hip_fail(e, "sliced lane setup");
        runs.resize((size_t)b->serve_sliced_active);
        const unsigned hc = std::thread::hardware_concurrency();
        const int np = std::max(1, std::min(6, (int)(hc ? hc : 2) / 2));
        for (int i = 0; i < np; i++) planners.emplace_back([this] { plan_loop(); });
        return DNX_OK;
    }
    void plan_loop()
    {
        for (;;) {
            int r;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || (!jobs.empty() && (int)plans.size() + building < AHEAD); });
                if (stop) return;
                r = jobs.front();
                jobs.pop_front();
                building++;
            }
            const ServeReq q = get(r);
            auto pl = std::make_unique<SlicedPlan>();
            const bool ok = solo_plan_units(*q.hp, b->cus, 3072, *pl);
            std::lock_guard<std::mutex> lk(mu);
            building--;
            if (ok) {
                plans[r] = std::move(pl);
            } else {
                failed.insert(r);
                cv.notify_one();  // one fewer held
            }
        }
    }
    // planners must be done with every request before any is released
    void stop_planners()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv.notify_all();
        for (auto &t : planners) t.join();
        planners.clear();
    }
    bool wants(const HostProblem &h) const { return on && h.uniform3 && h.m >= b->serve_sliced; }
    void route(int r)
    {
        queue.push_back(r);
        {
            std::lock_guard<std::mutex> lk(mu);
            jobs.push_back(r);
        }
        cv.notify_one();
    }
    int active() const
    {
        int k = 0;
        for (auto &R : runs) k += R.r >= 0;
        return k;
    }
    // whether the plan of request r is done: 1 ready, -1 does not apply, 0
    // not yet
    int plan_state(int r, std::unique_ptr<SlicedPlan> &out)
    {
        std::lock_guard<std::mutex> lk(mu);
        if (failed.count(r)) {
            failed.erase(r);
            return -1;
        }
        auto it = plans.find(r);
        if (it == plans.end()) return 0;
        out = std::move(it->second);
        plans.erase(it);
        cv.notify_one();  // room for one more plan
        return 1;
    }
    // places request r (plan pl) in run slot R; DNX_E_BUDGET when the arena
    // has no room for it now
    int start(Run &R, int r, std::unique_ptr<SlicedPlan> &pl, int64_t cap)
    {
        const ServeReq q = get(r);
        const HostProblem &h = *q.hp;
        const SoloLayout L = solo_layout(h, *pl);
        size_t off;
        if (!b->arena.alloc(L.total, off)) return DNX_E_BUDGET;
        hipError_t e;
        if (up_pending) {
            if ((e = hipEventSynchronize(up_ev)) != hipSuccess) return hip_fail(e, "sliced lane staging wait");
            up_pending = false;
        }
        if (stage_size < L.upload) {
            if (stage) (void)hipHostFree(stage);
            stage = nullptr;
            stage_size = 0;
            if ((e = hipHostMalloc((void **)&stage, L.upload, hipHostMallocDefault)) != hipSuccess) {
                b->arena.release(off, L.total);
                return hip_fail(e, "sliced lane staging");
            }
            stage_size = L.upload;
        }
        solo_stage(h, *pl, q.x0, cap, L, stage);
        unsigned char *base = (unsigned char *)b->arena_base + off;
        R.r = r;
        R.off = off;
        R.size = L.total;
        R.L = L;
        R.base = base;
        R.P = solo_devprob(h, *pl, L, base);
        R.prm = dev_params(&q.prm);
        R.desc = solo_desc(h, *pl, L, base, R.prm);
        R.pl = std::move(pl);
        R.attained = 0;
        R.inflight = 0;
        R.ended = false;
        memset(&R.last, 0, sizeof(R.last));
        if ((e = hipMemcpyAsync(base, stage, L.upload, hipMemcpyHostToDevice, st)) != hipSuccess ||
            (e = hipEventRecord(up_ev, st)) != hipSuccess)
            return hip_fail(e, "sliced lane upload");
        up_pending = true;
        launch_load_state(DNX_FP32, R.P, base + L.xstage, base + L.vd, base + L.xs, base + L.xl, st);
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "sliced lane state load");
        return DNX_OK;
    }
    static int status_of(const Ctl &c)
    {
        if (c.status != DNX_RUNNING) return c.status;
        return c.done[c.steps % 3];
    }
    // Issues a chunk if a chunk buffer is free and some resident request can
    // advance: members in order (below serve_las_steps steps by least
    // attained service, then by least expected work, m exp(-1.5 m/n)) up to
    // the lane's clause budget, at least one.
    int issue()
    {
        Chunk &C = chunks[next_chunk];
        if (C.busy) return DNX_OK;
        const long long T = b->serve_las_steps;
        std::vector<int> cand;
        for (int k = 0; k < (int)runs.size(); k++)
            if (runs[(size_t)k].r >= 0 && !runs[(size_t)k].ended) cand.push_back(k);
        if (cand.empty()) return DNX_OK;
        auto fg = [&](int k) { return T <= 0 || runs[(size_t)k].attained < T; };
        auto work = [&](int k) {
            const Run &R = runs[(size_t)k];
            return (double)R.P.m * std::exp(-1.5 * (double)R.P.m / (double)std::max(R.P.n, 1));
        };
        std::stable_sort(cand.begin(), cand.end(), [&](int x, int y) {
            const bool fx = fg(x), fy = fg(y);
            if (fx != fy) return fx;
            if (fx) return runs[(size_t)x].attained < runs[(size_t)y].attained;
            return work(x) < work(y);
        });
        long long budget = 0;
        C.members.clear();
        for (int k : cand) {
            const long long mk = runs[(size_t)k].P.m;
            if (!C.members.empty() && budget + mk > b->serve_active_clauses) break;
            C.members.push_back(k);
            budget += mk;
        }
        // descriptors and maps, staged in pinned memory, one copy up
        const int nm = (int)C.members.size();
        std::vector<int2> mc, mv;
        int region = 1;
        for (int j = 0; j < nm; j++) {
            const Run &R = runs[(size_t)C.members[(size_t)j]];
            for (int u = 0; u < R.P.nunits; u++) mc.push_back(make_int2(j, u));
            for (int v = 0; v * SX_F < R.P.n; v++) mv.push_back(make_int2(j, v));
            region = std::max(region, R.P.sx_region);
        }
        const size_t dbytes = al256(sizeof(SXDesc) * (size_t)nm), cbytes = al256(sizeof(int2) * mc.size()),
                     vbytes = al256(sizeof(int2) * (mv.empty() ? 1 : mv.size())),
                     hbytes = al256(sizeof(int) * HEADW * (size_t)nm);
        const size_t need = dbytes + cbytes + vbytes + hbytes;
        hipError_t e;
        int rc;
        if (C.h_size < need) {
            if (C.h_stage) (void)hipHostFree(C.h_stage);
            C.h_stage = nullptr;
            C.h_size = 0;
            const size_t want = std::max(need, (size_t)1 << 20);
            if ((e = hipHostMalloc((void **)&C.h_stage, want, hipHostMallocDefault)) != hipSuccess)
                return hip_fail(e, "sliced lane maps");
            C.h_size = want;
        }
        auto grow = [&](DevBuf &d, size_t bytes) { return d.size >= bytes ? DNX_OK : d.alloc(std::max(bytes, d.size * 2)); };
        if ((rc = grow(C.desc, dbytes)) || (rc = grow(C.mapc, cbytes)) || (rc = grow(C.mapv, vbytes)) ||
            (rc = grow(C.heads, hbytes)))
            return rc;
        unsigned char *h = C.h_stage;
        for (int j = 0; j < nm; j++) ((SXDesc *)h)[j] = runs[(size_t)C.members[(size_t)j]].desc;
        memcpy(h + dbytes, mc.data(), sizeof(int2) * mc.size());
        if (!mv.empty()) memcpy(h + dbytes + cbytes, mv.data(), sizeof(int2) * mv.size());
        if ((e = hipMemcpyAsync(C.desc.ptr, h, sizeof(SXDesc) * (size_t)nm, hipMemcpyHostToDevice, st)) != hipSuccess ||
            (e = hipMemcpyAsync(C.mapc.ptr, h + dbytes, sizeof(int2) * mc.size(), hipMemcpyHostToDevice, st)) !=
                hipSuccess ||
            (!mv.empty() && (e = hipMemcpyAsync(C.mapv.ptr, h + dbytes + cbytes, sizeof(int2) * mv.size(),
                                                hipMemcpyHostToDevice, st)) != hipSuccess))
            return hip_fail(e, "sliced lane maps");
        for (int i = 0; i < CHUNK; i++)
            launch_step_sliced_multi((const SXDesc *)C.desc.ptr, C.mapc.ptr, (int)mc.size(), C.mapv.ptr,
                                     (int)mv.size(), region, i % 3, st);
        launch_ctl_gather((const SXDesc *)C.desc.ptr, nm, HEADW, (int *)C.heads.ptr, st);
        C.heads_off = dbytes + cbytes + vbytes;
        if ((e = hipGetLastError()) != hipSuccess ||
            (e = hipMemcpyAsync(h + C.heads_off, C.heads.ptr, sizeof(int) * HEADW * (size_t)nm,
                                hipMemcpyDeviceToHost, st)) != hipSuccess ||
            (e = hipEventRecord(C.ev, st)) != hipSuccess)
            return hip_fail(e, "sliced lane chunk");
        for (int k : C.members) runs[(size_t)k].inflight++;
        C.busy = true;
        next_chunk ^= 1;
        return DNX_OK;
    }
    // Polls the oldest chunk in flight; fills `ended` with runs that ended
    // and are in no chunk still in flight.  Returns 1 when a chunk completed.
    int poll(std::vector<int> &ended, int *rc)
    {
        *rc = DNX_OK;
        // the older chunk in flight: next_chunk when both are, else the other
        Chunk &C = chunks[next_chunk].busy ? chunks[next_chunk] : chunks[next_chunk ^ 1];
        if (!C.busy) return 0;
        const hipError_t q = hipEventQuery(C.ev);
        if (q == hipErrorNotReady) return 0;
        if (q != hipSuccess) {
            *rc = hip_fail(q, "sliced lane poll");
            return 0;
        }
        const size_t nm = C.members.size();
        const int *heads = (const int *)(C.h_stage + C.heads_off);
        for (size_t j = 0; j < nm; j++) {
            Run &R = runs[(size_t)C.members[j]];
            memcpy(&R.last, heads + j * HEADW, solo_ctl_head());
            R.attained = R.last.steps;
            R.inflight--;
            if (status_of(R.last) != DNX_RUNNING) R.ended = true;
            if (R.ended && R.inflight == 0) ended.push_back(C.members[j]);
        }
        C.busy = false;
        return 1;
    }
    void release_run(Run &R)
    {
        b->arena.release(R.off, R.size);
        R.r = -1;
        R.pl.reset();
    }
    ~SlicedLane()
    {
        stop_planners();
        if (st) (void)hipStreamSynchronize(st);
        for (auto &c : chunks) {
            if (c.ev) (void)hipEventDestroy(c.ev);
            if (c.h_stage) (void)hipHostFree(c.h_stage);
        }
        if (stage) (void)hipHostFree(stage);
        if (up_ev) (void)hipEventDestroy(up_ev);
        if (st) (void)hipStreamDestroy(st);
    }
};

// The continuous-batching loop.  Requests become visible at their arrival
// time and are admitted first come, first served as soon as a slot and an
// arena range are free.  Two lanes advance what is resident, each on its own
// streams and polled through events, so admission and retirement never wait
// for the other lane:
//   narrow  instances below `big` clauses, fused-kernel launches of `spl`
//           steps (one workgroup per instance);
//   wide    larger instances, graph-captured chunks of wide steps (their
//           clauses and variables spread over many workgroups).
// An instance is retired the moment a lane observes that it ended, which
// frees its memory for the queue.  `get(i)` returns request i (it may block
// while the request is still being prepared); `release(i)` says its host
// copy is no longer needed.
static int serve_loop(dnx_batch *b, int count, const double *arrival, int64_t cap, int spl, double deadline,
                      const std::function<ServeReq(int)> &get, const std::function<void(int)> &release,
                      const ServeOut &out, double *makespan)
{
    using clk = std::chrono::steady_clock;
    const clk::time_point T0 = clk::now();
    auto now_s = [&]() { return std::chrono::duration<double>(clk::now() - T0).count(); };
    // host time of a device clock reading: of three stamps, the one whose
    // call returned soonest; the reading is taken near the call's end (a
    // copy and a wait follow it), so the end is used: times to solution are
    // then never early, at most a few tens of microseconds late
    uint64_t tick0 = 0;
    double rate = 1e8, host0 = 0.0, best = 1e30;
    int rc;
    for (int k = 0; k < 3; k++) {
        uint64_t tk = 0;
        const double hb = now_s();
        if ((rc = dnx_batch_stamp(b, &tk, &rate))) return rc;
        const double he = now_s();
        if (he - hb < best) {
            best = he - hb;
            tick0 = tk;
            host0 = he;
        }
    }
    const long long big =
        b->wide_clauses > 0 || b->wide_frac > 0 ? std::max<long long>(b->serve_big, 1) : LLONG_MAX;
    hipStream_t st = b->ctx->stream;
    hipError_t e;
    // lane bookkeeping
    std::vector<int> slot_req((size_t)b->nslots, -1);
    std::vector<char> slot_wide((size_t)b->nslots, 0);
    std::vector<int> narrow_run, wide_set;  // members of each lane
    // the slots each lane's launch in flight covers: a lane judges only
    // these from its snapshot, because a slot admitted later may still show
    // its previous occupant's final status there
    std::vector<int> narrow_launched, wide_launched;
    std::vector<int> wide_active;                          // members the wide graph covers
    std::vector<long long> attained((size_t)b->nslots, 0);  // steps done per slot
    bool narrow_busy = false, wide_busy = false, wide_dirty = false;
    hipStream_t wst = nullptr;
    hipEvent_t ev_n = nullptr, ev_w = nullptr, ev_adm = nullptr;
    hipGraphExec_t wgraph = nullptr;
    int nbc = 0, nbv = 0;
    struct Cleanup {
        hipStream_t &wst;
        hipEvent_t &a, &b2, &c;
        hipGraphExec_t &g;
        ~Cleanup()
        {
            if (wst) (void)hipStreamSynchronize(wst);
            if (g) (void)hipGraphExecDestroy(g);
            if (a) (void)hipEventDestroy(a);
            if (b2) (void)hipEventDestroy(b2);
            if (c) (void)hipEventDestroy(c);
            if (wst) (void)hipStreamDestroy(wst);
        }
    } cleanup{wst, ev_n, ev_w, ev_adm, wgraph};
    if ((e = hipStreamCreateWithFlags(&wst, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&ev_n, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&ev_w, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&ev_adm, hipEventDisableTiming)) != hipSuccess)
        return hip_fail(e, "serve setup");
    BDyn *hn = b->h_dyn, *hw = b->h_dyn + b->nslots;  // lane snapshots
    for (auto &c : b->served_by) c = 0;
    SlicedLane sliced(b, get);
    if ((rc = sliced.init())) return rc;
    std::deque<int> queue;
    int next = 0, finished = 0;
    std::vector<int8_t> assign;

    // ends slot s's request: time-to-solution from the device tick, the
    // exact clause check of its sign assignment (copied on the lane's own
    // stream, which is idle), then the slot and its arena range are freed.
    // Free slots are in no work list, so the device status record is left
    // as it is.
    std::vector<BF2> vbuf;
    DevBuf xbuf;
    std::vector<float> xf;
    auto finish = [&](int s, const BDyn &d, hipStream_t ls) -> int {
        const int r = slot_req[(size_t)s];
        int stt = d.status;
        out.steps[r] = d.steps;
        if (out.t) out.t[r] = d.t;
        out.tts[r] = host0 + (double)(d.t_done - tick0) / rate - arrival[r];
        SlotHost &sl = b->slots[(size_t)s];
        if ((out.x && out.x[r]) || (out.assign && out.assign[r])) {
            // the final state as the instance leaves (the verified state when
            // solved: an ended instance has no pending step), on the lane's
            // own stream, which is idle
            const size_t N = (size_t)sl.n + 2 * (size_t)sl.m;
            if (xbuf.size < sizeof(float) * (N ? N : 1)) {
                const int ra = xbuf.alloc(sizeof(float) * (N ? N : 1));
                if (ra) return ra;
            }
            launch_batch_materialize((const BInst *)b->b_inst.ptr + s, (const BDyn *)b->b_dyn.ptr + s, sl.n, sl.m,
                                     (float *)xbuf.ptr, ls);
            xf.resize(N ? N : 1);
            hipError_t ee = hipGetLastError();
            if (ee == hipSuccess && N)
                ee = hipMemcpyAsync(xf.data(), xbuf.ptr, sizeof(float) * N, hipMemcpyDeviceToHost, ls);
            if (ee == hipSuccess) ee = hipStreamSynchronize(ls);
            if (ee != hipSuccess) return hip_fail(ee, "served state copy");
            if (out.x && out.x[r])
                for (size_t i = 0; i < N; i++) out.x[r][i] = (double)xf[i];
            if (out.assign && out.assign[r])
                for (int v = 0; v < sl.n; v++) out.assign[r][v] = xf[(size_t)v] >= 0.0f ? 1 : -1;
        }
        if (stt == DNX_SAT) {
            vbuf.resize((size_t)(sl.n ? sl.n : 1));
            hipError_t ee = sl.n ? hipMemcpyAsync(vbuf.data(), b->h_inst[(size_t)s].vd, sizeof(BF2) * (size_t)sl.n,
                                                  hipMemcpyDeviceToHost, ls)
                                 : hipSuccess;
            if (ee == hipSuccess) ee = hipStreamSynchronize(ls);
            if (ee != hipSuccess) return hip_fail(ee, "assignment copy");
            // against the request's own clause table (released below)
            const HostProblem &hq = *get(r).hp;
            int bad = 0;
            for (int c = 0; c < sl.m; c++) {
                bool ok = false;
                for (int q = hq.cp[(size_t)c]; q < hq.cp[(size_t)c + 1] && !ok; q++) {
                    const int32_t lit = hq.lits[(size_t)q];
                    const bool pos = vbuf[(size_t)std::abs(lit) - 1].v >= 0.0f;
                    ok = (lit > 0) == pos;
                }
                bad += !ok;
            }
            if (bad != 0) stt = DNX_FAILED;
        }
        out.status[r] = stt;
        if (out.where) out.where[r] = DNX_SERVE_ENDED;
        b->served_by[slot_wide[(size_t)s] ? 1 : 0]++;
        b->arena.release(sl.off, sl.size);
        sl = SlotHost();
        b->free_slots.push_back(s);
        slot_req[(size_t)s] = -1;
        release(r);
        finished++;
        return DNX_OK;
    };
    // ends a sliced-lane run: its final state read back on the lane's stream (the
    // verified state when solved), the exact clause check against the
    // request's own table, then its range is freed.  Time to solution is
    // taken by the host clock when the lane sees the end.
    auto finish_sliced = [&](SlicedLane::Run &R) -> int {
        const int r = R.r;
        const Ctl &c = R.last;
        int stt = SlicedLane::status_of(c);
        const HostProblem &hq = *get(r).hp;
        const size_t n = (size_t)hq.n, N = n + 2 * (size_t)hq.m;
        const bool want_x = out.x && out.x[r];
        launch_materialize(DNX_FP32, R.P, R.prm, R.base + R.L.vd, R.base + R.L.xs, R.base + R.L.xl,
                           (Ctl *)(R.base + R.L.ctl), R.base + R.L.xstage, sliced.st);
        xf.resize(N ? N : 1);
        const size_t rd = want_x ? N : n;
        hipError_t ee = hipGetLastError();
        if (ee == hipSuccess && rd)
            ee = hipMemcpyAsync(xf.data(), R.base + R.L.xstage, sizeof(float) * rd, hipMemcpyDeviceToHost, sliced.st);
        if (ee == hipSuccess) ee = hipStreamSynchronize(sliced.st);
        if (ee != hipSuccess) return hip_fail(ee, "sliced-lane state copy");
        if (stt == DNX_SAT) {
            std::vector<int8_t> a(n ? n : 1);
            for (size_t v = 0; v < n; v++) a[v] = xf[v] >= 0.0f ? 1 : -1;
            if (count_unsat(hq, a.data()) != 0) stt = DNX_FAILED;
        }
        out.status[r] = stt;
        out.steps[r] = c.steps;
        if (out.t) out.t[r] = c.t;
        out.tts[r] = now_s() - arrival[r];
        if (want_x)
            for (size_t i = 0; i < N; i++) out.x[r][i] = (double)xf[i];
        if (out.assign && out.assign[r])
            for (size_t v = 0; v < n; v++) out.assign[r][v] = xf[v] >= 0.0f ? 1 : -1;
        if (out.where) out.where[r] = DNX_SERVE_ENDED;
        b->served_by[2]++;
        sliced.release_run(R);
        release(r);
        finished++;
        return DNX_OK;
    };

    for (int i = 0; i < count; i++) {
        out.status[i] = DNX_UNKNOWN;
        out.steps[i] = 0;
        out.tts[i] = -1.0;
        out.wait[i] = -1.0;
        if (out.where) out.where[i] = DNX_SERVE_NOT_ARRIVED;
    }
    while (finished < count) {
        bool progressed = false;
        const double now = now_s();
        if (deadline > 0 && now > deadline) {
            // requests still queued, running or not yet arrived stay
            // DNX_UNKNOWN with no time-to-solution; where they were and the
            // steps they had completed are reported
            if (wst) (void)hipStreamSynchronize(wst);
            (void)hipStreamSynchronize(st);
            if (sliced.st) (void)hipStreamSynchronize(sliced.st);
            sliced.stop_planners();
            for (auto &R : sliced.runs) {
                if (R.r < 0) continue;
                out.steps[R.r] = R.attained;
                if (out.where) out.where[R.r] = DNX_SERVE_SLICED;
                const int r = R.r;
                sliced.release_run(R);
                release(r);
            }
            for (int r : sliced.queue) {
                if (out.where) out.where[r] = DNX_SERVE_SLICED_WAIT;
                release(r);
            }
            sliced.queue.clear();
            for (int s = 0; s < b->nslots; s++) {
                if (slot_req[(size_t)s] < 0) continue;
                const int r = slot_req[(size_t)s];
                out.steps[r] = attained[(size_t)s];
                if (out.where)
                    out.where[r] = !slot_wide[(size_t)s] ? DNX_SERVE_NARROW
                                   : std::binary_search(wide_active.begin(), wide_active.end(), s)
                                       ? DNX_SERVE_WIDE
                                       : DNX_SERVE_WIDE_WAIT;
                SlotHost &sl = b->slots[(size_t)s];
                b->arena.release(sl.off, sl.size);
                sl = SlotHost();
                b->free_slots.push_back(s);
                slot_req[(size_t)s] = -1;
                release(r);
            }
            for (int r : queue) {
                if (out.where) out.where[r] = DNX_SERVE_QUEUED;
                release(r);
            }
            for (int r = next; r < count; r++) release(r);
            break;
        }
        while (next < count && arrival[next] <= now) {
            if (sliced.on && sliced.wants(*get(next).hp)) {
                if (out.where) out.where[next] = DNX_SERVE_SLICED_WAIT;
                sliced.route(next++);
                continue;
            }
            if (out.where) out.where[next] = DNX_SERVE_QUEUED;
            queue.push_back(next++);
        }
        // ---- sliced lane: start requests whose plans are ready, retire the
        // ones that ended, keep two chunks queued
        if (sliced.on) {
            for (auto &R : sliced.runs) {
                if (R.r >= 0 || sliced.queue.empty()) continue;
                const int r = sliced.queue.front();
                std::unique_ptr<SlicedPlan> pl;
                const int ps = sliced.plan_state(r, pl);
                if (ps == 0) break;
                sliced.queue.pop_front();
                progressed = true;
                if (ps < 0) {  // the exchange does not apply: the other lanes take it
                    if (out.where) out.where[r] = DNX_SERVE_QUEUED;
                    queue.push_back(r);
                    continue;
                }
                rc = sliced.start(R, r, pl, cap);
                if (rc == DNX_E_BUDGET) {
                    rc = DNX_OK;
                    if (narrow_run.empty() && wide_set.empty() && sliced.active() == 0) {
                        // larger than the whole arena: it can never run
                        out.status[r] = DNX_FAILED;
                        out.wait[r] = now_s() - arrival[r];
                        if (out.where) out.where[r] = DNX_SERVE_ENDED;
                        b->served_by[3]++;
                        release(r);
                        finished++;
                        continue;
                    }
                    // no room now: back to the front with its plan kept
                    sliced.queue.push_front(r);
                    std::lock_guard<std::mutex> lk(sliced.mu);
                    sliced.plans[r] = std::move(pl);
                    break;
                }
                if (rc) return rc;
                out.wait[r] = now_s() - arrival[r];
                if (out.where) out.where[r] = DNX_SERVE_SLICED;
            }
            for (;;) {
                std::vector<int> ended;
                int prc = DNX_OK;
                const int got = sliced.poll(ended, &prc);
                if (prc) return prc;
                if (!got) break;
                progressed = true;
                for (int k : ended)
                    if ((rc = finish_sliced(sliced.runs[(size_t)k]))) return rc;
            }
            for (int i = 0; i < 2; i++)
                if ((rc = sliced.issue())) return rc;
        }
        // ---- admission -------------------------------------------------
        while (!queue.empty()) {
            const int r = queue.front();
            ServeReq q = get(r);
            int slot = -1;
            rc = dnx_batch_admit_host(b, *q.hp, &q.prm, q.x0, cap, 1, &slot);
            if (rc == DNX_E_BUDGET) {
                if (narrow_run.empty() && wide_set.empty() && sliced.active() == 0) {
                    // larger than the whole arena: it can never run
                    queue.pop_front();
                    out.status[r] = DNX_FAILED;
                    out.steps[r] = 0;
                    out.tts[r] = -1.0;
                    out.wait[r] = now_s() - arrival[r];
                    if (out.where) out.where[r] = DNX_SERVE_ENDED;
                    b->served_by[3]++;
                    release(r);
                    finished++;
                    continue;
                }
                break;
            }
            if (rc) return rc;
            slot_req[(size_t)slot] = r;
            attained[(size_t)slot] = 0;
            out.wait[r] = now_s() - arrival[r];
            if (q.hp->m >= big) {
                slot_wide[(size_t)slot] = 1;
                wide_set.push_back(slot);
                wide_dirty = true;
            } else {
                slot_wide[(size_t)slot] = 0;
                narrow_run.push_back(slot);
            }
            queue.pop_front();
            progressed = true;
        }
        // ---- narrow lane -----------------------------------------------
        if (narrow_busy && hipEventQuery(ev_n) == hipSuccess) {
            narrow_busy = false;
            progressed = true;
            for (int s : narrow_launched) {
                if (slot_req[(size_t)s] < 0) continue;
                attained[(size_t)s] = hn[s].steps;
                if (hn[s].status == DNX_RUNNING) continue;
                if ((rc = finish(s, hn[s], st))) return rc;
                narrow_run.erase(std::find(narrow_run.begin(), narrow_run.end(), s));
            }
        }
        if (!narrow_busy && !narrow_run.empty()) {
            narrow_launched = narrow_run;
            b->buf = 0;
            if ((rc = batch_enqueue_slots(b, spl, narrow_launched))) return rc;
            if ((e = hipMemcpyAsync(hn, b->b_dyn.ptr, sizeof(BDyn) * (size_t)b->nslots, hipMemcpyDeviceToHost, st)) !=
                    hipSuccess ||
                (e = hipEventRecord(ev_n, st)) != hipSuccess)
                return hip_fail(e, "narrow poll");
            narrow_busy = progressed = true;
        }
        // ---- wide lane -------------------------------------------------
        if (wide_busy && hipEventQuery(ev_w) == hipSuccess) {
            wide_busy = false;
            progressed = true;
            for (int s : wide_launched) {
                attained[(size_t)s] = hw[s].steps;
                if (hw[s].status == DNX_RUNNING) continue;
                if ((rc = finish(s, hw[s], wst))) return rc;
                wide_set.erase(std::find(wide_set.begin(), wide_set.end(), s));
                wide_dirty = true;
            }
        }
        if (!wide_busy && !wide_set.empty()) {
            // Two classes, then the lane's clause budget in that order; the
            // members left out keep their state and wait.  Members below
            // serve_las_steps steps: least attained service (new arrivals
            // start at once).  Members past it: least expected work first,
            // m * exp(-1.5 m/n) -- random 3-SAT step counts fall about 4.5x
            // per unit of clause ratio above the threshold (~4.27) -- so that
            // under overload the lane completes what it can instead of
            // advancing every long run equally.
            std::vector<int> active = wide_set;
            const long long T = b->serve_las_steps;
            auto fg = [&](int x) { return T <= 0 || attained[(size_t)x] < T; };
            auto work = [&](int x) {
                const SlotHost &sl = b->slots[(size_t)x];
                return (double)sl.m * std::exp(-1.5 * (double)sl.m / (double)std::max(sl.n, 1));
            };
            std::stable_sort(active.begin(), active.end(), [&](int x, int y) {
                const bool fx = fg(x), fy = fg(y);
                if (fx != fy) return fx;
                if (fx) return attained[(size_t)x] < attained[(size_t)y];
                return work(x) < work(y);
            });
            long long budget = 0;
            size_t keep = 0;
            while (keep < active.size() &&
                   (keep == 0 || budget + b->slots[(size_t)active[keep]].m <= b->serve_active_clauses)) {
                budget += b->slots[(size_t)active[keep]].m;
                keep++;
            }
            active.resize(keep);
            std::sort(active.begin(), active.end());
            if (active != wide_active) wide_dirty = true;
            if (wide_dirty || !wgraph) {
                wide_active = active;
                // new maps and graph for the active members, entered from
                // their status records (current after every chunk)
                std::vector<int> bmap, vmap;
                for (int s : wide_active) {
                    const SlotHost &sl = b->slots[(size_t)s];
                    for (int j = 0; j * WIDE_BLOCK < sl.m; j++) {
                        bmap.push_back(s);
                        bmap.push_back(j);
                    }
                    for (int j = 0; j * WIDE_BLOCK < sl.n; j++) {
                        vmap.push_back(s);
                        vmap.push_back(j);
                    }
                }
                nbc = (int)bmap.size() / 2;
                nbv = (int)vmap.size() / 2;
                auto grow = [&](DevBuf &buf, size_t bytes) { return buf.size >= bytes ? DNX_OK : buf.alloc(bytes); };
                if ((rc = grow(b->b_wlist, sizeof(int) * wide_active.size())) ||
                    (rc = grow(b->b_bmap, sizeof(int) * (bmap.size() + 2))) ||
                    (rc = grow(b->b_vmap, sizeof(int) * (vmap.size() + 2))))
                    return rc;
                if ((e = hipMemcpy(b->b_wlist.ptr, wide_active.data(), sizeof(int) * wide_active.size(),
                                   hipMemcpyHostToDevice)) != hipSuccess ||
                    (nbc && (e = hipMemcpy(b->b_bmap.ptr, bmap.data(), sizeof(int) * bmap.size(),
                                           hipMemcpyHostToDevice)) != hipSuccess) ||
                    (nbv && (e = hipMemcpy(b->b_vmap.ptr, vmap.data(), sizeof(int) * vmap.size(),
                                           hipMemcpyHostToDevice)) != hipSuccess))
                    return hip_fail(e, "wide map upload");
                if (wgraph) {
                    (void)hipGraphExecDestroy(wgraph);
                    wgraph = nullptr;
                }
                const int chunk = (b->serve_chunk + 2) / 3 * 3;
                hipGraph_t g;
                if ((e = hipStreamBeginCapture(wst, hipStreamCaptureModeThreadLocal)) != hipSuccess)
                    return hip_fail(e, "wide capture");
                for (int k = 0; k < chunk; k++)
                    launch_wide_step((const BInst *)b->b_inst.ptr, (Ctl *)b->b_wctl.ptr, (const int *)b->b_bmap.ptr,
                                     nbc, (const int *)b->b_vmap.ptr, nbv, k % 3, wst);
                if ((e = hipStreamEndCapture(wst, &g)) != hipSuccess) return hip_fail(e, "wide capture end");
                e = hipGraphInstantiate(&wgraph, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (e != hipSuccess) return hip_fail(e, "wide graph");
                launch_wide_enter((const BDyn *)b->b_dyn.ptr, (Ctl *)b->b_wctl.ptr, (const int *)b->b_wlist.ptr,
                                  (int)wide_active.size(), 0, wst);
                wide_dirty = false;
            }
            wide_launched = wide_active;
            if ((e = hipGraphLaunch(wgraph, wst)) != hipSuccess) return hip_fail(e, "wide launch");
            launch_wide_sync((BDyn *)b->b_dyn.ptr, (Ctl *)b->b_wctl.ptr, (const int *)b->b_wlist.ptr,
                             (int)wide_active.size(), 2, wst);
            if ((e = hipMemcpyAsync(hw, b->b_dyn.ptr, sizeof(BDyn) * (size_t)b->nslots, hipMemcpyDeviceToHost, wst)) !=
                    hipSuccess ||
                (e = hipEventRecord(ev_w, wst)) != hipSuccess)
                return hip_fail(e, "wide poll");
            wide_busy = progressed = true;
        }
        if (!progressed) {
            // nothing completed: sleep a little, or until the next arrival
            double dt = 50e-6;
            if (narrow_run.empty() && wide_set.empty() && sliced.active() == 0 && sliced.queue.empty() && next < count)
                dt = std::min(0.01, arrival[next] - now_s());
            if (dt > 0) std::this_thread::sleep_for(std::chrono::duration<double>(dt));
        }
    }
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "serve drain");
    if (makespan) *makespan = now_s();
    return DNX_OK;
}

}  // namespace dnx

extern "C" {

int dnx_serve_trace(dnx_batch *b, int count, dnx_problem *const *probs, const dnx_params *prms,
                    const double *const *x0s, const double *arrival_s, int64_t step_cap, int steps_per_launch,
                    double deadline_s, int *status, int64_t *steps, double *tts_s, double *wait_s, double *makespan_s,
                    double *t_out, int8_t *const *assign_out, double *const *x_out, int *where_out)
{
    if (!b || count < 0 || (count && (!probs || !x0s || !arrival_s || !status || !steps || !tts_s || !wait_s)) ||
        steps_per_launch <= 0)
        return DNX_E_ARG;
    for (int i = 1; i < count; i++)
        if (arrival_s[i] < arrival_s[i - 1]) {
            set_error("arrival times must be non-decreasing");
            return DNX_E_ARG;
        }
    auto get = [&](int i) {
        ServeReq q;
        q.hp = &probs[i]->hp;
        q.x0 = x0s[i];
        if (prms)
            q.prm = prms[i];
        else
            dnx_params_default(probs[i]->hp.m, 10000, &q.prm);
        return q;
    };
    ServeOut out{status, steps, tts_s, wait_s};
    out.t = t_out;
    out.assign = assign_out;
    out.x = x_out;
    out.where = where_out;
    return serve_loop(b, count, arrival_s, step_cap, steps_per_launch, deadline_s, get, [](int) {}, out, makespan_s);
}

// Served planted stream: Poisson arrivals at rate_per_s over duration_s
// (at most max_requests), n log-uniform in [n_min, n_max], ratio uniform in
// [r_min, r_max].  Instances are generated by worker threads ahead of their
// arrival time and freed at retirement.
int dnx_serve_planted(dnx_batch *b, uint64_t seed, double rate_per_s, double duration_s, int n_min, int n_max,
                      double r_min, double r_max, double p1, double p2, int64_t step_cap, int steps_per_launch,
                      double deadline_s, int max_requests, int gen_threads, int *count_out, double *arrival_out,
                      int *n_out, int *m_out, int *status, int64_t *steps, double *tts_s, double *wait_s,
                      double *makespan_s, int *where_out)
{
    if (!b || rate_per_s <= 0 || duration_s <= 0 || n_min < 3 || n_max < n_min || r_min <= 0 || r_max < r_min ||
        max_requests <= 0 || !count_out || !arrival_out || !n_out || !m_out || !status || !steps || !tts_s ||
        !wait_s || steps_per_launch <= 0)
        return DNX_E_ARG;
    // the trace: arrival times and instance sizes from one seeded stream
    SplitMix64 rng(seed);
    std::vector<uint64_t> iseed;
    int count = 0;
    double t = 0.0;
    for (;;) {
        t += -std::log(1.0 - rng.unit()) / rate_per_s;
        if (t > duration_s || count >= max_requests) break;
        const double ln = std::log((double)n_min) + rng.unit() * (std::log((double)n_max) - std::log((double)n_min));
        const int n = std::max(n_min, std::min(n_max, (int)std::lround(std::exp(ln))));
        const double r = r_min + rng.unit() * (r_max - r_min);
        arrival_out[count] = t;
        n_out[count] = n;
        m_out[count] = (int)std::lround(r * n);
        iseed.push_back(rng.next());
        count++;
    }
    *count_out = count;
    struct Slot {
        std::unique_ptr<HostProblem> hp;
        std::vector<double> x0;
        bool ready = false;
    };
    std::vector<Slot> reqs((size_t)count);
    std::mutex mu;
    std::condition_variable cv;
    std::atomic<int> next_gen{0};
    std::atomic<int> released{0};
    std::atomic<bool> stop{false};
    const double lookahead_s = 2.0;  // generate at most this far ahead of arrival
    using clk = std::chrono::steady_clock;
    const clk::time_point T0 = clk::now();
    auto worker = [&]() {
        for (;;) {
            const int i = next_gen.fetch_add(1);
            if (i >= count || stop.load()) return;
            for (;;) {  // keep host memory bounded: stay near real time
                const double now = std::chrono::duration<double>(clk::now() - T0).count();
                if (arrival_out[i] - now < lookahead_s || stop.load()) break;
                std::this_thread::sleep_for(std::chrono::milliseconds(5));
            }
            auto hp = std::make_unique<HostProblem>();
            hp->n = n_out[i];
            hp->m = m_out[i];
            hp->lits.resize((size_t)hp->m * 3);
            generate_planted(iseed[(size_t)i], hp->n, hp->m, p1, p2, hp->lits.data(), nullptr);
            hp->cp.resize((size_t)hp->m + 1);
            for (int c = 0; c <= hp->m; c++) hp->cp[(size_t)c] = 3 * c;
            finalize_problem(*hp);
            std::vector<double> x0((size_t)hp->n + 2 * (size_t)hp->m);
            init_uniform(iseed[(size_t)i], hp->n, hp->m, 0.5, 1.0, x0.data());
            {
                std::lock_guard<std::mutex> lk(mu);
                reqs[(size_t)i].hp = std::move(hp);
                reqs[(size_t)i].x0 = std::move(x0);
                reqs[(size_t)i].ready = true;
            }
            cv.notify_all();
        }
    };
    const int nt = std::max(1, std::min(gen_threads > 0 ? gen_threads : 4, 32));
    std::vector<std::thread> th;
    for (int k = 0; k < nt; k++) th.emplace_back(worker);
    auto get = [&](int i) {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return reqs[(size_t)i].ready; });
        ServeReq q;
        q.hp = reqs[(size_t)i].hp.get();
        q.x0 = reqs[(size_t)i].x0.data();
        dnx_params_default(q.hp->m, 10000, &q.prm);
        return q;
    };
    auto release = [&](int i) {
        std::lock_guard<std::mutex> lk(mu);
        reqs[(size_t)i].hp.reset();
        std::vector<double>().swap(reqs[(size_t)i].x0);
        released++;
    };
    ServeOut out{status, steps, tts_s, wait_s};
    out.where = where_out;
    int rc = serve_loop(b, count, arrival_out, step_cap, steps_per_launch, deadline_s, get, release, out, makespan_s);
    stop = true;
    for (auto &x : th) x.join();
    return rc;
}


}  // extern "C"

// ---------------------------------------------------------------------------
// Autotuner.  Candidates are timed on synthetic planted instances drawn from
// the caller's size distribution, each in fixed-step mode (every instance
// takes exactly `steps` steps, so every candidate does the same work):
//   1. fused-kernel shapes: workgroup size of modes 0 and 1 and the LDS
//      budget that admits an instance to mode 0, scored by compulsory bytes
//      per second over the whole sample;
//   2. the wide-mode threshold: the k largest instances (k = 1, 4, 16, 64)
//      are stepped narrow and wide; the threshold is the largest total
//      clause count at which wide was faster.
// The winners are applied to `b`, which must be empty.
// ---------------------------------------------------------------------------
namespace dnx {

struct TuneInst {
    HostProblem hp;
    std::vector<double> x0;
};

static int tune_run(dnx_ctx *ctx, const std::vector<TuneInst> &inst, const std::vector<int> &pick, int bs0, int cap0,
                    int bs1, long long wide, int steps, double *ms_out)
{
    size_t bytes = 1 << 20;
    for (int i : pick) {
        const HostProblem &h = inst[(size_t)i].hp;
        bytes += layout_of(h.n, h.m, h.nnz(), !h.uniform3, h.has_unit, KC_MAX).total;
    }
    dnx_batch *t = nullptr;
    int rc = dnx_batch_open(ctx, DNX_FP32, (int)pick.size(), (int64_t)bytes, &t);
    if (rc) return rc;
    struct Close {
        dnx_batch *t;
        ~Close() { dnx_batch_close(t); }
    } closer{t};
    // the fused kernel's shapes alone: no epochs
    if ((rc = dnx_batch_set_shape(t, BM_LDS_ALL, bs0, cap0)) || (rc = dnx_batch_set_shape(t, BM_LDS_V, bs1, 1 << 30)) ||
        (rc = dnx_batch_set_wide(t, wide, 30)) || (rc = dnx_batch_set_epoch(t, 0, 256, 1024)))
        return rc;
    for (int i : pick) {
        dnx_params prm;
        dnx_params_default(inst[(size_t)i].hp.m, 10000, &prm);
        int slot;
        if ((rc = dnx_batch_admit_host(t, inst[(size_t)i].hp, &prm, inst[(size_t)i].x0.data(), steps, 0, &slot)))
            return rc;
    }
    double ms[NBUCKET];
    int64_t nl[NBUCKET];
    if ((rc = dnx_batch_timing(t, 1, nullptr, nullptr, nullptr))) return rc;
    // one launch runs the fixed steps (steps is a multiple of the wide chunk)
    if ((rc = dnx_batch_run(t, steps, 0.0, nullptr))) return rc;
    if ((rc = dnx_batch_timing(t, 0, ms, nl, nullptr))) return rc;
    double tot = 0.0;
    for (int md = 0; md < NBUCKET; md++) tot += ms[md];
    *ms_out = tot;
    return DNX_OK;
}

// Times `steps` fixed steps of the picked instances in gang mode from the
// first step (the tail threshold set past every clause count), with gangs
// asking for `cpw` clauses per workgroup; ms_out is the device time of the
// gang (or, when gangs do not apply, wide) launches.
static int tune_gang_run(dnx_ctx *ctx, const std::vector<TuneInst> &inst, const std::vector<int> &pick, int cpw,
                         int steps, double *ms_out)
{
    size_t bytes = 1 << 20;
    for (int i : pick) {
        const HostProblem &h = inst[(size_t)i].hp;
        bytes += layout_of(h.n, h.m, h.nnz(), !h.uniform3, h.has_unit, KC_MAX).total;
    }
    dnx_batch *t = nullptr;
    int rc = dnx_batch_open(ctx, DNX_FP32, (int)pick.size(), (int64_t)bytes, &t);
    if (rc) return rc;
    struct Close {
        dnx_batch *t;
        ~Close() { dnx_batch_close(t); }
    } closer{t};
    if ((rc = dnx_batch_set_gang(t, 256, cpw, steps)) ||
        (rc = dnx_batch_set_wide(t, std::numeric_limits<int64_t>::max(), 30)) || (rc = dnx_batch_set_epoch(t, 0, 256, 1024)))
        return rc;
    for (int i : pick) {
        dnx_params prm;
        dnx_params_default(inst[(size_t)i].hp.m, 10000, &prm);
        int slot;
        if ((rc = dnx_batch_admit_host(t, inst[(size_t)i].hp, &prm, inst[(size_t)i].x0.data(), steps, 0, &slot)))
            return rc;
    }
    double ms[NBUCKET];
    int64_t nl[NBUCKET];
    if ((rc = dnx_batch_timing(t, 1, nullptr, nullptr, nullptr))) return rc;
    if ((rc = dnx_batch_run(t, steps, 0.0, nullptr))) return rc;
    if ((rc = dnx_batch_timing(t, 0, ms, nl, nullptr))) return rc;
    double tot = 0.0;
    for (int md = 0; md < NBUCKET; md++) tot += ms[md];
    *ms_out = tot;
    return DNX_OK;
}

}  // namespace dnx

extern "C" {

int dnx_batch_get_gang(dnx_batch *b, int *max_blocks, int *clauses_per_workgroup, int *steps_per_launch)
{
    if (!b) return DNX_E_ARG;
    if (max_blocks) *max_blocks = b->gang_blocks;
    if (clauses_per_workgroup) *clauses_per_workgroup = b->gang_cpw;
    if (steps_per_launch) *steps_per_launch = b->gang_steps;
    return DNX_OK;
}

int dnx_batch_tune(dnx_batch *b, int n_min, int n_max, double r_min, double r_max, int sample, int steps,
                   uint64_t seed, double *report, int report_len, int *report_count)
{
    if (!b || n_min < 3 || n_max < n_min || r_min <= 0 || r_max < r_min || sample <= 0 || steps <= 0)
        return DNX_E_ARG;
    for (auto &sl : b->slots)
        if (sl.used) {
            set_error("tune an empty batch");
            return DNX_E_STATE;
        }
    steps = (steps + 29) / 30 * 30;
    // the sample
    SplitMix64 rng(seed);
    std::vector<TuneInst> inst((size_t)sample);
    for (auto &ti : inst) {
        const double ln = std::log((double)n_min) + rng.unit() * (std::log((double)n_max) - std::log((double)n_min));
        const int n = std::max(n_min, std::min(n_max, (int)std::lround(std::exp(ln))));
        const int m = (int)std::lround((r_min + rng.unit() * (r_max - r_min)) * n);
        const uint64_t s = rng.next();
        ti.hp.n = n;
        ti.hp.m = m;
        ti.hp.lits.resize((size_t)m * 3);
        generate_planted(s, n, m, 0.46, 0.46, ti.hp.lits.data(), nullptr);
        ti.hp.cp.resize((size_t)m + 1);
        for (int c = 0; c <= m; c++) ti.hp.cp[(size_t)c] = 3 * c;
        finalize_problem(ti.hp);
        ti.x0.resize((size_t)n + 2 * (size_t)m);
        init_uniform(s, n, m, 0.5, 1.0, ti.x0.data());
    }
    double alg = 0.0;
    for (auto &ti : inst) alg += (double)steps * (28.0 * ti.hp.m + 8.0 * ti.hp.n);
    std::vector<int> all((size_t)sample);
    for (int i = 0; i < sample; i++) all[(size_t)i] = i;
    int nrep = 0;
    auto rep = [&](double a, double c, double d, double e, double f) {
        if (report && nrep + 5 <= report_len) {
            report[nrep] = a;
            report[nrep + 1] = c;
            report[nrep + 2] = d;
            report[nrep + 3] = e;
            report[nrep + 4] = f;
        }
        nrep += 5;
    };
    // 1. fused-kernel shapes
    const int bss[3] = {256, 512, 1024};
    const int caps[4] = {0, 48 * 1024, 96 * 1024, b->lds_max};
    double best = -1.0;
    int b0 = b->shape[BM_LDS_ALL].bs, c0 = b->shape[BM_LDS_ALL].lds_cap, b1 = b->shape[BM_LDS_V].bs;
    for (int cap : caps) {
        if (cap > b->lds_max) continue;
        for (int bs : bss) {
            double ms = 0.0;
            int rc = tune_run(b->ctx, inst, all, bs, cap, bs, 0, steps, &ms);
            if (rc) return rc;
            const double gbps = alg / (ms * 1e-3) / 1e9;
            rep(0, bs, cap, ms, gbps);
            if (gbps > best) {
                best = gbps;
                b0 = bs;
                c0 = cap;
                b1 = bs;
            }
        }
    }
    // mode-1 workgroup size on its own, with the chosen mode-0 shape
    for (int bs : bss) {
        if (bs == b1) continue;
        double ms = 0.0;
        int rc = tune_run(b->ctx, inst, all, b0, c0, bs, 0, steps, &ms);
        if (rc) return rc;
        const double gbps = alg / (ms * 1e-3) / 1e9;
        rep(1, bs, c0, ms, gbps);
        if (gbps > best) {
            best = gbps;
            b1 = bs;
        }
    }
    b->shape[BM_LDS_ALL].bs = b0;
    b->shape[BM_LDS_ALL].lds_cap = c0;
    b->shape[BM_LDS_V].bs = b1;
    // 2. the tail's shape: clauses per gang workgroup, timed on the sample's
    // eight largest instances advanced in gangs from their first step
    if (b->gang_blocks > 0) {
        std::vector<int> big = all;
        std::stable_sort(big.begin(), big.end(),
                         [&](int x, int y) { return inst[(size_t)x].hp.m > inst[(size_t)y].hp.m; });
        big.resize(std::min<size_t>(big.size(), 8));
        double gbest = -1.0;
        int cbest = b->gang_cpw;
        double galg = 0.0;
        for (int i : big) galg += (double)steps * (28.0 * inst[(size_t)i].hp.m + 8.0 * inst[(size_t)i].hp.n);
        for (int cpw : {256, 384, 768}) {
            double ms = 0.0;
            int rc = tune_gang_run(b->ctx, inst, big, cpw, steps, &ms);
            if (rc) return rc;
            const double gbps = ms > 0 ? galg / (ms * 1e-3) / 1e9 : 0.0;
            rep(2, cpw, 0, ms, gbps);
            if (gbps > gbest) {
                gbest = gbps;
                cbest = cpw;
            }
        }
        b->gang_cpw = cbest;
    }
    if (report_count) *report_count = nrep;
    return DNX_OK;
}

int dnx_batch_get_shape(dnx_batch *b, int mode, int *block, int *lds_cap, int64_t *wide_clauses)
{
    if (!b || mode < 0 || mode >= BM_COUNT) return DNX_E_ARG;
    if (block) *block = b->shape[mode].bs;
    if (lds_cap) *lds_cap = b->shape[mode].lds_cap;
    if (wide_clauses) *wide_clauses = b->wide_clauses;
    return DNX_OK;
}

}  // extern "C"

namespace dnx {

// FP32 portfolio: the `threads` copies are slots of one batch, advanced
// together (wide mode at every size, so a large instance still spreads over
// the whole device); the run ends at the first status sync at which a copy
// is satisfied, and the winner is the satisfied copy with the fewest steps
// (lowest thread index on a tie).
int portfolio_batched(dnx_problem *p, const dnx_params *prm, int seed, int threads, int64_t step_cap,
                      double walltime_s, int8_t *assign_out, int64_t *steps_out, double *t_out, int *status,
                      int *winner)
{
    const HostProblem &h = p->hp;
    const Layout L = layout_of(h.n, h.m, h.nnz(), !h.uniform3, h.has_unit, KC_MAX);
    dnx_batch *b = nullptr;
    int rc = dnx_batch_open(p->ctx, DNX_FP32, threads, (int64_t)((L.total + 4096) * (size_t)threads + (1 << 20)), &b);
    if (rc) return rc;
    struct Close {
        dnx_batch *b;
        ~Close() { dnx_batch_close(b); }
    } guard{b};
    b->group_stop = true;
    b->wide_clauses = std::numeric_limits<long long>::max();
    std::vector<int> slot((size_t)threads);
    std::vector<double> x((size_t)h.n + 2 * (size_t)h.m + 1);
    for (int i = 0; i < threads; i++) {
        init_reference_thread(h, seed, i, x.data());
        dnx_params pi = *prm;
        if (i > 0) pi.alpha = prm->alpha / i;  // dynex.cc:2657
        if ((rc = dnx_batch_admit_host(b, h, &pi, x.data(), step_cap, 1, &slot[(size_t)i]))) return rc;
    }
    int remaining = 0;
    if ((rc = dnx_batch_run(b, 32, walltime_s, &remaining))) return rc;
    std::vector<int> st((size_t)threads);
    std::vector<int64_t> steps((size_t)threads);
    std::vector<double> t((size_t)threads);
    const BDyn *d = b->h_dyn;
    int best = -1;
    for (int i = 0; i < threads; i++) {
        const BDyn &q = d[slot[(size_t)i]];
        if (q.status == DNX_SAT && (best < 0 || q.steps < d[slot[(size_t)best]].steps)) best = i;
    }
    *status = DNX_UNKNOWN;
    if (winner) *winner = best;
    const int show = best >= 0 ? best : 0;
    if (steps_out) *steps_out = d[slot[(size_t)show]].steps;
    if (t_out) *t_out = d[slot[(size_t)show]].t;
    if (best < 0) return DNX_OK;
    int unsat = 0;
    std::vector<int8_t> a((size_t)(h.n ? h.n : 1));
    if ((rc = dnx_batch_assignment(b, slot[(size_t)best], a.data(), &unsat))) return rc;
    if (unsat) {
        set_error("portfolio: the winning copy's assignment does not verify");
        *status = DNX_FAILED;
        return DNX_OK;
    }
    if (assign_out) std::copy(a.begin(), a.begin() + h.n, assign_out);
    *status = DNX_SAT;
    return DNX_OK;
}

}  // namespace dnx
