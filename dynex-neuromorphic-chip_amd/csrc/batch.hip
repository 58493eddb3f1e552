// batch.hip -- the fused step kernel of the batched engine (FP32).
//
// One workgroup owns one instance at a time and advances it up to K adaptive
// steps entirely inside the kernel: clause pass, workgroup barrier, variable
// pass, workgroup max of the ladder level, barrier, next step.  No grid-wide
// synchronisation exists, so thousands of uneven instances run side by side
// and each keeps its own step size, time and step count.  Workgroups take
// instances from a list (largest first) through one atomic ticket per
// instance.  The arithmetic is the single-instance FP32 arithmetic
// (dmm_math.hpp) with the same clause-order fold, so an instance computes the
// same bits alone, in a batch of any composition, and on the single-instance
// path.
#include <hip/hip_runtime.h>

#include "batch.hpp"
#include "dmm_math.hpp"

#pragma clang fp contract(off)

namespace dnx {

// The descriptor's pointers arrive as generic pointers (they are loaded from
// device memory, not from the kernel arguments), so the compiler would access
// them with FLAT instructions, which count against the LDS counter as well
// and make every LDS wait also wait for device-memory traffic.  They all
// point into the device arena: the kernels read the descriptor through this
// view, whose pointers are typed as global memory.
#define GLB __attribute__((address_space(1)))
struct GInst {
    int n, m;
    int mode;
    int flags;
    GLB const int32_t *lits;
    GLB const int32_t *cp;
    GLB const int32_t *lpos;
    GLB const uint32_t *pk;
    GLB const int32_t *gpos;
    GLB const int32_t *eidx;
    GLB const int32_t *row_ptr;
    GLB const int8_t *unit;
    GLB float *xs, *xl;
    GLB unsigned long long *vd;  // (V, dV) pairs as 8-byte words
    GLB float *e;
    GLB const int32_t *cbound;
    GLB const int32_t *cbase;
    GLB float *cst;
    int nchunk;
    float alpha, beta, gamma, delta, eps, zeta, xlmax;
};
__device__ __forceinline__ GInst global_view(const BInst &B)
{
    GInst I;
    I.n = B.n;
    I.m = B.m;
    I.mode = B.mode;
    I.flags = B.flags;
    I.lits = (GLB const int32_t *)B.lits;
    I.cp = (GLB const int32_t *)B.cp;
    I.lpos = (GLB const int32_t *)B.lpos;
    I.pk = (GLB const uint32_t *)B.pk;
    I.gpos = (GLB const int32_t *)B.gpos;
    I.eidx = (GLB const int32_t *)B.eidx;
    I.row_ptr = (GLB const int32_t *)B.row_ptr;
    I.unit = (GLB const int8_t *)B.unit;
    I.xs = (GLB float *)B.xs;
    I.xl = (GLB float *)B.xl;
    I.vd = (GLB unsigned long long *)B.vd;
    I.e = (GLB float *)B.e;
    I.cbound = (GLB const int32_t *)B.cbound;
    I.cbase = (GLB const int32_t *)B.cbase;
    I.cst = (GLB float *)B.cst;
    I.nchunk = B.nchunk;
    I.alpha = B.alpha;
    I.beta = B.beta;
    I.gamma = B.gamma;
    I.delta = B.delta;
    I.eps = B.eps;
    I.zeta = B.zeta;
    I.xlmax = B.xlmax;
    return I;
}
// one (V, dV) pair of the view: V in the low word, dV in the high word
__device__ __forceinline__ BF2 vdr(const GInst &I, int i)
{
    const unsigned long long u = I.vd[i];
    BF2 r;
    r.v = __uint_as_float((unsigned)u);
    r.d = __uint_as_float((unsigned)(u >> 32));
    return r;
}
__device__ __forceinline__ void vdw(const GInst &I, int i, BF2 o)
{
    I.vd[i] = (unsigned long long)__float_as_uint(o.v) | ((unsigned long long)__float_as_uint(o.d) << 32);
}

// Four consecutive clauses 4g .. 4g+3 of a mode-0 3-SAT instance in the
// steady state: packed literal words, memories, stored C.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
struct CGroup {
    uint32_t W[4][3];
    float XS[4], XL[4], CX[4];
};

// Group g's clauses inside [c_lo, c_hi): a whole group moves as 16-byte
// accesses at one 32-bit offset from each array's base (every array starts
// 256-byte aligned), which leaves almost no address arithmetic; a group cut
// by the range's ends takes its clauses one at a time (zeros outside).
__device__ __forceinline__ void load_group(const GInst &I, int g, int c_lo, int c_hi, CGroup &G)
{
    const int c4 = g << 2;
    if (c4 >= c_lo && c4 + 4 <= c_hi) {
        const uint32_t o16 = (uint32_t)g * 16u, o48 = (uint32_t)g * 48u;
        GLB const char *pkb = (GLB const char *)I.pk;
        const u32x4 w0 = *(GLB const u32x4 *)(pkb + o48);
        const u32x4 w1 = *(GLB const u32x4 *)(pkb + o48 + 16u);
        const u32x4 w2 = *(GLB const u32x4 *)(pkb + o48 + 32u);
        const f32x4 a = *(GLB const f32x4 *)((GLB const char *)I.xs + o16);
        const f32x4 b = *(GLB const f32x4 *)((GLB const char *)I.xl + o16);
        const f32x4 q = *(GLB const f32x4 *)((GLB const char *)I.cst + o16);
        G.W[0][0] = w0.x;
        G.W[0][1] = w0.y;
        G.W[0][2] = w0.z;
        G.W[1][0] = w0.w;
        G.W[1][1] = w1.x;
        G.W[1][2] = w1.y;
        G.W[2][0] = w1.z;
        G.W[2][1] = w1.w;
        G.W[2][2] = w2.x;
        G.W[3][0] = w2.y;
        G.W[3][1] = w2.z;
        G.W[3][2] = w2.w;
        G.XS[0] = a.x;
        G.XS[1] = a.y;
        G.XS[2] = a.z;
        G.XS[3] = a.w;
        G.XL[0] = b.x;
        G.XL[1] = b.y;
        G.XL[2] = b.z;
        G.XL[3] = b.w;
        G.CX[0] = q.x;
        G.CX[1] = q.y;
        G.CX[2] = q.z;
        G.CX[3] = q.w;
    } else {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int c = c4 + u;
            G.W[u][0] = G.W[u][1] = G.W[u][2] = 0u;
            G.XS[u] = G.XL[u] = G.CX[u] = 0.0f;
            if (c >= c_lo && c < c_hi) {
                G.W[u][0] = I.pk[3 * c];
                G.W[u][1] = I.pk[3 * c + 1];
                G.W[u][2] = I.pk[3 * c + 2];
                G.XS[u] = I.xs[c];
                G.XL[u] = I.xl[c];
                G.CX[u] = I.cst[c];
            }
        }
    }
}

// Writes back group g's memories and stored C inside [c_lo, c_hi).
__device__ __forceinline__ void store_group(const GInst &I, int g, int c_lo, int c_hi, const CGroup &G)
{
    const int c4 = g << 2;
    if (c4 >= c_lo && c4 + 4 <= c_hi) {
        const uint32_t o16 = (uint32_t)g * 16u;
        *(GLB f32x4 *)((GLB char *)I.xs + o16) = f32x4{G.XS[0], G.XS[1], G.XS[2], G.XS[3]};
        *(GLB f32x4 *)((GLB char *)I.xl + o16) = f32x4{G.XL[0], G.XL[1], G.XL[2], G.XL[3]};
        *(GLB f32x4 *)((GLB char *)I.cst + o16) = f32x4{G.CX[0], G.CX[1], G.CX[2], G.CX[3]};
    } else {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int c = c4 + u;
            if (c >= c_lo && c < c_hi) {
                I.xs[c] = G.XS[u];
                I.xl[c] = G.XL[u];
                I.cst[c] = G.CX[u];
            }
        }
    }
}

// One workgroup's work loop: takes instances from the list and advances each
// up to K steps (or, in a flow, until it ends or the flow stops).
template <int MODE, int BS>
__device__ __forceinline__ void flow_worker(const BInst *__restrict__ insts, BDyn *__restrict__ dyn,
                                            const int *__restrict__ work, int nwork, int *__restrict__ head,
                                            int K, BFlow *__restrict__ flow)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_item;
    __shared__ int s_lev[2];
    __shared__ int s_cb[KC_MAX + 1];  // mode 0: chunk bounds of the current instance
    __shared__ int s_stop;             // flow: the stop flag as read in this step
    const int tid = threadIdx.x;
    for (;;) {
        __syncthreads();  // previous item finished with LDS and s_item
        if (tid == 0) {
            const bool stopped = flow && __hip_atomic_load(&flow->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int item = stopped ? nwork : atomicAdd(head, 1);
            // a workgroup that finds the list empty while the flow runs is
            // idle from now on: once idle_trigger of them are, the flow stops
            // (the idle CUs can then serve the running instances as gangs)
            if (item >= nwork && !stopped && flow && flow->idle_trigger > 0 &&
                __hip_atomic_fetch_add(&flow->idle, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 >=
                    flow->idle_trigger)
                __hip_atomic_store(&flow->stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_item = item;
            s_lev[0] = 0;
            s_lev[1] = 0;
        }
        __syncthreads();
        const int it = s_item;
        if (it >= nwork) return;
        const int slot = __builtin_amdgcn_readfirstlane(work[it]);
        BDyn d = dyn[slot];
        if (d.status != DNX_RUNNING) continue;
        const GInst I = global_view(insts[slot]);
        const int n = I.n, m = I.m;
        const bool general = (I.flags & BF_GENERAL) != 0;
        PT<float> p;
        p.alpha = I.alpha;
        p.beta = I.beta;
        p.gamma = I.gamma;
        p.delta = I.delta;
        p.eps = I.eps;
        p.zeta = I.zeta;
        p.xlmax = I.xlmax;

        // Modes 0 and 1 keep V of every variable in LDS (lv) and beside it
        // one float per variable (ld): with several chunks (mode 0) the
        // variable's partial fold between chunks, and after the variable pass
        // its new dV, which puts V of the next state in place and is written
        // back with V, as the pending form, when the launch ends.  The dV of
        // the state a launch starts from stays in device memory, which this
        // kernel does not write before the launch ends: only a launch's first
        // step (V of the state not yet in place) reads it, through pvr.  Mode
        // 2 keeps the (V, dV) pairs in device memory.
        float *lv = (float *)smem, *ld = (float *)(smem + lds_var_bytes(n) / 2);
        auto pvr = [&](int i) -> BF2 { return vdr(I, i); };
        auto pvv = [&](int i) -> float {
            if constexpr (MODE == BM_GLOBAL) return vdr(I, i).v;
            else return lv[i];
        };
        // mode 0 keeps the contributions of one chunk of consecutive clauses
        // at a time in LDS (one chunk when they all fit)
        const int nch = MODE == BM_LDS_ALL ? I.nchunk : 1;
        constexpr int UV = 2;  // chunk folds: variables whose slot ranges one thread loads at once
        float *se = (float *)(smem + lds_var_bytes(n));
        if (MODE == BM_LDS_ALL)
            for (int i = tid; i <= nch; i += BS) s_cb[i] = I.cbound[i];
        if (MODE != BM_GLOBAL) {
            for (int v = tid; v < n; v += BS) lv[v] = vdr(I, v).v;
            __syncthreads();
        }
        // the variable pass's slot ranges (the last chunk's with several
        // chunks) and unit pins of each thread's first two variables never
        // change: they are read once per instance, and instances of up to
        // 2 BS variables keep them in registers
        auto rowp = MODE == BM_LDS_ALL ? I.cbase + (size_t)(nch - 1) * (n + 1) : I.row_ptr;
        int RB0 = 0, RE0 = 0, RB1 = 0, RE1 = 0, PIN01 = 0;
        if (tid < n) {
            RB0 = rowp[tid];
            RE0 = rowp[tid + 1];
            PIN01 |= (I.flags & BF_UNIT) && I.unit[tid] != 0;
        }
        if (tid + BS < n) {
            RB1 = rowp[tid + BS];
            RE1 = rowp[tid + BS + 1];
            PIN01 |= ((I.flags & BF_UNIT) && I.unit[tid + BS] != 0) << 1;
        }
        bool pend = d.plevel >= 0;
        // mat: V of the current state is already in place (every step of a
        // launch after its first), and cst holds each clause's C of the
        // previous state, so the clause pass needs neither h*dV nor a second
        // evaluation of C for the memory update
        bool mat = false;
        int plev = d.plevel;
        long long steps = d.steps;
        double t = d.t;
        int done = 0;
        for (int k = 0; k < K; k++) {
            // flow: the stop flag is requested now and judged at the end of
            // the step, so its load travels during the step
            const int fl = flow && tid == 0 ? __hip_atomic_load(&flow->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
            const float h = pend ? level_h<float>(plev) : 0.0f;
            if (pend) {
                steps += 1;
                t += level_dt(plev);
            }
            // ---- clause pass --------------------------------------------
            // steady: V of this state in place, and the memories of the state
            // being updated clamped (it is not the initial state)
            const bool steady = mat && steps >= 2;
            int unsolved = 0;
            auto prevV = [&](int v) { return pvr(v).v; };
            auto curV = [&](int v) {
                if (mat) return pvv(v);
                BF2 a = pvr(v);
                return pend ? clampV(a.v + h * a.d) : a.v;
            };
            // mode 0, 3-SAT steady state: a thread's first clause group of
            // the next chunk, requested before the current chunk's fold
            CGroup pre;
            bool have_pre = false;
            for (int ch = 0; ch < nch; ch++) {
                const int c_lo = MODE == BM_LDS_ALL ? s_cb[ch] : 0;
                const int c_hi = MODE == BM_LDS_ALL ? s_cb[ch + 1] : m;
                if (MODE == BM_LDS_ALL && !general && steady) {
                    // 3-SAT in the steady state, mode 0: V in place and
                    // clamped, the memories clamped by the previous update,
                    // C of the previous state stored (the steady-state forms
                    // return the same bits as the general path below).  One
                    // word per literal packs its sign, variable and LDS slot.
                    // (Buffer-resource addressing of the device arrays
                    // measured slower than these global accesses.)
                    // C >= 1/2 anywhere means unsolved.  C lies in [+0, 1]
                    // here, where the order of the bit patterns as unsigned
                    // ints is that of the values: an integer max (a float max
                    // would first quiet its running operand every clause).
                    uint32_t cmaxb = 0u;
                    // V's LDS base and each extracted variable as values the
                    // compiler cannot fold, so that a literal's V address is
                    // a bit-field extract and one shift-and-add (folded, the
                    // compiler makes it a shift, a mask and an add)
                    typedef __attribute__((address_space(3))) const float LdsF;
                    uint32_t lvb = (uint32_t)(uintptr_t)(LdsF *)lv;
                    asm volatile("" : "+s"(lvb));  // no instruction: hides the constant
                    auto vat = [&](uint32_t w) {
                        uint32_t v = __builtin_amdgcn_ubfe(w, 16, 15);
                        asm volatile("" : "+v"(v));  // keeps the extract one instruction
                        return *(LdsF *)(uintptr_t)(lvb + (v << 2));
                    };
                    // Each thread takes four consecutive clauses 4g .. 4g+3 at
                    // a time (load_group / store_group).
                    const int g_end = (c_hi + 3) >> 2;
                    bool first = true;
                    for (int g = (c_lo >> 2) + tid; g < g_end; g += BS) {
                        const int c4 = g << 2;
                        CGroup G;
                        if (first && have_pre)
                            G = pre;
                        else
                            load_group(I, g, c_lo, c_hi, G);
                        first = false;
                        auto one = [&](int u) {
                            const int32_t l0 = (int32_t)G.W[u][0], l1 = (int32_t)G.W[u][1], l2 = (int32_t)G.W[u][2];
                            float xs = G.XS[u], xl = G.XL[u];
                            const float V0 = vat(G.W[u][0]), V1 = vat(G.W[u][1]), V2 = vat(G.W[u][2]);
                            memory_update_steady(p, G.CX[u], h, xs, xl);
                            G.XS[u] = xs;
                            G.XL[u] = xl;
                            auto emit = [&](int j, float val) { se[G.W[u][j] & 0xffffu] = val; };
                            const float C = clause3_steady(p, l0, l1, l2, V0, V1, V2, xs, xl, emit);
                            G.CX[u] = C;
                            cmaxb = max(cmaxb, __float_as_uint(C));
                        };
                        // 3-SAT chunks start at multiples of four clauses:
                        // only an instance's last group can be partial
                        if (c4 >= c_lo && c4 + 4 <= c_hi) {
#pragma unroll
                            for (int u = 0; u < 4; u++) one(u);
                        } else {
#pragma unroll
                            for (int u = 0; u < 4; u++)
                                if (c4 + u >= c_lo && c4 + u < c_hi) one(u);
                        }
                        store_group(I, g, c_lo, c_hi, G);
                    }
                    // the next chunk's first group of this thread, requested
                    // before this chunk's fold so that its loads travel
                    // during the fold (chunk ch does not touch the next
                    // chunk's clauses)
                    have_pre = false;
                    if (ch + 1 < nch) {
                        const int lo1 = s_cb[ch + 1], hi1 = s_cb[ch + 2];
                        const int g1 = (lo1 >> 2) + tid;
                        if (g1 < ((hi1 + 3) >> 2)) {
                            load_group(I, g1, lo1, hi1, pre);
                            have_pre = true;
                        }
                    }
                    unsolved |= cmaxb >= 0x3f000000u;  // 0.5f
                } else if (!general && steady) {
                    // the same for modes 1 and 2 (signed literals,
                    // contributions in device memory)
                    constexpr int U = 4;
                    for (int cb = c_lo + tid; cb < c_hi; cb += U * BS) {
                        int32_t L[U][3];
                        float XS[U], XL[U], CX[U];
#pragma unroll
                        for (int u = 0; u < U; u++) {
                            const int c = cb + u * BS;
                            if (c < c_hi) {
                                L[u][0] = I.lits[3 * c];
                                L[u][1] = I.lits[3 * c + 1];
                                L[u][2] = I.lits[3 * c + 2];
                                XS[u] = I.xs[c];
                                XL[u] = I.xl[c];
                                CX[u] = I.cst[c];
                            }
                        }
#pragma unroll
                        for (int u = 0; u < U; u++) {
                            const int c = cb + u * BS;
                            if (c >= c_hi) break;
                            const int32_t l0 = L[u][0], l1 = L[u][1], l2 = L[u][2];
                            float xs = XS[u], xl = XL[u];
                            const float V0 = pvv(abs(l0) - 1), V1 = pvv(abs(l1) - 1), V2 = pvv(abs(l2) - 1);
                            memory_update_steady(p, CX[u], h, xs, xl);
                            I.xs[c] = xs;
                            I.xl[c] = xl;
                            auto ep = I.e + 3 * c;
                            auto emit = [&](int j, float val) { ep[j] = val; };
                            const float C = clause3_steady(p, l0, l1, l2, V0, V1, V2, xs, xl, emit);
                            I.cst[c] = C;
                            unsolved |= !(C < 0.5f);
                        }
                    }
                } else if (!general) {
                    // 3-SAT outside the steady state (a launch's first steps):
                    // each thread takes its clauses two at a time, with both
                    // clauses' literals, memories, stored C (and LDS slots)
                    // loaded before either is evaluated.
                    constexpr int U = 2;
                    for (int cb = c_lo + tid; cb < c_hi; cb += U * BS) {
                        int32_t L[U][3], P[U][3];
                        float XS[U], XL[U], CX[U];
#pragma unroll
                        for (int u = 0; u < U; u++) {
                            const int c = cb + u * BS;
                            if (c < c_hi) {
                                L[u][0] = I.lits[3 * c];
                                L[u][1] = I.lits[3 * c + 1];
                                L[u][2] = I.lits[3 * c + 2];
                                XS[u] = I.xs[c];
                                XL[u] = I.xl[c];
                                if (mat) CX[u] = I.cst[c];
                                if (MODE == BM_LDS_ALL) {
                                    P[u][0] = I.lpos[3 * c];
                                    P[u][1] = I.lpos[3 * c + 1];
                                    P[u][2] = I.lpos[3 * c + 2];
                                }
                            }
                        }
#pragma unroll
                        for (int u = 0; u < U; u++) {
                            const int c = cb + u * BS;
                            if (c >= c_hi) break;
                            const int32_t l0 = L[u][0], l1 = L[u][1], l2 = L[u][2];
                            float xs = XS[u], xl = XL[u];
                            float r0, r1, r2;
                            if (mat) {
                                // V of this state is in place; the previous
                                // state's C was stored by the previous pass
                                r0 = pvv(abs(l0) - 1);
                                r1 = pvv(abs(l1) - 1);
                                r2 = pvv(abs(l2) - 1);
                                memory_update(p, CX[u], h, xs, xl);
                                I.xs[c] = xs;
                                I.xl[c] = xl;
                            } else {
                                // one (V, dV) pair read per literal serves both
                                // the memory update and the clause
                                const BF2 a0 = pvr(abs(l0) - 1), a1 = pvr(abs(l1) - 1), a2 = pvr(abs(l2) - 1);
                                r0 = a0.v;
                                r1 = a1.v;
                                r2 = a2.v;
                                if (pend) {
                                    memory_update(p, clause3_C_vals<float>(l0, l1, l2, r0, r1, r2), h, xs, xl);
                                    I.xs[c] = xs;
                                    I.xl[c] = xl;
                                    r0 = clampV(a0.v + h * a0.d);
                                    r1 = clampV(a1.v + h * a1.d);
                                    r2 = clampV(a2.v + h * a2.d);
                                }
                            }
                            const float cxs = clamp01(xs), cxl = clampXl(xl, p.xlmax);
                            bool sat;
                            float C;
                            if (MODE == BM_LDS_ALL) {
                                auto emit = [&](int j, float val) { se[P[u][j]] = val; };
                                C = clause3_vals<float>(p, l0, l1, l2, r0, r1, r2, cxs, cxl, emit, sat);
                            } else {
                                auto ep = I.e + 3 * c;
                                auto emit = [&](int j, float val) { ep[j] = val; };
                                C = clause3_vals<float>(p, l0, l1, l2, r0, r1, r2, cxs, cxl, emit, sat);
                            }
                            I.cst[c] = C;
                            unsolved |= !(C < 0.5f) || !sat;
                        }
                    }
                } else {
                    for (int c = c_lo + tid; c < c_hi; c += BS) {
                        const int b = I.cp[c];
                        const int kk = I.cp[c + 1] - b;
                        auto cl = I.lits + b;
                        float xs = I.xs[c], xl = I.xl[c];
                        if (pend) {
                            float Cp = mat ? I.cst[c] : clauseK_C<float>(cl, kk, prevV);
                            memory_update(p, Cp, h, xs, xl);
                            I.xs[c] = xs;
                            I.xl[c] = xl;
                        }
                        const float cxs = clamp01(xs), cxl = clampXl(xl, p.xlmax);
                        bool sat;
                        float C;
                        if (MODE == BM_LDS_ALL) {
                            auto lp = I.lpos + b;
                            auto emit = [&](int j, float val) { se[lp[j]] = val; };
                            C = clauseK<float>(p, cl, kk, cxs, cxl, curV, emit, sat);
                        } else {
                            auto ep = I.e + b;
                            auto emit = [&](int j, float val) { ep[j] = val; };
                            C = clauseK<float>(p, cl, kk, cxs, cxl, curV, emit, sat);
                        }
                        I.cst[c] = C;
                        unsolved |= !(C < 0.5f) || !sat;
                    }
                }
                if (ch + 1 < nch) {
                    // fold this chunk's contributions (clause order within each
                    // variable's slots), then free se for the next chunk.  Each
                    // thread's next UV variables' slot ranges are loaded
                    // together, the first ones before the barrier so that the
                    // loads travel while the other waves finish the chunk.
                    auto cbase = I.cbase + (size_t)ch * (n + 1);
                    int RB[UV], RE[UV];
                    auto rows = [&](int vb) {
#pragma unroll
                        for (int u = 0; u < UV; u++) {
                            const int v = vb + u * BS;
                            RB[u] = v < n ? cbase[v] : 0;
                            RE[u] = v < n ? cbase[v + 1] : 0;
                        }
                    };
                    rows(tid);
                    __syncthreads();
                    for (int vb = tid; vb < n; vb += UV * BS) {
                        if (vb != tid) rows(vb);
#pragma unroll
                        for (int u = 0; u < UV; u++) {
                            const int v = vb + u * BS;
                            if (v >= n) break;
                            ld[v] = lds_fold(se, RB[u], RE[u], ch ? ld[v] : 0.0f);
                        }
                    }
                    __syncthreads();
                }
            }
            const bool any_unsolved = __syncthreads_or(unsolved);
            if (d.stop_on_solve && !any_unsolved)
                done = DNX_SAT;
            else if (steps >= d.cap)
                done = DNX_UNKNOWN;
            // ---- variable pass ------------------------------------------
            // (one variable at a time per thread: batching the slot-range
            // loads of several variables measured slower)
            int fmax = 0;
            auto var_one = [&](int v, int rb, int re, bool pinned) {
                const float V = pend && !mat ? curV(v) : pvv(v);
                float dV = 0.0f;
                if (!done) {
                    if (MODE == BM_LDS_ALL) {
                        dV = lds_fold(se, rb, re, nch > 1 ? ld[v] : 0.0f);
                    } else {
                        dV = fold_gather<float>(I.e, I.eidx, rb, re);
                    }
                    if (pinned) dV = 0.0f;  // dynex.cc:306
                    fmax = max(fmax, fail_level(V, dV));
                }
                if constexpr (MODE == BM_GLOBAL) {
                    BF2 o;
                    o.v = V;
                    o.d = dV;
                    vdw(I, v, o);
                } else {
                    lv[v] = V;
                    ld[v] = dV;
                }
            };
            if (n <= 2 * BS) {
                // slot ranges and pins from registers: in mode 0 this pass
                // then touches no device memory and does not wait for the
                // clause pass's stores to complete
                if (tid < n) var_one(tid, RB0, RE0, PIN01 & 1);
                if (tid + BS < n) var_one(tid + BS, RB1, RE1, (PIN01 >> 1) & 1);
            } else {
                for (int v = tid; v < n; v += BS)
                    var_one(v, rowp[v], rowp[v + 1], (I.flags & BF_UNIT) && I.unit[v] != 0);
            }
            if (done) break;
            const int wm = wave_max(fmax);
            if ((tid & 63) == 0 && wm > 0) atomicMax(&s_lev[k & 1], wm);
            if (tid == 0) s_stop = fl;
            __syncthreads();
            plev = s_lev[k & 1];
            pend = true;
            if (tid == 0) s_lev[(k + 1) & 1] = 0;
            if (flow && s_stop) break;
            if (k + 1 < K) {
                // another step follows in this launch: put V of the next
                // state in place (the last step of a launch leaves the pending
                // form, V and dV with the level, which every path resumes from)
                const float hn = level_h<float>(plev);
                for (int v = tid; v < n; v += BS) {
                    if constexpr (MODE == BM_GLOBAL) {
                        BF2 a = vdr(I, v);
                        a.v = clampV(a.v + hn * a.d);
                        vdw(I, v, a);
                    } else {
                        lv[v] = clampV(lv[v] + hn * ld[v]);
                    }
                }
                __syncthreads();
                mat = true;
            }
        }
        __syncthreads();
        if (MODE != BM_GLOBAL)
            for (int v = tid; v < n; v += BS) {
                BF2 o;
                o.v = lv[v];
                o.d = ld[v];
                vdw(I, v, o);
            }
        if (tid == 0) {
            BDyn o = d;
            o.steps = steps;
            o.t = t;
            if (done) {
                o.status = done;
                o.plevel = -1;
                o.t_done = wall_clock64();
            } else {
                o.plevel = plev;
            }
            dyn[slot] = o;
            if (flow && done) {
                const long long left =
                    __hip_atomic_fetch_add(&flow->live, -(long long)m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - m;
                if (left <= flow->threshold)
                    __hip_atomic_store(&flow->stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

template <int MODE, int BS>
__global__ __launch_bounds__(BS) void k_batch(const BInst *__restrict__ insts, BDyn *__restrict__ dyn,
                                              const int *__restrict__ work, int nwork, int *__restrict__ head,
                                              int K, BFlow *__restrict__ flow)
{
    flow_worker<MODE, BS>(insts, dyn, work, nwork, head, K, flow);
}

// ---------------------------------------------------------------------------
// Wide mode.  When few instances remain, one workgroup per instance leaves
// most of the chip idle and each instance latency-bound.  Wide mode advances
// the remaining instances together with one clause kernel and one variable
// kernel per step, spread over the union of their clauses and variables: a
// block map names each block's instance and its local block index.  Each
// instance keeps its own control block (the single-instance Ctl protocol,
// indexed by a rotation r shared by all) and so its own h, t and step
// count; the arithmetic is the single-instance FP32 arithmetic.
// ---------------------------------------------------------------------------
constexpr int WBS = 256;
constexpr int WCAPW = 2048;

// BDyn -> Ctl for the listed slots; pending levels go to rotation (r0+2)%3.
__global__ __launch_bounds__(256) void k_wide_enter(const BDyn *__restrict__ dyn, Ctl *__restrict__ wctl,
                                                    const int *__restrict__ list, int count, int r0)
{
    const int i = blockIdx.x;
    if (i >= count) return;
    const int slot = list[i];
    Ctl *c = &wctl[slot];
    int *f = &c->flags[0][0][0];
    for (int j = threadIdx.x; j < 3 * FG_COUNT * FLAG_GROUP; j += blockDim.x) f[j] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        const BDyn d = dyn[slot];
        c->status = d.status;
        c->stop_on_solve = d.stop_on_solve;
        c->steps = d.steps;
        c->cap = d.cap;
        c->t = d.t;
        c->t_done = 0;
        c->track_loc = 0;
        for (int s = 0; s < 3; s++) {
            c->stepped[s] = 0;
            c->done[s] = 0;
        }
        if (d.plevel >= 0) {
            const int ps = (r0 + 2) % 3;
            c->stepped[ps] = 1;
            for (int L = 0; L < d.plevel; L++) c->flags[ps][FG_LEVEL + L][0] = 1;
        }
    }
}

// Ctl -> BDyn after a chunk whose last step used rotation rl.
__global__ __launch_bounds__(64) void k_wide_sync(BDyn *__restrict__ dyn, Ctl *__restrict__ wctl,
                                                  const int *__restrict__ list, int count, int rl)
{
    const int i = blockIdx.x;
    if (i >= count) return;
    const int slot = list[i];
    BDyn o = dyn[slot];
    if (o.status != DNX_RUNNING) return;
    Ctl *c = &wctl[slot];
    const int lev = dev_level(c, rl);
    if (threadIdx.x == 0) {
        int st = c->status;
        if (st == DNX_RUNNING && c->done[rl]) st = c->done[rl];
        o.steps = c->steps;
        o.t = c->t;
        if (st != DNX_RUNNING) {
            c->status = st;
            o.status = st;
            o.plevel = -1;
            o.t_done = c->t_done;
        } else {
            o.plevel = c->stepped[rl] ? lev : -1;
        }
        dyn[slot] = o;
    }
}

__global__ __launch_bounds__(WBS) void k_wide_clause(const BInst *__restrict__ insts, Ctl *__restrict__ wctl,
                                                     const int2 *__restrict__ bmap, int nblk, int r)
{
    const int b = blockIdx.x;
    if (b >= nblk) return;
    const int2 bm = bmap[b];
    const int slot = __builtin_amdgcn_readfirstlane(bm.x), lb = __builtin_amdgcn_readfirstlane(bm.y);
    Ctl *__restrict__ ctl = &wctl[slot];
    if (ctl->status != DNX_RUNNING) return;
    const int ps = (r + 2) % 3;
    const int done = ctl->done[ps];
    if (done) {
        if (lb == 0 && threadIdx.x == 0) ctl->status = done;
        return;
    }
    const GInst I = global_view(insts[slot]);
    const bool pend = ctl->stepped[ps] != 0;
    const int plev = dev_level(ctl, ps);
    const float h = pend ? level_h<float>(plev) : 0.0f;
    if (pend && lb == 0 && threadIdx.x == 0) {
        ctl->t += level_dt(plev);
        ctl->steps += 1;
    }
    PT<float> p;
    p.alpha = I.alpha;
    p.beta = I.beta;
    p.gamma = I.gamma;
    p.delta = I.delta;
    p.eps = I.eps;
    p.zeta = I.zeta;
    p.xlmax = I.xlmax;
    const int c = lb * WBS + threadIdx.x;
    int unsolved = 0;
    if (!(I.flags & BF_GENERAL) && c < I.m) {
        const int32_t l0 = I.lits[3 * c], l1 = I.lits[3 * c + 1], l2 = I.lits[3 * c + 2];
        const int32_t g0 = I.gpos[3 * c], g1 = I.gpos[3 * c + 1], g2 = I.gpos[3 * c + 2];
        const BF2 a0 = vdr(I, abs(l0) - 1), a1 = vdr(I, abs(l1) - 1), a2 = vdr(I, abs(l2) - 1);
        float xs = I.xs[c], xl = I.xl[c];
        float r0 = a0.v, r1 = a1.v, r2 = a2.v;
        if (pend) {
            memory_update(p, clause3_C_vals<float>(l0, l1, l2, r0, r1, r2), h, xs, xl);
            I.xs[c] = xs;
            I.xl[c] = xl;
            r0 = clampV(a0.v + h * a0.d);
            r1 = clampV(a1.v + h * a1.d);
            r2 = clampV(a2.v + h * a2.d);
        }
        // contributions go to their variable-major slots, so the variable
        // kernel reads each variable's slots as one contiguous range
        auto emit = [&](int j, float val) { I.e[j == 0 ? g0 : j == 1 ? g1 : g2] = val; };
        bool sat;
        float C = clause3_vals<float>(p, l0, l1, l2, r0, r1, r2, clamp01(xs), clampXl(xl, p.xlmax), emit, sat);
        unsolved = !(C < 0.5f) || !sat;
    } else if ((I.flags & BF_GENERAL) && c < I.m) {
        const int bb = I.cp[c];
        const int k = I.cp[c + 1] - bb;
        auto cl = I.lits + bb;
        float xs = I.xs[c], xl = I.xl[c];
        auto prevV = [&](int v) { return vdr(I, v).v; };
        auto curV = [&](int v) {
            BF2 a = vdr(I, v);
            return pend ? clampV(a.v + h * a.d) : a.v;
        };
        if (pend) {
            memory_update(p, clauseK_C<float>(cl, k, prevV), h, xs, xl);
            I.xs[c] = xs;
            I.xl[c] = xl;
        }
        auto gp = I.gpos + bb;
        auto emit = [&](int j, float val) { I.e[gp[j]] = val; };
        bool sat;
        float C = clauseK<float>(p, cl, k, clamp01(xs), clampXl(xl, p.xlmax), curV, emit, sat);
        unsolved = !(C < 0.5f) || !sat;
    }
    if (__syncthreads_or(unsolved) && threadIdx.x == 0) raise_flag(ctl->flags[r][FG_UNSOLVED]);
}

__global__ __launch_bounds__(WBS) void k_wide_var(const BInst *__restrict__ insts, Ctl *__restrict__ wctl,
                                                  const int2 *__restrict__ vmap, int nblk, int r)
{
    __shared__ float stage[WBS / 64][WCAPW];
    __shared__ int sh[WBS / 64];
    const int b = blockIdx.x;
    if (b >= nblk) return;
    const int2 bm = vmap[b];
    const int slot = __builtin_amdgcn_readfirstlane(bm.x), lb = __builtin_amdgcn_readfirstlane(bm.y);
    Ctl *__restrict__ ctl = &wctl[slot];
    // the descriptor is read beside the control words (it is always valid
    // memory); its arrays only once the instance is known to run
    const GInst I = global_view(insts[slot]);
    if (ctl->status != DNX_RUNNING) return;
    const int ps = (r + 2) % 3;
    if (ctl->done[ps]) return;
    const int w = threadIdx.x >> 6;
    const int v0 = (lb * (WBS / 64) + w) * 64;
    const int v = v0 + (threadIdx.x & 63);
    const bool vok = v < I.n;
    // everything of the lane's variable is requested at once
    const int rb = vok ? I.row_ptr[v] : 0, re = vok ? I.row_ptr[v + 1] : 0;
    const BF2 a = vok ? vdr(I, v) : BF2{0.0f, 0.0f};
    const bool pinned = vok && (I.flags & BF_UNIT) && I.unit[v] != 0;
    const bool pend = ctl->stepped[ps] != 0;
    const float h = pend ? level_h<float>(dev_level(ctl, ps)) : 0.0f;
    const bool solved = ctl->stop_on_solve && !any_flag(ctl->flags[r][FG_UNSOLVED]);
    const int done = solved ? DNX_SAT : (ctl->steps >= ctl->cap ? DNX_UNKNOWN : 0);
    // contributions are variable-major in wide mode (k_wide_clause)
    const float dVf =
        wave_fold_vm<WCAPW>([&](int q) { return I.e[q]; }, rb, re, v0 < I.n ? min(63, I.n - 1 - v0) : 0, stage[w]);
    int f = 0;
    if (vok) {
        const float V = pend ? clampV(a.v + h * a.d) : a.v;
        float dV = 0.0f;
        if (!done) {
            dV = dVf;
            if (pinned) dV = 0.0f;  // dynex.cc:306
            f = fail_level(V, dV);
        }
        BF2 o;
        o.v = V;
        o.d = dV;
        vdw(I, v, o);
    }
    const int ns = (r + 1) % 3;
    if (done) {
        if (lb == 0) {
            if (threadIdx.x == 0) {
                ctl->done[r] = done;
                ctl->t_done = wall_clock64();
            }
            clear_slot(ctl, ns);
        }
        return;
    }
    const int bmx = block_max_int<WBS>(f, sh);
    if (threadIdx.x == 0)
        for (int L = 0; L < bmx; L++) raise_flag(ctl->flags[r][FG_LEVEL + L]);
    if (lb == 0) {
        if (threadIdx.x == 0) ctl->stepped[r] = 1;
        clear_slot(ctl, ns);
    }
}

void launch_wide_enter(const BDyn *dyn, Ctl *wctl, const int *list, int count, int r0, void *stream)
{
    if (count > 0) k_wide_enter<<<count, 256, 0, (hipStream_t)stream>>>(dyn, wctl, list, count, r0);
}

void launch_wide_sync(BDyn *dyn, Ctl *wctl, const int *list, int count, int rl, void *stream)
{
    if (count > 0) k_wide_sync<<<count, 64, 0, (hipStream_t)stream>>>(dyn, wctl, list, count, rl);
}

void launch_wide_step(const BInst *insts, Ctl *wctl, const int *bmap, int nbc, const int *vmap, int nbv, int r,
                      void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (nbc > 0) k_wide_clause<<<nbc, WBS, 0, st>>>(insts, wctl, (const int2 *)bmap, nbc, r);
    if (nbv > 0) k_wide_var<<<nbv, WBS, 0, st>>>(insts, wctl, (const int2 *)vmap, nbv, r);
}

// ---------------------------------------------------------------------------
// Gang mode.  For the tail of a batch -- a few instances, each needing
// thousands of steps -- a per-step kernel pair pays two launch gaps and a
// chain of dependent device-memory round trips per step.  Here each running
// instance is advanced by a gang of G workgroups (G <= GANG_MAX) for K steps
// inside one cooperative launch (every workgroup resident at once).  A
// workgroup owns a contiguous range of clauses, whose literals, transpose
// slots and memories stay in LDS for the whole launch, and a range of
// variables.  A step is: clause pass (contributions to their variable-major
// slots of e[]), gang barrier carrying the OR of the "unsolved" verdicts,
// variable pass (contiguous fold, (V, dV) pairs, ladder level), gang barrier
// carrying the maximum level.  The arithmetic is the single-instance FP32
// arithmetic with the same clause-order fold.
//
// All workgroups of a gang run on one XCD (the host places gang g's
// workgroups at block indices congruent to its XCD modulo 8, the order in
// which the dispatcher hands workgroups to XCDs; every workgroup checks its
// XCD id before the first step), so the XCD's L2 is the one point where
// they see each other's data.  A barrier is then: every wave waits until
// its stores have reached L2 (vmcnt 0), workgroup `rank` publishes
// (epoch << 16 | value) in its own 128-byte line of slot (epoch & 1), one
// wave polls the gang's G lines (one lane each) until every epoch has
// arrived and takes the OR / max of the values.  Loads of data that other
// workgroups of the gang write during the launch (V/dV pairs,
// contributions) bypass the CU's L1 (agent-scope relaxed loads).  No
// agent-scope fence is needed: those write back and invalidate the whole
// L2, which measured 3-5x slower per barrier.  Two line slots suffice: no
// workgroup can publish epoch e+2 before every workgroup has published e+1,
// which is after it finished reading epoch e.  Epochs grow across launches
// (the host passes each launch a fresh base), so the lines never need
// clearing.  Every wait is bounded: a wait that exceeds GANG_SPIN polls sets
// *err = 1 and the launch ends on every workgroup.
// ---------------------------------------------------------------------------
constexpr int GBS = 512;
constexpr int GANG_SPIN = 1 << 21;

__device__ __forceinline__ int xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20); }

__device__ __forceinline__ unsigned long long *gang_line(unsigned long long *words, int gang, int s, int rank)
{
    return words + (((size_t)gang * 2 + s) * GANG_MAX + rank) * GANG_LINE;
}

// (V, dV) pair and contribution loads from L2 (bypassing the CU's L1)
__device__ __forceinline__ BF2 vdr_l2(const GInst &I, int i)
{
    const unsigned long long u =
        __hip_atomic_load((unsigned long long *)(I.vd + i), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    BF2 r;
    r.v = __uint_as_float((unsigned)u);
    r.d = __uint_as_float((unsigned)(u >> 32));
    return r;
}
__device__ __forceinline__ float er_l2(const GInst &I, int q)
{
    return __hip_atomic_load((float *)(I.e + q), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// All threads of the workgroup call this.  Returns false when the wait
// timed out (or another workgroup's did); `res` is the gang-wide OR (op 0)
// or max (op 1) of `value`.
__device__ __forceinline__ bool gang_barrier(unsigned long long *words, int gang, int G, int rank,
                                             unsigned long long epoch, int value, int op, int &res, int *err,
                                             const int *lvalue = nullptr)
{
    __shared__ int s_res, s_ok;
    // this wave's stores of the step have reached the XCD's L2
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    __syncthreads();
    const int s = (int)(epoch & 1);
    // lvalue: the workgroup's value, gathered in LDS before this barrier.
    // The line goes out as a plain 8-byte store (one access; the L1 writes
    // through to the XCD's L2, where the other workgroups' agent-scope loads
    // find it): measured 0.2 us one way against 0.5 us for an agent-scope
    // store.  The asm statement keeps the compiler from moving it.
    if (threadIdx.x == 0) {
        const unsigned long long line = (epoch << 16) | (unsigned)((lvalue ? *lvalue : value) & 0xffff);
        asm volatile("global_store_dwordx2 %0, %1, off" ::"v"(gang_line(words, gang, s, rank)), "v"(line)
                     : "memory");
    }
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        bool ok = false;
        int r = 0;
        for (int it = 0; it < GANG_SPIN; it++) {
            const unsigned long long w =
                lane < G ? __hip_atomic_load(gang_line(words, gang, s, lane), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                         : (epoch << 16);
            if (__all((w >> 16) >= epoch)) {
                const int v = lane < G ? (int)(w & 0xffff) : 0;
                r = op ? wave_max(v) : (__ballot(v != 0) != 0ull ? 1 : 0);
                ok = true;
                break;
            }
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
            __builtin_amdgcn_s_sleep(1);
        }
        if (lane == 0) {
            s_ok = ok;
            s_res = r;
            if (!ok) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
    res = s_res;
    return s_ok != 0;
}

// LDS of one gang workgroup: its clauses' literals, transpose slots and
// memories, and the per-wave staging of the fold.
template <int GB, int CW>
struct GangLds {
    int32_t lits[3 * GANG_CPW];
    int32_t gpos[3 * GANG_CPW];
    float xs[GANG_CPW], xl[GANG_CPW];
    float cst[GANG_CPW];  // each clause's C of the state the previous step evaluated
    float stage[GB / 64][CW];
};
constexpr int GANG_LDS = (int)sizeof(GangLds<GBS, WCAPW>);
// the epoch launch's workgroups: 1024 threads (the fused kernel's mode-0
// shape); its gang members stage up to 1024 contributions per wave (a wave
// whose variables have more folds them from device memory, same order)
constexpr int EBS = 1024, ECAPW = 1024;
constexpr int EGANG_LDS = (int)sizeof(GangLds<EBS, ECAPW>);

// Workgroup `rank` of gang `gang` (G workgroups) advances instance `slot` by
// up to K steps.
template <int GBS, int CW>
__device__ __forceinline__ void gang_run(const BInst *__restrict__ insts, BDyn *__restrict__ dyn, int slot, int rank,
                                         int G, int gang, unsigned long long *__restrict__ words,
                                         unsigned long long base, int K, int *__restrict__ err, GangLds<GBS, CW> &L)
{
    __shared__ int sh[GBS / 64];
    const int tid = threadIdx.x;
    int32_t *s_lits = L.lits, *s_gpos = L.gpos;
    float *s_xs = L.xs, *s_xl = L.xl, *s_cst = L.cst;
    auto &stage = L.stage;
    const BDyn d = dyn[slot];
    if (d.status != DNX_RUNNING) return;
    const GInst I = global_view(insts[slot]);
    const int n = I.n, m = I.m;
    const bool general = (I.flags & BF_GENERAL) != 0;
    PT<float> p;
    p.alpha = I.alpha;
    p.beta = I.beta;
    p.gamma = I.gamma;
    p.delta = I.delta;
    p.eps = I.eps;
    p.zeta = I.zeta;
    p.xlmax = I.xlmax;
    // this workgroup's clauses [c0, c0 + cn) and variables [vb, vb + Vb)
    const int Cb = (m + G - 1) / G;
    const int c0 = min(m, rank * Cb);
    const int cn = min(m, c0 + Cb) - c0;
    // variables per workgroup in whole waves (not whole workgroups: with
    // n / G below the workgroup size, most of a large gang's workgroups
    // would have no variable and a few would fold for all of them)
    const int Vb = ((n + G - 1) / G + 63) / 64 * 64;
    const int vb = rank * Vb;
    if (!general)
        for (int i = tid; i < 3 * cn; i += GBS) {
            s_lits[i] = I.lits[3 * c0 + i];
            s_gpos[i] = I.gpos[3 * c0 + i];
        }
    for (int i = tid; i < cn; i += GBS) {
        s_xs[i] = I.xs[c0 + i];
        s_xl[i] = I.xl[c0 + i];
    }
    __syncthreads();
    // slot range and (V, dV) pair of each thread's first variable: the range
    // is read once per launch, the pair kept from the thread's own last write
    int RB0 = 0, RE0 = 0;
    BF2 PA{0.0f, 0.0f};
    {
        const int v = vb + tid;
        if (v < n && v < vb + Vb) {
            RB0 = I.row_ptr[v];
            RE0 = I.row_ptr[v + 1];
            PA = vdr(I, v);
        }
    }
    bool pend = d.plevel >= 0;
    int plev = d.plevel;
    long long steps = d.steps;
    double t = d.t;
    int done = 0;
    unsigned long long epoch = base;
    for (int k = 0; k < K; k++) {
        const float h = pend ? level_h<float>(plev) : 0.0f;
        if (pend) {
            steps += 1;
            t += level_dt(plev);
        }
        // ---- clause pass ----------------------------------------------------
        int unsolved = 0;
        // steady: a step of this launch before this one stored each clause's
        // C, and the memories come out of a memory update (clamped), so the
        // steady-state forms apply (V is clamped below) -- the same bits
        const bool steady = k >= 1 && steps >= 2;
        if (!general) {
            // 3-SAT: the (V, dV) pairs of all of a thread's clauses (CPT =
            // GANG_CPW / GBS of them) are requested before any is evaluated.
            // Every clause slot is evaluated on every path (indices clamped
            // into the range; a slot past the range computes the last clause
            // again and stores nothing), and a thread's contributions are
            // stored after all its clauses: the compiler's waits then count a
            // fixed number of accesses, and no clause waits for the stores of
            // the one before it
            constexpr int CPT = GANG_CPW / GBS;
            static_assert(CPT * GBS == GANG_CPW, "one pass of CPT clauses per thread");
            if (tid < cn) {
                BF2 A[CPT][3];
                int ic[CPT];
#pragma unroll
                for (int u = 0; u < CPT; u++) {
                    ic[u] = min(tid + u * GBS, cn - 1);
#pragma unroll
                    for (int j = 0; j < 3; j++) A[u][j] = vdr_l2(I, abs(s_lits[3 * ic[u] + j]) - 1);
                }
                float E[CPT][3];
                int Q[CPT][3];
#pragma unroll
                for (int u = 0; u < CPT; u++) {
                    const int c = ic[u];
                    const bool ok = tid + u * GBS < cn;
                    const int32_t l0 = s_lits[3 * c], l1 = s_lits[3 * c + 1], l2 = s_lits[3 * c + 2];
                    const BF2 a0 = A[u][0], a1 = A[u][1], a2 = A[u][2];
                    float xs = s_xs[c], xl = s_xl[c];
                    auto emit = [&](int j, float val) { E[u][j] = val; };
                    float C;
                    bool sat = true;
                    if (steady) {
                        memory_update_steady(p, s_cst[c], h, xs, xl);
                        C = clause3_steady(p, l0, l1, l2, clampV(a0.v + h * a0.d), clampV(a1.v + h * a1.d),
                                           clampV(a2.v + h * a2.d), xs, xl, emit);
                    } else {
                        float r0 = a0.v, r1 = a1.v, r2 = a2.v;
                        if (pend) {
                            memory_update(p, clause3_C_vals<float>(l0, l1, l2, r0, r1, r2), h, xs, xl);
                            r0 = clampV(a0.v + h * a0.d);
                            r1 = clampV(a1.v + h * a1.d);
                            r2 = clampV(a2.v + h * a2.d);
                        }
                        C = clause3_vals<float>(p, l0, l1, l2, r0, r1, r2, clamp01(xs), clampXl(xl, p.xlmax), emit,
                                                sat);
                    }
#pragma unroll
                    for (int j = 0; j < 3; j++) Q[u][j] = s_gpos[3 * c + j];
                    if (ok) {
                        if (steady || pend) {
                            s_xs[c] = xs;
                            s_xl[c] = xl;
                        }
                        s_cst[c] = C;
                        unsolved |= !(C < 0.5f) || !sat;
                    }
                }
#pragma unroll
                for (int u = 0; u < CPT; u++)
                    if (tid + u * GBS < cn)
#pragma unroll
                        for (int j = 0; j < 3; j++) I.e[Q[u][j]] = E[u][j];
            }
        }
        // any clause length (instances with general clauses)
        for (int i = tid; general && i < cn; i += GBS) {
            const int c = c0 + i;
            float xs = s_xs[i], xl = s_xl[i];
            const int bb = I.cp[c];
            const int kk = I.cp[c + 1] - bb;
            auto cl = I.lits + bb;
            auto prevV = [&](int v) { return vdr_l2(I, v).v; };
            auto curV = [&](int v) {
                BF2 a = vdr_l2(I, v);
                return pend ? clampV(a.v + h * a.d) : a.v;
            };
            if (pend) {
                memory_update(p, clauseK_C<float>(cl, kk, prevV), h, xs, xl);
                s_xs[i] = xs;
                s_xl[i] = xl;
            }
            auto gp = I.gpos + bb;
            auto emit = [&](int j, float val) { I.e[gp[j]] = val; };
            bool sat;
            const float C = clauseK<float>(p, cl, kk, clamp01(xs), clampXl(xl, p.xlmax), curV, emit, sat);
            unsolved |= !(C < 0.5f) || !sat;
        }
        int any_unsolved = 0;
        if (!gang_barrier(words, gang, G, rank, ++epoch, __syncthreads_or(unsolved), 0, any_unsolved, err)) return;
        if (d.stop_on_solve && !any_unsolved)
            done = DNX_SAT;
        else if (steps >= d.cap)
            done = DNX_UNKNOWN;
        // ---- variable pass --------------------------------------------------
        int fmax = 0;
        const int w = tid >> 6;
        // every wave takes the same number of rounds (the fold stages through
        // LDS behind a workgroup barrier); a wave past the range folds nothing
        for (int r0 = 0; r0 < Vb; r0 += GBS) {
            const int v0 = vb + r0 + w * 64;
            const bool wok = r0 + w * 64 < Vb;
            const int v = v0 + (tid & 63);
            const bool vok = wok && v < n;
            const bool first = r0 == 0;
            const int rb = first ? RB0 : vok ? I.row_ptr[v] : 0, re = first ? RE0 : vok ? I.row_ptr[v + 1] : 0;
            const BF2 a = first ? PA : vok ? vdr_l2(I, v) : BF2{0.0f, 0.0f};  // requested before the fold
            const float dVf = wave_fold_vm<CW>([&](int q) { return er_l2(I, q); }, rb, re,
                                                  wok && v0 < n ? min(63, n - 1 - v0) : 0, stage[w]);
            if (vok) {
                const float V = pend ? clampV(a.v + h * a.d) : a.v;
                float dV = 0.0f;
                if (!done) {
                    dV = dVf;
                    if ((I.flags & BF_UNIT) && I.unit[v] != 0) dV = 0.0f;  // dynex.cc:306
                    fmax = max(fmax, fail_level(V, dV));
                }
                BF2 o;
                o.v = V;
                o.d = dV;
                vdw(I, v, o);
                if (first) PA = o;
            }
        }
        if (done) break;
        const int bmx = block_max_int<GBS>(fmax, sh);
        if (tid == 0) sh[0] = bmx;
        __syncthreads();
        const int wg_level = sh[0];
        if (!gang_barrier(words, gang, G, rank, ++epoch, wg_level, 1, plev, err)) return;
        pend = true;
    }
    __syncthreads();
    for (int i = tid; i < cn; i += GBS) {
        I.xs[c0 + i] = s_xs[i];
        I.xl[c0 + i] = s_xl[i];
    }
    if (rank == 0 && tid == 0) {
        BDyn o = d;
        o.steps = steps;
        o.t = t;
        if (done) {
            o.status = done;
            o.plevel = -1;
            o.t_done = wall_clock64();
        } else {
            o.plevel = pend ? plev : -1;
        }
        dyn[slot] = o;
    }
}

// ---------------------------------------------------------------------------
// Regrouped gang (3-SAT).  gang_run's steps and barriers -- workgroup r of G
// evaluates clauses [r Cb, r Cb + cn), owns variables [r Vb, r Vb + vn) and
// publishes their (V, dV) pairs in vd; a gang barrier follows each pass --
// with the contributions regrouped: writer A stages its clauses'
// contributions in LDS by owner B (region (A, B); inside it, in the order of
// B's variable-major slots) and writes its area as one contiguous run, and
// owner B reads its G regions (one contiguous run each) into LDS, where each
// owned variable's contributions lie together in slot (clause) order: the
// fold is the sum gang_run and the fused kernel form, in the same order.
// gang_run writes each contribution to its own transpose slot instead (3m
// scattered 4-byte stores per step, each its own L2 request) and gathers
// them back through the transpose.  The layout is derived at each launch from
// the transpose (eidx, row_ptr): owners count and rank their slots by writer
// and publish the ranks and region sizes, and the launch's first gang barrier
// makes them visible.
// ---------------------------------------------------------------------------
struct RxLds {
    int32_t lits[3 * GANG_CPW];          // the workgroup's clauses' literals
    uint16_t ldst[3 * GANG_CPW];         // each literal's place in the writer's area
    uint16_t sdst[RX_SLOTS];             // owner: the slot (from its first) each gathered value folds as
    float xs[GANG_CPW], xl[GANG_CPW], cst[GANG_CPW];
    float sv[RX_VCAP], sd[RX_VCAP];      // owned variables' V and dV
    int pre[GANG_MAX + 1];               // writer: start of region (this, B) in its area
    int soff[GANG_MAX + 1];              // owner: start of region (A, this) in its gathered values
    int rst[GANG_MAX];                   // owner: start of region (A, this) in the scratch
    union {
        int cnt[RX_NCAP + GANG_MAX];       // set-up: contributions per (writer, owned variable)
        struct {
            float stage[RX_SLOTS];         // clause pass: the area; variable pass: the gathered regions
            int src[RX_SLOTS];             // owner: scratch index of each gathered value
        } s;
    } u;
};
constexpr int RX_LDS = (int)sizeof(RxLds);

__device__ __forceinline__ float rx_ld(GLB const float *p)
{
    return __hip_atomic_load((float *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int rx_ldi(GLB const int *p)
{
    return __hip_atomic_load((int *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Workgroup `rank` of regrouped gang `gang` (G workgroups) advances instance
// `slot` by up to K steps, with gang_run's bits.  Returns false, having
// changed nothing, when the instance does not fit the tables (every
// workgroup of the gang judges this alike, from row_ptr at the block
// boundaries); the caller then runs gang_run.  Barrier epochs: base + 1 for
// the set-up, base + 2 + 2k and base + 3 + 2k for step k.
template <int GBS>
__device__ __forceinline__ bool gang_rx(const BInst *__restrict__ insts, BDyn *__restrict__ dyn, int slot, int rank,
                                        int G, int gang, unsigned long long *__restrict__ words,
                                        unsigned long long base, int K, int *__restrict__ err,
                                        unsigned char *__restrict__ scr, RxLds &L)
{
    constexpr int NW = GBS / 64;
    constexpr int CPT = GANG_CPW / GBS;     // clauses per thread
    constexpr int RV = RX_VCAP / GBS;       // owned variables per thread
    constexpr int NR = RX_SLOTS / GBS;      // gathered values per thread
    static_assert(CPT * GBS == GANG_CPW && RV * GBS == RX_VCAP && NR * GBS == RX_SLOTS, "whole rounds");
    __shared__ int s_lv;  // the workgroup's largest ladder level of the step
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const BDyn d = dyn[slot];
    if (d.status != DNX_RUNNING) return true;
    const GInst I = global_view(insts[slot]);
    const int n = I.n, m = I.m;
    if (I.flags & BF_GENERAL) return false;
    const int Cb = (m + G - 1) / G, c0 = min(m, rank * Cb), cn = min(m, c0 + Cb) - c0;
    const int Vb = rx_vblock(n, G), vb = min(n, rank * Vb), vn = min(n, vb + Vb) - vb;
    {
        int over = 0;
        if (tid < G) {
            const int b0 = min(n, tid * Vb), b1 = min(n, b0 + Vb);
            over = I.row_ptr[b1] - I.row_ptr[b0] > RX_SLOTS;
        }
        if (__syncthreads_or(over) || Vb > RX_VCAP || n > RX_NCAP || Cb > GANG_CPW) return false;
    }
    const RxLayout lay = rx_layout(n, m, G);
    GLB float *ew = (GLB float *)(scr + lay.ew);
    GLB int *rankt = (GLB int *)(scr + lay.rank);
    GLB int *szt = (GLB int *)(scr + lay.szt);
    PT<float> p;
    p.alpha = I.alpha;
    p.beta = I.beta;
    p.gamma = I.gamma;
    p.delta = I.delta;
    p.eps = I.eps;
    p.zeta = I.zeta;
    p.xlmax = I.xlmax;

    // ---- set-up: the clauses, memories and owned (V, dV) into LDS ----------
    for (int i = tid; i < 3 * cn; i += GBS) L.lits[i] = I.lits[3 * c0 + i];
    for (int i = tid; i < cn; i += GBS) {
        L.xs[i] = I.xs[c0 + i];
        L.xl[i] = I.xl[c0 + i];
    }
    for (int i = tid; i < vn; i += GBS) {
        const BF2 a = vdr(I, vb + i);
        L.sv[i] = a.v;
        L.sd[i] = a.d;
    }
    for (int i = tid; i < G * Vb; i += GBS) L.u.cnt[i] = 0;
    const int qb = I.row_ptr[vb];
    int RB[RV], RE[RV], PIN = 0;  // owned variables' slot ranges and unit pins
#pragma unroll
    for (int r = 0; r < RV; r++) {
        const int vl = tid + r * GBS;
        RB[r] = vl < vn ? I.row_ptr[vb + vl] : 0;
        RE[r] = vl < vn ? I.row_ptr[vb + vl + 1] : 0;
        PIN |= (vl < vn && (I.flags & BF_UNIT) && I.unit[vb + vl] != 0) << r;
    }
    __syncthreads();
    // owner: contributions to each owned variable from each writer (a
    // variable's slots run in clause order, so its writers do not decrease)
#pragma unroll
    for (int r = 0; r < RV; r++)
        for (int q = RB[r]; q < RE[r]; q++) L.u.cnt[((I.eidx[q] / 3) / Cb) * Vb + tid + r * GBS]++;
    __syncthreads();
    // exclusive scan of each writer's row over the owned variables; the
    // row's total is the size of region (writer, this owner)
    for (int A = w; A < G; A += NW) {
        int carry = 0;
        for (int x0 = 0; x0 < Vb; x0 += 64) {
            const int x = x0 + lane < Vb ? L.u.cnt[A * Vb + x0 + lane] : 0;
            int inc = x;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(inc, o);
                if (lane >= o) inc += y;
            }
            if (x0 + lane < Vb) L.u.cnt[A * Vb + x0 + lane] = carry + inc - x;
            carry += __shfl(inc, 63);
        }
        if (lane == 0) {
            L.soff[A] = carry;  // region size for now
            szt[A * G + rank] = carry;
        }
    }
    __syncthreads();
    if (tid == 0) {
        int s = 0;
        for (int A = 0; A < G; A++) {
            const int z = L.soff[A];
            L.soff[A] = s;
            s += z;
        }
        L.soff[G] = s;
    }
    __syncthreads();
    // each owned slot's rank in its region (for its writer), and the slot
    // each gathered value folds as
#pragma unroll
    for (int r = 0; r < RV; r++) {
        int pa = -1, k = 0;
        for (int q = RB[r]; q < RE[r]; q++) {
            const int A = (I.eidx[q] / 3) / Cb;
            k = A == pa ? k + 1 : 0;
            pa = A;
            const int rk = L.u.cnt[A * Vb + tid + r * GBS] + k;
            rankt[q] = rk;
            L.sdst[L.soff[A] + rk] = (uint16_t)(q - qb);
        }
    }
    unsigned long long epoch = base;
    int dummy;
    if (!gang_barrier(words, gang, G, rank, ++epoch, 0, 0, dummy, err)) return true;
    // writer: region starts in its area; each literal's place in it
    if (tid == 0) {
        int s = 0;
        for (int B = 0; B < G; B++) {
            L.pre[B] = s;
            s += rx_ldi(szt + rank * G + B);
        }
        L.pre[G] = s;
    }
    // owner: where region (A, this) starts in the scratch
    if (tid < G) {
        int s = 0;
        for (int B = 0; B < rank; B++) s += rx_ldi(szt + tid * G + B);
        L.rst[tid] = tid * 3 * Cb + s;
    }
    __syncthreads();
    if (L.pre[G] != 3 * cn) {
        if (tid == 0) __hip_atomic_store(err, 4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return true;
    }
    for (int i = tid; i < 3 * cn; i += GBS)
        L.ldst[i] = (uint16_t)(L.pre[(abs(L.lits[i]) - 1) / Vb] + rx_ldi(rankt + I.gpos[3 * c0 + i]));
    for (int A = w; A < G; A += NW)
        for (int j = lane; j < L.soff[A + 1] - L.soff[A]; j += 64) L.u.s.src[L.soff[A] + j] = L.rst[A] + j;
    const int totB = L.soff[G];
    __syncthreads();

    // ---- steps ------------------------------------------------------------
    bool pend = d.plevel >= 0;
    int plev = d.plevel;
    long long steps = d.steps;
    double t = d.t;
    int done = 0;
    for (int k = 0; k < K; k++) {
        const float h = pend ? level_h<float>(plev) : 0.0f;
        if (pend) {
            steps += 1;
            t += level_dt(plev);
        }
        // ---- clause pass (gang_run's arithmetic) ----------------------------
        int unsolved = 0;
        const bool steady = k >= 1 && steps >= 2;
        if (tid < cn) {
            // rounds past the workgroup's clauses are skipped (the test is
            // uniform); inside the last one a thread past them evaluates the
            // last clause again and keeps nothing
            BF2 A[CPT][3];
            int ic[CPT];
#pragma unroll
            for (int u = 0; u < CPT; u++) {
                ic[u] = min(tid + u * GBS, cn - 1);
                if (u * GBS < cn)
#pragma unroll
                    for (int j = 0; j < 3; j++) A[u][j] = vdr_l2(I, abs(L.lits[3 * ic[u] + j]) - 1);
            }
#pragma unroll
            for (int u = 0; u < CPT; u++) {
                if (u * GBS >= cn) continue;
                const int c = ic[u];
                const bool ok = tid + u * GBS < cn;
                const int32_t l0 = L.lits[3 * c], l1 = L.lits[3 * c + 1], l2 = L.lits[3 * c + 2];
                const BF2 a0 = A[u][0], a1 = A[u][1], a2 = A[u][2];
                float xs = L.xs[c], xl = L.xl[c];
                float E[3];
                auto emit = [&](int j, float val) { E[j] = val; };
                float C;
                bool sat = true;
                if (steady) {
                    memory_update_steady(p, L.cst[c], h, xs, xl);
                    C = clause3_steady(p, l0, l1, l2, clampV(a0.v + h * a0.d), clampV(a1.v + h * a1.d),
                                       clampV(a2.v + h * a2.d), xs, xl, emit);
                } else {
                    float r0 = a0.v, r1 = a1.v, r2 = a2.v;
                    if (pend) {
                        memory_update(p, clause3_C_vals<float>(l0, l1, l2, r0, r1, r2), h, xs, xl);
                        r0 = clampV(a0.v + h * a0.d);
                        r1 = clampV(a1.v + h * a1.d);
                        r2 = clampV(a2.v + h * a2.d);
                    }
                    C = clause3_vals<float>(p, l0, l1, l2, r0, r1, r2, clamp01(xs), clampXl(xl, p.xlmax), emit,
                                            sat);
                }
                if (ok) {
                    if (steady || pend) {
                        L.xs[c] = xs;
                        L.xl[c] = xl;
                    }
                    L.cst[c] = C;
                    unsolved |= !(C < 0.5f) || !sat;
#pragma unroll
                    for (int j = 0; j < 3; j++) L.u.s.stage[L.ldst[3 * c + j]] = E[j];
                }
            }
        }
        const int wg_unsolved = __syncthreads_or(unsolved);
        // the area, one contiguous run
        for (int i = tid; i < 3 * cn; i += GBS) ew[(size_t)rank * 3 * Cb + i] = L.u.s.stage[i];
        int any_unsolved = 0;
        if (!gang_barrier(words, gang, G, rank, ++epoch, wg_unsolved, 0, any_unsolved, err)) return true;
        if (d.stop_on_solve && !any_unsolved)
            done = DNX_SAT;
        else if (steps >= d.cap)
            done = DNX_UNKNOWN;
        // ---- variable pass --------------------------------------------------
        if (tid == 0) s_lv = 0;  // read only after barrier B's first workgroup barrier
        {
            float x[NR];
#pragma unroll
            for (int u = 0; u < NR; u++) {
                const int i = tid + u * GBS;
                x[u] = i < totB ? rx_ld(ew + L.u.s.src[i]) : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < NR; u++) {
                const int i = tid + u * GBS;
                if (i < totB) L.u.s.stage[L.sdst[i]] = x[u];
            }
        }
        __syncthreads();
        int fmax = 0;
#pragma unroll
        for (int r = 0; r < RV; r++) {
            const int vl = tid + r * GBS;
            if (vl < vn) {
                // the variable's contributions lie together in slot order:
                // eight loads at a time, summed in that order
                float acc = 0.0f;
                const float *sg = L.u.s.stage + (RB[r] - qb);
                const int cnt = RE[r] - RB[r];
                for (int j0 = 0; j0 < cnt; j0 += 8) {
                    float y[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) y[u] = j0 + u < cnt ? sg[j0 + u] : 0.0f;
#pragma unroll
                    for (int u = 0; u < 8; u++)
                        if (j0 + u < cnt) acc = acc + y[u];
                }
                const float V = pend ? clampV(L.sv[vl] + h * L.sd[vl]) : L.sv[vl];
                float dV = 0.0f;
                if (!done) {
                    dV = acc;
                    if ((PIN >> r) & 1) dV = 0.0f;  // dynex.cc:306
                    fmax = max(fmax, fail_level(V, dV));
                }
                L.sv[vl] = V;
                L.sd[vl] = dV;
                BF2 o;
                o.v = V;
                o.d = dV;
                vdw(I, vb + vl, o);
            }
        }
        if (done) break;
        const int wm = wave_max(fmax);
        if (lane == 0 && wm > 0) atomicMax(&s_lv, wm);
        if (!gang_barrier(words, gang, G, rank, ++epoch, 0, 1, plev, err, &s_lv)) return true;
        pend = true;
    }
    // ---- write-back ----------------------------------------------------------
    __syncthreads();
    for (int i = tid; i < cn; i += GBS) {
        I.xs[c0 + i] = L.xs[i];
        I.xl[c0 + i] = L.xl[i];
    }
    if (rank == 0 && tid == 0) {
        BDyn o = d;
        o.steps = steps;
        o.t = t;
        if (done) {
            o.status = done;
            o.plevel = -1;
            o.t_done = wall_clock64();
        } else {
            o.plevel = pend ? plev : -1;
        }
        dyn[slot] = o;
    }
    return true;
}

// map[b] = (slot, rank, G, gang) for block b (slot -1: idle block).  Before
// the first step every block checks that it runs on XCD b % 8; one that does
// not sets *err = 2, and since every block passes a launch-wide arrival count
// (*arrive, compared against arrive_base + nblk) before touching any state,
// a misplaced launch ends with every instance unchanged.
__global__ __launch_bounds__(GBS) void k_gang(const BInst *__restrict__ insts, BDyn *__restrict__ dyn,
                                             const int4 *__restrict__ map, int nblk,
                                             unsigned long long *__restrict__ words, unsigned long long base, int K,
                                             int *__restrict__ err, int *__restrict__ arrive, int arrive_base)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_go;
    const int b = blockIdx.x;
    if (b >= nblk) return;
    const int tid = threadIdx.x;
    if (tid == 0) {
        if (xcc_id() != (b & 7)) __hip_atomic_store(err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        __hip_atomic_fetch_add(arrive, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int it = 0;
        while (__hip_atomic_load(arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - arrive_base < nblk &&
               ++it < GANG_SPIN)
            __builtin_amdgcn_s_sleep(1);
        __threadfence();
        if (it >= GANG_SPIN) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_go = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    }
    __syncthreads();
    if (!s_go) return;
    const int4 gm = map[b];
    const int slot = __builtin_amdgcn_readfirstlane(gm.x), rank = __builtin_amdgcn_readfirstlane(gm.y);
    const int G = __builtin_amdgcn_readfirstlane(gm.z), gang = __builtin_amdgcn_readfirstlane(gm.w);
    if (slot < 0) return;
    gang_run<GBS, WCAPW>(insts, dyn, slot, rank, G, gang, words, base, K, err, *(GangLds<GBS, WCAPW> *)smem);
}

// ---------------------------------------------------------------------------
// Epoch launch: the tail of a batch as one cooperative launch of one
// workgroup per CU, in which gangs and single-workgroup instances share the
// chip.  A workgroup takes its role from the XCD it runs on and the order in
// which workgroups start there (flow->xcd_rank): the first gcount[x]
// workgroups on XCD x are the gang members listed at gmap[x * GANG_MAX + r]
// as (slot, rank, G, gang); every other workgroup is a flow worker, which
// takes instances from the work list and runs each until it ends or the flow
// stops.  Roles are thus placed by where the workgroups actually run, not by
// their block index.  A gang member first waits (bounded) until all of its
// XCD's gang members have started (*err = 3 if they do not: the gangs then
// change nothing); the gangs advance their instances up to K steps, and the
// last gang workgroup to finish raises the flow's stop, after which the flow
// workers leave at the end of their current step.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(EBS) void k_epoch(const BInst *__restrict__ insts, BDyn *__restrict__ dyn,
                                              const int4 *__restrict__ gmap, const int *__restrict__ gcount,
                                              unsigned long long *__restrict__ words, unsigned long long base, int K,
                                              int *__restrict__ err, const int *__restrict__ work, int nwork,
                                              int *__restrict__ head, BFlow *__restrict__ flow,
                                              unsigned char *__restrict__ gscr,
                                              const unsigned long long *__restrict__ goff)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int s_role;
    const int tid = threadIdx.x;
    if (tid == 0) {
        const int x = xcc_id();
        int role = -1;
        if (x >= 0 && x < GANG_XCD) {
            const int gx = gcount[x];
            if (gx > 0) {
                const int r = __hip_atomic_fetch_add(&flow->xcd_rank[x], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (r < gx) {
                    role = x * GANG_MAX + r;
                    // every gang member of this XCD has started before any
                    // of them touches an instance
                    int it = 0;
                    while (__hip_atomic_load(&flow->xcd_rank[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < gx &&
                           ++it < GANG_SPIN)
                        __builtin_amdgcn_s_sleep(1);
                    if (it >= GANG_SPIN) __hip_atomic_store(err, 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) role = -2;
                }
            }
        }
        s_role = role;
    }
    __syncthreads();
    const int role = s_role;
    if (role == -1) {
        flow_worker<BM_LDS_ALL, EBS>(insts, dyn, work, nwork, head, INT_MAX, flow);
        return;
    }
    if (role >= 0) {
        const int4 gm = gmap[role];
        const int slot = __builtin_amdgcn_readfirstlane(gm.x), rank = __builtin_amdgcn_readfirstlane(gm.y);
        const int gz = __builtin_amdgcn_readfirstlane(gm.z), gang = __builtin_amdgcn_readfirstlane(gm.w);
        const int G = gz & 0xffff;
        // a regrouped gang that does not fit its tables runs as gang_run
        if (slot >= 0 && (!(gz >> 16) || !gang_rx<EBS>(insts, dyn, slot, rank, G, gang, words, base, K, err,
                                                        gscr + goff[gang], *(RxLds *)smem))) {
            __syncthreads();
            gang_run<EBS, ECAPW>(insts, dyn, slot, rank, G, gang, words, base, K, err, *(GangLds<EBS, ECAPW> *)smem);
        }
    }
    __syncthreads();
    if (tid == 0 &&
        __hip_atomic_fetch_add(&flow->gangs_done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 >= flow->gangs)
        __hip_atomic_store(&flow->stop, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_batch_zero(int *p, int count)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
}

__global__ void k_batch_stamp(unsigned long long *p)
{
    if (threadIdx.x == 0) *p = wall_clock64();
}

// reference-layout state of one slot with its pending step applied
__global__ __launch_bounds__(256) void k_batch_materialize(const BInst *__restrict__ ip, const BDyn *__restrict__ dp,
                                                           float *__restrict__ xo)
{
    const GInst I = global_view(*ip);
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int n = I.n, m = I.m;
    const int plev = dp->status == DNX_RUNNING ? dp->plevel : -1;
    const bool pend = plev >= 0;
    const float h = pend ? level_h<float>(plev) : 0.0f;
    if (i < n) {
        BF2 a = vdr(I, i);
        xo[i] = pend ? clampV(a.v + h * a.d) : a.v;
    }
    if (i < m) {
        float xs = I.xs[i], xl = I.xl[i];
        if (pend) {
            PT<float> p;
            p.alpha = I.alpha;
            p.beta = I.beta;
            p.gamma = I.gamma;
            p.delta = I.delta;
            p.eps = I.eps;
            p.zeta = I.zeta;
            p.xlmax = I.xlmax;
            const bool general = (I.flags & BF_GENERAL) != 0;
            const int b = general ? I.cp[i] : 3 * i;
            auto getV = [&](int v) { return vdr(I, v).v; };
            float Cp = general ? clauseK_C<float>(I.lits + b, I.cp[i + 1] - b, getV)
                               : clause3_C<float>(I.lits + b, getV);
            memory_update(p, Cp, h, xs, xl);
        }
        xo[n + i] = xs;
        xo[n + m + i] = xl;
    }
}

// Dynamic LDS above 64 KB must be allowed per kernel, and the setting is
// per device: set[dev] keeps the largest size allowed on each device.
constexpr int MAX_DEV_ATTR = 64;
static void lds_attr(const void *fn, int lds, int (&set)[MAX_DEV_ATTR])
{
    if (lds <= 65536) return;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEV_ATTR) dev = 0;
    if (lds > set[dev]) {
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        set[dev] = lds;
    }
}

int gang_occupancy()
{
    static int set[MAX_DEV_ATTR] = {};
    lds_attr((const void *)k_gang, GANG_LDS, set);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_gang, GBS, GANG_LDS) != hipSuccess) return 0;
    return nb;
}

int launch_gang(const BInst *insts, BDyn *dyn, const int *map, int nblk, unsigned long long *words,
                unsigned long long base, int K, int *err, int *arrive, int arrive_base, void *stream)
{
    if (nblk <= 0) return 0;
    static int set[MAX_DEV_ATTR] = {};
    lds_attr((const void *)k_gang, GANG_LDS, set);
    const int4 *mp = (const int4 *)map;
    void *args[] = {(void *)&insts, (void *)&dyn, (void *)&mp, (void *)&nblk, (void *)&words, (void *)&base,
                    (void *)&K, (void *)&err, (void *)&arrive, (void *)&arrive_base};
    return (int)hipLaunchCooperativeKernel((const void *)k_gang, dim3(nblk), dim3(GBS), args, GANG_LDS,
                                           (hipStream_t)stream);
}

int epoch_lds(int flow_lds) { return max(flow_lds, max(EGANG_LDS, RX_LDS)); }

int epoch_occupancy(int flow_lds)
{
    static int set[MAX_DEV_ATTR] = {};
    const int lds = epoch_lds(flow_lds);
    lds_attr((const void *)k_epoch, lds, set);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_epoch, EBS, lds) != hipSuccess) return 0;
    return nb;
}

int launch_epoch(const BInst *insts, BDyn *dyn, const int *gmap, const int *gcount, unsigned long long *words,
                 unsigned long long base, int K, int *err, const int *work, int nwork, int *head, BFlow *flow,
                 unsigned char *gscr, const unsigned long long *goff, int flow_lds, int grid, void *stream)
{
    static int set[MAX_DEV_ATTR] = {};
    const int lds = epoch_lds(flow_lds);
    lds_attr((const void *)k_epoch, lds, set);
    const int4 *gm = (const int4 *)gmap;
    void *args[] = {(void *)&insts, (void *)&dyn,  (void *)&gm,   (void *)&gcount, (void *)&words,
                    (void *)&base,  (void *)&K,    (void *)&err,  (void *)&work,   (void *)&nwork,
                    (void *)&head,  (void *)&flow, (void *)&gscr, (void *)&goff};
    return (int)hipLaunchCooperativeKernel((const void *)k_epoch, dim3(grid), dim3(EBS), args, lds,
                                           (hipStream_t)stream);
}

template <int MODE, int BS>
static void launch_one(int lds, int grid, const BInst *insts, BDyn *dyn, const int *work, int nwork, int *head,
                       int K, BFlow *flow, hipStream_t st)
{
    static int lds_set[MAX_DEV_ATTR] = {};
    lds_attr((const void *)k_batch<MODE, BS>, lds, lds_set);
    k_batch<MODE, BS><<<grid, BS, lds, st>>>(insts, dyn, work, nwork, head, K, flow);
}

template <int MODE, int BS>
static int occ_one(int lds)
{
    if (lds > 65536)
        (void)hipFuncSetAttribute((const void *)k_batch<MODE, BS>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch<MODE, BS>, BS, lds) != hipSuccess) return 0;
    return nb;
}

template <int MODE>
static void launch_mode(int bs, int lds, int grid, const BInst *insts, BDyn *dyn, const int *work, int nwork,
                        int *head, int K, BFlow *flow, hipStream_t st)
{
    switch (bs) {
        case 128: launch_one<MODE, 128>(lds, grid, insts, dyn, work, nwork, head, K, flow, st); break;
        case 256: launch_one<MODE, 256>(lds, grid, insts, dyn, work, nwork, head, K, flow, st); break;
        case 512: launch_one<MODE, 512>(lds, grid, insts, dyn, work, nwork, head, K, flow, st); break;
        default: launch_one<MODE, 1024>(lds, grid, insts, dyn, work, nwork, head, K, flow, st); break;
    }
}

template <int MODE>
static int occ_mode(int bs, int lds)
{
    switch (bs) {
        case 128: return occ_one<MODE, 128>(lds);
        case 256: return occ_one<MODE, 256>(lds);
        case 512: return occ_one<MODE, 512>(lds);
        default: return occ_one<MODE, 1024>(lds);
    }
}

int batch_max_threads() { return 1024; }

int batch_occupancy(int mode, int bs, int lds)
{
    if (mode == BM_LDS_ALL) return occ_mode<BM_LDS_ALL>(bs, lds);
    if (mode == BM_LDS_V) return occ_mode<BM_LDS_V>(bs, lds);
    return occ_mode<BM_GLOBAL>(bs, 0);
}

void launch_batch(int mode, int bs, int lds_bytes, int grid, const BInst *insts, BDyn *dyn, const int *work,
                  int nwork, int *head, int ksteps, BFlow *flow, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (nwork <= 0 || grid <= 0) return;
    if (mode == BM_LDS_ALL)
        launch_mode<BM_LDS_ALL>(bs, lds_bytes, grid, insts, dyn, work, nwork, head, ksteps, flow, st);
    else if (mode == BM_LDS_V)
        launch_mode<BM_LDS_V>(bs, lds_bytes, grid, insts, dyn, work, nwork, head, ksteps, flow, st);
    else
        launch_mode<BM_GLOBAL>(bs, 0, grid, insts, dyn, work, nwork, head, ksteps, flow, st);
}

void launch_batch_zero(int *p, int count, void *stream)
{
    if (count > 0) k_batch_zero<<<(count + 255) / 256, 256, 0, (hipStream_t)stream>>>(p, count);
}

void launch_batch_stamp(unsigned long long *p, void *stream)
{
    k_batch_stamp<<<1, 64, 0, (hipStream_t)stream>>>(p);
}

// ADMIT_PARTS workgroups per record, each taking every ADMIT_PARTS-th
// 256-lane stripe of the record's arrays.
constexpr int ADMIT_PARTS = 16;

__global__ __launch_bounds__(256) void k_admit_place(const AdmitRec *__restrict__ recs,
                                                     const unsigned char *__restrict__ stage,
                                                     unsigned char *__restrict__ arena, BInst *__restrict__ inst,
                                                     BDyn *__restrict__ dyn)
{
    const AdmitRec *R = recs + blockIdx.x / ADMIT_PARTS;
    const long long t = (long long)(blockIdx.x % ADMIT_PARTS) * 256 + threadIdx.x;
    constexpr long long S = (long long)ADMIT_PARTS * 256;
    const unsigned char *src = stage + R->src;
    unsigned char *dst = arena + R->dst;
    const long long n16 = (long long)(R->prefix / 16), nnz = R->nnz;
    for (long long i = t; i < n16; i += S) ((uint4 *)dst)[i] = ((const uint4 *)src)[i];
    const int32_t *lits = (const int32_t *)(src + R->lits);
    const int32_t *pos = (const int32_t *)(src + R->pos);
    const int32_t *lp = (const int32_t *)(src + R->lpos);
    int32_t *eidx = (int32_t *)(dst + R->eidx);
    // pos is a permutation of [0, nnz): the scatter covers eidx[0, nnz)
    // once, the rest of the region is zeroed
    for (long long q = t; q < (long long)R->eidx_words; q += S) {
        if (q < nnz)
            eidx[pos[q]] = (int32_t)q;
        else
            eidx[q] = 0;
    }
    if (R->pk) {
        uint32_t *pk = (uint32_t *)(dst + R->pk);
        for (long long q = t; q < (long long)R->pk_words; q += S) {
            uint32_t w = 0u;
            if (R->fill_pk && q < nnz) {
                const int32_t l = lits[q];
                w = (l < 0 ? 0x80000000u : 0u) | ((uint32_t)(abs(l) - 1) << 16) | (uint32_t)lp[q];
            }
            pk[q] = w;
        }
    }
    float *e = (float *)(dst + R->e);
    for (long long q = t; q < (long long)R->e_words; q += S) e[q] = 0.0f;
    float *cst = (float *)(dst + R->cst);
    for (long long q = t; q < (long long)R->cst_words; q += S) cst[q] = 0.0f;
    if (t == 0) {
        inst[R->slot] = R->inst;
        dyn[R->slot] = R->dyn;
    }
}

void launch_admit_place(const AdmitRec *recs, int nrec, const unsigned char *stage, unsigned char *arena,
                        BInst *inst, BDyn *dyn, void *stream)
{
    if (nrec > 0)
        k_admit_place<<<(unsigned)nrec * ADMIT_PARTS, 256, 0, (hipStream_t)stream>>>(recs, stage, arena, inst, dyn);
}

void launch_batch_materialize(const BInst *inst, const BDyn *dyn, int n, int m, float *x_out, void *stream)
{
    int mx = n > m ? n : m;
    if (mx > 0) k_batch_materialize<<<(mx + 255) / 256, 256, 0, (hipStream_t)stream>>>(inst, dyn, x_out);
}

}  // namespace dnx
