// kernels.hip -- gfx950 kernels for the memcomputing SAT integrator.
//
// The reference evaluates the circuit ODE with a clause loop that pushes each
// literal's contribution onto a per-variable heap vector, then sorts and sums
// every vector (dynex.cc:362-530), and steps with an explicit Euler ladder
// (dynex.cc:301-327, 1387-1417).  Here one step of the resident state is two
// data-parallel passes with no host round trip:
//
//   clause pass   one lane per clause.  Applies the previous step's update to
//                 the clause's Xs/Xl (recomputing C of the previous state from
//                 the gathered [V, dV] pairs, so no per-clause derivative is
//                 stored), evaluates C/G/R for the new state and scatters one
//                 contribution per literal into the variable-major slot array
//                 e[] through the precomputed transpose (no atomics).  Counts
//                 loc (C >= 1/2) and the exact sign-check failures with wave
//                 ballots.
//   variable pass one lane per variable.  Materialises V of the new state,
//                 folds its contiguous contribution segment (ascending-sorted
//                 in FP64 so the sum is bit-identical to the reference), and
//                 computes the exact ladder level f in {0..4}: the step is
//                 accepted at h = 2^-(3+f) where f is the largest per-variable
//                 level, because the ladder's fail test is monotone in h.
//
// All floating-point expressions keep the reference's evaluation order and
// are compiled without contraction (no FMA), so FP64 results match a strict
// IEEE build of the reference bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include "dmm_math.hpp"

#pragma clang fp contract(off)

namespace dnx {

constexpr int CBS = 256;  // clause-pass block size

// ---------------------------------------------------------------------------
// Resident-state step: clause pass.
// ---------------------------------------------------------------------------
template <typename T, bool U3>
__global__ __launch_bounds__(CBS) void k_step_clause(DevProb P, PT<T> p, const Pair<T> *__restrict__ vd,
                                                     T *__restrict__ Xs, T *__restrict__ Xl,
                                                     T *__restrict__ e, Ctl *__restrict__ ctl, int slot)
{
    if (ctl->status != DNX_RUNNING) return;
    const int ps = (slot + 2) % 3;
    const int done = ctl->done[ps];
    if (done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) ctl->status = done;
        return;
    }
    const bool pend = ctl->stepped[ps] != 0;
    const int plev = dev_level(ctl, ps);
    const T h = pend ? level_h<T>(plev) : T(0.0);
    if (pend && blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->t += level_dt(plev);
        ctl->steps += 1;
    }
    const int c = blockIdx.x * CBS + threadIdx.x;
    int unsolved = 0, locu = 0;
    if (U3 && c < P.m) {
        // one gather of each [V, dV] pair serves both the previous state's C
        // (lazy memory update) and the new state's V
        const int32_t l0 = P.lits[3 * c], l1 = P.lits[3 * c + 1], l2 = P.lits[3 * c + 2];
        const Pair<T> a0 = vd[abs(l0) - 1], a1 = vd[abs(l1) - 1], a2 = vd[abs(l2) - 1];
        T xs = Xs[c], xl = Xl[c];
        T r0 = a0.v, r1 = a1.v, r2 = a2.v;
        if (pend) {
            memory_update(p, clause3_C_vals<T>(l0, l1, l2, r0, r1, r2), h, xs, xl);
            Xs[c] = xs;
            Xl[c] = xl;
            r0 = clampV(a0.v + h * a0.d);
            r1 = clampV(a1.v + h * a1.d);
            r2 = clampV(a2.v + h * a2.d);
        }
        T *ep = e + 3 * c;
        auto emit = [&](int j, T val) { ep[j] = val; };
        bool sat;
        T C = clause3_vals<T>(p, l0, l1, l2, r0, r1, r2, clamp01(xs), clampXl(xl, p.xlmax), emit, sat);
        locu = !(C < T(0.5));
        unsolved = locu || !sat;
    } else if (!U3 && c < P.m) {
        const int b = P.cp[c];
        const int k = P.cp[c + 1] - b;
        const int32_t *cl = P.lits + b;
        T xs = Xs[c], xl = Xl[c];
        auto prevV = [&](int v) { return vd[v].v; };
        auto curV = [&](int v) {
            Pair<T> a = vd[v];
            return pend ? clampV(a.v + h * a.d) : a.v;
        };
        if (pend) {
            T Cp = clauseK_C<T>(cl, k, prevV);
            memory_update(p, Cp, h, xs, xl);
            Xs[c] = xs;
            Xl[c] = xl;
        }
        const T cxs = clamp01(xs), cxl = clampXl(xl, p.xlmax);
        T *ep = e + b;
        auto emit = [&](int j, T val) { ep[j] = val; };
        bool sat;
        T C = clauseK<T>(p, cl, k, cxs, cxl, curV, emit, sat);
        locu = !(C < T(0.5));  // dynex.cc:499
        unsolved = locu || !sat;
    }
    if (ctl->track_loc) {
        __shared__ int shl[CBS / 64];
        add_loc<CBS>(ctl, slot, locu, shl);
    }
    if (__syncthreads_or(unsolved) && threadIdx.x == 0) raise_flag(ctl->flags[slot][FG_UNSOLVED]);
}

// ---------------------------------------------------------------------------
// Resident-state step, 3-SAT with the grouped exchange: clause pass.
//
// Each wave takes chunks of 64 consecutive clauses (GX_CHUNK), GX_U chunks at
// a time with every load of those chunks issued before any is used.  A
// literal's contribution goes to e[gbase[chunk][g] + offset]: the
// contributions are stored group by group (group g = the XCD whose
// variable-pass workgroups fold them), so the fold on each XCD gathers from
// a region of 1/8 of e that stays in that XCD's L2, instead of from all of e
// (one line from beyond L2 per 4-byte value).  Inside a chunk the 64 lanes'
// stores of one group land in one run of consecutive positions.  The control
// words are read once per workgroup, all requested before any is used.
// ---------------------------------------------------------------------------
constexpr int GX_U = 2;
constexpr int GX_BS = 256;

__device__ __forceinline__ int gx_var(uint32_t w) { return (int)((w >> 11) & 0xFFFFFu); }
__device__ __forceinline__ int gx_lit(uint32_t w) { return (w >> 31) ? -1 : 1; }  // sign as a +-1 literal

template <typename T>
__global__ __launch_bounds__(GX_BS) void k_gx_clause(int m, const uint32_t *__restrict__ gw,
                                                     const int32_t *__restrict__ gbase, PT<T> p,
                                                     const Pair<T> *__restrict__ vd, T *__restrict__ Xs,
                                                     T *__restrict__ Xl, T *__restrict__ e, Ctl *__restrict__ ctl,
                                                     int slot)
{
    const int ps = (slot + 2) % 3;
    const int status = ctl->status, done = ctl->done[ps], stepped = ctl->stepped[ps], track = ctl->track_loc;
    const int plev = dev_level(ctl, ps);
    if (status != DNX_RUNNING) return;
    if (done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) ctl->status = done;
        return;
    }
    const bool pend = stepped != 0;
    const T h = pend ? level_h<T>(plev) : T(0.0);
    if (pend && blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->t += level_dt(plev);
        ctl->steps += 1;
    }
    const int lane = threadIdx.x & 63;
    const int nw = gridDim.x * (GX_BS / 64);
    const int nchunk = (m + GX_CHUNK - 1) / GX_CHUNK;
    int unsolved = 0, locu = 0;
    for (int k0 = blockIdx.x * (GX_BS / 64) + (threadIdx.x >> 6); k0 < nchunk; k0 += GX_U * nw) {
        uint32_t W[GX_U][3];
        T XS[GX_U], XL[GX_U];
        int GB[GX_U];
        bool ok[GX_U];
#pragma unroll
        for (int u = 0; u < GX_U; u++) {
            const int k = k0 + u * nw;
            const int c = k * GX_CHUNK + lane;
            ok[u] = k < nchunk && c < m;
            W[u][0] = W[u][1] = W[u][2] = 0u;
            XS[u] = XL[u] = T(0.0);
            GB[u] = 0;
            if (ok[u]) {
                W[u][0] = gw[3 * c];
                W[u][1] = gw[3 * c + 1];
                W[u][2] = gw[3 * c + 2];
                XS[u] = Xs[c];
                XL[u] = Xl[c];
            }
            if (k < nchunk && lane < GX_GROUPS) GB[u] = gbase[k * GX_GROUPS + lane];
        }
        Pair<T> A[GX_U][3];
#pragma unroll
        for (int u = 0; u < GX_U; u++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                A[u][j].v = A[u][j].d = T(0.0);
                if (ok[u]) A[u][j] = vd[gx_var(W[u][j])];
            }
#pragma unroll
        for (int u = 0; u < GX_U; u++) {
            // positions of the three contributions: the run start of each
            // literal's group in this chunk (lane g holds group g's), plus
            // the literal's offset
            int pos[3];
#pragma unroll
            for (int j = 0; j < 3; j++)
                pos[j] = __shfl(GB[u], (int)((W[u][j] >> 8) & 7u), 64) + (int)(W[u][j] & 0xFFu);
            if (!ok[u]) continue;
            const int c = (k0 + u * nw) * GX_CHUNK + lane;
            const int32_t l0 = gx_lit(W[u][0]), l1 = gx_lit(W[u][1]), l2 = gx_lit(W[u][2]);
            T xs = XS[u], xl = XL[u];
            T r0 = A[u][0].v, r1 = A[u][1].v, r2 = A[u][2].v;
            if (pend) {
                memory_update(p, clause3_C_vals<T>(l0, l1, l2, r0, r1, r2), h, xs, xl);
                Xs[c] = xs;
                Xl[c] = xl;
                r0 = clampV(A[u][0].v + h * A[u][0].d);
                r1 = clampV(A[u][1].v + h * A[u][1].d);
                r2 = clampV(A[u][2].v + h * A[u][2].d);
            }
            auto emit = [&](int j, T val) { e[pos[j]] = val; };
            bool sat;
            const T C = clause3_vals<T>(p, l0, l1, l2, r0, r1, r2, clamp01(xs), clampXl(xl, p.xlmax), emit, sat);
            const int lu = !(C < T(0.5));
            locu += lu;
            unsolved |= lu || !sat;
        }
    }
    if (track) {
        __shared__ int shl[GX_BS / 64];
        add_loc_count<GX_BS>(ctl, slot, locu, shl);
    }
    if (__syncthreads_or(unsolved) && threadIdx.x == 0) raise_flag(ctl->flags[slot][FG_UNSOLVED]);
}

// ---------------------------------------------------------------------------
// Resident-state step, 3-SAT with the sliced exchange (see sliced_tables).
//
// Clause pass: one workgroup per unit.  The unit's clauses read V only from
// its (at most three) slices, which the workgroup first copies into LDS --
// V of the state being evaluated: the candidate of the pending step's level
// that the fold wrote (vnext), or the stored V at a launch's very first
// state -- so the three V reads of a clause are LDS reads instead of
// gathers from device memory.  The memory update uses the clause's C of the
// previous state, which the previous pass stored (cst).  Each sub-batch's
// contributions are staged in LDS ordered by destination fold block and
// leave as one contiguous run per block (coalesced stores).
//
// Fold: one workgroup per block of SX_F variables.  Its slots form one
// contiguous region of e; the workgroup reads the region (coalesced),
// places each value at its variable-major slot in LDS (sq), then each
// thread folds its variable's slots in literal order -- sorted ascending
// first in FP64 -- and writes (V, dV) and the four candidates of V of the
// next state.
// ---------------------------------------------------------------------------
template <typename T>
struct SXShape;
template <>
struct SXShape<float> {
    static constexpr int BS = 1024, S = 10000, SUB = 1024;
};
template <>
struct SXShape<double> {
    static constexpr int BS = 512, S = 5000, SUB = 512;
};

// The clause pass of one unit `ub` of one instance (the kernels below call
// it with their block index, or with the block's entry of a map over many
// instances).
template <typename T>
__device__ __forceinline__ void sx_clause_body(int n, int vstride, int m, long long esink,
                                               const uint32_t *__restrict__ sw, const int32_t *__restrict__ units,
                                               const int32_t *__restrict__ sbrun, const int2 *__restrict__ runs,
                                               PT<T> p, const Pair<T> *__restrict__ vd,
                                               const T *__restrict__ vnext, T *__restrict__ Xs, T *__restrict__ Xl,
                                               T *__restrict__ cst, T *__restrict__ e, Ctl *__restrict__ ctl,
                                               int slot, int ub)
{
    constexpr int BS = SXShape<T>::BS, S = SXShape<T>::S, SUB = SXShape<T>::SUB;
    static_assert(SUB == BS, "one clause per thread and sub-batch");
    static_assert((S * sizeof(T)) % 16 == 0, "slices of whole 16-byte words");
    // V of the unit's slices; two staging buffers (the write-out of one
    // sub-batch runs beside the evaluation of the next) and three run tables
    __shared__ __align__(16) T sv[3 * S];
    __shared__ T st[2][3 * SUB];
    __shared__ unsigned char smap[2][3 * SUB];
    __shared__ int2 sr[3][SX_MAXRUNS];
    __shared__ int s_sbr[65];
    const int tid = threadIdx.x;
    // The control words are requested first: loads complete in the order
    // they were issued, so the test of the status waits for them alone.
    const int ps = (slot + 2) % 3;
    const int status = ctl->status, done = ctl->done[ps], stepped = ctl->stepped[ps], track = ctl->track_loc;
    const int plev = dev_level(ctl, ps);
    const long long steps0 = ctl->steps;
    const int32_t *U = units + 8 * ub;
    const int c0 = U[0], c1 = U[1], nsl = U[2], sl0 = U[3], sl1 = U[4], sl2 = U[5], sb0 = U[6], run0 = U[7];
    // sub-batches of equal size (the planner's sx_unit_sub), at most SUB
    const int nsb = sx_unit_batches(c1 - c0, SUB), sub = sx_unit_sub(c1 - c0, SUB);
    // Every device-memory access of the sub-batch loop is made on every path
    // (indices clamped into the arrays, stores of lanes without a clause sent
    // to the sink entry one past the end of Xs, Xl, cst and e; the run and
    // sub-batch tables carry padding), so the waits for the next sub-batch's
    // operands count a fixed number of later accesses and leave stores in
    // flight.
    struct Ops {
        uint32_t w[3];
        T xs, xl, cp;
        int2 run;
    };
    // indices past the unit's last clause are clamped to it (not to the
    // last clause of the problem): the lanes without a clause, and the
    // look-ahead past the last sub-batch, then re-read the unit's own lines
    // instead of pulling the next unit's, which another XCD reads anyway
    const int c_last = min(max(c0, c1 - 1), m - 1);
    auto load_ops = [&](int cb, int r0, Ops &o) {
        const int li = min(cb + tid, c_last);
        o.run = runs[r0 + min(tid, SX_MAXRUNS - 1)];
        o.w[0] = sw[3 * li];
        o.w[1] = sw[3 * li + 1];
        o.w[2] = sw[3 * li + 2];
        o.xs = Xs[li];
        o.xl = Xl[li];
        o.cp = cst[li];
    };
    // the operands of the first sub-batch
    Ops cur;
    load_ops(c0, run0, cur);
    // the unit's sub-batch table (first run of each sub-batch), into LDS
    const int sbr = sbrun[sb0 + min(tid, nsb + 1)];
    const int nr0 = sbrun[sb0 + 1] - run0;
    // V of the unit's slices at ladder level 0, requested before the control
    // words are known (the level is 0 at almost every step; a step at another
    // level requests its own level's candidates below): 16-byte words from
    // the fold's candidates (each level's array padded to whole words)
    typedef T T16 __attribute__((ext_vector_type(16 / sizeof(T))));
    constexpr int WPS = S * (int)sizeof(T) / 16;  // 16-byte words per slice
    constexpr int SR = (WPS + BS - 1) / BS;
    const int wmax = vstride * (int)sizeof(T) / 16 - 1;
    T16 x[3][SR];
    auto stage_loads = [&](int lev) {
        const T16 *src = (const T16 *)(vnext + (size_t)lev * vstride);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int base = (k == 0 ? sl0 : k == 1 ? sl1 : sl2);
#pragma unroll
            for (int u = 0; u < SR; u++) x[k][u] = src[min(max(base, 0) * WPS + tid + u * BS, wmax)];
        }
    };
    stage_loads(0);
    if (status != DNX_RUNNING) return;
    if (done) {
        if (ub == 0 && tid == 0) ctl->status = done;
        return;
    }
    const bool pend = stepped != 0;
    const T h = pend ? level_h<T>(plev) : T(0.0);
    // FP32 from the third state on: V comes clamped from the fold and the
    // memories from a memory update, so the steady-state forms apply (the
    // same bits with fewer instructions, see clause3_steady)
    const bool steady = sizeof(T) == 4 && pend && steps0 >= 1;
    if (pend && ub == 0 && tid == 0) {
        ctl->t += level_dt(plev);
        ctl->steps += 1;
    }
    // V of this state over the unit's slices: every load of a thread's share
    // requested before the first LDS store; the candidates of the pending
    // step's level, or the stored V at a launch's very first state
    if (pend) {
        if (plev != 0) stage_loads(plev < 3 ? plev : 3);
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int u = 0; u < SR; u++) {
                const int i = tid + u * BS;
                if (k < nsl && i < WPS) *(T16 *)(sv + k * S + i * (16 / (int)sizeof(T))) = x[k][u];
            }
    } else {
        constexpr int SRV = (S + BS - 1) / BS;
        T xv[3][SRV];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int base = (k == 0 ? sl0 : k == 1 ? sl1 : sl2) * S;
#pragma unroll
            for (int u = 0; u < SRV; u++) xv[k][u] = vd[min(max(base, 0) + tid + u * BS, n - 1)].v;
        }
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int u = 0; u < SRV; u++) {
                const int i = tid + u * BS;
                if (k < nsl && i < S) sv[k * S + i] = xv[k][u];
            }
    }
    if (tid <= nsb + 1) s_sbr[tid] = sbr;
    if (tid < nr0) sr[0][tid] = cur.run;
    __syncthreads();  // V staged, the sub-batch table and the first runs in LDS
    int unsolved = 0, locu = 0;
    // Iteration j evaluates sub-batch j into staging buffer j & 1 and writes
    // out sub-batch j - 1 from the other buffer; one barrier per iteration.
    // The operands of sub-batch j + 1 are requested at iteration j.
    for (int j = 0; j <= nsb; j++) {
        const int b = j & 1;
        Ops nxt;
        load_ops(c0 + (j + 1) * sub, s_sbr[min(j + 1, nsb)], nxt);
        if (j < nsb) {
            const int cb = c0 + j * sub, ce = min(c1, cb + sub);
            const int2 *R = sr[j % 3];
            const bool ok = cb + tid < ce;
            T xs = cur.xs, xl = cur.xl, C = T(0.0);
            if (ok) {
                const uint32_t w0 = cur.w[0], w1 = cur.w[1], w2 = cur.w[2];
                // staging slot of each literal: its run's start plus its
                // offset; the slot also records the run for the write-out
                int sp[3];
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    const uint32_t w = q == 0 ? w0 : q == 1 ? w1 : w2;
                    const int rid = (int)((w >> 7) & 0xFFu);
                    sp[q] = (R[rid].x >> 16) + (int)(w & 0x7Fu);
                    smap[b][sp[q]] = (unsigned char)rid;
                }
                const int32_t l0 = (w0 >> 31) ? -1 : 1, l1 = (w1 >> 31) ? -1 : 1, l2 = (w2 >> 31) ? -1 : 1;
                const T r0v = sv[(w0 >> 15) & 0xFFFFu], r1v = sv[(w1 >> 15) & 0xFFFFu],
                        r2v = sv[(w2 >> 15) & 0xFFFFu];
                auto emit = [&](int q, T val) { st[b][sp[q]] = val; };
                bool sat = true;
                if constexpr (sizeof(T) == 4) {
                    if (steady) {
                        memory_update_steady(p, cur.cp, h, xs, xl);
                        // only bit 31 (the sign) of the words is read
                        C = clause3_steady(p, (int32_t)w0, (int32_t)w1, (int32_t)w2, r0v, r1v, r2v, xs, xl, emit);
                    }
                }
                if (!steady) {
                    if (pend) memory_update(p, cur.cp, h, xs, xl);
                    C = clause3_vals<T>(p, l0, l1, l2, r0v, r1v, r2v, clamp01(xs), clampXl(xl, p.xlmax), emit,
                                        sat);
                }
                const int lu = !(C < T(0.5));
                locu += lu;
                unsolved |= lu || !sat;
            }
            // a state without a pending step writes its memories back unchanged
            const int si = ok ? cb + tid : m;
            Xs[si] = xs;
            Xl[si] = xl;
            cst[si] = C;
        }
        if (j >= 1) {
            // sub-batch j - 1: one contiguous run per fold block
            const int cb = c0 + (j - 1) * sub, cnt = 3 * (min(c1, cb + sub) - cb);
            const int2 *R = sr[(j - 1) % 3];
#pragma unroll
            for (int u = 0; u < 3; u++) {
                const int k = tid + u * BS;
                const int kk = k < cnt ? k : 0;
                const int2 r = R[smap[b ^ 1][kk]];
                e[k < cnt ? (long long)r.y + (kk - (r.x >> 16)) : esink] = st[b ^ 1][kk];
            }
        }
        // the runs of sub-batch j + 1 (sub-batch j - 1's table is free now
        // that this iteration's write-out has read it)
        if (j + 1 < nsb && tid < s_sbr[j + 2] - s_sbr[j + 1]) sr[(j + 1) % 3][tid] = nxt.run;
        __syncthreads();
        cur = nxt;
    }
    if (track) {
        __shared__ int shl[BS / 64];
        add_loc_count<BS>(ctl, slot, locu, shl);
    }
    if (__syncthreads_or(unsolved) && tid == 0) raise_flag(ctl->flags[slot][FG_UNSOLVED]);
}

template <typename T>
__global__ __launch_bounds__(SXShape<T>::BS) void k_sx_clause(int n, int vstride, int m, long long esink,
                                                             const uint32_t *__restrict__ sw,
                                                             const int32_t *__restrict__ units,
                                                             const int32_t *__restrict__ sbrun,
                                                             const int2 *__restrict__ runs, PT<T> p,
                                                             const Pair<T> *__restrict__ vd,
                                                             const T *__restrict__ vnext, T *__restrict__ Xs,
                                                             T *__restrict__ Xl, T *__restrict__ cst,
                                                             T *__restrict__ e, Ctl *__restrict__ ctl, int slot)
{
    sx_clause_body<T>(n, vstride, m, esink, sw, units, sbrun, runs, p, vd, vnext, Xs, Xl, cst, e, ctl, slot,
                      (int)blockIdx.x);
}

// Many instances in one launch: block b runs unit map[b].y of instance
// map[b].x.
template <typename T>
__global__ __launch_bounds__(SXShape<T>::BS) void k_sx_clause_multi(const SXDesc *__restrict__ d,
                                                                   const int2 *__restrict__ map, int slot)
{
    const int2 mb = map[blockIdx.x];
    const SXDesc &D = d[mb.x];
    sx_clause_body<T>(D.n, D.vstride, D.m, (long long)((D.n + SX_F - 1) / SX_F) * D.stride, D.sw, D.units, D.sbrun,
                      (const int2 *)D.runs,
                      to_pt<T>(D.prm), (const Pair<T> *)D.vd, (const T *)D.vnext, (T *)D.xs, (T *)D.xl,
                      (T *)D.cst, (T *)D.e, D.ctl, slot, mb.y);
}

// The variable pass of fold block `vb` of one instance.  The block's region
// of e is [vb stride, (vb + 1) stride): every region is padded to the same
// length (its padding's slots name the sink slot `stride`), so the region's
// first loads go out with the control words, before anything is known.
template <typename T, bool SORT>
__device__ __forceinline__ void sx_var_body(int n, int vstride, int stride, const int32_t *__restrict__ row_ptr,
                                            const uint16_t *__restrict__ sq, const T *__restrict__ e,
                                            Pair<T> *__restrict__ vd, T *__restrict__ vnext,
                                            const int8_t *__restrict__ unit, Ctl *__restrict__ ctl, int slot, int vb)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T *sv = (T *)smem;
    __shared__ int sh[SX_F / 64];
    const int tid = threadIdx.x;
    // SU values per thread in flight (one round for the usual region sizes)
    constexpr int SU = sizeof(T) == 4 ? 32 : 16;
    const long long s0 = (long long)vb * stride;
    T x[SU];
    int q[SU];
#pragma unroll
    for (int u = 0; u < SU; u++) {
        const int k = tid + u * SX_F;
        q[u] = k < stride ? (int)sq[s0 + k] : stride;
        x[u] = k < stride ? e[s0 + k] : T(0.0);
    }
    // everything else that does not depend on another load: the control
    // words, the block's first slot, this thread's slot range and (V, dV)
    // pair
    const int ps = (slot + 2) % 3;
    const int status = ctl->status, pdone = ctl->done[ps], stepped = ctl->stepped[ps];
    const int plev = dev_level(ctl, ps);
    const bool unsolved_any = any_flag(ctl->flags[slot][FG_UNSOLVED]);
    const int stop = ctl->stop_on_solve, track = ctl->track_loc;
    const long long steps = ctl->steps, cap = ctl->cap;
    const int v0 = vb * SX_F;
    const int v = v0 + tid;
    const int rb0 = row_ptr[v0];
    const int rb = v < n ? row_ptr[v] : 0, re = v < n ? row_ptr[v + 1] : 0;
    Pair<T> a;
    a.v = a.d = T(0.0);
    if (v < n) a = vd[v];
    if (status != DNX_RUNNING || pdone) return;
    const bool pend = stepped != 0;
    const T h = pend ? level_h<T>(plev) : T(0.0);
    const bool solved = stop && !unsolved_any;
    const int done = solved ? DNX_SAT : (steps >= cap ? DNX_UNKNOWN : 0);
    if (vb == 0 && tid < 64 && track) fold_loc(ctl, slot);
    if (!done) {
#pragma unroll
        for (int u = 0; u < SU; u++)
            if (tid + u * SX_F < stride) sv[q[u]] = x[u];
        // regions longer than SU workgroups' worth: further rounds
        for (int k0 = tid + SU * SX_F; k0 < stride; k0 += SU * SX_F) {
#pragma unroll
            for (int u = 0; u < SU; u++) {
                const int k = k0 + u * SX_F;
                q[u] = k < stride ? (int)sq[s0 + k] : stride;
                x[u] = k < stride ? e[s0 + k] : T(0.0);
            }
#pragma unroll
            for (int u = 0; u < SU; u++)
                if (k0 + u * SX_F < stride) sv[q[u]] = x[u];
        }
    }
    __syncthreads();
    int f = 0;
    // the wave's largest segment, from every lane (before any lane branches)
    const int dmax = SORT ? wave_max(re - rb) : 0;
    if (v < n) {
        const T V = pend ? clampV(a.v + h * a.d) : a.v;
        T dV = T(0.0);
        if (!done) {
            T *sg = sv + (rb - rb0);
            const int deg = re - rb;
            T acc = T(0.0);
            if (SORT) {
                // dynex.cc:527-530: ascending sort, left fold from +0
                acc = sorted_sum_lds<T>(sg, deg, dmax);
            } else {
                for (int j = 0; j < deg; j++) acc = acc + sg[j];
            }
            dV = acc;
            if (unit && unit[v] != 0) dV = T(0.0);  // dynex.cc:306
            f = fail_level(V, dV);
#pragma unroll
            for (int L = 0; L < 4; L++) vnext[(size_t)L * vstride + v] = clampV(V + level_h<T>(L) * dV);
        }
        Pair<T> o;
        o.v = V;
        o.d = dV;
        vd[v] = o;
    }
    const int ns = (slot + 1) % 3;
    if (done) {
        if (vb == 0) {
            if (tid == 0) ctl->done[slot] = done;
            clear_slot(ctl, ns);
        }
        return;
    }
    const int bm = block_max_int<SX_F>(f, sh);
    if (tid == 0)
        for (int L = 0; L < bm; L++) raise_flag(ctl->flags[slot][FG_LEVEL + L]);
    if (vb == 0) {
        if (tid == 0) ctl->stepped[slot] = 1;
        clear_slot(ctl, ns);
    }
}

template <typename T, bool SORT>
__global__ __launch_bounds__(SX_F) void k_sx_var(int n, int vstride, int stride, const int32_t *__restrict__ row_ptr,
                                                 const uint16_t *__restrict__ sq, const T *__restrict__ e,
                                                 Pair<T> *__restrict__ vd, T *__restrict__ vnext,
                                                 const int8_t *__restrict__ unit, Ctl *__restrict__ ctl, int slot)
{
    sx_var_body<T, SORT>(n, vstride, stride, row_ptr, sq, e, vd, vnext, unit, ctl, slot, (int)blockIdx.x);
}

// Many instances in one launch: block b folds block map[b].y of instance
// map[b].x (dynamic LDS sized for the largest region among them).
template <typename T, bool SORT>
__global__ __launch_bounds__(SX_F) void k_sx_var_multi(const SXDesc *__restrict__ d, const int2 *__restrict__ map,
                                                       int slot)
{
    const int2 mb = map[blockIdx.x];
    const SXDesc &D = d[mb.x];
    sx_var_body<T, SORT>(D.n, D.vstride, D.stride, D.row_ptr, D.sq, (const T *)D.e, (Pair<T> *)D.vd, (T *)D.vnext,
                         D.unit, D.ctl, slot, mb.y);
}

// ---------------------------------------------------------------------------
// Resident-state step: variable pass.
// ---------------------------------------------------------------------------
template <typename T, bool SORT, int BS, int CAPW>
__global__ __launch_bounds__(BS) void k_step_var(DevProb P, Pair<T> *__restrict__ vd, const T *__restrict__ e,
                                                 Ctl *__restrict__ ctl, int slot)
{
    __shared__ T stage[BS / 64][CAPW];
    __shared__ int sh[BS / 64];
    if (ctl->status != DNX_RUNNING) return;
    const int ps = (slot + 2) % 3;
    if (ctl->done[ps]) return;
    const bool pend = ctl->stepped[ps] != 0;
    const T h = pend ? level_h<T>(dev_level(ctl, ps)) : T(0.0);
    const bool solved = ctl->stop_on_solve && !any_flag(ctl->flags[slot][FG_UNSOLVED]);
    const int done = solved ? DNX_SAT : (ctl->steps >= ctl->cap ? DNX_UNKNOWN : 0);
    if (blockIdx.x == 0 && threadIdx.x < 64 && ctl->track_loc) fold_loc(ctl, slot);
    const int w = threadIdx.x >> 6;
    const int v0 = (blockIdx.x * (BS / 64) + w) * 64;
    const int v = v0 + (threadIdx.x & 63);
    const T dVf = wave_fold<T, SORT, CAPW>(e, P.eidx, P.row_ptr, P.n, v0, !done, stage[w], nullptr);
    int f = 0;
    if (v < P.n) {
        Pair<T> a = vd[v];
        T V = pend ? clampV(a.v + h * a.d) : a.v;
        T dV = T(0.0);
        if (!done) {
            dV = dVf;
            if (P.unit && P.unit[v] != 0) dV = T(0.0);  // pinned (dynex.cc:306)
            f = fail_level(V, dV);
        }
        Pair<T> o;
        o.v = V;
        o.d = dV;
        vd[v] = o;
    }
    const int ns = (slot + 1) % 3;
    if (done) {
        if (blockIdx.x == 0) {
            if (threadIdx.x == 0) ctl->done[slot] = done;
            clear_slot(ctl, ns);
        }
        return;
    }
    int bm = block_max_int<BS>(f, sh);
    if (threadIdx.x == 0)
        for (int L = 0; L < bm; L++) raise_flag(ctl->flags[slot][FG_LEVEL + L]);
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) ctl->stepped[slot] = 1;
        clear_slot(ctl, ns);
    }
}

// ---------------------------------------------------------------------------
// Function-level kernels (dnx_rhs, dnx_euler) on the reference layout.
// ---------------------------------------------------------------------------
template <typename T, bool U3>
__global__ __launch_bounds__(CBS) void k_rhs_clause(DevProb P, PT<T> p, const T *__restrict__ x,
                                                    T *__restrict__ dxdt, T *__restrict__ e,
                                                    int *__restrict__ lpart, double *__restrict__ epart)
{
    __shared__ int sh[CBS / 64];
    __shared__ double shd[CBS / 64];
    const int c = blockIdx.x * CBS + threadIdx.x;
    const int n = P.n, m = P.m;
    int unsat = 0;
    double en = 0.0;
    if (c < m) {
        const int b = U3 ? 3 * c : P.cp[c];
        const int k = U3 ? 3 : P.cp[c + 1] - b;
        const int32_t *cl = P.lits + b;
        const T xs = clamp01(x[n + c]);                    // dynex.cc:366
        const T xl = clampXl(x[n + m + c], p.xlmax);       // dynex.cc:367
        T *ep = e + b;
        auto getV = [&](int v) { return x[v]; };
        auto emit = [&](int j, T val) { ep[j] = val; };
        bool sat;
        T C = U3 ? clause3<T>(p, cl, xs, xl, getV, emit, sat)
                 : clauseK<T>(p, cl, k, xs, xl, getV, emit, sat);
        T dxs = p.beta * (xs + p.eps);
        dxdt[n + c] = dxs * (C - p.gamma);        // dynex.cc:503
        dxdt[n + m + c] = p.alpha * (C - p.delta);  // dynex.cc:506
        unsat = !(C < T(0.5));
        en = (double)C;
    }
    int bu = block_sum_int<CBS>(unsat, sh);
    en = wave_sum(en);
    if ((threadIdx.x & 63) == 0) shd[threadIdx.x >> 6] = en;
    __syncthreads();
    if (threadIdx.x == 0) {
        lpart[blockIdx.x] = bu;
        double s = 0.0;
        for (int i = 0; i < CBS / 64; i++) s += shd[i];
        epart[blockIdx.x] = s;
    }
}

template <typename T, bool SORT, int BS, int CAPW>
__global__ __launch_bounds__(BS) void k_rhs_var(DevProb P, const T *__restrict__ e, T *__restrict__ dxdt)
{
    __shared__ T stage[BS / 64][CAPW];
    const int w = threadIdx.x >> 6;
    const int v0 = (blockIdx.x * (BS / 64) + w) * 64;
    const int v = v0 + (threadIdx.x & 63);
    const T d = wave_fold<T, SORT, CAPW>(e, P.eidx, P.row_ptr, P.n, v0, true, stage[w], nullptr);
    if (v < P.n) dxdt[v] = d;
}

template <typename T>
__global__ __launch_bounds__(256) void k_euler(DevProb P, PT<T> p, const T *__restrict__ x,
                                               const T *__restrict__ dxdt, T h, T *__restrict__ xo)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long n = P.n, m = P.m;
    if (i >= n + 2 * m) return;
    T v = x[i] + h * dxdt[i];  // dynex.cc:305, 311, 317
    if (i < n) {
        if (P.unit && P.unit[i] != 0) v = T(P.unit[i]);
        v = clampV(v);
    } else if (i < n + m) {
        v = clamp01(v);
    } else {
        v = clampXl(v, p.xlmax);
    }
    xo[i] = v;
}

// state -> reference layout with the pending update applied
template <typename T, bool U3>
__global__ __launch_bounds__(256) void k_materialize(DevProb P, PT<T> p, const Pair<T> *__restrict__ vd,
                                                     const T *__restrict__ Xs, const T *__restrict__ Xl,
                                                     const Ctl *__restrict__ ctl, T *__restrict__ xo)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int n = P.n, m = P.m;
    int plev = -1;
    if (ctl->status == DNX_RUNNING) {
        const int s = (int)(ctl->steps % 3);
        if (!ctl->done[s] && ctl->stepped[s]) plev = dev_level(ctl, s);
    }
    const bool pend = plev >= 0;
    const T h = pend ? level_h<T>(plev) : T(0.0);
    if (i < n) {
        Pair<T> a = vd[i];
        xo[i] = pend ? clampV(a.v + h * a.d) : a.v;
    }
    if (i < m) {
        const int pi = P.pidx ? P.pidx[i] : i;  // sliced exchange: processing order
        T xs = Xs[pi], xl = Xl[pi];
        if (pend) {
            const int b = U3 ? 3 * i : P.cp[i];
            const int k = U3 ? 3 : P.cp[i + 1] - b;
            auto getV = [&](int v) { return vd[v].v; };
            T Cp = U3 ? clause3_C<T>(P.lits + b, getV) : clauseK_C<T>(P.lits + b, k, getV);
            memory_update(p, Cp, h, xs, xl);
        }
        xo[n + i] = xs;
        xo[n + m + i] = xl;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_load_state(DevProb P, const T *__restrict__ x, Pair<T> *__restrict__ vd,
                                                    T *__restrict__ Xs, T *__restrict__ Xl)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int n = P.n, m = P.m;
    if (i < n) {
        Pair<T> a;
        a.v = (P.unit && P.unit[i] != 0) ? T(P.unit[i]) : x[i];
        a.d = T(0.0);
        vd[i] = a;
    }
    if (i < m) {
        const int pi = P.pidx ? P.pidx[i] : i;  // sliced exchange: processing order
        Xs[pi] = x[n + i];
        Xl[pi] = x[n + m + i];
    }
}

// Plain streaming copy (16 bytes per lane per iteration): the known-size
// reference for calibrating memory-traffic counters.
__global__ __launch_bounds__(256) void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, long long count)
{
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < count; i += (long long)gridDim.x * 256)
        dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
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

static inline unsigned nblk(long long count, int bs) { return (unsigned)((count + bs - 1) / bs); }

int rhs_clause_blocks(int m) { return (int)nblk(m > 0 ? m : 1, CBS); }

void launch_copy(const void *src, void *dst, long long bytes, void *stream)
{
    const long long count = bytes / 16;
    if (count > 0) k_copy<<<4096, 256, 0, (hipStream_t)stream>>>((const uint4 *)src, (uint4 *)dst, count);
}

template <typename T>
static void rhs_impl(const DevProb &P, const DevParams &prm, const void *x, void *dxdt, void *e, int *lpart,
                     double *epart, hipStream_t st)
{
    PT<T> p = to_pt<T>(prm);
    constexpr bool SORT = sizeof(T) == 8;
    unsigned gc = nblk(P.m > 0 ? P.m : 1, CBS);
    if (P.cp == nullptr)
        k_rhs_clause<T, true><<<gc, CBS, 0, st>>>(P, p, (const T *)x, (T *)dxdt, (T *)e, lpart, epart);
    else
        k_rhs_clause<T, false><<<gc, CBS, 0, st>>>(P, p, (const T *)x, (T *)dxdt, (T *)e, lpart, epart);
    constexpr int BS = SORT ? 128 : 256;
    if (P.n > 0) k_rhs_var<T, SORT, BS, 2048><<<nblk(P.n, BS), BS, 0, st>>>(P, (const T *)e, (T *)dxdt);
}

void launch_rhs(int prec, const DevProb &P, const DevParams &prm, const void *x, void *dxdt, void *e, int *loc_part,
                double *energy_part, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (prec == DNX_FP64)
        rhs_impl<double>(P, prm, x, dxdt, e, loc_part, energy_part, st);
    else
        rhs_impl<float>(P, prm, x, dxdt, e, loc_part, energy_part, st);
}

void launch_euler(int prec, const DevProb &P, const DevParams &prm, const void *x, const void *dxdt, double h,
                  void *x_out, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    long long N = (long long)P.n + 2LL * P.m;
    if (N == 0) return;
    if (prec == DNX_FP64)
        k_euler<double><<<nblk(N, 256), 256, 0, st>>>(P, to_pt<double>(prm), (const double *)x,
                                                       (const double *)dxdt, h, (double *)x_out);
    else
        k_euler<float><<<nblk(N, 256), 256, 0, st>>>(P, to_pt<float>(prm), (const float *)x,
                                                     (const float *)dxdt, (float)h, (float *)x_out);
}

template <typename T>
static void step_impl(const DevProb &P, const DevParams &prm, void *vd, void *xs, void *xl, void *e, Ctl *ctl,
                      int slot, hipStream_t st)
{
    PT<T> p = to_pt<T>(prm);
    constexpr bool SORT = sizeof(T) == 8;
    constexpr int BS = SORT ? 128 : 256;
    if (P.cp == nullptr && P.gw) {
        // grouped exchange: GX_U chunks per wave and pass, at most 8
        // workgroups per CU's worth of waves for the whole clause set
        const long long nchunk = ((long long)P.m + GX_CHUNK - 1) / GX_CHUNK;
        const long long waves = (nchunk + GX_U - 1) / GX_U;
        const unsigned gc = (unsigned)std::max(1LL, (waves + GX_BS / 64 - 1) / (GX_BS / 64));
        k_gx_clause<T><<<gc, GX_BS, 0, st>>>(P.m, P.gw, P.gbase, p, (const Pair<T> *)vd, (T *)xs, (T *)xl, (T *)e,
                                             ctl, slot);
        DevProb Q = P;
        Q.eidx = P.geidx;
        k_step_var<T, SORT, BS, 2048><<<nblk(P.n > 0 ? P.n : 1, BS), BS, 0, st>>>(Q, (Pair<T> *)vd, (const T *)e,
                                                                                 ctl, slot);
        return;
    }
    unsigned gc = nblk(P.m > 0 ? P.m : 1, CBS);
    if (P.cp == nullptr)
        k_step_clause<T, true><<<gc, CBS, 0, st>>>(P, p, (const Pair<T> *)vd, (T *)xs, (T *)xl, (T *)e, ctl,
                                                   slot);
    else
        k_step_clause<T, false><<<gc, CBS, 0, st>>>(P, p, (const Pair<T> *)vd, (T *)xs, (T *)xl, (T *)e, ctl,
                                                    slot);
    k_step_var<T, SORT, BS, 2048><<<nblk(P.n > 0 ? P.n : 1, BS), BS, 0, st>>>(P, (Pair<T> *)vd, (const T *)e, ctl,
                                                                             slot);
}

template <typename T>
static void step_sliced_impl(const DevProb &P, const DevParams &prm, void *vd, void *xs, void *xl, void *e,
                             void *vnext, void *cst, Ctl *ctl, int slot, hipStream_t st)
{
    constexpr bool SORT = sizeof(T) == 8;
    // the clause pass's sink entry follows the fold blocks' padded regions
    const long long esink = (long long)nblk(P.n > 0 ? P.n : 1, SX_F) * P.sx_region;
    k_sx_clause<T><<<P.nunits, SXShape<T>::BS, 0, st>>>(P.n, sx_vstride(P.n), P.m, esink, P.sw,
                                                        P.units, P.sbrun, (const int2 *)P.runs,
                                                        to_pt<T>(prm), (const Pair<T> *)vd, (const T *)vnext,
                                                        (T *)xs, (T *)xl, (T *)cst, (T *)e, ctl, slot);
    // the block's region and its sink slot
    const int lds = (int)sizeof(T) * (P.sx_region + 1);
    static int lds_set[MAX_DEV_ATTR] = {};
    lds_attr((const void *)k_sx_var<T, SORT>, lds, lds_set);
    k_sx_var<T, SORT><<<nblk(P.n > 0 ? P.n : 1, SX_F), SX_F, lds, st>>>(P.n, sx_vstride(P.n), P.sx_region,
                                                                      P.row_ptr, P.sq,
                                                                      (const T *)e, (Pair<T> *)vd, (T *)vnext,
                                                                      P.unit, ctl, slot);
}

// out[k] <- the first `words` ints of instance k's control block
__global__ __launch_bounds__(64) void k_ctl_gather(const SXDesc *__restrict__ d, int count, int words,
                                                   int *__restrict__ out)
{
    const int k = blockIdx.x;
    if (k >= count) return;
    const int *src = (const int *)d[k].ctl;
    for (int i = threadIdx.x; i < words; i += 64) out[(size_t)k * words + i] = src[i];
}

void launch_ctl_gather(const SXDesc *d, int count, int words, int *out, void *stream)
{
    if (count > 0) k_ctl_gather<<<count, 64, 0, (hipStream_t)stream>>>(d, count, words, out);
}

void launch_step_sliced_multi(const SXDesc *d, const void *map_c, int nclause, const void *map_v, int nvar,
                              int lds_region, int slot, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (nclause > 0)
        k_sx_clause_multi<float><<<nclause, SXShape<float>::BS, 0, st>>>(d, (const int2 *)map_c, slot);
    const int lds = (int)sizeof(float) * (lds_region + 1);
    static int lds_set[MAX_DEV_ATTR] = {};
    lds_attr((const void *)k_sx_var_multi<float, false>, lds, lds_set);
    if (nvar > 0) k_sx_var_multi<float, false><<<nvar, SX_F, lds, st>>>(d, (const int2 *)map_v, slot);
}

void launch_step_sliced(int prec, const DevProb &P, const DevParams &prm, void *vd, void *xs, void *xl, void *e,
                        void *vnext, void *cst, Ctl *ctl, int slot, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (prec == DNX_FP64)
        step_sliced_impl<double>(P, prm, vd, xs, xl, e, vnext, cst, ctl, slot, st);
    else
        step_sliced_impl<float>(P, prm, vd, xs, xl, e, vnext, cst, ctl, slot, st);
}

void launch_step(int prec, const DevProb &P, const DevParams &prm, void *vd, void *xs, void *xl, void *e, Ctl *ctl,
                 int slot, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (prec == DNX_FP64)
        step_impl<double>(P, prm, vd, xs, xl, e, ctl, slot, st);
    else
        step_impl<float>(P, prm, vd, xs, xl, e, ctl, slot, st);
}

template <typename T>
static void materialize_impl(const DevProb &P, const DevParams &prm, const void *vd, const void *xs,
                             const void *xl, const Ctl *ctl, void *x_out, hipStream_t st)
{
    int mx = P.n > P.m ? P.n : P.m;
    if (mx == 0) return;
    if (P.cp == nullptr)
        k_materialize<T, true><<<nblk(mx, 256), 256, 0, st>>>(P, to_pt<T>(prm), (const Pair<T> *)vd,
                                                               (const T *)xs, (const T *)xl, ctl, (T *)x_out);
    else
        k_materialize<T, false><<<nblk(mx, 256), 256, 0, st>>>(P, to_pt<T>(prm), (const Pair<T> *)vd,
                                                                (const T *)xs, (const T *)xl, ctl, (T *)x_out);
}

void launch_materialize(int prec, const DevProb &P, const DevParams &prm, const void *vd, const void *xs,
                        const void *xl, const Ctl *ctl, void *x_out, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (prec == DNX_FP64)
        materialize_impl<double>(P, prm, vd, xs, xl, ctl, x_out, st);
    else
        materialize_impl<float>(P, prm, vd, xs, xl, ctl, x_out, st);
}

void launch_load_state(int prec, const DevProb &P, const void *x, void *vd, void *xs, void *xl, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    int mx = P.n > P.m ? P.n : P.m;
    if (mx == 0) return;
    if (prec == DNX_FP64)
        k_load_state<double><<<nblk(mx, 256), 256, 0, st>>>(P, (const double *)x, (Pair<double> *)vd,
                                                             (double *)xs, (double *)xl);
    else
        k_load_state<float><<<nblk(mx, 256), 256, 0, st>>>(P, (const float *)x, (Pair<float> *)vd,
                                                           (float *)xs, (float *)xl);
}

}  // namespace dnx
