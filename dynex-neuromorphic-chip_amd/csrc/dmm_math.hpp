// dmm_math.hpp -- device-side arithmetic of the memcomputing circuit shared
// by the single-instance kernels (kernels.hip) and the batched engine
// (batch.hip): clause equations, memory updates, the ladder level and the
// wave/block reductions.  Every expression keeps the reference's evaluation
// order; the including files are compiled without FP contraction.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

#include "dnx.h"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace dnx {

template <typename T>
struct alignas(2 * sizeof(T)) Pair {
    T v, d;
};

template <typename T>
struct PT {
    T alpha, beta, gamma, delta, eps, zeta, xlmax;
};

template <typename T>
__host__ __device__ static inline PT<T> to_pt(const DevParams &p)
{
    PT<T> r;
    r.alpha = (T)p.alpha;
    r.beta = (T)p.beta;
    r.gamma = (T)p.gamma;
    r.delta = (T)p.delta;
    r.eps = (T)p.epsilon;
    r.zeta = (T)p.zeta;
    r.xlmax = (T)p.xl_max;
    return r;
}

// Clamps and min.  The reference's compare-and-select (boost clamp,
// dynex.cc:313-324) equals the hardware min/max on every value the
// integrator can hold: the state is finite, and the one case where they part
// (a signed zero against a zero bound) cannot occur -- V's bounds are +-1, and
// the clause terms i = 1 - q V are never -0 (1 - 1 = +0).  clamp01 keeps the
// compare form, since Xs could in principle start at -0.
__device__ __forceinline__ float dmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float dmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double dmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ double dmax(double a, double b) { return fmax(a, b); }

template <typename T>
__device__ __forceinline__ T clampV(T v)
{
    return dmin(dmax(v, T(-1.0)), T(1.0));
}
template <typename T>
__device__ __forceinline__ T clamp01(T v)
{
    if (v < T(0.0)) v = T(0.0);
    if (v > T(1.0)) v = T(1.0);
    return v;
}
template <typename T>
__device__ __forceinline__ T clampXl(T v, T xlmax)
{
    return dmin(dmax(v, T(1.0)), xlmax);
}
template <typename T>
__device__ __forceinline__ T tmin(T a, T b)
{
    return dmin(a, b);
}

// step size of ladder level L (the fallback level 4 keeps the 1/64 state)
template <typename T>
__device__ __forceinline__ T level_h(int L)
{
    return L == 0 ? T(0.125) : L == 1 ? T(0.0625) : L == 2 ? T(0.03125) : T(0.015625);
}
__device__ __forceinline__ double level_dt(int L)
{
    return L == 0 ? 0.125 : L == 1 ? 0.0625 : L == 2 ? 0.03125 : L == 3 ? 0.015625 : 0.0078125;
}

// Smallest ladder level at which |clamp(V + h dV) - V| < 1 (dynex.cc:1403-1406).
template <typename T>
__device__ __forceinline__ int fail_level(T V, T dV)
{
#pragma unroll
    for (int L = 0; L < 4; L++) {
        T h = level_h<T>(L);
        T Vn = clampV(V + h * dV);
        if (!(fabs(Vn - V) >= T(1))) return L;
    }
    return 4;
}

// Xs/Xl Euler update of one clause from its previous state and previous C
// (dynex.cc:503, 506 for the derivatives; 313-324 for update and clamps).
template <typename T>
__device__ __forceinline__ void memory_update(const PT<T> &p, T Cprev, T h, T &xs, T &xl)
{
    T Xs = clamp01(xs);
    T dxs = p.beta * (Xs + p.eps);
    dxs = dxs * (Cprev - p.gamma);
    T dxl = p.alpha * (Cprev - p.delta);
    xs = clamp01(xs + h * dxs);
    xl = clampXl(xl + h * dxl, p.xlmax);
}

// Contribution of one literal: Xl*Xs*G + (1 + zeta*Xl)*(1 - Xs)*R
// (dynex.cc:400-402), in the reference's evaluation order.
template <typename T>
__device__ __forceinline__ T contrib(const PT<T> &p, T Xs, T Xl, T G, T R)
{
    T a = Xl * Xs;
    a = a * G;
    T b = T(1.0) + p.zeta * Xl;
    b = b * (T(1.0) - Xs);
    b = b * R;
    return a + b;
}

// The exact sign check of one literal (dynex.cc:674-687) on the raw value,
// in the state's own precision (a float's sign compares the same as its
// double's).
template <typename T>
__device__ __forceinline__ bool lit_true(int32_t lit, T x)
{
    return (x > T(0.0) && lit > 0) || (x < T(0.0) && lit < 0);
}

// ---------------------------------------------------------------------------
// Clause evaluation.  getV(var0) returns the raw state value of a variable;
// emit(j, value) receives literal j's contribution.  Returns C and sets
// `sat` to the exact sign check of the raw values (dynex.cc:674-687).
// ---------------------------------------------------------------------------
// r0..r2 are the raw state values of the clause's variables.
template <typename T, typename Emit>
__device__ __forceinline__ T clause3_vals(const PT<T> &p, int32_t l0, int32_t l1, int32_t l2, T r0, T r1, T r2, T Xs,
                                          T Xl, Emit emit, bool &sat)
{
    sat = lit_true(l0, r0) || lit_true(l1, r1) || lit_true(l2, r2);
    T q0 = l0 > 0 ? T(1) : T(-1), q1 = l1 > 0 ? T(1) : T(-1), q2 = l2 > 0 ? T(1) : T(-1);
    T V0 = clampV(r0), V1 = clampV(r1), V2 = clampV(r2);
    T i0 = T(1.0) - q0 * V0, i1 = T(1.0) - q1 * V1, i2 = T(1.0) - q2 * V2;
    T C = tmin(i0, tmin(i1, i2));
    C = C / T(2.0);
    C = clamp01(C);
    T G0 = q0 * tmin(i1, i2) / T(2.0);
    T G1 = q1 * tmin(i0, i2) / T(2.0);
    T G2 = q2 * tmin(i0, i1) / T(2.0);
    T R0 = (C == i0 / T(2.0)) ? (q0 - V0) / T(2.0) : T(0.0);
    T R1 = (C == i1 / T(2.0)) ? (q1 - V1) / T(2.0) : T(0.0);
    T R2 = (C == i2 / T(2.0)) ? (q2 - V2) / T(2.0) : T(0.0);
    emit(0, contrib(p, Xs, Xl, G0, R0));
    emit(1, contrib(p, Xs, Xl, G1, R1));
    emit(2, contrib(p, Xs, Xl, G2, R2));
    return C;
}

template <typename T>
__device__ __forceinline__ T clause3_C_vals(int32_t l0, int32_t l1, int32_t l2, T r0, T r1, T r2)
{
    T i0 = T(1.0) - (l0 > 0 ? T(1) : T(-1)) * clampV(r0);
    T i1 = T(1.0) - (l1 > 0 ? T(1) : T(-1)) * clampV(r1);
    T i2 = T(1.0) - (l2 > 0 ? T(1) : T(-1)) * clampV(r2);
    return clamp01(tmin(i0, tmin(i1, i2)) / T(2.0));
}

template <typename T, typename GetV, typename Emit>
__device__ __forceinline__ T clause3v(const PT<T> &p, int32_t l0, int32_t l1, int32_t l2, T Xs, T Xl, GetV getV,
                                      Emit emit, bool &sat)
{
    return clause3_vals<T>(p, l0, l1, l2, getV(abs(l0) - 1), getV(abs(l1) - 1), getV(abs(l2) - 1), Xs, Xl, emit,
                           sat);
}

template <typename T, typename GetV, typename Emit, typename CL>
__device__ __forceinline__ T clause3(const PT<T> &p, CL cl, T Xs, T Xl, GetV getV, Emit emit, bool &sat)
{
    return clause3v<T>(p, cl[0], cl[1], cl[2], Xs, Xl, getV, emit, sat);
}

// C of a 3-literal clause only (for the lazy memory update).
template <typename T, typename GetV>
__device__ __forceinline__ T clause3v_C(int32_t l0, int32_t l1, int32_t l2, GetV getV)
{
    return clause3_C_vals<T>(l0, l1, l2, getV(abs(l0) - 1), getV(abs(l1) - 1), getV(abs(l2) - 1));
}
template <typename T, typename GetV, typename CL>
__device__ __forceinline__ T clause3_C(CL cl, GetV getV)
{
    return clause3v_C<T>(cl[0], cl[1], cl[2], getV);
}

// Any clause length (dynex.cc:420-494 for k=2 and k>=4, 517 for k=1; the
// empty clause follows the k-SAT branch with min = INT_MAX).  Two passes over
// the literals keep everything in registers: the first finds the smallest
// i_l (and its first index) and the smallest of the rest; the second emits.
template <typename T, typename GetV, typename Emit, typename CL>
__device__ __forceinline__ T clauseK(const PT<T> &p, CL cl, int k, T Xs, T Xl, GetV getV, Emit emit, bool &sat)
{
    sat = false;
    if (k == 1) {
        sat = lit_true(cl[0], getV(abs(cl[0]) - 1));
        emit(0, T(0.0));
        return T(0.0);
    }
    T min1 = T(INT_MAX), min2 = T(INT_MAX);
    int idx1 = -1;
    for (int j = 0; j < k; j++) {
        T r = getV(abs(cl[j]) - 1);
        sat = sat || lit_true(cl[j], r);
        T I = T(1.0) - (cl[j] > 0 ? T(1) : T(-1)) * clampV(r);
        if (I < min1) {
            if (idx1 >= 0) min2 = tmin(min2, min1);
            min1 = I;
            idx1 = j;
        } else {
            min2 = tmin(min2, I);
        }
    }
    // dynex.cc:391-392 and 431-432 clamp C for k = 3 and k = 2; the k-SAT
    // branch (dynex.cc:457-470) does not, so the empty clause keeps
    // C = INT_MAX / 2
    T C = min1 / T(2.0);
    if (k == 2 || k == 3) C = clamp01(C);
    for (int j = 0; j < k; j++) {
        T q = cl[j] > 0 ? T(1) : T(-1);
        T V = clampV(getV(abs(cl[j]) - 1));
        T I = T(1.0) - q * V;
        T G = q * (j == idx1 ? min2 : min1) / T(2.0);
        T R = (C == I / T(2.0)) ? (q - V) / T(2.0) : T(0.0);
        emit(j, contrib(p, Xs, Xl, G, R));
    }
    return C;
}

template <typename T, typename GetV, typename CL>
__device__ __forceinline__ T clauseK_C(CL cl, int k, GetV getV)
{
    if (k == 1) return T(0.0);
    T min1 = T(INT_MAX);
    for (int j = 0; j < k; j++) {
        T I = T(1.0) - (cl[j] > 0 ? T(1) : T(-1)) * clampV(getV(abs(cl[j]) - 1));
        min1 = tmin(min1, I);
    }
    T C = min1 / T(2.0);
    if (k == 2 || k == 3) C = clamp01(C);  // as clauseK
    return C;
}

// ---------------------------------------------------------------------------
// Steady-state forms (FP32).  Valid when every V read lies in [-1, 1] (V is
// already clamped) and Xs, Xl come out of a memory update (already clamped).
// On such inputs they return the bits of memory_update<float> and
// clause3_vals<float>, with fewer instructions:
//  * q = +-1, so q V is exact and 1 - q V == fma(-q, V, 1) (one rounding of
//    the same exact value);
//  * i = 1 - q V lies in {0} U [2^-24, 2] (1 - x is exact for x in [1/2, 1]
//    and rounds to a multiple of 2^-24 below), so C = min(i) / 2 lies in
//    [0, 1] and its clamp01 is the identity; halving an i, a min of them or
//    q - V is exact, so C == i_j / 2 iff min(i) == i_j, q m / 2 == (q/2) m and
//    (q - V) / 2 == fma(-1/2, V, q/2), signed zeros included;
//  * C < 1/2 means some i_j < 1, i.e. q_j V_j > 0, so that literal's sign
//    check holds: a clause with C < 1/2 is satisfied by the signs.
// Only bit 31 of l0..l2 is read (the literal's sign), so they may be signed
// literals or packed literal words.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float signed_const(int32_t lit, uint32_t magnitude_bits)
{
    return __builtin_bit_cast(float, ((uint32_t)lit & 0x80000000u) | magnitude_bits);
}

__device__ __forceinline__ void memory_update_steady(const PT<float> &p, float Cprev, float h, float &xs, float &xl)
{
    float dxs = p.beta * (xs + p.eps);
    dxs = dxs * (Cprev - p.gamma);
    float dxl = p.alpha * (Cprev - p.delta);
    // clamp01 as one median of three: a memory that came out of an update is
    // never -0 (x + y is -0 only for two -0 terms, and clamp01 maps any
    // negative value to +0), the one input the two forms could tell apart
    xs = __builtin_amdgcn_fmed3f(xs + h * dxs, 0.0f, 1.0f);
    xl = clampXl(xl + h * dxl, p.xlmax);
}

template <typename Emit>
__device__ __forceinline__ float clause3_steady(const PT<float> &p, int32_t l0, int32_t l1, int32_t l2, float V0,
                                                float V1, float V2, float Xs, float Xl, Emit emit)
{
    const float q0 = signed_const(l0, 0x3f800000u), q1 = signed_const(l1, 0x3f800000u),
                q2 = signed_const(l2, 0x3f800000u);
    const float h0 = signed_const(l0, 0x3f000000u), h1 = signed_const(l1, 0x3f000000u),
                h2 = signed_const(l2, 0x3f000000u);
    const float i0 = __builtin_fmaf(-q0, V0, 1.0f), i1 = __builtin_fmaf(-q1, V1, 1.0f),
                i2 = __builtin_fmaf(-q2, V2, 1.0f);
    const float m12 = dmin(i1, i2), m02 = dmin(i0, i2), m01 = dmin(i0, i1);
    const float mn = dmin(i0, m12);
    const float R0 = mn == i0 ? __builtin_fmaf(-0.5f, V0, h0) : 0.0f;
    const float R1 = mn == i1 ? __builtin_fmaf(-0.5f, V1, h1) : 0.0f;
    const float R2 = mn == i2 ? __builtin_fmaf(-0.5f, V2, h2) : 0.0f;
    emit(0, contrib(p, Xs, Xl, h0 * m12, R0));
    emit(1, contrib(p, Xs, Xl, h1 * m02, R1));
    emit(2, contrib(p, Xs, Xl, h2 * m01, R2));
    return mn * 0.5f;
}

// s plus LDS slots [rb, re) in slot order: the reads go out eight at a time,
// the additions stay one after another (a slot past re is neither read nor
// added).
__device__ __forceinline__ float lds_fold(const float *se, int rb, int re, float s)
{
    for (int q = rb; q < re; q += 8) {
        float x[8];
#pragma unroll
        for (int u = 0; u < 8; u++) x[u] = q + u < re ? se[q + u] : 0.0f;
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (q + u < re) s = s + x[u];
    }
    return s;
}

// Clause-order fold of the contributions gathered through the transpose:
// slots [rb, re) of a variable name literal positions eidx[q] whose values
// sit in e[].  Eight independent gathers are kept in flight; the additions
// stay in slot order.
template <typename T, typename EP, typename IP>
__device__ __forceinline__ T fold_gather(EP e, IP eidx, int rb, int re)
{
    T acc = T(0.0);
    for (int q = rb; q < re; q += 8) {
        // eight gathers in flight, the last round predicated
        int i[8];
        T x[8];
#pragma unroll
        for (int u = 0; u < 8; u++) i[u] = q + u < re ? eidx[q + u] : -1;
#pragma unroll
        for (int u = 0; u < 8; u++) x[u] = i[u] >= 0 ? e[i[u]] : T(0.0);
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (q + u < re) acc = acc + x[u];
    }
    return acc;
}

// ---------------------------------------------------------------------------
// wave / block reductions (64-wide wavefronts)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int BS>
__device__ __forceinline__ int block_sum_int(int v, int *sh)
{
    // v is 0/1 per lane: ballot + popcount per wave
    unsigned long long b = __ballot(v);
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = __popcll(b);
    __syncthreads();
    int r = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < BS / 64; i++) r += sh[i];
    return r;  // valid on thread 0
}

template <int BS>
__device__ __forceinline__ int block_max_int(int v, int *sh)
{
    v = wave_max(v);
    int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    int r = -1;
    if (threadIdx.x == 0)
        for (int i = 0; i < BS / 64; i++) r = max(r, sh[i]);
    return r;  // valid on thread 0
}

// Flag groups (see Ctl): a workgroup raises its own entry with a plain
// store; every wave reads all 64 entries and ballots.
__device__ __forceinline__ void raise_flag(int *group)
{
    const int b = blockIdx.x;
    group[(b & 7) * 32 + ((b >> 3) & 7)] = 1;
}
__device__ __forceinline__ bool any_flag(const int *group)
{
    const int l = threadIdx.x & 63;
    return __ballot(group[(l & 7) * 32 + (l >> 3)] != 0) != 0ull;
}
// loc tracking: workgroup b adds its count of clauses with C >= 1/2 to
// entry (b & 7) * 32 + ((b >> 3) & 7) of the slot's counters (a workgroup
// whose count is 0 adds nothing); called by every thread of the workgroup
template <int BS>
__device__ __forceinline__ void add_loc(Ctl *c, int s, int unsat, int *sh)
{
    const int cnt = block_sum_int<BS>(unsat, sh);
    if (threadIdx.x == 0 && cnt > 0) {
        const int b = blockIdx.x;
        atomicAdd(&c->locc[s][(b & 7) * 32 + ((b >> 3) & 7)], cnt);
    }
}
// the same for a per-thread count of clauses (a thread that evaluates
// several clauses)
template <int BS>
__device__ __forceinline__ void add_loc_count(Ctl *c, int s, int count, int *sh)
{
    int v = count;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int cnt = 0;
        for (int i = 0; i < BS / 64; i++) cnt += sh[i];
        if (cnt > 0) {
            const int b = blockIdx.x;
            atomicAdd(&c->locc[s][(b & 7) * 32 + ((b >> 3) & 7)], cnt);
        }
    }
}
// sums slot s's counters into loc_last and loc_min (one wave of one
// workgroup calls it)
__device__ __forceinline__ void fold_loc(Ctl *c, int s)
{
    const int l = threadIdx.x & 63;
    int v = c->locc[s][(l & 7) * 32 + (l >> 3)];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (l == 0) {
        c->loc_last = v;
        if (v < c->loc_min) c->loc_min = v;
    }
}

__device__ __forceinline__ int dev_level(const Ctl *c, int s)
{
    int L = 0;
#pragma unroll
    for (int g = 0; g < 4; g++) L += any_flag(c->flags[s][FG_LEVEL + g]) ? 1 : 0;
    return L;
}
// clears slot s (called by all threads of one workgroup)
__device__ __forceinline__ void clear_slot(Ctl *c, int s)
{
    for (int i = threadIdx.x; i < FG_COUNT * FLAG_USED; i += blockDim.x) {
        const int g = i / FLAG_USED, l = i % FLAG_USED;
        c->flags[s][g][(l & 7) * 32 + (l >> 3)] = 0;
    }
    for (int l = threadIdx.x; l < FLAG_USED; l += blockDim.x) c->locc[s][(l & 7) * 32 + (l >> 3)] = 0;
    if (threadIdx.x == 0) {
        c->stepped[s] = 0;
        c->done[s] = 0;
    }
}

// ---------------------------------------------------------------------------
// Per-variable fold.  The variable owns transpose slots [b, b+deg); slot q
// holds the literal position eidx[q] whose contribution e[] carries (the
// clause pass writes e in literal order, coalesced).
// ---------------------------------------------------------------------------
template <typename T, bool SORT, int BS, int CAP, typename EP, typename IP>
__device__ __forceinline__ T fold_var(EP e, IP eidx, int b, int deg, T *buf)
{
    T acc = T(0.0);
    if (!SORT) return fold_gather<T>(e, eidx, b, b + deg);
    // dynex.cc:527-530: ascending std::sort then left fold from 0.  Zero
    // contributions (which the reference never pushes) do not change an
    // ascending fold from +0, so every slot is folded.
    const int t = threadIdx.x;
    if (CAP > 0 && deg <= CAP) {
        // stage the gathered values (eight loads in flight), then sort
        for (int j = 0; j < deg; j += 8) {
            int i[8];
#pragma unroll
            for (int u = 0; u < 8; u++) i[u] = j + u < deg ? eidx[b + j + u] : -1;
            T x[8];
#pragma unroll
            for (int u = 0; u < 8; u++) x[u] = i[u] >= 0 ? e[i[u]] : T(0.0);
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (j + u < deg) buf[(j + u) * BS + t] = x[u];
        }
        for (int j = 1; j < deg; j++) {
            T x = buf[j * BS + t];
            int i = j;
            while (i > 0 && buf[(i - 1) * BS + t] > x) {
                buf[i * BS + t] = buf[(i - 1) * BS + t];
                i--;
            }
            buf[i * BS + t] = x;
        }
        for (int j = 0; j < deg; j++) acc = acc + buf[j * BS + t];
    } else {
        // rare high-degree variable: selection of the next larger value by
        // repeated scans (ties counted), no scratch needed
        T last = T(0.0);
        int taken = 0;
        bool first = true;
        while (taken < deg) {
            T best = T(0.0);
            bool have = false;
            for (int j = 0; j < deg; j++) {
                T x = e[eidx[b + j]];
                if ((first || x > last) && (!have || x < best)) {
                    best = x;
                    have = true;
                }
            }
            int cnt = 0;
            for (int j = 0; j < deg; j++) cnt += (e[eidx[b + j]] == best);
            for (int r = 0; r < cnt; r++) acc = acc + best;
            taken += cnt;
            last = best;
            first = false;
        }
    }
    return acc;
}

// Ascending left fold from +0 of s[0, deg) with deg <= NB (a power of two),
// through registers: the values, +0 past deg, go through a bitonic sorting
// network (every compare-exchange pair fixed at compile time), then are
// summed in order.  This is dynex.cc:527-530's std::sort then accumulate:
// equal values are the same bits except for signed zeros, and zeros -- the
// pads included -- never change an ascending fold from +0 (the running sum
// is never -0), so every ascending order gives the same bits.
template <typename T, int NB>
__device__ __forceinline__ T sorted_sum_net(const T *s, int deg)
{
    T r[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) r[j] = j < deg ? s[j] : T(0.0);
#pragma unroll
    for (int k = 2; k <= NB; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < NB; i++) {
                const int l = i ^ j;
                if (l > i) {
                    const T lo = fmin(r[i], r[l]), hi = fmax(r[i], r[l]);
                    if ((i & k) == 0) {
                        r[i] = lo;
                        r[l] = hi;
                    } else {
                        r[i] = hi;
                        r[l] = lo;
                    }
                }
            }
    T acc = T(0.0);
#pragma unroll
    for (int j = 0; j < NB; j++) acc = acc + r[j];
    return acc;
}

// Ascending fold of a lane's LDS segment s[0, deg): a sorting network sized
// for the wave's largest segment (dmax, uniform across the wave), insertion
// sort in place beyond 64.
template <typename T>
__device__ __forceinline__ T sorted_sum_lds(T *s, int deg, int dmax)
{
    if (dmax <= 8) return sorted_sum_net<T, 8>(s, deg);
    if (dmax <= 16) return sorted_sum_net<T, 16>(s, deg);
    if (dmax <= 32) return sorted_sum_net<T, 32>(s, deg);
    if (dmax <= 64) return sorted_sum_net<T, 64>(s, deg);
    for (int j = 1; j < deg; j++) {
        const T x = s[j];
        int k = j;
        while (k > 0 && s[k - 1] > x) {
            s[k] = s[k - 1];
            k--;
        }
        s[k] = x;
    }
    T acc = T(0.0);
    for (int j = 0; j < deg; j++) acc = acc + s[j];
    return acc;
}

// Wave-cooperative fold: a wave owns 64 consecutive variables, whose
// transpose slots are contiguous.  The wave stages all their contributions
// in LDS (coalesced index loads, independent gathers, eight per lane in
// flight), then each lane folds its own segment in slot order -- sorted
// first in FP64 -- exactly as a lane-serial fold would.  A wave whose
// segment exceeds the staging area folds from device memory instead.
// Must be reached by every thread of the block (it synchronises).
template <typename T, bool SORT, int CAPW, typename EP, typename IP>
__device__ __forceinline__ T wave_fold(EP e, IP eidx, IP row_ptr, int n, int v0, bool active, T *stage,
                                       T *sortbuf_unused)
{
    (void)sortbuf_unused;
    const int lane = threadIdx.x & 63;
    const int v = v0 + lane;
    const int vend = min(v0 + 64, n);
    const int sb = v0 < n ? row_ptr[v0] : 0;
    const int total = v0 < n ? row_ptr[vend] - sb : 0;
    const bool fits = total <= CAPW;
    if (active && fits) {
        // eight gathers per lane in flight, the last round predicated rather
        // than finished one gather at a time
        for (int q = lane; q < total; q += 8 * 64) {
            int i[8];
#pragma unroll
            for (int u = 0; u < 8; u++) i[u] = q + u * 64 < total ? eidx[sb + q + u * 64] : -1;
            T x[8];
#pragma unroll
            for (int u = 0; u < 8; u++) x[u] = i[u] >= 0 ? e[i[u]] : T(0.0);
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (q + u * 64 < total) stage[q + u * 64] = x[u];
        }
    }
    __syncthreads();
    T acc = T(0.0);
    const int rb = v < n ? row_ptr[v] : 0, re = v < n ? row_ptr[v + 1] : 0;
    // the wave's largest segment, from every lane (before any lane branches)
    const int dmax = SORT ? wave_max(re - rb) : 0;
    if (active && v < n) {
        if (fits) {
            T *s = stage + (rb - sb);
            const int deg = re - rb;
            if (SORT) {
                // dynex.cc:527-530: ascending sort, left fold from +0
                acc = sorted_sum_lds<T>(s, deg, dmax);
            } else {
                for (int j = 0; j < deg; j++) acc = acc + s[j];
            }
        } else if (SORT) {
            acc = fold_var<T, true, 64, 0>(e, eidx, rb, re - rb, nullptr);
        } else {
            acc = fold_gather<T>(e, eidx, rb, re);
        }
    }
    return acc;
}

// Wave-cooperative fold over contributions stored variable-major (slot q of
// the transpose at ld(q)), as wide and gang mode write them: a wave's 64
// variables own one contiguous range, staged in LDS with coalesced loads (no
// index array to wait for), then each lane folds its own slots in order.
// rb/re are the lane's slot range (re = rb = 0 past the last variable),
// `last` the index of the wave's last valid lane.  Must be reached by every
// thread of the block.
template <int CAPW, typename LD>
__device__ __forceinline__ float wave_fold_vm(LD ld, int rb, int re, int last, float *stage)
{
    const int lane = threadIdx.x & 63;
    const int sb = __builtin_amdgcn_readfirstlane(rb);
    const int total = __builtin_amdgcn_readlane(re, last) - sb;
    const bool fits = total <= CAPW;
    if (fits) {
        for (int q = lane; q < total; q += 16 * 64) {
            float x[16];
#pragma unroll
            for (int u = 0; u < 16; u++) x[u] = q + u * 64 < total ? ld(sb + q + u * 64) : 0.0f;
#pragma unroll
            for (int u = 0; u < 16; u++)
                if (q + u * 64 < total) stage[q + u * 64] = x[u];
        }
    }
    __syncthreads();
    float acc = 0.0f;
    if (fits) {
        const float *s = stage + (rb - sb);
        for (int j = 0; j < re - rb; j++) acc = acc + s[j];
    } else {
        for (int q = rb; q < re; q++) acc = acc + ld(q);
    }
    return acc;
}

}  // namespace dnx
