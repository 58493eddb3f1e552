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
static inline PT<T> to_pt(const DevParams &p)
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

template <typename T>
__device__ __forceinline__ T clampV(T v)
{
    if (v < T(-1.0)) v = T(-1.0);
    if (v > T(1.0)) v = T(1.0);
    return v;
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
    if (v < T(1.0)) v = T(1.0);
    if (v > xlmax) v = xlmax;
    return v;
}
template <typename T>
__device__ __forceinline__ T tmin(T a, T b)
{
    return b < a ? b : a;
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

__device__ __forceinline__ bool lit_true(int32_t lit, double x)
{
    return (x > 0.0 && lit > 0) || (x < 0.0 && lit < 0);
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
    sat = lit_true(l0, (double)r0) || lit_true(l1, (double)r1) || lit_true(l2, (double)r2);
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

template <typename T, typename GetV, typename Emit>
__device__ __forceinline__ T clause3(const PT<T> &p, const int32_t *cl, T Xs, T Xl, GetV getV, Emit emit, bool &sat)
{
    return clause3v<T>(p, cl[0], cl[1], cl[2], Xs, Xl, getV, emit, sat);
}

// C of a 3-literal clause only (for the lazy memory update).
template <typename T, typename GetV>
__device__ __forceinline__ T clause3v_C(int32_t l0, int32_t l1, int32_t l2, GetV getV)
{
    return clause3_C_vals<T>(l0, l1, l2, getV(abs(l0) - 1), getV(abs(l1) - 1), getV(abs(l2) - 1));
}
template <typename T, typename GetV>
__device__ __forceinline__ T clause3_C(const int32_t *cl, GetV getV)
{
    return clause3v_C<T>(cl[0], cl[1], cl[2], getV);
}

// Any clause length (dynex.cc:420-494 for k=2 and k>=4, 517 for k=1; the
// empty clause follows the k-SAT branch with min = INT_MAX).  Two passes over
// the literals keep everything in registers: the first finds the smallest
// i_l (and its first index) and the smallest of the rest; the second emits.
template <typename T, typename GetV, typename Emit>
__device__ __forceinline__ T clauseK(const PT<T> &p, const int32_t *cl, int k, T Xs, T Xl,
                                     GetV getV, Emit emit, bool &sat)
{
    sat = false;
    if (k == 1) {
        sat = lit_true(cl[0], (double)getV(abs(cl[0]) - 1));
        emit(0, T(0.0));
        return T(0.0);
    }
    T min1 = T(INT_MAX), min2 = T(INT_MAX);
    int idx1 = -1;
    for (int j = 0; j < k; j++) {
        T r = getV(abs(cl[j]) - 1);
        sat = sat || lit_true(cl[j], (double)r);
        T I = T(1.0) - (cl[j] > 0 ? T(1) : T(-1)) * clampV(r);
        if (I < min1) {
            if (idx1 >= 0) min2 = tmin(min2, min1);
            min1 = I;
            idx1 = j;
        } else {
            min2 = tmin(min2, I);
        }
    }
    T C = min1 / T(2.0);
    if (k <= 3) C = clamp01(C);
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

template <typename T, typename GetV>
__device__ __forceinline__ T clauseK_C(const int32_t *cl, int k, GetV getV)
{
    if (k == 1) return T(0.0);
    T min1 = T(INT_MAX);
    for (int j = 0; j < k; j++) {
        T I = T(1.0) - (cl[j] > 0 ? T(1) : T(-1)) * clampV(getV(abs(cl[j]) - 1));
        min1 = tmin(min1, I);
    }
    T C = min1 / T(2.0);
    if (k <= 3) C = clamp01(C);
    return C;
}

// Clause-order fold of the contributions gathered through the transpose:
// slots [rb, re) of a variable name literal positions eidx[q] whose values
// sit in e[].  Eight independent gathers are kept in flight; the additions
// stay in slot order.
template <typename T>
__device__ __forceinline__ T fold_gather(const T *__restrict__ e, const int32_t *__restrict__ eidx, int rb, int re)
{
    T acc = T(0.0);
    int q = rb;
    for (; q + 8 <= re; q += 8) {
        int i[8];
        T x[8];
#pragma unroll
        for (int u = 0; u < 8; u++) i[u] = eidx[q + u];
#pragma unroll
        for (int u = 0; u < 8; u++) x[u] = e[i[u]];
#pragma unroll
        for (int u = 0; u < 8; u++) acc = acc + x[u];
    }
    for (; q < re; q++) acc = acc + e[eidx[q]];
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

}  // namespace dnx
