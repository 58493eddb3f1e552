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

#include <climits>

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
template <typename T, typename GetV, typename Emit>
__device__ __forceinline__ T clause3(const PT<T> &p, const int32_t *cl, T Xs, T Xl, GetV getV,
                                     Emit emit, bool &sat)
{
    int32_t l0 = cl[0], l1 = cl[1], l2 = cl[2];
    T r0 = getV(abs(l0) - 1), r1 = getV(abs(l1) - 1), r2 = getV(abs(l2) - 1);
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

// C of a 3-literal clause only (for the lazy memory update).
template <typename T, typename GetV>
__device__ __forceinline__ T clause3_C(const int32_t *cl, GetV getV)
{
    T V0 = clampV(getV(abs(cl[0]) - 1)), V1 = clampV(getV(abs(cl[1]) - 1)),
      V2 = clampV(getV(abs(cl[2]) - 1));
    T i0 = T(1.0) - (cl[0] > 0 ? T(1) : T(-1)) * V0;
    T i1 = T(1.0) - (cl[1] > 0 ? T(1) : T(-1)) * V1;
    T i2 = T(1.0) - (cl[2] > 0 ? T(1) : T(-1)) * V2;
    return clamp01(tmin(i0, tmin(i1, i2)) / T(2.0));
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

// ---------------------------------------------------------------------------
// Per-variable fold of the contribution segment.
// ---------------------------------------------------------------------------
template <typename T, bool SORT, int BS, int CAP>
__device__ __forceinline__ T fold_segment(T *seg, int deg, T *buf)
{
    T acc = T(0.0);
    if (!SORT) {
        for (int j = 0; j < deg; j++) acc = acc + seg[j];
        return acc;
    }
    // dynex.cc:527-530: ascending std::sort then left fold from 0.  Zero
    // contributions (which the reference never pushes) do not change an
    // ascending fold from +0, so every slot is folded.
    const int t = threadIdx.x;
    if (deg <= CAP) {
        for (int j = 0; j < deg; j++) {
            T x = seg[j];
            int i = j;
            while (i > 0 && buf[(i - 1) * BS + t] > x) {
                buf[i * BS + t] = buf[(i - 1) * BS + t];
                i--;
            }
            buf[i * BS + t] = x;
        }
        for (int j = 0; j < deg; j++) acc = acc + buf[j * BS + t];
    } else {
        // rare high-degree variable: insertion sort in place (the segment is
        // private to this variable and rewritten every step)
        for (int j = 1; j < deg; j++) {
            T x = seg[j];
            int i = j;
            while (i > 0 && seg[i - 1] > x) {
                seg[i] = seg[i - 1];
                i--;
            }
            seg[i] = x;
        }
        for (int j = 0; j < deg; j++) acc = acc + seg[j];
    }
    return acc;
}

constexpr int CBS = 256;  // clause-pass block size
constexpr int VBS = 64;   // variable-pass block size
constexpr int SORT_CAP = 64;

// ---------------------------------------------------------------------------
// Resident-state step: clause pass.
// ---------------------------------------------------------------------------
template <typename T, bool U3>
__global__ __launch_bounds__(CBS) void k_step_clause(DevProb P, PT<T> p, const Pair<T> *__restrict__ vd,
                                                     T *__restrict__ Xs, T *__restrict__ Xl,
                                                     T *__restrict__ e, Ctl *__restrict__ ctl, int slot)
{
    __shared__ int sh[CBS / 64];
    if (ctl->status != DNX_RUNNING) return;
    const int ps = (slot + 2) % 3;
    const int done = ctl->done[ps];
    if (done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) ctl->status = done;
        return;
    }
    const int plev = ctl->level[ps];
    const bool pend = plev >= 0;
    const T h = pend ? level_h<T>(plev) : T(0.0);
    if (pend && blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->t += level_dt(plev);
        ctl->steps += 1;
    }
    const int c = blockIdx.x * CBS + threadIdx.x;
    int unsat = 0, bad = 0;
    if (c < P.m) {
        const int b = U3 ? 3 * c : P.cp[c];
        const int k = U3 ? 3 : P.cp[c + 1] - b;
        const int32_t *cl = P.lits + b;
        T xs = Xs[c], xl = Xl[c];
        auto prevV = [&](int v) { return vd[v].v; };
        auto curV = [&](int v) {
            Pair<T> a = vd[v];
            return pend ? clampV(a.v + h * a.d) : a.v;
        };
        if (pend) {
            T Cp = U3 ? clause3_C<T>(cl, prevV) : clauseK_C<T>(cl, k, prevV);
            memory_update(p, Cp, h, xs, xl);
            Xs[c] = xs;
            Xl[c] = xl;
        }
        const T cxs = clamp01(xs), cxl = clampXl(xl, p.xlmax);
        const int32_t *lp = P.lpos + b;
        auto emit = [&](int j, T val) { e[lp[j]] = val; };
        bool sat;
        T C = U3 ? clause3<T>(p, cl, cxs, cxl, curV, emit, sat)
                 : clauseK<T>(p, cl, k, cxs, cxl, curV, emit, sat);
        unsat = !(C < T(0.5));
        bad = !sat;
    }
    int bu = block_sum_int<CBS>(unsat, sh);
    __syncthreads();
    int bb = block_sum_int<CBS>(bad, sh);
    if (threadIdx.x == 0) {
        if (bu) atomicAdd(&ctl->loc[slot], bu);
        if (bb) atomicAdd(&ctl->vbad[slot], bb);
    }
}

// ---------------------------------------------------------------------------
// Resident-state step: variable pass.
// ---------------------------------------------------------------------------
template <typename T, bool SORT>
__global__ __launch_bounds__(VBS) void k_step_var(DevProb P, Pair<T> *__restrict__ vd,
                                                  T *__restrict__ e, Ctl *__restrict__ ctl, int slot)
{
    __shared__ T buf[SORT ? SORT_CAP * VBS : 1];
    __shared__ int sh[VBS / 64];
    if (ctl->status != DNX_RUNNING) return;
    const int ps = (slot + 2) % 3;
    if (ctl->done[ps]) return;
    const int plev = ctl->level[ps];
    const bool pend = plev >= 0;
    const T h = pend ? level_h<T>(plev) : T(0.0);
    const bool solved = ctl->stop_on_solve && ctl->loc[slot] == 0 && ctl->vbad[slot] == 0;
    const int done = solved ? DNX_SAT : (ctl->steps >= ctl->cap ? DNX_UNKNOWN : 0);
    const int v = blockIdx.x * VBS + threadIdx.x;
    int f = -1;
    if (v < P.n) {
        Pair<T> a = vd[v];
        T V = pend ? clampV(a.v + h * a.d) : a.v;
        T dV = T(0.0);
        if (!done) {
            const int b = P.row_ptr[v];
            dV = fold_segment<T, SORT, VBS, SORT_CAP>(e + b, P.row_ptr[v + 1] - b, buf);
            if (P.unit && P.unit[v] != 0) dV = T(0.0);  // pinned (dynex.cc:306)
            f = fail_level(V, dV);
        }
        Pair<T> o;
        o.v = V;
        o.d = dV;
        vd[v] = o;
    }
    if (done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            ctl->done[slot] = done;
            const int ns = (slot + 1) % 3;
            ctl->loc[ns] = 0;
            ctl->vbad[ns] = 0;
            ctl->level[ns] = -1;
            ctl->done[ns] = 0;
        }
        return;
    }
    int bm = block_max_int<VBS>(f, sh);
    if (threadIdx.x == 0) {
        if (bm > 0 || blockIdx.x == 0) atomicMax(&ctl->level[slot], bm);
        if (blockIdx.x == 0) {
            const int ns = (slot + 1) % 3;
            ctl->loc[ns] = 0;
            ctl->vbad[ns] = 0;
            ctl->level[ns] = -1;
            ctl->done[ns] = 0;
        }
    }
}

// ---------------------------------------------------------------------------
// Function-level kernels (dnx_rhs, dnx_euler) on the reference layout.
// ---------------------------------------------------------------------------
template <typename T, bool U3>
__global__ __launch_bounds__(CBS) void k_rhs_clause(DevProb P, PT<T> p, const T *__restrict__ x,
                                                    T *__restrict__ dxdt, T *__restrict__ e,
                                                    int *__restrict__ loc, double *__restrict__ epart)
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
        const int32_t *lp = P.lpos + b;
        auto getV = [&](int v) { return x[v]; };
        auto emit = [&](int j, T val) { e[lp[j]] = val; };
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
        if (bu) atomicAdd(loc, bu);
        double s = 0.0;
        for (int i = 0; i < CBS / 64; i++) s += shd[i];
        epart[blockIdx.x] = s;
    }
}

template <typename T, bool SORT>
__global__ __launch_bounds__(VBS) void k_rhs_var(DevProb P, T *__restrict__ e, T *__restrict__ dxdt)
{
    __shared__ T buf[SORT ? SORT_CAP * VBS : 1];
    const int v = blockIdx.x * VBS + threadIdx.x;
    if (v >= P.n) return;
    const int b = P.row_ptr[v];
    dxdt[v] = fold_segment<T, SORT, VBS, SORT_CAP>(e + b, P.row_ptr[v + 1] - b, buf);
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
        if (!ctl->done[s]) plev = ctl->level[s];
    }
    const bool pend = plev >= 0;
    const T h = pend ? level_h<T>(plev) : T(0.0);
    if (i < n) {
        Pair<T> a = vd[i];
        xo[i] = pend ? clampV(a.v + h * a.d) : a.v;
    }
    if (i < m) {
        T xs = Xs[i], xl = Xl[i];
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
        Xs[i] = x[n + i];
        Xl[i] = x[n + m + i];
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned nblk(long long count, int bs) { return (unsigned)((count + bs - 1) / bs); }

int rhs_clause_blocks(int m) { return (int)nblk(m > 0 ? m : 1, CBS); }

template <typename T>
static void rhs_impl(const DevProb &P, const DevParams &prm, const void *x, void *dxdt, void *e, int *loc,
                     double *epart, hipStream_t st)
{
    PT<T> p = to_pt<T>(prm);
    constexpr bool SORT = sizeof(T) == 8;
    unsigned gc = nblk(P.m > 0 ? P.m : 1, CBS);
    if (P.cp == nullptr)
        k_rhs_clause<T, true><<<gc, CBS, 0, st>>>(P, p, (const T *)x, (T *)dxdt, (T *)e, loc, epart);
    else
        k_rhs_clause<T, false><<<gc, CBS, 0, st>>>(P, p, (const T *)x, (T *)dxdt, (T *)e, loc, epart);
    if (P.n > 0) k_rhs_var<T, SORT><<<nblk(P.n, VBS), VBS, 0, st>>>(P, (T *)e, (T *)dxdt);
}

void launch_rhs(int prec, const DevProb &P, const DevParams &prm, const void *x, void *dxdt, void *e, int *loc,
                double *energy_part, int, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (prec == DNX_FP64)
        rhs_impl<double>(P, prm, x, dxdt, e, loc, energy_part, st);
    else
        rhs_impl<float>(P, prm, x, dxdt, e, loc, energy_part, st);
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
    unsigned gc = nblk(P.m > 0 ? P.m : 1, CBS);
    if (P.cp == nullptr)
        k_step_clause<T, true><<<gc, CBS, 0, st>>>(P, p, (const Pair<T> *)vd, (T *)xs, (T *)xl, (T *)e, ctl,
                                                   slot);
    else
        k_step_clause<T, false><<<gc, CBS, 0, st>>>(P, p, (const Pair<T> *)vd, (T *)xs, (T *)xl, (T *)e, ctl,
                                                    slot);
    k_step_var<T, SORT><<<nblk(P.n > 0 ? P.n : 1, VBS), VBS, 0, st>>>(P, (Pair<T> *)vd, (T *)e, ctl, slot);
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
