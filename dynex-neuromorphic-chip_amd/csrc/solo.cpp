// solo.cpp -- layout, staging and device view of a solo instance (solo.hpp).
#include "solo.hpp"

#include <algorithm>
#include <climits>
#include <cstddef>
#include <cstring>

namespace dnx {

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

bool solo_plan(const HostProblem &h, int cus, SlicedPlan &pl)
{
    if (!h.uniform3 || h.m <= 0) return false;
    if (!sliced_tables(h, sx_slice(DNX_FP32), sx_sub(DNX_FP32), SX_F, cus, sx_max_rounds(DNX_FP32),
                       sx_region_cap(DNX_FP32), pl))
        return false;
    for (size_t k = 0; k + 1 < pl.sb_run.size(); k++)
        if (pl.sb_run[k + 1] - pl.sb_run[k] > SX_MAXRUNS) return false;
    return true;
}

bool solo_plan_units(const HostProblem &h, int cus, int clauses_per_unit, SlicedPlan &pl)
{
    if (!h.uniform3 || h.m <= 0) return false;
    const int want = std::max(1, std::min(cus, (h.m + clauses_per_unit - 1) / clauses_per_unit));
    if (want < cus && solo_plan(h, want, pl)) return true;
    return solo_plan(h, cus, pl);
}

SXDesc solo_desc(const HostProblem &h, const SlicedPlan &pl, const SoloLayout &L, unsigned char *base,
                 const DevParams &prm)
{
    SXDesc d;
    memset(&d, 0, sizeof(d));
    d.n = h.n;
    d.m = h.m;
    d.vstride = sx_vstride(h.n > 0 ? h.n : 1);
    d.stride = pl.stride;
    d.sw = (const uint32_t *)(base + L.sw);
    d.units = (const int32_t *)(base + L.units);
    d.sbrun = (const int32_t *)(base + L.sbrun);
    d.runs = (const int32_t *)(base + L.runs);
    d.row_ptr = (const int32_t *)(base + L.row);
    d.sq = (const uint16_t *)(base + L.sq);
    d.unit = h.has_unit ? (const int8_t *)(base + L.unit) : nullptr;
    d.vd = base + L.vd;
    d.vnext = base + L.vnext;
    d.xs = base + L.xs;
    d.xl = base + L.xl;
    d.cst = base + L.cst;
    d.e = base + L.e;
    d.ctl = (Ctl *)(base + L.ctl);
    d.prm = prm;
    return d;
}

SoloLayout solo_layout(const HostProblem &h, const SlicedPlan &pl)
{
    SoloLayout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o += al256(bytes > 0 ? bytes : 1);
        return at;
    };
    const size_t n = (size_t)h.n, m = (size_t)h.m, nnz = (size_t)h.nnz();
    L.lits = take(4 * nnz);
    L.row = take(4 * (n + 1));
    L.unit = take(h.has_unit ? n : 0);
    L.sw = take(4 * pl.gw.size());
    L.units = take(4 * pl.units.size());
    L.sbrun = take(4 * pl.sb_run.size());
    L.runs = take(4 * pl.runs.size());
    L.sq = take(2 * pl.q.size());
    L.pidx = take(4 * pl.pidx.size());
    L.ctl = take(sizeof(Ctl));
    // the initial state in FP32; later the materialised state read back
    L.xstage = take(4 * (n + 2 * m));
    L.upload = o;
    // one entry past the end of Xs, Xl, e and the stored C: the clause pass
    // sends the stores of lanes without a clause there
    L.vd = take(2 * 4 * (n > 0 ? n : 1));
    L.xs = take(4 * (m + 1));
    L.xl = take(4 * (m + 1));
    L.e = take(4 * ((size_t)std::max<long long>((long long)nnz, pl.e_len) + 1));
    L.vnext = take(4 * 4 * (size_t)sx_vstride(h.n > 0 ? h.n : 1));
    L.cst = take(4 * (m + 1));
    L.total = o;
    return L;
}

void solo_stage(const HostProblem &h, const SlicedPlan &pl, const double *x0, int64_t cap, const SoloLayout &L,
                unsigned char *stage)
{
    memset(stage, 0, L.upload);
    auto put = [&](size_t at, const void *src, size_t bytes) {
        if (bytes) memcpy(stage + at, src, bytes);
    };
    const size_t n = (size_t)h.n, m = (size_t)h.m;
    put(L.lits, h.lits.data(), 4 * h.lits.size());
    put(L.row, h.row_ptr.data(), 4 * h.row_ptr.size());
    if (h.has_unit) put(L.unit, h.unit.data(), n);
    put(L.sw, pl.gw.data(), 4 * pl.gw.size());
    put(L.units, pl.units.data(), 4 * pl.units.size());
    put(L.sbrun, pl.sb_run.data(), 4 * pl.sb_run.size());
    put(L.runs, pl.runs.data(), 4 * pl.runs.size());
    put(L.sq, pl.q.data(), 2 * pl.q.size());
    put(L.pidx, pl.pidx.data(), 4 * pl.pidx.size());
    Ctl *c = (Ctl *)(stage + L.ctl);
    c->status = DNX_RUNNING;
    c->stop_on_solve = 1;
    c->cap = cap;
    c->loc_min = c->loc_last = h.m;
    float *x = (float *)(stage + L.xstage);
    for (size_t i = 0; i < n + 2 * m; i++) x[i] = (float)x0[i];
}

DevProb solo_devprob(const HostProblem &h, const SlicedPlan &pl, const SoloLayout &L, unsigned char *base)
{
    DevProb d;
    memset(&d, 0, sizeof(d));
    d.n = h.n;
    d.m = h.m;
    d.max_deg = h.max_deg;
    d.cp = nullptr;
    d.lits = (const int32_t *)(base + L.lits);
    d.eidx = nullptr;
    d.row_ptr = (const int32_t *)(base + L.row);
    d.unit = h.has_unit ? (const int8_t *)(base + L.unit) : nullptr;
    d.sw = (const uint32_t *)(base + L.sw);
    d.units = (const int32_t *)(base + L.units);
    d.sbrun = (const int32_t *)(base + L.sbrun);
    d.runs = (const int32_t *)(base + L.runs);
    d.sq = (const uint16_t *)(base + L.sq);
    d.pidx = (const int32_t *)(base + L.pidx);
    d.nunits = (int)pl.units.size() / 8;
    d.sx_region = pl.stride;
    return d;
}

size_t solo_ctl_head() { return offsetof(Ctl, track_loc); }

}  // namespace dnx
