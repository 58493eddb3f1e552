// solo.hpp -- one large 3-SAT instance advanced on the whole device with the
// sliced exchange, its tables and state held in one contiguous device range
// (a range of a batch's arena): the serving loop's solo lane.  The plan, the
// kernels and the step sequence are the single-instance path's (a resident
// state in DNX_FP32 with DNX_EXCHANGE_SLICED), so a request computes the same
// bits here as it does solved alone.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "dnx.h"
#include "host.hpp"
#include "kernels.hpp"

namespace dnx {

// The sliced plan with the single-instance path's parameters for FP32 on a
// device of `cus` compute units; false when the exchange does not apply.
bool solo_plan(const HostProblem &h, int cus, SlicedPlan &pl);
// The same with units of about `clauses_per_unit` clauses (for many
// instances sharing one launch), else the single-instance plan.
bool solo_plan_units(const HostProblem &h, int cus, int clauses_per_unit, SlicedPlan &pl);

// Byte layout of a solo instance inside its range.  [0, upload) is built on
// the host and copied in one transfer; the rest is device-only state.
struct SoloLayout {
    size_t lits, row, unit, sw, units, sbrun, runs, sq, pidx, ctl, xstage, upload;
    size_t vd, xs, xl, e, vnext, cst, total;
};
SoloLayout solo_layout(const HostProblem &h, const SlicedPlan &pl);

// Writes [0, L.upload) for the instance: its tables, the initial control
// block (running, stop at the first verified state, state `cap` the last
// evaluated) and the initial state x0 (reference layout) in FP32.
void solo_stage(const HostProblem &h, const SlicedPlan &pl, const double *x0, int64_t cap, const SoloLayout &L,
                unsigned char *stage);

// The device view of the instance placed at `base`.
DevProb solo_devprob(const HostProblem &h, const SlicedPlan &pl, const SoloLayout &L, unsigned char *base);

// The instance's entry of a many-instance sliced step.
SXDesc solo_desc(const HostProblem &h, const SlicedPlan &pl, const SoloLayout &L, unsigned char *base,
                 const DevParams &prm);

// Bytes of the control block's head that carry the run's status, step count,
// time and per-slot verdicts (what a poll needs).
size_t solo_ctl_head();

}  // namespace dnx
