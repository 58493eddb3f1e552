// kernels.hpp -- device data structures and launch wrappers shared by the
// C ABI (api.cpp) and the kernels (kernels.hip).  No HIP headers here: the
// launchers take a hipStream_t through an opaque pointer.
#pragma once

#include <cstdint>

namespace dnx {

// Device copy of a problem.  cp == nullptr when every clause has 3 literals
// (clause c then owns literal positions 3c..3c+2).
struct DevProb {
    int n, m;
    int max_deg;
    const int32_t *cp;
    const int32_t *lits;     // signed 1-based literal per position
    const int32_t *lpos;     // variable-major slot of each literal position
    const int32_t *row_ptr;  // n+1
    const int8_t *unit;      // n pins, or nullptr when there are none
};

struct DevParams {
    double alpha, beta, gamma, delta, epsilon, zeta, xl_max;
};

// Per-instance control block.  Three rotating slots carry one step's
// reductions: step s writes slot s%3 (loc and vbad from the clause pass,
// level from the variable pass); the clause pass of step s+1 reads slot s%3
// to apply step s's update lazily; the variable pass of step s clears slot
// (s+1)%3 for step s+1.
struct Ctl {
    int status;         // 0 running, else DNX_SAT / DNX_UNKNOWN
    int stop_on_solve;  // 1: stop at the first verified state
    long long steps;    // index of the last state the clause pass produced
    long long cap;      // step cap: state `cap` is the last one evaluated
    double t;           // simulated time of that state
    int loc[3];         // clauses with C >= 1/2
    int vbad[3];        // clauses with no strictly true literal
    int level[3];       // max fail level, -1 when no step was taken
    int done[3];        // DNX_SAT / DNX_UNKNOWN when the state ended the run
    int pad[4];
};

// ---- launchers (kernels.hip) ---------------------------------------------
// prec: 0 = FP64 (sorted fold), 1 = FP32 (clause-order fold)
void launch_rhs(int prec, const DevProb &P, const DevParams &prm, const void *x,
                void *dxdt, void *e, int *loc, double *energy_part, int nblk_clause,
                void *stream);
int rhs_clause_blocks(int m);

void launch_euler(int prec, const DevProb &P, const DevParams &prm, const void *x,
                  const void *dxdt, double h, void *x_out, void *stream);

// one adaptive step of the resident state (clause pass + variable pass)
void launch_step(int prec, const DevProb &P, const DevParams &prm, void *vd, void *xs,
                 void *xl, void *e, Ctl *ctl, int slot, void *stream);

// state -> x_out (reference layout, precision T) with any pending update
// applied; pins applied.  `tmpv` holds n values of T.
void launch_materialize(int prec, const DevProb &P, const DevParams &prm,
                        const void *vd, const void *xs, const void *xl, const Ctl *ctl,
                        void *x_out, void *stream);

// x (reference layout, T) -> resident layout; pins applied to V.
void launch_load_state(int prec, const DevProb &P, const void *x, void *vd, void *xs,
                       void *xl, void *stream);

}  // namespace dnx
