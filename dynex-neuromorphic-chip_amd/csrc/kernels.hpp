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
    const int32_t *eidx;     // literal position of each variable-major slot
    const int32_t *row_ptr;  // n+1
    const int8_t *unit;      // n pins, or nullptr when there are none
    // grouped exchange of the resident-state step (3-SAT only, else null):
    // contributions are stored grouped by the XCD whose variable-pass
    // workgroups fold them (see grouped_tables), so each XCD's fold gathers
    // from a region that stays in its L2.  gw: one word per literal (sign,
    // variable, group, offset in its chunk's run); gbase: per 64-clause chunk
    // the start of its run in each group; geidx: the e position of each
    // variable-major slot (replaces eidx in the fold).
    const uint32_t *gw;
    const int32_t *gbase;
    const int32_t *geidx;
    // sliced exchange (3-SAT only, else null; see sliced_tables): sw one
    // word per literal in processing order (sign, LDS index of its variable
    // among the unit's slices, staging slot), units 8 ints per unit, sbrun
    // the first run of each sub-batch, runs (staging start << 16 | length,
    // e position), sq the variable-major slot of each e position, pidx the
    // processing position of each clause (Xs, Xl and the stored C are kept
    // in processing order)
    const uint32_t *sw;
    const int32_t *units;
    const int32_t *sbrun;
    const int32_t *runs;
    const uint16_t *sq;
    const int32_t *pidx;
    int nunits;
    int sx_region;  // slots per fold block's region of e (every region padded to it)
};

// Sliced exchange shapes: variables per slice (three slices of V in LDS),
// clauses per sub-batch, variables per fold block.
constexpr int SX_F = 256;
constexpr int SX_MAXRUNS = 255;  // runs of one sub-batch (three slices of fold blocks)
inline int sx_slice(int prec) { return prec == 0 ? 5000 : 10000; }
inline int sx_sub(int prec) { return prec == 0 ? 512 : 1024; }
inline int sx_region_cap(int prec) { return prec == 0 ? 20000 : 40000; }
inline int sx_max_rounds(int) { return 2; }
// A unit of `size` clauses is taken in sx_unit_batches sub-batches of equal
// size (at most `sub_max`, the workgroup's threads), sx_unit_sub clauses each
// except the last: one more sub-batch of a few clauses would cost a whole
// iteration of the clause pass.  (constexpr: usable in the kernels too.)
constexpr int sx_unit_batches(int size, int sub_max) { return (size + sub_max - 1) / sub_max; }
constexpr int sx_unit_sub(int size, int sub_max)
{
    return sx_unit_batches(size, sub_max) > 0 ? (size + sx_unit_batches(size, sub_max) - 1) / sx_unit_batches(size, sub_max)
                                              : sub_max;
}
// elements per ladder level in the fold's candidate array: whole 16-byte words
inline int sx_vstride(int n) { return (n + 3) / 4 * 4; }

// Packed literal word of the grouped exchange: bit 31 the sign (set for a
// negative literal), bits 30..11 the 0-based variable, bits 10..8 the group
// (XCD) of the variable's fold, bits 7..0 the literal's offset in its chunk's
// run of that group.
constexpr int GX_CHUNK = 64;       // clauses per chunk (one wave's clauses)
constexpr int GX_GROUPS = 8;       // XCDs
constexpr int GX_MAX_VARS = 1 << 20;
inline int gx_var_block(int prec) { return prec == 0 ? 128 : 256; }  // variable-pass workgroup size

struct DevParams {
    double alpha, beta, gamma, delta, epsilon, zeta, xl_max;
};

struct Ctl;

// One instance of a sliced step that advances many instances in one launch
// pair (the sliced tables and state of a resident instance; see DevProb).
struct SXDesc {
    int n, m, vstride, stride;  // stride: the plan's slots per fold-block region of e
    const uint32_t *sw;
    const int32_t *units, *sbrun, *runs, *row_ptr;
    const uint16_t *sq;
    const int8_t *unit;
    void *vd, *vnext, *xs, *xl, *cst, *e;
    Ctl *ctl;
    DevParams prm;
};

// Per-instance control block.  Three rotating slots carry one step's
// verdicts: step s writes slot s%3 (the clause pass raises the `unsolved`
// flag group, the variable pass the level groups and `stepped`); the clause
// pass of step s+1 reads slot s%3 to apply step s's update lazily; the
// variable pass of step s clears slot (s+1)%3 for step s+1.
//
// A flag group is 8 cache lines of 32 ints.  Workgroup b raises entry
// (b & 7) * 32 + ((b >> 3) & 7) with a plain store: workgroups are dealt to
// the 8 XCDs round-robin, so each line is written only from one XCD's L2 and
// no store contends across the chip.  A reader ORs the 64 used entries with
// one load per lane and a wave ballot.
constexpr int FLAG_GROUP = 256;   // ints per group (8 lines x 32)
constexpr int FLAG_USED = 64;     // entries a group actually uses
constexpr int FG_UNSOLVED = 0;    // some clause has C >= 1/2 or no true literal
constexpr int FG_LEVEL = 1;       // groups 1..4: some variable needs level >= L
constexpr int FG_COUNT = 5;

struct Ctl {
    int status;         // 0 running, else DNX_SAT / DNX_UNKNOWN
    int stop_on_solve;  // 1: stop at the first verified state
    long long steps;    // index of the last state the clause pass produced
    long long cap;      // step cap: state `cap` is the last one evaluated
    double t;           // simulated time of that state
    int stepped[3];     // the variable pass produced a level (a step is pending)
    int done[3];        // DNX_SAT / DNX_UNKNOWN when the state ended the run
    unsigned long long t_done;  // device clock when a run ended (batched wide mode)
    // loc tracking (single-instance path, when track_loc != 0): the clause
    // pass adds each workgroup's count of clauses with C >= 1/2 to one of 64
    // counters of its slot (spread over 8 lines like a flag group), and the
    // variable pass sums them into loc_last and keeps the lowest in loc_min
    // (the reference's `global`, dynex.cc:1188, 584-600)
    int track_loc;
    int loc_min;
    int loc_last;
    int pad_loc;
    int locc[3][FLAG_GROUP];
    int flags[3][FG_COUNT][FLAG_GROUP];
};

inline int flag_entry(int i) { return (i & 7) * 32 + (i >> 3); }  // i in [0, 64)
inline bool ctl_any(const Ctl &c, int s, int g)
{
    for (int i = 0; i < FLAG_USED; i++)
        if (c.flags[s][g][flag_entry(i)]) return true;
    return false;
}
// level of a slot: the level groups are prefix-closed
inline int ctl_level(const Ctl &c, int s)
{
    int L = 0;
    for (int g = 0; g < 4; g++) L += ctl_any(c, s, FG_LEVEL + g) ? 1 : 0;
    return L;
}

// ---- launchers (kernels.hip) ---------------------------------------------
// prec: 0 = FP64 (sorted fold), 1 = FP32 (clause-order fold)
void launch_rhs(int prec, const DevProb &P, const DevParams &prm, const void *x,
                void *dxdt, void *e, int *loc_part, double *energy_part, void *stream);
int rhs_clause_blocks(int m);

// streaming copy of `bytes` (multiple of 16) from src to dst
void launch_copy(const void *src, void *dst, long long bytes, void *stream);

void launch_euler(int prec, const DevProb &P, const DevParams &prm, const void *x,
                  const void *dxdt, double h, void *x_out, void *stream);

// one adaptive step of the resident state (clause pass + variable pass)
void launch_step(int prec, const DevProb &P, const DevParams &prm, void *vd, void *xs,
                 void *xl, void *e, Ctl *ctl, int slot, void *stream);
// the same with the sliced exchange: vnext holds 4 n candidate V values (one
// per ladder level) written by the fold, cst each clause's last C
// FP32 sliced step of many instances: nclause blocks (map_c: instance, unit)
// then nvar blocks (map_v: instance, fold block), at step slot `slot`;
// lds_region the largest region stride among them
void launch_step_sliced_multi(const SXDesc *d, const void *map_c, int nclause, const void *map_v, int nvar,
                              int lds_region, int slot, void *stream);
// copies the first `words` ints of each instance's control block to out
// (count x words ints)
void launch_ctl_gather(const SXDesc *d, int count, int words, int *out, void *stream);
void launch_step_sliced(int prec, const DevProb &P, const DevParams &prm, void *vd, void *xs, void *xl,
                        void *e, void *vnext, void *cst, Ctl *ctl, int slot, void *stream);

// state -> x_out (reference layout, precision T) with any pending update
// applied; pins applied.  `tmpv` holds n values of T.
void launch_materialize(int prec, const DevProb &P, const DevParams &prm,
                        const void *vd, const void *xs, const void *xl, const Ctl *ctl,
                        void *x_out, void *stream);

// x (reference layout, T) -> resident layout; pins applied to V.
void launch_load_state(int prec, const DevProb &P, const void *x, void *vd, void *xs,
                       void *xl, void *stream);

}  // namespace dnx
