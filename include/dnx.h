/*
 * dnx.h -- C ABI of libdnx, the MI355X-native memcomputing SAT integrator.
 *
 * The reference has no plugin or FFI layer: its in-process boundary is two
 * C++ functions that pass std::vector<double> by value and publish results
 * through globals, and its process boundary is the `dynex -i file.cnf` CLI.
 * Each entry point below names the reference code it replaces.  Common rules:
 *   - every function returns an int status: DNX_OK (0) or a negative DNX_E_*
 *     code; dnx_last_error() returns a message for the calling thread's last
 *     failure.  Nothing calls exit() or kills a thread (the reference calls
 *     pthread_exit from inside its RHS, dynex.cc:340-345).
 *   - buffers are caller-owned host memory; sizes are explicit.
 *   - state vectors at this boundary use the reference layout
 *     x = [V(0..n-1) | Xs(0..m-1) | Xl(0..m-1)] in double (dynex.cc:1205-1210);
 *     the device keeps its own structure-of-arrays layout.
 *   - a dnx_ctx owns one HIP stream on one device; a context and the objects
 *     made from it must not be used from two threads at once.
 */
#ifndef DNX_H
#define DNX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define DNX_OK 0
#define DNX_E_ARG -1     /* bad argument */
#define DNX_E_IO -2      /* file cannot be opened or written */
#define DNX_E_PARSE -3   /* malformed DIMACS */
#define DNX_E_TOOLONG -4 /* clause longer than DNX_MAX_LITS (dynex.cc:2388) */
#define DNX_E_HIP -5     /* HIP runtime failure */
#define DNX_E_NOMEM -6   /* device or host allocation failed */
#define DNX_E_BUDGET -7  /* request exceeds the configured memory budget */
#define DNX_E_STATE -8   /* call not valid in the object's current state */

#define DNX_MAX_LITS 25 /* MAX_LITS_SYSTEM, dynex.cc:69 */

/* ---- precision ------------------------------------------------------- */
#define DNX_FP64 0 /* reference arithmetic: double, sorted per-variable fold */
#define DNX_FP32 1 /* float state and RHS, clause-order fold, double time   */

/* ---- solve status ---------------------------------------------------- */
#define DNX_RUNNING 0
#define DNX_SAT 1     /* assignment passed the exact integer clause check */
#define DNX_UNKNOWN 2 /* step cap or wall-clock limit reached (dynex.cc:1519-1521) */
#define DNX_FAILED 3

/* ---- adaptive step levels -------------------------------------------- */
/* level L in 0..3: the step was accepted at h = 2^-(3+L);
 * level 4: the 1/64 trial also failed; its state is kept and t advances by
 * 1/128 (dynex.cc:1394-1417). */

typedef struct dnx_ctx dnx_ctx;
typedef struct dnx_problem dnx_problem;
typedef struct dnx_state dnx_state;

/* Equation constants.  Defaults (dnx_params_default) are the reference's
 * (dynex.cc:132-140): alpha 2.5, beta 20, gamma 0.25, delta 0.05,
 * epsilon 0.1, zeta 0.1.  xl_max is the already-scaled bound on Xl; the
 * reference computes it as 10000*m in wrapping int arithmetic and uses
 * INT_MAX when that is not positive (dynex.cc:2358-2359). */
typedef struct {
    double alpha, beta, gamma, delta, epsilon, zeta;
    int64_t xl_max;
} dnx_params;

/* Thread-local message for the last failing call. */
const char *dnx_last_error(void);

/* Fill *p with the reference defaults for a problem with m clauses;
 * xl_max_flag is the -m value (10000 by default, dynex.cc:141). */
int dnx_params_default(int m, int xl_max_flag, dnx_params *p);

/* ---- context --------------------------------------------------------- */
/* Opens device `device` and creates the context's stream. */
int dnx_open(int device, dnx_ctx **out);
int dnx_close(dnx_ctx *ctx);

/* ---- problems -------------------------------------------------------- */
/* A NULL ctx makes a host-only problem: it can be inspected, checked and
 * used for initial states, and the device entry points reject it with
 * DNX_E_STATE.
 *
 * DIMACS reader with the reference's token rules (dynex.cc:2306-2392):
 * tokens up to "p", then " cnf n m", clauses terminated by "0", "c" skips a
 * line.  A missing file is DNX_E_IO (the reference crashes, dynex.cc:2308). */
int dnx_problem_from_dimacs(dnx_ctx *ctx, const char *path, dnx_problem **out);

/* From a clause table in CSR form: clause c has literals
 * lits[clause_ptr[c] .. clause_ptr[c+1]), signed, 1-based
 * (the reference's cls/clauseSizes, dynex.cc:178-183). */
int dnx_problem_from_lits(dnx_ctx *ctx, int n, int m, const int32_t *clause_ptr,
                          const int32_t *lits, dnx_problem **out);
int dnx_problem_info(const dnx_problem *p, int *n, int *m, int64_t *nnz);

/* How the resident-state step of a 3-SAT problem brings V to its clauses
 * and each literal's contribution to its variable's fold (no counterpart in
 * the reference, whose clause loop reads x directly and appends to
 * per-variable vectors, dynex.cc:362-410, 527-530); the bits are the same
 * either way.
 *   DNX_EXCHANGE_SLICED (default): clauses regrouped into units whose
 *     variables lie in at most three slices of V, which the unit's
 *     workgroup holds in LDS; contributions leave as one contiguous run per
 *     fold block (3-SAT whose slice sets need at most two rounds of units;
 *     otherwise the grouped exchange);
 *   DNX_EXCHANGE_GROUPED: V gathered from device memory, contributions
 *     stored grouped by the XCD that folds the variable, so each XCD's fold
 *     reads a region that stays in its L2 (3-SAT with at most 2^20
 *     variables; otherwise the literal order);
 *   DNX_EXCHANGE_LITERAL: contributions in literal order, gathered through
 *     the transpose.
 * Applies to states created afterwards. */
#define DNX_EXCHANGE_LITERAL 0
#define DNX_EXCHANGE_GROUPED 1
#define DNX_EXCHANGE_SLICED 2
int dnx_problem_set_exchange(dnx_problem *p, int exchange);
/* Copies the clause table back out (clause_ptr[m+1], lits[nnz]). */
int dnx_problem_clauses(const dnx_problem *p, int32_t *clause_ptr, int32_t *lits);
int dnx_problem_free(dnx_problem *p);

/* ---- function-level entry points (parity surface) -------------------- */
/* dmm_generate_dxdt (dynex.cc:332-530): dxdt = RHS(x); *loc = number of
 * clauses with C >= 1/2; *energy = sum of C.  In DNX_FP64 the result is
 * bit-identical to the reference's strict-IEEE evaluation. */
int dnx_rhs(dnx_problem *p, int prec, const dnx_params *prm, const double *x,
            double *dxdt, int *loc, double *energy);

/* dmm_dostep (dynex.cc:301-327): x_out = clamp(x + h*dxdt) with unit pins. */
int dnx_euler(dnx_problem *p, int prec, const dnx_params *prm, const double *x,
              const double *dxdt, double h, double *x_out);

/* dnx_rhs on device buffers with no device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops that must not call hipMalloc).  x (n+2m values,
 * reference layout) and dxdt are device pointers in the compute type (double
 * for DNX_FP64, float for DNX_FP32); loc (int) and energy (double) are device
 * pointers or NULL and receive what dnx_rhs returns through its pointers,
 * bit-identical.  The work is queued on `stream` (a hipStream_t; NULL: the
 * context's stream) and the call returns without waiting.  The per-literal
 * contributions and per-workgroup partial sums live in one device array of
 * DNX_NOALLOC_SCRATCH_BYTES per GPU (looked up when a context is opened on
 * it, held until the process ends): a problem needing more is refused with
 * DNX_E_BUDGET (about nnz <= 8.3M in DNX_FP64, 16.7M in DNX_FP32).  Every call
 * on one GPU shares that array and it keeps the last call's contents: calls
 * that overlap on one GPU (two streams, two host threads) race. */
#define DNX_NOALLOC_SCRATCH_BYTES 67108864 /* 64 MiB */
int dnx_rhs_device_noalloc(dnx_problem *p, int prec, const dnx_params *prm, const void *x,
                           void *dxdt, int *loc, double *energy, void *stream);

/* ---- resident state (the mode-6 loop of dynex.cc:1381-1427) ---------- */
/* Uploads x0 (reference layout; unit-clause pins are applied to V as the
 * reference does at dynex.cc:2694-2699) and keeps the state on the device. */
int dnx_state_create(dnx_problem *p, int prec, const dnx_params *prm,
                     const double *x0, dnx_state **out);
int dnx_state_free(dnx_state *s);

/* loc tracking, for the reference's `global` (the lowest loc reached,
 * dynex.cc:190, 1188, 584-600, printed at 1500-1504 and 2864): with enable
 * != 0 every state the run evaluates has its loc (clauses with C >= 1/2,
 * dynex.cc:499) counted on the device and the lowest kept; enabling resets
 * the lowest to m.  Off by default (it adds one atomic per clause workgroup
 * per step). */
int dnx_state_track_loc(dnx_state *s, int enable);

/* The lowest loc over the states evaluated since tracking was enabled (m if
 * none) and the loc of the last evaluated state. */
int dnx_state_best_loc(dnx_state *s, int *best_loc, int *last_loc);

/* Advances nsteps adaptive steps without stopping at a solution (fixed-step
 * trajectory mode).  No host synchronisation inside. */
int dnx_state_step(dnx_state *s, int nsteps);

/* Advances nsteps steps in fixed-step mode like dnx_state_step and reports
 * the device time they took (HIP events on the context stream), in ms. */
int dnx_state_time(dnx_state *s, int nsteps, double *ms);

/* Reads the current state: x (reference layout, may be NULL), simulated
 * time, accepted-step count, the level of the last step, and loc of the
 * last evaluated state. */
int dnx_state_read(dnx_state *s, double *x, double *t, int64_t *steps,
                   int *last_level, int *loc);

/* Runs the mode-6 loop until the evaluated state passes the exact clause
 * check (DNX_SAT), step_cap states have been stepped past, or walltime_s
 * seconds pass (DNX_UNKNOWN; walltime_s <= 0 means no limit).  assign_out[n]
 * receives +1/-1 from the sign of the verified V (x >= 0 positive, as the
 * reference's solution file, dynex.cc:693-699).  *steps_out counts the
 * accepted steps before the verified state (the counter the reference
 * prints at that RHS, dynex.cc:549). */
int dnx_state_solve(dnx_state *s, int64_t step_cap, double walltime_s,
                    int8_t *assign_out, int64_t *steps_out, double *t_out,
                    int *status);

/* One-call form of create + solve + free (the reference's dmm(),
 * dynex.cc:1182). */
int dnx_solve(dnx_problem *p, int prec, const dnx_params *prm, const double *x0,
              int64_t step_cap, double walltime_s, int8_t *assign_out,
              int64_t *steps_out, double *t_out, int *status);

/* The reference's thread portfolio (-w, dynex.cc:2731-2801): `threads`
 * copies with the reference's per-thread initial states
 * (dnx_init_reference_thread) and alpha/i for thread i > 0 (dynex.cc:2657)
 * advance in rounds of equal step counts.  The first round in which a copy
 * verifies ends the run; the winner is the verified copy with the fewest
 * steps (lowest index on a tie) -- deterministic, where the reference takes
 * whichever thread finishes first in wall-clock time.  *winner = its index. */
int dnx_portfolio_solve(dnx_problem *p, int prec, const dnx_params *prm, int seed, int threads,
                        int64_t step_cap, double walltime_s, int8_t *assign_out,
                        int64_t *steps_out, double *t_out, int *status, int *winner);

/* ---- batched engine -------------------------------------------------- */
/* Many independent instances advanced together on one device (no
 * counterpart in the reference, which runs one instance per thread; its
 * thread portfolio, dynex.cc:2731-2801, is one use).  Each instance sits in a
 * slot and keeps its own step size, simulated time and step count; its
 * results are bit-identical to solving it alone (dnx_solve with DNX_FP32).
 * Instances are packed into one device arena sized by a memory budget. */
typedef struct dnx_batch dnx_batch;

/* Opens an engine with `slots` instance slots and a device arena of
 * arena_bytes (<= 0: 80% of device memory).  prec must be DNX_FP32. */
int dnx_batch_open(dnx_ctx *ctx, int prec, int slots, int64_t arena_bytes, dnx_batch **out);
int dnx_batch_close(dnx_batch *b);

/* Execution shape per mode (0: V/dV and contributions in LDS, 1: V/dV in
 * LDS, 2: device memory): workgroup size (128/256/512/1024) and the LDS
 * bytes an instance may use to qualify for the mode.  Only while empty. */
int dnx_batch_set_shape(dnx_batch *b, int mode, int block, int lds_cap);

/* Uploads an instance (a host-only problem is fine) with initial state x0
 * (reference layout; unit pins applied) into a free slot.  The run ends at
 * the first verified state (stop_on_solve != 0) or at state step_cap.
 * DNX_E_BUDGET when no slot or arena range is free. */
int dnx_batch_admit(dnx_batch *b, const dnx_problem *p, const dnx_params *prm, const double *x0,
                    int64_t step_cap, int stop_on_solve, int *slot_out);

/* `count` admissions in one call: the same slots and contents as count
 * dnx_batch_admit calls in order (slots_out[i] is instance i's), with the
 * tables staged by host threads into pinned chunks, each uploaded in one
 * copy and placed by one kernel that also derives the transposed literal
 * index, and a single wait at the end.  prms may be NULL (defaults for each
 * instance's m, dnx_params_default with -m 10000).  All or none: on
 * DNX_E_BUDGET (too few slots, or the arena cannot hold them all) nothing
 * is admitted. */
int dnx_batch_admit_many(dnx_batch *b, int count, const dnx_problem *const *probs,
                         const dnx_params *prms, const double *const *x0s, int64_t step_cap,
                         int stop_on_solve, int *slots_out);

/* Enqueues one launch per execution shape that advances every running slot
 * by up to `steps` steps.  Returns without waiting. */
int dnx_batch_advance(dnx_batch *b, int steps);

/* Waits for queued work and copies the per-slot table (arrays of length
 * `slots`; any may be NULL): status (-1 free slot), steps of the evaluated
 * state, simulated time, and the device wall-clock tick at which the
 * instance ended (see dnx_batch_stamp). */
int dnx_batch_poll(dnx_batch *b, int *status, int64_t *steps, double *t, uint64_t *t_done_ticks);

/* Advances until no slot runs (or walltime_s passes when > 0), launching
 * steps_per_launch steps at a time.  *remaining = slots still running. */
int dnx_batch_run(dnx_batch *b, int steps_per_launch, double walltime_s, int *remaining);

/* Current device wall-clock tick and its rate. */
int dnx_batch_stamp(dnx_batch *b, uint64_t *ticks, double *ticks_per_s);

/* State of a slot in the reference layout (pending step applied). */
int dnx_batch_read_state(dnx_batch *b, int slot, double *x, int *status, int64_t *steps, double *t);

/* Sign assignment (+1/-1, x >= 0 positive) of a slot and, when unsat is
 * not NULL, the number of clauses it leaves unsatisfied (exact integer check
 * over the clause table the slot holds, read back from the device; the
 * batch keeps no host copy of admitted problems). */
int dnx_batch_assignment(dnx_batch *b, int slot, int8_t *assign, int *unsat);

/* Frees a slot and its arena range. */
int dnx_batch_retire(dnx_batch *b, int slot);

int dnx_batch_info(dnx_batch *b, int slot, int *n, int *m, int *mode);

/* Wide mode: when the running instances hold at most max_clauses clauses
 * in total, dnx_batch_run advances them with per-step kernels spread over
 * all of their clauses and variables (many workgroups per instance) instead
 * of one workgroup per instance, syncing status every steps_per_sync steps.
 * 0 disables it.  Results are bit-identical either way. */
int dnx_batch_set_wide(dnx_batch *b, int64_t max_clauses, int steps_per_sync);

/* The tail split (default on): when wide mode would be entered, the running
 * instances whose whole state fits in LDS in one piece keep running in the
 * fused kernel -- on a stream of their own, until each ends -- while wide
 * mode advances the others beside them.  Off: wide mode takes them all.
 * Results are bit-identical either way. */
int dnx_batch_set_tail_split(dnx_batch *b, int enable);

/* Order in which the fused kernel's workgroups take instances: those of at
 * least large_clauses clauses first, largest first (the longest ones start
 * at once), then the others smallest first (short instances end early
 * instead of waiting behind long ones).  0 (the default): all largest
 * first.  No counterpart in the reference, which solves one instance per
 * process; results are bit-identical in any order. */
int dnx_batch_set_order(dnx_batch *b, int64_t large_clauses);

/* Flow launches (default on): while every running instance is held in LDS
 * (mode 0) and dnx_batch_run has neither a wall-clock limit nor a group
 * stop, one fused launch keeps each instance on its workgroup until it ends
 * -- largest instances first, the next instance as soon as a workgroup
 * frees -- and ends once the running instances hold no more clauses than
 * the wide-mode threshold; steps_per_launch then only applies to the other
 * launches.  Off: every fused launch advances each instance at most
 * steps_per_launch steps.  Results are bit-identical either way. */
int dnx_batch_set_flow(dnx_batch *b, int enable);

/* The elastic tail (default on: min_clauses 6,000, steps 1,024,
 * clauses_per_workgroup 768): while every
 * running instance is held in LDS (mode 0) and flow launches apply, the
 * clause threshold below is not used.  Instead, once fewer instances run
 * than the device has CUs, the spare CUs go to the largest running
 * instances of at least min_clauses clauses as gangs (at least the
 * workgroups that hold their clauses in LDS and one per
 * clauses_per_workgroup clauses -- a smaller gang steps no faster than one
 * workgroup -- grown towards dnx_batch_set_gang's clauses_per_workgroup
 * while CUs are spare), in cooperative epoch launches
 * that keep every other running instance on a workgroup of its own; each
 * epoch ends once its gangs have taken `steps` steps, and the next is
 * planned from the statuses.  The flow launch before the first epoch stops
 * as soon as enough workgroups have run out of instances for the largest
 * instance's gang.  min_clauses = 0 turns it off (the clause threshold of
 * dnx_batch_set_wide_fraction then applies).  No counterpart in the
 * reference; results are bit-identical either way.  On the 4,096-instance
 * mixed workload a batch takes 273-278 ms with it (regrouped gangs), 372
 * ms with the threshold tail. */
int dnx_batch_set_epoch(dnx_batch *b, int64_t min_clauses, int steps, int clauses_per_workgroup);

/* How epoch gangs of 3-SAT instances exchange contributions (default 1).
 * 1: regrouped -- each gang workgroup stages its clauses' contributions in
 * LDS grouped by the workgroup that folds them and writes them as one
 * contiguous run, and each folding workgroup reads its share as one run per
 * writer; 0: every contribution goes to its own transpose slot and is
 * gathered back through the transpose.  Instances with other clause lengths
 * always take the transpose.  No counterpart in the reference; results are
 * bit-identical either way. */
int dnx_batch_set_gang_regroup(dnx_batch *b, int enable);

/* The tail threshold relative to the run: once the running instances hold
 * at most `fraction` of the clauses that were running when dnx_batch_run
 * began -- and, with gang mode on, at most 75% of the clauses its workgroup
 * budget holds (2048 per workgroup) -- gang mode (when the running set fits
 * it) or else wide mode takes over (the default: 0.0125, 300 steps per
 * wide-mode sync). */
int dnx_batch_set_wide_fraction(dnx_batch *b, double fraction, int steps_per_sync);

/* Serving: requests with at least `clauses` clauses run in the wide lane
 * (default 60,000; needs wide mode enabled).  Each wide chunk (30 steps,
 * then a status sync) advances members in the order dnx_batch_set_serve_order
 * sets, up to active_clauses clauses in total (default 6,000,000). */
int dnx_batch_set_serve_big(dnx_batch *b, int64_t clauses, int64_t active_clauses);

/* Serving: order of the wide lane's members for its active set.  Members
 * that have run fewer than las_steps steps come first, least attained
 * service among them; the rest follow by least expected work, m times
 * exp(-1.5 m/n) (random 3-SAT needs fewer steps the further its clause
 * ratio lies above the threshold), so an overloaded lane finishes what it
 * can instead of advancing every long run equally.  las_steps = 0 orders
 * every member by least attained service.  Default 1024. */
int dnx_batch_set_serve_order(dnx_batch *b, int64_t las_steps);

/* Serving: 3-SAT requests with at least min_clauses clauses (default
 * 60,000; 0 turns the lane off) run in the sliced lane: the single-instance
 * path's sliced exchange (each request's plan built on host threads after
 * arrival, units of about 3,000 clauses), their state resident in the
 * batch's arena, at most max_active (1-4096, default 1024) resident at a
 * time.  Each 30-step chunk advances an active set of them together, one
 * launch pair per step over all their units and fold blocks, chosen in the
 * wide lane's order (least attained service below
 * dnx_batch_set_serve_order's las_steps, then least expected work) up to
 * dnx_batch_set_serve_big's active_clauses.  A request whose plan does not
 * apply goes to the other lanes.  Results are the bits of solving it
 * alone. */
int dnx_batch_set_serve_sliced(dnx_batch *b, int64_t min_clauses, int max_active);

/* Requests the last serving call (dnx_serve_trace / dnx_serve_planted)
 * retired in each lane: counts[0] narrow, [1] wide, [2] sliced, [3] larger
 * than the arena (DNX_FAILED). */
int dnx_batch_served_by(dnx_batch *b, int64_t *counts);

/* Autotuner: times candidate execution shapes on `sample` synthetic planted
 * instances (n log-uniform in [n_min, n_max], clause ratio uniform in
 * [r_min, r_max]) for `steps` fixed steps each -- workgroup size of modes 0
 * and 1 and the LDS budget of mode 0 (the fused kernel's clause tile), then
 * the tail's clauses per gang workgroup (256 / 384 / 768, on the sample's
 * eight largest instances in gangs) -- and applies the fastest to the empty
 * batch b.  report (may be NULL) receives report_len doubles: one 5-tuple per
 * measurement (kind 0/1/2 = mode-0 shape / mode-1 block / gang clauses per
 * workgroup, block or clauses, lds_cap, ms, GB/s); *report_count is the
 * number of doubles written. */
int dnx_batch_tune(dnx_batch *b, int n_min, int n_max, double r_min, double r_max, int sample,
                   int steps, uint64_t seed, double *report, int report_len, int *report_count);

/* The current shape of a mode and the wide-mode threshold. */
int dnx_batch_get_shape(dnx_batch *b, int mode, int *block, int *lds_cap, int64_t *wide_clauses);

/* Gang mode: the tail of a run (the instances left when wide mode would be
 * taken, or -- with the tail split -- the wide-mode instances left once the
 * fused side launch has ended) is advanced by gangs of up to 32 cooperating
 * workgroups per instance, all of one gang on one XCD, steps_per_launch
 * steps per cooperative launch, about clauses_per_workgroup clauses per
 * workgroup (defaults 256, 384, 512; max_blocks 0 disables it).  When the
 * running instances ask for more than max_blocks workgroups (or more than
 * can be resident at once), every gang is cut to a common size, never below
 * what holds its clauses in LDS (2048 per workgroup); if even that does not
 * fit, wide mode runs until enough instances have ended.  Results are
 * bit-identical to the other modes. */
int dnx_batch_set_gang(dnx_batch *b, int max_blocks, int clauses_per_workgroup, int steps_per_launch);
/* The gang settings in force (what dnx_batch_set_gang or the autotuner set). */
int dnx_batch_get_gang(dnx_batch *b, int *max_blocks, int *clauses_per_workgroup, int *steps_per_launch);

/* Waits for queued work, then reports the summed device time (ms) and the
 * count of timed launches since the last reset, in seven buckets: modes 0-2
 * of the fused kernel, wide mode, gang mode, the fused kernel's tail launch
 * beside wide mode, and epoch launches (arrays of 7).  alg_bytes (may be
 * NULL) receives the compulsory bytes of the steps those launches took
 * inside dnx_batch_run: per instance step, 4 per literal + 16 per clause +
 * 8 per variable.  enable = 1/0 turns launch timing on/off and resets the
 * sums (-1: report only). */
int dnx_batch_timing(dnx_batch *b, int enable, double *ms, int64_t *launches, double *alg_bytes);
int dnx_batch_memory(dnx_batch *b, int64_t *arena_bytes, int64_t *free_bytes);

/* ---- serving ---------------------------------------------------------- */
/* Continuous batching over a batch engine (no counterpart in the reference).
 * Request i becomes visible arrival_s[i] seconds after the call starts
 * (non-decreasing); it is admitted as soon as a slot and an arena range are
 * free (first come, first served), advanced alongside everything resident in
 * launches of steps_per_launch steps (a new request waits for the running
 * launch to end: 128-256 keeps that wait to a few ms), and retired the
 * moment its state
 * verifies or reaches step_cap, freeing its memory for the queue.  Per
 * request: status (DNX_SAT only after the exact clause check; DNX_FAILED for
 * a request larger than the arena), steps, time-to-solution from arrival to
 * retirement by the device clock, and queueing delay before admission.
 * prms may be NULL (reference defaults).  With deadline_s > 0 the loop
 * stops deadline_s seconds after the call starts; requests not finished by
 * then stay DNX_UNKNOWN with time-to-solution -1.  Optional outputs (NULL to
 * skip; a NULL entry skips one request): t_out[i] the simulated time of the
 * request's last state, assign_out[i] (n entries) its sign assignment and
 * x_out[i] (n + 2m entries) its final state in the reference layout, read
 * as the request retires. */
int dnx_serve_trace(dnx_batch *b, int count, dnx_problem *const *probs, const dnx_params *prms,
                    const double *const *x0s, const double *arrival_s, int64_t step_cap,
                    int steps_per_launch, double deadline_s, int *status, int64_t *steps,
                    double *tts_s, double *wait_s, double *makespan_s, double *t_out,
                    int8_t *const *assign_out, double *const *x_out, int *where_out);

/* where_out[i] (optional, both serving calls): where request i was when the
 * loop ended.  For a request still DNX_UNKNOWN at the deadline, steps[i] is
 * then the steps it had completed by the last status the loop read. */
#define DNX_SERVE_ENDED 0       /* retired: SAT, step cap, or DNX_FAILED */
#define DNX_SERVE_QUEUED 1      /* arrived, never admitted (no slot or arena range free) */
#define DNX_SERVE_NARROW 2      /* running in the fused-kernel lane */
#define DNX_SERVE_WIDE 3        /* running in the wide lane, in its active set */
#define DNX_SERVE_WIDE_WAIT 4   /* resident in the wide lane, waiting for its active set */
#define DNX_SERVE_NOT_ARRIVED 5 /* arrival after the deadline */
#define DNX_SERVE_SLICED 6      /* resident in the sliced lane (dnx_batch_set_serve_sliced) */
#define DNX_SERVE_SLICED_WAIT 7 /* routed to the sliced lane, plan or room not ready */

/* The served planted stream: Poisson arrivals at rate_per_s for duration_s
 * (at most max_requests), n log-uniform in [n_min, n_max], clause ratio
 * uniform in [r_min, r_max], literal truth counts as dnx_generate_planted,
 * v ~ U[-1,1), Xs = 0.5, Xl = 1.  gen_threads host threads build instances
 * ahead of their arrival.  Outputs are arrays of max_requests entries;
 * *count_out requests were issued. */
int dnx_serve_planted(dnx_batch *b, uint64_t seed, double rate_per_s, double duration_s, int n_min,
                      int n_max, double r_min, double r_max, double p1, double p2, int64_t step_cap,
                      int steps_per_launch, double deadline_s, int max_requests, int gen_threads,
                      int *count_out, double *arrival_out, int *n_out, int *m_out, int *status,
                      int64_t *steps, double *tts_s, double *wait_s, double *makespan_s,
                      int *where_out);

/* ---- measurement ------------------------------------------------------ */
/* Streams `bytes` (rounded down to a multiple of 16) from one device buffer
 * to another `repeats` times with a plain copy kernel and reports the mean
 * device time per copy (ms).  The known-size reference for memory-traffic
 * counters and for the achievable streaming bandwidth. */
int dnx_device_copy(dnx_ctx *ctx, int64_t bytes, int repeats, double *ms);

/* ---- host utilities -------------------------------------------------- */
/* Initial state of the reference's thread 0 for seed `seed`
 * (dynex.cc:2643-2700): seed 0 sets every entry to 1.0; other seeds draw V
 * from std::uniform_int_distribution<int>(-1,1) over std::mt19937(seed) with
 * Xs = 0 and Xl = 1; seed -1 / -2 set V to 0 / -1; unit clauses pin V. */
int dnx_init_reference(const dnx_problem *p, int seed, double *x0);

/* Initial state of thread `tid` of a multi-thread run (dynex.cc:2643-2700):
 * with seed 0, threads 0/1/2 set every entry to 1/0/-1; every other thread
 * draws V from one mt19937(seed) stream shared in thread order (Xs = 0,
 * Xl = 1).  Thread 0 is dnx_init_reference. */
int dnx_init_reference_thread(const dnx_problem *p, int seed, int tid, double *x0);

/* Planted random k=3 instance (no counterpart in the reference, which reads
 * files only): sigma uniform +-1, three distinct variables per clause, the
 * number t of literals true under sigma drawn with P(1)=p1, P(2)=p2,
 * P(3)=1-p1-p2, the true ones a uniform subset.  lits_out[3m];
 * planted_out[n] (may be NULL) gets sigma. */
int dnx_generate_planted(uint64_t seed, int n, int m, double p1, double p2,
                         int32_t *lits_out, int8_t *planted_out);

/* Initial state v ~ U[-1,1), Xs = xs0, Xl = xl0 from a seeded stream. */
int dnx_init_uniform(uint64_t seed, int n, int m, double xs0, double xl0,
                     double *x0);

/* Writes a DIMACS file for a clause table. */
int dnx_write_dimacs(const char *path, int n, int m, const int32_t *clause_ptr,
                     const int32_t *lits, const char *comment);

/* Exact integer clause check of an assignment (+1/-1 per variable):
 * returns the number of unsatisfied clauses through *unsat. */
int dnx_check_assignment(const dnx_problem *p, const int8_t *assign, int *unsat);

#ifdef __cplusplus
}
#endif
#endif /* DNX_H */
