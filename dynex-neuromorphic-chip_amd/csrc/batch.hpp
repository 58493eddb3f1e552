// batch.hpp -- device structures of the batched engine, shared by the
// engine (batch_engine.cpp) and its kernels (batch.hip).  No HIP headers.
#pragma once

#include <cstddef>
#include <cstdint>

#include "kernels.hpp"

namespace dnx {

// Execution shape of one instance inside the fused step kernel.
//   0  V (and dV) and the contribution array live in LDS
//   1  V (and dV) live in LDS, contributions in device memory
//   2  both in device memory (instances too large for LDS)
enum BatchMode { BM_LDS_ALL = 0, BM_LDS_V = 1, BM_GLOBAL = 2, BM_COUNT = 3 };
constexpr int KC_MAX = 16;  // mode 0: chunks of consecutive clauses per instance at most

// LDS bytes of the fused kernel's per-variable arrays in modes 0 and 1: V,
// and one float beside it (the partial fold between chunks, then the new dV),
// each rounded to 16 bytes; mode 0's chunk contributions follow them.
constexpr size_t lds_var_bytes(int n) { return 2 * (((size_t)n * 4 + 15) / 16 * 16); }

struct alignas(8) BF2 {
    float v, d;
};

// Read-only descriptor of an admitted instance (one per slot).
struct BInst {
    int n, m;
    int mode;
    int flags;                // BF_GENERAL: clause_ptr valid; BF_UNIT: pins
    const int32_t *lits;      // signed 1-based literals, clause order
    const int32_t *cp;        // m+1 offsets (general clause lengths only)
    const int32_t *lpos;      // mode 0: LDS slot of each literal within its chunk
    const uint32_t *pk;       // mode 0, 3-SAT: literal words, sign << 31 | var << 16 | LDS slot
    const int32_t *gpos;      // variable-major (transpose) slot of each literal
    const int32_t *eidx;      // literal of each variable-major slot (global e)
    const int32_t *row_ptr;   // n+1
    const int8_t *unit;       // n pins
    float *xs, *xl;           // m each
    BF2 *vd;                  // n resident [V, dV] pairs
    float *e;                 // nnz contributions (modes 1, 2, wide, gang)
    // mode 0: nchunk chunks of consecutive clauses [cbound[k], cbound[k+1]);
    // chunk k's LDS slots of variable v are [cbase[k(n+1)+v], cbase[k(n+1)+v+1])
    // (one chunk: cbase == row_ptr)
    const int32_t *cbound;
    const int32_t *cbase;
    float *cst;               // fused kernel: each clause's C of the last state it evaluated
    int nchunk, pad0;
    float alpha, beta, gamma, delta, eps, zeta, xlmax, pad;
};
enum { BF_GENERAL = 1, BF_UNIT = 2 };

// Mutable per-slot state.  `steps`, `t` and `plevel` follow the same lazy
// convention as the single-instance control block: the stored V/Xs/Xl are
// state `steps`; when plevel >= 0 the step from it at that ladder level has
// been decided and is applied at the start of the next launch.
struct BDyn {
    int status;               // DNX_RUNNING / DNX_SAT / DNX_UNKNOWN, -1 free slot
    int plevel;
    int stop_on_solve;
    int loc;                  // unused by the kernel, kept for alignment
    long long steps;
    long long cap;
    double t;
    unsigned long long t_done;  // device wall clock when the instance ended
};

// Flow control of a fused launch whose workgroups keep their instances
// until they end (instead of K steps each): an instance that ends takes its
// clauses off `live`, and the one that brings live to `threshold` or below
// raises `stop`; every workgroup then leaves after its current step, with
// its instance's state written back, and takes no further instance.  With
// idle_trigger > 0, `idle` counts the workgroups that found the list empty,
// and the one that brings it to idle_trigger raises `stop` too.  In an epoch
// launch (see launch_epoch) the gang workgroups count themselves out in
// `gangs_done`, and the last of `gangs` to finish raises `stop`.
struct BFlow {
    long long live;
    long long threshold;
    int stop;
    int idle;
    int idle_trigger;
    int gangs;
    int gangs_done;
    int xcd_rank[8];  // epoch launch: workgroups that have started on each XCD
    int pad[15];
};
static_assert(sizeof(BFlow) == 128, "two cache lines");

// one launch of the fused kernel over `nwork` slots listed in `work`: K steps
// per instance, or (flow != nullptr) each instance until it ends or the flow
// stops
void launch_batch(int mode, int bs, int lds_bytes, int grid, const BInst *insts, BDyn *dyn,
                  const int *work, int nwork, int *head, int ksteps, BFlow *flow, void *stream);
void launch_batch_zero(int *p, int count, void *stream);
void launch_batch_stamp(unsigned long long *p, void *stream);
// materialised state of one slot in the reference layout (float)
void launch_batch_materialize(const BInst *inst, const BDyn *dyn, int n, int m, float *x_out, void *stream);

// Admission: one instance's staged tables and where they go.  The host stages
// a prefix of the instance's arena range (tables only the host can build);
// the device copies it into place and derives the rest: the variable-major
// literal index (eidx[pos[q]] = q), mode 0's packed literal words, and
// zeroed contribution and stored-C arrays.  Offsets are bytes; `src` is
// relative to the staging buffer, `dst` to the arena, the others to the
// instance's range; word counts cover whole 256-byte-aligned regions.
struct AdmitRec {
    unsigned long long src, dst, prefix;  // prefix: staged bytes (a multiple of 256)
    long long nnz;
    int slot, fill_pk;                    // fill_pk: mode-0 3-SAT packed words
    unsigned long long lits, pos, lpos;   // staged: literals, global transpose slots, pk's LDS slots
    unsigned long long pk, eidx, e, cst;  // derived (pk 0: none)
    unsigned long long pk_words, eidx_words, e_words, cst_words;
    BInst inst;
    BDyn dyn;
};
// places `nrec` records (each record's tables, then its descriptor and
// status entry at its slot)
void launch_admit_place(const AdmitRec *recs, int nrec, const unsigned char *stage, unsigned char *arena,
                        BInst *inst, BDyn *dyn, void *stream);
int batch_max_threads();

// wide mode (see batch.hip); maps hold (slot, local block) pairs
constexpr int WIDE_BLOCK = 256;  // clauses or variables per wide block
void launch_wide_enter(const BDyn *dyn, Ctl *wctl, const int *list, int count, int r0, void *stream);
void launch_wide_sync(BDyn *dyn, Ctl *wctl, const int *list, int count, int rl, void *stream);
void launch_wide_step(const BInst *insts, Ctl *wctl, const int *bmap, int nbc, const int *vmap, int nbv, int r,
                      void *stream);
// resident workgroups per CU for a shape (0 if it cannot launch)
int batch_occupancy(int mode, int bs, int lds_bytes);

// gang mode (see batch.hip): map holds (slot, rank, G, gang) per workgroup,
// every gang's workgroups at block indices congruent to one XCD modulo 8;
// words holds 2 * GANG_MAX lines of GANG_LINE u64 per gang
constexpr int GANG_MAX = 32;     // workgroups per gang
constexpr int GANG_CPW = 2048;   // clauses per workgroup at most (LDS)
constexpr int GANG_LINE = 16;    // u64 per barrier line (128 bytes)
int gang_occupancy();
// cooperative launch of K steps; returns the hipError_t of the launch
constexpr int GANG_XCD = 8;      // XCDs the dispatcher deals workgroups to, in turn
// *err: 1 a barrier wait timed out, 2 a workgroup ran on another XCD than its
// gang (the launch then changed nothing); *arrive counts arrivals across
// launches (this launch waits for arrive_base + nblk)
int launch_gang(const BInst *insts, BDyn *dyn, const int *map, int nblk, unsigned long long *words,
                unsigned long long base, int K, int *err, int *arrive, int arrive_base, void *stream);

// Regrouped gangs (3-SAT; see gang_rx in batch.hip): per-gang scratch.  A
// workgroup evaluates at most GANG_CPW clauses, folds at most RX_SLOTS
// contributions and owns at most RX_VCAP variables; the instance has at most
// RX_NCAP variables.
constexpr int RX_SLOTS = 3 * GANG_CPW;
constexpr int RX_VCAP = 2048;
constexpr int RX_NCAP = 8192;
struct RxLayout {
    unsigned long long ew, rank, szt, total;  // byte offsets in the gang's scratch, and its size
};
constexpr unsigned long long rx_align(unsigned long long x) { return (x + 255) / 256 * 256; }
// ew: G writer areas of 3 Cb contributions (Cb = ceil(m / G) clauses per
// workgroup); rank: 3m ints (each variable-major slot's place in its region);
// szt: G x G ints (region sizes)
constexpr RxLayout rx_layout(int n, int m, int G)
{
    const unsigned long long cb = (unsigned long long)((m + G - 1) / G);
    const unsigned long long ew = 0, rk = ew + rx_align(4ull * (unsigned long long)G * 3ull * cb);
    const unsigned long long sz = rk + rx_align(4ull * 3ull * (unsigned long long)m);
    (void)n;
    return RxLayout{ew, rk, sz, sz + rx_align(4ull * (unsigned long long)G * (unsigned long long)G)};
}
// variables a regrouped gang's workgroup owns
constexpr int rx_vblock(int n, int G) { return (n + G - 1) / G; }

// epoch launch (see batch.hip): `grid` workgroups (one per CU), gangs listed
// per XCD in gmap (GANG_XCD x GANG_MAX entries of (slot, rank, G | regrouped
// << 16, gang)) with gcount[x] members on XCD x, flow workers over `work`;
// flow->gangs holds the total of gcount, and flow's counters start at zero.
// A regrouped gang uses the rx_layout scratch at gscr + goff[gang].  *err: 1 a
// gang barrier timed out, 3 not every gang member started (the gangs changed
// nothing), 4 a regrouped gang's tables disagreed.  Returns the hipError_t of
// the launch.
int epoch_occupancy(int flow_lds);
int launch_epoch(const BInst *insts, BDyn *dyn, const int *gmap, const int *gcount, unsigned long long *words,
                 unsigned long long base, int K, int *err, const int *work, int nwork, int *head, BFlow *flow,
                 unsigned char *gscr, const unsigned long long *goff, int flow_lds, int grid, void *stream);

}  // namespace dnx
