// host.hpp -- host-side problem model shared by the C ABI, the batch engine
// and the CLI.  Everything here is plain C++17 (no HIP types).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "dnx.h"

namespace dnx {

// Clause table in CSR form plus the variable-major transpose the kernels
// gather through (it replaces the reference's padded `occurrence` table,
// dynex.cc:2412-2422, and the CUDA file's atomics, dynex.cu:244-256).
struct HostProblem {
    int n = 0, m = 0;
    std::vector<int32_t> cp;    // m+1 clause offsets into lits
    std::vector<int32_t> lits;  // signed 1-based literals (dynex.cc:2380)
    std::vector<int8_t> unit;   // n: pin of V from unit clauses (dynex.cc:2458-2469)
    bool has_unit = false;
    bool uniform3 = false;      // every clause has exactly 3 literals
    int max_deg = 0;
    // transpose: var v owns slots [row_ptr[v], row_ptr[v+1]) in ascending
    // literal-position order; lpos[p] is the slot of literal position p.
    std::vector<int32_t> row_ptr;  // n+1
    std::vector<int32_t> lpos;     // nnz
    std::vector<int32_t> eidx;     // nnz: literal position of each slot (inverse of lpos)

    int64_t nnz() const { return (int64_t)lits.size(); }
};

// Fills unit/uniform3/row_ptr/lpos from n, m, cp, lits.  Returns DNX_OK or
// DNX_E_ARG when a literal is 0 or out of range.
int finalize_problem(HostProblem &p);

// DIMACS reader with the reference's token rules (dynex.cc:2306-2392).
int parse_dimacs(const char *path, HostProblem &p, std::string &err);

int write_dimacs(const char *path, int n, int m, const int32_t *cp,
                 const int32_t *lits, const char *comment);

// Reference thread-0 initial state (dynex.cc:2643-2700).
void init_reference(const HostProblem &p, int seed, double *x0);
// Initial state of thread `tid` of a multi-thread run (same lines).
void init_reference_thread(const HostProblem &p, int seed, int tid, double *x0);

// splitmix64 stream: the portable generator for synthetic workloads.
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    // uniform integer in [0, bound) by 64x64->128 multiply-shift
    uint32_t below(uint32_t bound) {
        return (uint32_t)(((unsigned __int128)next() * bound) >> 64);
    }
    // uniform double in [0, 1) with 53 random bits
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

void generate_planted(uint64_t seed, int n, int m, double p1, double p2,
                      int32_t *lits, int8_t *planted);
void init_uniform(uint64_t seed, int n, int m, double xs0, double xl0, double *x0);

int count_unsat(const HostProblem &p, const int8_t *assign);

// Tables of the grouped exchange (3-SAT, n <= GX_MAX_VARS): variable v is
// folded by the variable-pass workgroup v / vb, which runs on XCD
// (v / vb) % 8 (workgroups are dealt to the 8 XCDs in turn), so group
// g = (v / vb) & 7.  Contributions are stored group by group, each group in
// literal order; gw/gbase let the clause pass find a literal's position,
// geidx gives the fold each slot's position.  False when not 3-SAT or too
// many variables.
bool grouped_tables(const HostProblem &p, int vb, std::vector<uint32_t> &gw, std::vector<int32_t> &gbase,
                    std::vector<int32_t> &geidx);

// Tables of the sliced exchange (3-SAT).  The variables are cut into slices
// of S; every clause touches at most three.  Clauses are regrouped into
// units, each holding clauses whose variables all lie in the same (at most
// three) slices, about m / cus clauses per unit and at most max_rounds x cus
// units (one workgroup per compute unit at a time), so
// one workgroup per unit can hold the unit's slices of V in LDS and read V
// from there.  A unit's clauses are taken SUB at a time (a sub-batch); a
// sub-batch's contributions are staged in LDS ordered by destination fold
// block (F variables) and written out as one contiguous run per block.  e
// is block-major: fold block b owns positions [row_ptr[bF], row_ptr[bF+F])
// and its runs arrive there in processing order; q gives, for each
// position, the variable-major slot (relative to the block's first slot)
// of the literal stored there.
struct SlicedPlan {
    int S = 0, SUB = 0, F = 0;
    int max_region = 0;             // slots of the largest fold block
    // every fold block's region of e has the same length, stride (max_region
    // rounded up to four): block b owns [b stride, (b + 1) stride), its
    // padding's q entries name slot `stride` (a sink), and e holds
    // e_len = nblocks x stride entries plus the clause pass's sink after them
    int stride = 0;
    long long e_len = 0;
    // 3 per clause, processing order: sign << 31 | LDS index << 15 | run << 7
    // | offset in the run (the run's index within its sub-batch)
    std::vector<uint32_t> gw;
    std::vector<int32_t> units;     // 8 per unit: c0, c1, nslices, slice0..2, first sub-batch, first run
    std::vector<int32_t> sb_run;    // per sub-batch, its first run; one more entry at the end
    std::vector<int32_t> runs;      // 2 per run: staging start << 16 | length, e position
    std::vector<uint16_t> q;        // per e position
    std::vector<int32_t> pidx;      // per clause, its processing position
};
bool sliced_tables(const HostProblem &p, int S, int SUB, int F, int cus, int max_rounds, int region_cap,
                   SlicedPlan &out);

// Thread-local error message used by dnx_last_error().
void set_error(const std::string &msg);

}  // namespace dnx
