/*
 * dnx_oracle.c -- CPU restatement of the reference integrator's hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests, the
 * smoke() entry and the bench's CPU-baseline leg load it to compare against
 * or to time as the reference's CPU path.  The product (libdnx) never links,
 * loads or calls it.
 *
 * Built with  gcc -O2 -ffp-contract=off  (see oracle/Makefile) so every
 * floating-point operation happens in the order written, without fused
 * multiply-add: the reference's own README build uses -Ofast (README.md:24),
 * which permits reassociation; strict IEEE evaluation of the source order is
 * the definition pinned here.
 *
 * Contents and the reference lines each part follows:
 *   orc_dimacs_*     DIMACS reader            dynex.cc:2306-2392
 *   orc_xl_max       xl_max scaling           dynex.cc:2358-2359
 *   orc_unit_vars    unit-clause pins         dynex.cc:2458-2469
 *   orc_init_ref     initial state, thread 0  dynex.cc:2643-2700
 *   orc_mt19937_*    std::mt19937 (C++ [rand.eng.mers], the engine used at
 *                    dynex.cc:2636) and libstdc++-11's
 *                    uniform_int_distribution (Lemire's nearly divisionless
 *                    method, bits/uniform_int_dist.h) used at dynex.cc:2668
 *   orc_rhs_*        dmm_generate_dxdt        dynex.cc:332-530
 *   orc_dostep_*     dmm_dostep               dynex.cc:301-327
 *   orc_step_*       adaptive ladder          dynex.cc:1387-1417
 *   orc_verify_*     exact clause check       dynex.cc:672-687
 *   orc_solve_*      dmm() mode-6 driver      dynex.cc:1381-1427
 */
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAX_LITS 25 /* MAX_LITS_SYSTEM, dynex.cc:69 */

typedef struct {
    double alpha, beta, gamma, delta, epsilon, zeta;
    long long xl_max; /* the reference's int after scaling by m */
} orc_params;

/* ------------------------------------------------------------------ */
/* DIMACS reader (dynex.cc:2306-2392).                                 */
/*                                                                     */
/* Tokens are skipped until one equals "p"; " cnf %i %i" then gives n   */
/* and m.  Each clause is read token by token until a token equal to   */
/* "0"; a token "c" skips the rest of its line; every other token is   */
/* atoi'd and stored, non-zero values count as literals.  Errors the   */
/* reference does not check (missing file, EOF, variable out of range) */
/* are reported as negative codes here instead of crashing.            */
/* ------------------------------------------------------------------ */
#define ORC_E_IO -2
#define ORC_E_PARSE -3
#define ORC_E_TOOLONG -4

static int orc_read_header(FILE *f, int *n, int *m)
{
    char buf[4096];
    for (;;) {
        if (fscanf(f, "%4095s", buf) != 1) return ORC_E_PARSE;
        if (strcmp(buf, "p") == 0) break;
    }
    if (fscanf(f, " cnf %i %i", n, m) != 2) return ORC_E_PARSE;
    if (*n < 0 || *m < 0) return ORC_E_PARSE;
    return 0;
}

/* Reads the clauses.  When cp/lits are NULL only *nnz is computed. */
static int orc_read_clauses(FILE *f, int n, int m, long long *nnz, int *cp,
                            int *lits)
{
    char buf[4096];
    long long pos = 0;
    for (int c = 0; c < m; c++) {
        int k = 0;
        if (cp) cp[c] = (int)pos;
        for (;;) {
            if (fscanf(f, "%4095s", buf) != 1) return ORC_E_PARSE;
            if (strcmp(buf, "c") == 0) {
                /* dynex.cc:2372-2375: skip the rest of the line; the
                 * loop condition then tests the skipped text against "0" */
                int ch;
                while ((ch = fgetc(f)) != EOF && ch != '\n') {
                }
                continue;
            }
            int lit = atoi(buf);
            if (lit != 0) {
                if (abs(lit) > n) return ORC_E_PARSE;
                if (k >= ORC_MAX_LITS) return ORC_E_TOOLONG; /* dynex.cc:2388 */
                if (lits) lits[pos] = lit;
                pos++;
                k++;
            }
            if (strcmp(buf, "0") == 0) break;
        }
    }
    if (cp) cp[m] = (int)pos;
    *nnz = pos;
    return 0;
}

int orc_dimacs_header(const char *path, int *n, int *m, long long *nnz)
{
    FILE *f = fopen(path, "r");
    if (!f) return ORC_E_IO;
    int rc = orc_read_header(f, n, m);
    if (rc == 0) rc = orc_read_clauses(f, *n, *m, nnz, NULL, NULL);
    fclose(f);
    return rc;
}

int orc_dimacs_read(const char *path, int n, int m, int *cp, int *lits)
{
    FILE *f = fopen(path, "r");
    if (!f) return ORC_E_IO;
    int n2, m2;
    long long nnz;
    int rc = orc_read_header(f, &n2, &m2);
    if (rc == 0 && (n2 != n || m2 != m)) rc = ORC_E_PARSE;
    if (rc == 0) rc = orc_read_clauses(f, n, m, &nnz, cp, lits);
    fclose(f);
    return rc;
}

/* dynex.cc:2358-2359: xl_max = xl_max*m in int arithmetic (wrapping), and
 * INT_MAX when the product is not positive. */
long long orc_xl_max(int xl_max_flag, int m)
{
    uint32_t w = (uint32_t)xl_max_flag * (uint32_t)m;
    int32_t s = (int32_t)w;
    if (s <= 0) s = INT_MAX;
    return (long long)s;
}

/* dynex.cc:2458-2469: unit[v-1] = sign of the last unit clause on v. */
void orc_unit_vars(int n, int m, const int *cp, const int *lits,
                   signed char *unit)
{
    memset(unit, 0, (size_t)n);
    for (int c = 0; c < m; c++) {
        if (cp[c + 1] - cp[c] == 1) {
            int lit = lits[cp[c]];
            unit[abs(lit) - 1] = lit > 0 ? 1 : -1;
        }
    }
}

/* ------------------------------------------------------------------ */
/* std::mt19937 (32-bit Mersenne Twister, C++ [rand.eng.mers]).        */
/* ------------------------------------------------------------------ */
typedef struct {
    uint32_t mt[624];
    int idx;
} orc_mt19937;

void orc_mt19937_seed(orc_mt19937 *g, uint32_t s)
{
    g->mt[0] = s;
    for (int i = 1; i < 624; i++)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->idx = 624;
}

uint32_t orc_mt19937_next(orc_mt19937 *g)
{
    if (g->idx >= 624) {
        for (int i = 0; i < 624; i++) {
            uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            g->mt[i] = g->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        g->idx = 0;
    }
    uint32_t y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* libstdc++-11 uniform_int_distribution<int>(a,b) for a 32-bit engine:
 * Lemire's nearly divisionless reduction onto [0, b-a+1). */
int orc_uniform_int(orc_mt19937 *g, int a, int b)
{
    uint32_t range = (uint32_t)b - (uint32_t)a + 1u;
    uint64_t prod = (uint64_t)orc_mt19937_next(g) * range;
    uint32_t low = (uint32_t)prod;
    if (low < range) {
        uint32_t thr = (uint32_t)(-range) % range;
        while (low < thr) {
            prod = (uint64_t)orc_mt19937_next(g) * range;
            low = (uint32_t)prod;
        }
    }
    return (int)((uint32_t)a + (uint32_t)(prod >> 32));
}

/* Test hook: first `count` draws of uniform_int(a,b) from mt19937(seed). */
void orc_uniform_int_stream(int seed, int a, int b, int count, int *out)
{
    orc_mt19937 g;
    orc_mt19937_seed(&g, (uint32_t)seed);
    for (int i = 0; i < count; i++) out[i] = orc_uniform_int(&g, a, b);
}

uint32_t orc_mt19937_nth(uint32_t seed, int count)
{
    orc_mt19937 g;
    orc_mt19937_seed(&g, seed);
    uint32_t y = 0;
    for (int i = 0; i < count; i++) y = orc_mt19937_next(&g);
    return y;
}

/* Initial state of thread 0 (dynex.cc:2643-2700), reference layout.
 * seed 0: every entry of [V|Xs|Xl] is 1.0.  Any other seed: V drawn from
 * uniform_int(-1,1) on mt19937(seed), Xs = 0, Xl = 1; seed -1 then sets V
 * to 0 and seed -2 to -1.  Unit clauses override V. */
void orc_init_ref(int n, int m, const signed char *unit, int seed, double *x)
{
    size_t N = (size_t)n + 2 * (size_t)m;
    if (seed == 0) {
        for (size_t j = 0; j < N; j++) x[j] = 1.0;
    } else {
        orc_mt19937 g;
        orc_mt19937_seed(&g, (uint32_t)seed);
        for (int j = 0; j < n; j++) x[j] = (double)orc_uniform_int(&g, -1, 1);
        for (int j = n; j < n + m; j++) x[j] = 0.0;
        for (size_t j = (size_t)n + m; j < N; j++) x[j] = 1.0;
    }
    if (seed == -1)
        for (int j = 0; j < n; j++) x[j] = 0.0;
    if (seed == -2)
        for (int j = 0; j < n; j++) x[j] = -1.0;
    if (unit)
        for (int j = 0; j < n; j++)
            if (unit[j] != 0) x[j] = (double)unit[j];
}

/* ------------------------------------------------------------------ */
/* Precision instances of the hot path.                                */
/* ------------------------------------------------------------------ */
#define T double
#define SFX f64
#define MINF fmin
#define ORC_FABS fabs
#define ORC_SORTED 1
#include "oracle_core.inc"
#undef T
#undef SFX
#undef MINF
#undef ORC_FABS
#undef ORC_SORTED

#define T float
#define SFX f32
#define MINF fminf
#define ORC_FABS fabsf
#define ORC_SORTED 0
#include "oracle_core.inc"
#undef T
#undef SFX
#undef MINF
#undef ORC_FABS
#undef ORC_SORTED
