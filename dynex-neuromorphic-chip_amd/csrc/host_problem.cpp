// host_problem.cpp -- DIMACS I/O, CSR transpose, initial states and the
// planted-instance generator.
#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "host.hpp"

namespace dnx {

int finalize_problem(HostProblem &p)
{
    const int n = p.n, m = p.m;
    if ((int)p.cp.size() != m + 1 || p.cp[0] != 0 || (int64_t)p.cp[m] != p.nnz())
        return DNX_E_ARG;
    p.unit.assign((size_t)n, 0);
    p.has_unit = false;
    p.uniform3 = true;
    std::vector<int32_t> deg((size_t)n + 1, 0);
    for (int c = 0; c < m; c++) {
        int k = p.cp[c + 1] - p.cp[c];
        if (k < 0 || k > DNX_MAX_LITS) return k < 0 ? DNX_E_ARG : DNX_E_TOOLONG;
        if (k != 3) p.uniform3 = false;
        for (int q = p.cp[c]; q < p.cp[c + 1]; q++) {
            int32_t lit = p.lits[(size_t)q];
            if (lit == 0 || lit > n || lit < -n) return DNX_E_ARG;
            deg[(size_t)std::abs(lit) - 1]++;
        }
        if (k == 1) {  // dynex.cc:2458-2469: the last unit clause on v wins
            int32_t lit = p.lits[(size_t)p.cp[c]];
            p.unit[(size_t)std::abs(lit) - 1] = lit > 0 ? 1 : -1;
            p.has_unit = true;
        }
    }
    p.row_ptr.assign((size_t)n + 1, 0);
    p.max_deg = 0;
    for (int v = 0; v < n; v++) {
        p.row_ptr[(size_t)v + 1] = p.row_ptr[(size_t)v] + deg[(size_t)v];
        if (deg[(size_t)v] > p.max_deg) p.max_deg = deg[(size_t)v];
    }
    // counting sort of literal positions by variable; ascending p inside a
    // variable, i.e. clause order then slot order
    std::vector<int32_t> fill(p.row_ptr.begin(), p.row_ptr.end() - 1);
    p.lpos.assign((size_t)p.nnz(), 0);
    for (int64_t q = 0; q < p.nnz(); q++) {
        int v = std::abs(p.lits[(size_t)q]) - 1;
        p.lpos[(size_t)q] = fill[(size_t)v]++;
    }
    p.eidx.assign((size_t)p.nnz(), 0);
    for (int64_t q = 0; q < p.nnz(); q++) p.eidx[(size_t)p.lpos[(size_t)q]] = (int32_t)q;
    return DNX_OK;
}

// Reads one whitespace-delimited token; returns false at EOF.
static bool next_token(FILE *f, std::string &tok)
{
    tok.clear();
    int ch;
    while ((ch = fgetc(f)) != EOF && (ch == ' ' || ch == '\t' || ch == '\n' ||
                                      ch == '\r' || ch == '\v' || ch == '\f')) {
    }
    if (ch == EOF) return false;
    do {
        tok.push_back((char)ch);
    } while ((ch = fgetc(f)) != EOF && !(ch == ' ' || ch == '\t' || ch == '\n' ||
                                         ch == '\r' || ch == '\v' || ch == '\f'));
    if (ch != EOF) ungetc(ch, f);
    return true;
}

int parse_dimacs(const char *path, HostProblem &p, std::string &err)
{
    FILE *f = fopen(path, "r");
    if (!f) {
        err = std::string("cannot open ") + path + ": " + strerror(errno);
        return DNX_E_IO;
    }
    std::string tok;
    // dynex.cc:2314-2351: skip tokens (comments, key=/plain=/cipher= tags)
    // up to the "p" token; the tags only drive display and are ignored.
    for (;;) {
        if (!next_token(f, tok)) {
            fclose(f);
            err = "no 'p cnf' header";
            return DNX_E_PARSE;
        }
        if (tok == "p") break;
    }
    int n = 0, m = 0;
    if (fscanf(f, " cnf %i %i", &n, &m) != 2 || n < 0 || m < 0) {  // dynex.cc:2352
        fclose(f);
        err = "malformed 'p cnf n m' header";
        return DNX_E_PARSE;
    }
    p = HostProblem();
    p.n = n;
    p.m = m;
    p.cp.reserve((size_t)m + 1);
    p.lits.reserve((size_t)m * 3);
    p.cp.push_back(0);
    // dynex.cc:2369-2392: tokens until "0"; "c" skips the rest of its line.
    for (int c = 0; c < m; c++) {
        int k = 0;
        for (;;) {
            if (!next_token(f, tok)) {
                fclose(f);
                err = "unexpected end of file in clause " + std::to_string(c);
                return DNX_E_PARSE;
            }
            if (tok == "c") {
                int ch;
                while ((ch = fgetc(f)) != EOF && ch != '\n') {
                }
                continue;
            }
            long lit = strtol(tok.c_str(), nullptr, 10);
            if (lit != 0) {
                if (lit > n || lit < -(long)n) {
                    fclose(f);
                    err = "literal " + tok + " out of range in clause " + std::to_string(c);
                    return DNX_E_PARSE;
                }
                if (k >= DNX_MAX_LITS) {
                    fclose(f);
                    err = "clause " + std::to_string(c) + " has more than " +
                          std::to_string(DNX_MAX_LITS) + " literals";
                    return DNX_E_TOOLONG;
                }
                p.lits.push_back((int32_t)lit);
                k++;
            }
            if (tok == "0") break;
        }
        p.cp.push_back((int32_t)p.lits.size());
    }
    fclose(f);
    int rc = finalize_problem(p);
    if (rc != DNX_OK) err = "invalid clause table";
    return rc;
}

int write_dimacs(const char *path, int n, int m, const int32_t *cp,
                 const int32_t *lits, const char *comment)
{
    FILE *f = fopen(path, "w");
    if (!f) return DNX_E_IO;
    if (comment && comment[0]) fprintf(f, "c %s\n", comment);
    fprintf(f, "p cnf %d %d\n", n, m);
    for (int c = 0; c < m; c++) {
        for (int q = cp[c]; q < cp[c + 1]; q++) fprintf(f, "%d ", lits[q]);
        fprintf(f, "0\n");
    }
    return fclose(f) == 0 ? DNX_OK : DNX_E_IO;
}

// Thread `tid` of a `-w` run (dynex.cc:2643-2700): threads 0, 1, 2 of seed 0
// set every entry to 1, 0, -1; every other thread draws its V from one
// mt19937(seed) stream shared in thread order, with Xs = 0 and Xl = 1.
void init_reference_thread(const HostProblem &p, int seed, int tid, double *x)
{
    const size_t n = (size_t)p.n, m = (size_t)p.m, N = n + 2 * m;
    std::mt19937 gen((std::mt19937::result_type)seed);
    std::uniform_int_distribution<int> rand_v(-1, 1);
    for (int i = 0; i <= tid; i++) {
        if (i <= 2 && seed == 0) {
            if (i == tid) {
                const double c = i == 0 ? 1.0 : (i == 1 ? 0.0 : -1.0);
                for (size_t j = 0; j < N; j++) x[j] = c;
            }
        } else {
            for (size_t j = 0; j < n; j++) {
                const double v = (double)rand_v(gen);
                if (i == tid) x[j] = v;
            }
            if (i == tid) {
                for (size_t j = n; j < n + m; j++) x[j] = 0.0;
                for (size_t j = n + m; j < N; j++) x[j] = 1.0;
            }
        }
    }
    if (seed == -1)
        for (size_t j = 0; j < n; j++) x[j] = 0.0;
    if (seed == -2)
        for (size_t j = 0; j < n; j++) x[j] = -1.0;
    for (size_t j = 0; j < n; j++)
        if (p.unit[j] != 0) x[j] = (double)p.unit[j];
}

void init_reference(const HostProblem &p, int seed, double *x)
{
    const size_t n = (size_t)p.n, m = (size_t)p.m, N = n + 2 * m;
    if (seed == 0) {  // dynex.cc:2661-2663, thread 0
        for (size_t j = 0; j < N; j++) x[j] = 1.0;
    } else {          // dynex.cc:2665-2670
        std::mt19937 gen((std::mt19937::result_type)seed);
        std::uniform_int_distribution<int> rand_v(-1, 1);
        for (size_t j = 0; j < n; j++) x[j] = (double)rand_v(gen);
        for (size_t j = n; j < n + m; j++) x[j] = 0.0;
        for (size_t j = n + m; j < N; j++) x[j] = 1.0;
    }
    if (seed == -1)
        for (size_t j = 0; j < n; j++) x[j] = 0.0;  // dynex.cc:2673-2675
    if (seed == -2)
        for (size_t j = 0; j < n; j++) x[j] = -1.0;  // dynex.cc:2677-2679
    for (size_t j = 0; j < n; j++)  // dynex.cc:2694-2699
        if (p.unit[j] != 0) x[j] = (double)p.unit[j];
}

// The planted generator.  Per clause: three distinct variables by rejection,
// t from one uniform draw, then which slots are true: t=1 picks the true slot,
// t=2 picks the false slot, t=3 makes all true.  A literal is true under
// sigma when its sign equals sigma of its variable.
void generate_planted(uint64_t seed, int n, int m, double p1, double p2,
                      int32_t *lits, int8_t *planted)
{
    SplitMix64 rng(seed);
    std::vector<int8_t> sigma((size_t)n);
    for (int v = 0; v < n; v++) sigma[(size_t)v] = (rng.next() >> 63) ? 1 : -1;
    for (int c = 0; c < m; c++) {
        uint32_t a = rng.below((uint32_t)n), b, d;
        do {
            b = rng.below((uint32_t)n);
        } while (b == a);
        do {
            d = rng.below((uint32_t)n);
        } while (d == a || d == b);
        double u = rng.unit();
        int t = (u < p1) ? 1 : (u < p1 + p2) ? 2 : 3;
        uint32_t s = rng.below(3);
        uint32_t vars[3] = {a, b, d};
        for (int j = 0; j < 3; j++) {
            bool tru = (t == 3) || (t == 1 && (uint32_t)j == s) || (t == 2 && (uint32_t)j != s);
            int sg = sigma[vars[j]] * (tru ? 1 : -1);
            lits[(size_t)c * 3 + (size_t)j] = sg * (int32_t)(vars[j] + 1);
        }
    }
    if (planted)
        for (int v = 0; v < n; v++) planted[v] = sigma[(size_t)v];
}

void init_uniform(uint64_t seed, int n, int m, double xs0, double xl0, double *x)
{
    SplitMix64 rng(seed ^ 0x5bd1e9955bd1e995ULL);
    for (int v = 0; v < n; v++) x[v] = 2.0 * rng.unit() - 1.0;
    for (int c = 0; c < m; c++) x[(size_t)n + c] = xs0;
    for (int c = 0; c < m; c++) x[(size_t)n + m + c] = xl0;
}

int count_unsat(const HostProblem &p, const int8_t *assign)
{
    int bad = 0;
    for (int c = 0; c < p.m; c++) {
        bool sat = false;
        for (int q = p.cp[c]; q < p.cp[c + 1]; q++) {
            int32_t lit = p.lits[(size_t)q];
            int8_t a = assign[std::abs(lit) - 1];
            if ((lit > 0 && a > 0) || (lit < 0 && a < 0)) {
                sat = true;
                break;
            }
        }
        if (!sat) bad++;
    }
    return bad;
}

}  // namespace dnx
