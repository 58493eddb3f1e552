// host_problem.cpp -- DIMACS I/O, CSR transpose, initial states and the
// planted-instance generator.
#include <algorithm>
#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "host.hpp"
#include "kernels.hpp"

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

bool grouped_tables(const HostProblem &p, int vb, std::vector<uint32_t> &gw, std::vector<int32_t> &gbase,
                    std::vector<int32_t> &geidx)
{
    if (!p.uniform3 || p.n > GX_MAX_VARS || vb <= 0) return false;
    const int64_t nnz = p.nnz();
    const int64_t nchunk = ((int64_t)p.m + GX_CHUNK - 1) / GX_CHUNK;
    auto group = [&](int64_t q) { return (int)(((std::abs(p.lits[(size_t)q]) - 1) / vb) & (GX_GROUPS - 1)); };
    int64_t cur[GX_GROUPS] = {0};
    for (int64_t q = 0; q < nnz; q++) cur[group(q)]++;
    for (int g = GX_GROUPS - 1, s = (int)nnz; g >= 0; g--) {  // group starts
        s -= (int)cur[g];
        cur[g] = s;
    }
    gw.assign((size_t)nnz, 0u);
    gbase.assign((size_t)(nchunk * GX_GROUPS), 0);
    std::vector<int32_t> epos((size_t)nnz);
    for (int64_t k = 0; k < nchunk; k++) {
        int cnt[GX_GROUPS] = {0};
        for (int g = 0; g < GX_GROUPS; g++) gbase[(size_t)(k * GX_GROUPS + g)] = (int32_t)cur[g];
        const int64_t q1 = std::min<int64_t>(nnz, (k + 1) * 3 * GX_CHUNK);
        for (int64_t q = k * 3 * GX_CHUNK; q < q1; q++) {
            const int32_t l = p.lits[(size_t)q];
            const int g = group(q), off = cnt[g]++;
            gw[(size_t)q] = (l < 0 ? 0x80000000u : 0u) | ((uint32_t)(std::abs(l) - 1) << 11) | ((uint32_t)g << 8) |
                            (uint32_t)off;
            epos[(size_t)q] = (int32_t)(cur[g] + off);
        }
        for (int g = 0; g < GX_GROUPS; g++) cur[g] += cnt[g];
    }
    geidx.resize((size_t)nnz);
    for (int64_t s = 0; s < nnz; s++) geidx[(size_t)s] = epos[(size_t)p.eidx[(size_t)s]];
    return true;
}

bool sliced_tables(const HostProblem &p, int S, int SUB, int F, int cus, int max_rounds, int region_cap,
                   SlicedPlan &out)
{
    if (!p.uniform3 || p.m <= 0 || S <= 0 || SUB <= 0 || SUB > 1365 || F <= 0 || cus <= 0) return false;
    const int n = p.n, m = p.m;
    const int NS = (n + S - 1) / S;
    // fold blocks: every block's slots must fit the fold's LDS and q's 16 bits
    int max_region = 0;
    for (int v0 = 0; v0 < n; v0 += F) {
        const int r = p.row_ptr[(size_t)std::min(n, v0 + F)] - p.row_ptr[(size_t)v0];
        max_region = std::max(max_region, r);
    }
    // the padded regions' sink slot (stride) must fit q's 16 bits
    if (max_region > region_cap || max_region > 65528) return false;
    // the distinct slices of each clause, as a sorted triple (-1 padded)
    struct Key {
        int s[3];
        bool operator<(const Key &o) const
        {
            return s[0] != o.s[0] ? s[0] < o.s[0] : s[1] != o.s[1] ? s[1] < o.s[1] : s[2] < o.s[2];
        }
        bool operator==(const Key &o) const { return s[0] == o.s[0] && s[1] == o.s[1] && s[2] == o.s[2]; }
    };
    // each variable's slice and offset in it (no divisions per literal)
    std::vector<int32_t> vsl((size_t)n), voff((size_t)n);
    for (int v = 0; v < n; v++) {
        vsl[(size_t)v] = v / S;
        voff[(size_t)v] = v % S;
    }
    std::vector<Key> key((size_t)m);
    for (int c = 0; c < m; c++) {
        int a[3];
        for (int j = 0; j < 3; j++) a[j] = vsl[(size_t)(std::abs(p.lits[(size_t)3 * c + j]) - 1)];
        std::sort(a, a + 3);
        Key k{{a[0], -1, -1}};
        int ns = 1;
        for (int j = 1; j < 3; j++)
            if (a[j] != k.s[ns - 1]) k.s[ns++] = a[j];
        key[(size_t)c] = k;
    }
    // clause sets by slice set, each in clause order; largest first.  The
    // clauses are ordered by key, stably: a counting sort over the key's
    // lexicographic rank when the key space is small, else a stable sort
    std::vector<int> ord((size_t)m);
    const int64_t R = (int64_t)NS + 1, space = R * R * R;
    if (space <= ((int64_t)1 << 22)) {
        auto rank = [&](const Key &k) { return ((int64_t)(k.s[0] + 1) * R + (k.s[1] + 1)) * R + (k.s[2] + 1); };
        std::vector<int> start((size_t)space + 1, 0);
        for (int c = 0; c < m; c++) start[(size_t)rank(key[(size_t)c]) + 1]++;
        for (int64_t r = 0; r < space; r++) start[(size_t)r + 1] += start[(size_t)r];
        for (int c = 0; c < m; c++) ord[(size_t)start[(size_t)rank(key[(size_t)c])]++] = c;
    } else {
        for (int c = 0; c < m; c++) ord[(size_t)c] = c;
        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return key[(size_t)x] < key[(size_t)y]; });
    }
    struct Set {
        Key k;
        int begin, size;
    };
    std::vector<Set> sets;
    for (int i = 0; i < m;) {
        int j = i;
        while (j < m && key[(size_t)ord[(size_t)j]] == key[(size_t)ord[(size_t)i]]) j++;
        sets.push_back({key[(size_t)ord[(size_t)i]], i, j - i});
        i = j;
    }
    // three-slice sets first (only a unit of exactly their slices can take
    // them), then the two- and one-slice sets, which fit many units' room;
    // each class largest first (n = 100k, m = 800k: the largest unit 3,173
    // clauses instead of 3,700 when all sets went largest first)
    auto nsl_of = [](const Set &x) { return (x.k.s[1] >= 0) + (x.k.s[2] >= 0); };
    std::stable_sort(sets.begin(), sets.end(), [&](const Set &x, const Set &y) {
        return nsl_of(x) != nsl_of(y) ? nsl_of(x) > nsl_of(y) : x.size > y.size;
    });
    // units of at most T clauses whose slice union has at most three
    // slices: each set first fills the free room of units it is compatible
    // with, then opens new ones; T grows until at most `cus` units result
    struct Unit {
        int sl[3];
        int ns = 0, size = 0;
        std::vector<std::pair<int, int>> parts;  // (set, count)
    };
    auto merge = [](const Unit &u, const Key &k, int *out3) {
        int nsl = u.ns;
        for (int j = 0; j < u.ns; j++) out3[j] = u.sl[j];
        for (int j = 0; j < 3 && k.s[j] >= 0; j++) {
            bool have = false;
            for (int i = 0; i < nsl; i++) have |= out3[i] == k.s[j];
            if (!have) {
                if (nsl == 3) return -1;
                out3[nsl++] = k.s[j];
            }
        }
        return nsl;
    };
    auto form = [&](int T, std::vector<Unit> &U) {
        U.clear();
        for (int si = 0; si < (int)sets.size(); si++) {
            int rem = sets[(size_t)si].size;
            for (auto &u : U) {
                if (rem == 0) break;
                if (u.size >= T) continue;
                int sl[3];
                const int ns = merge(u, sets[(size_t)si].k, sl);
                if (ns < 0) continue;
                const int take = std::min(rem, T - u.size);
                for (int j = 0; j < ns; j++) u.sl[j] = sl[j];
                u.ns = ns;
                u.size += take;
                u.parts.push_back({si, take});
                rem -= take;
            }
            while (rem > 0) {
                Unit u;
                u.ns = merge(u, sets[(size_t)si].k, u.sl);
                const int take = std::min(rem, T);
                u.size = take;
                u.parts.push_back({si, take});
                rem -= take;
                U.push_back(std::move(u));
            }
        }
    };
    // every three-slice set needs a unit of its own: with more of them than
    // max_rounds rounds of units, staging V would cost more than it saves
    int triples = 0;
    for (const Set &st : sets) triples += st.k.s[2] >= 0;
    if (triples > max_rounds * cus) return false;
    // as few rounds of units as the three-slice sets allow
    const int target = std::max(1, (triples + cus - 1) / cus) * cus;
    std::vector<Unit> U;
    int T = std::max((m + target - 1) / target, std::min(m, SUB / 4));
    for (int it = 0; it < 400; it++) {
        form(T, U);
        if ((int)U.size() <= target || T >= m) break;
        T = std::min(m, T + std::max(1, T / 64));
    }
    if ((int)U.size() > max_rounds * cus) return false;
    for (const Unit &u : U)
        if ((u.size + SUB - 1) / SUB > 63) return false;  // the clause pass keeps a unit's sub-batch table in LDS
    // processing order, units and sub-batches
    out = SlicedPlan();
    out.S = S;
    out.SUB = SUB;
    out.F = F;
    out.max_region = max_region;
    out.stride = (max_region + 3) / 4 * 4;
    const int nblk = (n + F - 1) / F;
    out.e_len = (long long)nblk * out.stride;
    if (out.e_len >= INT_MAX) return false;  // e positions are 32-bit in the run table
    std::vector<int> cursor(sets.size(), 0);
    std::vector<int> porder;
    porder.reserve((size_t)m);
    int nsb = 0;
    for (const Unit &u : U) {
        const int c0 = (int)porder.size();
        for (auto &pt : u.parts) {
            const Set &st = sets[(size_t)pt.first];
            for (int k = 0; k < pt.second; k++)
                porder.push_back(ord[(size_t)st.begin + cursor[(size_t)pt.first] + k]);
            cursor[(size_t)pt.first] += pt.second;
        }
        const int c1 = (int)porder.size();
        int sl[3] = {u.sl[0], u.ns > 1 ? u.sl[1] : -1, u.ns > 2 ? u.sl[2] : -1};
        out.units.insert(out.units.end(), {c0, c1, u.ns, sl[0], sl[1], sl[2], nsb, 0});  // [7]: see below
        nsb += sx_unit_batches(c1 - c0, SUB);
    }
    out.pidx.assign((size_t)m, 0);
    for (int i = 0; i < m; i++) out.pidx[(size_t)porder[(size_t)i]] = i;
    // per sub-batch: literals ordered by destination fold block (then by
    // processing position), one run per block
    out.gw.assign((size_t)3 * m, 0u);
    out.q.assign((size_t)out.e_len, (uint16_t)out.stride);
    out.sb_run.clear();
    out.runs.clear();
    std::vector<int64_t> fill((size_t)(n + F - 1) / F + 1);
    for (size_t b = 0; b + 1 < fill.size(); b++) fill[b] = (int64_t)b * out.stride;
    std::vector<int> lit_blk((size_t)3 * SUB), cnt, order3((size_t)3 * SUB);
    std::vector<int64_t> lit_q((size_t)3 * SUB);  // literal position of each staged literal
    for (size_t ui = 0; ui < U.size(); ui++) {
        const int *ud = &out.units[ui * 8];
        const int c0 = ud[0], c1 = ud[1], nsl = ud[2];
        const int sub = sx_unit_sub(c1 - c0, SUB);
        for (int cb = c0; cb < c1; cb += sub) {
            const int ce = std::min(c1, cb + sub), cntl = 3 * (ce - cb);
            if (cb == c0) out.units[ui * 8 + 7] = (int)out.runs.size() / 2;  // the unit's first run
            out.sb_run.push_back((int)out.runs.size() / 2);
            // stable counting sort of the sub-batch's literals by block
            int bmin = INT_MAX, bmax = -1;
            for (int i = 0; i < cntl; i++) {
                const int c = porder[(size_t)cb + i / 3];
                const int64_t lq = (int64_t)3 * c + i % 3;
                const int v = std::abs(p.lits[(size_t)lq]) - 1;
                lit_q[(size_t)i] = lq;
                lit_blk[(size_t)i] = v / F;
                bmin = std::min(bmin, v / F);
                bmax = std::max(bmax, v / F);
            }
            cnt.assign((size_t)(bmax - bmin + 2), 0);
            for (int i = 0; i < cntl; i++) cnt[(size_t)(lit_blk[(size_t)i] - bmin + 1)]++;
            std::vector<int> start(cnt.size());
            for (size_t b = 1; b < cnt.size(); b++) start[b] = start[b - 1] + cnt[b];
            std::vector<int> pos(start.begin(), start.end());
            for (int i = 0; i < cntl; i++) order3[(size_t)(pos[(size_t)(lit_blk[(size_t)i] - bmin)]++)] = i;
            // runs: (staging start, length) and their e positions; a
            // block's run is cut into pieces of at most 128 (a literal's word
            // holds its piece and its offset in the piece)
            for (int b = bmin; b <= bmax; b++) {
                const int bls = start[(size_t)(b - bmin)], blen = cnt[(size_t)(b - bmin + 1)];
                for (int p0 = 0; p0 < blen; p0 += 128) {
                const int ls = bls + p0, len = std::min(128, blen - p0);
                const int rid = (int)out.runs.size() / 2 - out.sb_run.back();
                if (rid >= SX_MAXRUNS) return false;
                out.runs.push_back((ls << 16) | len);
                out.runs.push_back((int32_t)fill[(size_t)b]);
                const int rb0 = p.row_ptr[(size_t)b * F];
                for (int k = 0; k < len; k++) {
                    const int i = order3[(size_t)ls + k];
                    const int64_t lq = lit_q[(size_t)i];  // literal position
                    out.q[(size_t)fill[(size_t)b] + k] = (uint16_t)(p.lpos[(size_t)lq] - rb0);
                    // the literal's word: sign, LDS index of its variable
                    // (slice index within the unit, offset in the slice),
                    // staging slot
                    const int v = std::abs(p.lits[(size_t)lq]) - 1;
                    int si = 0;
                    while (si < nsl && ud[3 + si] != vsl[(size_t)v]) si++;
                    const uint32_t lidx = (uint32_t)(si * S + voff[(size_t)v]);
                    out.gw[(size_t)3 * (cb + i / 3) + i % 3] = (p.lits[(size_t)lq] < 0 ? 0x80000000u : 0u) |
                                                             (lidx << 15) | ((uint32_t)rid << 7) | (uint32_t)k;
                }
                fill[(size_t)b] += len;
                }
            }
        }
    }
    out.sb_run.push_back((int)out.runs.size() / 2);
    // padding: the clause pass also reads the tables one sub-batch past a
    // unit's last and a whole workgroup's worth of runs (and uses neither)
    out.sb_run.push_back(out.sb_run.back());
    out.runs.insert(out.runs.end(), (size_t)2 * 256, 0);
    (void)NS;
    return true;
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
