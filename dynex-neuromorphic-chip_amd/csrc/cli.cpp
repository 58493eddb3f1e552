// cli.cpp -- `dynex`: the drop-in command line of the reference
// (dynex.cc:2204-2296 flags, 662-700, 1494-1521 and 2864 output) over libdnx.
//
//   dynex -i file.cnf [-l seed] [-c alpha] [-b beta] [-n gamma] [-h delta]
//         [-j epsilon] [-k zeta] [-m xl_max] [-z maxsteps] [-u walltime]
//         [-q 0|1] [-w threads] [-P 64|32] [-G device]
//
// Flags take "-x value" or "-x=value" (dynex.cc:2208-2233).  -w runs the
// reference's thread portfolio on the device (dnx_portfolio_solve).  -P
// (precision) and -G (device) are additions; the reference's tuning,
// integrator-mode and partable flags (-t -g -r -v -s -o -e -p -x -y -d -a -f)
// are accepted and ignored: only the default adaptive Euler mode is on this
// path.  That mode has no step cap (dynex.cc:1387 loops until solved; the
// reference's maxsteps, dynex.cc:154, only feeds its odeint modes), so -z is
// unbounded by default and, when given, caps the adaptive loop (an addition).
//
// Output follows the reference line for line:
//   solved:  "c [i] T=<t> SOLUTION FOUND", the signed literals of the
//            verified state (x > 0 positive, x < 0 negative, 0 omitted),
//            "c [i] VERIFYING...", "c [i] SAT (VERIFIED)" (dynex.cc:662-690),
//            then TIME SPENT, BEST LOC (dynex.cc:1499-1504) and the
//            "s [i] SATISFIABLE" / "v ..." lines (dynex.cc:1507-1514).  The
//            reference prints the first four even under -q 1 and clears
//            quiet after them (dynex.cc:663, 716), so the rest follows too.
//   -z cap:  TIME SPENT, BEST LOC and "s [i] UNKNOWN" (dynex.cc:1519-1521).
//   -u:      nothing for the thread: the reference's RHS ends the thread when
//            the wall clock passes the limit (dynex.cc:335-345).
// and always ends with `global`, the lowest loc reached (dynex.cc:2864).
#include <getopt.h>

#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "dnx.h"

static const char *arg(char *a)
{
    if (a && a[0] == '=') return a + 1;  // dynex.cc:2208: "-x=value"
    return a;
}

static void usage(const char *prog)
{
    printf("c Syntax: %s -i file.cnf [-l seed] [-c alpha] [-b beta] [-n gamma] [-h delta]\n"
           "c            [-j epsilon] [-k zeta] [-m xl_max] [-z maxsteps] [-u walltime]\n"
           "c            [-q 0|1] [-w threads] [-P 64|32] [-G device]\n",
           prog);
}

int main(int argc, char **argv)
{
    std::string input;
    int seed = 0, xlm = 10000, quiet = 0, prec_bits = 64, device = 0, threads = 1;
    double alpha = 2.5, beta = 20.0, gamma = 25.0 / 100.0, delta = 5.0 / 100.0, eps = 1.0 / 10.0,
           zeta = 1.0 / 10.0, walltime = 0.0, maxsteps = -1.0;
    int c;
    while ((c = getopt(argc, argv, "r:v:w:u:q:p:b:c:n:h:j:k:l:m:e:f:g:s:a:z:x:y:d:o:t:i:P:G:")) != -1) {
        const char *a = arg(optarg);
        switch (c) {
            case 'i': input = a; break;
            case 'l': seed = atoi(a); break;
            case 'c': alpha = atof(a); break;
            case 'b': beta = atof(a); break;
            case 'n': gamma = atof(a); break;
            case 'h': delta = atof(a); break;
            case 'j': eps = atof(a); break;
            case 'k': zeta = atof(a); break;
            case 'm': xlm = atoi(a); break;
            case 'z': maxsteps = atof(a); break;
            case 'u': walltime = atof(a); break;
            case 'q': quiet = atoi(a); break;
            case 'P': prec_bits = atoi(a); break;
            case 'G': device = atoi(a); break;
            case 'w': threads = atoi(a); break;
            case 'r': case 'v': case 'p': case 'e': case 'f': case 'g':
            case 's': case 'a': case 'x': case 'y': case 'd': case 'o': case 't':
                break;
            default: usage(argv[0]); return EXIT_FAILURE;
        }
    }
    if (input.empty()) {
        usage(argv[0]);
        return EXIT_FAILURE;
    }
    const int prec = prec_bits == 32 ? DNX_FP32 : DNX_FP64;
    if (!quiet) {
        printf("c --------------------------------------------\n");
        printf("c MEMCOMPUTING SAT INTEGRATOR (MI355X)        \n");
        printf("c --------------------------------------------\n");
        printf("c INSTANCE  : %s\n", input.c_str());
    }
    dnx_ctx *ctx = nullptr;
    if (dnx_open(device, &ctx) != DNX_OK) {
        fprintf(stderr, "c ERROR: %s\n", dnx_last_error());
        return EXIT_FAILURE;
    }
    dnx_problem *prob = nullptr;
    int rc = dnx_problem_from_dimacs(ctx, input.c_str(), &prob);
    if (rc != DNX_OK) {
        printf("c ERROR: %s\n", dnx_last_error());
        dnx_close(ctx);
        return EXIT_FAILURE;
    }
    int n = 0, m = 0;
    int64_t nnz = 0;
    dnx_problem_info(prob, &n, &m, &nnz);
    if (!quiet) {
        printf("c VARIABLES : %d\n", n);
        printf("c CLAUSES   : %d\n", m);
        printf("c RATIO     : %lf\n", m > 0 && n > 0 ? (double)m / n : 0.0);
        printf("c PRECISION : FP%d\n", prec_bits == 32 ? 32 : 64);
    }
    dnx_params prm;
    dnx_params_default(m, xlm, &prm);
    prm.alpha = alpha;
    prm.beta = beta;
    prm.gamma = gamma;
    prm.delta = delta;
    prm.epsilon = eps;
    prm.zeta = zeta;
    std::vector<double> x0((size_t)n + 2 * (size_t)m);
    dnx_init_reference(prob, seed, x0.data());
    std::vector<int8_t> assign((size_t)n);
    // no cap unless -z is given (mode 6 has none); a value beyond int64 is no cap
    const int64_t cap = maxsteps < 0 || maxsteps >= 9.2e18 ? INT64_MAX : (int64_t)maxsteps;
    int64_t steps = 0;
    double t = 0.0;
    int status = DNX_UNKNOWN;
    auto t0 = std::chrono::steady_clock::now();
    int winner = 0, best_loc = m;
    std::vector<double> xv;  // the verified state, for the literal listing
    if (threads > 1) {
        rc = dnx_portfolio_solve(prob, prec, &prm, seed, threads, cap, walltime, assign.data(), &steps, &t, &status,
                                 &winner);
        if (winner < 0) winner = 0;
        // the copies' loc is not tracked: their lowest is known only when one verifies
        if (rc == DNX_OK && status == DNX_SAT) best_loc = 0;
    } else {
        dnx_state *st = nullptr;
        rc = dnx_state_create(prob, prec, &prm, x0.data(), &st);
        if (rc == DNX_OK) rc = dnx_state_track_loc(st, 1);
        if (rc == DNX_OK) rc = dnx_state_solve(st, cap, walltime, assign.data(), &steps, &t, &status);
        if (rc == DNX_OK) rc = dnx_state_best_loc(st, &best_loc, nullptr);
        if (rc == DNX_OK && status == DNX_SAT) {
            xv.resize((size_t)n + 2 * (size_t)m);
            rc = dnx_state_read(st, xv.data(), nullptr, nullptr, nullptr, nullptr);
        }
        if (st) dnx_state_free(st);
    }
    double spent = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc != DNX_OK) {
        printf("c ERROR: %s\n", dnx_last_error());
        dnx_problem_free(prob);
        dnx_close(ctx);
        return EXIT_FAILURE;
    }
    const int global = status == DNX_SAT ? 0 : best_loc;
    const bool walltime_end = status != DNX_SAT && steps < cap;
    if (status == DNX_SAT) {
        // dynex.cc:662-690: printed whatever -q says, which it then clears
        printf("\nc [%d] T=%.2f SOLUTION FOUND (%lld)\n", winner, t, (long long)steps);
        for (int i = 0; i < n; i++) {
            const double v = xv.empty() ? (double)assign[(size_t)i] : xv[(size_t)i];
            if (v > 0) printf("%d ", i + 1);
            if (v < 0) printf("%d ", -(i + 1));
        }
        printf("\nc [%d] VERIFYING...\n", winner);
        printf("c [%d] SAT (VERIFIED)\n", winner);
        quiet = 0;
        std::string sol = input + ".solution.txt";  // dynex.cc:2477-2480, 693-699
        FILE *fs = fopen(sol.c_str(), "w");
        if (fs) {
            for (int i = 0; i < n; i++) fprintf(fs, "%d, ", assign[(size_t)i] > 0 ? i + 1 : -(i + 1));
            fclose(fs);
        }
    }
    if (!walltime_end && !quiet) {
        printf("c [%d] TIME SPENT: %.5fs\n", winner, spent);  // dynex.cc:1499
        // dynex.cc:1500-1504: the thread's lowest loc against the global
        // one; with one thread (or the winner of a portfolio) they agree
        printf("c [%d] BEST LOC = %d (GLOBAL %d) **\n", winner, global, global);
    }
    if (status == DNX_SAT) {
        if (!quiet) printf("\ns [%d] SATISFIABLE", winner);  // dynex.cc:1509-1514
        for (int i = 0; i < n; i++) {
            if (i % 20 == 0)
                if (!quiet) printf("\nv ");
            if (!quiet) printf("%i ", assign[(size_t)i] > 0 ? i + 1 : -(i + 1));
        }
        if (!quiet) printf("0\n");
    } else if (!walltime_end && !quiet) {
        printf("\ns [%d] UNKNOWN\n", winner);  // dynex.cc:1520
    }
    printf("%d", global);  // dynex.cc:2864
    fflush(stdout);
    dnx_problem_free(prob);
    dnx_close(ctx);
    return 0;
}
