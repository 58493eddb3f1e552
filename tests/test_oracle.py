"""Pins the CPU oracle (oracle/dnx_oracle.c) to facts that do not come from
the oracle itself: hand-derived closed forms of the clause equations
(dynex.cc:362-506), an independent left fold of the per-variable
contributions (dynex.cc:527-530), the Euler clamps (dynex.cc:301-327), the
five outcomes of the adaptive ladder (dynex.cc:1387-1417), an independent
integer clause check, brute-force satisfiability on tiny formulas, the
reference's DIMACS fixtures, the C++ standard's mt19937 known answer, and
libstdc++'s own uniform_int_distribution compiled with g++."""
import functools
import itertools
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import CNF_NAMES, cnf_path, planted_cp, random_state

REL = 1e-12


def one_clause(lits, n=None):
    lits = list(lits)
    n = n or max(abs(l) for l in lits)
    return O.Problem(n, 1, np.array([0, len(lits)], np.int32), np.array(lits, np.int32))


def state(v, xs, xl):
    return np.array(list(v) + list(np.atleast_1d(xs)) + list(np.atleast_1d(xl)), np.float64)


def close(a, b):
    return abs(a - b) <= REL * max(1.0, abs(b))


# ---------------------------------------------------------------------------
# clause equations: closed forms (dynex.cc:371-418, 420-454, 457-494, 503-506)
# ---------------------------------------------------------------------------
def test_clause_all_satisfied_corner():
    # V = +1 on every positive literal: i = 0, C = 0, G = 0, R = (1-1)/2 = 0
    p = one_clause([1, 2, 3])
    prm = O.params(1)
    d, loc, en = O.rhs(p, prm, state([1, 1, 1], 1.0, 1.0))
    assert list(d[:3]) == [0.0, 0.0, 0.0]
    assert loc == 0 and en == 0.0
    assert close(d[3], 20 * 1.1 * -0.25)  # beta (Xs+eps)(C-gamma)
    assert close(d[4], 2.5 * -0.05)       # alpha (C-delta)


def test_clause_all_false_corner_ties():
    # every literal false at a corner: i = 2 each, C = 1, all tie -> all get R
    p = one_clause([1, -2, 3])
    d, loc, _ = O.rhs(p, O.params(1), state([-1, 1, -1], 0.5, 2.0))
    # G = q*2/2 = q; R = (q - V)/2 = q; contrib = Xl Xs q + (1+0.1*2)(0.5) q = 1.6 q
    assert close(d[0], 1.6) and close(d[1], -1.6) and close(d[2], 1.6)
    assert loc == 1
    assert close(d[3], 20 * 0.6 * 0.75) and close(d[4], 2.5 * 0.95)


def test_clause_distinct_minimum():
    p = one_clause([1, 2, 3])
    d, loc, en = O.rhs(p, O.params(1), state([0.5, 0.0, -0.5], 0.5, 1.0))
    # i = (0.5, 1, 1.5): C = 0.25, G = (0.5, 0.25, 0.25), R = (0.25, 0, 0)
    # contrib = 0.5 G + 1.1*0.5 R
    assert close(d[0], 0.5 * 0.5 + 0.55 * 0.25)
    assert close(d[1], 0.125) and close(d[2], 0.125)
    assert loc == 0 and close(en, 0.25)


def test_clause_two_way_tie():
    p = one_clause([1, 2, 3])
    d, _, _ = O.rhs(p, O.params(1), state([0.5, 0.5, -0.5], 0.5, 1.0))
    # i = (0.5, 0.5, 1.5): both minima get R = 0.25; G = 0.25 for all
    assert close(d[0], 0.2625) and close(d[1], 0.2625) and close(d[2], 0.125)


def test_clause_k2():
    p = one_clause([1, -2])
    d, _, _ = O.rhs(p, O.params(1), state([0.5, 0.5], 0.5, 1.0))
    # i = (0.5, 1.5): C = .25, G0 = i1/2 = .75, G1 = -i0/2 = -.25, R0 = .25
    assert close(d[0], 0.5 * 0.75 + 0.55 * 0.25) and close(d[1], -0.125)


def test_clause_k5():
    p = one_clause([1, 2, 3, 4, 5])
    d, loc, _ = O.rhs(p, O.params(1), state([0.5, 0.9, 0.1, 0.5, -1.0], 0.5, 1.0))
    # i = (0.5, 0.1, 0.9, 0.5, 2): argmin is literal 2; G = min of the others / 2
    exp = [0.5 * 0.05, 0.5 * 0.25 + 0.55 * 0.05, 0.5 * 0.05, 0.5 * 0.05, 0.5 * 0.05]
    for a, b in zip(d[:5], exp):
        assert abs(a - b) < 1e-15
    assert loc == 0


def test_clause_unit_and_empty():
    # k = 1: no clause math, C = 0 (dynex.cc:517), counted satisfied
    p = O.Problem(2, 2, np.array([0, 1, 1], np.int32), np.array([2], np.int32))
    x = state([0.3, -0.4], [0.5, 0.5], [1.0, 1.0])
    d, loc, _ = O.rhs(p, O.params(2), x)
    assert d[0] == 0.0 and d[1] == 0.0
    assert close(d[2], 20 * 0.6 * -0.25)
    # the empty clause takes the k-SAT branch with min = INT_MAX: C = INT_MAX/2
    assert close(d[5], 2.5 * (2147483647.0 / 2 - 0.05))
    assert loc == 1


def test_clause_loc_matches_integer_recount():
    """loc counts C >= 1/2, i.e. clauses where every literal has q*V <= 0."""
    rng = np.random.default_rng(3)
    n, m = 40, 200
    lits = np.concatenate([rng.choice(np.arange(1, n + 1), 3, replace=False) * rng.choice([-1, 1], 3)
                           for _ in range(m)]).astype(np.int32)
    p = O.Problem(n, m, planted_cp(m), lits)
    for s in range(5):
        x = random_state(rng, n, m)
        x[:n] = np.where(rng.random(n) < 0.3, np.sign(x[:n]), x[:n])
        _, loc, _ = O.rhs(p, O.params(m), x)
        V = np.clip(x[:n], -1, 1)
        q = np.sign(lits).reshape(m, 3)
        vv = V[np.abs(lits) - 1].reshape(m, 3)
        assert loc == int(np.all(q * vv <= 0, axis=1).sum())


# ---------------------------------------------------------------------------
# clause -> variable accumulation (dynex.cc:408-410, 527-530)
# ---------------------------------------------------------------------------
def _per_clause_contribs(p, prm, x, prec):
    """Each clause evaluated alone: its contributions to its variables."""
    out = []
    n, m = p.n, p.m
    for c in range(m):
        cl = p.lits[p.cp[c]:p.cp[c + 1]]
        q = O.Problem(n, 1, np.array([0, len(cl)], np.int32), cl)
        xc = np.concatenate([x[:n], [x[n + c]], [x[n + m + c]]])
        d, _, _ = O.rhs(q, prm, xc, prec)
        out.append({abs(int(l)) - 1: d[abs(int(l)) - 1] for l in cl})
    return out


def test_fold_is_ascending_sorted_left_fold_f64():
    rng = np.random.default_rng(11)
    n, m = 12, 60
    lits = np.concatenate([rng.choice(np.arange(1, n + 1), 3, replace=False) * rng.choice([-1, 1], 3)
                           for _ in range(m)]).astype(np.int32)
    p = O.Problem(n, m, planted_cp(m), lits)
    prm = O.params(m)
    x = random_state(rng, n, m, xl_hi=1e4)
    d, _, _ = O.rhs(p, prm, x)
    per = _per_clause_contribs(p, prm, x, 64)
    order_matters = 0
    for v in range(n):
        vals = [cc[v] for cc in per if v in cc and cc[v] != 0.0]
        srt = functools.reduce(lambda a, b: a + b, sorted(vals), 0.0)
        raw = functools.reduce(lambda a, b: a + b, vals, 0.0)
        assert d[v] == srt
        order_matters += srt != raw
    assert order_matters > 0  # the data can tell sorted from clause order


def test_fold_f32_is_clause_order():
    rng = np.random.default_rng(12)
    n, m = 12, 60
    lits = np.concatenate([rng.choice(np.arange(1, n + 1), 3, replace=False) * rng.choice([-1, 1], 3)
                           for _ in range(m)]).astype(np.int32)
    p = O.Problem(n, m, planted_cp(m), lits)
    prm = O.params(m)
    x = random_state(rng, n, m, xl_hi=1e4).astype(np.float32)
    d, _, _ = O.rhs(p, prm, x, prec=32)
    per = _per_clause_contribs(p, prm, x, 32)
    differs = 0
    for v in range(n):
        vals = [np.float32(cc[v]) for cc in per if v in cc]
        acc = np.float32(0.0)
        for a in vals:
            acc = np.float32(acc + a)
        srt = np.float32(0.0)
        for a in sorted(vals):
            srt = np.float32(srt + a)
        assert d[v] == acc
        differs += srt != acc
    assert differs > 0


# ---------------------------------------------------------------------------
# Euler step and clamps (dynex.cc:301-327)
# ---------------------------------------------------------------------------
def test_dostep_clamps_and_pins():
    # var 3 is pinned to -1 by a unit clause
    cp = np.array([0, 3, 4], np.int32)
    p = O.Problem(3, 2, cp, np.array([1, 2, 3, -3], np.int32))
    prm = O.params(2)
    x = state([0.5, -0.5, 0.7], [0.9, 0.1], [1.5, prm.xl_max - 1.0])
    d = state([10.0, -10.0, 5.0], [2.0, -2.0], [-8.0, 100.0])
    out = O.dostep(p, prm, x, d, 0.125)
    assert list(out) == [1.0, -1.0, -1.0, 1.0, 0.0, 1.0, float(prm.xl_max)]
    out = O.dostep(p, prm, x, d * 0.01, 0.125)
    assert out[0] == 0.5 + 0.125 * 0.1 and out[3] == 0.9 + 0.125 * 0.02


def test_xl_max_scaling():
    # dynex.cc:2358-2359: int product, INT_MAX when not positive (wraps)
    assert O.lib().orc_xl_max(10000, 800) == 8_000_000
    assert O.lib().orc_xl_max(10000, 800_000) == 2**31 - 1
    assert O.lib().orc_xl_max(10000, 430_000) == (10000 * 430_000) % 2**32


# ---------------------------------------------------------------------------
# adaptive ladder (dynex.cc:1387-1417)
# ---------------------------------------------------------------------------
def _ladder_check(p, prm, x0):
    x = x0.copy()
    t, lev, loc, bad, _ = O.step(p, prm, x, 0.0)
    d, loc2, _ = O.rhs(p, prm, x0)
    assert loc == loc2
    n = p.n
    hs = [0.125, 0.0625, 0.03125, 0.015625]

    def fails(h):
        tr = O.dostep(p, prm, x0, d, h)
        return bool(np.any(np.abs(tr[:n] - x0[:n]) >= 1))

    assert 0 <= lev <= 4
    for L in range(min(lev, 4)):
        assert fails(hs[L])  # every larger step was rejected
    if lev < 4:
        assert not fails(hs[lev])
        assert np.array_equal(x, O.dostep(p, prm, x0, d, hs[lev]))
        assert t == hs[lev]
    else:  # fallback: the 1/64 trial is kept, t advances by 1/128
        assert np.array_equal(x, O.dostep(p, prm, x0, d, 0.015625))
        assert t == 0.0078125
    return lev


def test_ladder_levels_cover_all_outcomes():
    rng = np.random.default_rng(5)
    n, m = 30, 150
    lits = np.concatenate([rng.choice(np.arange(1, n + 1), 3, replace=False) * rng.choice([-1, 1], 3)
                           for _ in range(m)]).astype(np.int32)
    p = O.Problem(n, m, planted_cp(m), lits)
    prm = O.params(m)
    seen = set()
    for xl in (1.0, 4.0, 12.0, 30.0, 60.0, 500.0):
        for s in range(4):
            x0 = random_state(rng, n, m)
            x0[n:n + m] = 1.0
            x0[n + m:] = xl * rng.uniform(0.8, 1.2, m)
            seen.add(_ladder_check(p, prm, x0))
    assert seen == {0, 1, 2, 3, 4}


def test_step_invariants_and_time():
    rng = np.random.default_rng(8)
    n, m = 50, 400
    lits = np.concatenate([rng.choice(np.arange(1, n + 1), 3, replace=False) * rng.choice([-1, 1], 3)
                           for _ in range(m)]).astype(np.int32)
    p = O.Problem(n, m, planted_cp(m), lits)
    prm = O.params(m)
    x = random_state(rng, n, m)
    t = 0.0
    for _ in range(200):
        t_new, lev, _, _, _ = O.step(p, prm, x, t)
        dt = t_new - t
        assert dt in (0.125, 0.0625, 0.03125, 0.015625, 0.0078125)
        assert (dt == 0.0078125) == (lev == 4)
        t = t_new
        assert np.all(np.abs(x[:n]) <= 1) and np.all((x[n:n + m] >= 0) & (x[n:n + m] <= 1))
        assert np.all((x[n + m:] >= 1) & (x[n + m:] <= prm.xl_max))


# ---------------------------------------------------------------------------
# exact check and brute force (dynex.cc:662-703)
# ---------------------------------------------------------------------------
def py_unsat(cp, lits, x):
    bad = 0
    for c in range(len(cp) - 1):
        ok = any((x[abs(l) - 1] > 0 and l > 0) or (x[abs(l) - 1] < 0 and l < 0) for l in lits[cp[c]:cp[c + 1]])
        bad += not ok
    return bad


def test_verify_matches_integer_check():
    rng = np.random.default_rng(1)
    n, m = 20, 80
    lits = np.concatenate([rng.choice(np.arange(1, n + 1), 3, replace=False) * rng.choice([-1, 1], 3)
                           for _ in range(m)]).astype(np.int32)
    p = O.Problem(n, m, planted_cp(m), lits)
    for _ in range(20):
        x = np.concatenate([rng.choice([-1.0, -0.2, 0.0, 0.3, 1.0], n), np.zeros(2 * m)])
        assert O.verify(p, x) == py_unsat(p.cp, p.lits, x)


@pytest.mark.parametrize("seed", range(6))
def test_bruteforce_small(seed):
    rng = np.random.default_rng(100 + seed)
    n = 10
    m = int(rng.choice([30, 43, 50]))
    lits = np.concatenate([rng.choice(np.arange(1, n + 1), 3, replace=False) * rng.choice([-1, 1], 3)
                           for _ in range(m)]).astype(np.int32)
    cp = planted_cp(m)
    sat_any = any(py_unsat(cp, lits, np.array(bits) * 2.0 - 1.0) == 0
                  for bits in itertools.product([0, 1], repeat=n))
    p = O.Problem(n, m, cp, lits)
    x = O.init_ref(p, 0)
    steps, t, status = O.solve(p, O.params(m), x, 3000)
    if status == 1:
        assert sat_any
        assert py_unsat(cp, lits, x[:n]) == 0
    else:
        assert status == 2 and steps == 3000


# ---------------------------------------------------------------------------
# DIMACS reader (dynex.cc:2306-2392) against the reference's fixtures
# ---------------------------------------------------------------------------
def py_read_dimacs(path):
    toks, clauses, planted = [], [], None
    with open(path) as f:
        lines = f.read().splitlines()
    if lines and lines[0].startswith("c ["):
        planted = [int(t) for t in lines[0][3:].split() if t not in ("]",)]
    body = [ln for ln in lines if not ln.startswith("c")]
    hdr = [ln for ln in body if ln.startswith("p")][0].split()
    n, m = int(hdr[2]), int(hdr[3])
    for ln in body:
        if ln.startswith("p"):
            continue
        toks += [int(t) for t in ln.split()]
    cur = []
    for t in toks:
        if t == 0:
            clauses.append(cur)
            cur = []
        else:
            cur.append(t)
    return n, m, clauses, planted


@pytest.mark.parametrize("nv", sorted(CNF_NAMES))
def test_dimacs_fixtures(cnf_dir, nv):
    path = cnf_path(cnf_dir, nv)
    p = O.Problem.from_dimacs(path)
    n, m, clauses, planted = py_read_dimacs(path)
    assert (p.n, p.m) == (n, m) == (nv, 8 * nv)
    assert len(clauses) == m and all(len(c) == 3 for c in clauses)
    assert np.array_equal(p.cp, planted_cp(m))
    assert np.array_equal(p.lits, np.array(sum(clauses, []), np.int32))
    if planted is not None:  # the planted assignment on line 1 satisfies every clause
        assert len(planted) == n
        a = np.zeros(n)
        for l in planted:
            a[abs(l) - 1] = 1.0 if l > 0 else -1.0
        assert py_unsat(p.cp, p.lits, a) == 0


def test_dimacs_edge_cases(tmp_path):
    f = tmp_path / "e.cnf"
    f.write_text("c hello\nc key=0101\np cnf 5 4\n1 -2 0\nc inside\n3\n 4 5 0 -1 0\n2 3 4 5 -1 0\n")
    p = O.Problem.from_dimacs(str(f))
    assert (p.n, p.m) == (5, 4)
    assert list(p.cp) == [0, 2, 5, 6, 11]
    assert list(p.lits) == [1, -2, 3, 4, 5, -1, 2, 3, 4, 5, -1]
    assert list(p.unit) == [-1, 0, 0, 0, 0]
    g = tmp_path / "long.cnf"
    g.write_text("p cnf 30 1\n" + " ".join(str(i) for i in range(1, 27)) + " 0\n")
    with pytest.raises(ValueError):
        O.Problem.from_dimacs(str(g))
    with pytest.raises(ValueError):
        O.Problem.from_dimacs(str(tmp_path / "missing.cnf"))


# ---------------------------------------------------------------------------
# initial state (dynex.cc:2643-2700) and its random streams
# ---------------------------------------------------------------------------
def test_mt19937_known_answer():
    # C++ [rand.predef]: the 10000th invocation of a default-constructed
    # std::mt19937 (seed 5489) produces 4123659995
    assert O.mt19937_nth(5489, 10000) == 4123659995


CPP_UNIFORM = r"""
#include <cstdio>
#include <cstdlib>
#include <random>
int main(int argc, char** argv) {
    std::mt19937 g(atoi(argv[1]));
    std::uniform_int_distribution<int> d(atoi(argv[2]), atoi(argv[3]));
    int n = atoi(argv[4]);
    for (int i = 0; i < n; i++) printf("%d\n", d(g));
}
"""


@pytest.fixture(scope="module")
def cpp_uniform(tmp_path_factory):
    d = tmp_path_factory.mktemp("cpp")
    src = d / "u.cpp"
    src.write_text(CPP_UNIFORM)
    exe = d / "u"
    subprocess.run(["g++", "-O1", "-std=c++17", str(src), "-o", str(exe)], check=True)
    return str(exe)


@pytest.mark.parametrize("seed,a,b", [(1, -1, 1), (42, -1, 1), (-1, -1, 1), (7, 0, 9), (123456, -5, 1000)])
def test_uniform_int_matches_libstdcxx(cpp_uniform, seed, a, b):
    out = subprocess.run([cpp_uniform, str(seed), str(a), str(b), "3000"], check=True, capture_output=True,
                         text=True).stdout.split()
    assert list(O.uniform_int_stream(seed, a, b, 3000)) == [int(v) for v in out]


def test_init_reference_seeds():
    cp = np.array([0, 3, 6, 7], np.int32)
    lits = np.array([1, 2, 3, -4, 5, 6, -2], np.int32)  # var 2 pinned to -1
    p = O.Problem(6, 3, cp, lits)
    x = O.init_ref(p, 0)
    assert np.all(x[[0, 2, 3, 4, 5]] == 1.0) and x[1] == -1.0 and np.all(x[6:] == 1.0)
    x = O.init_ref(p, 9)
    draws = O.uniform_int_stream(9, -1, 1, 6).astype(float)
    draws[1] = -1.0
    assert np.array_equal(x[:6], draws) and np.all(x[6:9] == 0.0) and np.all(x[9:] == 1.0)
    assert np.all(O.init_ref(p, -1)[[0, 2, 3, 4, 5]] == 0.0)
    assert np.all(O.init_ref(p, -2)[:6] == -1.0)


def test_fixture_n100_solves_from_seed0(cnf_dir):
    p = O.Problem.from_dimacs(cnf_path(cnf_dir, 100))
    x = O.init_ref(p, 0)
    steps, t, status = O.solve(p, O.params(p.m), x, 100000)
    assert status == 1
    assert py_unsat(p.cp, p.lits, x[:p.n]) == 0
    assert 0 < t <= steps * 0.125
