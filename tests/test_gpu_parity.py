"""GPU parity of the HIP path (through the C ABI) against the CPU oracle.

FP64 is held to bit-identity with the oracle, which restates the reference's
strict-IEEE arithmetic (dynex.cc).  FP32 is held to bit-identity with the
oracle's float instance (same operation order, clause-order fold), and its
deviation from the FP64 trajectory is reported."""
import os
import subprocess

import numpy as np
import pytest

import dnx
import oracle as O
from conftest import CNF_NAMES, cnf_path, planted_cp, random_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = dnx.Context(0)
    yield c
    c.close()


def gen(n, m, seed):
    lits = dnx.generate_planted(seed, n, m)
    return planted_cp(m), lits


def mixed_problem(seed=4):
    """k in {0,1,2,3,4,5}, a duplicated literal, an unused variable, unit
    pins.  The empty clause (a lone "0" in DIMACS, dynex.cc:2369-2386) takes
    the reference's k-SAT branch, whose C = INT_MAX/2 is not clamped
    (dynex.cc:457-470)."""
    rng = np.random.default_rng(seed)
    n = 60
    cls = []
    for c in range(300):
        k = int(rng.choice([1, 2, 3, 3, 3, 4, 5], p=[0.02, 0.2, 0.2, 0.2, 0.18, 0.1, 0.1]))
        vs = rng.choice(np.arange(1, n), k, replace=False)  # variable n never used
        cls.append(list(vs * rng.choice([-1, 1], k)))
    cls.append([5, 5, -7])
    cls.insert(17, [])
    cls.insert(200, [])
    cp = np.cumsum([0] + [len(c) for c in cls]).astype(np.int32)
    return n, cp, np.array(sum(cls, []), np.int32)


def both(ctx, n, cp, lits):
    return ctx.problem_from_lits(n, cp, lits), O.Problem(n, len(cp) - 1, cp, lits)


def bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64) if a.dtype == np.float64 else a,
                          np.asarray(b).view(np.uint64) if b.dtype == np.float64 else b)


# ---------------------------------------------------------------------------
# function-level parity: RHS (dynex.cc:332-530) and Euler (dynex.cc:301-327)
# ---------------------------------------------------------------------------
def rhs_cases(ctx, cnf_dir):
    for nv in (100, 1000, 10000):
        path = cnf_path(cnf_dir, nv)
        yield f"fixture{nv}", ctx.problem_from_dimacs(path), O.Problem.from_dimacs(path)
    n, cp, lits = mixed_problem()
    yield ("mixed",) + both(ctx, n, cp, lits)
    cp, lits = gen(3000, 24000, 2)
    yield ("planted3000",) + both(ctx, 3000, cp, lits)


def test_rhs_fp64_bitexact(ctx, cnf_dir):
    rng = np.random.default_rng(0)
    for name, gp, op in rhs_cases(ctx, cnf_dir):
        prm = dnx.params(gp.m)
        oprm = O.params(op.m)
        states = [random_state(rng, gp.n, gp.m, xl_hi=50.0) for _ in range(3)]
        wild = random_state(rng, gp.n, gp.m)
        wild[:gp.n] *= 1.5  # out-of-range values exercise the read clamps
        wild[gp.n:gp.n + gp.m] = rng.uniform(-0.5, 1.5, gp.m)
        wild[gp.n + gp.m:] = rng.uniform(0.0, 3.0, gp.m)
        states.append(wild)
        for x in states:
            d, loc, en = gp.rhs(prm, x, dnx.FP64)
            d0, loc0, en0 = O.rhs(op, oprm, x)
            assert loc == loc0, name
            assert bits_equal(d, d0), (name, np.flatnonzero(d != d0)[:10])
            assert abs(en - en0) <= 1e-9 * max(1.0, abs(en0))
        gp.free()


def test_rhs_fp32_bitexact(ctx, cnf_dir):
    rng = np.random.default_rng(1)
    for name, gp, op in rhs_cases(ctx, cnf_dir):
        prm = dnx.params(gp.m)
        oprm = O.params(op.m)
        for _ in range(3):
            x = random_state(rng, gp.n, gp.m, xl_hi=50.0).astype(np.float32)
            d, loc, en = gp.rhs(prm, x.astype(np.float64), dnx.FP32)
            d0, loc0, _ = O.rhs(op, oprm, x, prec=32)
            assert loc == loc0, name
            assert np.array_equal(d.astype(np.float32).view(np.uint32), d0.view(np.uint32)), name
        gp.free()


def test_euler_bitexact(ctx):
    n, cp, lits = mixed_problem(9)
    gp, op = both(ctx, n, cp, lits)
    m = gp.m
    prm, oprm = dnx.params(m), O.params(m)
    rng = np.random.default_rng(2)
    for h in (0.125, 0.0625, 0.03125, 0.015625):
        x = random_state(rng, n, m)
        d = rng.normal(0, 8, n + 2 * m)
        assert bits_equal(gp.euler(prm, x, d, h, dnx.FP64), O.dostep(op, oprm, x, d, h))
        x32, d32 = x.astype(np.float32), d.astype(np.float32)
        out = gp.euler(prm, x32.astype(np.float64), d32.astype(np.float64), h, dnx.FP32).astype(np.float32)
        assert np.array_equal(out, O.dostep(op, oprm, x32, d32, h, prec=32))
    gp.free()


# ---------------------------------------------------------------------------
# resident trajectories: the mode-6 loop (dynex.cc:1387-1417), fixed steps
# ---------------------------------------------------------------------------
def run_traj(gp, op, x0, checkpoints, prec, exchange=dnx.EXCHANGE_SLICED):
    """Steps the GPU state and the oracle to each checkpoint and compares."""
    m = gp.m
    prm, oprm = dnx.params(m), O.params(m)
    gp.set_exchange(exchange)
    st = gp.state(prm, x0, dnx.FP64 if prec == 64 else dnx.FP32)
    xo = x0.copy() if prec == 64 else x0.astype(np.float32)
    if op.has_unit:  # the reference pins V at initialisation (dynex.cc:2694-2699)
        xo[:op.n] = np.where(op.unit != 0, op.unit, xo[:op.n])
    t_o, done = 0.0, 0
    dev = []
    try:
        for cpnt in checkpoints:
            k = cpnt - done
            st.step(k)
            tt, lev, loc = O.trajectory(op, oprm, xo, k, prec=prec)
            t_o += tt
            r = st.read()
            assert r["steps"] == cpnt
            assert r["t"] == t_o
            assert r["level"] == lev[-1]
            if prec == 64:
                assert bits_equal(r["x"], xo), (cpnt, np.flatnonzero(r["x"] != xo)[:8])
            else:
                assert np.array_equal(r["x"].astype(np.float32), xo), cpnt
            done = cpnt
            dev.append(r["x"])
    finally:
        st.free()
    return dev


@pytest.mark.parametrize("exchange", [dnx.EXCHANGE_SLICED, dnx.EXCHANGE_GROUPED, dnx.EXCHANGE_LITERAL])
@pytest.mark.parametrize("prec", [64, 32])
def test_trajectory_c1_1000_steps(ctx, prec, exchange):
    """n=100, m=800: 1000 steps bit-identical to the oracle with each exchange
    (V from LDS slices and contributions in runs per fold block; grouped by
    folding XCD; literal order)."""
    cp, lits = gen(100, 800, 11)
    gp, op = both(ctx, 100, cp, lits)
    x0 = dnx.init_uniform(11, 100, 800)
    run_traj(gp, op, x0, [1, 2, 10, 100, 500, 1000], prec, exchange)
    gp.free()


@pytest.mark.parametrize("exchange", [dnx.EXCHANGE_SLICED, dnx.EXCHANGE_GROUPED])
@pytest.mark.parametrize("prec", [64, 32])
def test_trajectory_exchange_ragged_shapes(ctx, prec, exchange):
    """Clause counts that end inside a chunk or sub-batch, variable counts that
    end inside a fold block or slice, variables with no clause, many units
    and slice sets, a duplicated literal: each exchange against the oracle."""
    cases = [(3, 1), (7, 65), (300, 1031), (2051, 9001), (5000, 3000), (30000, 130000)]
    for seed, (n, m) in enumerate(cases):
        cp, lits = gen(n, m, 300 + seed)
        if seed == 2:
            lits[3] = lits[4]
        gp, op = both(ctx, n, cp, lits)
        x0 = dnx.init_uniform(300 + seed, n, m)
        run_traj(gp, op, x0, [1, 3, 40], prec, exchange)
        gp.free()


@pytest.mark.parametrize("prec", [64, 32])
def test_trajectory_mixed_k(ctx, prec):
    n, cp, lits = mixed_problem(5)
    gp, op = both(ctx, n, cp, lits)
    x0 = random_state(np.random.default_rng(6), n, len(cp) - 1)
    run_traj(gp, op, x0, [1, 3, 50, 300], prec)
    gp.free()


def test_trajectory_fixture_seed0(ctx, cnf_dir):
    path = cnf_path(cnf_dir, 1000)
    gp, op = ctx.problem_from_dimacs(path), O.Problem.from_dimacs(path)
    x0 = O.init_ref(op, 0)
    run_traj(gp, op, x0, [1, 5, 60, 200], 64)
    gp.free()


def test_trajectory_c2_fp64_1000_steps(ctx):
    """n=10,000, m=80,000: 1000 steps bit-identical to the oracle (well inside
    the 1e-9 relative bound the path is specified to)."""
    cp, lits = gen(10000, 80000, 21)
    gp, op = both(ctx, 10000, cp, lits)
    x0 = dnx.init_uniform(21, 10000, 80000)
    run_traj(gp, op, x0, [1, 10, 100, 1000], 64)
    gp.free()


@pytest.mark.parametrize("prec", [64, 32])
def test_trajectory_c3_full_size(ctx, prec):
    """n=100,000, m=800,000 (the largest single-instance size): 100 steps
    bit-identical to the oracle, compared after 1, 4, 30 and 100 steps, with
    the default exchange (sliced in FP32; FP64 needs more slices than two
    rounds of units allow and takes the grouped one); and the grouped and
    literal-order exchanges to step 4."""
    cp, lits = gen(100000, 800000, 23)
    gp, op = both(ctx, 100000, cp, lits)
    x0 = dnx.init_uniform(23, 100000, 800000)
    run_traj(gp, op, x0, [1, 4, 30, 100], prec)
    run_traj(gp, op, x0, [1, 4], prec, dnx.EXCHANGE_GROUPED)
    run_traj(gp, op, x0, [1, 4], prec, dnx.EXCHANGE_LITERAL)
    gp.free()


def test_trajectory_c2_fp32_vs_fp64(ctx):
    """FP32 is bit-identical to the float oracle for 1,000 steps; its
    distance from the FP64 trajectory is reported (not a parity bound)."""
    cp, lits = gen(10000, 80000, 22)
    gp, op = both(ctx, 10000, cp, lits)
    x0 = dnx.init_uniform(22, 10000, 80000)
    marks = [10, 100, 1000]
    x32 = run_traj(gp, op, x0, marks, 32)
    s64 = gp.state(dnx.params(80000), x0, dnx.FP64)
    try:
        done = 0
        for k, x in zip(marks, x32):
            s64.step(k - done)
            done = k
            ref = s64.read()["x"]
            dv = np.max(np.abs(x[:10000] - ref[:10000]))
            print(f"step {k}: max |V32 - V64| = {dv:.3e}")
    finally:
        s64.free()
    gp.free()


@pytest.mark.parametrize("prec", [64, 32])
def test_loc_tracking_matches_oracle(ctx, prec):
    """The device-side loc count of every evaluated state (the reference's
    `global`, dynex.cc:1188, 584-600) against the oracle's per-state loc,
    including an empty clause, which is never satisfied."""
    n, cp, lits = mixed_problem(8)
    gp, op = both(ctx, n, cp, lits)
    m = gp.m
    x0 = random_state(np.random.default_rng(9), n, m)
    st = gp.state(dnx.params(m), x0, dnx.FP64 if prec == 64 else dnx.FP32)
    xo = x0.copy() if prec == 64 else x0.astype(np.float32)
    xo[:n] = np.where(op.unit != 0, op.unit, xo[:n])
    try:
        assert st.best_loc() == (m, m)  # nothing evaluated yet
        st.track_loc(True)
        _, _, loc = O.trajectory(op, O.params(m), xo, 90, prec=prec)
        for k in (1, 2, 30, 90):
            st.step(k - (0 if k == 1 else {2: 1, 30: 2, 90: 30}[k]))
            assert st.best_loc() == (int(loc[:k].min()), int(loc[k - 1])), k
        assert loc.min() >= 2  # the two empty clauses
    finally:
        st.free()
    gp.free()


# ---------------------------------------------------------------------------
# solving: stop at the first verified state (dynex.cc:662-703)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("nv", sorted(CNF_NAMES))
def test_solve_fixtures_fp64_matches_oracle(ctx, cnf_dir, nv):
    path = cnf_path(cnf_dir, nv)
    gp, op = ctx.problem_from_dimacs(path), O.Problem.from_dimacs(path)
    x0 = O.init_ref(op, 0)
    r = gp.solve(dnx.params(gp.m), x0, dnx.FP64, cap=100000)
    assert r["status"] == dnx.SAT
    assert gp.check_assignment(r["assign"]) == 0
    if nv <= 1000:
        xo = x0.copy()
        steps, t, status = O.solve(op, O.params(op.m), xo, 100000)
        assert status == 1 and r["steps"] == steps and r["t"] == t
        assert np.array_equal(r["assign"], np.where(xo[:op.n] >= 0, 1, -1))
    gp.free()


@pytest.mark.parametrize("prec", [64, 32])
def test_solve_generated_matches_oracle(ctx, prec):
    for seed, (n, m) in enumerate([(100, 800), (500, 2150), (2000, 12000)]):
        cp, lits = gen(n, m, 100 + seed)
        gp, op = both(ctx, n, cp, lits)
        x0 = dnx.init_uniform(100 + seed, n, m)
        r = gp.solve(dnx.params(m), x0, dnx.FP64 if prec == 64 else dnx.FP32, cap=20000)
        xo = x0.copy() if prec == 64 else x0.astype(np.float32)
        steps, t, status = O.solve(op, O.params(m), xo, 20000, prec=prec)
        assert r["status"] == (dnx.SAT if status == 1 else dnx.UNKNOWN)
        assert r["steps"] == steps and r["t"] == t
        if status == 1:
            assert gp.check_assignment(r["assign"]) == 0
        gp.free()


def test_step_cap_reports_unknown(ctx):
    cp, lits = gen(10000, 80000, 5)
    gp = ctx.problem_from_lits(10000, cp, lits)
    x0 = dnx.init_uniform(5, 10000, 80000)
    for cap in (0, 1, 4, 31):
        r = gp.solve(dnx.params(80000), x0, dnx.FP64, cap=cap)
        assert r["status"] == dnx.UNKNOWN and r["steps"] == cap
    gp.free()


# ---------------------------------------------------------------------------
# portfolio (-w, dynex.cc:2643-2700 initial states, 2657 alpha/i, 2731-2801)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("seed", [0, 3])
def test_portfolio_winner_is_first_single_solve(ctx, cnf_dir, prec, seed):
    """Every copy advances in lockstep, so the winner is the copy whose own
    solve needs the fewest steps (lowest index on a tie), with that solve's
    step count, time and assignment."""
    gp = ctx.problem_from_dimacs(cnf_path(cnf_dir, 200))
    pr = dnx.FP64 if prec == 64 else dnx.FP32
    threads, cap = 5, 30000
    single = []
    for i in range(threads):
        prm = dnx.params(gp.m)
        if i > 0:
            prm.alpha = prm.alpha / i
        single.append(gp.solve(prm, gp.init_reference_thread(seed, i), pr, cap=cap))
    r = gp.portfolio_solve(dnx.params(gp.m), seed=seed, threads=threads, prec=pr, cap=cap)
    sat = [i for i in range(threads) if single[i]["status"] == dnx.SAT]
    assert sat, [s["status"] for s in single]
    best = min(sat, key=lambda i: (single[i]["steps"], i))
    assert r["status"] == dnx.SAT and r["winner"] == best
    assert r["steps"] == single[best]["steps"] and r["t"] == single[best]["t"]
    assert np.array_equal(r["assign"], single[best]["assign"])
    assert gp.check_assignment(r["assign"]) == 0
    gp.free()


def test_cli_portfolio(cnf_dir, tmp_path):
    src = cnf_path(cnf_dir, 100)
    path = str(tmp_path / os.path.basename(src))
    with open(src) as f, open(path, "w") as g:
        g.write(f.read())
    exe = os.path.join(dnx.HERE, "lib", "dynex")
    r = subprocess.run([exe, "-i", path, "-w", "4"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "SATISFIABLE" in r.stdout and "SAT (VERIFIED)" in r.stdout
    assert r.stdout.rstrip().endswith("0")


@pytest.mark.parametrize("quiet", [0, 1])
def test_cli_solves_fixture(cnf_dir, tmp_path, quiet):
    """The reference's solved output (dynex.cc:662-690, 1499-1514, 2864):
    SOLUTION FOUND, the signed literals of the verified state, VERIFYING,
    SAT (VERIFIED), TIME SPENT, BEST LOC, the v lines and global = 0; the
    reference prints all of it under -q 1 as well (it clears quiet when it
    finds the solution, dynex.cc:663, 716)."""
    src = cnf_path(cnf_dir, 200)
    path = str(tmp_path / os.path.basename(src))
    with open(src) as f, open(path, "w") as g:
        g.write(f.read())
    exe = os.path.join(dnx.HERE, "lib", "dynex")
    r = subprocess.run([exe, "-i", path, "-q", str(quiet)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    assert "s [0] SATISFIABLE" in r.stdout and "c [0] SAT (VERIFIED)" in lines
    k = [i for i, ln in enumerate(lines) if ln.startswith("c [0] T=") and "SOLUTION FOUND" in ln]
    assert len(k) == 1
    listing = [int(t) for t in lines[k[0] + 1].split()]
    assert lines[k[0] + 2] == "c [0] VERIFYING..."
    assert "c [0] BEST LOC = 0 (GLOBAL 0) **" in lines
    vals = [int(t) for ln in lines if ln.startswith("v ") for t in ln[2:].split()]
    assert vals[-1] == 0 and len(vals) == 201
    sol = open(path + ".solution.txt").read().replace(",", " ").split()
    op = O.Problem.from_dimacs(path)
    a = np.zeros(op.n)
    for t in sol:
        a[abs(int(t)) - 1] = 1.0 if int(t) > 0 else -1.0
    assert O.verify(op, np.concatenate([a, np.zeros(2 * op.m)])) == 0
    # the listing names each variable whose V is not zero with its sign, and
    # agrees with the solution file and the v lines
    assert len(set(abs(v) for v in listing)) == len(listing)
    for v in listing:
        assert np.sign(v) == a[abs(v) - 1] and vals[abs(v) - 1] == v
    assert r.stdout.rstrip().endswith("0") and lines[-1].endswith("0")


def unsat_file(cnf_dir, tmp_path):
    """The n=100 fixture plus an empty clause: never satisfiable, and its
    loc never drops below 1."""
    src = cnf_path(cnf_dir, 100)
    txt = open(src).read().replace("p cnf 100 800", "p cnf 100 801")
    path = str(tmp_path / "with_empty_clause.cnf")
    with open(path, "w") as g:
        g.write(txt.rstrip("\n") + "\n0\n")
    return path


def test_cli_step_cap_unknown_and_global(cnf_dir, tmp_path):
    """-z caps the adaptive loop: UNKNOWN, BEST LOC and `global` equal the
    lowest loc the oracle's trajectory reaches over states 0..cap."""
    path = unsat_file(cnf_dir, tmp_path)
    exe = os.path.join(dnx.HERE, "lib", "dynex")
    r = subprocess.run([exe, "-i", path, "-z", "60"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    op = O.Problem.from_dimacs(path)
    assert op.m == 801 and op.cp[800] == op.cp[801]
    x = O.init_ref(op, 0)
    _, _, loc = O.trajectory(op, O.params(op.m), x, 61)
    g = int(loc.min())
    assert g >= 1
    lines = r.stdout.splitlines()
    assert "s [0] UNKNOWN" in lines and "SOLUTION FOUND" not in r.stdout
    assert f"c [0] BEST LOC = {g} (GLOBAL {g}) **" in lines
    assert lines[-1].endswith(str(g)) and r.stdout.rstrip().split()[-1] == str(g)


def test_cli_walltime_ends_silently(cnf_dir, tmp_path):
    """-u: the reference's RHS ends the thread once the wall clock passes the
    limit (dynex.cc:335-345), so no TIME SPENT or s line follows; only
    `global`, the lowest loc reached, is printed (dynex.cc:2864)."""
    path = unsat_file(cnf_dir, tmp_path)
    exe = os.path.join(dnx.HERE, "lib", "dynex")
    r = subprocess.run([exe, "-i", path, "-u", "0.5"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "TIME SPENT" not in r.stdout and "s [0]" not in r.stdout and "SOLUTION" not in r.stdout
    g = int(r.stdout.rstrip().split()[-1])
    assert 1 <= g <= 801
