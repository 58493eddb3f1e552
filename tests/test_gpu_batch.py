"""Batched engine (FP32) on the GPU: every instance must compute the same bits
as the float oracle alone, whatever the batch around it, the execution shape
or the number of steps per launch."""
import numpy as np
import pytest

import dnx
import oracle as O
from conftest import planted_cp

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    """Equal float32 arrays bit for bit (signed zeros included)."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.fixture(scope="module")
def ctx():
    c = dnx.Context(0)
    yield c
    c.close()


def make_set(seed, count, nmin=60, nmax=2500, rmin=4.6, rmax=8.0, with_general=True):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        m = int(round(rng.uniform(rmin, rmax) * n))
        lits = dnx.generate_planted(seed * 1000 + i, n, m)
        cp = planted_cp(m)
        if with_general and i % 7 == 3:
            # turn a few clauses into k=2, k=1 and k=4 clauses
            cl = [list(lits[3 * c:3 * c + 3]) for c in range(m)]
            cl[0] = cl[0][:2]
            cl[1] = cl[1][:1]
            cl[2] = cl[2] + [-cl[3][0]]
            if i % 14 == 3:
                cl[4] = []  # the empty clause: C = INT_MAX/2, never solved (dynex.cc:457-470)
            cp = np.cumsum([0] + [len(c) for c in cl]).astype(np.int32)
            lits = np.array(sum(cl, []), np.int32)
        x0 = dnx.init_uniform(seed * 1000 + i, n, m)
        out.append((n, cp, lits, x0))
    return out


def oracle_solve(n, cp, lits, x0, cap):
    op = O.Problem(n, len(cp) - 1, cp, lits)
    x = x0.astype(np.float32)
    if op.has_unit:
        x[:n] = np.where(op.unit != 0, op.unit, x[:n])
    steps, t, status = O.solve(op, O.params(op.m), x, cap, prec=32)
    return steps, t, status, x


def run_batch(ctx, inst, cap, spl, shapes=None, stop=True, wide=0, gang=None, split=False, flow=True, order=0,
              many=False, epoch=None, regroup=True):
    """gang = (max_blocks, clauses_per_workgroup, steps_per_launch) enables
    gang mode for the tail (the instances left once wide mode applies);
    split keeps the single-chunk LDS instances of the tail in the fused
    kernel beside wide mode; flow runs the fused launch until the wide-mode
    threshold instead of spl steps per launch; order > 0 starts instances of
    at least that many clauses first and the others smallest first; many
    admits the whole set in one dnx_batch_admit_many call ("defaults": with
    prms NULL); epoch = (min_clauses, steps) turns the elastic tail on (off
    otherwise); regroup picks the exchange of its 3-SAT gangs."""
    b = dnx.Batch(ctx, slots=len(inst) + 2, arena_bytes=1 << 30)
    b.set_wide(wide, 30)
    b.set_tail_split(split)
    b.set_flow(flow)
    b.set_order(order)
    if gang:
        b.set_gang(*gang)
    else:
        b.set_gang(0)
    b.set_epoch(*(epoch or (0, 256)))
    b.set_gang_regroup(regroup)
    if shapes:
        for md, (bs, lds) in shapes.items():
            b.set_shape(md, bs, lds)
    slots = []
    if many:
        hps = [dnx.host_problem(n, cp, lits) for n, cp, lits, x0 in inst]
        prms = None if many == "defaults" else [dnx.params(len(cp) - 1) for n, cp, lits, x0 in inst]
        slots = b.admit_many(hps, [x0 for n, cp, lits, x0 in inst], cap=cap, stop_on_solve=stop, prms=prms)
        for hp in hps:
            hp.free()
    else:
        for n, cp, lits, x0 in inst:
            hp = dnx.host_problem(n, cp, lits)
            slots.append(b.admit(hp, dnx.params(len(cp) - 1), x0, cap=cap, stop_on_solve=stop))
            hp.free()
    b.run(steps_per_launch=spl)
    res = b.poll()
    out = []
    for (n, cp, lits, x0), s in zip(inst, slots):
        m = len(cp) - 1
        st = b.read_state(s, n, m)
        a, unsat = b.assignment(s, n)
        out.append(dict(status=int(res["status"][s]), steps=int(res["steps"][s]), t=float(res["t"][s]),
                        x=st["x"].astype(np.float32), assign=a, unsat=unsat, mode=b.info(s)[2]))
    b.close()
    return out


@pytest.fixture(scope="module")
def inst_set():
    return make_set(7, 36)


@pytest.fixture(scope="module")
def oracle_results(inst_set):
    return [oracle_solve(n, cp, lits, x0, 1500) for n, cp, lits, x0 in inst_set]


@pytest.mark.parametrize("many", [False, True, "defaults"])
def test_batch_matches_oracle(ctx, inst_set, oracle_results, many):
    out = run_batch(ctx, inst_set, 1500, spl=37, many=many)
    modes = set()
    for r, (steps, t, status, xo), (n, cp, lits, x0) in zip(out, oracle_results, inst_set):
        modes.add(r["mode"])
        assert r["status"] == (dnx.SAT if status == 1 else dnx.UNKNOWN)
        assert r["steps"] == steps and r["t"] == t
        assert same_bits(r["x"], xo)
        if status == 1:
            assert r["unsat"] == 0
    # with the default 48 KiB mode-0 budget the larger instances hold their
    # contributions in LDS chunk by chunk; mode 1 is covered by the shapes below
    assert 0 in modes


@pytest.mark.parametrize("flow,wide", [(False, 0), (True, 0), (True, 60000), (True, 1 << 40)])
def test_flow_gives_identical_bits(ctx, inst_set, oracle_results, flow, wide):
    """Flow launches (each instance kept on its workgroup until it ends or
    the running clauses fall to the wide-mode threshold) against launches of
    a fixed number of steps: the same bits, with and without wide mode."""
    out = run_batch(ctx, inst_set, 1500, spl=29, flow=flow, wide=wide)
    for r, (steps, t, status, xo) in zip(out, oracle_results):
        assert r["status"] == (dnx.SAT if status == 1 else dnx.UNKNOWN)
        assert r["steps"] == steps and r["t"] == t
        assert same_bits(r["x"], xo)


@pytest.mark.parametrize("shapes", [
    {0: (128, 0), 1: (128, 0), 2: (256, 0)},        # everything in device memory
    {0: (512, 0), 1: (1024, 65536), 2: (1024, 0)},  # V/dV in LDS
    {0: (1024, 160 * 1024), 1: (256, 0), 2: (512, 0)},  # as much in LDS as fits
])
def test_shapes_give_identical_bits(ctx, inst_set, oracle_results, shapes):
    out = run_batch(ctx, inst_set, 1500, spl=64, shapes=shapes)
    for r, (steps, t, status, xo) in zip(out, oracle_results):
        assert r["steps"] == steps and r["t"] == t
        assert same_bits(r["x"], xo)


@pytest.mark.parametrize("wide", [1 << 40, 60000])
def test_wide_mode_gives_identical_bits(ctx, inst_set, oracle_results, wide):
    """Wide mode from the start (all instances), and switching to it once the
    running instances hold at most 60,000 clauses."""
    out = run_batch(ctx, inst_set, 1500, spl=40, wide=wide)
    for r, (steps, t, status, xo) in zip(out, oracle_results):
        assert r["status"] == (dnx.SAT if status == 1 else dnx.UNKNOWN)
        assert r["steps"] == steps and r["t"] == t
        assert same_bits(r["x"], xo)
        if status == 1:
            assert r["unsat"] == 0


@pytest.mark.parametrize("wide,gang", [
    (1 << 40, (512, 256, 37)),     # gangs of up to 32 workgroups, relaunch every 37 steps
    (1 << 40, (128, 2048, 128)),   # few large workgroups
    (60000, (128, 1024, 64)),      # fused kernel, then wide mode, then gangs
    (1 << 40, (48, 128, 50)),      # budget below the running set's wish: gangs cut to a common level
])
def test_gang_mode_gives_identical_bits(ctx, inst_set, oracle_results, wide, gang):
    out = run_batch(ctx, inst_set, 1500, spl=40, wide=wide, gang=gang)
    for r, (steps, t, status, xo) in zip(out, oracle_results):
        assert r["status"] == (dnx.SAT if status == 1 else dnx.UNKNOWN)
        assert r["steps"] == steps and r["t"] == t
        assert same_bits(r["x"], xo)
        if status == 1:
            assert r["unsat"] == 0



@pytest.mark.parametrize("regroup", [True, False])
@pytest.mark.parametrize("epoch", [(2000, 32), (6000, 7, 256), (1, 256, 4096)])
def test_epoch_gives_identical_bits(ctx, inst_set, oracle_results, epoch, regroup):
    """The elastic tail: fewer instances than CUs run from the start, so
    every epoch gives the spare CUs to the instances of at least min_clauses
    clauses as gangs and keeps the others on single workgroups; gangs are
    re-planned every `steps` gang steps (instances move between gangs of
    other sizes and single workgroups).  3-SAT gangs regroup their
    contributions by folding workgroup or exchange them through the
    transpose; the instances with other clause lengths always take the
    transpose.  Every instance computes the oracle's bits."""
    out = run_batch(ctx, inst_set, 1500, spl=37, gang=(256, 384, 512), flow=True, epoch=epoch, regroup=regroup)
    for (n, cp, lits, x0), o, (steps, t, status, xo) in zip(inst_set, out, oracle_results):
        assert o["status"] == status and o["steps"] == steps and o["t"] == t, (n, o["steps"], steps)
        assert same_bits(o["x"], xo), n


@pytest.mark.parametrize("regroup", [True, False])
def test_epoch_gangs_wider_than_their_instances(ctx, regroup):
    """Gangs of 32 workgroups over instances of 200-256 clauses and 46-60
    variables: several workgroups of each gang evaluate no clause and own no
    variable (their regions are empty); every instance computes the oracle's
    bits for 300 steps."""
    inst = []
    for i, (n, m) in enumerate([(46, 200), (60, 256), (53, 231)]):
        lits = dnx.generate_planted(777 + i, n, m)
        inst.append((n, planted_cp(m), lits, dnx.init_uniform(777 + i, n, m)))
    out = run_batch(ctx, inst, 300, spl=37, gang=(256, 1, 512), flow=True, epoch=(1, 64, 64), stop=False,
                    regroup=regroup)
    for (n, cp, lits, x0), o in zip(inst, out):
        op = O.Problem(n, len(cp) - 1, cp, lits)
        x = x0.astype(np.float32)
        t, _, _ = O.trajectory(op, O.params(op.m), x, 300, prec=32)
        assert o["status"] == dnx.UNKNOWN and o["steps"] == 300 and o["t"] == t, (n, o["steps"])
        assert same_bits(o["x"], x), n


def test_epoch_after_flow_with_more_instances_than_cus(ctx):
    """More instances than CUs: the first flow launch stops once enough of
    its workgroups found the list empty for the largest instance's gang, and
    epochs take over; the bits equal the oracle's for every instance."""
    inst = make_set(13, 300, nmin=60, nmax=300, with_general=False)
    big = make_set(17, 3, nmin=2300, nmax=2600, rmin=4.4, rmax=4.6, with_general=False)
    inst = big[:1] + inst + big[1:]
    ref = [oracle_solve(n, cp, lits, x0, 700) for n, cp, lits, x0 in inst]
    out = run_batch(ctx, inst, 700, spl=37, gang=(256, 384, 512), flow=True, epoch=(5000, 64), order=5000)
    for (n, cp, lits, x0), o, (steps, t, status, xo) in zip(inst, out, ref):
        assert o["status"] == status and o["steps"] == steps and o["t"] == t, (n, o["steps"], steps)
        assert same_bits(o["x"], xo), n


@pytest.mark.parametrize("flow,order", [(True, 6000), (False, 1 << 40)])
def test_order_gives_identical_bits(ctx, inst_set, oracle_results, flow, order):
    """Large instances first then the others smallest first (flow launch),
    and every instance smallest first (launches of a fixed step count)."""
    out = run_batch(ctx, inst_set, 1500, spl=33, flow=flow, order=order)
    for r, (steps, t, status, xo) in zip(out, oracle_results):
        assert r["status"] == (dnx.SAT if status == 1 else dnx.UNKNOWN)
        assert r["steps"] == steps and r["t"] == t
        assert same_bits(r["x"], xo)
        if status == 1:
            assert r["unsat"] == 0


@pytest.mark.parametrize("wide", [1 << 40, 60000])
def test_tail_split_gives_identical_bits(ctx, inst_set, oracle_results, wide):
    """The tail split from the start, and from 60,000 running clauses on."""
    out = run_batch(ctx, inst_set, 1500, spl=40, wide=wide, split=True)
    for r, (steps, t, status, xo) in zip(out, oracle_results):
        assert r["status"] == (dnx.SAT if status == 1 else dnx.UNKNOWN)
        assert r["steps"] == steps and r["t"] == t
        assert same_bits(r["x"], xo)
        if status == 1:
            assert r["unsat"] == 0


def test_tail_split_fixed_steps(ctx):
    inst = make_set(15, 6, nmin=100, nmax=3000, with_general=True)
    out = run_batch(ctx, inst, 45, spl=25, stop=False, wide=1 << 40, split=True)
    for r, (n, cp, lits, x0) in zip(out, inst):
        op = O.Problem(n, len(cp) - 1, cp, lits)
        x = x0.astype(np.float32)
        if op.has_unit:
            x[:n] = np.where(op.unit != 0, op.unit, x[:n])
        O.trajectory(op, O.params(op.m), x, 45, prec=32)
        assert r["status"] == dnx.UNKNOWN and r["steps"] == 45
        assert same_bits(r["x"], x)


def test_gang_fixed_step_trajectory(ctx):
    """Fixed steps (no stop on solve) with k=1, 2, 4 clauses and unit pins,
    ending inside a launch and across launch boundaries."""
    inst = make_set(14, 6, nmin=100, nmax=3000, with_general=True)
    out = run_batch(ctx, inst, 45, spl=25, stop=False, wide=1 << 40, gang=(512, 300, 16))
    for r, (n, cp, lits, x0) in zip(out, inst):
        op = O.Problem(n, len(cp) - 1, cp, lits)
        x = x0.astype(np.float32)
        if op.has_unit:
            x[:n] = np.where(op.unit != 0, op.unit, x[:n])
        O.trajectory(op, O.params(op.m), x, 45, prec=32)
        assert r["status"] == dnx.UNKNOWN and r["steps"] == 45
        assert same_bits(r["x"], x)


def test_full_mixed_workload(ctx):
    """The benchmark's workload at full size (4096 planted instances, n
    log-uniform in [100, 5000], ratio U[4.3, 8]) with the default policy:
    every instance ends SAT with an assignment that passes the exact clause
    check, and sampled instances compute the same steps and time as the
    single-instance FP32 path and the float oracle."""
    rng = np.random.default_rng(4096)
    count = 4096
    ns = np.exp(rng.uniform(np.log(100), np.log(5000), count)).round().astype(int)
    rs = rng.uniform(4.3, 8.0, count)
    inst = []
    for i in range(count):
        n = int(ns[i])
        m = int(round(rs[i] * n))
        s = 4096 * 100003 + i
        inst.append((n, planted_cp(m), dnx.generate_planted(s, n, m), dnx.init_uniform(s, n, m)))
    b = dnx.Batch(ctx, slots=count)
    hps = [dnx.host_problem(n, cp, lits) for n, cp, lits, x0 in inst]
    slots = b.admit_many(hps, [x0 for n, cp, lits, x0 in inst], cap=100000)
    for hp in hps:
        hp.free()
    b.run(steps_per_launch=2048)
    res = b.poll()
    assert np.all(res["status"][slots] == dnx.SAT)
    for (n, cp, lits, x0), s in zip(inst, slots):
        assert b.assignment(s, n)[1] == 0
    steps = res["steps"][slots]
    # six short instances; the two that ran longest (on single workgroups:
    # below the elastic tail's 12,000 clauses); and the two longest of at
    # least 12,000 clauses, which end in gangs beside single-workgroup
    # instances -- these four also compare their final state bit for bit
    pick = [i for i in np.random.default_rng(1).permutation(count) if steps[i] < 1500][:6]
    order = np.argsort(steps)
    longest = [int(i) for i in order[-2:]]
    longest += [int(i) for i in order[::-1] if len(inst[i][1]) - 1 >= 12000][:2]
    states = {i: b.read_state(slots[i], inst[i][0], len(inst[i][1]) - 1)["x"].astype(np.float32) for i in longest}
    b.close()
    for i in longest:
        n, cp, lits, x0 = inst[i]
        o_steps, o_t, o_status, xo = oracle_solve(n, cp, lits, x0, 100000)
        assert o_status == 1 and o_steps == steps[i] and o_t == res["t"][slots][i], (n, o_steps, steps[i])
        assert same_bits(states[i], xo), n
    for i in pick:
        n, cp, lits, x0 = inst[i]
        gp = ctx.problem_from_lits(n, cp, lits)
        r = gp.solve(dnx.params(len(cp) - 1), x0, dnx.FP32, cap=100000)
        gp.free()
        assert r["steps"] == steps[i]
        o_steps, o_t, o_status, _ = oracle_solve(n, cp, lits, x0, 100000)
        assert o_status == 1 and o_steps == steps[i]


def test_lds_edge_instance(ctx):
    """An instance whose all-in-LDS form needs 163,700 bytes -- under the
    device's 160 KiB, but not with the fused kernel's static LDS on top --
    must be cut into two LDS chunks (the launch stays within the device
    limit) and compute the oracle's bits."""
    n, m = 1000, 12975
    lits = dnx.generate_planted(4242, n, m)
    cp = planted_cp(m)
    x0 = dnx.init_uniform(4242, n, m)
    assert 8 * n + 12 * m == 163700
    out = run_batch(ctx, [(n, cp, lits, x0)], 300, spl=64, shapes={0: (1024, 160 * 1024)})[0]
    steps, t, status, xo = oracle_solve(n, cp, lits, x0, 300)
    assert out["mode"] == 0
    assert out["steps"] == steps and out["t"] == t
    assert same_bits(out["x"], xo)


def test_wide_fixed_step_trajectory(ctx):
    inst = make_set(12, 5, nmin=100, nmax=3000, with_general=True)
    out = run_batch(ctx, inst, 45, spl=25, stop=False, wide=1 << 40)
    for r, (n, cp, lits, x0) in zip(out, inst):
        op = O.Problem(n, len(cp) - 1, cp, lits)
        x = x0.astype(np.float32)
        if op.has_unit:
            x[:n] = np.where(op.unit != 0, op.unit, x[:n])
        O.trajectory(op, O.params(op.m), x, 45, prec=32)
        assert r["status"] == dnx.UNKNOWN and r["steps"] == 45
        assert same_bits(r["x"], x)


def test_alone_equals_in_the_mix(ctx, inst_set):
    mix = run_batch(ctx, inst_set, 1500, spl=101)
    for i in (0, 5, 17, 35):
        alone = run_batch(ctx, [inst_set[i]], 1500, spl=9)[0]
        assert alone["steps"] == mix[i]["steps"] and alone["t"] == mix[i]["t"]
        assert same_bits(alone["x"], mix[i]["x"])


def test_fixed_step_trajectory(ctx):
    inst = make_set(11, 6, nmin=100, nmax=3000, with_general=True)
    out = run_batch(ctx, inst, 60, spl=25, stop=False)
    for r, (n, cp, lits, x0) in zip(out, inst):
        op = O.Problem(n, len(cp) - 1, cp, lits)
        x = x0.astype(np.float32)
        if op.has_unit:
            x[:n] = np.where(op.unit != 0, op.unit, x[:n])
        O.trajectory(op, O.params(op.m), x, 60, prec=32)
        assert r["status"] == dnx.UNKNOWN and r["steps"] == 60
        assert same_bits(r["x"], x)


def test_single_instance_path_agrees(ctx, inst_set):
    """The batched engine and dnx_solve(FP32) are the same computation."""
    out = run_batch(ctx, inst_set[:8], 1500, spl=50)
    for r, (n, cp, lits, x0) in zip(out, inst_set[:8]):
        gp = ctx.problem_from_lits(n, cp, lits)
        s = gp.solve(dnx.params(len(cp) - 1), x0, dnx.FP32, cap=1500)
        gp.free()
        assert s["steps"] == r["steps"] and s["t"] == r["t"]
        if s["status"] == dnx.SAT:
            assert np.array_equal(s["assign"], r["assign"])


@pytest.mark.parametrize("many", [False, True])
def test_large_instance_admission(ctx, many):
    """Instances of ~20 MB and ~45 MB arena ranges between small ones (the
    larger's host-built part, ~22 MB, exceeds the 16 MB admission staging
    chunk, which then grows; admitted together, the set spans several
    chunks): each instance's fixed-step trajectory equals the single-instance
    FP32 path's, bit for bit (the large ones run with their state in device
    memory)."""
    inst = []
    for i, (n, m) in enumerate([(300, 1350), (60000, 258000), (500, 2300), (150000, 640000), (700, 3100)]):
        inst.append((n, planted_cp(m), dnx.generate_planted(900 + i, n, m), dnx.init_uniform(900 + i, n, m)))
    out = run_batch(ctx, inst, 40, spl=20, stop=False, many=many)
    for r, (n, cp, lits, x0) in zip(out, inst):
        gp = ctx.problem_from_lits(n, cp, lits)
        st = gp.state(dnx.params(len(cp) - 1), x0, dnx.FP32)
        st.step(40)
        s = st.read()
        st.free()
        gp.free()
        assert r["status"] == dnx.UNKNOWN and r["steps"] == 40 and s["steps"] == 40
        assert r["t"] == s["t"]
        assert same_bits(r["x"], s["x"].astype(np.float32))


def test_retire_readmit_and_budget(ctx):
    inst = make_set(13, 10, nmin=200, nmax=2000)
    b = dnx.Batch(ctx, slots=4, arena_bytes=8 << 20)
    arena, free0 = b.memory()
    assert free0 == arena
    slots = []
    for n, cp, lits, x0 in inst[:4]:
        hp = dnx.host_problem(n, cp, lits)
        slots.append(b.admit(hp, dnx.params(len(cp) - 1), x0, cap=1500))
        hp.free()
    n, cp, lits, x0 = inst[4]
    hp = dnx.host_problem(n, cp, lits)
    with pytest.raises(dnx.DnxError) as e:
        b.admit(hp, dnx.params(len(cp) - 1), x0)
    assert e.value.code == dnx.E_BUDGET  # no free slot
    b.run(64)
    for s in slots:
        b.retire(s)
    assert b.memory()[1] == arena
    s = b.admit(hp, dnx.params(len(cp) - 1), x0, cap=1500)
    b.run(64)
    st = b.poll()["status"][s]
    assert st in (dnx.SAT, dnx.UNKNOWN)
    hp.free()
    big = dnx.generate_planted(1, 50000, 400000)
    hp = dnx.host_problem(50000, planted_cp(400000), big)
    with pytest.raises(dnx.DnxError) as e:
        b.admit(hp, dnx.params(400000), dnx.init_uniform(1, 50000, 400000))
    assert e.value.code == dnx.E_BUDGET  # larger than the arena
    hp.free()
    b.close()


def test_admit_many_edge_instances(ctx):
    """One call admitting degenerate and mixed instances -- a single variable
    with a unit clause (pinned), k-SAT with an empty clause, unit pins among
    3-SAT clauses, an instance with no clauses -- computes the float oracle's
    bits, as single admissions do."""
    insts = []
    # one variable, one unit clause: V pinned, solved at once
    insts.append((1, np.array([0, 1], np.int32), np.array([1], np.int32)))
    # k-SAT with an empty clause (never solved) and a duplicated literal
    cl = [[1, -2, 3], [], [2, 2, -4], [4, 5], [-1, -3, -5, 2]]
    insts.append((5, np.cumsum([0] + [len(c) for c in cl]).astype(np.int32), np.array(sum(cl, []), np.int32)))
    # 3-SAT plus unit clauses (pins)
    lits = dnx.generate_planted(31, 200, 900)
    cl = [list(lits[3 * c:3 * c + 3]) for c in range(900)] + [[7], [-11]]
    insts.append((200, np.cumsum([0] + [len(c) for c in cl]).astype(np.int32), np.array(sum(cl, []), np.int32)))
    # variables but no clauses: the empty formula is satisfied at state 0
    insts.append((3, np.zeros(1, np.int32), np.zeros(0, np.int32)))
    full = [(n, cp, lits, dnx.init_uniform(500 + i, n, len(cp) - 1)) for i, (n, cp, lits) in enumerate(insts)]
    for many in (False, True):
        out = run_batch(ctx, full, 300, spl=16, many=many)
        for r, (n, cp, lits, x0) in zip(out, full):
            steps, t, status, xo = oracle_solve(n, cp, lits, x0, 300)
            assert r["status"] == (dnx.SAT if status == 1 else dnx.UNKNOWN)
            assert r["steps"] == steps and r["t"] == t
            assert same_bits(r["x"], xo)


def test_admit_many_all_or_none(ctx):
    """dnx_batch_admit_many admits all or nothing: with too few free slots or
    too little arena for the whole set it fails with DNX_E_BUDGET and leaves
    the slots and memory as they were; the set then fits after a retire."""
    inst = make_set(21, 6, nmin=200, nmax=1500)
    hps = [dnx.host_problem(n, cp, lits) for n, cp, lits, x0 in inst]
    x0s = [x0 for n, cp, lits, x0 in inst]
    b = dnx.Batch(ctx, slots=5, arena_bytes=64 << 20)
    try:
        arena, free0 = b.memory()
        with pytest.raises(dnx.DnxError) as e:
            b.admit_many(hps, x0s, cap=1500)
        assert e.value.code == dnx.E_BUDGET  # 6 instances, 5 slots
        assert b.memory()[1] == free0
        first = b.admit_many(hps[:1], x0s[:1], cap=1500)
        with pytest.raises(dnx.DnxError) as e:
            b.admit_many(hps[1:], x0s[1:], cap=1500)
        assert e.value.code == dnx.E_BUDGET  # 5 more, 4 free slots
        b.retire(first[0])
        assert b.memory()[1] == free0
        slots = b.admit_many(hps[1:], x0s[1:], cap=1500)
        assert len(set(slots)) == 5
        b.run(64)
        res = b.poll()
        for (n, cp, lits, x0), s in zip(inst[1:], slots):
            steps, t, status, xo = oracle_solve(n, cp, lits, x0, 1500)
            assert res["steps"][s] == steps and res["t"][s] == t
            assert same_bits(b.read_state(s, n, len(cp) - 1)["x"], xo)
    finally:
        b.close()
    small = dnx.Batch(ctx, slots=8, arena_bytes=2 << 20)
    try:
        free1 = small.memory()[1]
        with pytest.raises(dnx.DnxError) as e:
            small.admit_many(hps, x0s, cap=1500)
        assert e.value.code == dnx.E_BUDGET  # the set needs more than 2 MiB
        assert small.memory()[1] == free1
    finally:
        small.close()
        for hp in hps:
            hp.free()


def check_served(r, i, steps, t, status, xo, n):
    """Request i of a served trace against the instance solved alone: status,
    steps, simulated time, the final state bit for bit and the assignment."""
    assert r["status"][i] == (dnx.SAT if status == 1 else dnx.UNKNOWN), i
    assert r["steps"][i] == steps and r["t"][i] == t, i
    assert same_bits(r["x"][i], xo), i
    assert np.array_equal(r["assign"][i], np.where(np.asarray(xo[:n]) >= 0, 1, -1).astype(np.int8)), i
    assert r["tts"][i] > 0 and r["wait"][i] >= 0


@pytest.mark.parametrize("big", [1 << 40, 8000])
def test_serving_results_equal_solo_runs(ctx, inst_set, oracle_results, big):
    """Requests arriving while others run are admitted as slots and memory
    free up; every served result equals the instance solved alone (the float
    oracle: steps, time, final state bits, assignment), whether it ran in the
    narrow lane or (clauses >= big) in the wide lane."""
    k = 24
    probs = [dnx.host_problem(n, cp, lits) for n, cp, lits, _ in inst_set[:k]]
    x0s = [x0 for _, _, _, x0 in inst_set[:k]]
    arrivals = np.sort(np.random.default_rng(3).uniform(0.0, 0.25, k))
    # few slots and a small arena force queueing and slot reuse
    b = dnx.Batch(ctx, slots=6, arena_bytes=6 << 20)
    b.set_serve_big(big)
    r = b.serve(probs, x0s, arrivals, cap=1500, steps_per_launch=16, results=True)
    b.close()
    for p in probs:
        p.free()
    for i in range(k):
        steps, t, status, xo = oracle_results[i]
        check_served(r, i, steps, t, status, xo, inst_set[i][0])
    assert r["makespan"] >= arrivals[-1]
    assert np.max(r["wait"]) > 0  # some requests had to queue


@pytest.mark.parametrize("lane", ["sliced", "wide"])
def test_serving_above_default_wide_threshold(ctx, inst_set, oracle_results, lane):
    """A request of 63,000 clauses (above the 60,000-clause defaults) runs in
    the sliced lane by default, in the wide lane with the sliced lane off,
    while small requests arrive around it in the narrow lane; each result
    equals its solo run (the float oracle)."""
    n, m = 9000, 63000
    big = (n, planted_cp(m), dnx.generate_planted(6300, n, m), dnx.init_uniform(6300, n, m))
    insts = list(inst_set[:6]) + [big] + list(inst_set[6:12])
    probs = [dnx.host_problem(nn, cp, lits) for nn, cp, lits, _ in insts]
    arrivals = np.sort(np.random.default_rng(4).uniform(0.0, 0.1, len(insts)))
    b = dnx.Batch(ctx, slots=8)
    if lane == "wide":
        b.set_serve_sliced(0)
    r = b.serve(probs, [x0 for *_, x0 in insts], arrivals, cap=1500, steps_per_launch=16, results=True)
    lanes = b.served_by()
    b.close()
    assert lanes[lane] == 1 and lanes["narrow"] == len(insts) - 1
    for p in probs:
        p.free()
    ref = oracle_results[:6] + [oracle_solve(*big, 1500)] + oracle_results[6:12]
    for i, ((nn, *_), (steps, t, status, xo)) in enumerate(zip(insts, ref)):
        check_served(r, i, steps, t, status, xo, nn)
    assert r["status"][6] == dnx.SAT


def test_serving_sliced_lane_equals_solo_runs(ctx, inst_set, oracle_results):
    """3-SAT requests of at least 5,000 clauses run in the sliced lane (the
    single-instance path's sliced exchange, several requests per launch)
    while the others run in the narrow lane around them; every result equals
    the float oracle (steps, time, final state bits, assignment), and the
    lane counts say which lane retired each."""
    insts = list(inst_set[:16])
    probs = [dnx.host_problem(nn, cp, lits) for nn, cp, lits, _ in insts]
    arrivals = np.sort(np.random.default_rng(8).uniform(0.0, 0.05, len(insts)))
    b = dnx.Batch(ctx, slots=8)
    b.set_serve_sliced(5000, 2)
    r = b.serve(probs, [x0 for *_, x0 in insts], arrivals, cap=1500, steps_per_launch=16, results=True)
    lanes = b.served_by()
    b.close()
    for p in probs:
        p.free()
    for i, ((nn, *_), (steps, t, status, xo)) in enumerate(zip(insts, oracle_results[:16])):
        check_served(r, i, steps, t, status, xo, nn)
    expect = sum(1 for nn, cp, lits, _ in insts if len(cp) - 1 >= 5000 and len(lits) == 3 * (len(cp) - 1))
    print("lanes", lanes, "expected sliced", expect)
    assert expect > 0 and lanes["sliced"] == expect
    assert lanes["narrow"] + lanes["sliced"] == len(insts)


def unsat3(n, m, seed):
    """A planted 3-SAT instance made unsatisfiable by two 3-literal clauses,
    (x1 x1 x1) and (-x1 -x1 -x1): every clause still has three literals."""
    lits = dnx.generate_planted(seed, n, m - 2)
    lits = np.concatenate([lits, [1, 1, 1, -1, -1, -1]]).astype(np.int32)
    return n, planted_cp(m), lits, dnx.init_uniform(seed, n, m)


def unsolvable(n, m, seed):
    """A planted instance plus one empty clause: it never verifies, so it
    runs until its step cap."""
    lits = dnx.generate_planted(seed, n, m)
    cp = np.concatenate([planted_cp(m), [3 * m]]).astype(np.int32)
    return n, cp, lits, dnx.init_uniform(seed, n, m + 1)


def test_serve_deadline_reports_where(ctx):
    """Requests still unfinished at the deadline stay DNX_UNKNOWN with no
    time-to-solution, and report where they were: three in the narrow lane,
    two in the wide lane (a budget that holds one at a time, so one runs and
    one waits), one never admitted (no slot), one arriving after the
    deadline, two in the sliced lane (room for one: one runs, one waits); the
    steps they had completed come back too."""
    # narrow x3, sliced, wide x2, sliced, queued, late (arrival order)
    insts = [unsolvable(1500, 6000, 70 + i) for i in range(3)] + [unsat3(5000, 25000, 95)]
    insts += [unsolvable(2500, 9000, 80 + i) for i in range(2)] + [unsat3(5000, 25000, 96)]
    insts += [unsolvable(800, 3000, 90), unsolvable(800, 3000, 91)]
    arrivals = np.array([0.0, 0.0, 0.0, 0.0, 0.01, 0.01, 0.02, 0.05, 5.0])
    probs = [dnx.host_problem(n, cp, lits) for n, cp, lits, _ in insts]
    b = dnx.Batch(ctx, slots=5)
    b.set_serve_big(8000, 10000)
    b.set_serve_sliced(20000, 1)
    r = b.serve(probs, [x0 for *_, x0 in insts], arrivals, cap=10 ** 9, steps_per_launch=64, deadline=1.0)
    b.close()
    for p in probs:
        p.free()
    assert np.all(r["status"] == dnx.UNKNOWN) and np.all(r["tts"] == -1.0)
    w = r["where"]
    assert list(w[:3]) == [dnx.SERVE_NARROW] * 3
    assert set(w[4:6]) <= {dnx.SERVE_WIDE, dnx.SERVE_WIDE_WAIT} and dnx.SERVE_WIDE in set(w[4:6])
    assert w[3] == dnx.SERVE_SLICED and w[6] == dnx.SERVE_SLICED_WAIT  # room for one
    assert w[7] == dnx.SERVE_QUEUED and w[8] == dnx.SERVE_NOT_ARRIVED
    assert np.all(r["steps"][:3] > 0) and r["steps"][7] == 0 and r["steps"][8] == 0
    assert r["steps"][4:6].max() > 0 and r["steps"][3] > 0 and r["steps"][6] == 0
    print("where", w, "steps", r["steps"])


def test_serve_planted_stream(ctx):
    b = dnx.Batch(ctx, slots=64)
    r = b.serve_planted(seed=5, rate=200.0, duration=0.5, n_min=100, n_max=3000, cap=20000, steps_per_launch=32,
                        max_requests=1000, gen_threads=4)
    b.close()
    assert 60 <= r["count"] <= 150
    assert np.all(np.diff(r["arrival"]) >= 0)
    assert np.all((r["status"] == dnx.SAT) | (r["status"] == dnx.UNKNOWN))
    assert np.all(r["tts"] > 0)
    assert np.all(r["where"] == dnx.SERVE_ENDED)
    print("served", r["count"], "p50 tts ms", np.percentile(r["tts"], 50) * 1e3,
          "p99", np.percentile(r["tts"], 99) * 1e3)


def test_autotuner_applies_a_shape(ctx, inst_set, oracle_results):
    b = dnx.Batch(ctx, slots=64, arena_bytes=1 << 30)
    rep = b.tune(n_min=100, n_max=2000, sample=48, steps=30)
    assert rep.shape[1] == 5 and len(rep) >= 12
    assert np.all(rep[:, 3] > 0)
    bs0, cap0, wide = b.get_shape(0)
    assert bs0 in (256, 512, 1024) and wide >= 0
    gang_rows = rep[rep[:, 0] == 2]
    assert len(gang_rows) == 3 and set(gang_rows[:, 1]) == {256, 384, 768}
    assert b.get_gang()[1] in (256, 384, 768)
    print("tuned: mode0", b.get_shape(0), "mode1", b.get_shape(1), "gang", b.get_gang())
    # the tuned batch computes the same bits
    slots = []
    for n, cp, lits, x0 in inst_set[:10]:
        hp = dnx.host_problem(n, cp, lits)
        slots.append(b.admit(hp, dnx.params(len(cp) - 1), x0, cap=1500))
        hp.free()
    b.run(64)
    res = b.poll()
    for s, (steps, t, status, xo), (n, cp, lits, x0) in zip(slots, oracle_results[:10], inst_set[:10]):
        assert res["steps"][s] == steps and res["t"][s] == t
        assert same_bits(b.read_state(s, n, len(cp) - 1)["x"].astype(np.float32), xo)
    b.close()
