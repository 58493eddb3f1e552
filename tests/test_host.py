"""CPU-side tests of libdnx: the library loads and exports every declared
symbol, and its host logic (DIMACS reader, reference initial state, planted
generator, exact clause check, DIMACS writer, parameters) agrees with the
oracle and with the generator's stated distribution.  No GPU calls."""
import ctypes as C
import subprocess

import numpy as np
import pytest

import dnx
import oracle as O
from conftest import CNF_NAMES, cnf_path, planted_cp

ROOT_BIN = dnx.os.path.join(dnx.HERE, "lib", "dynex")


def host_problem_dimacs(path):
    h = C.c_void_p()
    dnx.check(dnx.lib().dnx_problem_from_dimacs(None, path.encode(), C.byref(h)))
    return dnx.Problem._wrap(None, h)


def host_problem(n, cp, lits):
    h = C.c_void_p()
    lits = np.ascontiguousarray(lits, np.int32)
    if lits.size == 0:
        lits = np.zeros(1, np.int32)
    dnx.check(dnx.lib().dnx_problem_from_lits(None, n, len(cp) - 1, np.ascontiguousarray(cp, np.int32), lits,
                                              C.byref(h)))
    return dnx.Problem._wrap(None, h)


def test_library_exports_declared_symbols():
    names = dnx.declared_symbols()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", dnx.LIB_PATH], check=True, capture_output=True,
                         text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [s for s in names if s not in exported]
    assert not missing, missing
    L = dnx.lib()
    for s in names:
        assert hasattr(L, s)


def test_params_match_reference_defaults():
    for m in (1, 800, 80000, 214748, 214749, 430000, 800000):
        p = dnx.params(m)
        assert (p.alpha, p.beta, p.gamma, p.delta, p.epsilon, p.zeta) == (2.5, 20.0, 0.25, 0.05, 0.1, 0.1)
        assert p.xl_max == O.lib().orc_xl_max(10000, m)


@pytest.mark.parametrize("nv", sorted(CNF_NAMES))
def test_dimacs_reader_matches_oracle(cnf_dir, nv):
    path = cnf_path(cnf_dir, nv)
    p = host_problem_dimacs(path)
    q = O.Problem.from_dimacs(path)
    cp, lits = p.clauses()
    assert (p.n, p.m) == (q.n, q.m)
    assert np.array_equal(cp, q.cp) and np.array_equal(lits, q.lits)
    p.free()


def test_dimacs_reader_edge_cases(tmp_path):
    f = tmp_path / "e.cnf"
    f.write_text("c hello\nc key=0101\np cnf 5 4\n1 -2 0\nc inside\n3\n 4 5 0 -1 0\n2 3 4 5 -1 0\n")
    p = host_problem_dimacs(str(f))
    q = O.Problem.from_dimacs(str(f))
    cp, lits = p.clauses()
    assert np.array_equal(cp, q.cp) and np.array_equal(lits, q.lits)
    g = tmp_path / "long.cnf"
    g.write_text("p cnf 30 1\n" + " ".join(str(i) for i in range(1, 27)) + " 0\n")
    with pytest.raises(dnx.DnxError) as e:
        host_problem_dimacs(str(g))
    assert e.value.code == dnx.E_TOOLONG
    with pytest.raises(dnx.DnxError) as e:
        host_problem_dimacs(str(tmp_path / "missing.cnf"))
    assert e.value.code == dnx.E_IO
    h = tmp_path / "range.cnf"
    h.write_text("p cnf 3 1\n1 2 7 0\n")
    with pytest.raises(dnx.DnxError) as e:
        host_problem_dimacs(str(h))
    assert e.value.code == dnx.E_PARSE


def test_from_lits_rejects_bad_tables():
    with pytest.raises(dnx.DnxError):
        host_problem(3, [0, 3], [1, 0, 2])
    with pytest.raises(dnx.DnxError):
        host_problem(3, [0, 3], [1, 2, 4])


@pytest.mark.parametrize("seed", [0, 1, 7, 42, -1, -2, 99991])
def test_init_reference_matches_oracle(seed):
    cp = np.array([0, 3, 6, 7, 9], np.int32)
    lits = np.array([1, 2, 3, -4, 5, 6, -2, 3, -6], np.int32)
    p = host_problem(6, cp, lits)
    q = O.Problem(6, 4, cp, lits)
    assert np.array_equal(p.init_reference(seed), O.init_ref(q, seed))


@pytest.mark.parametrize("seed", [0, 5, 1234, -1, -2])
def test_init_reference_thread_blocks(seed):
    """dynex.cc:2643-2700: one mt19937(seed) stream serves all threads in
    thread order; with seed 0 threads 0-2 are constant (1, 0, -1 over every
    entry) and draw nothing; every other thread takes the next n draws of
    uniform_int(-1, 1) (pinned to libstdc++ in test_oracle) as V, Xs = 0,
    Xl = 1; seeds -1/-2 override V; unit clauses pin V last."""
    n, m = 6, 3
    cp = np.array([0, 3, 6, 7], np.int32)
    lits = np.array([1, 2, 3, -4, 5, 6, -2], np.int32)  # var 2 pinned to -1
    p = host_problem(n, cp, lits)
    threads = 9
    draws = O.uniform_int_stream(seed, -1, 1, n * threads).astype(float)
    block = 0
    for tid in range(threads):
        x = p.init_reference_thread(seed, tid)
        if seed == 0 and tid <= 2:
            want = np.full(n + 2 * m, (1.0, 0.0, -1.0)[tid])
        else:
            want = np.concatenate([draws[block * n:(block + 1) * n], np.zeros(m), np.ones(m)])
            block += 1
        if seed == -1:
            want[:n] = 0.0
        if seed == -2:
            want[:n] = -1.0
        want[1] = -1.0
        assert np.array_equal(x, want), tid
    assert np.array_equal(p.init_reference_thread(seed, 0), p.init_reference(seed))


def test_device_entry_points_reject_host_only_problem():
    p = host_problem(3, [0, 3], [1, 2, 3])
    x = np.zeros(5)
    with pytest.raises(dnx.DnxError) as e:
        p.rhs(dnx.params(1), x)
    assert e.value.code == dnx.E_STATE


def test_generator_distribution():
    n, m = 2000, 40000
    lits, planted = dnx.generate_planted(7, n, m, with_planted=True)
    L = lits.reshape(m, 3)
    v = np.abs(L)
    assert v.min() >= 1 and v.max() <= n
    assert np.all((v[:, 0] != v[:, 1]) & (v[:, 0] != v[:, 2]) & (v[:, 1] != v[:, 2]))
    tru = (np.sign(L) == planted[v - 1]).sum(axis=1)
    assert tru.min() >= 1  # planted assignment satisfies every clause
    frac = np.bincount(tru, minlength=4)[1:] / m
    # binomial standard error at m=40000 is ~0.0025
    assert abs(frac[0] - 0.46) < 0.012 and abs(frac[1] - 0.46) < 0.012 and abs(frac[2] - 0.08) < 0.008
    # which slot is true is uniform; variables are uniform
    one = L[tru == 1]
    slot = np.argmax(np.sign(one) == planted[np.abs(one) - 1], axis=1)
    assert np.all(np.abs(np.bincount(slot, minlength=3) / len(slot) - 1 / 3) < 0.02)
    deg = np.bincount(v.ravel(), minlength=n + 1)[1:]
    assert abs(deg.mean() - 3 * m / n) < 1e-9 and deg.std() < 2 * np.sqrt(3 * m / n)
    assert abs((planted > 0).mean() - 0.5) < 0.05
    # deterministic per seed, different across seeds
    assert np.array_equal(lits, dnx.generate_planted(7, n, m))
    assert not np.array_equal(lits, dnx.generate_planted(8, n, m))


def test_check_assignment_and_writer(tmp_path):
    n, m = 300, 2400
    lits, planted = dnx.generate_planted(3, n, m, with_planted=True)
    cp = planted_cp(m)
    p = host_problem(n, cp, lits)
    assert p.check_assignment(planted) == 0
    flipped = -planted
    q = O.Problem(n, m, cp, lits)
    assert p.check_assignment(flipped) == O.verify(q, np.concatenate([flipped.astype(float), np.zeros(2 * m)]))
    f = tmp_path / "w.cnf"
    dnx.write_dimacs(str(f), n, cp, lits, "planted seed 3")
    r = host_problem_dimacs(str(f))
    cp2, lits2 = r.clauses()
    assert np.array_equal(cp2, cp) and np.array_equal(lits2, lits)


def test_init_uniform():
    x = dnx.init_uniform(5, 1000, 3000)
    assert np.all((x[:1000] >= -1) & (x[:1000] < 1)) and abs(x[:1000].mean()) < 0.1
    assert np.all(x[1000:4000] == 0.5) and np.all(x[4000:] == 1.0)
    assert np.array_equal(x, dnx.init_uniform(5, 1000, 3000))


def test_cli_usage_without_input():
    r = subprocess.run([ROOT_BIN], capture_output=True, text=True)
    assert r.returncode != 0 and "Syntax" in r.stdout


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus above 1 (or a multi-rank launch)
    exits with a message before any device work."""
    import os
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for args, env in (["--gpus", "2"], {}), ([], {"WORLD_SIZE": "2"}):
        r = subprocess.run([sys.executable, os.path.join(root, "bench.py")] + args, capture_output=True, text=True,
                           timeout=120, env={**os.environ, **env})
        assert r.returncode == 2
        assert "one GPU" in r.stderr and r.stdout == ""
