"""dnx_rhs_device_noalloc (the RHS on caller-owned device buffers, its scratch
in the library's fixed device array) against dnx_rhs, which allocates its
scratch per call: dxdt, loc and energy bit-identical, FP64 and FP32, on two
calls in a row with the second problem smaller than the first (the array
keeps the first call's contents), on the context's stream and on a stream of
the caller's.  Device buffers come from the HIP runtime libdnx itself uses."""
import ctypes as C
import os

import numpy as np
import pytest

import dnx
from conftest import planted_cp, random_state

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


@pytest.fixture(scope="module")
def ctx():
    c = dnx.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def hip(ctx):
    dnx.lib()
    h = C.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD | os.RTLD_LAZY)  # the runtime libdnx loaded
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    return h


class DevArray:
    def __init__(self, hip, nbytes):
        self.hip, self.ptr = hip, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), max(nbytes, 8)) == 0
        self.addr = self.ptr.value

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert self.hip.hipMemcpy(self.addr, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, dtype, count):
        out = np.zeros(count, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.addr, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.addr)


def mixed_problem(seed=7):
    """k in {0..5}, a duplicated literal, an empty clause."""
    rng = np.random.default_rng(seed)
    n, cls = 50, []
    for _ in range(200):
        k = int(rng.choice([1, 2, 3, 4, 5], p=[0.05, 0.25, 0.4, 0.2, 0.1]))
        vs = rng.choice(np.arange(1, n + 1), k, replace=False)
        cls.append(list(vs * rng.choice([-1, 1], k)))
    cls.append([4, 4, -9])
    cls.insert(33, [])
    cp = np.cumsum([0] + [len(c) for c in cls]).astype(np.int32)
    return n, cp, np.array(sum(cls, []), np.int32)


def run_noalloc(hip, p, prm, x, prec, stream=None, totals=True):
    dt = np.float64 if prec == dnx.FP64 else np.float32
    N = len(x)
    bx, bd, bt = DevArray(hip, N * 8), DevArray(hip, N * 8), DevArray(hip, 16)
    try:
        bx.put(x.astype(dt))
        bd.put(np.full(N, np.nan, dt))
        bt.put(np.array([-1.0, -1.0]))  # [energy | loc, pad]
        p.rhs_device_noalloc(prm, bx.addr, bd.addr, (bt.addr + 8) if totals else None,
                             bt.addr if totals else None, stream, prec)
        if stream is None:
            assert hip.hipDeviceSynchronize() == 0
        else:
            assert hip.hipStreamSynchronize(stream) == 0
        d = bd.get(dt, N).astype(np.float64)
        raw = bt.get(np.float64, 2)
        energy = float(raw[0])
        loc = int(raw[1:].view(np.int32)[0])
        return d, loc, energy
    finally:
        bx.free(), bd.free(), bt.free()


def check_against_rhs(hip, p, prm, x, prec, stream=None):
    d0, loc0, en0 = p.rhs(prm, x, prec)
    d1, loc1, en1 = run_noalloc(hip, p, prm, x, prec, stream)
    print(f"noalloc n={p.n} m={p.m} prec={prec} loc={loc1} energy={en1!r}", flush=True)
    assert np.array_equal(d0.view(np.uint64), d1.view(np.uint64))
    assert loc1 == loc0
    assert np.float64(en1).view(np.uint64) == np.float64(en0).view(np.uint64)


@pytest.mark.parametrize("prec", [dnx.FP64, dnx.FP32])
def test_noalloc_matches_rhs(ctx, hip, prec):
    rng = np.random.default_rng(11 + prec)
    n1, m1 = 20000, 85000
    big = ctx.problem_from_lits(n1, planted_cp(m1), dnx.generate_planted(5, n1, m1))
    n2, cp2, lits2 = mixed_problem()
    small = ctx.problem_from_lits(n2, cp2, lits2)
    try:
        for p in (big, small):  # second call smaller than the first
            prm = dnx.params(p.m)
            x = random_state(rng, p.n, p.m, xl_hi=50.0)
            check_against_rhs(hip, p, prm, x, prec)
        s = C.c_void_p()
        assert hip.hipStreamCreate(C.byref(s)) == 0
        try:
            for p in (big, small):  # the same on a stream of the caller's
                prm = dnx.params(p.m)
                x = random_state(rng, p.n, p.m)
                check_against_rhs(hip, p, prm, x, prec, s.value)
        finally:
            hip.hipStreamDestroy(s.value)
        # without loc/energy outputs dxdt is the same
        prm = dnx.params(small.m)
        x = random_state(rng, small.n, small.m)
        d0, _, _ = small.rhs(prm, x, prec)
        d1, _, _ = run_noalloc(hip, small, prm, x, prec, totals=False)
        assert np.array_equal(d0.view(np.uint64), d1.view(np.uint64))
    finally:
        big.free()
        small.free()


def test_noalloc_rejects_oversize(ctx, hip):
    """nnz = 9M: 72 MB of FP64 contributions exceed the 64 MiB array (refused
    before anything is queued); the FP32 form (36 MB) fits and matches."""
    n, m = 1_000_000, 3_000_000
    p = ctx.problem_from_lits(n, planted_cp(m), dnx.generate_planted(9, n, m))
    N = n + 2 * m
    bx, bd = DevArray(hip, N * 8), DevArray(hip, N * 8)
    try:
        prm = dnx.params(m)
        with pytest.raises(dnx.DnxError) as ei:
            p.rhs_device_noalloc(prm, bx.addr, bd.addr, None, None, None, dnx.FP64)
        assert ei.value.code == dnx.E_BUDGET
        assert "scratch" in str(ei.value)
        bx.free(), bd.free()
        bx = bd = None
        x = random_state(np.random.default_rng(3), n, m)
        check_against_rhs(hip, p, prm, x, dnx.FP32)
    finally:
        for b in (bx, bd):
            if b is not None:
                b.free()
        p.free()
