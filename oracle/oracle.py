"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: this module is imported by tests/, by
``__graft_entry__.smoke()`` as the checker, and by ``bench.py``'s
``cpu_baseline`` leg, which times it as the reference's CPU path.  The product
library never calls it.

The oracle restates the reference integrator (dynex.cc) in strict-IEEE C; the
file-by-line mapping is in ``dnx_oracle.c``'s header.  Problems are passed as a
CSR clause table ``(n, m, cp[m+1], lits[nnz])`` with signed 1-based literals;
states in the reference layout ``x = [V(n) | Xs(m) | Xl(m)]``.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
I8P = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
F32P = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


class Params(C.Structure):
    _fields_ = [("alpha", C.c_double), ("beta", C.c_double), ("gamma", C.c_double),
                ("delta", C.c_double), ("epsilon", C.c_double), ("zeta", C.c_double),
                ("xl_max", C.c_longlong)]


def lib():
    """Load liboracle.so, building it with make when it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        import subprocess
        subprocess.run(["make", "-C", _HERE, "-s"], check=True)
    L = C.CDLL(path)
    for sfx, FP in (("f64", F64P), ("f32", F32P)):
        getattr(L, "orc_rhs_" + sfx).argtypes = [C.c_int, C.c_int, I32P, I32P, C.POINTER(Params), FP, FP,
                                                 C.POINTER(C.c_double)]
        getattr(L, "orc_rhs_" + sfx).restype = C.c_int
        ft = C.c_double if sfx == "f64" else C.c_float
        getattr(L, "orc_dostep_" + sfx).argtypes = [C.c_int, C.c_int, C.c_void_p, C.POINTER(Params), FP, FP, ft, FP]
        getattr(L, "orc_verify_" + sfx).argtypes = [C.c_int, C.c_int, I32P, I32P, FP]
        getattr(L, "orc_verify_" + sfx).restype = C.c_int
        getattr(L, "orc_step_" + sfx).argtypes = [C.c_int, C.c_int, I32P, I32P, C.c_void_p, C.POINTER(Params), FP,
                                                  C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                  C.POINTER(C.c_int), C.c_int]
        getattr(L, "orc_step_" + sfx).restype = C.c_int
        getattr(L, "orc_solve_" + sfx).argtypes = [C.c_int, C.c_int, I32P, I32P, C.c_void_p, C.POINTER(Params), FP,
                                                   C.c_longlong, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        getattr(L, "orc_solve_" + sfx).restype = C.c_longlong
        getattr(L, "orc_trajectory_" + sfx).argtypes = [C.c_int, C.c_int, I32P, I32P, C.c_void_p, C.POINTER(Params),
                                                        FP, C.c_int, C.POINTER(C.c_double), I32P, I32P]
    L.orc_dimacs_header.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_longlong)]
    L.orc_dimacs_read.argtypes = [C.c_char_p, C.c_int, C.c_int, I32P, I32P]
    L.orc_xl_max.argtypes = [C.c_int, C.c_int]
    L.orc_xl_max.restype = C.c_longlong
    L.orc_unit_vars.argtypes = [C.c_int, C.c_int, I32P, I32P, I8P]
    L.orc_init_ref.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, F64P]
    L.orc_uniform_int_stream.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, I32P]
    L.orc_mt19937_nth.argtypes = [C.c_uint32, C.c_int]
    L.orc_mt19937_nth.restype = C.c_uint32
    _LIB = L
    return L


class Problem:
    """CSR clause table as held by the oracle."""

    def __init__(self, n, m, cp, lits):
        self.n, self.m = int(n), int(m)
        self.cp = np.ascontiguousarray(cp, dtype=np.int32)
        self.lits = np.ascontiguousarray(lits, dtype=np.int32)
        self.unit = np.zeros(self.n, dtype=np.int8)
        if self.n:
            lib().orc_unit_vars(self.n, self.m, self.cp, self.lits, self.unit)
        self.has_unit = bool(self.unit.any())

    @classmethod
    def from_dimacs(cls, path):
        L = lib()
        n, m, nnz = C.c_int(), C.c_int(), C.c_longlong()
        rc = L.orc_dimacs_header(path.encode(), C.byref(n), C.byref(m), C.byref(nnz))
        if rc:
            raise ValueError(f"oracle DIMACS reader: code {rc}")
        cp = np.zeros(m.value + 1, np.int32)
        lits = np.zeros(max(nnz.value, 1), np.int32)
        rc = L.orc_dimacs_read(path.encode(), n.value, m.value, cp, lits)
        if rc:
            raise ValueError(f"oracle DIMACS reader: code {rc}")
        return cls(n.value, m.value, cp, lits[:nnz.value])

    def _unit_ptr(self):
        return self.unit.ctypes.data_as(C.c_void_p) if self.has_unit else None


def params(m, xl_max_flag=10000, **kw):
    p = Params(2.5, 20.0, 25.0 / 100.0, 5.0 / 100.0, 1.0 / 10.0, 1.0 / 10.0, lib().orc_xl_max(xl_max_flag, m))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _dt(prec):
    return (np.float64, "f64") if prec == 64 else (np.float32, "f32")


def rhs(prob, prm, x, prec=64):
    """dmm_generate_dxdt: returns (dxdt, loc, energy)."""
    dt, sfx = _dt(prec)
    x = np.ascontiguousarray(x, dtype=dt)
    d = np.zeros_like(x)
    en = C.c_double()
    loc = getattr(lib(), "orc_rhs_" + sfx)(prob.n, prob.m, prob.cp, prob.lits, C.byref(prm), x, d, C.byref(en))
    return d, loc, en.value


def dostep(prob, prm, x, dxdt, h, prec=64):
    dt, sfx = _dt(prec)
    x = np.ascontiguousarray(x, dtype=dt)
    dxdt = np.ascontiguousarray(dxdt, dtype=dt)
    out = np.zeros_like(x)
    getattr(lib(), "orc_dostep_" + sfx)(prob.n, prob.m, prob._unit_ptr(), C.byref(prm), x, dxdt, h, out)
    return out


def verify(prob, x, prec=64):
    dt, sfx = _dt(prec)
    return getattr(lib(), "orc_verify_" + sfx)(prob.n, prob.m, prob.cp, prob.lits, np.ascontiguousarray(x, dt))


def step(prob, prm, x, t=0.0, prec=64, stop_on_solve=False):
    """One mode-6 iteration in place; returns (t, level, loc, vbad, solved)."""
    dt, sfx = _dt(prec)
    assert x.dtype == dt and x.flags.c_contiguous
    tt, lev, loc, bad = C.c_double(t), C.c_int(), C.c_int(), C.c_int()
    s = getattr(lib(), "orc_step_" + sfx)(prob.n, prob.m, prob.cp, prob.lits, prob._unit_ptr(), C.byref(prm), x,
                                         C.byref(tt), C.byref(lev), C.byref(loc), C.byref(bad), int(stop_on_solve))
    return tt.value, lev.value, loc.value, bad.value, bool(s)


def trajectory(prob, prm, x, nsteps, prec=64):
    """nsteps fixed ladder steps in place; returns (t, levels, locs)."""
    dt, sfx = _dt(prec)
    assert x.dtype == dt and x.flags.c_contiguous
    lev = np.zeros(max(nsteps, 1), np.int32)
    loc = np.zeros(max(nsteps, 1), np.int32)
    tt = C.c_double(0.0)
    getattr(lib(), "orc_trajectory_" + sfx)(prob.n, prob.m, prob.cp, prob.lits, prob._unit_ptr(), C.byref(prm), x,
                                           nsteps, C.byref(tt), lev, loc)
    return tt.value, lev[:nsteps], loc[:nsteps]


def solve(prob, prm, x, cap, prec=64):
    """dmm() mode-6 loop in place; returns (steps, t, status) with status 1 SAT, 2 cap."""
    dt, sfx = _dt(prec)
    assert x.dtype == dt and x.flags.c_contiguous
    tt, st = C.c_double(), C.c_int()
    steps = getattr(lib(), "orc_solve_" + sfx)(prob.n, prob.m, prob.cp, prob.lits, prob._unit_ptr(), C.byref(prm), x,
                                              cap, C.byref(tt), C.byref(st))
    return steps, tt.value, st.value


def init_ref(prob, seed):
    x = np.zeros(prob.n + 2 * prob.m, np.float64)
    lib().orc_init_ref(prob.n, prob.m, prob._unit_ptr(), seed, x)
    return x


def uniform_int_stream(seed, a, b, count):
    out = np.zeros(count, np.int32)
    lib().orc_uniform_int_stream(seed, a, b, count, out)
    return out


def mt19937_nth(seed, count):
    return lib().orc_mt19937_nth(seed, count)
