"""Python view of libdnx (the C ABI in include/dnx.h) over ctypes.

Used by the test suite and the benchmark to drive the HIP path.  There is no
fallback: when ``lib/libdnx.so`` is missing, or no GPU is visible, the calls
raise instead of computing anything on the CPU.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libdnx.so")
INCLUDE_H = os.path.join(os.path.dirname(HERE), "include", "dnx.h")

OK, E_ARG, E_IO, E_PARSE, E_TOOLONG, E_HIP, E_NOMEM, E_BUDGET, E_STATE = 0, -1, -2, -3, -4, -5, -6, -7, -8
EXCHANGE_LITERAL, EXCHANGE_GROUPED, EXCHANGE_SLICED = 0, 1, 2
FP64, FP32 = 0, 1
RUNNING, SAT, UNKNOWN, FAILED = 0, 1, 2, 3
SERVE_ENDED, SERVE_QUEUED, SERVE_NARROW, SERVE_WIDE, SERVE_WIDE_WAIT, SERVE_NOT_ARRIVED = 0, 1, 2, 3, 4, 5
SERVE_SLICED, SERVE_SLICED_WAIT = 6, 7

I32P = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
I8P = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
I64P = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
F64P = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class Params(C.Structure):
    _fields_ = [("alpha", C.c_double), ("beta", C.c_double), ("gamma", C.c_double),
                ("delta", C.c_double), ("epsilon", C.c_double), ("zeta", C.c_double),
                ("xl_max", C.c_int64)]


class DnxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dnx error {code}: {msg}")
        self.code = code


_LIB = None


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise DnxError(E_STATE, f"{LIB_PATH} is not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, ip, dp = C.c_void_p, C.c_int, C.c_double
    pp = C.POINTER(C.c_void_p)
    L.dnx_last_error.restype = C.c_char_p
    L.dnx_params_default.argtypes = [ip, ip, C.POINTER(Params)]
    L.dnx_open.argtypes = [ip, pp]
    L.dnx_close.argtypes = [vp]
    L.dnx_problem_from_dimacs.argtypes = [vp, C.c_char_p, pp]
    L.dnx_problem_from_lits.argtypes = [vp, ip, ip, I32P, I32P, pp]
    L.dnx_problem_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64)]
    L.dnx_problem_clauses.argtypes = [vp, I32P, I32P]
    L.dnx_problem_free.argtypes = [vp]
    L.dnx_rhs.argtypes = [vp, ip, C.POINTER(Params), F64P, F64P, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.dnx_euler.argtypes = [vp, ip, C.POINTER(Params), F64P, F64P, dp, F64P]
    L.dnx_rhs_device_noalloc.argtypes = [vp, ip, C.POINTER(Params), vp, vp, vp, vp, vp]
    L.dnx_state_create.argtypes = [vp, ip, C.POINTER(Params), F64P, pp]
    L.dnx_state_free.argtypes = [vp]
    L.dnx_state_step.argtypes = [vp, ip]
    L.dnx_state_read.argtypes = [vp, vp, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int)]
    L.dnx_state_solve.argtypes = [vp, C.c_int64, dp, vp, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                  C.POINTER(C.c_int)]
    L.dnx_solve.argtypes = [vp, ip, C.POINTER(Params), F64P, C.c_int64, dp, vp, C.POINTER(C.c_int64),
                            C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.dnx_init_reference.argtypes = [vp, ip, F64P]
    L.dnx_init_reference_thread.argtypes = [vp, ip, ip, F64P]
    L.dnx_portfolio_solve.argtypes = [vp, ip, C.POINTER(Params), ip, ip, C.c_int64, dp, vp, C.POINTER(C.c_int64),
                                      C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.dnx_generate_planted.argtypes = [C.c_uint64, ip, ip, dp, dp, I32P, vp]
    L.dnx_init_uniform.argtypes = [C.c_uint64, ip, ip, dp, dp, F64P]
    L.dnx_write_dimacs.argtypes = [C.c_char_p, ip, ip, I32P, I32P, C.c_char_p]
    L.dnx_check_assignment.argtypes = [vp, I8P, C.POINTER(C.c_int)]
    L.dnx_device_copy.argtypes = [vp, C.c_int64, ip, C.POINTER(C.c_double)]
    L.dnx_batch_open.argtypes = [vp, ip, ip, C.c_int64, pp]
    L.dnx_batch_close.argtypes = [vp]
    L.dnx_batch_set_shape.argtypes = [vp, ip, ip, ip]
    L.dnx_batch_admit.argtypes = [vp, vp, C.POINTER(Params), F64P, C.c_int64, ip, C.POINTER(C.c_int)]
    L.dnx_batch_admit_many.argtypes = [vp, ip, pp, vp, pp, C.c_int64, ip, I32P]
    L.dnx_batch_set_serve_order.argtypes = [vp, C.c_int64]
    L.dnx_batch_get_gang.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.dnx_batch_set_serve_sliced.argtypes = [vp, C.c_int64, ip]
    L.dnx_batch_served_by.argtypes = [vp, I64P]
    L.dnx_batch_advance.argtypes = [vp, ip]
    L.dnx_batch_poll.argtypes = [vp, vp, vp, vp, vp]
    L.dnx_batch_run.argtypes = [vp, ip, dp, C.POINTER(C.c_int)]
    L.dnx_batch_stamp.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]
    L.dnx_batch_read_state.argtypes = [vp, ip, vp, C.POINTER(C.c_int), C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    L.dnx_batch_assignment.argtypes = [vp, ip, I8P, C.POINTER(C.c_int)]
    L.dnx_batch_retire.argtypes = [vp, ip]
    L.dnx_batch_info.argtypes = [vp, ip, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.dnx_batch_memory.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.dnx_batch_timing.argtypes = [vp, ip, vp, vp, vp]
    L.dnx_batch_set_wide.argtypes = [vp, C.c_int64, ip]
    L.dnx_batch_set_serve_big.argtypes = [vp, C.c_int64, C.c_int64]
    L.dnx_batch_set_gang.argtypes = [vp, ip, ip, ip]
    L.dnx_batch_set_wide_fraction.argtypes = [vp, dp, ip]
    L.dnx_batch_set_tail_split.argtypes = [vp, ip]
    L.dnx_batch_set_order.argtypes = [vp, C.c_int64]
    L.dnx_batch_set_flow.argtypes = [vp, ip]
    L.dnx_batch_set_epoch.argtypes = [vp, C.c_int64, ip, ip]
    L.dnx_batch_set_gang_regroup.argtypes = [vp, ip]
    L.dnx_batch_tune.argtypes = [vp, ip, ip, dp, dp, ip, ip, C.c_uint64, vp, ip, C.POINTER(C.c_int)]
    L.dnx_batch_get_shape.argtypes = [vp, ip, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64)]
    L.dnx_state_time.argtypes = [vp, ip, C.POINTER(C.c_double)]
    L.dnx_serve_trace.argtypes = [vp, ip, vp, vp, vp, F64P, C.c_int64, ip, dp, vp, vp, vp, vp,
                                  C.POINTER(C.c_double), vp, vp, vp, vp]
    L.dnx_state_track_loc.argtypes = [vp, ip]
    L.dnx_problem_set_exchange.argtypes = [vp, ip]
    L.dnx_state_best_loc.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.dnx_serve_planted.argtypes = [vp, C.c_uint64, dp, dp, ip, ip, dp, dp, dp, dp, C.c_int64, ip, dp, ip, ip,
                                    C.POINTER(C.c_int), vp, vp, vp, vp, vp, vp, vp, C.POINTER(C.c_double), vp]
    _LIB = L
    return L


def check(rc):
    if rc != OK:
        raise DnxError(rc, lib().dnx_last_error().decode(errors="replace"))
    return rc


def declared_symbols():
    """Names of the functions include/dnx.h declares."""
    import re
    txt = open(INCLUDE_H).read()
    return sorted(set(re.findall(r"\b(dnx_[a-z0-9_]+)\s*\(", txt)))


def params(m, xl_max_flag=10000, **kw):
    p = Params()
    check(lib().dnx_params_default(m, xl_max_flag, C.byref(p)))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


# ---------------------------------------------------------------------------
# host utilities (no GPU needed)
# ---------------------------------------------------------------------------
def generate_planted(seed, n, m, p1=0.46, p2=0.46, with_planted=False):
    lits = np.zeros(3 * m, np.int32)
    planted = np.zeros(n, np.int8)
    check(lib().dnx_generate_planted(seed, n, m, p1, p2, lits, planted.ctypes.data_as(C.c_void_p)))
    return (lits, planted) if with_planted else lits


def init_uniform(seed, n, m, xs0=0.5, xl0=1.0):
    x = np.zeros(n + 2 * m, np.float64)
    check(lib().dnx_init_uniform(seed, n, m, xs0, xl0, x))
    return x


def host_problem(n, cp, lits):
    """A problem without a device copy (for batch admission and checks)."""
    h = C.c_void_p()
    cp = np.ascontiguousarray(cp, np.int32)
    lits = np.ascontiguousarray(lits, np.int32)
    if lits.size == 0:
        lits = np.zeros(1, np.int32)
    check(lib().dnx_problem_from_lits(None, int(n), len(cp) - 1, cp, lits, C.byref(h)))
    return Problem._wrap(None, h)


def write_dimacs(path, n, cp, lits, comment=""):
    m = len(cp) - 1
    check(lib().dnx_write_dimacs(path.encode(), n, m, np.ascontiguousarray(cp, np.int32),
                                 np.ascontiguousarray(lits, np.int32), comment.encode()))


# ---------------------------------------------------------------------------
# GPU objects
# ---------------------------------------------------------------------------
class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().dnx_open(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().dnx_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_copy(self, nbytes, repeats=1):
        """Mean device ms of one plain copy of nbytes (known-size reference)."""
        ms = C.c_double()
        check(lib().dnx_device_copy(self.h, int(nbytes), int(repeats), C.byref(ms)))
        return ms.value

    def problem_from_lits(self, n, cp, lits):
        return Problem(self, n, cp, lits)

    def problem_from_dimacs(self, path):
        h = C.c_void_p()
        check(lib().dnx_problem_from_dimacs(self.h, path.encode(), C.byref(h)))
        return Problem._wrap(self, h)


class Problem:
    def __init__(self, ctx, n, cp, lits):
        cp = np.ascontiguousarray(cp, np.int32)
        lits = np.ascontiguousarray(lits, np.int32)
        if lits.size == 0:
            lits = np.zeros(1, np.int32)
        h = C.c_void_p()
        check(lib().dnx_problem_from_lits(ctx.h, int(n), len(cp) - 1, cp, lits, C.byref(h)))
        self._init(ctx, h)

    @classmethod
    def _wrap(cls, ctx, h):
        o = cls.__new__(cls)
        o._init(ctx, h)
        return o

    def _init(self, ctx, h):
        self.ctx, self.h = ctx, h
        n, m, nnz = C.c_int(), C.c_int(), C.c_int64()
        check(lib().dnx_problem_info(h, C.byref(n), C.byref(m), C.byref(nnz)))
        self.n, self.m, self.nnz = n.value, m.value, nnz.value

    def set_exchange(self, exchange):
        """EXCHANGE_SLICED (default), EXCHANGE_GROUPED or EXCHANGE_LITERAL for states created afterwards."""
        check(lib().dnx_problem_set_exchange(self.h, int(exchange)))

    def clauses(self):
        cp = np.zeros(self.m + 1, np.int32)
        lits = np.zeros(max(self.nnz, 1), np.int32)
        check(lib().dnx_problem_clauses(self.h, cp, lits))
        return cp, lits[:self.nnz]

    def free(self):
        if self.h:
            lib().dnx_problem_free(self.h)
            self.h = None

    def rhs(self, prm, x, prec=FP64):
        x = np.ascontiguousarray(x, np.float64)
        d = np.zeros_like(x)
        loc, en = C.c_int(), C.c_double()
        check(lib().dnx_rhs(self.h, prec, C.byref(prm), x, d, C.byref(loc), C.byref(en)))
        return d, loc.value, en.value

    def rhs_device_noalloc(self, prm, x_ptr, dxdt_ptr, loc_ptr=None, energy_ptr=None, stream=None, prec=FP64):
        """dnx_rhs_device_noalloc: device addresses (ints) in the compute type;
        queued on `stream` (a hipStream_t address, None: the context's stream)."""
        check(lib().dnx_rhs_device_noalloc(self.h, prec, C.byref(prm), x_ptr, dxdt_ptr, loc_ptr, energy_ptr,
                                           stream))

    def euler(self, prm, x, dxdt, h, prec=FP64):
        x = np.ascontiguousarray(x, np.float64)
        out = np.zeros_like(x)
        check(lib().dnx_euler(self.h, prec, C.byref(prm), x, np.ascontiguousarray(dxdt, np.float64), h, out))
        return out

    def init_reference(self, seed=0):
        x = np.zeros(self.n + 2 * self.m, np.float64)
        check(lib().dnx_init_reference(self.h, seed, x))
        return x

    def init_reference_thread(self, seed, tid):
        x = np.zeros(self.n + 2 * self.m, np.float64)
        check(lib().dnx_init_reference_thread(self.h, seed, tid, x))
        return x

    def portfolio_solve(self, prm, seed=0, threads=4, prec=FP64, cap=100000, walltime=0.0):
        a = np.zeros(max(self.n, 1), np.int8)
        steps, t, st, w = C.c_int64(), C.c_double(), C.c_int(), C.c_int()
        check(lib().dnx_portfolio_solve(self.h, prec, C.byref(prm), seed, threads, cap, walltime,
                                        a.ctypes.data_as(C.c_void_p), C.byref(steps), C.byref(t), C.byref(st),
                                        C.byref(w)))
        return dict(status=st.value, steps=steps.value, t=t.value, assign=a[:self.n], winner=w.value)

    def check_assignment(self, assign):
        u = C.c_int()
        check(lib().dnx_check_assignment(self.h, np.ascontiguousarray(assign, np.int8), C.byref(u)))
        return u.value

    def solve(self, prm, x0, prec=FP64, cap=100000, walltime=0.0):
        a = np.zeros(max(self.n, 1), np.int8)
        steps, t, st = C.c_int64(), C.c_double(), C.c_int()
        check(lib().dnx_solve(self.h, prec, C.byref(prm), np.ascontiguousarray(x0, np.float64), cap, walltime,
                              a.ctypes.data_as(C.c_void_p), C.byref(steps), C.byref(t), C.byref(st)))
        return dict(status=st.value, steps=steps.value, t=t.value, assign=a[:self.n])

    def state(self, prm, x0, prec=FP64):
        return State(self, prm, x0, prec)


class State:
    def __init__(self, prob, prm, x0, prec):
        h = C.c_void_p()
        check(lib().dnx_state_create(prob.h, prec, C.byref(prm), np.ascontiguousarray(x0, np.float64), C.byref(h)))
        self.prob, self.h, self.prec = prob, h, prec

    def free(self):
        if self.h:
            lib().dnx_state_free(self.h)
            self.h = None

    def step(self, nsteps):
        check(lib().dnx_state_step(self.h, nsteps))

    def read(self):
        x = np.zeros(self.prob.n + 2 * self.prob.m, np.float64)
        t, steps, lev, loc = C.c_double(), C.c_int64(), C.c_int(), C.c_int()
        check(lib().dnx_state_read(self.h, x.ctypes.data_as(C.c_void_p), C.byref(t), C.byref(steps), C.byref(lev),
                                   C.byref(loc)))
        return dict(x=x, t=t.value, steps=steps.value, level=lev.value, loc=loc.value)

    def time_steps(self, nsteps):
        ms = C.c_double()
        check(lib().dnx_state_time(self.h, int(nsteps), C.byref(ms)))
        return ms.value

    def track_loc(self, enable=True):
        check(lib().dnx_state_track_loc(self.h, int(bool(enable))))

    def best_loc(self):
        """(lowest loc since tracking began, loc of the last evaluated state)"""
        b, l = C.c_int(), C.c_int()
        check(lib().dnx_state_best_loc(self.h, C.byref(b), C.byref(l)))
        return b.value, l.value

    def solve(self, cap=100000, walltime=0.0):
        a = np.zeros(max(self.prob.n, 1), np.int8)
        steps, t, st = C.c_int64(), C.c_double(), C.c_int()
        check(lib().dnx_state_solve(self.h, cap, walltime, a.ctypes.data_as(C.c_void_p), C.byref(steps),
                                    C.byref(t), C.byref(st)))
        return dict(status=st.value, steps=steps.value, t=t.value, assign=a[:self.prob.n])


class Batch:
    """Slots of independent instances advanced together (FP32)."""

    def __init__(self, ctx, slots, arena_bytes=0):
        h = C.c_void_p()
        check(lib().dnx_batch_open(ctx.h, FP32, int(slots), int(arena_bytes), C.byref(h)))
        self.h, self.slots = h, int(slots)

    def close(self):
        if self.h:
            lib().dnx_batch_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_shape(self, mode, block, lds_cap=0):
        check(lib().dnx_batch_set_shape(self.h, mode, block, lds_cap))

    def admit(self, prob, prm, x0, cap=100000, stop_on_solve=True):
        s = C.c_int()
        check(lib().dnx_batch_admit(self.h, prob.h, C.byref(prm), np.ascontiguousarray(x0, np.float64), int(cap),
                                    int(bool(stop_on_solve)), C.byref(s)))
        return s.value

    def admit_many(self, probs, x0s, cap=100000, stop_on_solve=True, prms=None):
        """dnx_batch_admit_many: the slots of probs[i] in order; prms None
        means each instance's defaults."""
        k = len(probs)
        x0s = [np.ascontiguousarray(x, np.float64) for x in x0s]
        if len(x0s) != k or (prms is not None and len(prms) != k):
            raise DnxError(E_ARG, "probs, x0s and prms differ in length")
        ph = (C.c_void_p * k)(*[p.h for p in probs])
        xh = (C.c_void_p * k)(*[x.ctypes.data for x in x0s])
        pa = (Params * k)(*prms) if prms is not None else None
        slots = np.zeros(k, np.int32)
        check(lib().dnx_batch_admit_many(self.h, k, ph, pa, xh, int(cap), int(bool(stop_on_solve)), slots))
        return [int(s) for s in slots]

    def advance(self, steps):
        check(lib().dnx_batch_advance(self.h, int(steps)))

    def run(self, steps_per_launch=256, walltime=0.0):
        r = C.c_int()
        check(lib().dnx_batch_run(self.h, int(steps_per_launch), float(walltime), C.byref(r)))
        return r.value

    def poll(self):
        S = self.slots
        st = np.zeros(S, np.int32)
        steps = np.zeros(S, np.int64)
        t = np.zeros(S, np.float64)
        td = np.zeros(S, np.uint64)
        check(lib().dnx_batch_poll(self.h, st.ctypes.data_as(C.c_void_p), steps.ctypes.data_as(C.c_void_p),
                                   t.ctypes.data_as(C.c_void_p), td.ctypes.data_as(C.c_void_p)))
        return dict(status=st, steps=steps, t=t, t_done=td)

    def stamp(self):
        tk, rate = C.c_uint64(), C.c_double()
        check(lib().dnx_batch_stamp(self.h, C.byref(tk), C.byref(rate)))
        return tk.value, rate.value

    def read_state(self, slot, n, m):
        x = np.zeros(n + 2 * m, np.float64)
        st, steps, t = C.c_int(), C.c_int64(), C.c_double()
        check(lib().dnx_batch_read_state(self.h, slot, x.ctypes.data_as(C.c_void_p), C.byref(st), C.byref(steps),
                                         C.byref(t)))
        return dict(x=x, status=st.value, steps=steps.value, t=t.value)

    def assignment(self, slot, n):
        a = np.zeros(max(n, 1), np.int8)
        u = C.c_int()
        check(lib().dnx_batch_assignment(self.h, slot, a, C.byref(u)))
        return a[:n], u.value

    def retire(self, slot):
        check(lib().dnx_batch_retire(self.h, slot))

    def info(self, slot):
        n, m, md = C.c_int(), C.c_int(), C.c_int()
        check(lib().dnx_batch_info(self.h, slot, C.byref(n), C.byref(m), C.byref(md)))
        return n.value, m.value, md.value

    def tune(self, n_min=100, n_max=5000, r_min=4.3, r_max=8.0, sample=512, steps=60, seed=1):
        """Autotune execution shapes; returns the measurements as rows."""
        rep = np.zeros(512, np.float64)
        cnt = C.c_int()
        check(lib().dnx_batch_tune(self.h, int(n_min), int(n_max), float(r_min), float(r_max), int(sample),
                                   int(steps), int(seed), rep.ctypes.data_as(C.c_void_p), rep.size, C.byref(cnt)))
        return rep[:min(cnt.value, rep.size)].reshape(-1, 5)

    def get_shape(self, mode):
        bs, cap, wide = C.c_int(), C.c_int(), C.c_int64()
        check(lib().dnx_batch_get_shape(self.h, int(mode), C.byref(bs), C.byref(cap), C.byref(wide)))
        return bs.value, cap.value, wide.value

    def get_gang(self):
        mb, cpw, k = C.c_int(), C.c_int(), C.c_int()
        check(lib().dnx_batch_get_gang(self.h, C.byref(mb), C.byref(cpw), C.byref(k)))
        return mb.value, cpw.value, k.value

    def set_gang(self, max_blocks=256, clauses_per_workgroup=384, steps_per_launch=512):
        check(lib().dnx_batch_set_gang(self.h, int(max_blocks), int(clauses_per_workgroup), int(steps_per_launch)))

    def set_serve_big(self, clauses, active_clauses=6000000):
        check(lib().dnx_batch_set_serve_big(self.h, int(clauses), int(active_clauses)))

    def set_serve_order(self, las_steps):
        check(lib().dnx_batch_set_serve_order(self.h, int(las_steps)))

    def set_serve_sliced(self, min_clauses, max_active=1024):
        check(lib().dnx_batch_set_serve_sliced(self.h, int(min_clauses), int(max_active)))

    def served_by(self):
        """Requests the last serving call retired per lane: narrow, wide, sliced, never runnable."""
        c = np.zeros(4, np.int64)
        check(lib().dnx_batch_served_by(self.h, c))
        return dict(narrow=int(c[0]), wide=int(c[1]), sliced=int(c[2]), too_large=int(c[3]))

    def set_wide(self, max_clauses, steps_per_sync=300):
        check(lib().dnx_batch_set_wide(self.h, int(max_clauses), int(steps_per_sync)))

    def set_tail_split(self, enable=True):
        check(lib().dnx_batch_set_tail_split(self.h, int(bool(enable))))

    def set_order(self, large_clauses=0):
        check(lib().dnx_batch_set_order(self.h, int(large_clauses)))

    def set_flow(self, enable=True):
        check(lib().dnx_batch_set_flow(self.h, int(bool(enable))))

    def set_epoch(self, min_clauses=6000, steps=1024, clauses_per_workgroup=768):
        """Elastic tail: spare CUs serve the largest running instances (at least min_clauses clauses) as
        gangs of at most clauses_per_workgroup clauses per workgroup beside the single-workgroup ones, in
        epochs of `steps` gang steps; min_clauses 0 turns it off."""
        check(lib().dnx_batch_set_epoch(self.h, int(min_clauses), int(steps), int(clauses_per_workgroup)))

    def set_gang_regroup(self, enable=True):
        """Epoch gangs of 3-SAT instances regroup their contributions by folding workgroup (True) or
        exchange them through the transpose."""
        check(lib().dnx_batch_set_gang_regroup(self.h, int(bool(enable))))

    def set_wide_fraction(self, fraction=0.0125, steps_per_sync=300):
        check(lib().dnx_batch_set_wide_fraction(self.h, float(fraction), int(steps_per_sync)))

    def timing(self, enable=-1, with_bytes=False):
        """Summed device ms and launch counts per fused-kernel mode (0-2), wide mode (3), gang mode (4),
        the fused kernel's tail launch beside wide mode (5) and epoch launches (6); with_bytes adds the
        compulsory bytes of the steps those launches took."""
        ms = np.zeros(7, np.float64)
        nl = np.zeros(7, np.int64)
        ab = np.zeros(7, np.float64)
        check(lib().dnx_batch_timing(self.h, int(enable), ms.ctypes.data_as(C.c_void_p), nl.ctypes.data_as(C.c_void_p),
                                     ab.ctypes.data_as(C.c_void_p)))
        return (ms, nl, ab) if with_bytes else (ms, nl)

    def memory(self):
        a, f = C.c_int64(), C.c_int64()
        check(lib().dnx_batch_memory(self.h, C.byref(a), C.byref(f)))
        return a.value, f.value

    def serve(self, probs, x0s, arrivals, cap=100000, steps_per_launch=64, deadline=0.0, results=False):
        """Continuous batching over a request trace; see dnx_serve_trace.  With results=True each request's
        simulated time, sign assignment and final state come back too (t, assign, x)."""
        k = len(probs)
        ph = (C.c_void_p * max(k, 1))(*[p.h for p in probs])
        xs = [np.ascontiguousarray(x, np.float64) for x in x0s]
        xp = (C.c_void_p * max(k, 1))(*[x.ctypes.data for x in xs])
        arr = np.ascontiguousarray(arrivals, np.float64)
        st = np.zeros(max(k, 1), np.int32)
        steps = np.zeros(max(k, 1), np.int64)
        tts = np.zeros(max(k, 1), np.float64)
        wait = np.zeros(max(k, 1), np.float64)
        tt = np.zeros(max(k, 1), np.float64)
        wh = np.zeros(max(k, 1), np.int32)
        mk = C.c_double()
        asg = [np.zeros(max(p.n, 1), np.int8) for p in probs] if results else []
        xo = [np.zeros(max(p.n + 2 * p.m, 1), np.float64) for p in probs] if results else []
        ap = (C.c_void_p * max(k, 1))(*[a.ctypes.data for a in asg]) if results else None
        xop = (C.c_void_p * max(k, 1))(*[x.ctypes.data for x in xo]) if results else None
        check(lib().dnx_serve_trace(self.h, k, ph, None, xp, arr if k else np.zeros(1), int(cap), int(steps_per_launch),
                                    float(deadline), st.ctypes.data_as(C.c_void_p), steps.ctypes.data_as(C.c_void_p),
                                    tts.ctypes.data_as(C.c_void_p), wait.ctypes.data_as(C.c_void_p), C.byref(mk),
                                    tt.ctypes.data_as(C.c_void_p), ap, xop, wh.ctypes.data_as(C.c_void_p)))
        out = dict(status=st[:k], steps=steps[:k], tts=tts[:k], wait=wait[:k], makespan=mk.value, t=tt[:k],
                   where=wh[:k])
        if results:
            out["assign"] = [a[:p.n] for a, p in zip(asg, probs)]
            out["x"] = [x[:p.n + 2 * p.m] for x, p in zip(xo, probs)]
        return out

    def serve_planted(self, seed, rate, duration, n_min=100, n_max=100000, r_min=4.3, r_max=8.0, p1=0.46, p2=0.46,
                      cap=100000, steps_per_launch=64, max_requests=100000, gen_threads=8, deadline=0.0):
        M = int(max_requests)
        cnt = C.c_int()
        arr = np.zeros(M, np.float64)
        ns = np.zeros(M, np.int32)
        ms = np.zeros(M, np.int32)
        st = np.zeros(M, np.int32)
        steps = np.zeros(M, np.int64)
        tts = np.zeros(M, np.float64)
        wait = np.zeros(M, np.float64)
        wh = np.zeros(M, np.int32)
        mk = C.c_double()
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        check(lib().dnx_serve_planted(self.h, int(seed), float(rate), float(duration), int(n_min), int(n_max),
                                      float(r_min), float(r_max), float(p1), float(p2), int(cap),
                                      int(steps_per_launch), float(deadline), M, int(gen_threads), C.byref(cnt),
                                      ptr(arr), ptr(ns),
                                      ptr(ms), ptr(st), ptr(steps), ptr(tts), ptr(wait), C.byref(mk), ptr(wh)))
        k = cnt.value
        return dict(count=k, arrival=arr[:k], n=ns[:k], m=ms[:k], status=st[:k], steps=steps[:k], tts=tts[:k],
                    wait=wait[:k], makespan=mk.value, where=wh[:k])
