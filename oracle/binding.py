"""ctypes binding of oracle/liboracle.so (the CPU FP64 oracle).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and
the CPU-baseline leg of bench.py.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from yuzen_amd.capi import YzCase, YzCzCase, YzTerm

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    _lib = load_from(LIB)
    return _lib


def set_fault(k, m=1):
    """the oracle's fault injection: the film's Newton-pass solves k ..
    k + m - 1 from now break down before they iterate (k = 0: off)"""
    load().yzo_set_fault(int(k), int(m))


def threads(n=0):
    """the OpenMP threads of the oracle's row loops: n > 0 sets them;
    returns the number in force"""
    return load().yzo_threads(int(n))


def load_from(path):
    """the oracle's C API from a library at `path` (the tests load copies
    built with a planted mistake this way)"""
    lib = C.CDLL(path)
    P = C.POINTER
    pd, pi_ = P(C.c_double), P(C.c_int)
    lib.yzo_create.restype = C.c_void_p
    lib.yzo_create.argtypes = [P(YzCase), pd, pd, pd, C.c_double]
    lib.yzo_destroy.argtypes = [C.c_void_p]
    lib.yzo_set_lin_mode.argtypes = [C.c_void_p, C.c_int]
    lib.yzo_n.argtypes = [C.c_void_p]
    lib.yzo_threads.argtypes = [C.c_int]
    lib.yzo_set_fault.argtypes = [C.c_long, C.c_long]
    lib.yzo_nnz.argtypes = [C.c_void_p]
    lib.yzo_advance.argtypes = [C.c_void_p, pi_, pi_, pd]
    lib.yzo_get_h.argtypes = [C.c_void_p, pd]
    lib.yzo_set_h.argtypes = [C.c_void_p, pd]
    lib.yzo_clock.argtypes = [C.c_void_p, pd, pd, pd]
    lib.yzo_pattern.argtypes = [C.c_void_p, pi_, pi_, C.c_char_p]
    lib.yzo_assemble.argtypes = [C.c_void_p, pd, pd, C.c_double, pd, pd, pd, pd, pd]
    lib.yzo_linsolve.argtypes = [C.c_int, C.c_int, pi_, pi_, pd, pd, pd, C.c_int,
                                 C.c_int, C.c_int, C.c_double, pi_, pd]
    lib.yzo_mobility.argtypes = [C.c_double, C.c_double, C.c_double, pd]
    lib.yzo_evap.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, pd]
    lib.yzo_velocity.argtypes = [C.c_void_p, pd, pd, C.c_int, pd, pd]
    lib.yzo_velocity_terms.argtypes = [C.c_void_p, pd, pd, pd, pd]
    lib.yzo_energy.argtypes = [C.c_void_p, pd, pd, pd, pd]
    lib.yzo_source.argtypes = [C.c_void_p, C.c_double, pd]
    lib.yzo_limiter.argtypes = [pd, C.c_double, C.c_int, pd]
    lib.yzo_get_c.argtypes = [C.c_void_p, pd]
    lib.yzo_face_fluxes.argtypes = [C.c_void_p, pd, pd, pd, pd]
    lib.yzo_conc_stats.argtypes = [C.c_void_p, pd]
    lib.yzo_conc_face.argtypes = [C.c_int, pd, C.c_double, C.c_int, pd]
    lib.yzo_conc_settle.argtypes = [C.c_void_p]
    lib.yzo_conc_settle.restype = C.c_longlong
    lib.yzo_conc_nnz.argtypes = [C.c_void_p]
    lib.yzo_conc_system.argtypes = [C.c_void_p, pi_, pi_, pd, pd]
    lib.yzo_cz_case_defaults.argtypes = [P(YzCzCase)]
    lib.yzo_cz_create.restype = C.c_void_p
    lib.yzo_cz_create.argtypes = [P(YzCzCase), pd, pd, pd, C.c_double]
    lib.yzo_cz_destroy.argtypes = [C.c_void_p]
    lib.yzo_cz_advance.argtypes = [C.c_void_p, pi_, pi_, pd]
    lib.yzo_cz_get_h.argtypes = [C.c_void_p, pd]
    lib.yzo_cz_get_c.argtypes = [C.c_void_p, pd]
    lib.yzo_cz_volume_fluxes.argtypes = [C.c_void_p, pd, pd, pd]
    lib.yzo_cz_stats.argtypes = [C.c_void_p, pd]
    lib.yzo_cz_settle.argtypes = [C.c_void_p]
    lib.yzo_cz_clock.argtypes = [C.c_void_p, pd]
    lib.yzo_cz_settle.restype = C.c_longlong
    lib.yzo_cz_velocity.argtypes = [C.c_void_p, pd, pd, pd, pd]
    lib.yzo_cz_redistribute.restype = C.c_double
    lib.yzo_cz_redistribute.argtypes = [C.c_void_p, pd, pd, pd, pd]
    lib.yzo_case_defaults.argtypes = [P(YzCase)]
    lib.yzo_pass_log.argtypes = [C.c_void_p, pd, pd, pd, C.c_int]
    lib.yzo_check.argtypes = [C.c_void_p, P(YzTerm), pd]
    lib.yzo_script_create.restype = C.c_void_p
    lib.yzo_script_create.argtypes = [P(YzCase), C.c_int, pd, C.c_double, pd, C.c_int, pd, C.c_int,
                                      pd, C.c_int, pi_, C.c_int]
    lib.yzo_script_info.argtypes = [C.c_void_p, pi_, P(C.c_longlong), pd, C.c_int]
    return lib


def mobility(lam, h1, h2):
    out = np.zeros(6)
    load().yzo_mobility(lam, h1, h2, _pd(out))
    return out


def limiter(f, sign, vanleer=0):
    """the limited face of four heights: [value, downwind weights w0..w3,
    value again, derivative g0..g3]; Koren, or van Leer (vanleer 1)"""
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(10)
    load().yzo_limiter(_pd(f), sign, int(vanleer), _pd(out))
    return out


def conc_face(scheme, f, sign, rowpos=-1):
    """FaceValue and FaceWeights (rowpos: the row's cell among LL 0, L 1, R 2,
    RR 3; a limited scheme's weights depend on it) of a concentration scheme"""
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.empty(5)
    load().yzo_conc_face(scheme, _pd(f), sign, int(rowpos), _pd(out))
    return out


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _pi(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class OracleStep:
    """The oracle's time stepper over one case."""

    # the linear solves: the reference's algebraic multigrid, or the
    # device's structured one (the same error vectors as the device's solves
    # up to rounding, so Newton counts compare pass for pass)
    LIN_ALGEBRAIC, LIN_STRUCTURED = 0, 1

    def __init__(self, case, h, s, c, dt0, lin_mode=0, lib=None):
        self.lib = lib or load()
        self.case = case.copy()
        h = np.ascontiguousarray(h, dtype=np.float64)
        s = np.ascontiguousarray(s, dtype=np.float64)
        c = np.ascontiguousarray(c, dtype=np.float64)
        self.p = self.lib.yzo_create(C.byref(self.case), _pd(h), _pd(s), _pd(c), dt0)
        self.n = self.lib.yzo_n(self.p)
        self.nnz = self.lib.yzo_nnz(self.p)
        self.lib.yzo_set_lin_mode(self.p, int(lin_mode))

    def __del__(self):
        if getattr(self, "p", None):
            self.lib.yzo_destroy(self.p)
            self.p = None

    def advance(self):
        nk, lin, r = C.c_int(), C.c_int(), C.c_double()
        rc = self.lib.yzo_advance(self.p, C.byref(nk), C.byref(lin), C.byref(r))
        return rc, nk.value, lin.value, r.value

    def h(self):
        out = np.empty(self.n)
        self.lib.yzo_get_h(self.p, _pd(out))
        return out

    def set_h(self, h):
        h = np.ascontiguousarray(h, dtype=np.float64)
        self.lib.yzo_set_h(self.p, _pd(h))

    def c(self):
        out = np.empty(self.n)
        self.lib.yzo_get_c(self.p, _pd(out))
        return out

    def face_fluxes(self, h, c):
        h = np.ascontiguousarray(h, dtype=np.float64)
        c = np.ascontiguousarray(c, dtype=np.float64)
        F, r = np.zeros(2 * self.n), np.zeros(self.n)
        self.lib.yzo_face_fluxes(self.p, _pd(h), _pd(c), _pd(F), _pd(r))
        return F.reshape(2, self.n), r

    def velocity(self, h, c, bc=True):
        """the face velocities of the field (h, c): (vx, vy), padded, vx on
        the Xm face and vy on the Ym face of each cell; bc False leaves out
        the boundary rows"""
        h = np.ascontiguousarray(h, dtype=np.float64)
        c = np.ascontiguousarray(c, dtype=np.float64)
        vx, vy = np.zeros(self.n), np.zeros(self.n)
        self.lib.yzo_velocity(self.p, _pd(h), _pd(c), 1 if bc else 0, _pd(vx), _pd(vy))
        return vx, vy

    def velocity_terms(self, h, c):
        """the .vts's velocity fields of (h, c): (vx, vy), each (6, n), the
        mixed field, then the capillary, Marangoni, vertical-gravity,
        in-plane gravity and centrifugal terms' own, with boundary rows"""
        h = np.ascontiguousarray(h, dtype=np.float64)
        c = np.ascontiguousarray(c, dtype=np.float64)
        vx, vy = np.zeros(6 * self.n), np.zeros(6 * self.n)
        self.lib.yzo_velocity_terms(self.p, _pd(h), _pd(c), _pd(vx), _pd(vy))
        return vx.reshape(6, self.n), vy.reshape(6, self.n)

    def energy(self, h, c):
        """(F, D, phi): the free energy, the power of the face fluxes against
        the potential, and the potential of the field (h, c)"""
        h = np.ascontiguousarray(h, dtype=np.float64)
        c = np.ascontiguousarray(c, dtype=np.float64)
        out, phi = np.zeros(2), np.zeros(self.n)
        self.lib.yzo_energy(self.p, _pd(h), _pd(c), _pd(out), _pd(phi))
        return out[0], out[1], phi

    def conc_system(self):
        """the last concentration system: (rowptr, col, val, rhs), unscaled"""
        nnz = self.lib.yzo_conc_nnz(self.p)
        rp = np.empty(self.n + 1, dtype=np.int32)
        col = np.empty(nnz, dtype=np.int32)
        val, rhs = np.empty(nnz), np.empty(self.n)
        self.lib.yzo_conc_system(self.p, _pi(rp), _pi(col), _pd(val), _pd(rhs))
        return rp, col, val, rhs

    def conc_settle(self):
        """the solves that completed the steps of a limited concentration
        scheme, over the run"""
        return int(self.lib.yzo_conc_settle(self.p))

    def conc_stats(self):
        out = np.zeros(5)
        self.lib.yzo_conc_stats(self.p, _pd(out))
        return out

    def clock(self):
        t, a, b = C.c_double(), C.c_double(), C.c_double()
        self.lib.yzo_clock(self.p, C.byref(t), C.byref(a), C.byref(b))
        return t.value, a.value, b.value

    def check(self, t):
        """the run's termination tests after the last step: (status, info8)"""
        info = np.zeros(8)
        st = self.lib.yzo_check(self.p, C.byref(t), _pd(info))
        return st, info

    def pass_log(self):
        """the passes of the last step's accepted attempt: (correction,
        linear tolerance, linear relative residual) arrays"""
        a, b, r = np.zeros(256), np.zeros(256), np.zeros(256)
        m = self.lib.yzo_pass_log(self.p, _pd(a), _pd(b), _pd(r), 256)
        return a[:m].copy(), b[:m].copy(), r[:m].copy()

    def pattern(self):
        rp = np.empty(self.n + 1, dtype=np.int32)
        col = np.empty(self.nnz, dtype=np.int32)
        bc = C.create_string_buffer(self.n)
        self.lib.yzo_pattern(self.p, _pi(rp), _pi(col), bc)
        return rp, col, np.frombuffer(bc.raw, dtype=np.int8).astype(bool)

    def assemble(self, h_old, h, dt):
        n, nnz = self.n, self.nnz
        A, J, IA = np.empty(nnz), np.empty(nnz), np.empty(nnz)
        ax, rhs = np.empty(n), np.empty(n)
        h_old = np.ascontiguousarray(h_old, dtype=np.float64)
        h = np.ascontiguousarray(h, dtype=np.float64)
        self.lib.yzo_assemble(self.p, _pd(h_old), _pd(h), dt, _pd(A), _pd(J), _pd(IA), _pd(ax), _pd(rhs))
        return A, J, IA, ax, rhs

    def evap(self, h, c, dt):
        """(Ev E(h, c) of the law alone, EvapAt's capped rate at step dt)"""
        out = np.empty(2)
        self.lib.yzo_evap(self.p, h, c, dt, _pd(out))
        return out[0], out[1]

    def source(self, t):
        """the nozzle's supply of every cell of the padded grid at time t"""
        out = np.empty(self.n)
        self.lib.yzo_source(self.p, t, _pd(out))
        return out


def linsolve(rowptr, col, val, rhs, x0, precond=1, scaling=1, maxit=5000, eps=1e-13):
    lib = load()
    n = len(rowptr) - 1
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    x = np.array(x0, dtype=np.float64)
    it, r = C.c_int(), C.c_double()
    rc = lib.yzo_linsolve(n, len(col), _pi(rowptr), _pi(col), _pd(val), _pd(rhs), _pd(x),
                          precond, scaling, maxit, eps, C.byref(it), C.byref(r))
    return rc, x, it.value, r.value


class OracleCZ:
    """The oracle's ConcZ stepper over one film (yz_cz_oracle.cpp)."""

    def __init__(self, case, h, s, c, dt, lib=None):
        self.lib = lib or load()
        self.case = case.copy()
        self.n2 = (case.nx + 4) * (case.ny + 4)
        self.n3 = self.n2 * (case.nz + 4)
        h = np.ascontiguousarray(h, dtype=np.float64)
        s = np.ascontiguousarray(s, dtype=np.float64)
        c = np.ascontiguousarray(c, dtype=np.float64)
        assert h.size == self.n2 and s.size == self.n2 and c.size == self.n3
        self.p = self.lib.yzo_cz_create(C.byref(self.case), _pd(h), _pd(s), _pd(c), dt)

    def __del__(self):
        if getattr(self, "p", None):
            self.lib.yzo_cz_destroy(self.p)
            self.p = None

    def advance(self):
        lin, ps, r = C.c_int(), C.c_int(), C.c_double()
        rc = self.lib.yzo_cz_advance(self.p, C.byref(lin), C.byref(ps), C.byref(r))
        return rc, lin.value, ps.value, r.value

    def h(self):
        out = np.empty(self.n2)
        self.lib.yzo_cz_get_h(self.p, _pd(out))
        return out

    def c(self):
        out = np.empty(self.n3)
        self.lib.yzo_cz_get_c(self.p, _pd(out))
        return out

    def volume_fluxes(self):
        qx, qy, w = np.empty(self.n3), np.empty(self.n3), np.empty(self.n3)
        self.lib.yzo_cz_volume_fluxes(self.p, _pd(qx), _pd(qy), _pd(w))
        return qx, qy, w

    def stats(self):
        out = np.empty(8)
        self.lib.yzo_cz_stats(self.p, _pd(out))
        return out

    def clock(self):
        """(dt now, dt old, trials the step control rejected)"""
        out = np.empty(3)
        self.lib.yzo_cz_clock(self.p, _pd(out))
        return out

    def settle(self):
        """SettleLimiter's solves over the run (a limited scheme)"""
        return int(self.lib.yzo_cz_settle(self.p))

    def velocity(self, h, c):
        h = np.ascontiguousarray(h, dtype=np.float64)
        c = np.ascontiguousarray(c, dtype=np.float64)
        ux, uy = np.empty(self.n3), np.empty(self.n3)
        self.lib.yzo_cz_velocity(self.p, _pd(h), _pd(c), _pd(ux), _pd(uy))
        return ux, uy

    def redistribute(self, h, c):
        """RedistributeExcess then RedistributeDeficit on a copy of c:
        (c, excess moved, deficit moved, deficit made up)"""
        h = np.ascontiguousarray(h, dtype=np.float64)
        c = np.array(c, dtype=np.float64)
        mv, ad = C.c_double(), C.c_double()
        ex = self.lib.yzo_cz_redistribute(self.p, _pd(h), _pd(c), C.byref(mv), C.byref(ad))
        return c, ex, mv.value, ad.value


def case_defaults(lib=None):
    """the oracle's default case (its Newton and step-control numerics are
    the reference's defaults)"""
    c = YzCase()
    (lib if lib is not None else load()).yzo_case_defaults(C.byref(c))
    return c


class ScriptedNewton:
    """The oracle's Newton loop and step control over n rows, with each
    pass's correction, coupled-side correction, right-hand-side norm and
    solve return code taken from scripts instead of an assembled system
    (the reference's ScriptedNK, yuzen/tests/test_newton_convergence.cpp:
    38-84). Entry k of a script is used by the k-th solve of the run
    (counted over every attempt and step); the last entry repeats."""

    def __init__(self, case, phi0, dt0, dH=(), dC=(), rn=(), rc=(), lib=None):
        self.lib = lib if lib is not None else load()
        self.case = case.copy()
        phi0 = np.ascontiguousarray(phi0, dtype=np.float64)
        self.n = phi0.size
        self._keep = [np.ascontiguousarray(v, dtype=np.float64) for v in (dH, dC, rn)]
        rc = np.ascontiguousarray(rc, dtype=np.int32)
        self._keep.append(rc)
        a, b, r = self._keep[:3]
        self.p = self.lib.yzo_script_create(C.byref(self.case), self.n, _pd(phi0), dt0,
                                            _pd(a), a.size, _pd(b), b.size, _pd(r), r.size,
                                            _pi(rc), rc.size)

    def __del__(self):
        if getattr(self, "p", None):
            self.lib.yzo_destroy(self.p)
            self.p = None

    def advance(self):
        """one time step: (rc, Newton passes of the accepted attempt)"""
        nk, lin, r = C.c_int(), C.c_int(), C.c_double()
        rc = self.lib.yzo_advance(self.p, C.byref(nk), C.byref(lin), C.byref(r))
        return rc, nk.value

    def phi(self):
        out = np.empty(self.n)
        self.lib.yzo_get_h(self.p, _pd(out))
        return out

    def clock(self):
        """(time, dt of the next step, dt of the last step)"""
        t, a, b = C.c_double(), C.c_double(), C.c_double()
        self.lib.yzo_clock(self.p, C.byref(t), C.byref(a), C.byref(b))
        return t.value, a.value, b.value

    def info(self):
        """(solves so far, stagnation cuts, the tolerance of every solve)"""
        calls, cuts = C.c_int(), C.c_longlong()
        eta = np.zeros(4096)
        m = self.lib.yzo_script_info(self.p, C.byref(calls), C.byref(cuts), _pd(eta), eta.size)
        return calls.value, cuts.value, eta[:min(m, eta.size)].copy()
