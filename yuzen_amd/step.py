"""Python handle on the device step (ctypes over libyuzen_hip.so).

``HipStep`` is the film-height equation of one case resident on the GPU: it
is created from a :class:`yuzen_amd.capi.YzCase` and the padded fields,
advances one implicit step per ``advance()`` (the reference's
``C_NewtonKantorovich::TimeAdvance`` contract: 0 accepted, 1 failed at a
fixed step, 2 device error) and hands back the film. ``linsolve`` is the
one-system seam with the contract of ``ILinearSolverBackend::Solve``.
Nothing here computes: every call goes to the HIP library, and a missing
library raises.
"""
import ctypes as C
import os

import numpy as np

from . import capi


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _pf(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def _pi(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class HipStep:
    def __init__(self, case, h, s, c, dt0):
        self.lib = capi.load()
        self.case = case.copy()
        self.dim = case.dim
        nyp = case.ny + 4 if case.dim == 2 else 1
        self.n = (case.nx + 4) * nyp
        self.N = case.nx * (case.ny if case.dim == 2 else 1)
        self.ns = 13 if case.dim == 2 else 5
        h = np.ascontiguousarray(h, dtype=np.float64)
        s = np.ascontiguousarray(s, dtype=np.float64)
        c = np.ascontiguousarray(c, dtype=np.float64)
        assert h.size == self.n and s.size == self.n and c.size == self.n
        self.p = self.lib.yz_step_create(C.byref(self.case), _pd(h), _pd(s), _pd(c), dt0)
        if not self.p:
            raise RuntimeError("yz_step_create: " + capi.last_error())

    def close(self):
        if getattr(self, "p", None):
            self.lib.yz_step_destroy(self.p)
            self.p = None

    def __del__(self):
        self.close()

    def save_state(self, path):
        """The restart state file (yz_step_save_state)."""
        _state_call(self.lib.yz_step_save_state, self.p, path, "yz_step_save_state")

    def load_state(self, path):
        """Continue from a restart state file (yz_step_load_state)."""
        _state_call(self.lib.yz_step_load_state, self.p, path, "yz_step_load_state")

    def advance(self):
        nk, lin, r = C.c_int(), C.c_int(), C.c_double()
        rc = self.lib.yz_step_advance(self.p, C.byref(nk), C.byref(lin), C.byref(r))
        if rc == 2:
            raise RuntimeError("yz_step_advance: " + capi.last_error())
        return rc, nk.value, lin.value, r.value

    def h(self):
        out = np.empty(self.n)
        if self.lib.yz_step_get_h(self.p, _pd(out)):
            raise RuntimeError(capi.last_error())
        return out

    def velocity(self):
        """the face velocities of the current film (CalVelocity): (ux, uy),
        padded, ux on the Xm face and uy on the Ym face of each cell"""
        ux, uy = np.empty(self.n), np.zeros(self.n)
        if self.lib.yz_step_velocity(self.p, _pd(ux), _pd(uy)):
            raise RuntimeError(capi.last_error())
        return ux, uy

    def velocity_terms(self):
        """the .vts's velocity fields of the current film: (ux6, uy6), each
        (6, n): the mixed field, then the capillary, Marangoni,
        vertical-gravity, in-plane gravity and centrifugal terms' own"""
        ux, uy = np.empty(6 * self.n), np.zeros(6 * self.n)
        if self.lib.yz_step_velocity_terms(self.p, _pd(ux), _pd(uy)):
            raise RuntimeError(capi.last_error())
        return ux.reshape(6, self.n), uy.reshape(6, self.n)

    def field_info(self):
        """FieldInfo of the current film: (volume, solute, mean c, max h+s,
        min h+s, max h, min h, max c, min c)"""
        out = np.zeros(9)
        if self.lib.yz_step_field_info(self.p, _pd(out)):
            raise RuntimeError(capi.last_error())
        return tuple(out)

    def series_open(self, directory, dimensioned=False, L0=1.0, H0=1.0, T0=1.0, sigma0=1.0, energy=False):
        """open the run's History.dat, CalcInfo.dat (and Energy.dat) in
        directory; every step from now on adds to the run's sums"""
        sc = capi.YzSeriesScales(1 if dimensioned else 0, L0, H0, T0, sigma0, 1 if energy else 0)
        if self.lib.yz_step_series_open(self.p, os.fsencode(directory), C.byref(sc)):
            raise RuntimeError(capi.last_error())

    def series_write(self):
        """append one row to each open time series"""
        if self.lib.yz_step_series_write(self.p):
            raise RuntimeError(capi.last_error())

    def energy(self):
        """(F, D): the free energy of the current film and the power of its
        face fluxes against the potential"""
        F, D = C.c_double(), C.c_double()
        if self.lib.yz_step_energy(self.p, C.byref(F), C.byref(D)):
            raise RuntimeError(capi.last_error())
        return F.value, D.value

    def balances(self):
        """(volume, solute, evaporated, supplied): the film's volume and
        solute, and the volume the last step evaporated and the nozzle
        supplied"""
        out = np.zeros(4)
        if self.lib.yz_step_balances(self.p, _pd(out)):
            raise RuntimeError(capi.last_error())
        return tuple(out)

    def write_dat(self, path, L0=1.0, H0=1.0, U0=1.0, dimensioned=False):
        """the reference's ASCII field file of the current film"""
        if self.lib.yz_step_write_dat(self.p, os.fsencode(path), L0, H0, U0, 1 if dimensioned else 0):
            raise RuntimeError(capi.last_error())

    def write_vts(self, path, L0=1.0, H0=1.0, U0=1.0, T0=1.0, dimensioned=False):
        """the reference's VTK structured-grid file of the current film"""
        if self.lib.yz_step_write_vts(self.p, os.fsencode(path), L0, H0, U0, T0, 1 if dimensioned else 0):
            raise RuntimeError(capi.last_error())

    def set_h(self, h):
        h = np.ascontiguousarray(h, dtype=np.float64)
        if self.lib.yz_step_set_h(self.p, _pd(h)):
            raise RuntimeError(capi.last_error())

    def c(self):
        """the concentration of the state (the frozen input with conc_model 0)"""
        out = np.empty(self.n)
        if self.lib.yz_step_get_c(self.p, _pd(out)):
            raise RuntimeError(capi.last_error())
        return out

    def face_fluxes(self, h, c):
        """FaceVolumeFluxes of film h with concentration c: (F[2, n], r[n])"""
        h = np.ascontiguousarray(h, dtype=np.float64)
        c = np.ascontiguousarray(c, dtype=np.float64)
        F, r = np.zeros(2 * self.n), np.zeros(self.n)
        if self.lib.yz_step_face_fluxes(self.p, _pd(h), _pd(c), _pd(F), _pd(r)):
            raise RuntimeError(capi.last_error())
        return F.reshape(2, self.n), r

    def conc_stats(self):
        """[solves, their Krylov iterations, solves that needed the multigrid,
        solute moved above 1, below 0 (last step)]"""
        out = np.zeros(5)
        self.lib.yz_step_conc_stats(self.p, _pd(out))
        return out

    def clock(self):
        t, a, b = C.c_double(), C.c_double(), C.c_double()
        self.lib.yz_step_clock(self.p, C.byref(t), C.byref(a), C.byref(b))
        return t.value, a.value, b.value

    def pass_log(self):
        """the passes of the last step's accepted attempt: (correction,
        linear tolerance, linear relative residual) arrays"""
        a, b, r = np.zeros(256), np.zeros(256), np.zeros(256)
        m = self.lib.yz_step_pass_log(self.p, _pd(a), _pd(b), _pd(r), 256, None)
        return a[:m].copy(), b[:m].copy(), r[:m].copy()

    def set_term(self, t):
        """the run's termination tests after every step (capi.term(...)), or None"""
        self._term = t
        if self.lib.yz_step_set_term(self.p, C.byref(t) if t is not None else None):
            raise RuntimeError("yz_step_set_term: " + capi.last_error())

    def term_status(self):
        """(status, info8) after the last step"""
        st = C.c_int()
        info = np.zeros(8)
        self.lib.yz_step_term_status(self.p, C.byref(st), _pd(info))
        return st.value, info

    def run(self, nsteps):
        """up to nsteps steps, stopping where a termination test says so:
        (rc, steps done, status)"""
        done, st = C.c_int(), C.c_int()
        rc = self.lib.yz_step_run(self.p, nsteps, C.byref(done), C.byref(st))
        if rc == 2:
            raise RuntimeError("yz_step_run: " + capi.last_error())
        return rc, done.value, st.value

    def stagnation_cuts(self):
        n = C.c_longlong()
        self.lib.yz_step_pass_log(self.p, None, None, None, 0, C.byref(n))
        return n.value

    def timing(self):
        out = np.zeros(6)
        self.lib.yz_step_timing(self.p, _pd(out))
        return out

    def kernel_time(self, enable=-1):
        """(launches, total ms) of the fine-level Krylov SpMV since the last
        reset; enable=1 resets and starts timing, 0 resets and stops."""
        n, ms = C.c_longlong(), C.c_double()
        self.lib.yz_step_kernel_time(self.p, enable, C.byref(n), C.byref(ms))
        return n.value, ms.value

    def assemble(self, h_old, h, dt):
        ia = np.empty(self.ns * self.N)
        ax = np.empty(self.n)
        rhs = np.empty(self.n)
        h_old = np.ascontiguousarray(h_old, dtype=np.float64)
        h = np.ascontiguousarray(h, dtype=np.float64)
        if self.lib.yz_step_assemble(self.p, _pd(h_old), _pd(h), dt, _pd(ia), _pd(ax), _pd(rhs)):
            raise RuntimeError(capi.last_error())
        return ia.reshape(self.ns, self.N), ax, rhs

    # ---- learned preconditioner (contexts with PC_LEARNED / PC_LEARNED_AMG)
    def _check(self, rc, what):
        if rc:
            raise RuntimeError(what + ": " + capi.last_error())

    def learn_params(self):
        out = np.empty(capi.LEARN_NPARAM, dtype=np.float32)
        self._check(self.lib.yz_learn_get_params(self.p, _pf(out)), "yz_learn_get_params")
        return out

    def learn_set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == capi.LEARN_NPARAM
        self._check(self.lib.yz_learn_set_params(self.p, _pf(p)), "yz_learn_set_params")

    def learn_sample(self):
        """The row-scaled FP32 IA planes (ns, N), features (8, N) and row
        scales (N) of the Newton system at the current film."""
        ia = np.empty((self.ns, self.N), dtype=np.float32)
        feat = np.empty((capi.LEARN_NFEAT, self.N), dtype=np.float32)
        ms = np.empty(self.N, dtype=np.float32)
        self._check(self.lib.yz_learn_sample(self.p, _pf(ia), _pf(feat), _pf(ms)), "yz_learn_sample")
        return ia, feat, ms

    def learn_loss_grad(self, nprobe, seed):
        g = np.empty(capi.LEARN_NPARAM, dtype=np.float32)
        loss = C.c_double()
        self._check(self.lib.yz_learn_loss_grad(self.p, nprobe, seed, C.byref(loss), _pf(g)),
                    "yz_learn_loss_grad")
        return loss.value, g

    def learn_train(self, iters, lr, nprobe):
        a, b = C.c_double(), C.c_double()
        self._check(self.lib.yz_learn_train(self.p, iters, lr, nprobe, C.byref(a), C.byref(b)),
                    "yz_learn_train")
        return a.value, b.value

    def learn_apply(self):
        M = np.empty((self.ns, self.N), dtype=np.float32)
        self._check(self.lib.yz_learn_apply(self.p, _pf(M)), "yz_learn_apply")
        return M

    def learn_stats(self):
        """fallbacks, Adam steps, training ms, first and last probe loss"""
        out = np.zeros(5)
        self._check(self.lib.yz_learn_stats(self.p, _pd(out)), "yz_learn_stats")
        return out


def _state_call(fn, p, path, what):
    rc = fn(p, os.fsencode(path))
    if rc != 0:
        raise RuntimeError(f"{what}: " + capi.last_error())


class LinSolver:
    """The linear-solve seam (one CRS system per call)."""

    def __init__(self):
        self.lib = capi.load()
        self.p = self.lib.yz_linsolve_create()
        if not self.p:
            raise RuntimeError("yz_linsolve_create failed")

    def __del__(self):
        if getattr(self, "p", None):
            self.lib.yz_linsolve_destroy(self.p)
            self.p = None

    def solve(self, rowptr, col, val, rhs, x0, precond=capi.PC_AMG, scaling=1, max_iter=5000,
              epsilon=1e-13, grid=None, conc=0):
        s = capi.YzLinSettings()
        s.precond, s.scaling, s.max_iter, s.epsilon = precond, scaling, max_iter, epsilon
        s.conc = conc
        if grid is not None:
            s.grid_dim, s.grid_nx, s.grid_ny = grid[0], grid[1], grid[2]
            for k in range(4):
                s.bc[k] = grid[3][k]
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.array(x0, dtype=np.float64)
        it, r = C.c_int(), C.c_double()
        rc = self.lib.yz_linsolve(self.p, len(rowptr) - 1, len(col), _pi(rowptr), _pi(col), _pd(val),
                                  _pd(rhs), _pd(x), C.byref(s), C.byref(it), C.byref(r))
        return rc, x, it.value, r.value


def linsolve_noalloc(rowptr, col, val, rhs, x0, precond=capi.PC_AMG, scaling=1, max_iter=5000, epsilon=1e-13,
                     conc=0):
    """yz_linsolve_noalloc: the seam's solve in the library's scratch (no
    allocation); (rc, x, iterations, relative residual)"""
    lib = capi.load()
    s = capi.YzLinSettings()
    s.precond, s.scaling, s.max_iter, s.epsilon = precond, scaling, max_iter, epsilon
    s.conc = conc
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    x = np.array(x0, dtype=np.float64)
    it, r = C.c_int(), C.c_double()
    rc = lib.yz_linsolve_noalloc(len(rowptr) - 1, len(col), _pi(rowptr), _pi(col), _pd(val), _pd(rhs), _pd(x),
                                 C.byref(s), C.byref(it), C.byref(r))
    return rc, x, it.value, r.value


class HipBatch:
    """Independent 2-D films advanced together (the batched sweep): one
    fixed implicit step per ``advance()`` for every film, their Krylov
    solves batched into one launch per kernel. ``films`` is a list of
    ``(case, h, s, c, dt0)``; the films share one network
    (``set_params``)."""

    def __init__(self, films):
        self.lib = capi.load()
        self.n = len(films)
        arr = (capi.YzCase * self.n)()
        hs, ss, cs, dts = [], [], [], []
        self.sizes = []
        for k, (case, h, s, c, dt0) in enumerate(films):
            arr[k] = case
            n = (case.nx + 4) * (case.ny + 4)
            self.sizes.append(n)
            for lst, a in ((hs, h), (ss, s), (cs, c)):
                a = np.ascontiguousarray(a, dtype=np.float64)
                assert a.size == n
                lst.append(a)
            dts.append(dt0)
        self._cases = arr
        h = np.concatenate(hs)
        s = np.concatenate(ss)
        c = np.concatenate(cs)
        dt = np.ascontiguousarray(dts, dtype=np.float64)
        self.p = self.lib.yz_batch_create(self.n, arr, _pd(h), _pd(s), _pd(c), _pd(dt))
        if not self.p:
            raise RuntimeError("yz_batch_create: " + capi.last_error())

    def close(self):
        if getattr(self, "p", None):
            self.lib.yz_batch_destroy(self.p)
            self.p = None

    def __del__(self):
        self.close()

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == capi.LEARN_NPARAM
        if self.lib.yz_batch_set_params(self.p, _pf(p)):
            raise RuntimeError("yz_batch_set_params: " + capi.last_error())

    def advance(self):
        """(rc, status[n], nk[n], lin[n]); rc 0 all accepted, 1 some film
        failed (status 1), device errors raise"""
        st = np.zeros(self.n, dtype=np.int32)
        nk = np.zeros(self.n, dtype=np.int32)
        lin = np.zeros(self.n, dtype=np.int32)
        rc = self.lib.yz_batch_advance(self.p, _pi(st), _pi(nk), _pi(lin))
        if rc == 2:
            raise RuntimeError("yz_batch_advance: " + capi.last_error())
        return rc, st, nk, lin

    def h(self, film):
        out = np.empty(self.sizes[film])
        if self.lib.yz_batch_get_h(self.p, film, _pd(out)):
            raise RuntimeError(capi.last_error())
        return out

    def stats(self):
        """last step's milliseconds, Newton passes, Krylov iterations;
        FP64 fallbacks so far"""
        out = np.zeros(4)
        self.lib.yz_batch_stats(self.p, _pd(out))
        return out


class HipCZ:
    """One ConcZ film on the device (yz_cz_*): a 2-D film and its
    concentration through the thickness, stepped as CZ_HeightEquation_2D::
    TimeAdvance steps it."""

    def __init__(self, case, h, s, c, dt):
        self.lib = capi.load()
        self.case = case.copy()
        self.n2 = (case.nx + 4) * (case.ny + 4)
        self.n3 = self.n2 * (case.nz + 4)
        h = np.ascontiguousarray(h, dtype=np.float64)
        s = np.ascontiguousarray(s, dtype=np.float64)
        c = np.ascontiguousarray(c, dtype=np.float64)
        assert h.size == self.n2 and s.size == self.n2 and c.size == self.n3
        self.p = self.lib.yz_cz_create(C.byref(self.case), _pd(h), _pd(s), _pd(c), dt)
        if not self.p:
            raise RuntimeError("yz_cz_create: " + capi.last_error())

    def close(self):
        if getattr(self, "p", None):
            self.lib.yz_cz_destroy(self.p)
            self.p = None

    def __del__(self):
        self.close()

    def advance(self):
        lin, ps, r = C.c_int(), C.c_int(), C.c_double()
        rc = self.lib.yz_cz_advance(self.p, C.byref(lin), C.byref(ps), C.byref(r))
        if rc < 0:
            raise RuntimeError("yz_cz_advance: " + capi.last_error())
        return rc, lin.value, ps.value, r.value

    def h(self):
        out = np.empty(self.n2)
        self.lib.yz_cz_get_h(self.p, _pd(out))
        return out

    def c(self):
        out = np.empty(self.n3)
        self.lib.yz_cz_get_c(self.p, _pd(out))
        return out

    def save_state(self, path):
        """The restart state file (yz_cz_save_state)."""
        _state_call(self.lib.yz_cz_save_state, self.p, path, "yz_cz_save_state")

    def load_state(self, path):
        """Continue from a restart state file (yz_cz_load_state)."""
        _state_call(self.lib.yz_cz_load_state, self.p, path, "yz_cz_load_state")

    def volume_fluxes(self):
        qx, qy, w = np.empty(self.n3), np.empty(self.n3), np.empty(self.n3)
        self.lib.yz_cz_volume_fluxes(self.p, _pd(qx), _pd(qy), _pd(w))
        return qx, qy, w

    def stats(self):
        out = np.empty(8)
        self.lib.yz_cz_stats(self.p, _pd(out))
        return out

    def clock(self):
        """(the next step, the last accepted step, trials rejected)"""
        out = np.empty(3)
        self.lib.yz_cz_clock(self.p, _pd(out))
        return out

    def learn_stats(self):
        """the learned film solve: refinement rounds, FP32 iterations,
        solves that FP64 AMG finished, training milliseconds"""
        out = np.empty(4)
        self.lib.yz_cz_learn_stats(self.p, _pd(out))
        return out
