"""The device step's failure paths against the oracle's, both driven by the
same fault injection (yz_set_fault / the oracle's set_fault: the film's
Newton-pass solves k .. k+m-1 break down before they iterate, the
counterpart of the reference's YUZEN_FAULT_AMGCL=k:m,
yuzen/src/numerics/LinearSolverBackend.cpp:628-645):

  * without step control: the step ends with 1 at the same pass, the clock
    does not move, and the next step matches the oracle's;
  * after an extrapolated guess: the retry from h^n with the extrapolation
    blocked (NewtonKantorovich.cpp:424-486);
  * under step control: the same cuts of the step (ImplicitSolver.cpp:
    144-190), the same step size after them, and a breakdown that persists
    ends the step with 1 below Numerics.TimeStep.Min on both;
  * the learned FP32 path: a breakdown of its solve takes the same failure
    path.

The oracle's side of each case is pinned in tests/test_oracle_fault.py.
"""
import numpy as np
import pytest

from oracle import binding as ob
from yuzen_amd import capi, cases
from yuzen_amd.step import HipStep

pytestmark = pytest.mark.gpu


def _both(c, h, s, cc, dt, plan):
    """advance the oracle and the device through plan: a list of (fault k,
    m) before each step (None: no fault); the results per step"""
    o = ob.OracleStep(c, h, s, cc, dt, lin_mode=ob.OracleStep.LIN_STRUCTURED)
    g = HipStep(c, h, s, cc, dt)
    out = []
    try:
        for f in plan:
            for side in ("oracle", "device"):
                k, m = f if f else (0, 1)
                ob.set_fault(k, m)
                capi.set_fault(k, m)
                r = o.advance() if side == "oracle" else g.advance()
                out.append((side, r, o.clock() if side == "oracle" else g.clock()))
            ob.set_fault(0)
            capi.set_fault(0)
            (_, ro, co), (_, rg, cg) = out[-2], out[-1]
            print(f"fault {f}: oracle {ro[:3]} {co}, device {rg[:3]} {cg}", flush=True)
            assert ro[0] == rg[0] and ro[1] == rg[1]
            assert co[0] == cg[0] or abs(co[0] - cg[0]) <= 1e-15 * abs(co[0])
            assert abs(co[1] - cg[1]) <= 1e-15 * co[1] and abs(co[2] - cg[2]) <= 1e-15 * max(co[2], 1e-300)
            ho, hg = o.h(), g.h()
            assert np.abs(ho - hg).max() <= 1e-10 * np.abs(ho).max()
    finally:
        ob.set_fault(0)
        capi.set_fault(0)
        g.close()
    return out


def test_breakdown_without_step_control():
    c, h, s, cc, dt, _ = cases.leveling_1d(nx=64)
    out = _both(c, h, s, cc, dt, [(2, 1), None, None])
    assert out[0][1][0] == 1 and out[1][1][0] == 1 and out[1][2][0] == 0.0


def test_breakdown_after_extrapolation_retries():
    c, h, s, cc, dt, _ = cases.leveling_1d(nx=64)
    _both(c, h, s, cc, dt, [None, (1, 1), None])
    c, h, s, cc, dt, _ = cases.allphys_2d(n=48, dt=2e-5)
    _both(c, h, s, cc, dt, [None, (1, 1), (2, 1)])


def test_breakdown_under_step_control():
    c, h, s, cc, dt, _ = cases.spinodal_2d(n=32)
    out = _both(c, h, s, cc, dt, [(1, 1), None, (2, 3), None])
    assert out[1][2][2] < dt
    out = _both(c, h, s, cc, dt, [(1, 100000)])
    assert out[1][1][0] == 1 and out[1][2][1] < c.dt_min


def test_breakdown_of_the_learned_fp32_solve():
    c, h, s, cc, dt, _ = cases.allphys_2d(n=64, dt=2e-6)
    c.precond, c.krylov_fp32 = capi.PC_LEARNED, 1
    c.learn_passes, c.learn_iters, c.learn_probes, c.learn_lr = 2, 40, 4, 2e-3
    g = HipStep(c, h, s, cc, dt)
    try:
        capi.set_fault(1, 1)
        r = g.advance()
        capi.set_fault(0)
        print(f"learned, fault at the first solve: {r[:3]} {g.clock()}", flush=True)
        assert r[0] == 1 and r[1] == 1 and g.clock()[0] == 0.0
        assert g.advance()[0] == 0
        o = ob.OracleStep(c, h, s, cc, dt, lin_mode=ob.OracleStep.LIN_STRUCTURED)
        assert o.advance()[0] == 0
        ho = o.h()
        assert np.abs(g.h() - ho).max() <= 1e-8 * np.abs(ho).max()
    finally:
        capi.set_fault(0)
        g.close()
