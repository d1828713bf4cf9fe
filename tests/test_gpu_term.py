"""The run's termination tests on the device (yz_step_set_term,
yz_step_term_status, yz_step_run) against the oracle's (CheckTermination,
yuzen/src/model/well_mixed/WellMixedModelBase.h:1567-1679), step by step:
the same status after every step and the same field figures (volume, mean
c, min h, max c, min c to 1e-12 relative, the change rate, the steady
count). Also the step control with its truncation criterion read back once
per step: a spinodal film under step control advanced by yz_step_run to
MaxSteps matches the oracle's film and clock."""
import numpy as np
import pytest

from oracle import binding as ob
from yuzen_amd import capi, cases
from yuzen_amd.step import HipStep

from .test_oracle_term import _flat, _set_cell

pytestmark = pytest.mark.gpu


def _pair_steps(c, h, s, cc, dt, t, steps, label):
    o = ob.OracleStep(c, h, s, cc, dt, lin_mode=ob.OracleStep.LIN_STRUCTURED)
    g = HipStep(c, h, s, cc, dt)
    g.set_term(t)
    got = []
    for k in range(steps):
        ro, rg = o.advance(), g.advance()
        assert ro[0] == 0 and rg[0] == 0
        so, io = o.check(t)
        sg, ig = g.term_status()
        print(f"{label} step {k + 1}: oracle {so} device {sg} info {ig[:6]}", flush=True)
        assert so == sg, f"step {k + 1}: oracle {so}, device {sg}"
        for j in range(5):
            assert abs(io[j] - ig[j]) <= 1e-12 * max(abs(io[j]), 1e-300), (j, io[j], ig[j])
        assert abs(io[5] - ig[5]) <= 1e-9 * max(abs(io[5]), 1e-300)
        assert io[6] == ig[6] and io[7] == ig[7]
        got.append(sg)
        if sg != capi.RUN_GOING:
            break
    g.close()
    return got


def test_steady_film_at_rest():
    c, h, s, cc = _flat()
    got = _pair_steps(c, h, s, cc, 1e-3, capi.term(steady_eps=1e-3, steady_steps=10), 20, "at rest")
    assert got[-1] == capi.RUN_STEADY and len(got) == 15


def test_leveling_film_not_steady_and_end():
    c, h, s, cc, dt, _ = cases.allphys_2d(n=24, L=1.0, dt=2e-5)
    got = _pair_steps(c, h, s, cc, dt, capi.term(steady_eps=1e-3, max_steps=6), 8, "all-physics")
    assert got == [capi.RUN_GOING] * 5 + [capi.RUN_END]


def test_dry_out_and_concentration_range():
    c, h, s, cc = _flat(conc=0.97)
    got = _pair_steps(c, h, s, cc, 1e-3, capi.term(steady_eps=0, dry=1, end_conc_value=0.95), 3, "dry")
    assert got == [capi.RUN_DRY]
    c, h, s, cc = _flat(n=12, conc=0.4, per=False)
    c.conc_model, c.Df = 1, 0.1
    cc = _set_cell(c, cc, 5, 5, -0.2)
    got = _pair_steps(c, h, s, cc, 1e-4, capi.term(steady_eps=0), 3, "conc")
    assert got == [capi.RUN_CONC_RANGE]


def test_negative_film():
    c, h, s, cc = _flat(n=16, per=False)
    c.capillary = 1
    X, Y = cases.cell_centres(c)
    h = (0.5 + 0.53 * np.cos(2 * np.pi * X) * np.cos(2 * np.pi * Y)).ravel()
    got = _pair_steps(c, h, s, cc, 1e-6, capi.term(steady_eps=0), 2, "negative")
    assert got[0] == capi.RUN_NEGATIVE_FILM


def test_run_under_step_control():
    c, h, s, cc, dt, _ = cases.spinodal_2d(n=32)
    steps = 12
    o = ob.OracleStep(c, h, s, cc, dt, lin_mode=ob.OracleStep.LIN_STRUCTURED)
    for _ in range(steps):
        assert o.advance()[0] == 0
    g = HipStep(c, h, s, cc, dt)
    g.set_term(capi.term(steady_eps=0, max_steps=steps))
    rc, done, status = g.run(steps + 5)
    print(f"run: rc {rc}, {done} steps, status {status}; dt {g.clock()[1]:.3e} (oracle {o.clock()[1]:.3e})")
    assert rc == 0 and done == steps and status == capi.RUN_END
    assert abs(g.clock()[1] - o.clock()[1]) <= 1e-15 * o.clock()[1]
    ho = o.h()
    assert np.abs(g.h() - ho).max() <= 1e-10 * np.abs(ho).max()
    g.close()
