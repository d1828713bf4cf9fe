"""The oracle's Newton control and step control, pinned.

The oracle's Newton loop (exits, damping, extrapolation and its retry, the
Eisenstat-Walker tolerance) and its step control (LocalTruncationCriterion,
TimeStepAdjustment) are driven with scripted corrections instead of
assembled systems, the way the reference's own test drives its loop
(yuzen/tests/test_newton_convergence.cpp:38-84, ScriptedNK: every row of
the correction set to the script's value, the coupled side's correction
likewise, entry k used by the k-th solve of the run, the last entry
repeated).

  * the nine cases of test_newton_convergence.cpp:110-386, with the
    reference's expected exits;
  * exits, damping and extrapolation worked by hand from
    NewtonKantorovich.cpp:157-486;
  * the Eisenstat-Walker sequence (choice 2, NewtonKantorovich.cpp:108-126)
    on scripted right-hand-side norms, worked by hand;
  * the truncation criterion and the halve/double rules
    (ImplicitSolver.cpp:144-190, 271-307; TimeSteps.cpp:38-61, 102-130) on
    hand-made three-level sequences.

Each check is then run against copies of the oracle with one mistake
planted (a dropped square, a wrong limit, a missing branch): every planted
mistake must make its check fail.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import binding as ob
from yuzen_amd import cases

TINY = 1e-20     # a correction counted as converged (the reference test's)
BIG = 1.0        # a correction still moving

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _case(lib, *, eps=1e-10, rel=1e-6, nk_max=30, nk_min=1, step_control=False, change=0.5,
          crit=1e-3, floor=0.0, keep=50, dt_min=1e-8, dt_max=1.0, ew=1, extrap=1, lin_eps=1e-13):
    c = ob.case_defaults(lib)
    c.nk_eps, c.nk_eps_rel, c.nk_max, c.nk_min = eps, rel, nk_max, nk_min
    c.step_control = 1 if step_control else 0
    c.tsc_change, c.tsc_criterion, c.tsc_floor, c.tsc_keep = change, crit, floor, keep
    c.dt_min, c.dt_max = dt_min, dt_max
    c.ew, c.extrap, c.lin_eps = ew, extrap, lin_eps
    return c


def _run_one_step(lib, dH, dC, min_iter):
    """RunOneStep (test_newton_convergence.cpp:87-104): 8 rows at 0, dt 0.1
    fixed, NewtonEpsilon 1e-10, NewtonMaxIterations 30"""
    s = ob.ScriptedNewton(_case(lib, nk_min=min_iter), np.zeros(8), 0.1, dH=dH, dC=dC, lib=lib)
    rc, it = s.advance()
    assert rc == 0
    return it


def _run_adj(lib, script, eps_abs, eps_rel, max_iter, phi0=None):
    """RunAdj (:201-221): 8 rows, dt 0.1 under step control (MinDt 1e-8,
    MaxDt 1, change factor 0.5, criterion at the solver's default 1e-3)"""
    phi0 = np.zeros(8) if phi0 is None else phi0
    s = ob.ScriptedNewton(_case(lib, eps=eps_abs, rel=eps_rel, nk_max=max_iter, step_control=True,
                                crit=1e-3), phi0, 0.1, dH=script, lib=lib)
    rc, it = s.advance()
    assert rc == 0
    return s, it


def _floor_script():
    """FloorScript (:165-193): quadratic convergence, then a floor near 1e-10"""
    return [2.108252e-03, 3.576681e-05, 7.291922e-07, 1.669462e-08, 3.868198e-10,
            9.020539e-11, 8.684096e-11, 1.567753e-10, 9.732429e-11, 3.120448e-11,
            4.350973e-11, 1.190828e-10, 2.507970e-11, 1.002455e-10, 4.546102e-11,
            7.090800e-11, 7.484886e-11, 9.904685e-11, 9.818438e-11, 1.619510e-10]


def _stall_then_converge(trials, per_trial, v):
    """StallThenConverge (:290-297)"""
    return [v] * (trials * per_trial) + [1e-3, TINY]


# ---------------------------------------------------------------------------
# the reference's nine cases (test_newton_convergence.cpp:110-386)
# ---------------------------------------------------------------------------

def check_coupled_side_moving(lib):
    """:110-123 -- the thickness converged throughout, the coupled side
    settles at the third pass: three passes"""
    assert _run_one_step(lib, [TINY], [BIG, BIG, TINY], 1) == 3


def check_both_small(lib):
    """:126-133 -- both corrections small: one pass"""
    assert _run_one_step(lib, [TINY], [TINY], 1) == 1


def check_thickness_alone(lib):
    """:136-150 -- no coupled correction: the thickness alone decides"""
    assert _run_one_step(lib, [BIG, BIG, TINY], [], 1) == 3


def check_min_iterations(lib):
    """:153-166 -- NewtonMinIterations 6 and 3 are run; 0 counts as 1"""
    assert _run_one_step(lib, [TINY], [TINY], 6) == 6
    assert _run_one_step(lib, [TINY], [TINY], 3) == 3
    assert _run_one_step(lib, [TINY], [TINY], 0) == 1


def check_relative_test(lib):
    """:224-241 -- NewtonEpsilon 1e-12 below the floor; the relative test
    (1e-6 x 2.108e-3 = 2.11e-9 > 3.868e-10) exits at the fifth pass and the
    step is not shrunk"""
    s, it = _run_adj(lib, _floor_script(), 1e-12, 1e-6, 50)
    assert it == 5
    assert s.clock()[1] == pytest.approx(0.1)


def check_stagnation_exit(lib):
    """:244-260 -- relative test off: the floor exits by stagnation, after
    the fifth pass and well before 50, and counts as solved (four decades
    below the first correction), so the step is not shrunk. Worked by hand
    from the script: the best value improves (below 0.9 x best) at passes
    6, 10 and 13 and not again; the sixth pass without improvement after
    pass 13 is pass 19."""
    s, it = _run_adj(lib, _floor_script(), 1e-12, 0.0, 50)
    assert 5 < it < 25
    assert it == 19
    assert s.clock()[1] == pytest.approx(0.1)
    assert s.info()[1] == 0


def check_stagnation_margin(lib):
    """A correction that falls by 1 % a pass (0.99^k) has not improved on the
    first by the required 10 % six passes later (0.99^6 = 0.94): the 7th
    solve ends the attempt, far from convergence, and the fixed step fails
    (taken as improvements, the passes would run to NewtonMaxIterations)"""
    s = ob.ScriptedNewton(_case(lib), np.ones(4), 0.1, dH=[0.99 ** k for k in range(40)], lib=lib)
    assert s.advance()[0] == 1
    assert s.info()[0] == 7


def check_constant_correction_cut(lib):
    """:300-317 -- a constant correction stops on stagnation at its 7th
    pass (NewtonMaxIterations 7 not reached), the step is cut to half and
    converges in 2 passes: 9 solves, one stagnation cut"""
    s, it = _run_adj(lib, _stall_then_converge(1, 7, BIG), 1e-10, 1e-6, 7)
    calls, cuts, _ = s.info()
    assert calls == 9
    assert it == 2
    assert s.clock()[1] == pytest.approx(0.05)
    assert cuts == 1


def check_never_drops(lib):
    """:320-334 -- two trials stagnate at the first correction; the third,
    at a quarter of the step, converges: 16 solves, two cuts"""
    s, it = _run_adj(lib, _stall_then_converge(2, 7, 1.0), 1e-12, 1e-6, 50)
    calls, cuts, _ = s.info()
    assert calls == 16
    assert s.clock()[1] == pytest.approx(0.025)
    assert cuts == 2


def check_floor_1000_eps(lib):
    """:337-386 -- on a film of 1e5 (eps|phi| = 2.2e-11) a correction that
    stops at 1e-10 is within 1000 eps|phi|: 8 passes, accepted, not cut;
    the same stagnation on a film that starts at 0 is cut"""
    dH = [1e-9] + [1e-10] * 10
    s, it = _run_adj(lib, dH, 1e-12, 1e-6, 50, phi0=np.full(8, 1e5))
    assert it == 8
    assert s.clock()[1] == pytest.approx(0.1)
    assert s.info()[1] == 0
    s2, _ = _run_adj(lib, dH + [1e-3, TINY], 1e-12, 1e-6, 50)
    assert s2.clock()[1] == pytest.approx(0.05)
    assert s2.info()[1] == 1


# ---------------------------------------------------------------------------
# exits, damping, extrapolation (NewtonKantorovich.cpp:157-486), by hand
# ---------------------------------------------------------------------------

def check_damping(lib):
    """A correction larger than the last one is applied at alpha = 1/2:
    1 + 0.01 + 0.02/2 = 1.02 after three passes (1.03 undamped)"""
    s = ob.ScriptedNewton(_case(lib), np.ones(4), 0.1, dH=[1e-2, 2e-2, TINY], lib=lib)
    assert s.advance() == (0, 3)
    assert np.allclose(s.phi(), 1.02, rtol=0, atol=1e-15)


def check_extrapolation(lib):
    """Fixed step: phi 1 -> 1.001 in the first step; the second starts from
    phi^n + (dt/dt_old)(phi^n - phi^{n-1}) = 1.002 and converges at once.
    With extrapolation off it stays at 1.001."""
    for on, want in ((1, 1.002), (0, 1.001)):
        s = ob.ScriptedNewton(_case(lib, extrap=on), np.ones(4), 0.1, dH=[1e-3, TINY, TINY], lib=lib)
        assert s.advance() == (0, 2)
        assert s.advance() == (0, 1)
        assert np.allclose(s.phi(), want, rtol=0, atol=1e-15)


def check_extrapolation_guard(lib):
    """A film that thins 1 -> 0.4 would extrapolate to -0.2: a cell with
    liquid left without any blocks the extrapolation, and the step starts
    from phi^n = 0.4"""
    s = ob.ScriptedNewton(_case(lib), np.ones(4), 0.1, dH=[-0.6, TINY, TINY], lib=lib)
    assert s.advance()[0] == 0
    assert s.advance() == (0, 1)
    assert np.allclose(s.phi(), 0.4, rtol=0, atol=1e-15)


def check_retry_from_phi_n(lib):
    """The attempt from the extrapolated film (1.002) stagnates at 1 per
    pass (7 passes); it is made again from phi^n = 1.001 at the same fixed
    step, without counting a cut: 1.001 + 1e-4 in 2 passes"""
    dH = [1e-3, TINY] + [1.0] * 7 + [1e-4, TINY]
    s = ob.ScriptedNewton(_case(lib), np.ones(4), 0.1, dH=dH, lib=lib)
    assert s.advance() == (0, 2)
    assert s.advance() == (0, 2)
    calls, cuts, _ = s.info()
    assert calls == 11
    assert cuts == 0
    assert np.allclose(s.phi(), 1.0011, rtol=0, atol=1e-15)


def check_linear_failure(lib):
    """A solve that fails (return code 1) ends the attempt with no
    correction: under step control the step is cut to 0.05 and solved
    again from phi^n; that is no stagnation cut. At a fixed step the
    step fails (FailedFixedStep)."""
    s = ob.ScriptedNewton(_case(lib, step_control=True, crit=0.01), np.ones(4), 0.1,
                          dH=[1e-3, 1e-3, TINY], rc=[0, 1, 0], lib=lib)
    assert s.advance() == (0, 1)
    assert s.clock()[0] == pytest.approx(0.05)
    calls, cuts, _ = s.info()
    assert calls == 3 and cuts == 0
    assert np.allclose(s.phi(), 1.0, rtol=0, atol=1e-15)
    f = ob.ScriptedNewton(_case(lib), np.ones(4), 0.1, dH=[1e-3, 1e-3, TINY], rc=[0, 1, 0], lib=lib)
    assert f.advance()[0] == 1


def check_fixed_step_stagnation(lib):
    """A stagnation far from convergence at a fixed step fails the step"""
    s = ob.ScriptedNewton(_case(lib), np.ones(4), 0.1, dH=[1.0] * 8, lib=lib)
    assert s.advance()[0] == 1


# ---------------------------------------------------------------------------
# Eisenstat-Walker (NewtonKantorovich.cpp:108-126), by hand
# ---------------------------------------------------------------------------

def check_ew_sequence(lib):
    """|Rhs| 1, 1e-2, 1e-6, 1e-6, 5e-7: eta = 1e-4 (first pass), 0.9 x
    (1e-2)^2 = 9e-5, 0.9 x (1e-4)^2 = 9e-9, 0.9 capped at 1e-4, 0.225
    capped at 1e-4. With the input tolerance 1e-6 the third is raised to
    it; with EW off every pass takes the input tolerance."""
    dH = [1e-2, 1e-4, 1e-8, 1e-9, TINY]
    rn = [1.0, 1e-2, 1e-6, 1e-6, 5e-7]
    for lin_eps, want in ((1e-13, [1e-4, 9e-5, 9e-9, 1e-4, 1e-4]),
                          (1e-6, [1e-4, 9e-5, 1e-6, 1e-4, 1e-4])):
        s = ob.ScriptedNewton(_case(lib, eps=1e-12, rel=0.0, lin_eps=lin_eps), np.ones(4), 0.1,
                              dH=dH, rn=rn, lib=lib)
        assert s.advance() == (0, 5)
        eta = s.info()[2]
        assert np.allclose(eta, want, rtol=1e-12, atol=0), eta
    s = ob.ScriptedNewton(_case(lib, eps=1e-12, rel=0.0, ew=0, lin_eps=3e-9), np.ones(4), 0.1,
                          dH=dH, rn=rn, lib=lib)
    s.advance()
    assert np.all(s.info()[2] == 3e-9)


def check_ew_new_step(lib):
    """The first pass of every step takes 1e-4, whatever the last step's
    |Rhs|: step 1 |Rhs| 1, 1e-3 -> 1e-4, 9e-7; step 2 |Rhs| 1e-9 -> 1e-4"""
    s = ob.ScriptedNewton(_case(lib), np.ones(4), 0.1, dH=[1e-3, TINY, TINY], rn=[1.0, 1e-3, 1e-9],
                          lib=lib)
    s.advance()
    s.advance()
    assert np.allclose(s.info()[2], [1e-4, 9e-7, 1e-4], rtol=1e-12, atol=0)


# ---------------------------------------------------------------------------
# step control (ImplicitSolver.cpp:144-190, 271-307), by hand
# ---------------------------------------------------------------------------

def check_ltc_first_step(lib):
    """C_TimeSteps::Build starts DT().old at the initial step, so the first
    step is judged too, with phi^{n-1} = phi^n. dt 0.1, criterion 0.01,
    change factor 0.3:
      attempt 1: 1 -> 1.02, E = 2 (0.1)(0.02)/(0.2) = 0.02, 1 - 2 < 0: cut
        to dt 0.03;
      attempt 2 (dt_old 0.1, dt 0.03): 1 -> 1.01,
        E = 2 (0.3)(0.1 x 1.01 + 0.03 - 0.13)/0.13 = 0.0046, accepted."""
    s = ob.ScriptedNewton(_case(lib, step_control=True, crit=0.01, change=0.3), np.ones(4), 0.1,
                          dH=[0.02, TINY, 0.01, TINY], lib=lib)
    assert s.advance() == (0, 2)
    t, dt_now, dt_old = s.clock()
    assert t == pytest.approx(0.03) and dt_now == pytest.approx(0.03) and dt_old == pytest.approx(0.03)
    assert np.allclose(s.phi(), 1.01, rtol=0, atol=1e-15)
    assert s.info()[0] == 4


def check_ltc_three_levels(lib):
    """Three levels with dt != dt_old. Criterion 0.01, change 0.5.
      step 1 (dt 0.1): 1 -> 1.0005, E = 5e-4: accepted;
      step 2 attempt 1 (dt 0.1): from the extrapolation 1.001 a correction
        0.02: E = 0.02/1.0005 (the second difference over |phi^n|): cut;
      attempt 2 (dt 0.05, dt_old 0.1): from 1.0005 + 0.5 x 5e-4 = 1.00075
        a correction 0.01: E = 2 (0.5)(0.1 x 1.01075 + 0.05 - 0.15 x
        1.0005)/(0.15 x 1.0005) = 0.01/1.50075 = 0.0067: accepted (the
        factor 2 dt/dt_old = 1 decides it: at 2 it would be 0.0133)."""
    s = ob.ScriptedNewton(_case(lib, step_control=True, crit=0.01), np.ones(4), 0.1,
                          dH=[5e-4, TINY, 0.02, TINY, 0.01, TINY], lib=lib)
    assert s.advance() == (0, 2)
    assert s.advance() == (0, 2)
    t, dt_now, dt_old = s.clock()
    assert t == pytest.approx(0.15) and dt_now == pytest.approx(0.05) and dt_old == pytest.approx(0.05)
    assert np.allclose(s.phi(), 1.01075, rtol=0, atol=1e-14)
    assert s.info()[0] == 6


def check_ltc_keep_steps(lib):
    """The step grows after `keep` steps with a criterion above 0.9; the
    count is not reset by a step in between, and the longer step is taken
    from the next step on, capped at MaxDt. keep 3, criterion 0.01,
    MaxDt 0.15:
      steps 1, 2: corrections 1e-5 (E = 1e-5): count 1, 2;
      step 3: correction 5e-3 (E = 0.005, criterion 0.5): count stays 2;
      step 4: correction 1e-5: count 3 -> dt 0.1/0.5 = 0.2, capped 0.15,
        for step 5; the clock after step 4 is 0.4;
      step 5 (dt 0.15): the extrapolation (ratio 1.5) is exact, 0.55."""
    s = ob.ScriptedNewton(_case(lib, step_control=True, crit=0.01, keep=3, dt_max=0.15), np.ones(4), 0.1,
                          dH=[1e-5, TINY, 1e-5, TINY, 5e-3, TINY, 1e-5, TINY, TINY], lib=lib)
    for k in range(3):
        assert s.advance()[0] == 0
        t, dt_now, _ = s.clock()
        assert t == pytest.approx(0.1 * (k + 1)) and dt_now == pytest.approx(0.1)
    assert s.advance()[0] == 0
    t, dt_now, dt_old = s.clock()
    assert t == pytest.approx(0.4) and dt_now == pytest.approx(0.15) and dt_old == pytest.approx(0.1)
    assert s.advance() == (0, 1)
    assert s.clock()[0] == pytest.approx(0.55)


def check_ltc_scale_floor(lib):
    """ScaleFloor 0.5: the weight of a thin cell (0.01) is 0.5 x max|phi|,
    so a uniform correction 2e-3 gives E = max(2e-3/1, 2e-3/0.5) = 4e-3
    (criterion 0.6, accepted); without the floor the thin cell's 0.2 would
    cut the step"""
    phi0 = np.array([1.0, 0.01, 1.0, 1.0])
    s = ob.ScriptedNewton(_case(lib, step_control=True, crit=0.01, floor=0.5), phi0, 0.1,
                          dH=[2e-3, TINY], lib=lib)
    assert s.advance() == (0, 2)
    assert s.clock()[0] == pytest.approx(0.1)
    assert np.allclose(s.phi(), phi0 + 2e-3, rtol=0, atol=1e-15)


def check_ltc_is_zero(lib):
    """Math::IsZero is |x| <= DBL_MIN: a cell of 1e-310 has a subnormal
    weight, (dt_old + dt) w <= DBL_MIN, and is left out of the estimate
    (dividing by it would cut the step)"""
    phi0 = np.array([1.0, 1e-310, 1.0, 1.0])
    s = ob.ScriptedNewton(_case(lib, step_control=True, crit=0.01), phi0, 0.1, dH=[1e-5, TINY], lib=lib)
    assert s.advance() == (0, 2)
    assert s.clock()[0] == pytest.approx(0.1)


def check_dt_min(lib):
    """A cut below MinDt fails the step: the first attempt stagnates after 7
    solves, the cut to 0.05 is below MinDt 0.06, and no second attempt is
    made"""
    s = ob.ScriptedNewton(_case(lib, step_control=True, dt_min=0.06), np.ones(4), 0.1, dH=[1.0] * 20,
                          lib=lib)
    assert s.advance()[0] == 1
    assert s.info()[0] == 7


CHECKS = [v for k, v in sorted(globals().items()) if k.startswith("check_")]


@pytest.mark.parametrize("check", CHECKS, ids=[f.__name__[6:] for f in CHECKS])
def test_oracle(check):
    check(ob.load())


# ---------------------------------------------------------------------------
# planted mistakes: (replaced text, replacement, the checks it must fail)
# ---------------------------------------------------------------------------
PLANTS = {
    "ew_no_square": ("eta = 0.9 * ratio * ratio;", "eta = 0.9 * ratio;", ["ew_sequence"]),
    "ew_cap": ("const double kEtaMax = 1.0e-4;\n        double eta = kEtaMax;",
               "const double kEtaMax = 1.0e-3;\n        double eta = kEtaMax;", ["ew_sequence"]),
    "ew_no_floor": ("if (eta < eq.cs.lin_eps) eta = eq.cs.lin_eps;\n        ew_eta", "ew_eta", ["ew_sequence"]),
    "ew_across_steps": ("if (nk_pass > 1 && ew_prev > 0)", "if (ew_prev > 0)", ["ew_new_step"]),
    "stall_limit": ("const int stall_limit = 5;", "const int stall_limit = 4;",
                    ["stagnation_exit", "constant_correction_cut"]),
    "stall_no_margin": ("if (resid_eq < resid_best * 0.9)", "if (resid_eq < resid_best)", ["stagnation_margin"]),
    "stall_no_four_decades": ("|| (resid_first > 0 && resid_eq < resid_first * 1.0e-4)", "",
                              ["stagnation_exit"]),
    "stall_no_eps_phi": ("\n                    || (resid_eq <= floorPhi);", ";", ["floor_1000_eps"]),
    "no_coupled": ("if (coupled > resid_eq) resid_eq = coupled;", "", ["coupled_side_moving"]),
    "no_damping": ("(resid_eq > resid_eq_old) ? 0.5 : 1.0", "1.0", ["damping"]),
    "no_min_iter": ("if (outer >= eq.cs.nk_min - 1) {", "if (true) {", ["min_iterations"]),
    "no_relative": ("resid_eq < resid_first * eq.cs.nk_eps_rel) { done = true; break; }",
                    "resid_eq < 0.0) { done = true; break; }", ["relative_test"]),
    "extrap_no_ratio": ("phi_guess[i] = phi_old[i] + r * (phi_old[i] - phi_old2[i]);",
                        "phi_guess[i] = phi_old[i] + (phi_old[i] - phi_old2[i]);", ["ltc_three_levels"]),
    "extrap_no_guard": ("if (!eq.A.bc[i] && phi_old[i] > 0 && !(gg > 0)) ok = false;", "",
                        ["extrapolation_guard"]),
    "no_retry": ("if (!ok && extrap_move > 0) {", "if (false) {", ["retry_from_phi_n"]),
    "retry_counts_cut": ("                stag_cuts = cuts0;\n", "", ["retry_from_phi_n"]),
    "linfail_ignored": ("if (ch != 0) { done = false; break; }", "", ["linear_failure"]),
    "ltc_skip_first": ("if (is_zero(dt_old)) return 1.0;", "if (first_step) return 1.0;",
                       ["ltc_first_step"]),
    "ltc_no_ratio": ("2.0 * (dt_now / dt_old) * (dt_old", "2.0 * (dt_old", ["ltc_three_levels"]),
    "ltc_no_floor": ("if (w < wfloor) w = wfloor;", "", ["ltc_scale_floor"]),
    "ltc_exact_zero": ("const double e = is_zero(den) ? 0.0", "const double e = (den == 0) ? 0.0",
                       ["ltc_is_zero"]),
    "ltc_count_reset": ("if (crit > 0.90 && dt_now < eq.cs.dt_max) {",
                        "if (crit <= 0.90) tsc_count = 0;\n                    if (crit > 0.90 && dt_now < eq.cs.dt_max) {",
                        ["ltc_keep_steps"]),
    "ltc_grow_now": ("dt_request = std::min(dt_now / eq.cs.tsc_change, eq.cs.dt_max);",
                     "dt_request = dt_now = std::min(dt_now / eq.cs.tsc_change, eq.cs.dt_max);",
                     ["ltc_keep_steps"]),
    "ltc_no_cap": ("std::min(dt_now / eq.cs.tsc_change, eq.cs.dt_max)", "dt_now / eq.cs.tsc_change",
                   ["ltc_keep_steps"]),
    "ltc_fixed_half": ("dt_now *= eq.cs.tsc_change;", "dt_now *= 0.5;", ["ltc_first_step"]),
    "no_dt_min": ("if (dt_now < eq.cs.dt_min) { lin = (int)lin_sum; return 1; }",
                  "if (dt_now < 1e-300) { lin = (int)lin_sum; return 1; }", ["dt_min"]),
}


def _build_plant(tmp, name, old, new):
    src = open(os.path.join(ROOT, "oracle", "yz_oracle.cpp")).read()
    assert src.count(old) == 1, f"plant {name}: the text to replace is not in the oracle once"
    d = os.path.join(tmp, name, "oracle")
    os.makedirs(d)
    with open(os.path.join(d, "yz_oracle.cpp"), "w") as f:
        f.write(src.replace(old, new))
    for part in ("yz_cz_oracle.cpp", "yz_struct_mg.h"):
        shutil.copy(os.path.join(ROOT, "oracle", part), d)
    os.symlink(os.path.join(ROOT, "include"), os.path.join(tmp, name, "include"))
    so = os.path.join(d, "liboracle.so")
    subprocess.check_call(["g++", "-O0", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-w",
                           "-o", so, os.path.join(d, "yz_oracle.cpp")])
    return so


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("plants"))
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        futs = {k: ex.submit(_build_plant, tmp, k, o, n) for k, (o, n, _) in PLANTS.items()}
        return {k: f.result() for k, f in futs.items()}


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistake_fails(planted, plant):
    lib = ob.load_from(planted[plant])
    by_name = {f.__name__[6:]: f for f in CHECKS}
    for name in PLANTS[plant][2]:
        with pytest.raises(AssertionError):
            by_name[name](lib)



def test_modified_newton_jacobian_interval():
    """JacobianUpdateInterval k (NewtonKantorovich.cpp:80-91): IA is formed
    on the passes 0, k, 2k, ... of each attempt and kept in between. On the
    nonlinear all-physics film the step converges to the same film (to the
    Newton tolerance) in more passes, more the staler the Jacobian"""
    c, h, s, cc, dt, _ = cases.allphys_2d(n=32, L=1.5, dt=2e-5)
    runs = {}
    for ji in (1, 2, 3):
        cj = c.copy()
        cj.jac_interval = ji
        o = ob.OracleStep(cj, h, s, cc, dt, lin_mode=ob.OracleStep.LIN_STRUCTURED)
        nks = []
        for _ in range(4):
            r = o.advance()
            assert r[0] == 0
            nks.append(r[1])
        runs[ji] = (np.array(nks), o.h())
    print({k: list(v[0]) for k, v in runs.items()})
    ref = runs[1][1]
    for ji in (2, 3):
        assert np.abs(runs[ji][1] - ref).max() <= 1e-10 * np.abs(ref).max()
        assert (runs[ji][0] >= runs[1][0]).all() and runs[ji][0].sum() > runs[1][0].sum()
    assert runs[3][0].sum() > runs[2][0].sum()
