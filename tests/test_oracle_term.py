"""The oracle's termination tests of a run (WellMixed_ModelBase::
CheckTermination, yuzen/src/model/well_mixed/WellMixedModelBase.h:1567-1679,
with FieldInfo over the real cells, IO/FieldInfo.cpp:52-136, and
ChangeRateOfVariable, numerics/ImplicitSolver.cpp:225-240), on films whose
outcome is known by hand:

  * a flat film at rest does not move: its change rate is 0 from the fifth
    step on (before it the rate counts as 1), so the steady count reaches 11
    > SteadyStateSteps 10 at step 15, and not before;
  * Numerics.MaxSteps 3 ends the run after step 3; EndTime 2.5 dt after
    step 3;
  * a real cell below -1e-3 ends it (diverged), one at -5e-4 does not; a
    film below 0 everywhere within that slack has a negative volume; a NaN
    ends it before anything else;
  * the dry-out test on the mean, max or min concentration against
    EndConcValue;
  * with the concentration solved, a cell below -0.1 at the step's start
    ends the run at its first step.

Each check also fails with a mistake planted in a copy of the oracle.
"""
import numpy as np
import pytest

from oracle import binding as ob
from yuzen_amd import capi, cases

from .test_oracle_newton import _build_plant


def _flat(n=16, h0=1.0, conc=0.5, per=True):
    c = capi.case_defaults()
    c.dim, c.nx, c.ny = 2, n, n
    c.xmin, c.xmax, c.ymin, c.ymax = 0.0, 1.0, 0.0, 1.0
    for k in range(4):
        c.bc[k] = capi.BC_PERIODIC if per else capi.BC_SYMMETRY
    nyp, nxp = cases.padded_shape(c)
    h = np.full(nyp * nxp, h0)
    s = np.zeros_like(h)
    cc = np.full_like(h, conc)
    return c, h, s, cc


def _set_cell(c, v, ix, iy, val):
    nyp, nxp = cases.padded_shape(c)
    a = v.reshape(nyp, nxp).copy()
    a[iy + 2, ix + 2] = val
    return a.ravel()


def check_steady_at_rest(lib):
    c, h, s, cc = _flat()
    o = ob.OracleStep(c, h, s, cc, 1e-3)
    t = capi.term(steady_eps=1e-3, steady_steps=10)
    got = []
    for k in range(16):
        assert o.advance()[0] == 0
        st, info = o.check(t)
        got.append(st)
    assert got[:14] == [capi.RUN_GOING] * 14
    assert got[14] == capi.RUN_STEADY
    assert info[6] == 12 and info[5] == 0.0


def check_end(lib):
    c, h, s, cc = _flat()
    for t, want in ((capi.term(max_steps=3, steady_eps=0), [0, 0, capi.RUN_END]),
                    (capi.term(end_time=2.5e-3, steady_eps=0), [0, 0, capi.RUN_END])):
        o = ob.OracleStep(c, h, s, cc, 1e-3)
        got = []
        for k in range(3):
            assert o.advance()[0] == 0
            got.append(o.check(t)[0])
        assert got == want


def check_divergence(lib):
    c, h, s, cc = _flat()
    o = ob.OracleStep(c, h, s, cc, 1e-3)
    t = capi.term(steady_eps=0)
    o.set_h(_set_cell(c, h, 3, 4, -2e-3))
    assert o.check(t)[0] == capi.RUN_NEGATIVE_FILM
    o.set_h(_set_cell(c, h, 3, 4, -5e-4))
    assert o.check(t)[0] == capi.RUN_GOING
    o.set_h(np.full_like(h, -5e-4))
    st, info = o.check(t)
    assert st == capi.RUN_NEGATIVE_VOLUME and info[0] < 0
    bad = _set_cell(c, h, 3, 4, np.nan)
    bad = _set_cell(c, bad, 5, 5, -1.0)
    o.set_h(bad)
    assert o.check(t)[0] == capi.RUN_NAN


def check_dry_out(lib):
    c, h, s, cc = _flat(conc=0.97)
    cc = _set_cell(c, cc, 2, 2, 0.90)
    o = ob.OracleStep(c, h, s, cc, 1e-3)
    assert o.advance()[0] == 0
    # mean c = 0.97 - 0.07 / 256 = 0.96973, max 0.97, min 0.90
    cases_ = [(0, 0.95, capi.RUN_DRY), (0, 0.9698, capi.RUN_GOING), (1, 0.96, capi.RUN_DRY),
              (1, 0.975, capi.RUN_GOING), (2, 0.89, capi.RUN_DRY), (2, 0.91, capi.RUN_GOING)]
    for typ, val, want in cases_:
        st, info = o.check(capi.term(steady_eps=0, dry=1, end_conc_type=typ, end_conc_value=val))
        assert st == want, (typ, val, st)
    assert abs(info[1] - (0.97 - 0.07 / 256)) < 1e-14
    st, _ = o.check(capi.term(steady_eps=0, dry=0, end_conc_value=0.5))
    assert st == capi.RUN_GOING


def check_conc_range_at_start(lib):
    c, h, s, cc = _flat(n=12, conc=0.4, per=False)
    c.conc_model, c.Df = 1, 0.1
    cc = _set_cell(c, cc, 5, 5, -0.2)
    o = ob.OracleStep(c, h, s, cc, 1e-4)
    assert o.advance()[0] == 0
    st, info = o.check(capi.term(steady_eps=0))
    assert st == capi.RUN_CONC_RANGE and info[4] == -0.2


CHECKS = [v for k, v in sorted(globals().items()) if k.startswith("check_")]


@pytest.mark.parametrize("check", CHECKS, ids=[f.__name__[6:] for f in CHECKS])
def test_oracle(check):
    check(None)


PLANTS = {
    "steady_ge": ("if (t.steady_eps > 0 && steady_count > t.steady_steps) return YZ_RUN_STEADY;",
                  "if (t.steady_eps > 0 && steady_count >= t.steady_steps) return YZ_RUN_STEADY;", "steady_at_rest"),
    "rate_from_start": ("if (steps_done < 5) return 1;", "if (steps_done < 1) return 1;", "steady_at_rest"),
    "max_steps_off_by_one": ("steps_done <= t.max_steps - 1", "steps_done <= t.max_steps", "end"),
    "film_slack": ("if (minH < -1.0e-3) return YZ_RUN_NEGATIVE_FILM;", "if (minH < -1.0e-2) return YZ_RUN_NEGATIVE_FILM;",
                   "divergence"),
    "no_start_conc": ("if (conc_on) minC = std::min(minC, min_start_c);", "", "conc_range_at_start"),
    "dry_on_max_always": ("const double v = t.end_conc_type == 0 ? ave : (t.end_conc_type == 1 ? maxC : minC);",
                          "const double v = maxC;", "dry_out"),
}


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("term_plants"))
    return {k: _build_plant(tmp, k, o, n) for k, (o, n, _) in PLANTS.items()}


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistake_fails(planted, plant, monkeypatch):
    lib = ob.load_from(planted[plant])
    monkeypatch.setattr(ob, "_lib", lib)
    by_name = {f.__name__[6:]: f for f in CHECKS}
    with pytest.raises(AssertionError):
        by_name[PLANTS[plant][2]](lib)
