"""Evaporation and the nozzle's supply in B of the well-mixed film step on
the device, through the C-ABI, against the CPU oracle (whose rate law and
supply match the reference's own code to the bit, and whose steps match
closed forms: tests/test_oracle_evap_source.py).

* Assembly: the Newton right-hand side with - dt B, 1-D and 2-D, to 1e-12
  of its scale.
* Steps: per step max |dh| / max h <= 1e-10 with equal Newton counts, on
  the all-physics film with a Raoult rate, its precursor ramp and a
  rotating nozzle, also where the solvent and half-film caps bind, under
  step control, and on the FP32 inner solve with the learned
  preconditioner (1e-9, as its other tests).
* The batched sweep with B: films advanced together reach, to the bit,
  what each reaches alone.
* B with the well-mixed concentration solved (evaporation and the nozzle
  at the local or the nozzle's concentration): film and c against the
  oracle step by step with equal Newton counts; a uniform c stays uniform
  under a nozzle supplying at the local c.
"""
import numpy as np
import pytest

from oracle import binding as ob
from yuzen_amd import capi, cases
from yuzen_amd.step import HipBatch, HipStep

from .helpers import TERMS, interior, smooth_case

pytestmark = pytest.mark.gpu


def _with_b(c, coef=50.0, evap=True, source=True):
    """a Raoult rate with its precursor ramp (h* = 0.15); a nozzle on a
    fast arm whose flow rises over the first steps"""
    if evap:
        c.evaporation, c.Ev, c.evap_coef = 1, 1.0, coef
        c.evap_func, c.precursor = capi.EVAP_RAOULT, 0.15
    if source:
        c.source, c.src_radius, c.src_rate, c.src_tau = 1, 0.12, 200.0, 2e-5
        c.src_arm_length, c.src_arm_x, c.src_arm_y = 0.2, 0.05, -0.02
        c.src_omega, c.src_angle0, c.src_ramp, c.src_plateau = 3000.0, 0.3, 1e-5, 10.0
        c.src_subregions = 4
    return c


DT = 1e-5


def _film(dim, coef=50.0):
    """The smooth all-physics film with B. Its passes are solved to
    lin_eps (no Eisenstat-Walker tolerance): the oracle's AMG solves these
    small systems exactly in one iteration, the device's BiCGSTAB stops at
    the tolerance, and a pass solved to EW's 1e-4 leaves an error of the
    size of the next correction, which can move the Newton exit by a pass
    (the films stay 1e-13 apart). The EW path is compared on the other
    cases here."""
    c, h, s, cc = smooth_case(dim, TERMS, slip=0.05)
    c.ew = 0
    return _with_b(c, coef), h, s, cc


@pytest.mark.parametrize("dim", [1, 2])
def test_assembly_with_b_matches_oracle(dim):
    print(f"assembly with B dim={dim}", flush=True)
    c, h, s, cc = _film(dim)
    o = ob.OracleStep(c, h, s, cc, DT)
    o.set_h(h)
    hf = o.h()
    rng = np.random.default_rng(5)
    h_old = hf * (1.0 + 0.01 * rng.standard_normal(hf.size))
    dt = 2 * DT
    rhs = o.assemble(h_old, hf, dt)[4]
    g = HipStep(c, h, s, cc, DT)
    grhs = g.assemble(h_old, hf, dt)[2]
    g.close()
    ri, gri = interior(c, rhs), interior(c, grhs)
    assert np.abs(ri - gri).max() <= 1e-12 * np.abs(ri).max()
    # B is a visible part of it
    c0, _, _, _ = smooth_case(dim, TERMS, slip=0.05)
    o0 = ob.OracleStep(c0, h, s, cc, DT)
    r0 = interior(c0, o0.assemble(h_old, hf, dt)[4])
    assert np.abs(ri - r0).max() > 3e-4 * np.abs(ri).max()


def _run_pair(c, h, s, cc, dt, steps, label, tol=1e-10, counts=True):
    o = ob.OracleStep(c, h, s, cc, dt, lin_mode=ob.OracleStep.LIN_STRUCTURED)
    g = HipStep(c, h, s, cc, dt)
    for k in range(steps):
        print(f"{label} step {k}", flush=True)
        ro, rg = o.advance(), g.advance()
        ho, hg = o.h(), g.h()
        rel = np.abs(ho - hg).max() / np.abs(ho).max()
        print(f"  oracle nk={ro[1]} lin={ro[2]}  hip nk={rg[1]} lin={rg[2]}  rel={rel:.2e}", flush=True)
        assert ro[0] == 0 and rg[0] == 0
        if counts:
            assert ro[1] == rg[1], "Newton counts differ"
        assert rel <= tol
        if c.step_control:
            assert abs(o.clock()[1] - g.clock()[1]) <= 1e-15 * o.clock()[1]
    g.close()
    return o


@pytest.mark.parametrize("dim", [1, 2])
def test_steps_match_oracle(dim):
    c, h, s, cc = _film(dim)
    o = _run_pair(c, h, s, cc, DT, 5, f"evaporation + nozzle dim={dim}")
    # B moved the film visibly: against the same steps without it
    c0 = smooth_case(dim, TERMS, slip=0.05)[0]
    c0.ew = 0
    o0 = ob.OracleStep(c0, h, s, cc, DT)
    for _ in range(5):
        assert o0.advance()[0] == 0
    assert np.abs(o.h() - o0.h()).max() > 1e-3


@pytest.mark.parametrize("dim", [1, 2])
def test_steps_with_binding_caps_match_oracle(dim):
    """coef 1e5 at dt 1e-5: on the thick part of the film, where the
    precursor ramp lets the rate grow, the caps of EvapAt bind at the
    start (then B depends on h, J does not carry it, and the Newton loop
    takes more passes)."""
    c, h, s, cc = _film(dim, coef=1e5)
    c.source = 0
    o = ob.OracleStep(c, h, s, cc, DT)
    rates = np.array([o.evap(hh, ci, DT) for hh, ci in zip(interior(c, h), interior(c, cc))])
    assert np.sum(rates[:, 1] < rates[:, 0]) >= 5
    _run_pair(c, h, s, cc, DT, 3, f"capped evaporation dim={dim}")


def test_step_control_with_b_matches_oracle():
    c, h, s, cc, dt, _ = cases.spinodal_2d(n=32)
    assert c.step_control
    _with_b(c, coef=0.2)
    c.src_arm_x = c.src_arm_y = 0.5 * (c.xmin + c.xmax)
    c.src_arm_length, c.src_radius = 0.25 * (c.xmax - c.xmin), 0.1 * (c.xmax - c.xmin)
    c.src_rate = 0.05
    _run_pair(c, h, s, cc, dt, 4, "spinodal 32^2 adaptive with B")


def test_fp32_learned_with_b_matches_oracle():
    c, h, s, cc, dt, _ = cases.allphys_2d(n=48, L=2.0, dt=2e-5)
    _with_b(c, coef=2.0)
    c.src_arm_x = c.src_arm_y = 1.0
    c.src_arm_length, c.src_radius, c.src_rate = 0.5, 0.3, 20.0
    c.precond, c.krylov_fp32 = capi.PC_LEARNED, 1
    _run_pair(c, h, s, cc, dt, 4, "learned FP32 48^2 with B", tol=1e-9, counts=False)


def test_batch_with_b_matches_films_alone():
    from .test_gpu_batch import _alone, _weights
    params = _weights()
    films = []
    for k, n in enumerate((40, 64)):
        c, h, s, cc, dt = cases.sweep_film(k, side=n)
        _with_b(c, coef=5.0)
        L = c.xmax - c.xmin
        c.src_arm_x, c.src_arm_y = c.xmin + 0.5 * L, c.ymin + 0.5 * (c.ymax - c.ymin)
        c.src_arm_length, c.src_radius, c.src_rate = 0.25 * L, 0.15 * L, 50.0
        films.append((c, h, s, cc, dt))
    steps = 3
    ref = _alone(films, params, steps)
    b = HipBatch(films)
    b.set_params(params)
    counts = [[] for _ in films]
    for _ in range(steps):
        rc, st, nk, lin = b.advance()
        assert rc == 0 and np.all(st == 0)
        for k in range(len(films)):
            counts[k].append((int(nk[k]), int(lin[k])))
    for k, (rcounts, rh, _) in enumerate(ref):
        print(f"film {k}: alone {rcounts}; batch {counts[k]}")
        assert counts[k] == rcounts
        assert np.array_equal(b.h(k), rh)
    b.close()


def _conc_pair(c, h, s, cc, dt, steps, label):
    """film and concentration of the device against the oracle's, step by
    step: the film to 1e-10, c to 1e-10 of its scale, equal Newton counts"""
    o = ob.OracleStep(c, h, s, cc, dt)
    g = HipStep(c, h, s, cc, dt)
    try:
        for k in range(steps):
            ro, rg = o.advance(), g.advance()
            ho, hg, co, cg = o.h(), g.h(), o.c(), g.c()
            rh = np.abs(ho - hg).max() / np.abs(ho).max()
            rc = np.abs(interior(c, co) - interior(c, cg)).max() / np.abs(interior(c, co)).max()
            print(f"{label} step {k}: oracle nk={ro[1]} hip nk={rg[1]} dh {rh:.2e} dc {rc:.2e}", flush=True)
            assert ro[0] == 0 and rg[0] == 0
            assert ro[1] == rg[1]
            assert rh <= 1e-10 and rc <= 1e-10
    finally:
        g.close()
    return o


@pytest.mark.parametrize("supplied", [False, True])
@pytest.mark.parametrize("dim", [1, 2])
def test_b_with_well_mixed_concentration_matches_oracle(dim, supplied):
    """evaporation and the nozzle with the concentration solved in every
    pass (ConcEquationBase.h:310-318, StepSources, ConsistentFluxes and
    Reconstruct with the sources); the oracle's balance and closed forms are
    pinned in tests/test_oracle_conc.py"""
    c, h, s, cc = _film(dim)
    c.conc_model, c.Df, c.diff_exp = 1, 0.05, 0.7
    c.src_conc_supplied, c.src_conc = (1, 0.9) if supplied else (0, 0.0)
    _conc_pair(c, h, s, cc, DT, 4, f"conc + B dim={dim} supplied={supplied}")


def test_uniform_concentration_stays_uniform_under_a_local_nozzle():
    c, h, s, cc = _film(2)
    c.evaporation = 0
    c.conc_model, c.Df, c.diff_exp = 1, 0.05, 0.7
    cu = np.full_like(cc, 0.37)
    g = HipStep(c, h, s, cu, DT)
    try:
        for _ in range(3):
            assert g.advance()[0] == 0
            assert np.abs(interior(c, g.c()) - 0.37).max() <= 1e-14
    finally:
        g.close()
