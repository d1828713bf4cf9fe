"""The ConcZ step on the device (yz_cz_*) through the C-ABI, against the
CPU oracle (oracle/yz_cz_oracle.cpp).

* Per step: the film to max |dh| / max h <= 1e-10 and the concentration to
  max |dc| <= 1e-10 of the oracle's, with the same number of coupling
  passes and the same outcome, on the all-physics film (Periodic, and
  Symmetry over a substrate relief: capillary, gravity, Marangoni flow
  with its coupling passes, evaporation, a viscosity that depends on c),
  on a film without the Marangoni term (the concentration's own predictor
  and midpoint stages), on a drying film whose top cells pass c = 1
  (RedistributeExcess) and on a front carried into a film with no solute
  (RedistributeDeficit within and between columns); a step whose coupling
  passes do not settle fails on both, with the state of its start kept.
* The volume fluxes of the last concentration solve against the oracle's.
* At the 256^2 x 32 film with evaporation and Marangoni flow, properties
  that do not depend on the size: the solute is kept to rounding, the
  volume falls by what evaporates, and without evaporation a uniform
  concentration stays uniform.
"""
import numpy as np
import pytest

from oracle import binding as ob
from tests.cz_cases import _allphys, _case, _grid, _real2, _real3, _solute
from yuzen_amd import capi, cases
from yuzen_amd.step import HipCZ

pytestmark = pytest.mark.gpu


def _pair(c, h, s, cc, dt, steps, label, bar=1e-10):
    o = ob.OracleCZ(c, h, s, cc, dt)
    g = HipCZ(c, h, s, cc, dt)
    worst_h = worst_c = 0.0
    for k in range(steps):
        print(f"{label}: step {k}", flush=True)
        ro, rg = o.advance(), g.advance()
        assert ro[0] == rg[0], (ro, rg)
        assert ro[2] == rg[2], f"step {k}: coupling passes {ro[2]} / {rg[2]}"
        ho, hg = o.h(), g.h()
        co, cg = o.c(), g.c()
        rh = np.abs(ho - hg).max() / np.abs(ho).max()
        rc = np.abs(co - cg).max()
        worst_h, worst_c = max(worst_h, rh), max(worst_c, rc)
        assert rh <= bar, f"step {k}: film {rh:.2e}"
        assert rc <= bar, f"step {k}: concentration {rc:.2e}"
    so, sg = o.stats(), g.stats()
    print(f"{label}: {steps} steps, worst film {worst_h:.2e}, concentration {worst_c:.2e}; "
          f"oracle {so}, device {sg}")
    return o, g


@pytest.mark.parametrize("per", [True, False])
def test_allphys_matches_oracle(per):
    c, h, s, cc = _allphys(per=per)
    o, g = _pair(c, h, s, cc, 0.02, 4, f"all-physics 12^2 x 8 {'periodic' if per else 'symmetry'}")
    assert g.stats()[4] > 0
    # the fluxes carry third differences of the film (1 / dx^3 = 27 here),
    # which the two films' solvers leave 1e-13 apart
    qo, qg = o.volume_fluxes(), g.volume_fluxes()
    for a, b in zip(qo, qg):
        scale = max(np.abs(a).max(), 1e-300)
        assert np.abs(a - b).max() <= 1e-8 * scale
    g.close()


@pytest.mark.parametrize("scheme", [capi.CONV_KOREN, capi.CONV_VANLEER], ids=["koren", "vanleer"])
def test_limited_schemes_match_oracle(scheme):
    """Koren and van Leer: the weights frozen at the field each stage is
    solved at, SettleLimiter's re-solves; the same coupling passes and
    concentration solves as the oracle, film and c within 1e-10, on the
    all-physics film and on a front carried along x"""
    c, h, s, cc = _allphys(per=False)
    c.conc_scheme = scheme
    o, g = _pair(c, h, s, cc, 0.02, 3, f"limited {scheme} all-physics 12^2 x 8")
    assert o.stats()[2] == g.stats()[2] and o.settle() > 0
    n, L = 16, 4.0
    c = _case(n=n, nz=4, L=L, per=True, inplane_gravity=1, Pgx=2.0, Df=1e-3)
    c.conc_scheme = scheme
    x = _grid(n, L)
    X, _ = np.meshgrid(x, x)
    h = np.ones_like(X).ravel()
    cc = np.broadcast_to(np.where((X > 1.0) & (X < 2.5), 0.6, 0.2)[None], (c.nz + 4, n + 4, n + 4)).copy()
    o, g = _pair(c, h, np.zeros_like(h), cc.ravel(), 0.05, 4, f"limited {scheme} front 16^2 x 4")
    assert o.stats()[2] == g.stats()[2]
    c3 = _real3(c, g.c())
    assert c3.min() >= 0.2 - 1e-13 and c3.max() <= 0.6 + 1e-13


def test_step_control_matches_oracle():
    """the step control: the same cuts and the same accepted steps as the
    oracle, films and c within 1e-10 after every step"""
    c, h, s, cc = _allphys(per=False)
    c.step_control, c.tsc_keep, c.tsc_criterion = 1, 2, 2e-4
    o, g = _pair(c, h, s, cc, 0.02, 6, "step control all-physics 12^2 x 8")
    co, cg = o.clock(), g.clock()
    print(f"oracle clock {co}, device clock {cg}")
    assert co[0] == cg[0] and co[1] == cg[1] and co[2] == cg[2] and co[2] >= 9


def test_without_marangoni_matches_oracle():
    c, h, s, cc = _allphys(per=False, n=14, nz=6)
    c.marangoni = 0
    _, g = _pair(c, h, s, cc, 0.02, 4, "no Marangoni 14^2 x 6")
    assert g.stats()[4] == 0
    g.close()


def test_excess_redistribution_matches_oracle():
    n, nz, L = 12, 8, 4.0
    x = _grid(n, L)
    X, Y = np.meshgrid(x, x)
    c = _case(n=n, nz=nz, L=L, evaporation=1, Ev=0.5, evap_func=capi.EVAP_POWERN, evap_power=0.0, As=0.5)
    h = 1.0 + 0.05 * np.cos(2 * np.pi * X / L)
    cc = np.full((nz + 4, n + 4, n + 4), 0.97) + 0.01 * np.sin(2 * np.pi * Y / L)[None]
    o, g = _pair(c, h.ravel(), np.zeros(h.size), cc.ravel(), 0.01, 5, "excess")
    assert g.stats()[5] > 0
    assert abs(o.stats()[5] - g.stats()[5]) <= 1e-10 * o.stats()[5]
    assert _real3(c, g.c()).max() <= 1.0
    g.close()


def test_deficit_redistribution_matches_oracle():
    n, nz, L = 12, 8, 4.0
    x = _grid(n, L)
    X, Y = np.meshgrid(x, x)
    c = _case(n=n, nz=nz, L=L, per=True, inplane_gravity=1, Pgx=3.0, Pgy=1.0, Df=0.001, As=0.05)
    h = 1.0 + 0.2 * np.cos(2 * np.pi * X / L)
    cc = np.where((X > 1.0) & (X < 2.0), 0.5, 0.0)[None].repeat(nz + 4, 0)
    o, g = _pair(c, h.ravel(), np.zeros(h.size), cc.ravel(), 0.2, 3, "deficit")
    assert g.stats()[6] > 0
    assert abs(o.stats()[6] - g.stats()[6]) <= 1e-10 * o.stats()[6]
    assert _real3(c, g.c()).min() >= 0.0
    g.close()


def test_unsettled_passes_fail_on_both_and_keep_the_start():
    n, nz, L = 12, 8, 4.0
    x = _grid(n, L)
    X, Y = np.meshgrid(x, x)
    c = _case(n=n, nz=nz, L=L, evaporation=1, Ev=2.0, evap_func=capi.EVAP_POWERN, evap_power=0.0, As=0.02,
              marangoni=1, Ma=0.5)
    h = (1.0 + 0.05 * np.cos(2 * np.pi * X / L)).ravel()
    cc = (np.full((nz + 4, n + 4, n + 4), 0.9) + 0.05 * np.sin(2 * np.pi * Y / L)[None]).ravel()
    o = ob.OracleCZ(c, h, np.zeros(h.size), cc, 0.05)
    g = HipCZ(c, h, np.zeros(h.size), cc, 0.05)
    ro, rg = o.advance(), g.advance()
    assert ro[0] == 3 and rg[0] == 3, (ro, rg)
    assert np.array_equal(g.h(), h) and np.array_equal(g.c(), cc)
    g.close()


def _big(evap=1):
    c, h, s, cc, _, _ = cases.concz_2d()
    c.evaporation = evap
    return c, h, s, cc


def test_concz256_size_solute_volume_and_uniformity():
    dt = 0.01
    c, h, s, cc = _big()
    g = HipCZ(c, h, s, cc, dt)
    sol0 = _solute(c, h, cc)
    v = _real2(c, h).sum()
    for k in range(3):
        print(f"256^2 x 32: step {k}", flush=True)
        rc, lin, passes, _ = g.advance()
        assert rc == 0
        H, C3 = g.h(), g.c()
        sol = _solute(c, H, C3)
        _, _, w = g.volume_fluxes()
        evap = _real2(c, w.reshape(c.nz + 4, c.ny + 4, c.nx + 4)[c.nz + 2].ravel()).sum()
        vn = _real2(c, H).sum()
        print(f"  linear iterations {lin}, passes {passes}, solute change {abs(sol - sol0) / sol0:.2e}, "
              f"volume balance {abs((v - vn) - dt * evap) / v:.2e}")
        assert abs(sol - sol0) <= 1e-12 * sol0
        assert abs((v - vn) - dt * evap) <= 1e-12 * v
        v = vn
    g.close()
    c, h, s, cc = _big(evap=0)
    cc = np.full_like(cc, 0.37)
    g = HipCZ(c, h, s, cc, dt)
    for k in range(2):
        assert g.advance()[0] == 0
    assert np.abs(_real3(c, g.c()) - 0.37).max() <= 1e-14
    g.close()


def test_learned_film_solve_matches_oracle():
    """The film's solve with the learned preconditioner (film_precond
    YZ_PC_LEARNED: FP32 BiCGSTAB with the network's approximate inverse
    inside FP64 iterative refinement to the same h_eps, FP64 AMG finishing
    a solve whose round fails to halve the residual) reaches the film the
    oracle's FP64 AMG reaches, with the same coupling passes."""
    c, h, s, cc = _allphys(per=True, n=40, nz=8)
    c.film_precond = capi.PC_LEARNED
    c.learn_passes, c.learn_iters = 2, 60
    _, g = _pair(c, h, s, cc, 0.02, 4, "learned film solve 40^2 x 8", bar=1e-9)
    ls = g.learn_stats()
    print(f"learned film solve: rounds {ls[0]:.0f}, FP32 iterations {ls[1]:.0f}, AMG finished {ls[2]:.0f}, "
          f"training {ls[3]:.1f} ms")
    assert ls[0] > 0 and ls[1] > 0
    g.close()


def test_film_precond_refused():
    c, h, s, cc = _allphys(per=True)
    c.film_precond = capi.PC_JACOBI
    with pytest.raises(Exception):
        HipCZ(c, h, s, cc, 0.02)
