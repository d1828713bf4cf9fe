"""The oracle's second way of solving its linear systems: the device's
structured smoothed-aggregation multigrid restated on the CPU (OracleStep
lin_mode 1), beside the reference's algebraic one (lin_mode 0).

Both solve the same systems to the backend contract's relative residual, so
the films they give agree far inside the parity bar whatever their
different error vectors do to the Newton counts; and the closed form of a
linear leveling mode holds in the structured mode as in the other. These pin
the structured solve as a solver of the reference's systems (folded ghost
rows, row scaling, the right-hand side's ghost offsets, the Jacobi try and
the multigrid of the concentration's solves); the GPU tests pin it against
the device's own solves through equal Newton counts."""
import numpy as np
import pytest

from yuzen_amd import capi, cases
from oracle import binding as ob

from .test_oracle_physics import amplitude, leveling_case

S, A = ob.OracleStep.LIN_STRUCTURED, ob.OracleStep.LIN_ALGEBRAIC


def _both(c, h, s, cc, dt, steps, bar):
    os_, oa = ob.OracleStep(c, h, s, cc, dt, lin_mode=S), ob.OracleStep(c, h, s, cc, dt, lin_mode=A)
    worst = 0.0
    for k in range(steps):
        rs, ra = os_.advance(), oa.advance()
        assert rs[0] == 0 and ra[0] == 0, (k, rs, ra)
        hs, ha = os_.h(), oa.h()
        rel = np.abs(hs - ha).max() / np.abs(ha).max()
        worst = max(worst, rel)
        assert rel <= bar, f"step {k}: {rel:.2e}"
        if c.step_control:
            assert os_.clock()[1] == oa.clock()[1]
    return worst, os_


def test_leveling_1d_symmetry_films_agree():
    c, h, s, cc, dt, _ = cases.leveling_1d(nx=512, steps=200)
    worst, _ = _both(c, h, s, cc, dt, 30, 1e-12)
    print(f"leveling 512, 30 steps: films {worst:.2e} apart")


def test_spinodal_periodic_step_control_films_agree():
    c, h, s, cc, dt, _ = cases.spinodal_2d(n=48)
    worst, _ = _both(c, h, s, cc, dt, 12, 1e-11)
    print(f"spinodal 48^2 periodic under step control, 12 steps: films {worst:.2e} apart")


def test_allphys_symmetry_with_concentration_films_agree():
    """all terms, Symmetry walls, the well-mixed concentration solved in
    every pass (its Jacobi try and multigrid through the structured solve)"""
    c, h, s, cc, dt, _ = cases.allphys_2d(n=40, L=1.5, dt=2e-5)
    c.conc_model = 1
    c.conc_diffusion = 1
    worst, _ = _both(c, h, s, cc, dt, 3, 1e-11)
    print(f"all-physics 40^2 with the concentration solved, 3 steps: films {worst:.2e} apart")


def test_linear_leveling_closed_form_structured():
    """the Crank-Nicolson factor of a small cos mode (test_oracle_physics's
    closed form) in the structured mode: Symmetry edges fold the mode onto
    itself, so the solve's folding of ghost rows is on the path"""
    n, dt, steps = 48, 5e-4, 6
    c, h, s, cc = leveling_case(n, periodic=False, amp=1e-6)
    o = ob.OracleStep(c, h, s, cc, dt, lin_mode=S)
    dx = 1.0 / n
    k = 2 * np.pi
    lam = (1.0 / 3.0) * (2 - 2 * np.cos(k * dx)) ** 2 / dx ** 4
    g = (1 - dt * lam / 2) / (1 + dt * lam / 2)
    a0 = amplitude(c, o.h())
    for _ in range(steps):
        assert o.advance()[0] == 0
    a1 = amplitude(c, o.h())
    assert abs(a1 / a0 - g ** steps) <= 1e-5 * abs(g ** steps), (a1 / a0, g ** steps)
