"""Device path against the CPU oracle (FP64 plain method).

Assembly: the rows of IA = I + theta dt J, A(h)h and the Newton right-hand
side, term by term, 1-D and 2-D, with and without slip, Symmetry and
Periodic edges, against the oracle to 1e-12 of each row's scale (the two
sum the same terms in orders that differ only in rounding).

Steps: per step max|dh| / max h <= 1e-10 with equal Newton counts.
"""
import numpy as np
import pytest

from yuzen_amd import capi, cases
from yuzen_amd.step import HipStep, LinSolver
from oracle import binding as ob

from .helpers import TERMS, smooth_case, csr_to_planes, interior

pytestmark = pytest.mark.gpu


def _assembly_case(dim, terms, slip, periodic):
    c, h, s, cc = smooth_case(dim, terms, slip=slip, periodic=periodic)
    o = ob.OracleStep(c, h, s, cc, 1e-3)
    o.set_h(h)
    hf = o.h()
    rng = np.random.default_rng(3)
    h_old = hf * (1.0 + 0.01 * rng.standard_normal(hf.size))
    dt = 2e-3
    A, J, IA, ax, rhs = o.assemble(h_old, hf, dt)
    rp, col, bc = o.pattern()
    ref_planes = csr_to_planes(c, rp, col, IA)
    g = HipStep(c, h, s, cc, 1e-3)
    ia, gax, grhs = g.assemble(h_old, hf, dt)
    g.close()
    return c, ref_planes, ia, ax, gax, rhs, grhs


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("slip", [0.0, 0.05])
@pytest.mark.parametrize("term", TERMS + ["all"])
def test_assembly_rows(dim, slip, term):
    terms = TERMS if term == "all" else [term]
    print(f"assembly dim={dim} slip={slip} term={term}", flush=True)
    c, ref, ia, ax, gax, rhs, grhs = _assembly_case(dim, terms, slip, periodic=False)
    scale = np.abs(ref).max(axis=0)
    assert np.all(np.abs(ia - ref) <= 1e-12 * scale + 1e-300), np.abs(ia - ref).max()
    axi, gaxi = interior(c, ax), interior(c, gax)
    assert np.abs(axi - gaxi).max() <= 1e-12 * max(np.abs(axi).max(), 1e-300)
    ri, gri = interior(c, rhs), interior(c, grhs)
    assert np.abs(ri - gri).max() <= 1e-12 * max(np.abs(ri).max(), 1e-300) + 1e-15


@pytest.mark.parametrize("dim", [1, 2])
def test_assembly_periodic(dim):
    print(f"assembly periodic dim={dim}", flush=True)
    c, ref, ia, ax, gax, rhs, grhs = _assembly_case(dim, TERMS, 0.05, periodic=True)
    scale = np.abs(ref).max(axis=0)
    assert np.all(np.abs(ia - ref) <= 1e-12 * scale)
    ri, gri = interior(c, rhs), interior(c, grhs)
    assert np.abs(ri - gri).max() <= 1e-12 * np.abs(ri).max()


def _run_pair(c, h, s, cc, dt, steps, label):
    o = ob.OracleStep(c, h, s, cc, dt)
    g = HipStep(c, h, s, cc, dt)
    worst = 0.0
    for k in range(steps):
        ro = o.advance()
        rg = g.advance()
        ho, hg = o.h(), g.h()
        rel = np.abs(ho - hg).max() / np.abs(ho).max()
        worst = max(worst, rel)
        print(f"{label} step {k}: oracle nk={ro[1]} lin={ro[2]}  hip nk={rg[1]} lin={rg[2]}  rel={rel:.2e}", flush=True)
        assert ro[0] == 0 and rg[0] == 0
        assert ro[1] == rg[1], "Newton counts differ"
        assert rel <= 1e-10
        if c.step_control:
            assert abs(o.clock()[1] - g.clock()[1]) <= 1e-15 * o.clock()[1]
    g.close()
    return worst


def test_step_leveling_1d():
    c, h, s, cc, dt, _ = cases.leveling_1d(nx=512)
    _run_pair(c, h, s, cc, dt, 8, "leveling_1d")


def test_step_allphys_2d_small():
    c, h, s, cc, dt, _ = cases.allphys_2d(n=48, L=2.0, dt=2e-5)
    _run_pair(c, h, s, cc, dt, 4, "allphys_2d")


def test_step_allphys_2d_zero_guess():
    """The plain FP64 path with every pass's solve started from zero
    (zero_guess 1, a departure from the reference's guess) against the
    oracle run the same way."""
    c, h, s, cc, dt, _ = cases.allphys_2d(n=48, L=2.0, dt=2e-5)
    c.zero_guess = 1
    _run_pair(c, h, s, cc, dt, 4, "allphys_2d zero guess")


def test_step_spinodal_2d_small():
    c, h, s, cc, dt, _ = cases.spinodal_2d(n=32)
    _run_pair(c, h, s, cc, dt, 4, "spinodal_2d")


def test_linsolve_seam_matches_oracle():
    c, h, s, cc = smooth_case(2, TERMS, slip=0.05)
    o = ob.OracleStep(c, h, s, cc, 1e-3)
    o.set_h(h)
    hf = o.h()
    A, J, IA, ax, rhs = o.assemble(hf * 1.01, hf, 5e-3)
    rp, col, bc = o.pattern()
    rc0, x0, it0, r0 = ob.linsolve(rp, col, IA, rhs, np.zeros(o.n), eps=1e-12)
    ls = LinSolver()
    for grid in (None, (2, c.nx, c.ny, [0, 0, 0, 0])):
        rc, x, it, r = ls.solve(rp, col, IA, rhs, np.zeros(o.n), epsilon=1e-12, grid=grid)
        print(f"linsolve grid={grid is not None}: rc={rc} it={it} resid={r:.2e} (oracle it={it0})", flush=True)
        assert rc == 0 and rc0 == 0
        assert np.abs(x - x0).max() <= 1e-9 * np.abs(x0).max()


@pytest.mark.parametrize("dim", [1, 2])
def test_linsolve_seam_takes_the_concentration_system(dim):
    """A concentration system of the reference's layout (the 9-point wide
    stencil in 2-D, ghost rows) through the seam with the concentration
    role: the grid read off the pattern, the solution to the tolerance."""
    if dim == 2:
        c, h, s, cc, dt, _ = cases.allphys_2d(n=40, L=1.5, dt=2e-5)
    else:
        c, h, s, cc, dt, _ = cases.leveling_1d(nx=128, steps=1)
        c.marangoni, c.As, c.Ma = 1, 1.0, 0.3
        X, _ = cases.cell_centres(c)
        cc = (0.5 + 0.2 * np.tanh((X - 4.0) / 0.5)).ravel()
    c.conc_model, c.Df = 1, 0.05
    o = ob.OracleStep(c, h, s, cc, dt)
    assert o.advance()[0] == 0
    rp, col, val, rhs = o.conc_system()
    rc0, x0, it0, r0 = ob.linsolve(rp, col, val, rhs, np.zeros(o.n), eps=1e-13)
    ls = LinSolver()
    rc, x, it, r = ls.solve(rp, col, val, rhs, np.zeros(o.n), epsilon=1e-13, conc=1)
    print(f"concentration seam {dim}-D: rc={rc} it={it} resid={r:.2e} (oracle it={it0})", flush=True)
    assert rc == 0 and rc0 == 0
    assert np.abs(x - x0).max() <= 1e-10 * max(np.abs(x0).max(), 1e-300)
    assert np.abs(x0).max() > 0
