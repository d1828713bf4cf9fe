"""Closed forms and invariants that pin the oracle's whole step (the
reference's own checks of the same step: docs/validation.md analytic
leveling and conservation; yuzen/tests/test_linear_solver.cpp for the
backend contract)."""
import numpy as np
import pytest

from yuzen_amd import capi, cases
from oracle import binding as ob


def leveling_case(n, periodic=True, amp=1e-3, wavelength=1.0):
    c = capi.case_defaults()
    c.dim, c.nx, c.ny = 1, n, 1
    c.xmin, c.xmax, c.ymin, c.ymax = 0.0, wavelength, -1.0, 1.0
    for k in range(4):
        c.bc[k] = capi.BC_PERIODIC if periodic else capi.BC_SYMMETRY
    c.capillary = 1
    c.nk_eps = 1e-13
    X, _ = cases.cell_centres(c)
    h = 1.0 + amp * np.cos(2 * np.pi * X / wavelength)
    return c, h.ravel(), np.zeros(h.size), np.full(h.size, 0.5)


def amplitude(c, h):
    X, _ = cases.cell_centres(c)
    x = X.ravel()[2:-2]
    return 2.0 * np.mean((h[2:-2] - 1.0) * np.cos(2 * np.pi * x))


def test_linear_leveling_discrete_crank_nicolson():
    """A small mode cos kx of a uniform film (h0 = sigma = mu = 1) is an
    eigenvector of the discrete capillary operator with eigenvalue
    lambda = (1/3)(2 - 2 cos k dx)^2 / dx^4; Crank-Nicolson multiplies its
    amplitude by (1 - dt lambda/2)/(1 + dt lambda/2) per step. The whole
    step -- assembly, Jacobian, Newton loop, linear solve -- must give that
    factor to the size of the nonlinear correction (amplitude 1e-6)."""
    n, dt, steps = 48, 5e-4, 6
    c, h, s, cc = leveling_case(n, amp=1e-6)
    o = ob.OracleStep(c, h, s, cc, dt)
    a0 = amplitude(c, o.h())
    for _ in range(steps):
        assert o.advance()[0] == 0
    dx = 1.0 / n
    k = 2 * np.pi
    lam = (2 - 2 * np.cos(k * dx)) ** 2 / dx ** 4 / 3.0
    g = (1 - 0.5 * dt * lam) / (1 + 0.5 * dt * lam)
    assert abs(amplitude(c, o.h()) / a0 - g ** steps) <= 1e-5 * abs(g ** steps)


def test_linear_leveling_converges_in_space():
    """The decay approaches the continuum exp(-k^4 t / 3) at second order
    in dx (dt small enough that the time error is far below)."""
    errs = {}
    for nn in (32, 64, 128):
        c, h, s, cc = leveling_case(nn, amp=1e-5)
        dt, steps = 2e-5, 50
        o = ob.OracleStep(c, h, s, cc, dt)
        a0 = amplitude(c, o.h())
        for _ in range(steps):
            assert o.advance()[0] == 0
        exact = np.exp(-((2 * np.pi) ** 4) * steps * dt / 3.0)
        errs[nn] = abs(amplitude(c, o.h()) / a0 - exact)
    o1, o2 = np.log2(errs[32] / errs[64]), np.log2(errs[64] / errs[128])
    assert 1.8 < o1 < 2.2 and 1.8 < o2 < 2.2, errs


def test_volume_conserved_closed_box():
    """Symmetry walls: the volume is a sum of face fluxes that cancel."""
    c, h, s, cc = leveling_case(64, periodic=False, amp=0.2)
    c.vertical_gravity, c.Bo = 1, 0.5
    c.disjoining, c.Dp, c.hstar = 1, cases.disjoining_dp(10.0, 0.05), 0.05
    c.inplane_gravity, c.Pgx = 1, 0.3
    c.slip = 0.05
    o = ob.OracleStep(c, h, s, cc, 1e-4)
    v0 = o.h()[2:-2].sum()
    for _ in range(5):
        assert o.advance()[0] == 0
    v1 = o.h()[2:-2].sum()
    assert abs(v1 - v0) <= 1e-11 * v0


def poisson_crs(n):
    """u'' = 2 on [-1, 1], u(+-1) = 0 (exact x^2 - 1): the reference's
    test_linear_solver.cpp problem, central differences, Dirichlet rows."""
    x = np.linspace(-1, 1, n)
    dx = x[1] - x[0]
    rp, col, val, b = [0], [], [], np.zeros(n)
    for i in range(n):
        if i == 0 or i == n - 1:
            col += [i]; val += [1.0]; b[i] = 0.0
        else:
            col += [i, i - 1, i + 1]; val += [-2.0 / dx ** 2, 1.0 / dx ** 2, 1.0 / dx ** 2]; b[i] = 2.0
        rp.append(len(col))
    return x, np.array(rp), np.array(col), np.array(val), b


@pytest.mark.parametrize("precond", [capi.PC_JACOBI, capi.PC_AMG])
def test_poisson_through_backend_contract(precond):
    x, rp, col, val, b = poisson_crs(202)
    val0, b0 = val.copy(), b.copy()
    rc, u, it, r = ob.linsolve(rp, col, val, b, np.zeros_like(b), precond=precond, eps=1e-12)
    assert rc == 0 and r <= 1e-12
    assert np.abs(u - (x ** 2 - 1)).max() < 1e-4
    assert np.array_equal(val, val0) and np.array_equal(b, b0)   # left as given
    # repeat solves are bit-identical
    rc2, u2, it2, r2 = ob.linsolve(rp, col, val, b, np.zeros_like(b), precond=precond, eps=1e-12)
    assert np.array_equal(u, u2) and it == it2
    # zero right-hand side: the trivial solution
    rc3, u3, it3, r3 = ob.linsolve(rp, col, val, np.zeros_like(b), np.ones_like(b), precond=precond)
    assert rc3 == 0 and np.all(u3 == 0)
