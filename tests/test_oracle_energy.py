"""The energy budget (Output.Energy) on the oracle.

* The free energy F and the potential phi against the reference's own
  EnergyBudget.h (FreeEnergy, Potential, run on the term fixtures by
  oracle/ref_golden.cpp): capillary, disjoining and vertical-gravity cases,
  1-D and 2-D, Symmetry and Periodic, to the bit.
* D, the power of the face fluxes against phi (FluxPower), combines phi
  with the face fluxes, which match the reference's FaceFlux
  (tests/test_oracle_golden.py). It is pinned by what it must satisfy: the
  discrete identity of the Crank-Nicolson step,
  (F^{n+1} - F^n)/dt + (D^n + D^{n+1})/2 = O(dt^2), on a closed film and on
  a periodic one with the capillary, disjoining and vertical-gravity terms
  and a uniform concentration (the residual over D falls by at least 3 when
  dt halves), with D > 0. A flux or a potential of the wrong sign, scale or
  face breaks it at O(1).
"""
import numpy as np
import pytest

from oracle import binding as ob

from .helpers import smooth_case
from .test_oracle_golden import MAN, TERMS, case_from


@pytest.mark.parametrize("m", [m for m in MAN["cases"] if m["term"] in ("capillary", "disjoining", "vertical_gravity")],
                         ids=lambda m: m["name"])
def test_free_energy_and_potential_match_reference(m):
    g = {k: TERMS[f"{m['name']}__{k}"] for k in ("h", "s", "c", "energy", "phi")}
    o = ob.OracleStep(case_from(m), g["h"], g["s"], g["c"], 1e-3)
    F, _, phi = o.energy(g["h"], g["c"])
    assert F == g["energy"][0]
    assert np.array_equal(phi, g["phi"])


def _cosine_film(dim):
    """a closed film even about its walls (a sum of cosines that the
    Symmetry ghosts continue smoothly)"""
    c, _, _, _ = smooth_case(dim, ["capillary", "disjoining", "vertical_gravity"])
    nxp, nyp = c.nx + 4, (c.ny + 4 if dim == 2 else 1)
    x = c.xmin + (np.arange(nxp) - 1.5) * (c.xmax - c.xmin) / c.nx
    y = c.ymin + (np.arange(nyp) - 1.5) * (c.ymax - c.ymin) / c.ny if dim == 2 else np.zeros(1)
    X, Y = np.meshgrid(x, y)
    Lx, Ly = c.xmax - c.xmin, c.ymax - c.ymin
    h = 0.6 + 0.25 * np.cos(np.pi * (X - c.xmin) / Lx)
    if dim == 2:
        h = h + 0.15 * np.cos(2 * np.pi * (Y - c.ymin) / Ly)
    s = 0.05 * np.cos(2 * np.pi * (X - c.xmin) / Lx)
    return c, h.ravel(), s.ravel()


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("periodic", [False, True])
def test_crank_nicolson_energy_identity(dim, periodic):
    res = []
    for dt in (4e-5, 2e-5, 1e-5):
        if periodic:
            c, h, s, _ = smooth_case(dim, ["capillary", "disjoining", "vertical_gravity"], periodic=True)
        else:
            c, h, s = _cosine_film(dim)
        c.ew, c.nk_eps = 0, 1e-13
        cc = np.full(h.size, 0.5)
        o = ob.OracleStep(c, h, s, cc, dt)
        F0, D0, _ = o.energy(o.h(), cc)
        worst = 0.0
        for _ in range(3):
            assert o.advance()[0] == 0
            F1, D1, _ = o.energy(o.h(), cc)
            assert D1 > 0 and F1 < F0
            worst = max(worst, abs((F1 - F0) / dt + 0.5 * (D0 + D1)) / D1)
            F0, D0 = F1, D1
        res.append(worst)
    print(dim, periodic, res)
    assert res[2] < 1e-2
    assert res[0] / res[1] > 3.0 and res[1] / res[2] > 3.0
