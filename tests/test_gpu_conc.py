"""The well-mixed concentration on the device (conc_model 1) through the
C-ABI, against the CPU oracle and the reference's own face fluxes.

* Face volume fluxes (FaceVolumeFluxes, every term's FaceFlux / SideRate):
  the device against the reference's golden values of every term, 1-D and
  2-D, with and without slip, Symmetry and Periodic, to 1e-12 of the scale.
* Steps: the device's film and concentration against the oracle's after
  every step, max |dh| / max h <= 1e-10 and max |dc| <= 1e-10, with the
  same Newton counts, for the all-physics film (Symmetry and Periodic),
  the 1-D leveling film with Marangoni flow, and a front near c = 1 that
  the UTOPIA face values overshoot (RedistributeExcess on both).
* The limited schemes (Koren, van Leer): the same bars and counts, and
  the same number of concentration solves (the step's completion repeated
  until the frozen limiter settles), on a front carried over a periodic
  line and on the all-physics film.
* The learned FP32 inner solve with the concentration: film and
  concentration against the oracle to 1e-8.
* At 1024^2 (the all-physics film with the concentration), properties
  that do not depend on the size: the solute sum over the cells is kept to
  rounding, and a uniform concentration stays uniform.
"""
import json
import os

import numpy as np
import pytest

from yuzen_amd import capi, cases
from yuzen_amd.step import HipStep
from oracle import binding as ob

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
MAN = json.load(open(os.path.join(GOLD, "manifest.json")))
TERMS = np.load(os.path.join(GOLD, "ref_terms.npz"))


def _case_from(m):
    c = capi.case_defaults()
    c.dim, c.nx, c.ny = m["dim"], m["nx"], m["ny"]
    c.xmin, c.xmax, c.ymin, c.ymax = m["xmin"], m["xmax"], m["ymin"], m["ymax"]
    for k in range(4):
        c.bc[k] = capi.BC_PERIODIC if m["periodic"] else capi.BC_SYMMETRY
    for t in ("capillary", "vertical_gravity", "inplane_gravity", "centrifugal", "marangoni", "disjoining"):
        setattr(c, t, 1 if m["term"] == t else 0)
    c.As, c.Ma, c.Bo, c.Pgx, c.Pgy, c.Sp, c.Dp = m["As"], m["Ma"], m["Bo"], m["Pgx"], m["Pgy"], m["Sp"], m["Dp"]
    c.hstar, c.dis_n, c.dis_m = m["hstar"], 3.0, 2.0
    c.slip, c.rot_x, c.rot_y = m["slip"], m["rot_x"], m["rot_y"]
    c.visc_c0, c.visc_exp, c.sig_c0 = m["visc_c0"], m["visc_exp"], m["sig_c0"]
    c.height_scheme = m.get("height_scheme", capi.HSCHEME_KOREN)
    return c


def _interior(c, v):
    nyp, nxp = cases.padded_shape(c)
    a = np.asarray(v).reshape(nyp, nxp)
    return a[2:-2, 2:-2] if c.dim == 2 else a[0, 2:-2]


def test_face_fluxes_match_reference_and_oracle():
    for m in MAN["cases"]:
        g = {k: TERMS[f"{m['name']}__{k}"] for k in ("h", "s", "c", "ff", "side")}
        c = _case_from(m)
        d = HipStep(c, g["h"], g["s"], g["c"], 1e-3)
        F, r = d.face_fluxes(g["h"], g["c"])
        o = ob.OracleStep(c, g["h"], g["s"], g["c"], 1e-3)
        Fo, ro = o.face_fluxes(g["h"], g["c"])
        d.close()
        ff = g["ff"]
        scale = max(np.abs(ff).max(), 1e-300)
        nxp = c.nx + 4
        n = len(g["h"])
        nyp = n // nxp
        for p in range(n):
            ix, iy = p % nxp, p // nxp
            if not (2 <= ix <= nxp - 3 and (c.dim == 1 or 2 <= iy <= nyp - 3)):
                continue
            for a in range(c.dim):
                lo = p - (1 if a == 0 else nxp)
                assert abs(F[a][p] - ff[p, 2 * a + 1]) <= 1e-12 * scale, (m["name"], p, a)
                assert abs(F[a][lo] - ff[p, 2 * a]) <= 1e-12 * scale, (m["name"], p, a, "low")
        assert np.abs(F[:c.dim] - Fo[:c.dim]).max() <= 1e-12 * scale, m["name"]
        assert np.abs(r - g["side"]).max() <= 1e-12 * max(np.abs(g["side"]).max(), 1e-300)
        assert np.abs(r - ro).max() <= 1e-12 * max(np.abs(ro).max(), 1e-300)


def _pair(c, h, s, cc, dt, steps, label, bar=1e-10, counts=True):
    o = ob.OracleStep(c, h, s, cc, dt, lin_mode=ob.OracleStep.LIN_STRUCTURED)
    g = HipStep(c, h, s, cc, dt)
    worst_h = worst_c = 0.0
    moved = 0.0
    for k in range(steps):
        print(f"{label}: step {k}", flush=True)
        ro, rg = o.advance(), g.advance()
        assert ro[0] == 0 and rg[0] == 0, (ro, rg)
        if counts:
            assert ro[1] == rg[1], f"step {k}: Newton counts {ro[1]} / {rg[1]}"
        ho, hg = o.h(), g.h()
        co, cg = o.c(), g.c()
        rh = np.abs(ho - hg).max() / np.abs(ho).max()
        rc = np.abs(co - cg).max()
        worst_h, worst_c = max(worst_h, rh), max(worst_c, rc)
        assert rh <= bar, f"step {k}: film {rh:.2e}"
        assert rc <= bar, f"step {k}: concentration {rc:.2e}"
        so, sg = o.conc_stats(), g.conc_stats()
        moved += sg[3] + sg[4]
        # the excess moved is a sum of differences of near-equal values: it
        # carries the film's rounding, so it is held to the film's bar
        assert abs(so[3] - sg[3]) <= bar * max(so[3], 1e-300) + 1e-300
        if counts:
            # the same concentration solves, the completing re-solves of a
            # limited scheme included
            assert so[0] == sg[0], f"step {k}: concentration solves {so[0]:.0f} / {sg[0]:.0f}"
    st = g.conc_stats()
    print(f"{label}: {steps} steps, worst film {worst_h:.2e}, concentration {worst_c:.2e}; "
          f"device solves {st[0]:.0f}, iterations {st[1]:.0f}, multigrid {st[2]:.0f}, moved {moved:.3e}")
    g.close()
    return moved


@pytest.mark.parametrize("periodic", [False, True])
def test_allphys_with_concentration_matches_oracle(periodic):
    c, h, s, cc, dt, _ = cases.allphys_2d(n=40, L=1.5, dt=2e-5)
    c.conc_model, c.Df, c.diff_exp = 1, 0.05, 0.7
    if periodic:
        # a film that is periodic itself: no front across the seam
        for k in range(4):
            c.bc[k] = capi.BC_PERIODIC
        c.centrifugal = 0
        X, Y = cases.cell_centres(c)
        h = (0.6 + 0.25 * np.sin(2 * np.pi * X / 1.5) * np.cos(2 * np.pi * Y / 1.5)).ravel()
    _pair(c, h, s, cc, dt, 6, f"all-physics 40^2 {'periodic' if periodic else 'symmetry'}")


def test_leveling_1d_with_marangoni_matches_oracle():
    c, h, s, cc, dt, _ = cases.leveling_1d(nx=128, steps=10)
    c.conc_model, c.Df = 1, 0.2
    c.marangoni, c.As, c.Ma = 1, 1.0, 0.3
    c.centrifugal, c.Sp, c.rot_x = 1, 0.4, 5.0
    X, _ = cases.cell_centres(c)
    cc = (0.5 + 0.2 * np.tanh((X - 4.0) / 0.5)).ravel()
    _pair(c, h, s, cc, dt, 10, "leveling 1-D with Marangoni")


def test_overshoot_is_redistributed_as_the_oracle_does():
    c = capi.case_defaults()
    c.dim, c.nx, c.ny = 1, 48, 1
    c.xmin, c.xmax, c.ymin, c.ymax = 0.0, 1.0, -1.0, 1.0
    for k in range(4):
        c.bc[k] = capi.BC_PERIODIC
    c.capillary = 1
    c.conc_model, c.conc_diffusion = 1, 0
    c.inplane_gravity, c.Pgx = 1, 2.0
    c.nk_eps = 1e-12
    X, _ = cases.cell_centres(c)
    h = (np.ones_like(X) + 0.1 * np.sin(2 * np.pi * X)).ravel()
    conc = np.where((X > 0.3) & (X < 0.6), 0.999, 0.2).ravel()
    moved = _pair(c, h, np.zeros_like(h), conc, 4e-3, 6, "overshoot 1-D")
    assert moved > 0.0


@pytest.mark.parametrize("scheme", [capi.CONV_KOREN, capi.CONV_VANLEER], ids=["koren", "vanleer"])
def test_limited_schemes_match_oracle(scheme):
    """Koren and van Leer: the weights frozen at the latest iterate, the
    step completed by re-solves until that settles; a front carried over a
    periodic line, and the all-physics film with Symmetry walls"""
    c = capi.case_defaults()
    c.dim, c.nx, c.ny = 1, 60, 1
    c.xmin, c.xmax, c.ymin, c.ymax = 0.0, 1.0, -1.0, 1.0
    for k in range(4):
        c.bc[k] = capi.BC_PERIODIC
    c.capillary = 1
    c.conc_model, c.conc_diffusion, c.conc_scheme = 1, 0, scheme
    c.inplane_gravity, c.Pgx = 1, 0.6
    X, _ = cases.cell_centres(c)
    h = (np.ones_like(X) + 0.1 * np.sin(2 * np.pi * X)).ravel()
    conc = np.where((X > 0.3) & (X < 0.6), 0.8, 0.2).ravel()
    _pair(c, h, np.zeros_like(h), conc, 2e-3, 8, f"limited {scheme} front 1-D")
    c, h, s, cc, dt, _ = cases.allphys_2d(n=40, L=1.5, dt=2e-5)
    c.conc_model, c.Df, c.diff_exp, c.conc_scheme = 1, 0.05, 0.7, scheme
    _pair(c, h, s, cc, dt, 4, f"limited {scheme} all-physics 40^2")


def test_learned_fp32_with_concentration():
    c, h, s, cc, dt, _ = cases.allphys_2d(n=48, L=2.0, dt=2e-5)
    c.conc_model, c.Df = 1, 0.05
    c.precond, c.krylov_fp32, c.learn_passes, c.learn_iters = capi.PC_LEARNED, 1, 1, 40
    _pair(c, h, s, cc, dt, 3, "learned FP32 48^2 with concentration", bar=1e-8, counts=False)


def test_full_size_solute_and_uniformity():
    c, h, s, cc, dt, _ = cases.allphys_2d(n=1024, dt=2e-6)
    c.conc_model, c.Df = 1, 0.05
    for label, conc in (("varying", cc), ("uniform", np.full_like(cc, 0.37))):
        g = HipStep(c, h, s, conc, dt)
        sol0 = (_interior(c, g.h()) * _interior(c, g.c())).sum()
        for k in range(2):
            print(f"1024^2 {label}: step {k}", flush=True)
            rc, nk, lin, _r = g.advance()
            assert rc == 0
            hi, ci = _interior(c, g.h()), _interior(c, g.c())
            sol = (hi * ci).sum()
            print(f"  Newton {nk}, Krylov {lin}, solute change {abs(sol - sol0) / sol0:.2e}, "
                  f"concentration stats {g.conc_stats()}")
            assert abs(sol - sol0) <= 1e-12 * sol0
            if label == "uniform":
                assert np.abs(ci - 0.37).max() <= 1e-14
        g.close()
