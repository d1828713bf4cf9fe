"""The CPU oracle against golden vectors written by the reference's own code.

tests/golden holds what the reference's term kernels (SetMatrixA,
SetJacobiMatrix, AddConservative; the upwinded terms with the Koren and
the van Leer limiter), Mobility, both limiters and
NonDimensionalNumbers compute on the shapes of its test_jacobian.cpp
harness (tests/golden/make_golden.py, built by oracle/build_ref.sh from the
reference sources). The oracle must reproduce them to rounding: 1e-12 of
each row's scale (the two sum the same terms in orders that differ).
"""
import json
import os

import numpy as np
import pytest

from yuzen_amd import capi, cases
from oracle import binding as ob

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
MAN = json.load(open(os.path.join(GOLD, "manifest.json")))
TERMS = np.load(os.path.join(GOLD, "ref_terms.npz"))
SCAL = np.load(os.path.join(GOLD, "ref_scalars.npz"))


def case_from(m):
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


def as_dict(rp, col, val, rows):
    out = {}
    for i in rows:
        for k in range(rp[i], rp[i + 1]):
            out[(i, int(col[k]))] = out.get((i, int(col[k])), 0.0) + float(val[k])
    return out


@pytest.mark.parametrize("m", MAN["cases"], ids=[m["name"] for m in MAN["cases"]])
def test_face_fluxes_match_reference(m):
    """FaceVolumeFluxes of one term against that term's own FaceFlux and
    SideRate: the high face of every real cell, the low face of a first
    cell (stored at its ghost row), and every other low face against the
    same face taken from the cell below it (the reference forms a face the
    same way whichever cell asks)."""
    g = {k: TERMS[f"{m['name']}__{k}"] for k in ("h", "s", "c", "ff", "side")}
    c = case_from(m)
    o = ob.OracleStep(c, g["h"], g["s"], g["c"], 1e-3)
    _, _, bc = o.pattern()
    F, r = o.face_fluxes(g["h"], g["c"])
    ff = g["ff"]
    scale = max(np.abs(ff).max(), 1e-300)
    nxp = c.nx + 4
    checked = 0
    for i in np.nonzero(~bc)[0]:
        for a in range(c.dim):
            lo = i - (1 if a == 0 else nxp)
            assert abs(F[a][i] - ff[i, 2 * a + 1]) <= 1e-12 * scale, (m["name"], i, a)
            assert abs(F[a][lo] - ff[i, 2 * a]) <= 1e-12 * scale, (m["name"], i, a, "low")
            checked += 2
    assert checked > 0
    assert np.abs(r - g["side"]).max() <= 1e-12 * max(np.abs(g["side"]).max(), 1e-300)


@pytest.mark.parametrize("m", MAN["cases"], ids=[m["name"] for m in MAN["cases"]])
def test_face_velocities_match_reference(m):
    """One term's face velocities (its CalVelocity, summed by
    Cal_MixedVelocity) against the reference's, on every face the term
    writes (the equation's boundary rows are not part of the term)."""
    g = {k: TERMS[f"{m['name']}__{k}"] for k in ("h", "s", "c", "vx", "vy")}
    c = case_from(m)
    o = ob.OracleStep(c, g["h"], g["s"], g["c"], 1e-3)
    vx, vy = o.velocity(g["h"], g["c"], bc=False)
    for got, want in ((vx, g["vx"]), (vy, g["vy"])):
        scale = max(np.abs(want).max(), 1e-300)
        assert np.abs(got - want).max() <= 1e-12 * scale, (m["name"], np.abs(got - want).max() / scale)
    assert np.count_nonzero(g["vx"]) > 0


def test_conc_schemes_match_reference():
    for row in SCAL["schemes"]:
        k, sign, f, v, w = int(row[0]), row[1], row[2:6], row[6], row[7:11]
        mine = ob.conc_face(k, f, sign)
        assert mine[0] == v and np.array_equal(mine[1:], w), (k, sign, mine, row)


def test_limited_conc_weights_match_reference():
    """Koren's and van Leer's FaceValue and FaceWeights with the row's cell
    upwind (L for a positive flux) and downwind, through zero differences"""
    t = SCAL["limited_weights"]
    forms = set()
    for row in t:
        vl, sign, rowpos, f, v, w = int(row[0]), row[1], int(row[2]), row[3:7], row[7], row[8:12]
        mine = ob.conc_face(capi.CONV_VANLEER if vl else capi.CONV_KOREN, f, sign, rowpos)
        assert mine[0] == v and np.array_equal(mine[1:], w), (row, mine)
        forms.add((vl, sign > 0, rowpos))
    assert len(forms) == 8


@pytest.mark.parametrize("m", MAN["cases"], ids=[m["name"] for m in MAN["cases"]])
def test_terms_match_reference(m):
    g = {k: TERMS[f"{m['name']}__{k}"] for k in ("rp", "col", "av", "rpj", "colj", "jv", "h", "s", "c", "ax")}
    c = case_from(m)
    o = ob.OracleStep(c, g["h"], g["s"], g["c"], 1e-3)
    rp, col, bc = o.pattern()
    A, J, IA, ax, rhs = o.assemble(g["h"], g["h"], 1e-3)
    interior = np.nonzero(~bc)[0]
    ga = as_dict(g["rp"], g["col"], g["av"], interior)
    gj = as_dict(g["rpj"], g["colj"], g["jv"], interior)
    oa = as_dict(rp, col, A, interior)
    oj = as_dict(rp, col, J, interior)
    for gold, mine, what in ((ga, oa, "A"), (gj, oj, "J")):
        rowscale = {}
        for (i, j), v in gold.items():
            rowscale[i] = max(rowscale.get(i, 0.0), abs(v))
        for key in set(gold) | set(mine):
            a, b = gold.get(key, 0.0), mine.get(key, 0.0)
            assert abs(a - b) <= 1e-12 * max(rowscale.get(key[0], 0.0), 1e-300), (what, key, a, b)
    gax = g["ax"][interior]
    assert np.abs(ax[interior] - gax).max() <= 1e-12 * max(np.abs(gax).max(), 1e-300)


def test_mobility_matches_reference():
    t = SCAL["mobility"]
    worst = 0.0
    for lam, h1, h2, F, F1, F2, D, dD, DperH in t:
        mine = ob.mobility(lam, h1, h2)
        ref = np.array([F, F1, F2, D, dD, DperH])
        worst = max(worst, float(np.max(np.abs(mine - ref) / np.maximum(np.abs(ref), 1e-300))))
    assert worst <= 1e-13


@pytest.mark.parametrize("vanleer", [0, 1], ids=["koren", "vanleer"])
def test_limiter_matches_reference(vanleer):
    t = SCAL["limiter_vl" if vanleer else "limiter"]
    for row in t:
        f, sign = row[0:4], row[4]
        mine = ob.limiter(f, sign, vanleer)
        ref = np.concatenate([[row[5]], row[6:10], [row[10]], row[11:15]])
        assert np.allclose(mine, ref, rtol=0, atol=1e-15), (row, mine)


def test_disjoining_number_matches_reference():
    for th, hstar, n, m, dp in SCAL["dp"]:
        mine = cases.disjoining_dp(th, hstar, n, m)
        assert abs(mine - dp) <= 1e-14 * abs(dp)
