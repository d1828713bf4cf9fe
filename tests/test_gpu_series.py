"""The run's time series on the device (yz_step_field_info,
yz_step_series_open / yz_step_series_write: History.dat, CalcInfo.dat,
Energy.dat, WellMixedModelBase.h:505-570, 768-832, 1339-1411) through the
C-ABI:

* FieldInfo against the reference's own FieldInfo::Visit (built from its
  sources; tests/golden/ref_neutral.npz "info") on the same fields, 1-D and
  2-D, to 1e-13; and against numpy over the real cells after steps;
* the files' header lines as the reference's OutputTimeline writes them,
  without and with dimensions;
* History.dat's rows: the volume, solute, extrema and the run's sums read
  back against FieldInfo and yz_step_balances step by step; VolErr closes
  to the linear solve's tolerance on a closed film that evaporates and is
  fed by the nozzle (2-D), and VolErr and ResinVolErr on the 1-D film that
  the centrifugal term flings sideways (c uniform), where the outflowed
  volume and solute carry the loss; the dimensioned columns are the dimensionless ones times the scales;
* CalcInfo.dat: each step's Newton passes and last correction as the step
  returned them, the next dt, and the change rate of the step from its two
  films;
* Energy.dat: the free energy and dissipation of each state against the
  context's own, the dissipated energy the running sum of the
  Crank-Nicolson dissipation, the step's balance (F^{n+1} - F^n)/dt +
  (D^n + D^{n+1})/2, small against D on a closed film (capillary,
  disjoining, vertical gravity).
"""
import os

import numpy as np
import pytest

from yuzen_amd import capi, cases
from yuzen_amd.step import HipStep

from .helpers import TERMS, smooth_case

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NEU = np.load(os.path.join(GOLD, "ref_neutral.npz"))


def _neutral_case(d):
    c = capi.case_defaults()
    if d == "2d":
        c.dim, c.nx, c.ny = 2, 18, 14
        c.xmin, c.xmax, c.ymin, c.ymax = -0.5, 0.5, -0.4, 0.4
    else:
        c.dim, c.nx, c.ny = 1, 28, 1
        c.xmin, c.xmax, c.ymin, c.ymax = -0.5, 0.5, -1.0, 1.0
    c.capillary = 1
    return c


@pytest.mark.parametrize("d", ["1d", "2d"])
def test_field_info_matches_reference(d):
    c = _neutral_case(d)
    h, s, cc = (NEU[f"{d}__{k}"] for k in ("h", "s", "c"))
    g = HipStep(c, h, s, cc, 1e-3)
    try:
        got = np.array(g.field_info())
    finally:
        g.close()
    want = NEU[f"{d}__info"]
    print(f"field info {d}: {got}", flush=True)
    assert np.all(np.abs(got - want) <= 1e-13 * np.abs(want)), got - want


def _read(path):
    with open(path) as f:
        head = f.readline().rstrip("\n")
        rows = [line.rstrip("\n").split("\t") for line in f]
    return head, rows


def _with_b(c):
    c.evaporation, c.Ev, c.evap_coef = 1, 1.0, 50.0
    c.evap_func, c.precursor = capi.EVAP_RAOULT, 0.15
    c.source, c.src_radius, c.src_rate, c.src_tau = 1, 0.12, 200.0, 2e-5
    c.src_arm_length, c.src_arm_x, c.src_arm_y = 0.2, 0.05, -0.02
    c.src_omega, c.src_angle0, c.src_ramp, c.src_plateau = 3000.0, 0.3, 1e-5, 10.0
    c.src_subregions = 4
    return c


def test_headers(tmp_path):
    for dim, dims in ((2, True), (1, True), (2, False)):
        c, h, s, cc = smooth_case(dim, TERMS)
        g = HipStep(c, h, s, cc, 1e-5)
        try:
            g.series_open(str(tmp_path), dimensioned=dims, L0=1e-3, H0=1e-5, T0=0.1, sigma0=0.03, energy=True)
        finally:
            g.close()
        uV = (" [m^3]" if dim == 2 else " [m^3/m]") if dims else " [-]"
        uH = " [m]" if dims else " [-]"
        uE = (" [J]" if dim == 2 else " [J/m]") if dims else " [-]"
        uP = (" [W]" if dim == 2 else " [W/m]") if dims else " [-]"
        t = "Time [s]" if dims else "Time "
        hist = [t, "Volume" + uV, "ResinVol" + uV, "AveC [-]", "EvaporatedVol" + uV, "OutflowedVol" + uV,
                "OutflowedResin" + uV, "VolErr [-]", "ResinVolErr [-]", "maxHS" + uH, "minHS" + uH, "maxH" + uH,
                "minH" + uH, "maxC [-]", "minC [-]", "SuppliedVol" + uV, "SuppliedResin" + uV]
        assert _read(tmp_path / "History.dat")[0] == "\t".join(hist)
        calc = [t, "NK_iter", "NK_resid", "BCG_iter", "BCG_resid", "deltaT", "ChangeRate"]
        assert _read(tmp_path / "CalcInfo.dat")[0] == "\t".join(calc)
        en = [t, "FreeEnergy" + uE, "Dissipation" + uP, "Dissipated" + uE, "Balance" + uP]
        assert _read(tmp_path / "Energy.dat")[0] == "\t".join(en)


def _run(c, h, s, cc, dt, steps, path, **sc):
    """steps of the context with a row after each; what the context said
    at each step alongside"""
    g = HipStep(c, h, s, cc, dt)
    seen = []
    try:
        g.series_open(str(path), **sc)
        info0 = g.field_info()
        for _ in range(steps):
            hold = g.h()
            r = g.advance()
            assert r[0] == 0
            g.series_write()
            seen.append(dict(r=r, info=g.field_info(), bal=g.balances(), clock=g.clock(), h=g.h(), hold=hold,
                             log=g.pass_log()))
    finally:
        g.close()
    return info0, seen


def test_history_closes_with_evaporation_and_nozzle(tmp_path):
    c, h, s, cc = smooth_case(2, TERMS, slip=0.05)
    c = _with_b(c)
    info0, seen = _run(c, h, s, cc, 1e-5, 6, tmp_path)
    _, rows = _read(tmp_path / "History.dat")
    assert len(rows) == 6
    evap = sup = 0.0
    for k, (row, st) in enumerate(zip(rows, seen)):
        v = np.array([float(x) for x in row])
        evap += st["bal"][2]
        sup += st["bal"][3]
        i = st["info"]
        assert abs(v[0] - st["clock"][0]) <= 5e-7 * st["clock"][0]
        for col, want in ((1, i[0]), (2, i[1]), (3, i[2]), (4, evap), (9, i[3]), (10, i[4]), (11, i[5]),
                          (12, i[6]), (13, i[7]), (14, i[8]), (15, sup)):
            assert abs(v[col] - want) <= 5e-7 * abs(want) + 1e-300, (k, col, v[col], want)
        assert v[5] == 0.0 and v[6] == 0.0
        # VolErr: what the film holds plus what evaporated less what the
        # nozzle gave, against the opening volume
        assert abs(v[7]) <= 1e-9, (k, v[7])
        print(f"step {k}: VolErr {v[7]:.2e} ResinVolErr {v[8]:.2e} evaporated {v[4]:.3e} supplied {v[15]:.3e}",
              flush=True)
    assert evap > 0 and sup > 0


def test_history_closes_with_sideways_outflow(tmp_path):
    # a uniform c: with c frozen, the solute moves with the film only where
    # c is uniform
    c, h, s, cc = smooth_case(1, ["capillary", "centrifugal"], slip=0.0)
    cc = np.full(h.size, 0.5)
    info0, seen = _run(c, h, s, cc, 1e-4, 6, tmp_path)
    _, rows = _read(tmp_path / "History.dat")
    vals = np.array([[float(x) for x in r] for r in rows])
    assert np.all(vals[:, 5] > 0) and np.all(np.diff(vals[:, 5]) > 0)
    # the film lost what the outflowed column holds
    assert np.all(np.abs(vals[:, 7]) <= 1e-9), vals[:, 7]
    # and its solute with it (the concentration frozen)
    assert np.all(np.abs(vals[:, 8]) <= 1e-9), vals[:, 8]
    assert abs(vals[-1, 1] - info0[0]) > 1e3 * 1e-9 * info0[0]


def test_dimensioned_columns_are_scaled(tmp_path):
    c, h, s, cc = smooth_case(2, TERMS, slip=0.05)
    c = _with_b(c)
    sc = dict(L0=1.3e-3, H0=2.5e-6, T0=0.0186)
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir(); b.mkdir()
    _run(c, h, s, cc, 1e-5, 3, a)
    _run(c, h, s, cc, 1e-5, 3, b, dimensioned=True, **sc)
    va = np.array([[float(x) for x in r] for r in _read(a / "History.dat")[1]])
    vb = np.array([[float(x) for x in r] for r in _read(b / "History.dat")[1]])
    vol = sc["L0"] ** 2 * sc["H0"]
    scale = np.array([sc["T0"], vol, vol, 1, vol, vol, vol, 1, 1, sc["H0"], sc["H0"], sc["H0"], sc["H0"], 1, 1,
                      vol, vol])
    assert np.allclose(vb, va * scale, rtol=2e-6, atol=0)


def test_calcinfo_rows(tmp_path):
    c, h, s, cc, dt, _ = cases.spinodal_2d(n=32)
    _, seen = _run(c, h, s, cc, dt, 8, tmp_path)
    _, rows = _read(tmp_path / "CalcInfo.dat")
    for k, (row, st) in enumerate(zip(rows, seen)):
        t, nk, res, bcg, bres, dtn, rate = row
        assert int(nk) == st["r"][1]
        assert abs(float(res) - st["r"][3]) <= 5e-7 * abs(st["r"][3])
        assert abs(float(dtn) - st["clock"][1]) <= 5e-7 * st["clock"][1]
        assert int(bcg) > 0 and 0 <= float(bres) <= 1e-3
        if k + 1 >= 5:
            want = np.abs(st["h"] - st["hold"]).max() / (np.abs(st["h"]).max() * st["clock"][2])
            assert abs(float(rate) - want) <= 5e-6 * want
        else:
            assert float(rate) == 1.0


def test_energy_rows(tmp_path):
    """the rows against the context's own energy between the steps: F and
    D of the state each step reached, the running sum of (D^n + D^{n+1})/2
    dt and the last step's balance; and the balance small against D on a
    closed film"""
    c, h, s = smooth_case(2, ["capillary", "disjoining", "vertical_gravity"], slip=0.0)[:3]
    c.ew, c.nk_eps = 0, 1e-13
    cc = np.full(h.size, 0.5)
    dt = 1e-6
    g = HipStep(c, h, s, cc, dt)
    F, D = [], []
    try:
        g.series_open(str(tmp_path), energy=True)
        f0, d0 = g.energy()
        for _ in range(5):
            assert g.advance()[0] == 0
            g.series_write()
            f, d = g.energy()
            F.append(f); D.append(d)
    finally:
        g.close()
    v = np.array([[float(x) for x in r] for r in _read(tmp_path / "Energy.dat")[1]])
    Fp, Dp = np.array([f0] + F[:-1]), np.array([d0] + D[:-1])
    assert np.allclose(v[:, 1], F, rtol=1e-14, atol=0) and np.allclose(v[:, 2], D, rtol=1e-14, atol=0)
    assert np.allclose(v[:, 3], np.cumsum(0.5 * (Dp + np.array(D)) * dt), rtol=1e-13, atol=0)
    bal = (np.array(F) - Fp) / dt + 0.5 * (Dp + np.array(D))
    assert np.allclose(v[:, 4], bal, rtol=1e-9, atol=1e-12 * np.abs(D).max())
    print(f"energy balance / D: {np.abs(bal) / np.array(D)}", flush=True)
    assert np.all(np.abs(bal) <= 1e-2 * np.array(D))
