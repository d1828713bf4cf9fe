"""The VTK structured-grid file (.vts) of the device's film
(yz_step_write_vts, OutputVTK::Write, Output.cpp:375-637) and the velocity
fields it holds (yz_step_velocity_terms) through the C-ABI:

* the six fields (the mixed one, the capillary, Marangoni, vertical-gravity,
  in-plane gravity and centrifugal terms' own, each with its boundary rows)
  against the oracle's (whose terms match the reference's CalVelocity to the
  bit and whose slots are pinned by tests/test_oracle_velocity.py): every
  face within 1e-12 of the field's largest velocity, 1-D and 2-D, Symmetry
  and Periodic, with and without slip; the mixed one is yz_step_velocity's
  to the bit;
* the file the step context writes is yz_write_vts's file (byte for byte
  pinned against the reference's writer in tests/test_fieldio.py) of the
  context's own fields, clock and velocity fields, after steps of the
  all-physics film, without and with dimensions.
"""
import numpy as np
import pytest

from oracle import binding as ob
from yuzen_amd import capi, cases
from yuzen_amd.step import HipStep

from .helpers import TERMS, smooth_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("slip", [0.0, 0.05])
def test_velocity_terms_match_oracle(dim, periodic, slip):
    print(f"velocity terms dim={dim} periodic={periodic} slip={slip}", flush=True)
    c, h, s, cc = smooth_case(dim, TERMS, slip=slip, periodic=periodic)
    o = ob.OracleStep(c, h, s, cc, 1e-3)
    g = HipStep(c, h, s, cc, 1e-3)
    try:
        ux, uy = g.velocity_terms()
        mx, my = g.velocity()
    finally:
        g.close()
    vx, vy = o.velocity_terms(h, cc)
    assert np.array_equal(ux[0], mx)
    if dim == 2:
        assert np.array_equal(uy[0], my)
    for f in range(6):
        comps = ((ux[f], vx[f]), (uy[f], vy[f])) if dim == 2 else ((ux[f], vx[f]),)
        for got, want in comps:
            scale = np.abs(want).max()
            assert scale > 0, f
            err = np.abs(got - want).max() / scale
            print(f"  field {f}: max |dv| / max |v| = {err:.2e}")
            assert err <= 1e-12


@pytest.mark.parametrize("dimensioned", [False, True])
def test_step_file_is_the_writer_of_its_fields(dimensioned, tmp_path):
    c, h, s, cc, dt, _ = cases.allphys_2d(n=48, L=2.0, dt=2e-5)
    g = HipStep(c, h, s, cc, dt)
    try:
        for _ in range(3):
            assert g.advance()[0] == 0
        scales = dict(L0=1.3e-3, H0=2.5e-6, U0=0.07, T0=0.0186, dimensioned=dimensioned)
        a = str(tmp_path / "step.vts")
        g.write_vts(a, **scales)
        ux, uy = g.velocity_terms()
        t = g.clock()[0]
        hg, cg = g.h(), g.c()
    finally:
        g.close()
    b = str(tmp_path / "host.vts")
    capi.write_vts(b, c, t, hg, s, cg, ux, uy, **scales)
    ra, rb = open(a, "rb").read(), open(b, "rb").read()
    assert ra == rb
    want = f'Time="{t * (0.0186 if dimensioned else 1.0):e}"'
    assert want.encode() in ra[:400]
    # the oracle's fields at the film the step reached: the same file to the
    # velocity solve's rounding
    o = ob.OracleStep(c, h, s, cc, dt)
    vx, vy = o.velocity_terms(hg, cg)
    for f in range(6):
        for got, want in ((ux[f], vx[f]), (uy[f], vy[f])):
            assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
