"""The film height's face schemes (Numerics.Convection.Height,
FiniteDifferences.cpp:22-228) on the device against the oracle, whose terms
match the reference's own for each scheme (tests/test_oracle_golden.py: the
"_vl" cases for van Leer, the "_hs<k>" cases for the fixed-coefficient
schemes, which each upwinded term assembles by its own code,
BodyForceKernel.h:263-545, SolutoCapillarityKernel.h:182-405): the rows of
the upwinded terms (in-plane gravity, the centrifugal force, Marangoni) and
all terms together, 1-D and 2-D, with and without slip, to 1e-12 of each
row's scale; the face velocities; whole steps with equal Newton counts to
1e-10; the learned FP32 path to 1e-8; a check that each scheme's rows
differ from Koren's where the film is not smooth (the switch takes effect);
and a scheme number outside the list refused.
"""
import numpy as np
import pytest

from oracle import binding as ob
from yuzen_amd import capi, cases
from yuzen_amd.step import HipStep

from .helpers import TERMS, smooth_case, csr_to_planes, interior

pytestmark = pytest.mark.gpu


def _rows(c, h, s, cc):
    o = ob.OracleStep(c, h, s, cc, 1e-3)
    o.set_h(h)
    hf = o.h()
    rng = np.random.default_rng(5)
    h_old = hf * (1.0 + 0.01 * rng.standard_normal(hf.size))
    A, J, IA, ax, rhs = o.assemble(h_old, hf, 2e-3)
    rp, col, bc = o.pattern()
    ref = csr_to_planes(c, rp, col, IA)
    g = HipStep(c, h, s, cc, 1e-3)
    ia, gax, grhs = g.assemble(h_old, hf, 2e-3)
    g.close()
    return ref, ia, interior(c, rhs), interior(c, grhs)


SCHEMES = {"vanleer": capi.HSCHEME_VANLEER, "upwind1": capi.HSCHEME_UPWIND1, "central2": capi.HSCHEME_CENTRAL2,
           "central4": capi.HSCHEME_CENTRAL4, "upwind2": capi.HSCHEME_UPWIND2, "utopia": capi.HSCHEME_UTOPIA,
           "quick": capi.HSCHEME_QUICK, "kk": capi.HSCHEME_KK}


@pytest.mark.parametrize("scheme", list(SCHEMES))
@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("slip", [0.0, 0.05])
@pytest.mark.parametrize("term", ["inplane_gravity", "centrifugal", "marangoni", "all"])
def test_scheme_rows(scheme, dim, slip, term):
    terms = TERMS if term == "all" else [term]
    c, h, s, cc = smooth_case(dim, terms, slip=slip)
    c.height_scheme = SCHEMES[scheme]
    print(f"{scheme} rows dim={dim} slip={slip} term={term}", flush=True)
    ref, ia, ri, gri = _rows(c, h, s, cc)
    scale = np.abs(ref).max(axis=0)
    assert np.all(np.abs(ia - ref) <= 1e-12 * scale + 1e-300), np.abs(ia - ref).max()
    assert np.abs(ri - gri).max() <= 1e-12 * max(np.abs(ri).max(), 1e-300) + 1e-15


@pytest.mark.parametrize("scheme", list(SCHEMES))
def test_scheme_differs_from_koren(scheme):
    c, h, s, cc = smooth_case(2, ["inplane_gravity", "marangoni", "centrifugal"], slip=0.0, nx=40, ny=30)
    rng = np.random.default_rng(9)
    h = h * (1.0 + 0.05 * rng.standard_normal(h.size))
    k = c.copy()
    v = c.copy()
    v.height_scheme = SCHEMES[scheme]
    rk, iak, _, _ = _rows(k, h, s, cc)
    rv, iav, _, _ = _rows(v, h, s, cc)
    assert np.abs(iak - iav).max() > 1e-6 * np.abs(iak).max()


@pytest.mark.parametrize("scheme", list(SCHEMES))
@pytest.mark.parametrize("dim", [1, 2])
def test_scheme_velocity(scheme, dim):
    c, h, s, cc = smooth_case(dim, TERMS, slip=0.05)
    c.height_scheme = SCHEMES[scheme]
    o = ob.OracleStep(c, h, s, cc, 1e-3)
    g = HipStep(c, h, s, cc, 1e-3)
    ux, uy = g.velocity()
    g.close()
    vx, vy = o.velocity(h, cc)
    for got, want in (((ux, vx), (uy, vy)) if dim == 2 else ((ux, vx),)):
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("scheme", ["vanleer", "utopia", "quick"])
@pytest.mark.parametrize("learned", [False, True])
def test_scheme_steps(scheme, learned):
    c, h, s, cc, dt, _ = cases.allphys_2d(n=48, L=2.0, dt=2e-5)
    c.height_scheme = SCHEMES[scheme]
    if learned:
        c.precond, c.krylov_fp32 = capi.PC_LEARNED, 1
        c.learn_passes, c.learn_iters, c.learn_probes, c.learn_lr = 2, 40, 4, 2e-3
    o = ob.OracleStep(c, h, s, cc, dt, lin_mode=ob.OracleStep.LIN_STRUCTURED)
    g = HipStep(c, h, s, cc, dt)
    for k in range(4):
        ro, rg = o.advance(), g.advance()
        ho, hg = o.h(), g.h()
        rel = np.abs(ho - hg).max() / np.abs(ho).max()
        print(f"{scheme} step {k} learned={learned}: oracle nk={ro[1]} hip nk={rg[1]} rel={rel:.2e}", flush=True)
        assert ro[0] == 0 and rg[0] == 0
        if not learned:
            assert ro[1] == rg[1]
        assert rel <= (1e-8 if learned else 1e-10)
    g.close()


def test_unknown_height_scheme_refused():
    c, h, s, cc = smooth_case(2, TERMS)
    for bad in (-1, 9):
        c.height_scheme = bad
        with pytest.raises(Exception):
            HipStep(c, h, s, cc, 1e-3)
