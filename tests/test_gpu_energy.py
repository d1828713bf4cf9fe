"""The energy budget on the device (yz_step_energy) through the C-ABI,
against the oracle (whose F and phi match the reference's EnergyBudget.h
to the bit and whose D satisfies the Crank-Nicolson energy identity:
tests/test_oracle_energy.py): F and D within 1e-12 of the oracle's at the
same film, 1-D and 2-D, Symmetry and Periodic, every term on (the power of
the fluxes that are not gradients included); and along device steps of a
closed film the identity holds to O(dt^2)."""
import numpy as np
import pytest

from oracle import binding as ob
from yuzen_amd.step import HipStep

from .helpers import TERMS, smooth_case
from .test_oracle_energy import _cosine_film

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("slip", [0.0, 0.05])
def test_energy_matches_oracle(dim, periodic, slip):
    c, h, s, cc = smooth_case(dim, TERMS, slip=slip, periodic=periodic)
    print(f"energy dim={dim} periodic={periodic} slip={slip}", flush=True)
    g = HipStep(c, h, s, cc, 1e-3)
    F, D = g.energy()
    hg = g.h()
    g.close()
    Fo, Do, _ = ob.OracleStep(c, h, s, cc, 1e-3).energy(hg, cc)
    print(f"  F {F:.15e} / {Fo:.15e}, D {D:.15e} / {Do:.15e}")
    assert abs(F - Fo) <= 1e-12 * abs(Fo)
    assert abs(D - Do) <= 1e-12 * abs(Do)


@pytest.mark.parametrize("dim", [1, 2])
def test_energy_identity_along_device_steps(dim):
    res = []
    for dt in (2e-5, 1e-5):
        c, h, s = _cosine_film(dim)
        c.ew, c.nk_eps = 0, 1e-13
        cc = np.full(h.size, 0.5)
        g = HipStep(c, h, s, cc, dt)
        F0, D0 = g.energy()
        worst = 0.0
        for _ in range(3):
            assert g.advance()[0] == 0
            F1, D1 = g.energy()
            assert D1 > 0 and F1 < F0
            worst = max(worst, abs((F1 - F0) / dt + 0.5 * (D0 + D1)) / D1)
            F0, D0 = F1, D1
        g.close()
        res.append(worst)
    print(dim, res)
    assert res[1] < 1e-2 and res[0] / res[1] > 3.0


@pytest.mark.parametrize("coef", [50.0, 1e5])
def test_volume_balance_closes(coef):
    """On a periodic film with evaporation (capped where coef = 1e5) and a
    nozzle, every step changes the volume by what the nozzle supplied less
    what evaporated, as yz_step_balances reports them, to the linear
    solve's tolerance; the volume and solute it reports are the film's."""
    from .test_gpu_evap_source import DT, _with_b
    c, h, s, cc = smooth_case(2, [t for t in TERMS if t != "centrifugal"], slip=0.05, periodic=True)
    _with_b(c, coef=coef)
    g = HipStep(c, h, s, cc, DT)
    v0 = g.balances()
    assert v0[2] == 0 and v0[3] == 0
    cm = (c.xmax - c.xmin) / c.nx * (c.ymax - c.ymin) / c.ny
    nxp, nyp = c.nx + 4, c.ny + 4
    for k in range(4):
        assert g.advance()[0] == 0
        v1 = g.balances()
        hg = g.h().reshape(nyp, nxp)[2:-2, 2:-2]
        cg = cc.reshape(nyp, nxp)[2:-2, 2:-2]
        gap = (v1[0] - v0[0]) - (v1[3] - v1[2])
        print(f"step {k}: dV {v1[0] - v0[0]:.6e} supplied {v1[3]:.6e} evaporated {v1[2]:.6e} gap {gap:.2e}")
        assert v1[2] > 0 and v1[3] > 0
        assert abs(gap) <= 1e-10 * v1[0]
        assert abs(v1[0] - hg.sum() * cm) <= 1e-12 * v1[0]
        assert abs(v1[1] - (hg * cg).sum() * cm) <= 1e-12 * v1[1]
        v0 = v1
    g.close()


def test_balances_after_load_state(tmp_path):
    """A context loaded from a state file has taken no step of its own: it
    reports no evaporated or supplied volume until it steps, and then the
    balance closes over that step as it does on the context that wrote the
    file."""
    from .test_gpu_evap_source import DT, _with_b
    c, h, s, cc = smooth_case(2, [t for t in TERMS if t != "centrifugal"], slip=0.05, periodic=True)
    _with_b(c, coef=50.0)
    g = HipStep(c, h, s, cc, DT)
    for _ in range(2):
        assert g.advance()[0] == 0
    assert min(g.balances()[2:]) > 0
    path = str(tmp_path / "b.state")
    g.save_state(path)
    g.close()
    r = HipStep(c, h, s, cc, DT)
    r.load_state(path)
    v0 = r.balances()
    assert v0[2] == 0 and v0[3] == 0
    assert r.advance()[0] == 0
    v1 = r.balances()
    gap = (v1[0] - v0[0]) - (v1[3] - v1[2])
    assert v1[2] > 0 and v1[3] > 0 and abs(gap) <= 1e-10 * v1[0]
    r.close()
