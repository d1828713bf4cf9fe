"""The restart state (field*.state) of the device step through the C-ABI.

* The reference's own writer (C_RestartState::Write, built from the
  reference's sources by oracle/build_ref.sh; the fixture is
  tests/golden/ref_state_wm1d.state) and the device step agree on the
  layout: a context loads that file, holds its film, and writes the same
  file back byte for byte.
* Continuation is exact: a run of 2k steps and a run of k steps saved,
  loaded into a fresh context and continued for k steps end with the same
  film (and concentration) to the bit: the plain FP64 2-D all-physics
  film, the same with the well-mixed concentration, a film under step
  control (the step and its count come back), the learned FP32 path
  (the network and its training state come back), and the ConcZ film with
  its coupling passes (the previous concentration the corrector
  extrapolates from comes back).
"""
import os

import numpy as np
import pytest

from yuzen_amd import capi, cases
from yuzen_amd.step import HipCZ, HipStep
from tests.cz_cases import _allphys

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "ref_state_wm1d.state")


def test_reference_state_file_loads_and_is_written_back_byte_for_byte(tmp_path):
    c = capi.case_defaults()
    c.dim, c.nx = 1, 8
    c.xmin, c.xmax = 0.0, 1.0
    h0 = np.ones(12)
    g = HipStep(c, h0, np.zeros(12), np.full(12, 0.5), 1e-3)
    g.load_state(FIXTURE)
    n = np.arange(12)
    want_h = 1.0 + 0.1 * np.sin(0.7 * n + 0.3)
    assert np.array_equal(g.h(), want_h)
    out = str(tmp_path / "again.state")
    g.save_state(out)
    assert open(out, "rb").read() == open(FIXTURE, "rb").read()
    g.close()


def _continuation(c, h, s, cc, dt, k, label, tmp_path, get_c=False):
    print(f"{label}: {2 * k} steps in one run", flush=True)
    a = HipStep(c, h, s, cc, dt)
    for _ in range(2 * k):
        assert a.advance()[0] == 0
    ha = a.h()
    ca = a.c() if get_c else None
    a.close()
    print(f"{label}: {k} steps, saved, loaded, {k} more", flush=True)
    b = HipStep(c, h, s, cc, dt)
    for _ in range(k):
        assert b.advance()[0] == 0
    path = str(tmp_path / "run.state")
    b.save_state(path)
    b.close()
    d = HipStep(c, np.ones_like(h), s, cc, dt)
    d.load_state(path)
    for _ in range(k):
        assert d.advance()[0] == 0
    assert np.array_equal(d.h(), ha), np.abs(d.h() - ha).max()
    if get_c:
        assert np.array_equal(d.c(), ca)
    d.close()


def test_continuation_plain_fp64(tmp_path):
    c, h, s, cc, dt, _ = cases.allphys_2d(n=40, L=1.5, dt=2e-5)
    _continuation(c, h, s, cc, dt, 2, "all-physics 40^2 FP64", tmp_path)


def test_continuation_with_concentration(tmp_path):
    c, h, s, cc, dt, _ = cases.allphys_2d(n=40, L=1.5, dt=2e-5)
    c.conc_model, c.Df = 1, 0.05
    _continuation(c, h, s, cc, dt, 2, "all-physics 40^2 with the concentration", tmp_path, get_c=True)


def test_continuation_under_step_control(tmp_path):
    c, h, s, cc, dt, _ = cases.spinodal_2d(n=32)
    assert c.step_control
    _continuation(c, h, s, cc, dt, 3, "spinodal 32^2 adaptive", tmp_path)


def test_continuation_learned(tmp_path):
    c, h, s, cc, dt, _ = cases.allphys_2d(n=48, L=2.0, dt=2e-5)
    # training spans the save: it runs on the first 6 Newton passes
    c.precond, c.krylov_fp32, c.learn_passes, c.learn_iters = capi.PC_LEARNED, 1, 6, 30
    _continuation(c, h, s, cc, dt, 1, "learned FP32 48^2", tmp_path)


@pytest.mark.parametrize("control", [False, True], ids=["fixed", "step_control"])
def test_continuation_concz(tmp_path, control):
    """4 steps in one run against 2, saved, loaded into a fresh context and
    2 more, to the bit; under the step control the clock (the next step, the
    last accepted one) and the lengthening count travel in the file"""
    c, h, s, cc = _allphys(per=False)
    if control:
        c.step_control, c.tsc_keep, c.tsc_criterion = 1, 2, 2e-4
    dt = 0.02
    a = HipCZ(c, h, s, cc, dt)
    for _ in range(4):
        assert a.advance()[0] == 0
    ha, ca, clk = a.h(), a.c(), a.clock()
    a.close()
    b = HipCZ(c, h, s, cc, dt)
    for _ in range(2):
        assert b.advance()[0] == 0
    path = str(tmp_path / "cz.state")
    b.save_state(path)
    b.close()
    d = HipCZ(c, np.ones_like(h), s, np.zeros_like(cc), dt)
    d.load_state(path)
    for _ in range(2):
        assert d.advance()[0] == 0
    assert np.array_equal(d.h(), ha) and np.array_equal(d.c(), ca)
    assert d.clock()[0] == clk[0] and d.clock()[1] == clk[1]
    d.close()
