"""The learned preconditioner on the device, through the C-ABI.

* The BF16 MFMA network and the probe-loss gradient against the FP64 oracle
  (oracle/mlp.py) on the solver's own systems: with the oracle rounding the
  matrix-core operands to BF16 as the device does, the results may differ
  only by FP32 accumulation order and, rarely, one operand rounded to the
  other BF16 neighbour (M: 99% of entries within 1e-6, all within 2e-4 of
  max |M|; loss 1e-5; 1e-4 relative L2 per parameter block); against the unrounded FP64 network the
  BF16 rounding itself shows (2e-2 for M, 3e-2 per gradient block).
* Training lowers the probe loss; the gradient is bit-reproducible.
* The FP32 inexact-Newton step with the learned preconditioner reaches the
  FP64 oracle's film: max |dh| / max h <= 1e-9 at Newton tolerance 1e-8
  (the Newton residual and update stay FP64; only the correction's solve
  is FP32).
* A breakdown or a stall of the FP32 solve falls back to the FP64 AMG
  solve and the step stays on the oracle's film.
* The gain: fewer Krylov iterations than the preconditioners the learned
  one replaces (SPAI0 as the multigrid's smoother; damped Jacobi as the
  one-pass preconditioner), and than the untrained network.
"""
import numpy as np
import pytest

from yuzen_amd import capi, cases
from yuzen_amd.step import HipStep
from oracle import binding as ob
from oracle import mlp as om

pytestmark = pytest.mark.gpu

BLOCKS = [("W1", om.W1, om.B1), ("b1", om.B1, om.W2), ("W2", om.W2, om.B2),
          ("b2", om.B2, om.W3), ("W3", om.W3, om.B3), ("b3", om.B3, om.NP)]


def _case(name):
    if name == "allphys":
        return cases.allphys_2d(n=40, L=2.0, dt=2e-5)
    return cases.spinodal_2d(n=32)


def _step(c, h, s, cc, dt, precond=capi.PC_LEARNED, fp32=1, **kw):
    c = c.copy()
    c.precond, c.krylov_fp32 = precond, fp32
    for k, v in kw.items():
        setattr(c, k, v)
    return c, HipStep(c, h, s, cc, dt)


@pytest.mark.parametrize("name", ["allphys", "spinodal"])
def test_network_forward_matches_oracle(name):
    c, h, s, cc, dt, _ = _case(name)
    c, g = _step(c, h, s, cc, dt, learn_passes=0)
    p = om.init_params(3, scale_out=0.05).astype(np.float32)
    g.learn_set_params(p)
    _, feat, ms = g.learn_sample()
    M = g.learn_apply()
    Me, _ = om.forward(p.astype(np.float64), feat, ms, emulate=True)
    Mx, _ = om.forward(p.astype(np.float64), feat, ms, emulate=False)
    d = np.abs(M - Me) / np.abs(Me).max()
    e1, q99 = d.max(), np.quantile(d, 0.99)
    e2 = np.abs(M - Mx).max() / np.abs(Mx).max()
    print(f"{name}: M vs BF16-emulated oracle max {e1:.2e} (99% {q99:.2e}), vs FP64 network {e2:.2e}")
    # FP32 accumulation order; where an activation sits at a BF16 rounding
    # tie the two round apart (one BF16 ulp, 2^-8, on one operand)
    assert q99 <= 1e-6
    assert e1 <= 2e-4
    assert e2 <= 2e-2
    g.close()


@pytest.mark.parametrize("name", ["allphys", "spinodal"])
def test_probe_loss_gradient_matches_oracle(name):
    c, h, s, cc, dt, _ = _case(name)
    c, g = _step(c, h, s, cc, dt, learn_passes=0)
    p = om.init_params(5, scale_out=0.05).astype(np.float32)
    g.learn_set_params(p)
    ia, feat, ms = g.learn_sample()
    cols = om.columns(c.nx, c.ny, c.bc[0] == capi.BC_PERIODIC, c.bc[2] == capi.BC_PERIODIC)
    loss, grad = g.learn_loss_grad(3, 17)
    le, ge = om.loss_grad(p.astype(np.float64), ia, feat, ms, cols, 3, 17, emulate=True)
    lx, gx = om.loss_grad(p.astype(np.float64), ia, feat, ms, cols, 3, 17, emulate=False)
    print(f"{name}: loss device {loss:.8e} oracle(BF16) {le:.8e} oracle(FP64) {lx:.8e}")
    assert abs(loss - le) <= 1e-5 * le
    for blk, a, b in BLOCKS:
        r1 = np.linalg.norm(grad[a:b] - ge[a:b]) / np.linalg.norm(ge[a:b])
        r2 = np.linalg.norm(grad[a:b] - gx[a:b]) / np.linalg.norm(gx[a:b])
        print(f"  {blk}: rel L2 vs BF16 oracle {r1:.2e}, vs FP64 {r2:.2e}")
        assert r1 <= 1e-4, blk
        assert r2 <= 3e-2, blk
    # padding entries stay exactly zero
    w1 = grad[om.W1:om.B1].reshape(64, 16)
    assert np.all(w1[:, om.N_IN:] == 0)
    assert np.all(grad[om.W3:om.B3].reshape(16, 64)[om.N_OUT:] == 0)
    # reproducible to the bit
    loss2, grad2 = g.learn_loss_grad(3, 17)
    assert loss2 == loss and np.array_equal(grad, grad2)
    g.close()


def test_training_lowers_probe_loss():
    c, h, s, cc, dt, _ = cases.allphys_2d(n=64, L=2.5, dt=2e-5)
    c, g = _step(c, h, s, cc, dt, learn_passes=0)
    g.learn_sample()
    l0, _ = g.learn_loss_grad(8, 99)
    first, last = g.learn_train(120, 2e-3, 4)
    l1, _ = g.learn_loss_grad(8, 99)
    print(f"probe loss {l0:.4e} -> {l1:.4e} (training {first:.4e} -> {last:.4e})")
    assert l1 < 0.5 * l0
    st = g.learn_stats()
    assert st[1] == 120 and st[2] > 0
    g.close()


def _run_vs_oracle(c, h, s, cc, dt, steps, label, tol=1e-9):
    o = ob.OracleStep(c, h, s, cc, dt)
    g = HipStep(c, h, s, cc, dt)
    for k in range(steps):
        ro = o.advance()
        rg = g.advance()
        ho, hg = o.h(), g.h()
        rel = np.abs(ho - hg).max() / np.abs(ho).max()
        print(f"{label} step {k}: oracle nk={ro[1]} lin={ro[2]}  hip nk={rg[1]} lin={rg[2]}  rel={rel:.2e}",
              flush=True)
        assert ro[0] == 0 and rg[0] == 0
        assert rel <= tol
    st = g.learn_stats()
    g.close()
    return st


@pytest.mark.parametrize("precond", [capi.PC_LEARNED, capi.PC_LEARNED_AMG, capi.PC_AMG])
def test_fp32_step_matches_fp64_oracle(precond):
    c, h, s, cc, dt, _ = cases.allphys_2d(n=48, L=2.0, dt=2e-5)
    c.precond, c.krylov_fp32 = precond, 1
    st = _run_vs_oracle(c, h, s, cc, dt, 4, f"fp32 precond={precond}")
    print("fallbacks", st[0])


def test_breakdown_falls_back_to_fp64_amg():
    # an all-zero network gives M = 0: the FP32 solve breaks down at once
    # (v = A M p = 0) and every pass is solved again in FP64
    c, h, s, cc, dt, _ = cases.allphys_2d(n=48, L=2.0, dt=2e-5)
    c.precond, c.krylov_fp32, c.learn_passes = capi.PC_LEARNED, 1, 0
    o = ob.OracleStep(c, h, s, cc, dt)
    g = HipStep(c, h, s, cc, dt)
    g.learn_set_params(np.zeros(capi.LEARN_NPARAM, dtype=np.float32))
    passes = 0
    for k in range(3):
        ro, rg = o.advance(), g.advance()
        passes += rg[1]
        rel = np.abs(o.h() - g.h()).max() / np.abs(o.h()).max()
        print(f"zero-M step {k}: nk {ro[1]}/{rg[1]} rel {rel:.2e}")
        assert rg[0] == 0 and rel <= 1e-9
    assert g.learn_stats()[0] == passes
    g.close()


def test_stall_test_falls_back():
    # a one-iteration stall window (the residual must halve every
    # iteration) with the untrained network
    c, h, s, cc, dt, _ = cases.allphys_2d(n=48, L=2.0, dt=2e-5)
    c.precond, c.krylov_fp32, c.learn_passes, c.stall_window = capi.PC_LEARNED, 1, 0, 1
    st = _run_vs_oracle(c, h, s, cc, dt, 2, "stall window 1")
    assert st[0] >= 1


def _krylov(c, h, s, cc, dt, steps, train_steps=1):
    g = HipStep(c, h, s, cc, dt)
    lin = []
    for k in range(steps):
        r = g.advance()
        assert r[0] == 0
        lin.append(r[2])
    st = g.learn_stats()
    g.close()
    return sum(lin[train_steps:]), st


def test_learned_preconditioner_gain():
    c, h, s, cc, dt, _ = cases.allphys_2d(n=256, dt=2e-6)
    c.krylov_fp32, c.stall_window = 1, 0
    runs = {}
    for name, pc, lp in [("spai0_amg", capi.PC_AMG, 0), ("learned_amg", capi.PC_LEARNED_AMG, 2),
                         ("jacobi", capi.PC_JACOBI, 0), ("learned", capi.PC_LEARNED, 2),
                         ("untrained", capi.PC_LEARNED, 0)]:
        cc_ = c.copy()
        cc_.precond, cc_.learn_passes = pc, lp
        runs[name], st = _krylov(cc_, h, s, cc, dt, 4)
        print(f"{name}: Krylov iterations over steps 1-3 {runs[name]}, fallbacks {st[0]:.0f}")
    assert runs["learned_amg"] < 0.8 * runs["spai0_amg"]
    assert runs["learned"] < 0.5 * runs["jacobi"]
    assert runs["learned"] < 0.8 * runs["untrained"]


def test_graph_replay_matches_kernel_by_kernel():
    # the fused iteration replayed from its captured two-iteration graph
    # gives the same films and counts, to the bit, as launching it kernel by
    # kernel, over steps with eager work (assembly, set-up, training)
    # between the replays
    c, h, s, cc, dt, _ = cases.allphys_2d(n=96, L=3.0, dt=1e-5)
    c.precond, c.krylov_fp32 = capi.PC_LEARNED, 1
    runs = []
    for graphs in (0, 1):
        cg = c.copy()
        cg.graphs = graphs
        g = HipStep(cg, h, s, cc, dt)
        counts = []
        for k in range(4):
            r = g.advance()
            assert r[0] == 0
            counts.append((r[1], r[2]))
        runs.append((counts, g.h()))
        g.close()
    print("kernel by kernel", runs[0][0], "graph", runs[1][0])
    assert runs[0][0] == runs[1][0]
    assert np.array_equal(runs[0][1], runs[1][1])
