"""The FP64 oracle of the learned preconditioner, pinned by the mathematics:
central finite differences of the loss, the dense Frobenius norm that the
probe loss estimates (exact over all sign vectors), dense stencil products,
the Symmetry/Periodic column folds and BF16 round-to-nearest-even."""
import itertools

import numpy as np
import pytest

from oracle import mlp as om


def _system(nx, ny, perx, pery, seed):
    rng = np.random.default_rng(seed)
    n = nx * ny
    A = rng.normal(0, 0.15, (13, n))
    A[om.CENTRE] = 1.0 + rng.uniform(0, 0.3, n)
    feat = rng.normal(0, 0.5, (8, n))
    mscale = rng.uniform(0.3, 1.0, n)
    return A, feat, mscale, om.columns(nx, ny, perx, pery)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 3.0e-3])
    r = om.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie, even neighbour below
    assert r[2] == 1.0 + 2.0 ** -6          # tie, even neighbour above
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 2.0 ** -8 * 3.0e-3   # half an ulp


def test_columns_fold_like_ghost_rows():
    # Symmetry: ghost -1 is cell 0, -2 is cell 1 (and mirrored at the far
    # edge); Periodic wraps
    nx, ny = 5, 4
    cs = om.columns(nx, ny, False, False)
    cp = om.columns(nx, ny, True, True)
    k_Mo = om.SLOTS_2D.index((-2, 0))
    k_oP = om.SLOTS_2D.index((0, 2))
    # cell (0, 0): x-2 -> cell 1 (Symmetry), cell nx-2 (Periodic)
    assert cs[k_Mo, 0] == 1 and cp[k_Mo, 0] == nx - 2
    # cell (0, ny-1): y+2 -> row ny-2 (Symmetry), row 1 (Periodic)
    i = (ny - 1) * nx
    assert cs[k_oP, i] == (ny - 2) * nx and cp[k_oP, i] == 1 * nx
    assert cs[om.CENTRE].tolist() == list(range(nx * ny))


def test_stencil_products_equal_dense():
    A, _, _, cols = _system(5, 4, True, False, 1)
    x = np.random.default_rng(2).normal(size=20)
    D = om.dense(A, cols)
    np.testing.assert_allclose(om.apply(A, cols, x), D @ x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(om.apply_t(A, cols, x), D.T @ x, rtol=1e-13, atol=1e-13)


def test_probe_is_balanced_and_seeded():
    v = om.probe(7, 0, 4096)
    assert set(np.unique(v)) == {-1.0, 1.0}
    assert abs(v.mean()) < 0.06
    assert not np.array_equal(v, om.probe(7, 1, 4096))
    assert not np.array_equal(v, om.probe(8, 0, 4096))
    np.testing.assert_array_equal(v, om.probe(7, 0, 4096))


def _loss_all_signs(params, A, feat, mscale, cols):
    """Mean of ||v - A M v||^2 / N over all 2^N sign vectors, by brute force."""
    n = A.shape[1]
    M, _ = om.forward(params, feat, mscale)
    tot = 0.0
    for s in itertools.product((-1.0, 1.0), repeat=n):
        v = np.array(s)
        e = v - om.apply(A, cols, om.apply(M, cols, v))
        tot += e @ e
    return tot / (2 ** n) / n


@pytest.mark.parametrize("per", [False, True])
def test_probe_loss_expectation_is_frobenius(per):
    # 3 x 3 grid: every slot folds, 512 sign vectors
    A, feat, mscale, cols = _system(3, 3, per, per, 3)
    p = om.init_params(4, scale_out=0.05)
    M, _ = om.forward(p, feat, mscale)
    R = np.eye(9) - om.dense(A, cols) @ om.dense(M, cols)
    frob = (R * R).sum() / 9
    assert _loss_all_signs(p, A, feat, mscale, cols) == pytest.approx(frob, rel=1e-12)


def test_single_probe_loss_against_dense():
    A, feat, mscale, cols = _system(6, 5, False, True, 5)
    p = om.init_params(6, scale_out=0.05)
    M, _ = om.forward(p, feat, mscale)
    v = om.probe(99, 0, 30)
    e = v - om.dense(A, cols) @ (om.dense(M, cols) @ v)
    loss, _ = om.loss_grad(p, A, feat, mscale, cols, 1, 99, want_grad=False)
    assert loss == pytest.approx(e @ e / 30, rel=1e-12)


@pytest.mark.parametrize("per", [False, True])
def test_gradient_matches_finite_differences(per):
    A, feat, mscale, cols = _system(6, 5, per, per, 7)
    p = om.init_params(8, scale_out=0.2)
    _, g = om.loss_grad(p, A, feat, mscale, cols, 2, 5)
    rng = np.random.default_rng(9)
    live = [i for i in range(om.NP)
            if not (om.W1 <= i < om.B1 and (i - om.W1) % 16 >= om.N_IN)
            and not (om.W3 <= i < om.B3 and (i - om.W3) // 64 >= om.N_OUT)
            and not (i >= om.B3 and i - om.B3 >= om.N_OUT)]
    idx = sorted(set(rng.choice(live, 300, replace=False).tolist())
                 | set(range(om.B3, om.B3 + om.N_OUT)))
    h = 1e-6
    fd = np.empty(len(idx))
    for j, i in enumerate(idx):
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = om.loss_grad(pp, A, feat, mscale, cols, 2, 5, want_grad=False)
        lm, _ = om.loss_grad(pm, A, feat, mscale, cols, 2, 5, want_grad=False)
        fd[j] = (lp - lm) / (2 * h)
    err = np.linalg.norm(fd - g[idx]) / np.linalg.norm(fd)
    assert err < 1e-6, err


def test_frobenius_gradient_over_all_signs():
    # over all sign vectors the mean gradient is that of ||I - A M||_F^2 / N,
    # whose derivative in M is -2 A^T (I - A M) / N (entries at the stencil
    # columns)
    A, feat, mscale, cols = _system(3, 3, False, False, 11)
    p = om.init_params(12, scale_out=0.05)
    M, (X, H1, H2, _) = om.forward(p, feat, mscale)
    Ad, Md = om.dense(A, cols), om.dense(M, cols)
    Gd = -2.0 * Ad.T @ (np.eye(9) - Ad @ Md) / 9
    # dL/dM slot by slot: a folded column collects several slots, each of
    # which sees the whole entry's derivative
    G = np.array([[Gd[i, cols[k, i]] for i in range(9)] for k in range(13)])
    # the network's gradient from G by the chain rule, written densely
    w1, b1, w2, b2, w3, b3 = om.unpack(p)
    dY = np.zeros((9, 16))
    dY[:, :13] = (G * mscale).T
    gw3 = dY.T @ H2
    dZ2 = (dY @ w3) * (H2 > 0)
    gw2 = dZ2.T @ H1
    dZ1 = (dZ2 @ w2) * (H1 > 0)
    gw1 = dZ1.T @ X
    # the probe gradient averaged over all 512 sign vectors
    acc = np.zeros(om.NP)
    for s in itertools.product((-1.0, 1.0), repeat=9):
        v = np.array(s)
        e = v - om.apply(A, cols, om.apply(M, cols, v))
        w = om.apply_t(A, cols, e)
        Gp = -2.0 / 9 * w[None, :] * v[cols]
        dYp = np.zeros((9, 16))
        dYp[:, :13] = (Gp * mscale).T
        acc[om.W3:om.B3] += (dYp.T @ H2).ravel()
    acc /= 512
    np.testing.assert_allclose(acc[om.W3:om.B3].reshape(16, 64)[:13], gw3[:13], rtol=1e-10, atol=1e-12)
    assert np.abs(gw2).sum() > 0 and np.abs(gw1).sum() > 0
