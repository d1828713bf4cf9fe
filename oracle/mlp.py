"""FP64 restatement of the learned preconditioner: the per-cell network,
the Rademacher probes, the probe loss and its gradient by back-propagation.

Test infrastructure: imported only by tests/. There is no counterpart in
the reference (the reference preconditions with AMGCL's SPAI0 smoother,
yuzen/src/numerics/LinearSolverBackend.cpp:662-679); the functions here are
pinned by the mathematics instead: the gradient against central finite
differences of the loss, the probe loss against the dense Frobenius norm
||I - A M||_F^2 / N (its expectation over Rademacher probes, which is exact
for the mean over all 2^N sign vectors), and the stencil products against
dense matrices built entry by entry.

``emulate=True`` rounds every matrix-core operand to BF16 (round to
nearest even) as the device kernels do, accumulating in FP64; the result
then differs from the device's only by FP32 accumulation order.
"""
import numpy as np

# packed parameter layout (yuzen_amd/csrc/yz_mlp.h)
N_IN, N_H, N_OUT = 8, 64, 13
W1, B1 = 0, 64 * 16
W2, B2 = B1 + 64, B1 + 64 + 64 * 64
W3 = B2 + 64
B3 = W3 + 16 * 64
NP = B3 + 16

# 13-point slots, in the order oM mm om pm Mo mo oo po Po mp op pp oP
SLOTS_2D = [(0, -2), (-1, -1), (0, -1), (1, -1), (-2, 0), (-1, 0), (0, 0),
            (1, 0), (2, 0), (-1, 1), (0, 1), (1, 1), (0, 2)]
CENTRE = 6


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[W1:B1].reshape(64, 16), p[B1:W2], p[W2:B2].reshape(64, 64),
            p[B2:W3], p[W3:B3].reshape(16, 64), p[B3:NP])


def pack(w1, b1, w2, b2, w3, b3):
    p = np.zeros(NP)
    p[W1:B1] = w1.ravel()
    p[B1:W2] = b1
    p[W2:B2] = w2.ravel()
    p[B2:W3] = b2
    p[W3:B3] = w3.ravel()
    p[B3:NP] = b3
    return p


def _fold(j, n, periodic):
    """Ghost index -> cell: mirrored at a Symmetry edge (-1 -> 0, -2 -> 1),
    wrapped at a Periodic one."""
    j = np.asarray(j)
    if periodic:
        return np.mod(j, n)
    j = np.where(j < 0, -1 - j, j)
    return np.where(j >= n, 2 * n - 1 - j, j)


def columns(nx, ny, perx, pery):
    """(13, N) column of every slot of every row, cell i = iy * nx + ix."""
    iy, ix = np.divmod(np.arange(nx * ny), nx)
    cols = np.empty((13, nx * ny), dtype=np.int64)
    for k, (ox, oy) in enumerate(SLOTS_2D):
        cols[k] = _fold(iy + oy, ny, pery) * nx + _fold(ix + ox, nx, perx)
    return cols


def apply(planes, cols, x):
    """(P x)_i = sum_k P[k, i] x[cols[k, i]]"""
    return (planes * x[cols]).sum(axis=0)


def apply_t(planes, cols, e):
    """(P^T e)_j = sum over (k, i) with cols[k, i] = j of P[k, i] e_i"""
    w = np.zeros(e.shape[0])
    np.add.at(w, cols.ravel(), (planes * e[None, :]).ravel())
    return w


def dense(planes, cols):
    n = planes.shape[1]
    D = np.zeros((n, n))
    for k in range(planes.shape[0]):
        np.add.at(D, (np.arange(n), cols[k]), planes[k])
    return D


def probe(seed, p, n):
    """Rademacher probe p of a seed: bit 16 of a 32-bit mix of
    (seed, p, cell); 1 -> +1."""
    m = np.uint32(0xFFFFFFFF)
    with np.errstate(over="ignore"):
        i = np.arange(n, dtype=np.uint64)
        h = ((np.uint64(seed) * np.uint64(0x9E3779B1)) & np.uint64(m))
        h = h ^ ((np.uint64(p) * np.uint64(0x85EBCA77)) & np.uint64(m))
        h = h ^ ((i * np.uint64(0xC2B2AE3D)) & np.uint64(m))
        h = h ^ (h >> np.uint64(15))
        h = (h * np.uint64(0x2C1B3C6D)) & np.uint64(m)
        h = h ^ (h >> np.uint64(12))
        h = (h * np.uint64(0x297A2D39)) & np.uint64(m)
        h = h ^ (h >> np.uint64(15))
    return np.where((h >> np.uint64(16)) & np.uint64(1), 1.0, -1.0)


def forward(params, feat, mscale, emulate=False):
    """M (13, N) and the activations. feat (8, N), mscale (N)."""
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    w1, b1, w2, b2, w3, b3 = unpack(params)
    n = feat.shape[1]
    X = np.zeros((n, 16))
    X[:, :N_IN] = np.asarray(feat, dtype=np.float64).T
    H1 = np.maximum(r(X) @ r(w1).T + b1, 0.0)
    H2 = np.maximum(r(H1) @ r(w2).T + b2, 0.0)
    Y = r(H2) @ r(w3).T + b3
    M = (Y[:, :N_OUT] * np.asarray(mscale, dtype=np.float64)[:, None]).T
    return M, (X, H1, H2, Y)


def loss_grad(params, A, feat, mscale, cols, nprobe, seed, emulate=False, want_grad=True):
    """Probe loss L = (1/(P N)) sum_p ||v_p - A M v_p||^2 and dL/dparams
    (padding entries zero)."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[1]
    M, (X, H1, H2, _) = forward(params, feat, mscale, emulate)
    G = np.zeros((13, n))
    loss = 0.0
    for p in range(nprobe):
        v = probe(seed, p, n)
        e = v - apply(A, cols, apply(M, cols, v))
        loss += float(e @ e)
        if want_grad:
            w = apply_t(A, cols, e)
            G += (-2.0 / (nprobe * n)) * w[None, :] * v[cols]
    loss /= nprobe * n
    if not want_grad:
        return loss, None
    r = bf16 if emulate else (lambda a: a)
    w1, b1, w2, b2, w3, b3 = unpack(params)
    dY = np.zeros((n, 16))
    dY[:, :N_OUT] = (G * np.asarray(mscale, dtype=np.float64)[None, :]).T
    gw3 = r(dY).T @ r(H2)
    gb3 = dY.sum(axis=0)
    dZ2 = (r(dY) @ r(w3)) * (H2 > 0)
    gw2 = r(dZ2).T @ r(H1)
    gb2 = dZ2.sum(axis=0)
    dZ1 = (r(dZ2) @ r(w2)) * (H1 > 0)
    gw1 = r(dZ1).T @ r(X)
    gb1 = dZ1.sum(axis=0)
    gw1[:, N_IN:] = 0.0
    gw3[N_OUT:, :] = 0.0
    gb3[N_OUT:] = 0.0
    return loss, pack(gw1, gb1, gw2, gb2, gw3, gb3)


def init_params(seed=0, scale_out=0.01):
    """Small random weights for tests (the device has its own initial set)."""
    rng = np.random.default_rng(seed)
    w1 = np.zeros((64, 16))
    w1[:, :N_IN] = rng.normal(0, np.sqrt(2.0 / N_IN), (64, N_IN))
    w2 = rng.normal(0, np.sqrt(2.0 / 64), (64, 64))
    w3 = np.zeros((16, 64))
    w3[:N_OUT] = rng.normal(0, scale_out, (N_OUT, 64))
    b1 = rng.normal(0, 0.1, 64)
    b2 = rng.normal(0, 0.1, 64)
    b3 = np.zeros(16)
    b3[:N_OUT] = rng.normal(0, 0.05, N_OUT)
    b3[CENTRE] = 1.0
    return pack(w1, b1, w2, b2, w3, b3)
