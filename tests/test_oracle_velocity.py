"""The film's face velocities (CalVelocity) on the oracle, with the
equation's boundary rows.

Each term's velocities match the reference's own CalVelocity to the bit
(tests/test_oracle_golden.py). The boundary rows (SetVelocityBC) live in the
reference's equation class, which cannot be built here (its linear solver
needs Eigen/AMGCL), so they are pinned by what they must give:
* on a Periodic film, moving the film by one cell moves every velocity the
  rows give by one cell, the ghost faces included (the rows copy the faces
  of the other end; a wrong index breaks this);
* at Symmetry walls the normal velocity through the wall face and the ghost
  faces beyond is zero, and the ghost column carries the tangential
  velocity of its mirror cell;
* the mixed field is the plain sum of the terms' fields in the reference's
  order.
"""
import numpy as np
import pytest

from oracle import binding as ob

from .helpers import TERMS, smooth_case


def _shift_case(dim, slip, axis):
    """a periodic film with every term but the centrifugal one (its arm is
    measured from a fixed point), and the same film moved one cell along
    axis"""
    c, _, _, _ = smooth_case(dim, [t for t in TERMS if t != "centrifugal"], slip=slip, periodic=True)
    nx, ny = c.nx, (c.ny if dim == 2 else 1)
    rng = np.random.default_rng(11)
    shape = (ny, nx) if dim == 2 else (nx,)
    real = {k: rng.uniform(lo, hi, shape) for k, lo, hi in (("h", 0.4, 1.2), ("s", -0.1, 0.1), ("c", 0.3, 0.7))}

    def pad(a):
        return np.pad(a, 2, mode="wrap").ravel() if dim == 2 else np.pad(a, 2, mode="wrap")

    one = {k: pad(v) for k, v in real.items()}
    two = {k: pad(np.roll(v, 1, axis=(1 if (dim == 2 and axis == 0) else 0))) for k, v in real.items()}
    return c, one, two


@pytest.mark.parametrize("dim,axis", [(1, 0), (2, 0), (2, 1)])
@pytest.mark.parametrize("slip", [0.0, 0.05])
def test_periodic_rows_move_with_the_film(dim, axis, slip):
    c, one, two = _shift_case(dim, slip, axis)
    o = ob.OracleStep(c, one["h"], one["s"], one["c"], 1e-3)
    v1 = o.velocity(one["h"], one["c"])
    o2 = ob.OracleStep(c, two["h"], two["s"], two["c"], 1e-3)
    v2 = o2.velocity(two["h"], two["c"])
    nxp = c.nx + 4
    if dim == 1:
        a, b = v1[0], v2[0]
        assert np.abs(a).max() > 0
        # faces 1 .. nxp-1 carry the rows' values; face 0 is set to 0
        assert np.array_equal(b[2:nxp], a[1:nxp - 1])
        return
    nyp = c.ny + 4
    for comp in (0, 1):
        a = v1[comp].reshape(nyp, nxp)
        b = v2[comp].reshape(nyp, nxp)
        assert np.abs(a).max() > 0
        # the faces the terms and the rows write: x on rows 1 .. nyp-2 and
        # columns 1 .. nxp-1, y on columns 1 .. nxp-2 and rows 1 .. nyp-1
        rows = slice(1, nyp - 1) if comp == 0 else slice(1, nyp)
        cols = slice(1, nxp) if comp == 0 else slice(1, nxp - 1)
        A, B = a[rows, cols], b[rows, cols]
        if axis == 0:
            assert np.array_equal(B[:, 1:], A[:, :-1]), (comp, np.abs(B[:, 1:] - A[:, :-1]).max())
        else:
            assert np.array_equal(B[1:, :], A[:-1, :]), (comp, np.abs(B[1:, :] - A[:-1, :]).max())


@pytest.mark.parametrize("dim", [1, 2])
def test_symmetry_walls(dim):
    c, h, s, cc = smooth_case(dim, TERMS, slip=0.05)
    o = ob.OracleStep(c, h, s, cc, 1e-3)
    vx, vy = o.velocity(h, cc)
    raw = o.velocity(h, cc, bc=False)
    nxp = c.nx + 4
    if dim == 1:
        assert np.all(vx[[0, 1, 2, nxp - 2, nxp - 1]] == 0)
        assert np.array_equal(vx[3:nxp - 2], raw[0][3:nxp - 2])
        return
    nyp = c.ny + 4
    X, Y = vx.reshape(nyp, nxp), vy.reshape(nyp, nxp)
    RX, RY = raw[0].reshape(nyp, nxp), raw[1].reshape(nyp, nxp)
    # no flow through the walls (the wall faces and the ghost faces beyond)
    assert np.all(X[1:nyp - 1, [0, 1, 2, nxp - 2, nxp - 1]] == 0)
    assert np.all(Y[[0, 1, 2, nyp - 2, nyp - 1], 1:nxp - 1] == 0)
    # the ghost column / row carries its mirror cell's tangential velocity
    assert np.array_equal(Y[3:nyp - 2, 1], Y[3:nyp - 2, 2])
    assert np.array_equal(Y[3:nyp - 2, nxp - 2], Y[3:nyp - 2, nxp - 3])
    assert np.array_equal(X[1, 3:nxp - 2], X[2, 3:nxp - 2])
    assert np.array_equal(X[nyp - 2, 3:nxp - 2], X[nyp - 3, 3:nxp - 2])
    # inside, the rows leave what the terms wrote
    assert np.array_equal(X[3:nyp - 3, 3:nxp - 2], RX[3:nyp - 3, 3:nxp - 2])
    assert np.array_equal(Y[3:nyp - 2, 3:nxp - 3], RY[3:nyp - 2, 3:nxp - 3])
    assert np.abs(RX).max() > 0 and np.abs(RY).max() > 0


@pytest.mark.parametrize("periodic", [False, True])
def test_mixed_is_the_sum_of_the_terms(periodic):
    """each term alone (the Marangoni term kept on with As = 0 where it is
    not the one, so that sigma(c) of the capillary mobility is the same),
    each with its boundary rows, summed in the reference's order, is the
    mixed field to the bit"""
    c, h, s, cc = smooth_case(2, TERMS, slip=0.05, periodic=periodic)
    o = ob.OracleStep(c, h, s, cc, 1e-3)
    vx, vy = o.velocity(h, cc)
    order = ["capillary", "marangoni", "vertical_gravity", "inplane_gravity", "centrifugal", "disjoining"]
    sx, sy = np.zeros_like(vx), np.zeros_like(vy)
    for t in order:
        ct, _, _, _ = smooth_case(2, TERMS, slip=0.05, periodic=periodic)
        for u in order:
            setattr(ct, u, 1 if u in (t, "marangoni") else 0)
        if t != "marangoni":
            ct.As = 0.0
        tx, ty = ob.OracleStep(ct, h, s, cc, 1e-3).velocity(h, cc)
        assert np.abs(tx).max() > 0
        sx, sy = sx + tx, sy + ty
    assert np.array_equal(vx, sx) and np.array_equal(vy, sy)


@pytest.mark.parametrize("dim", [1, 2])
def test_terms_fields_in_their_slots(dim):
    """the .vts's velocity fields (velocity_terms): the mixed field as
    velocity() gives it, then each of the capillary, Marangoni,
    vertical-gravity, in-plane gravity and centrifugal terms' own field, the
    field of a case with that term alone (Symmetry walls, where one pass of
    the boundary rows is two)"""
    c, h, s, cc = smooth_case(dim, TERMS, slip=0.05, periodic=False)
    o = ob.OracleStep(c, h, s, cc, 1e-3)
    vx, vy = o.velocity(h, cc)
    tx, ty = o.velocity_terms(h, cc)
    assert np.array_equal(tx[0], vx) and np.array_equal(ty[0], vy)
    order = ["capillary", "marangoni", "vertical_gravity", "inplane_gravity", "centrifugal", "disjoining"]
    for k, t in enumerate(order[:5]):
        ct, _, _, _ = smooth_case(dim, TERMS, slip=0.05, periodic=False)
        for u in order:
            setattr(ct, u, 1 if u in (t, "marangoni") else 0)
        if t != "marangoni":
            ct.As = 0.0
        ax, ay = ob.OracleStep(ct, h, s, cc, 1e-3).velocity(h, cc)
        assert np.abs(ax).max() > 0, t
        assert np.array_equal(tx[k + 1], ax), t
        if dim == 2:
            assert np.array_equal(ty[k + 1], ay), t
