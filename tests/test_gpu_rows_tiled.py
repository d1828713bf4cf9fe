"""The LDS-tiled row kernel of 2-D films (k_rows_tiled: tiles of 64 x 4
cells, every face's quantities formed once for both of its cells, g(h) of
the slip mobility and the disjoining powers once per point) against the
per-cell kernel (k_rows, every cell forming its own four faces): the same
rows to the bit. A face's quantities are symmetric in its two cells (the
mobility, its derivatives, the face viscosity and tension come out the same
whichever cell forms them), so sharing them changes no value.

Rows (IA, A(h)h, the right-hand side) of every term alone and all together,
with and without slip, Symmetry and Periodic edges, grids that fill no whole
tile, films with dry cells, and the 1024^2 all-physics film; then whole
steps (plain FP64 and learned FP32) advanced with either kernel.
"""
import numpy as np
import pytest

from yuzen_amd import capi, cases
from yuzen_amd.step import HipStep

from .helpers import TERMS, smooth_case

pytestmark = pytest.mark.gpu


def _rows(c, h, s, cc, tiled):
    was = capi.set_row_kernel(tiled)
    try:
        rng = np.random.default_rng(11)
        h_old = h * (1.0 + 0.01 * rng.standard_normal(h.size))
        g = HipStep(c, h, s, cc, 1e-3)
        out = g.assemble(h_old, h, 2e-3)
        g.close()
    finally:
        capi.set_row_kernel(was)
    return out


def _same_rows(c, h, s, cc, label):
    print(f"rows {label}: {c.nx} x {c.ny}", flush=True)
    a = _rows(c, h, s, cc, 0)
    b = _rows(c, h, s, cc, 1)
    for name, x, y in zip(("IA", "A(h)h", "Rhs"), a, b):
        assert np.array_equal(x, y), f"{label}: {name} differs by {np.nanmax(np.abs(x - y)):.3e}"
        assert np.all(np.isfinite(x)), f"{label}: {name} not finite"


@pytest.mark.parametrize("slip", [0.0, 0.05])
@pytest.mark.parametrize("term", TERMS + ["all"])
def test_rows_each_term(term, slip):
    terms = TERMS if term == "all" else [term]
    c, h, s, cc = smooth_case(2, terms, slip=slip, periodic=False, nx=70, ny=37)
    _same_rows(c, h, s, cc, f"{term} slip {slip}")


@pytest.mark.parametrize("periodic", [False, True])
def test_rows_all_terms_edges_and_ragged_tiles(periodic):
    for nx, ny in ((18, 14), (64, 4), (65, 5), (129, 67)):
        c, h, s, cc = smooth_case(2, TERMS, slip=0.05, periodic=periodic, nx=nx, ny=ny)
        _same_rows(c, h, s, cc, f"periodic {periodic}")


def test_rows_dry_cells():
    # a film that dips below zero in a band: the dry faces (mobility 0, no
    # disjoining) and the wet ones next to them
    for slip in (0.0, 0.05):
        c, h, s, cc = smooth_case(2, TERMS, slip=slip, periodic=False, nx=90, ny=41)
        h = h - 0.35
        _same_rows(c, h, s, cc, f"dry slip {slip}")


def test_rows_allphys_1024():
    c, h, s, cc, dt, _ = cases.allphys_2d(n=1024, dt=2e-6)
    _same_rows(c, h, s, cc, "all-physics 1024")


@pytest.mark.parametrize("learned", [False, True])
def test_steps_same_with_either_kernel(learned):
    c, h, s, cc, dt, _ = cases.allphys_2d(n=192, dt=2e-6)
    if learned:
        c.precond, c.krylov_fp32 = capi.PC_LEARNED, 1
        c.learn_passes, c.learn_iters, c.learn_probes, c.learn_lr = 2, 60, 4, 2e-3
    films, counts = [], []
    for tiled in (0, 1):
        was = capi.set_row_kernel(tiled)
        try:
            g = HipStep(c, h, s, cc, dt)
            counts.append([g.advance()[:3] for _ in range(3)])
            films.append(g.h())
            g.close()
        finally:
            capi.set_row_kernel(was)
    print(f"steps learned={learned}: {counts[1]}", flush=True)
    assert counts[0] == counts[1]
    assert np.array_equal(films[0], films[1])
