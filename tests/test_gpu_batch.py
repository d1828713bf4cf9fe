"""The batched sweep through the C-ABI: films advanced together in one
batch reach, to the bit, the films each one reaches in a step context of
its own with the same network (the batch changes how the work is
launched, not the arithmetic), with the same Newton and Krylov counts,
including a film whose FP32 solves fall back to the FP64 AMG solve while
the others do not."""
import numpy as np
import pytest

from yuzen_amd import capi, cases
from yuzen_amd.step import HipStep, HipBatch

pytestmark = pytest.mark.gpu

SIDES = (40, 64, 100, 130)   # partial tiles and whole ones (64 x 16 tiles)


def _weights():
    c, h, s, cc, dt = cases.sweep_film(0, side=96)
    g = HipStep(c, h, s, cc, dt)
    g.learn_sample()
    g.learn_train(60, 2e-3, 4)
    p = g.learn_params()
    g.close()
    return p


def _films(stall_film=None):
    films = []
    for k, n in enumerate(SIDES):
        c, h, s, cc, dt = cases.sweep_film(k, side=n)
        if k == stall_film:
            c.stall_window = 1
        films.append((c, h, s, cc, dt))
    return films


def _alone(films, params, steps):
    out = []
    for c, h, s, cc, dt in films:
        g = HipStep(c, h, s, cc, dt)
        g.learn_set_params(params)
        counts = []
        for _ in range(steps):
            rc, nk, lin, _ = g.advance()
            assert rc == 0
            counts.append((nk, lin))
        out.append((counts, g.h(), g.learn_stats()[0]))
        g.close()
    return out


@pytest.mark.parametrize("stall_film", [None, 1])
def test_batch_matches_films_alone(stall_film):
    params = _weights()
    films = _films(stall_film)
    steps = 3
    ref = _alone(films, params, steps)
    b = HipBatch(films)
    b.set_params(params)
    counts = [[] for _ in films]
    for _ in range(steps):
        rc, st, nk, lin = b.advance()
        assert rc == 0 and np.all(st == 0)
        for k in range(len(films)):
            counts[k].append((int(nk[k]), int(lin[k])))
    for k, (rcounts, rh, rfb) in enumerate(ref):
        print(f"film {k} ({SIDES[k]}^2): alone {rcounts} fallbacks {rfb:.0f}; batch {counts[k]}")
        assert counts[k] == rcounts
        assert np.array_equal(b.h(k), rh)
    fb = b.stats()[3]
    assert fb == sum(r[2] for r in ref)
    if stall_film is not None:
        assert ref[stall_film][2] >= steps   # every pass of that film fell back
    b.close()


def test_batch_graph_and_mixed_tiles_match_films_alone():
    """The batch with its iteration replayed from a captured HIP graph
    (graphs 1: the grid fixed for the solve, films that have finished keep
    blocks that return at once) and with films of different tile rows in one
    batch (8 for two sizes, 16 for the others: one launch per kernel and
    tile group) reaches, to the bit, what each film reaches alone with the
    same tile rows and plain launches."""
    params = _weights()
    films = _films()
    try:
        capi.set_tile_rows(SIDES[0], SIDES[0], 8)
        capi.set_tile_rows(SIDES[2], SIDES[2], 8)
        ref = _alone(films, params, 2)
        gfilms = []
        for c, h, s, cc, dt in films:
            c = c.copy()
            c.graphs = 1
            gfilms.append((c, h, s, cc, dt))
        b = HipBatch(gfilms)
        b.set_params(params)
        counts = [[] for _ in films]
        for _ in range(2):
            rc, st, nk, lin = b.advance()
            assert rc == 0 and np.all(st == 0)
            for k in range(len(films)):
                counts[k].append((int(nk[k]), int(lin[k])))
        for k, (rcounts, rh, _) in enumerate(ref):
            print(f"film {k} ({SIDES[k]}^2, tile rows {capi.tile_rows(SIDES[k], SIDES[k])}): alone {rcounts}, "
                  f"batch with graph {counts[k]}")
            assert counts[k] == rcounts
            assert np.array_equal(b.h(k), rh)
        b.close()
    finally:
        for n in SIDES:
            capi.set_tile_rows(n, n, 16)


def test_tile_tuner():
    """The tuner times both tile shapes on a synthetic system of the size and
    makes the faster one the process's choice for that size"""
    rows, us = capi.tune_tiles(192, 192, nfilms=3, iters=20)
    print(f"3 films of 192^2: {us[0]:.2f} us per iteration with 16 rows, {us[1]:.2f} with 8: chose {rows}")
    assert rows in (8, 16) and all(u > 0 for u in us)
    assert rows == (8 if us[1] < 0.97 * us[0] else 16)
    assert capi.tile_rows(192, 192) == rows
    capi.set_tile_rows(192, 192, 16)
    with pytest.raises(ValueError):
        capi.set_tile_rows(192, 192, 12)


def test_batch_refuses_what_it_cannot_take():
    c, h, s, cc, dt = cases.sweep_film(0, side=40)
    c.step_control = 1
    with pytest.raises(RuntimeError):
        HipBatch([(c, h, s, cc, dt)])


def test_sweep_of_256_films():
    """The parameter sweep at its own size: the 256 films of cases.sweep (sides 64 to
    512, 19.8 M cells) advanced one step together. Every film reaches, to
    the bit and with the same Newton and Krylov counts, what it reaches in
    a step context of its own with the same network; and one film of each
    side is checked against the CPU oracle's FP64 step at the bar of the
    learned FP32 path, max |dh| / max h <= 1e-8."""
    from oracle import binding as ob
    params = _weights()
    films = cases.sweep(256)
    for n in cases.SWEEP_SIDES:
        rows, us = capi.tune_tiles(n, n, nfilms=sum(1 for f in films if f[0].nx == n), iters=20)
        print(f"tile tuner {n}^2: {us[0]:.2f} / {us[1]:.2f} us per iteration with 16 / 8 rows: {rows}", flush=True)
    cells = sum(f[0].nx * f[0].ny for f in films)
    print(f"sweep: {len(films)} films, {cells} cells", flush=True)
    b = HipBatch(films)
    b.set_params(params)
    rc, st, nk, lin = b.advance()
    assert rc == 0 and np.all(st == 0)
    hb = [b.h(k) for k in range(len(films))]
    b.close()
    checked = {}
    for k, (c, h, s, cc, dt) in enumerate(films):
        g = HipStep(c, h, s, cc, dt)
        g.learn_set_params(params)
        r = g.advance()
        assert r[0] == 0
        assert (r[1], r[2]) == (int(nk[k]), int(lin[k])), f"film {k}: alone {r[1:3]}, batch {nk[k]} {lin[k]}"
        assert np.array_equal(g.h(), hb[k]), f"film {k} ({c.nx}^2) differs from its batch run"
        g.close()
        if c.nx not in checked:
            o = ob.OracleStep(c, h, s, cc, dt)
            ro = o.advance()
            assert ro[0] == 0
            ho = o.h()
            rel = float(np.abs(ho - hb[k]).max() / np.abs(ho).max())
            checked[c.nx] = rel
            print(f"film {k} ({c.nx}^2): device nk {nk[k]} lin {lin[k]}, oracle nk {ro[1]}; "
                  f"max|dh|/max h {rel:.2e}", flush=True)
            assert rel <= 1e-8
    assert sorted(checked) == sorted(cases.SWEEP_SIDES)
    for n in cases.SWEEP_SIDES:
        capi.set_tile_rows(n, n, 16)
