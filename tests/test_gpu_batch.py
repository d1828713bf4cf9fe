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


def test_batch_refuses_what_it_cannot_take():
    c, h, s, cc, dt = cases.sweep_film(0, side=40)
    c.step_control = 1
    with pytest.raises(RuntimeError):
        HipBatch([(c, h, s, cc, dt)])
