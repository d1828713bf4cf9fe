"""yz_linsolve_noalloc, the seam's solve in the library's scratch array
(no allocation), through the C-ABI: two calls in a row, the second on a
smaller system, each to the bit what yz_linsolve (which allocates its
arrays) gives on the same system, with the same iteration count; a system
larger than the scratch is refused with a message."""
import numpy as np
import pytest

from oracle import binding as ob
from yuzen_amd import capi, cases
from yuzen_amd.step import LinSolver, linsolve_noalloc

from .helpers import TERMS, smooth_case

pytestmark = pytest.mark.gpu


def _film_system(c, h, s, cc, dt):
    o = ob.OracleStep(c, h, s, cc, dt)
    o.set_h(h)
    hf = o.h()
    _, _, IA, _, rhs = o.assemble(hf * 1.001, hf, dt)
    rp, col, _ = o.pattern()
    return rp, col, IA, rhs, o.n


def test_noalloc_matches_the_allocating_seam_twice_then_smaller():
    big = cases.allphys_2d(n=96, L=3.0, dt=2e-5)
    small = smooth_case(2, TERMS, slip=0.05)
    systems = [
        ("96^2 all-physics film", _film_system(*big[:4], 2e-5), 0),
        ("18 x 14 all terms", _film_system(*small, 5e-3), 0),
    ]
    ls = LinSolver()
    for label, (rp, col, val, rhs, n), conc in systems:
        x0 = np.zeros(n)
        rc_a, x_a, it_a, r_a = ls.solve(rp, col, val, rhs, x0, epsilon=1e-12, conc=conc)
        rc_n, x_n, it_n, r_n = linsolve_noalloc(rp, col, val, rhs, x0, epsilon=1e-12, conc=conc)
        print(f"{label}: seam rc={rc_a} it={it_a} resid={r_a:.2e}; no-alloc rc={rc_n} it={it_n} resid={r_n:.2e}",
              flush=True)
        assert rc_a == 0 and rc_n == 0
        assert it_n == it_a and r_n == r_a
        assert np.array_equal(x_n, x_a)


def test_noalloc_refuses_a_system_beyond_its_scratch():
    c, h, s, cc, dt, _ = cases.allphys_2d(n=400, dt=2e-6)
    rp, col, val, rhs, n = _film_system(c, h, s, cc, dt)
    rc, x, it, r = linsolve_noalloc(rp, col, val, rhs, np.zeros(n), epsilon=1e-8)
    msg = capi.last_error()
    print(f"400^2: rc={rc}: {msg}")
    assert rc == 2 and "scratch" in msg
    # and the scratch still serves a system that fits
    small = smooth_case(2, TERMS)
    rp, col, val, rhs, n = _film_system(*small, 5e-3)
    rc, x, it, r = linsolve_noalloc(rp, col, val, rhs, np.zeros(n), epsilon=1e-12)
    assert rc == 0 and r <= 1e-12
