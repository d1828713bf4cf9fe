"""Device path against the CPU oracle (FP64 plain method).

Assembly: the rows of IA = I + theta dt J, A(h)h and the Newton right-hand
side, term by term, 1-D and 2-D, with and without slip, Symmetry and
Periodic edges, against the oracle to 1e-12 of each row's scale (the two
sum the same terms in orders that differ only in rounding).

Steps: per step max|dh| / max h <= 1e-10 with equal Newton counts.
"""
import numpy as np
import pytest

from yuzen_amd import capi, cases
from yuzen_amd.step import HipStep, LinSolver
from oracle import binding as ob

from .helpers import TERMS, smooth_case, csr_to_planes, interior

pytestmark = pytest.mark.gpu


def _assembly_case(dim, terms, slip, periodic):
    c, h, s, cc = smooth_case(dim, terms, slip=slip, periodic=periodic)
    o = ob.OracleStep(c, h, s, cc, 1e-3)
    o.set_h(h)
    hf = o.h()
    rng = np.random.default_rng(3)
    h_old = hf * (1.0 + 0.01 * rng.standard_normal(hf.size))
    dt = 2e-3
    A, J, IA, ax, rhs = o.assemble(h_old, hf, dt)
    rp, col, bc = o.pattern()
    ref_planes = csr_to_planes(c, rp, col, IA)
    g = HipStep(c, h, s, cc, 1e-3)
    ia, gax, grhs = g.assemble(h_old, hf, dt)
    g.close()
    return c, ref_planes, ia, ax, gax, rhs, grhs


@pytest.mark.parametrize("dim", [1, 2])
@pytest.mark.parametrize("slip", [0.0, 0.05])
@pytest.mark.parametrize("term", TERMS + ["all"])
def test_assembly_rows(dim, slip, term):
    terms = TERMS if term == "all" else [term]
    print(f"assembly dim={dim} slip={slip} term={term}", flush=True)
    c, ref, ia, ax, gax, rhs, grhs = _assembly_case(dim, terms, slip, periodic=False)
    scale = np.abs(ref).max(axis=0)
    assert np.all(np.abs(ia - ref) <= 1e-12 * scale + 1e-300), np.abs(ia - ref).max()
    axi, gaxi = interior(c, ax), interior(c, gax)
    assert np.abs(axi - gaxi).max() <= 1e-12 * max(np.abs(axi).max(), 1e-300)
    ri, gri = interior(c, rhs), interior(c, grhs)
    assert np.abs(ri - gri).max() <= 1e-12 * max(np.abs(ri).max(), 1e-300) + 1e-15


@pytest.mark.parametrize("dim", [1, 2])
def test_assembly_periodic(dim):
    print(f"assembly periodic dim={dim}", flush=True)
    c, ref, ia, ax, gax, rhs, grhs = _assembly_case(dim, TERMS, 0.05, periodic=True)
    scale = np.abs(ref).max(axis=0)
    assert np.all(np.abs(ia - ref) <= 1e-12 * scale)
    ri, gri = interior(c, rhs), interior(c, grhs)
    assert np.abs(ri - gri).max() <= 1e-12 * np.abs(ri).max()


def _run_pair(c, h, s, cc, dt, steps, label):
    o = ob.OracleStep(c, h, s, cc, dt, lin_mode=ob.OracleStep.LIN_STRUCTURED)
    g = HipStep(c, h, s, cc, dt)
    worst = 0.0
    for k in range(steps):
        ro = o.advance()
        rg = g.advance()
        ho, hg = o.h(), g.h()
        rel = np.abs(ho - hg).max() / np.abs(ho).max()
        worst = max(worst, rel)
        print(f"{label} step {k}: oracle nk={ro[1]} lin={ro[2]}  hip nk={rg[1]} lin={rg[2]}  rel={rel:.2e}", flush=True)
        assert ro[0] == 0 and rg[0] == 0
        assert ro[1] == rg[1], "Newton counts differ"
        assert rel <= 1e-10
        if c.step_control:
            assert abs(o.clock()[1] - g.clock()[1]) <= 1e-15 * o.clock()[1]
    g.close()
    return worst


def test_step_leveling_1d():
    c, h, s, cc, dt, _ = cases.leveling_1d(nx=512)
    _run_pair(c, h, s, cc, dt, 8, "leveling_1d")


def test_step_allphys_2d_small():
    c, h, s, cc, dt, _ = cases.allphys_2d(n=48, L=2.0, dt=2e-5)
    _run_pair(c, h, s, cc, dt, 4, "allphys_2d")


def test_step_allphys_2d_zero_guess():
    """The plain FP64 path with every pass's solve started from zero
    (zero_guess 1, a departure from the reference's guess) against the
    oracle run the same way."""
    c, h, s, cc, dt, _ = cases.allphys_2d(n=48, L=2.0, dt=2e-5)
    c.zero_guess = 1
    _run_pair(c, h, s, cc, dt, 4, "allphys_2d zero guess")


@pytest.mark.parametrize("interval", [2, 3])
def test_step_modified_newton(interval):
    """JacobianUpdateInterval 2 and 3 (NewtonKantorovich.cpp:80-91): IA, its
    row scale and its multigrid kept between the passes that form it; equal
    Newton counts with the oracle run the same way; the learned FP32 path
    keeps its approximate inverse too and holds the learned bar"""
    c, h, s, cc, dt, _ = cases.allphys_2d(n=48, L=2.0, dt=2e-5)
    c.jac_interval = interval
    _run_pair(c, h, s, cc, dt, 4, f"allphys_2d jacobian interval {interval}")
    cl = c.copy()
    cl.precond, cl.krylov_fp32, cl.learn_passes, cl.learn_iters = capi.PC_LEARNED, 1, 1, 40
    o = ob.OracleStep(c, h, s, cc, dt, lin_mode=ob.OracleStep.LIN_STRUCTURED)
    g = HipStep(cl, h, s, cc, dt)
    for k in range(3):
        ro, rg = o.advance(), g.advance()
        assert ro[0] == 0 and rg[0] == 0
        rel = np.abs(o.h() - g.h()).max() / np.abs(o.h()).max()
        print(f"learned FP32, interval {interval}, step {k}: oracle nk={ro[1]} hip nk={rg[1]} rel={rel:.2e}")
        assert rel <= 1e-8
    g.close()


def test_step_spinodal_2d_small():
    c, h, s, cc, dt, _ = cases.spinodal_2d(n=32)
    _run_pair(c, h, s, cc, dt, 4, "spinodal_2d")


def test_linsolve_seam_matches_oracle():
    c, h, s, cc = smooth_case(2, TERMS, slip=0.05)
    o = ob.OracleStep(c, h, s, cc, 1e-3)
    o.set_h(h)
    hf = o.h()
    A, J, IA, ax, rhs = o.assemble(hf * 1.01, hf, 5e-3)
    rp, col, bc = o.pattern()
    rc0, x0, it0, r0 = ob.linsolve(rp, col, IA, rhs, np.zeros(o.n), eps=1e-12)
    ls = LinSolver()
    for grid in (None, (2, c.nx, c.ny, [0, 0, 0, 0])):
        rc, x, it, r = ls.solve(rp, col, IA, rhs, np.zeros(o.n), epsilon=1e-12, grid=grid)
        print(f"linsolve grid={grid is not None}: rc={rc} it={it} resid={r:.2e} (oracle it={it0})", flush=True)
        assert rc == 0 and rc0 == 0
        assert np.abs(x - x0).max() <= 1e-9 * np.abs(x0).max()


@pytest.mark.parametrize("dim", [1, 2])
def test_linsolve_seam_takes_the_concentration_system(dim):
    """A concentration system of the reference's layout (the 9-point wide
    stencil in 2-D, ghost rows) through the seam with the concentration
    role: the grid read off the pattern, the solution to the tolerance."""
    if dim == 2:
        c, h, s, cc, dt, _ = cases.allphys_2d(n=40, L=1.5, dt=2e-5)
    else:
        c, h, s, cc, dt, _ = cases.leveling_1d(nx=128, steps=1)
        c.marangoni, c.As, c.Ma = 1, 1.0, 0.3
        X, _ = cases.cell_centres(c)
        cc = (0.5 + 0.2 * np.tanh((X - 4.0) / 0.5)).ravel()
    c.conc_model, c.Df = 1, 0.05
    o = ob.OracleStep(c, h, s, cc, dt)
    assert o.advance()[0] == 0
    rp, col, val, rhs = o.conc_system()
    rc0, x0, it0, r0 = ob.linsolve(rp, col, val, rhs, np.zeros(o.n), eps=1e-13)
    ls = LinSolver()
    rc, x, it, r = ls.solve(rp, col, val, rhs, np.zeros(o.n), epsilon=1e-13, conc=1)
    print(f"concentration seam {dim}-D: rc={rc} it={it} resid={r:.2e} (oracle it={it0})", flush=True)
    assert rc == 0 and rc0 == 0
    assert np.abs(x - x0).max() <= 1e-10 * max(np.abs(x0).max(), 1e-300)
    assert np.abs(x0).max() > 0


def _scaled_residual(rp, col, val, rhs, x):
    """||b - A x|| / ||b|| of the system with its rows divided by their
    diagonals (the scaling the seam's tolerance is on)"""
    n = len(rp) - 1
    r = np.empty(n)
    bs = np.empty(n)
    for i in range(n):
        sl = slice(rp[i], rp[i + 1])
        d = val[sl][col[sl] == i].sum()
        d = d if abs(d) >= 1e-30 else 1.0
        r[i] = (rhs[i] - val[sl] @ x[col[sl]]) / d
        bs[i] = rhs[i] / d
    return np.linalg.norm(r) / np.linalg.norm(bs)


def test_linsolve_seam_return_codes():
    """The backend contract's codes through the seam: 1 when the iteration
    cap is reached first, 2 at a breakdown (a zero matrix: (rh, v) = 0),
    and the matrix and right-hand side left as given. A matrix that is not
    a structured film or concentration system (a ghost row that no longer
    copies, an interior row reaching past the stencil) takes the general
    CRS path and is solved to the tolerance."""
    c, h, s, cc = smooth_case(2, TERMS, slip=0.05)
    o = ob.OracleStep(c, h, s, cc, 1e-3)
    o.set_h(h)
    hf = o.h()
    A, J, IA, ax, rhs = o.assemble(hf * 1.01, hf, 5e-3)
    rp, col, bc = o.pattern()
    IA0, rhs0 = IA.copy(), rhs.copy()
    ls = LinSolver()
    rc, x, it, r = ls.solve(rp, col, IA, rhs, np.zeros(o.n), epsilon=1e-14, max_iter=2)
    print(f"cap: rc={rc} it={it} resid={r:.2e}")
    assert rc == 1 and it == 2 and r > 1e-14
    assert np.array_equal(IA, IA0) and np.array_equal(rhs, rhs0)
    bad = IA.copy()
    g = int(np.flatnonzero(bc)[0])                       # a ghost row: not a copy row any more
    bad[rp[g]:rp[g + 1]] = [2.0, 1.0]
    rc, x, it, r = ls.solve(rp, col, bad, rhs, np.zeros(o.n), precond=capi.PC_JACOBI, epsilon=1e-10,
                            max_iter=20000)
    res = _scaled_residual(rp, col, bad, rhs, x)
    print(f"no copy row (general path): rc={rc} it={it} resid={r:.2e} (checked {res:.2e})")
    assert rc == 0 and res <= 1.01e-10
    zero = np.zeros_like(IA)
    rc, x, it, r = ls.solve(rp, col, zero, rhs, np.zeros(o.n), epsilon=1e-12)
    print(f"zero matrix: rc={rc}")
    assert rc == 2


@pytest.mark.parametrize("kind", ["conv_diff_3d", "random"])
@pytest.mark.parametrize("precond", [capi.PC_AMG, capi.PC_JACOBI], ids=["spai0", "jacobi"])
def test_linsolve_general_crs(kind, precond):
    """The seam's general path (no structured grid): a 3-D 7-point
    convection-diffusion system on 20^3 cells and a random sparse
    diagonally dominant one (about 12 entries a row, columns in any order),
    solved to the tolerance on the scaled system, the matrix left as given,
    and the same bits on a second call"""
    rng = np.random.default_rng(7)
    if kind == "conv_diff_3d":
        m = 20
        n = m ** 3
        rows = []
        for k in range(n):
            ix, iy, iz = k % m, (k // m) % m, k // (m * m)
            ent = {k: 6.0 + 0.5 * rng.random()}
            for d, (dx, dy, dz) in enumerate(((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))):
                jx, jy, jz = ix + dx, iy + dy, iz + dz
                if 0 <= jx < m and 0 <= jy < m and 0 <= jz < m:
                    ent[jx + m * (jy + m * jz)] = -1.0 + (0.3 if d == 0 else 0.0)
            rows.append(ent)
    else:
        n = 5000
        rows = []
        for k in range(n):
            cols = rng.choice(n, 11, replace=False)
            ent = {int(j): rng.uniform(-1, 1) for j in cols if j != k}
            ent[k] = sum(abs(v) for v in ent.values()) + 0.5
            rows.append(ent)
    rp = np.zeros(n + 1, dtype=np.int32)
    col, val = [], []
    for k, ent in enumerate(rows):
        items = list(ent.items())
        rng.shuffle(items)
        col += [j for j, _ in items]
        val += [v for _, v in items]
        rp[k + 1] = len(col)
    col, val = np.array(col, dtype=np.int32), np.array(val)
    rhs = rng.uniform(-1, 1, n)
    val0 = val.copy()
    ls = LinSolver()
    rc, x, it, r = ls.solve(rp, col, val, rhs, np.zeros(n), precond=precond, epsilon=1e-12)
    res = _scaled_residual(rp, col, val, rhs, x)
    print(f"{kind}: n={n} nnz={len(col)} rc={rc} it={it} resid={r:.2e} (checked {res:.2e})")
    assert rc == 0 and res <= 1.01e-12 and np.array_equal(val, val0)
    rc2, x2, it2, _ = ls.solve(rp, col, val, rhs, np.zeros(n), precond=precond, epsilon=1e-12)
    assert rc2 == 0 and it2 == it and np.array_equal(x, x2)


def test_linsolve_context_reused_across_shapes():
    """One context solving systems of the same size but another grid (the
    transposed shape) or other edges (periodic, then Symmetry) gives what a
    fresh context gives: its multigrid is rebuilt for the new operator."""
    ls = LinSolver()
    for nx, ny, per in ((18, 14, False), (14, 18, False), (14, 18, True), (14, 18, False)):
        c, h, s, cc = smooth_case(2, TERMS, periodic=per, nx=nx, ny=ny)
        o = ob.OracleStep(c, h, s, cc, 1e-3)
        o.set_h(h)
        hf = o.h()
        _, _, IA, _, rhs = o.assemble(hf * 1.01, hf, 5e-3)
        rp, col, _ = o.pattern()
        rc, x, it, _ = ls.solve(rp, col, IA, rhs, np.zeros(o.n), epsilon=1e-12)
        rc1, x1, it1, _ = LinSolver().solve(rp, col, IA, rhs, np.zeros(o.n), epsilon=1e-12)
        print(f"{nx}x{ny} periodic={per}: reused it={it}, fresh it={it1}")
        assert rc == 0 and rc1 == 0 and it == it1
        assert np.array_equal(x, x1)


def test_linsolve_seam_time_at_1024():
    """The seam on the 1024^2 all-physics film's IA (1 056 784 rows): the
    solution to the tolerance, and the time of a call (the CRS copied to the
    device, folded into planes there, the multigrid set up, the solve, x
    back) printed beside the solve alone."""
    import time
    c, h, s, cc, dt, _ = cases.allphys_2d(n=1024, dt=2e-6)
    o = ob.OracleStep(c, h, s, cc, dt)
    o.set_h(h)
    hf = o.h()
    _, _, IA, _, rhs = o.assemble(hf, hf, dt)
    rp, col, _ = o.pattern()
    ls = LinSolver()
    ls.solve(rp, col, IA, rhs, np.zeros(o.n), epsilon=1e-8)        # first call: allocations
    t = time.perf_counter()
    rc, x, it, r = ls.solve(rp, col, IA, rhs, np.zeros(o.n), epsilon=1e-8)
    el = time.perf_counter() - t
    print(f"seam at 1024^2: rc={rc} it={it} resid={r:.2e}, {1e3 * el:.1f} ms per call")
    assert rc == 0 and r <= 1e-8
