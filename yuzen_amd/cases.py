"""The benchmark and parity cases as concrete, seeded inputs.

Each builder returns ``(case, h, s, c, dt0, steps)``: a filled
:class:`yuzen_amd.capi.YzCase`, the padded film, substrate and
concentration arrays (n = (nx+4)(ny+4) rows, ghosts left 0: the step forms
them from the boundary conditions), the first step and the step count.

Noise comes from a counter-based generator keyed by (seed, ix, iy)
(splitmix64 of the key), so every consumer -- the device path and the CPU
oracle -- sees the same initial film bit for bit, whatever the order it is
generated in.
"""
import math

import numpy as np

from .capi import (YzCase, case_defaults, BC_SYMMETRY, BC_PERIODIC,
                   PC_AMG, PC_LEARNED, cz_case_defaults)

_M64 = (1 << 64) - 1


def _splitmix64(z):
    z = (z + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def uniform_pm1(seed, ix, iy):
    """U(-1, 1) for integer arrays ix, iy (interior cell numbers from 0),
    keyed by (seed, ix, iy)."""
    ix = np.asarray(ix, dtype=np.uint64)
    iy = np.asarray(iy, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) * np.uint64(0x100000001B3)) ^ (iy << np.uint64(32)) ^ ix
        z = key + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


def padded_shape(case):
    nxp = case.nx + 4
    nyp = case.ny + 4 if case.dim == 2 else 1
    return nyp, nxp


def cell_centres(case):
    """Cell-centre coordinates over the padded grid, as the grid builds
    them (two ghosts, spacing over the padded extent)."""
    nyp, nxp = padded_shape(case)
    dx0 = (case.xmax - case.xmin) / case.nx
    xlo = case.xmin - 2 * dx0
    dx = ((case.xmax + 2 * dx0) - xlo) / nxp
    x = xlo + np.arange(nxp) * dx + dx * 0.5
    if case.dim == 2:
        dy0 = (case.ymax - case.ymin) / case.ny
        ylo = case.ymin - 2 * dy0
        dy = ((case.ymax + 2 * dy0) - ylo) / nyp
        y = ylo + np.arange(nyp) * dy + dy * 0.5
    else:
        y = np.zeros(1)
    return np.meshgrid(x, y)


def interior_index(case):
    nyp, nxp = padded_shape(case)
    ix = np.arange(nxp) - 2
    iy = (np.arange(nyp) - 2) if case.dim == 2 else np.zeros(1, dtype=int)
    IX, IY = np.meshgrid(ix, iy)
    return IX, IY


def disjoining_dp(theta_deg, hstar, n=3.0, m=2.0, As=1.0):
    """Dp = As (n-1)(m-1) / (2 H* (n-m)) theta_e^2
    (reference NonDimensionalNumbers.cpp:103-118)."""
    th = theta_deg * math.pi / 180.0
    return As * (n - 1.0) * (m - 1.0) / (2.0 * hstar * (n - m)) * th * th


def leveling_1d(nx=512, steps=200):
    """1-D capillary leveling over a substrate step: N cells on x in
    [0, 10], s = 0.5 (1 + tanh((x - 5)/0.2)), film h = 1 on it, capillary
    pressure only, Symmetry (no-flux) ends, dt = 1e-3 fixed, Newton
    tolerance 1e-10, BiCGSTAB + AMG."""
    c = case_defaults()
    c.dim, c.nx, c.ny = 1, nx, 1
    c.xmin, c.xmax, c.ymin, c.ymax = 0.0, 10.0, -1.0, 1.0
    for k in range(4):
        c.bc[k] = BC_SYMMETRY
    c.capillary = 1
    c.nk_eps = 1e-10
    c.precond = PC_AMG
    X, _ = cell_centres(c)
    s = 0.5 * (1.0 + np.tanh((X - 5.0) / 0.2))
    h = np.ones_like(X)
    conc = np.full_like(X, 0.5)
    return c, h.ravel(), s.ravel(), conc.ravel(), 1e-3, steps


def spinodal_2d(n=224, steps=2000, seed=2024):
    """2-D spinodal dewetting: n x n periodic cells, side 0.64, h = 1 +
    0.01 U(-1, 1) per cell, capillary + disjoining Pi = (h*/h)^3 -
    (h*/h)^2 with h* = 0.01 and theta_e = 10 deg, dt adaptive from 1e-5."""
    c = case_defaults()
    c.dim, c.nx, c.ny = 2, n, n
    c.xmin, c.xmax, c.ymin, c.ymax = 0.0, 0.64, 0.0, 0.64
    for k in range(4):
        c.bc[k] = BC_PERIODIC
    c.capillary = 1
    c.disjoining = 1
    c.hstar, c.dis_n, c.dis_m = 0.01, 3.0, 2.0
    c.Dp = disjoining_dp(10.0, 0.01)
    c.step_control = 1
    c.tsc_change, c.tsc_criterion, c.tsc_keep = 0.5, 0.01, 50
    c.dt_min, c.dt_max = 1e-15, 1e-5 * 1e4
    IX, IY = interior_index(c)
    h = 1.0 + 0.01 * uniform_pm1(seed, np.clip(IX, 0, n - 1), np.clip(IY, 0, n - 1))
    s = np.zeros_like(h)
    conc = np.full_like(h, 0.5)
    return c, h.ravel(), s.ravel(), conc.ravel(), 1e-5, steps


def allphys_2d(n=1024, steps=20, seed=7, L=10.0, dt=1e-4, nk_eps=1e-8):
    """2-D all-physics film, the throughput case: n x n cells on a square
    of side L = 10 with Symmetry walls; capillary, disjoining (h* = 0.01,
    theta_e = 15 deg), vertical gravity Bo = 0.5, in-plane gravity
    Pgx = 0.1, Pgy = 0.05, spin Sp = 0.2 about the centre of the square,
    Marangoni As Ma = 0.1 (As = 1, Ma = 0.1) on a frozen concentration
    c = 0.5 + 0.1 sin(2 pi x / L) cos(2 pi y / L), viscosity
    mu = exp(2 (c - 0.5)), slip Lambda = 0.05; film
    h = 0.55 + 0.45 tanh((x cos30 + y sin30 - L/2)/0.05) + 1e-3 U(-1,1),
    flat substrate; fixed dt = 1e-4, Newton tolerance 1e-8."""
    c = case_defaults()
    c.dim, c.nx, c.ny = 2, n, n
    c.xmin, c.xmax, c.ymin, c.ymax = 0.0, L, 0.0, L
    for k in range(4):
        c.bc[k] = BC_SYMMETRY
    c.capillary = c.disjoining = c.vertical_gravity = 1
    c.inplane_gravity = c.centrifugal = c.marangoni = 1
    c.hstar, c.dis_n, c.dis_m = 0.01, 3.0, 2.0
    c.Dp = disjoining_dp(15.0, 0.01)
    c.Bo, c.Pgx, c.Pgy, c.Sp = 0.5, 0.1, 0.05, 0.2
    c.As, c.Ma = 1.0, 0.1
    c.slip = 0.05
    c.rot_x, c.rot_y = 0.5 * L, 0.5 * L
    c.visc_c0, c.visc_exp, c.sig_c0 = 0.5, 2.0, 0.5
    c.nk_eps = nk_eps
    X, Y = cell_centres(c)
    IX, IY = interior_index(c)
    a = math.radians(30.0)
    h = 0.55 + 0.45 * np.tanh((X * math.cos(a) + Y * math.sin(a) - 0.5 * L) / 0.05)
    h = h + 1e-3 * uniform_pm1(seed, np.clip(IX, 0, n - 1), np.clip(IY, 0, n - 1))
    s = np.zeros_like(h)
    conc = 0.5 + 0.1 * np.sin(2 * math.pi * X / L) * np.cos(2 * math.pi * Y / L)
    return c, h.ravel(), s.ravel(), conc.ravel(), dt, steps


SWEEP_SIDES = (64, 96, 128, 192, 256, 384, 512)


def sweep_film(k, seed=4242, side=None, dx=0.01, dt=2e-6, nk_eps=1e-8):
    """Film k of the batched parameter sweep: side N drawn uniformly from
    {64, 96, 128, 192, 256, 384, 512} (or the given side), theta_e ~
    U(5, 20) degrees (disjoining, h* = 0.01, n = 3, m = 2), spin Sp ~
    U(0, 0.5) about the square's centre, Bo ~ U(0, 1), capillary on;
    h = 1 + 0.02 U(-1, 1) per cell; square of side N dx (dx = 0.01),
    Symmetry walls, frozen uniform concentration; fixed dt 2e-6 (the cell
    size and step of the 1024^2 all-physics film), Newton
    tolerance 1e-8; FP32 Krylov with the learned preconditioner (the
    network shared by the batch: learn_passes 0). The draws for film k are
    keyed by (seed, k), the film's noise by (seed + 1 + k, ix, iy)."""
    def draw(j):
        return 0.5 * (uniform_pm1(seed, k, j) + 1.0)
    n = side if side is not None else SWEEP_SIDES[min(int(draw(0) * len(SWEEP_SIDES)), len(SWEEP_SIDES) - 1)]
    theta = 5.0 + 15.0 * draw(1)
    sp = 0.5 * draw(2)
    bo = draw(3)
    L = n * dx
    c = case_defaults()
    c.dim, c.nx, c.ny = 2, n, n
    c.xmin, c.xmax, c.ymin, c.ymax = 0.0, L, 0.0, L
    for b in range(4):
        c.bc[b] = BC_SYMMETRY
    c.capillary = c.disjoining = c.vertical_gravity = c.centrifugal = 1
    c.hstar, c.dis_n, c.dis_m = 0.01, 3.0, 2.0
    c.Dp = disjoining_dp(theta, 0.01)
    c.Bo, c.Sp = bo, sp
    c.rot_x, c.rot_y = 0.5 * L, 0.5 * L
    c.nk_eps = nk_eps
    c.precond, c.krylov_fp32, c.learn_passes = PC_LEARNED, 1, 0
    IX, IY = interior_index(c)
    h = 1.0 + 0.02 * uniform_pm1(seed + 1 + k, np.clip(IX, 0, n - 1), np.clip(IY, 0, n - 1))
    s = np.zeros_like(h)
    conc = np.full_like(h, 0.5)
    return c, h.ravel(), s.ravel(), conc.ravel(), dt


def sweep(n_films=256, seed=4242, **kw):
    """The batched parameter sweep: n_films films of sweep_film."""
    return [sweep_film(k, seed=seed, **kw) for k in range(n_films)]


def concz_2d(n=256, nz=32, L=10.0, dt=0.01, steps=500, c0=0.3):
    """The ConcZ film: n x n cells x nz levels through the film,
    periodic, h = 1 + 0.05 cos(2 pi x / L) cos(2 pi y / L) on a flat
    substrate, c0 = 0.3 uniform; capillary and Marangoni flow
    (sigma = 1 + Ma (c_s - c0), Ma = 1, on the surface concentration),
    mu = exp(5 (c - c0)), Df = 1 with a uniform diffusivity, Raoult
    evaporation with Ev = 0.05; As = Ev / (Pe Df) = 0.005 for Pe = 10
    (Pe = E0 h0 / D0, Ev = E0 T0 / h0, Df = D0 T0 / L0^2:
    NonDimensionalNumbers.cpp:119-123); fixed dt.

    Returns ``(case, h, s, c, dt, steps)``: h and s padded
    (n+4)^2, c over every cell of the 3-D grid with its ghosts."""
    c = cz_case_defaults()
    c.nx = c.ny = n
    c.nz = nz
    c.xmin, c.xmax, c.ymin, c.ymax = 0.0, L, 0.0, L
    for k in range(4):
        c.bc[k] = BC_PERIODIC
    c.capillary, c.marangoni, c.evaporation = 1, 1, 1
    c.As, c.Ma, c.sig_c0 = 0.005, 1.0, c0
    c.visc_c0, c.visc_exp = c0, 5.0
    c.Df, c.diff_c0, c.diff_exp = 1.0, c0, 0.0
    c.Ev = 0.05
    dx = L / n
    x = -2 * dx + (np.arange(n + 4) + 0.5) * dx
    X, Y = np.meshgrid(x, x)
    h = (1.0 + 0.05 * np.cos(2 * np.pi * X / L) * np.cos(2 * np.pi * Y / L)).ravel()
    cc = np.full((nz + 4) * (n + 4) ** 2, c0)
    return c, h, np.zeros_like(h), cc, dt, steps
