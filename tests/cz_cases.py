"""ConcZ films shared by the oracle and the device tests: grids with two
ghost layers per side, row = (n+4) iy + ix and (n+4)^2 kz + (n+4) iy + ix."""
import numpy as np

from yuzen_amd import capi


def _grid(n, L):
    dx = L / n
    return -2 * dx + (np.arange(n + 4) + 0.5) * dx


def _case(n=12, nz=8, L=4.0, per=True, **kw):
    c = capi.cz_case_defaults()
    c.nx = c.ny = n
    c.nz = nz
    c.xmin, c.xmax, c.ymin, c.ymax = 0.0, L, 0.0, L
    for k in range(4):
        c.bc[k] = capi.BC_PERIODIC if per else capi.BC_SYMMETRY
    c.marangoni = c.evaporation = 0
    c.As, c.Ma, c.sig_c0, c.visc_c0, c.visc_exp = 0.005, 0.0, 0.3, 0.3, 0.0
    c.Df, c.diff_c0, c.Ev = 1.0, 0.3, 0.05
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def _real3(c, a):
    return np.asarray(a).reshape(c.nz + 4, c.ny + 4, c.nx + 4)[2:-2, 2:-2, 2:-2]


def _real2(c, a):
    return np.asarray(a).reshape(c.ny + 4, c.nx + 4)[2:-2, 2:-2]


def _solute(c, h, cc):
    return (_real2(c, h) * _real3(c, cc).mean(axis=0)).sum()


def _allphys(n=12, nz=8, L=4.0, per=True, **kw):
    c = _case(n=n, nz=nz, L=L, per=per, marangoni=1, evaporation=1, Ma=1.0, As=0.05,
              visc_exp=5.0, vertical_gravity=1, Bo=0.3, inplane_gravity=1, Pgx=0.05, Pgy=-0.03, **kw)
    x = _grid(n, L)
    X, Y = np.meshgrid(x, x)
    h = 1.0 + 0.05 * np.cos(2 * np.pi * X / L) * np.cos(2 * np.pi * Y / L) + 0.03 * np.sin(2 * np.pi * Y / L)
    s = 0.02 * np.cos(2 * np.pi * X / L) ** 2 if not per else np.zeros_like(h)
    cc = np.full((nz + 4, n + 4, n + 4), 0.3) + 0.01 * np.sin(2 * np.pi * X / L)[None]
    return c, h.ravel(), s.ravel(), cc.ravel()
