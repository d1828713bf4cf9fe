"""ctypes mirror of include/yuzen_gpu.h and the loader of libyuzen_hip.so.

The structure layout below must match ``struct yz_case`` and
``struct yz_lin_settings`` field for field; ``tests/test_capi.py`` checks the
sizes against the compiled library.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libyuzen_hip.so")

BC_SYMMETRY, BC_PERIODIC, BC_PINNING, BC_FRESHINLET, BC_OUTLET = range(5)
PC_JACOBI, PC_AMG, PC_LEARNED, PC_LEARNED_AMG = range(4)
HSCHEME_KOREN, HSCHEME_VANLEER = 0, 1
HSCHEME_UPWIND1, HSCHEME_CENTRAL2, HSCHEME_CENTRAL4, HSCHEME_UPWIND2 = 2, 3, 4, 5
HSCHEME_UTOPIA, HSCHEME_QUICK, HSCHEME_KK = 6, 7, 8
(CONV_UPWIND1, CONV_CENTRAL2, CONV_CENTRAL4, CONV_UPWIND2, CONV_UTOPIA, CONV_QUICK,
 CONV_KK, CONV_KOREN, CONV_VANLEER) = range(9)
LEARN_NPARAM, LEARN_NFEAT, LEARN_NOUT = 6288, 8, 13
EVAP_RAOULT, EVAP_POWERN = 0, 1


class YzSeriesScales(C.Structure):
    _fields_ = [("dimensioned", C.c_int), ("L0", C.c_double), ("H0", C.c_double), ("T0", C.c_double),
                ("sigma0", C.c_double), ("energy", C.c_int)]


class YzCase(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
        ("xmin", C.c_double), ("xmax", C.c_double),
        ("ymin", C.c_double), ("ymax", C.c_double),
        ("bc", C.c_int * 4),
        ("capillary", C.c_int), ("vertical_gravity", C.c_int),
        ("inplane_gravity", C.c_int), ("centrifugal", C.c_int),
        ("marangoni", C.c_int), ("disjoining", C.c_int),
        ("Bo", C.c_double), ("Pgx", C.c_double), ("Pgy", C.c_double),
        ("Sp", C.c_double), ("As", C.c_double), ("Ma", C.c_double),
        ("Dp", C.c_double),
        ("hstar", C.c_double), ("dis_n", C.c_double), ("dis_m", C.c_double),
        ("slip", C.c_double), ("rot_x", C.c_double), ("rot_y", C.c_double),
        ("visc_c0", C.c_double), ("visc_exp", C.c_double),
        ("sig_c0", C.c_double),
        ("theta", C.c_double),
        ("nk_max", C.c_int), ("nk_min", C.c_int), ("jac_interval", C.c_int),
        ("nk_eps", C.c_double), ("nk_eps_rel", C.c_double),
        ("ew", C.c_int), ("extrap", C.c_int),
        ("lin_maxit", C.c_int), ("lin_eps", C.c_double),
        ("row_scaling", C.c_int), ("precond", C.c_int),
        ("krylov_fp32", C.c_int),
        ("step_control", C.c_int),
        ("tsc_change", C.c_double), ("tsc_criterion", C.c_double),
        ("tsc_floor", C.c_double), ("tsc_keep", C.c_int),
        ("dt_min", C.c_double), ("dt_max", C.c_double),
        ("learn_passes", C.c_int), ("learn_iters", C.c_int),
        ("learn_probes", C.c_int), ("learn_lr", C.c_double),
        ("stall_window", C.c_int), ("graphs", C.c_int),
        ("conc_model", C.c_int),
        ("conc_convection", C.c_int), ("conc_diffusion", C.c_int),
        ("conc_scheme", C.c_int),
        ("Df", C.c_double), ("diff_c0", C.c_double), ("diff_exp", C.c_double),
        ("conc_lin_maxit", C.c_int), ("conc_lin_eps", C.c_double),
        ("conc_row_scaling", C.c_int),
        ("zero_guess", C.c_int),
        ("evaporation", C.c_int),
        ("Ev", C.c_double), ("evap_coef", C.c_double),
        ("evap_func", C.c_int),
        ("evap_power", C.c_double), ("evap_concr", C.c_double),
        ("molar_resin", C.c_double), ("molar_solvent", C.c_double),
        ("precursor", C.c_double),
        ("source", C.c_int),
        ("src_radius", C.c_double), ("src_rate", C.c_double), ("src_tau", C.c_double),
        ("src_arm_length", C.c_double), ("src_arm_x", C.c_double), ("src_arm_y", C.c_double),
        ("src_omega", C.c_double), ("src_angle0", C.c_double), ("src_ramp", C.c_double),
        ("src_plateau", C.c_double),
        ("src_subregions", C.c_int),
        ("height_scheme", C.c_int),
        ("src_conc_supplied", C.c_int), ("src_conc", C.c_double),
    ]

    def copy(self):
        c = YzCase()
        C.pointer(c)[0] = self
        return c


class YzTerm(C.Structure):
    _fields_ = [
        ("end_time", C.c_double), ("max_steps", C.c_longlong),
        ("steady_eps", C.c_double), ("steady_steps", C.c_int),
        ("dry", C.c_int), ("end_conc_type", C.c_int), ("end_conc_value", C.c_double),
    ]


RUN_GOING, RUN_NAN, RUN_CONC_RANGE, RUN_NEGATIVE_FILM, RUN_NEGATIVE_VOLUME, RUN_END, RUN_DRY, RUN_STEADY = range(8)


def term(end_time=0.0, max_steps=0, steady_eps=1e-3, steady_steps=10, dry=0, end_conc_type=0, end_conc_value=0.95):
    """the run's termination tests (the reference's defaults:
    SteadyStateCriterion 1e-3, SteadyStateSteps 10, EndConcType mean,
    EndConcValue 0.95)"""
    t = YzTerm()
    t.end_time, t.max_steps, t.steady_eps, t.steady_steps = end_time, max_steps, steady_eps, steady_steps
    t.dry, t.end_conc_type, t.end_conc_value = dry, end_conc_type, end_conc_value
    return t


class YzLinSettings(C.Structure):
    _fields_ = [
        ("precond", C.c_int), ("scaling", C.c_int), ("max_iter", C.c_int),
        ("epsilon", C.c_double),
        ("grid_dim", C.c_int), ("grid_nx", C.c_int), ("grid_ny", C.c_int),
        ("bc", C.c_int * 4),
        ("conc", C.c_int),
    ]


def case_defaults():
    """The reference's defaults (Input.cpp getin defaults), as
    yz_case_defaults sets them."""
    c = YzCase()
    c.dim, c.nx, c.ny = 2, 16, 16
    c.xmin, c.xmax, c.ymin, c.ymax = 0.0, 1.0, 0.0, 1.0
    c.capillary = 1
    c.dis_n, c.dis_m, c.hstar = 3.0, 2.0, 0.01
    c.visc_c0, c.sig_c0 = 0.5, 0.5
    c.As = 1.0
    c.theta = 0.5
    c.nk_max, c.nk_min, c.jac_interval = 30, 1, 1
    c.nk_eps, c.nk_eps_rel = 1e-10, 1e-6
    c.ew, c.extrap = 1, 1
    c.lin_maxit, c.lin_eps, c.row_scaling, c.precond = 5000, 1e-13, 1, PC_AMG
    c.tsc_change, c.tsc_criterion, c.tsc_floor, c.tsc_keep = 0.5, 0.01, 0.0, 50
    c.dt_min, c.dt_max = 1e-15, 1e30
    c.learn_passes, c.learn_iters, c.learn_probes, c.learn_lr = 1, 100, 2, 4e-3
    c.stall_window = 40
    c.graphs = 0
    c.conc_model, c.conc_convection, c.conc_diffusion, c.conc_scheme = 0, 1, 1, CONV_UTOPIA
    c.Df, c.diff_c0, c.diff_exp = 1.0, 0.5, 0.0
    c.conc_lin_maxit, c.conc_lin_eps, c.conc_row_scaling = 5000, 1e-13, 1
    c.zero_guess = -1
    c.evaporation, c.Ev, c.evap_coef, c.evap_func, c.evap_power, c.evap_concr = 0, 1.0, 1.0, EVAP_RAOULT, 0.5, 0.0
    c.molar_resin, c.molar_solvent, c.precursor = 2200.0, 106.0, 0.0
    c.source, c.src_radius, c.src_rate, c.src_tau = 0, 1.0, 0.0, 0.0
    c.src_arm_length, c.src_arm_x, c.src_arm_y = 1.0, 0.0, 0.0
    c.src_omega, c.src_angle0, c.src_ramp, c.src_plateau, c.src_subregions = 0.0, 0.0, 0.0, 1e10, 1
    return c


class YzCzCase(C.Structure):
    """struct yz_cz_case: one ConcZ film (CZ_HeightEquation_2D,
    CZ_ConcEquation_2D)."""
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("xmin", C.c_double), ("xmax", C.c_double),
        ("ymin", C.c_double), ("ymax", C.c_double),
        ("bc", C.c_int * 4),
        ("capillary", C.c_int), ("vertical_gravity", C.c_int),
        ("inplane_gravity", C.c_int), ("centrifugal", C.c_int),
        ("marangoni", C.c_int), ("evaporation", C.c_int),
        ("As", C.c_double), ("Ma", C.c_double), ("Bo", C.c_double),
        ("Pgx", C.c_double), ("Pgy", C.c_double), ("Sp", C.c_double),
        ("rot_x", C.c_double), ("rot_y", C.c_double),
        ("sig_c0", C.c_double),
        ("visc_c0", C.c_double), ("visc_exp", C.c_double),
        ("Df", C.c_double), ("diff_c0", C.c_double), ("diff_exp", C.c_double),
        ("Ev", C.c_double), ("evap_coef", C.c_double),
        ("evap_func", C.c_int),
        ("evap_power", C.c_double), ("evap_concr", C.c_double),
        ("molar_resin", C.c_double), ("molar_solvent", C.c_double),
        ("precursor", C.c_double),
        ("conc_scheme", C.c_int),
        ("lin_maxit", C.c_int),
        ("h_eps", C.c_double), ("c_eps", C.c_double),
        ("film_precond", C.c_int),
        ("learn_passes", C.c_int), ("learn_iters", C.c_int), ("learn_probes", C.c_int),
        ("learn_lr", C.c_double),
        ("step_control", C.c_int),
        ("tsc_change", C.c_double), ("tsc_criterion", C.c_double), ("tsc_floor", C.c_double),
        ("tsc_keep", C.c_int),
        ("dt_min", C.c_double), ("dt_max", C.c_double),
    ]

    def copy(self):
        c = YzCzCase()
        C.pointer(c)[0] = self
        return c


def cz_case_defaults():
    """The reference's defaults for a ConcZ film (Input.cpp getin
    defaults: Fluid.MolarMass 2200 / 106, Raoult evaporation with
    coefficient 1, UTOPIA, LinearSolver MaxIterations 5000 and Epsilon
    1e-13), as yz_cz_case_defaults sets them."""
    c = YzCzCase()
    c.nx, c.ny, c.nz = 16, 16, 8
    c.xmin, c.xmax, c.ymin, c.ymax = 0.0, 1.0, 0.0, 1.0
    c.capillary = 1
    c.As, c.sig_c0, c.visc_c0, c.Df, c.diff_c0 = 1.0, 0.5, 0.5, 1.0, 0.5
    c.Ev, c.evap_coef, c.evap_func, c.evap_power = 1.0, 1.0, EVAP_RAOULT, 0.5
    c.molar_resin, c.molar_solvent = 2200.0, 106.0
    c.conc_scheme = CONV_UTOPIA
    c.lin_maxit, c.h_eps, c.c_eps = 5000, 1e-13, 1e-13
    c.film_precond = PC_AMG
    c.learn_passes, c.learn_iters, c.learn_probes, c.learn_lr = 3, 150, 4, 2e-3
    c.step_control, c.tsc_change, c.tsc_criterion, c.tsc_floor, c.tsc_keep = 0, 0.5, 0.01, 0.0, 50
    c.dt_min, c.dt_max = 1e-15, 1e30
    return c


_lib = None


def load():
    """Load the HIP library. There is no fallback: a missing library is an
    error, so nothing can silently run anywhere but on the device."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libyuzen_hip.so is not built (run __graft_entry__.build() or "
            "python -m yuzen_amd.build)")
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    d, i = C.c_double, C.c_int
    pd, pi_ = P(C.c_double), P(C.c_int)
    lib.yz_last_error.restype = C.c_char_p
    lib.yz_case_defaults.argtypes = [P(YzCase)]
    lib.yz_step_create.restype = C.c_void_p
    lib.yz_step_create.argtypes = [P(YzCase), pd, pd, pd, d]
    lib.yz_step_advance.argtypes = [C.c_void_p, pi_, pi_, pd]
    lib.yz_step_get_h.argtypes = [C.c_void_p, pd]
    lib.yz_step_velocity.argtypes = [C.c_void_p, pd, pd]
    lib.yz_write_dat.argtypes = [C.c_char_p, C.POINTER(YzCase), pd, pd, pd, pd, pd, d, d, d, C.c_int]
    lib.yz_read_dat.argtypes = [C.c_char_p, C.POINTER(YzCase), d, pd, pd, pd]
    lib.yz_step_write_dat.argtypes = [C.c_void_p, C.c_char_p, d, d, d, C.c_int]
    lib.yz_step_velocity_terms.argtypes = [C.c_void_p, pd, pd]
    lib.yz_step_field_info.argtypes = [C.c_void_p, pd]
    lib.yz_step_series_open.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(YzSeriesScales)]
    lib.yz_step_series_write.argtypes = [C.c_void_p]
    lib.yz_write_vts.argtypes = [C.c_char_p, C.POINTER(YzCase), d, pd, pd, pd, pd, pd, d, d, d, d, C.c_int]
    lib.yz_step_write_vts.argtypes = [C.c_void_p, C.c_char_p, d, d, d, d, C.c_int]
    lib.yz_step_energy.argtypes = [C.c_void_p, pd, pd]
    lib.yz_step_balances.argtypes = [C.c_void_p, pd]
    lib.yz_step_set_h.argtypes = [C.c_void_p, pd]
    lib.yz_step_clock.argtypes = [C.c_void_p, pd, pd, pd]
    lib.yz_step_set_term.argtypes = [C.c_void_p, C.POINTER(YzTerm)]
    lib.yz_step_term_status.argtypes = [C.c_void_p, C.POINTER(C.c_int), pd]
    lib.yz_step_run.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.yz_step_pass_log.argtypes = [C.c_void_p, pd, pd, pd, C.c_int, C.POINTER(C.c_longlong)]
    lib.yz_step_timing.argtypes = [C.c_void_p, pd]
    lib.yz_step_destroy.argtypes = [C.c_void_p]
    lib.yz_step_kernel_time.argtypes = [C.c_void_p, i, P(C.c_longlong), pd]
    lib.yz_step_assemble.argtypes = [C.c_void_p, pd, pd, d, pd, pd, pd]
    lib.yz_step_get_c.argtypes = [C.c_void_p, pd]
    lib.yz_step_face_fluxes.argtypes = [C.c_void_p, pd, pd, pd, pd]
    lib.yz_step_conc_stats.argtypes = [C.c_void_p, pd]
    lib.yz_step_save_state.argtypes = [C.c_void_p, C.c_char_p]
    lib.yz_step_load_state.argtypes = [C.c_void_p, C.c_char_p]
    lib.yz_linsolve_create.restype = C.c_void_p
    lib.yz_linsolve_destroy.argtypes = [C.c_void_p]
    lib.yz_linsolve.argtypes = [C.c_void_p, i, i, pi_, pi_, pd, pd, pd,
                                P(YzLinSettings), pi_, pd]
    lib.yz_linsolve_noalloc.argtypes = [i, i, pi_, pi_, pd, pd, pd, P(YzLinSettings), pi_, pd]
    lib.yz_memory_get.argtypes = [C.c_void_p, pd, i]
    lib.yz_memory_set.argtypes = [C.c_void_p, pd, i]
    pf = P(C.c_float)
    lib.yz_learn_get_params.argtypes = [C.c_void_p, pf]
    lib.yz_learn_set_params.argtypes = [C.c_void_p, pf]
    lib.yz_learn_sample.argtypes = [C.c_void_p, pf, pf, pf]
    lib.yz_learn_loss_grad.argtypes = [C.c_void_p, i, C.c_uint, pd, pf]
    lib.yz_learn_train.argtypes = [C.c_void_p, i, d, i, pd, pd]
    lib.yz_learn_apply.argtypes = [C.c_void_p, pf]
    lib.yz_learn_stats.argtypes = [C.c_void_p, pd]
    lib.yz_batch_create.restype = C.c_void_p
    lib.yz_batch_create.argtypes = [i, P(YzCase), pd, pd, pd, pd]
    lib.yz_batch_set_params.argtypes = [C.c_void_p, pf]
    lib.yz_batch_advance.argtypes = [C.c_void_p, pi_, pi_, pi_]
    lib.yz_batch_get_h.argtypes = [C.c_void_p, i, pd]
    lib.yz_batch_stats.argtypes = [C.c_void_p, pd]
    lib.yz_tune_tiles.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), pd]
    lib.yz_tile_rows.argtypes = [C.c_int, C.c_int]
    lib.yz_set_tile_rows.argtypes = [C.c_int, C.c_int, C.c_int]
    lib.yz_set_row_kernel.argtypes = [C.c_int]
    lib.yz_set_fault.argtypes = [C.c_long, C.c_long]
    lib.yz_batch_destroy.argtypes = [C.c_void_p]
    lib.yz_cz_case_defaults.argtypes = [P(YzCzCase)]
    lib.yz_cz_create.restype = C.c_void_p
    lib.yz_cz_create.argtypes = [P(YzCzCase), pd, pd, pd, d]
    lib.yz_cz_advance.argtypes = [C.c_void_p, pi_, pi_, pd]
    lib.yz_cz_get_h.argtypes = [C.c_void_p, pd]
    lib.yz_cz_get_c.argtypes = [C.c_void_p, pd]
    lib.yz_cz_volume_fluxes.argtypes = [C.c_void_p, pd, pd, pd]
    lib.yz_cz_stats.argtypes = [C.c_void_p, pd]
    lib.yz_cz_clock.argtypes = [C.c_void_p, pd]
    lib.yz_cz_learn_stats.argtypes = [C.c_void_p, pd]
    lib.yz_cz_destroy.argtypes = [C.c_void_p]
    lib.yz_cz_save_state.argtypes = [C.c_void_p, C.c_char_p]
    lib.yz_cz_load_state.argtypes = [C.c_void_p, C.c_char_p]
    lib.yz_sizeof_cz_case.restype = C.c_int
    lib.yz_sizeof_case.restype = C.c_int
    lib.yz_sizeof_lin_settings.restype = C.c_int
    _lib = lib
    return lib


def write_dat(path, case, h, s, c, ux, uy=None, L0=1.0, H0=1.0, U0=1.0, dimensioned=False):
    """the reference's ASCII field file of padded fields (host code)"""
    lib = load()
    keep = [np.ascontiguousarray(v, dtype=np.float64) for v in (h, s, c, ux, uy if uy is not None else ux)]
    ptrs = [v.ctypes.data_as(C.POINTER(C.c_double)) for v in keep]
    if lib.yz_write_dat(os.fsencode(path), C.byref(case), *ptrs, L0, H0, U0, 1 if dimensioned else 0):
        raise RuntimeError(last_error())


def write_vts(path, case, time, h, s, c, ux6, uy6=None, L0=1.0, H0=1.0, U0=1.0, T0=1.0, dimensioned=False):
    """the reference's VTK structured-grid file of padded fields (host
    code): ux6 / uy6 the six velocity fields (mixed, then the capillary,
    Marangoni, vertical-gravity, in-plane gravity and centrifugal terms'),
    (6, n) each; uy6 None in 1-D"""
    lib = load()
    keep = [np.ascontiguousarray(v, dtype=np.float64) for v in (h, s, c)]
    keep += [np.ascontiguousarray(v, dtype=np.float64).ravel() for v in (ux6, uy6 if uy6 is not None else ux6)]
    if case.dim == 2 and uy6 is None:
        raise ValueError("write_vts: a 2-D film needs uy6")
    ptrs = [v.ctypes.data_as(C.POINTER(C.c_double)) for v in keep]
    if case.dim == 1:
        ptrs[4] = None
    if lib.yz_write_vts(os.fsencode(path), C.byref(case), time, *ptrs, L0, H0, U0, T0, 1 if dimensioned else 0):
        raise RuntimeError(last_error())


def read_dat(path, case, h, s, c, H_coef=1.0):
    """an ASCII field file into copies of the padded h, s, c (host code)"""
    lib = load()
    h, s, c = (np.array(v, dtype=np.float64) for v in (h, s, c))
    if lib.yz_read_dat(os.fsencode(path), C.byref(case), H_coef, _arr_ptr(h), _arr_ptr(s), _arr_ptr(c)):
        raise RuntimeError(last_error())
    return h, s, c


def _arr_ptr(v):
    return v.ctypes.data_as(C.POINTER(C.c_double))


def last_error():
    lib = load()
    e = lib.yz_last_error()
    return e.decode() if e else ""


def tune_tiles(nx, ny, nfilms=1, iters=40):
    """the tile tuner for films of nx x ny cells, solved nfilms at a time:
    (the tile rows chosen, [us per iteration with 16 rows, with 8])"""
    lib = load()
    rows = C.c_int()
    us = (C.c_double * 2)()
    if lib.yz_tune_tiles(nx, ny, nfilms, iters, C.byref(rows), us):
        raise RuntimeError("yz_tune_tiles: " + last_error())
    return rows.value, [us[0], us[1]]


def tile_rows(nx, ny):
    return load().yz_tile_rows(nx, ny)


def set_fault(k, m=1):
    """fault injection: the film's Newton-pass solves k .. k + m - 1 from now
    break down before they iterate (k = 0: off)"""
    load().yz_set_fault(int(k), int(m))


def set_row_kernel(tiled):
    """the row kernel of 2-D films, process-wide: 1 LDS-tiled (the
    default), 0 per cell; returns the previous setting"""
    return load().yz_set_row_kernel(int(tiled))


def set_tile_rows(nx, ny, rows):
    if load().yz_set_tile_rows(nx, ny, rows):
        raise ValueError(f"tile rows {rows} not taken (8 or 16)")
