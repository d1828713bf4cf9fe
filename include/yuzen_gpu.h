/*
 * yuzen_gpu.h -- C-ABI of the MI355X implementation of Yuzen's implicit
 * film-height step (well-mixed model, 1-D "xh" and 2-D "xyh").
 *
 * Two seams are exported, both plain C: pointers, sizes, return codes.
 *
 *  (1) The linear-solve seam, one call per solve, with the contract of
 *      ILinearSolverBackend::Solve (reference yuzen/src/numerics/
 *      LinearSolverBackend.h:56-103): x is the initial guess and the
 *      output; the matrix and the right-hand side are left as given;
 *      return 0 converged, 1 iteration cap, 2 breakdown or NaN; resid is
 *      the relative residual ||b - A x|| / ||b|| of the (row-scaled)
 *      system. Nothing throws across the seam.
 *
 *  (2) The step seam, the device-resident equivalent of
 *      C_NewtonKantorovich::TimeAdvance (reference yuzen/src/numerics/
 *      NewtonKantorovich.cpp:424-486) for WellMixed_HeightEquationBase
 *      (reference yuzen/src/model/well_mixed/HeightEquationBase.h:48-1601):
 *      fields are uploaded once, every Newton pass (assembly of A and of
 *      the Jacobian, residual, Krylov solve, update) runs on the GPU, and
 *      only the Newton control scalars cross back to the host.
 *
 * Errors: every call returns an int status (0 success) or NULL; the text of
 * the last error of the calling thread is yz_last_error(). A context is
 * bound to one HIP stream and is not thread-safe.
 *
 * Field layout (both seams, and the CPU oracle): the reference's row map,
 * row = nxp * iy + ix with nxp = nx + 4 and two ghost layers per side
 * (reference docs/numerics.md:54-75, GridTraits.h WM2D_Grid::Index);
 * 1-D: row = ix, n = nx + 4.
 */
#ifndef YUZEN_GPU_H
#define YUZEN_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary types (reference yuzen/src/model/BoundaryCondition.h:40-47).
 * Symmetry and Periodic are implemented; the others are refused. */
enum { YZ_BC_SYMMETRY = 0, YZ_BC_PERIODIC = 1, YZ_BC_PINNING = 2,
       YZ_BC_FRESHINLET = 3, YZ_BC_OUTLET = 4 };

/* Preconditioner of the height equation's Krylov solve.
 * YZ_PC_AMG is what PreconditionType = "ILU" selects in the reference's
 * AMGCL backend (smoothed aggregation + SPAI0, LinearSolverBackend.cpp:
 * 662-679); YZ_PC_JACOBI is "None"/"Jacobi" (damped Jacobi).
 * No reference counterpart: YZ_PC_LEARNED is the learned 13-point
 * approximate inverse M alone; YZ_PC_LEARNED_AMG the multigrid with M as
 * its finest smoother. Both run the FP32 inner solve, with the FP64 AMG
 * solve as the fallback when the stall test trips. */
enum { YZ_PC_JACOBI = 0, YZ_PC_AMG = 1, YZ_PC_LEARNED = 2, YZ_PC_LEARNED_AMG = 3 };

/* Everything one well-mixed film-height case needs, already made
 * dimensionless (reference NonDimensionalNumbers::Build,
 * yuzen/src/model/NonDimensionalNumbers.cpp:43-141). */
/* evaporation laws (Physics.Evaporation.ConcFunc) */
enum { YZ_EVAP_RAOULT = 0, YZ_EVAP_POWERN = 1 };

typedef struct yz_case {
    int dim;                    /* 1: "xh", 2: "xyh" */
    int nx, ny;                 /* interior cells (ny = 1 in 1-D) */
    double xmin, xmax, ymin, ymax;
    int bc[4];                  /* Xmin, Xmax, Ymin, Ymax: YZ_BC_* */

    /* physics switches (Physics.*.Enable) */
    int capillary, vertical_gravity, inplane_gravity, centrifugal,
        marangoni, disjoining;
    /* groups */
    double Bo, Pgx, Pgy, Sp, As, Ma, Dp;
    double hstar, dis_n, dis_m;  /* Pi = (hstar/h)^n - (hstar/h)^m */
    double slip;                 /* Lambda = slip length / h0; 0: no slip */
    double rot_x, rot_y;         /* centre of rotation */
    double visc_c0, visc_exp;    /* mu(c) = exp(visc_exp (c - visc_c0)) */
    double sig_c0;               /* sigma(c) = 1 + Ma (c - sig_c0) */

    /* time integration and Newton (Numerics.HeightEquation) */
    double theta;                /* 0.5: Crank-Nicolson */
    int nk_max, nk_min, jac_interval;
    double nk_eps, nk_eps_rel;
    int ew, extrap;              /* YUZEN_OPT ew / extrap (both default on) */

    /* linear solve (Numerics.HeightEquation.LinearSolver) */
    int lin_maxit;
    double lin_eps;
    int row_scaling;
    int precond;                 /* YZ_PC_* */
    int krylov_fp32;             /* 0: FP64 Krylov; 1: FP32 inner solve */

    /* step control (Numerics.TimeStepControl) */
    int step_control;
    double tsc_change, tsc_criterion, tsc_floor;
    int tsc_keep;
    double dt_min, dt_max;

    /* learned preconditioner (no reference counterpart): the network is
     * trained in place on the first learn_passes Newton passes of the
     * context, learn_iters Adam steps of rate learn_lr on learn_probes
     * probe vectors each, and frozen afterwards (0 passes: use the weights
     * as loaded by yz_learn_set_params or the initial ones) */
    int learn_passes, learn_iters, learn_probes;
    double learn_lr;
    /* the FP32 inner solve's stall test: it gives up, and the pass is
     * solved again in FP64 with the AMG, when its residual has not halved
     * over stall_window iterations (0: no test) */
    int stall_window;
    /* 1: the fused FP32 iteration of the learned preconditioner is replayed
     * from a captured HIP graph (two iterations per graph); 0 (default):
     * launched kernel by kernel. Same arithmetic, same results to the bit;
     * at 1024^2 the replay measured 3 % slower than the launches on MI355X,
     * so it is for callers that capture or batch. */
    int graphs;

    /* the well-mixed concentration (Physics.ConcField, reference
     * Input.cpp:440-477): 0 (default) frozen, as Model = "None" runs it;
     * 1 solved with the film in every Newton pass and the step completed
     * with the fluxes of the step, as Model = "WM" runs it
     * (HeightEquationBase.h:1390-1580, ConcEquationBase.h:261-418) */
    int conc_model;
    int conc_convection, conc_diffusion;  /* EnableConvection, EnableDiffusion */
    int conc_scheme;             /* Numerics.Convection.Concentration: YZ_CONV_* */
    double Df;                   /* diffusion number D0 T0 / L0^2 */
    double diff_c0, diff_exp;    /* D(c) = exp(diff_exp (c - diff_c0)) */
    /* Numerics.ConcEquation.LinearSolver */
    int conc_lin_maxit;
    double conc_lin_eps;
    int conc_row_scaling;

    /* the initial guess of each Newton pass's linear solve: 0 the previous
     * pass's correction, as the reference passes _correctVec
     * (NewtonKantorovich.cpp:130); 1 zero; -1 (default) the reference's on
     * the FP64 path and zero on the FP32 inner solve (no reference
     * counterpart: each inexact solve starts afresh, which at 1024^2 takes
     * 28 % fewer iterations than starting from a correction up to 10^4
     * times larger than the one sought) */
    int zero_guess;
    /* B of the film's residual (HeightEquationBase.h:753-787): the
     * evaporation, a Crank-Nicolson mean of EvapAt at the start of the step
     * and at the Newton guess (:830-846: Ev E(h, c), at most half the film
     * and the solvent h (1 - c) per step; E from Evaporation.cpp:62-103,
     * uniform in space), less the nozzle's supply at the middle of the step
     * (SourceBase.cpp, WM2D/Source.cpp, WM1D/Source_1D.cpp). With the
     * concentration solved (conc_model 1) the concentration equation takes
     * them too: solvent-only evaporation and the supply at the local or the
     * nozzle's concentration (ConcEquationBase.h:310-318, StepSources,
     * HeightEquationBase.h:936-960). The batch takes both, per film, with c
     * frozen. */
    int evaporation;                /* Physics.Evaporation.Enable */
    double Ev, evap_coef;           /* E = Ev evap_coef f(c) g(h) */
    int evap_func;                  /* YZ_EVAP_RAOULT (default) or YZ_EVAP_POWERN */
    double evap_power, evap_concr;  /* PowerN: max(1 - c - concr, 0)^power */
    double molar_resin, molar_solvent;  /* Raoult: Fluid.MolarMass (2200, 106) */
    double precursor;               /* PrecursorRatio: the ramp (h - h*) / (19 h*) in [0, 1]; 0 off */
    int source;                     /* Physics.Source.Enable */
    double src_radius, src_rate, src_tau;   /* Radius, InletRate, TimeConstant (dimensionless) */
    double src_arm_length, src_arm_x, src_arm_y;   /* ArmLength, ArmCenter */
    double src_omega, src_angle0, src_ramp, src_plateau;   /* RotationRate, StartAngle, RampTime, PlateauTime */
    int src_subregions;             /* SubRegions */
    /* Numerics.Convection.Height (HeightEquationBase.h:609-616): the face
     * interpolation of the upwinded terms (in-plane gravity, the
     * centrifugal force, Marangoni): the limited YZ_HSCHEME_KOREN (the
     * default) or YZ_HSCHEME_VANLEER (FiniteDifferences.cpp:97-143), or one
     * of the fixed-coefficient schemes (:22-73), which each term assembles
     * by its own code (BodyForceKernel.h:263-545, SolutoCapillarityKernel.h:
     * 104-405) */
    int height_scheme;
    /* Physics.Source.ConcType (SourceBase.cpp:37): 0 "local", the supplied
     * liquid arrives at the concentration it lands on; 1 "supplied", at
     * src_conc (Physics.Source.Conc). Enters the well-mixed concentration
     * (conc_model 1) only. */
    int src_conc_supplied;
    double src_conc;
} yz_case;

/* The film height's face interpolation (the keys of FiniteDifferences.cpp
 * :214-228): the two limited schemes, then the fixed-coefficient ones */
enum { YZ_HSCHEME_KOREN = 0, YZ_HSCHEME_VANLEER = 1, YZ_HSCHEME_UPWIND1 = 2, YZ_HSCHEME_CENTRAL2 = 3,
       YZ_HSCHEME_CENTRAL4 = 4, YZ_HSCHEME_UPWIND2 = 5, YZ_HSCHEME_UTOPIA = 6, YZ_HSCHEME_QUICK = 7,
       YZ_HSCHEME_KK = 8 };
/* Face interpolation of the concentration's advection (reference
 * FiniteDifferences.cpp:22-93, keys :214-228): the unlimited schemes, then
 * the limited Koren and van Leer (their limiter frozen at the latest
 * iterate, FaceWeights :166-190; the step completed by solves repeated until
 * that settles, HeightEquationBase.h:1506-1525). UTOPIA is the reference's
 * default (Input.cpp:2302-2303). */
enum { YZ_CONV_UPWIND1 = 0, YZ_CONV_CENTRAL2 = 1, YZ_CONV_CENTRAL4 = 2, YZ_CONV_UPWIND2 = 3,
       YZ_CONV_UTOPIA = 4, YZ_CONV_QUICK = 5, YZ_CONV_KK = 6, YZ_CONV_KOREN = 7, YZ_CONV_VANLEER = 8 };

/* Fill c with the reference's defaults (Input.cpp getin defaults):
 * theta 0.5, NewtonMaxIterations 30, NewtonMinIterations 1,
 * NewtonEpsilon 1e-10, NewtonEpsilonRel 1e-6, JacobianUpdateInterval 1,
 * MaxIterations 5000, Epsilon 1e-13, RowScaling on, AMG, EW and
 * extrapolation on, step control off; learned preconditioner trained on
 * the first pass, 100 Adam steps of rate 4e-3 on 2 probes each; stall
 * window 40;
 * graphs off; concentration frozen (conc_model 0), and for conc_model 1:
 * convection and diffusion on, UTOPIA, Df 1, D(c) exponent 0 about 0.5,
 * the concentration solve's MaxIterations 5000, Epsilon 1e-13, RowScaling
 * on; zero_guess -1. */
void yz_case_defaults(yz_case *c);

const char *yz_last_error(void);
/* sizeof(yz_case) and sizeof(yz_lin_settings) as compiled, so that a
 * binding can check its mirror of the two structures */
int yz_sizeof_case(void);
int yz_sizeof_lin_settings(void);

/* ---- step seam --------------------------------------------------------
 * yz_step_create: h, s, c are padded arrays of n = (nx+4)(ny+4) rows
 * (1-D: nx+4); their ghost values are ignored and formed again from the
 * boundary conditions. dt0 is the first step (Numerics.InitialDt). */
typedef struct yz_step yz_step;
yz_step *yz_step_create(const yz_case *c, const double *h, const double *s,
                        const double *conc, double dt0);
/* One implicit step, TimeAdvance's contract: returns 0 on an accepted
 * step, 1 when the Newton iteration failed at a fixed step (the reference
 * stops the run: FailedFixedStep), 2 on a device error. nk_iter is the
 * number of Newton passes of the accepted attempt, lin_iter the Krylov
 * iterations summed over them, resid the last Newton correction. */
int yz_step_advance(yz_step *st, int *nk_iter, int *lin_iter, double *resid);
/* n doubles, ghosts included */
int yz_step_get_h(yz_step *st, double *h);
int yz_step_set_h(yz_step *st, const double *h);
/* clock: time reached, step now in force, step before it */
int yz_step_clock(yz_step *st, double *time, double *dt_now, double *dt_old);
/* the Newton passes of the last step's accepted attempt (what the
 * reference prints under YUZEN_DEBUG_RESIDUAL, NewtonKantorovich.cpp:
 * 262-270): per pass the correction the exit test took (the larger of the
 * film's and the coupled side's), the linear tolerance the pass's solve was
 * given (Eisenstat-Walker, :108-126) and the relative residual it reached;
 * up to cap of them; any array may be NULL. stag_cuts: the attempts ended by
 * a stagnation far from convergence so far (_stagnationCuts, :357).
 * Returns the number of passes, -1 on a null context. */
int yz_step_pass_log(yz_step *st, double *corr, double *eta, double *lin_resid, int cap, long long *stag_cuts);
/* The restart state (field*.state, yuzen/src/IO/RestartState.h:24-49):
 * yz_step_save_state writes what the next step depends on in the
 * reference's layout and keys (WellMixed_ModelBase::WriteState,
 * WellMixedModelBase.h:896-927: time, step, dt, adjustCount.height,
 * height.old, height.correction, solver.height, solver.conc, field.h/s/c),
 * plus keys the reference's reader ignores (conc.start; learn.* for the
 * learned preconditioner). yz_step_load_state puts a context created for the
 * same case in that state (BuildRestart, :1196-1221), so the next steps are
 * those of the run that wrote it, to the bit. 0 or 2 (yz_last_error). */
int yz_step_save_state(yz_step *st, const char *path);
int yz_step_load_state(yz_step *st, const char *path);

/* ---- the run's termination tests (WellMixed_ModelBase::CheckTermination,
 * yuzen/src/model/well_mixed/WellMixedModelBase.h:1567-1679, after every
 * accepted step; FieldInfo over the real cells, IO/FieldInfo.cpp:52-136;
 * ChangeRateOfVariable, numerics/ImplicitSolver.cpp:225-240). */
typedef struct yz_term {
    double end_time;      /* Numerics.EndTime (nondimensional, the run starts at 0); <= 0: none */
    long long max_steps;  /* Numerics.MaxSteps; 0: none */
    double steady_eps;    /* Numerics.SteadyStateCriterion; <= 0 switches the test off */
    int steady_steps;     /* Numerics.SteadyStateSteps */
    int dry;              /* Physics.Evaporation.Enable: the dry-out test is taken */
    int end_conc_type;    /* EndConcType: 0 mean, 1 max, 2 min */
    double end_conc_value;/* EndConcValue */
} yz_term;
/* why the run stops after a step, in the reference's order of the tests */
enum { YZ_RUN_GOING = 0, YZ_RUN_NAN = 1, YZ_RUN_CONC_RANGE = 2, YZ_RUN_NEGATIVE_FILM = 3,
       YZ_RUN_NEGATIVE_VOLUME = 4, YZ_RUN_END = 5, YZ_RUN_DRY = 6, YZ_RUN_STEADY = 7 };
/* The tests evaluated on the device at the end of every accepted step
 * from the call on (t NULL: off), their outcome carried back by the step's
 * own final synchronisation. The concentration's slack test ([0, 1] by
 * 0.1, with the lowest c of the step's start) applies when the
 * concentration is solved (conc_model 1). */
int yz_step_set_term(yz_step *st, const yz_term *t);
/* the outcome after the last step: status (YZ_RUN_*) and info8 =
 * {fluid volume, mean c, min h, max c, min c (with the step's start when
 * solved), change rate x span, steady count, steps} (either may be NULL;
 * no synchronisation) */
int yz_step_term_status(yz_step *st, int *status, double *info8);
/* advance up to nsteps steps, stopping after the first step whose tests
 * say so (yz_step_set_term) or that fails; *done gets the steps taken,
 * *status the tests' outcome. Returns yz_step_advance's code. */
int yz_step_run(yz_step *st, int nsteps, int *done, int *status);

/* per-phase timing of the last yz_step_advance, milliseconds (HIP events):
 * [0] assembly, [1] residual, [2] AMG setup, [3] Krylov, [4] update,
 * [5] whole step */
int yz_step_timing(yz_step *st, double *ms6);
/* launch timing of one kernel of the Krylov iteration of the case's own
 * solve: with the fused FP32 iteration of the learned preconditioner, its
 * first kernel (K1: the x and r update, p, ph = M p and (r, r)), every
 * launch that did its work (iterations 1 on), each timed on the device's
 * own constant-rate clock (wall_clock64) from the last block exit of the
 * K4 before it to its own last block exit (its duration on the busy
 * stream, dispatch included), the exits written by the kernels into slots
 * per iteration and block and reduced after the solve; otherwise the
 * fine-level stencil product (the SpMV
 * v = A p with its fused dot product) of every iteration between two HIP
 * event markers on the step's stream. Returns the launches timed and their
 * total milliseconds since the last reset; enable 1 resets and starts, 0
 * resets and stops, -1 only reads */
int yz_step_kernel_time(yz_step *st, int enable, long long *count, double *total_ms);
void yz_step_destroy(yz_step *st);

/* The film's face velocities at the context's film and concentration
 * (WellMixed_HeightEquationBase::CalVelocity, HeightEquationBase.h:
 * 1583-1601, which the reference runs at the end of each step for output):
 * the plain sum of the terms' velocities (LubFields.cpp:467-495) with the
 * equation's boundary rows (SetVelocityBC, WM2D/HeightEquation_2D.cpp:
 * 32-151, WM1D/HeightEquation_1D.cpp:28-86). Padded arrays of n doubles:
 * ux[row] on the Xm face of that row's cell (the reference's Yedge / Node of
 * the same index), uy[row] on its Ym face (Xedge); uy may be NULL in 1-D.
 * Faces neither the terms nor the boundary rows write are 0. */
int yz_step_velocity(yz_step *st, double *ux, double *uy);

/* The velocity fields the .vts writes (Field_WM1D_Out / Field_WM2D_Out,
 * LubFields.cpp:370-383, 832-846, as CalVelocity leaves them,
 * HeightEquationBase.h:1584-1601): six padded fields of n doubles each in
 * ux6 (and uy6; NULL in 1-D), in this order: the mixed velocity as
 * yz_step_velocity gives it, then the capillary, Marangoni, vertical-gravity,
 * in-plane gravity and centrifugal terms' own fields, each with the boundary
 * rows once; a term the case leaves out is 0. 0 / 2. */
int yz_step_velocity_terms(yz_step *st, double *ux6, double *uy6);

/* The energy budget of the context's film (Output.Energy: EnergyBudget.h
 * :52-166, FluxPower HeightEquationBase.h:275-291): F the free energy
 * (sigma/2 |grad(h+s)|^2 + Bo/2 (h+s)^2 + Dp w(h), summed over the real cells
 * and faces times the cell measure) and D the power of the film's face
 * volume fluxes against the potential phi = sigma lap(h+s) + Dp Pi(h) -
 * Bo (h+s), times the cell measure: with Crank-Nicolson,
 * (F^{n+1} - F^n)/dt + (D^n + D^{n+1})/2 = O(dt^2) for the capillary,
 * disjoining and vertical-gravity terms on a closed or periodic film.
 * The reference's Energy.dat writes F, D and that balance. 0 / 2. */
int yz_step_energy(yz_step *st, double *F, double *D);

/* FieldInfo of the context's film (FieldInfo.cpp:52-136; what History.dat
 * and the termination tests read): out9 = the volume and the solute (the
 * sums over the real cells of h and c h times the cell measure), the mean c
 * (solute over volume), the largest and smallest h+s, h and c of a real
 * cell. 0 / 2. */
int yz_step_field_info(yz_step *st, double *out9);

/* The scales of the run's time series (Output.Dimensioned and the scales of
 * the non-dimensional numbers, WellMixedModelBase.h:681-700, 796-806). */
typedef struct yz_series_scales {
    int dimensioned;        /* 0: every column dimensionless */
    double L0, H0, T0;      /* the length [m], height [m] and time [s] scales */
    double sigma0;          /* Fluid.SurfaceTension.BaseValue [N/m] (Energy.dat) */
    int energy;             /* Output.Energy.Enable: Energy.dat as well */
} yz_series_scales;

/* Opens the run's time series in the directory dir: History.dat and
 * CalcInfo.dat, and Energy.dat when sc->energy, each truncated and given
 * the reference's header line (OutputTimeline::Build, Output.cpp:1053-1094,
 * with the labels of WellMixedModelBase.h:520-568). From then on every step
 * of the context adds to the run's evaporated volume (B's evaporation of the
 * step), the nozzle's supplied volume and solute, the volume and solute the
 * 1-D centrifugal term flung sideways (its theta-weighted rate,
 * HeightEquationBase.h:379-383, 1457-1459) and, with Energy.dat, the energy
 * dissipated (WellMixedModelBase.h:1404-1411); the volume and solute at the
 * opening are those VolErr and ResinVolErr are taken against. Opening again
 * starts the sums anew. 0 / 2. Device reductions and a host read per step
 * while open. */
int yz_step_series_open(yz_step *st, const char *dir, const yz_series_scales *sc);
/* Appends one row to each open file at the context's time, as the
 * reference's FreqOutput writes them (WellMixedModelBase.h:768-832):
 * History.dat the volume, solute, mean c, evaporated, outflowed volume and
 * solute, VolErr, ResinVolErr, max/min h+s, max/min h, max/min c, supplied
 * volume and solute (6 significant digits); CalcInfo.dat the last step's
 * Newton passes and last correction, the last linear solve's iterations and
 * relative residual, the next step's dt and ChangeRateOfVariable
 * (ImplicitSolver.cpp:218-234); Energy.dat the free energy, the
 * dissipation, the energy dissipated since the opening and the last step's
 * balance (17 digits). 0 / 2. */
int yz_step_series_write(yz_step *st);

/* The volume balance of the context's film (History.dat's columns,
 * WellMixedModelBase.h:1339-1360): out4[0] the film's volume and [1] its
 * solute (sums of h and h c over the real cells times the cell measure),
 * [2] the volume the last step evaporated (B's Crank-Nicolson evaporation of
 * its last Newton pass, as capped, times the cell measure and that step's
 * dt: EvaporationRate, HeightEquationBase.h:359-369) and [3] the volume the
 * nozzle supplied in it (SuppliedVolume, :293-303); 0 before the context's
 * first step and after yz_step_load_state until it steps again (the state
 * file holds the film, not the last step's evaporation or supply).
 * On a closed or periodic film without the 1-D sideways outflow, the volume
 * changes over a step by [3] - [2], to the linear solve's tolerance. 0 / 2. */
int yz_step_balances(yz_step *st, double *out4);

/* The ASCII field file (.dat) of a well-mixed film, in the reference's
 * layout (OutputNeutral::Write of Field_WM1D_Out / Field_WM2D_Out,
 * Output.cpp:830-920, LubFields.cpp:204-290, 322-372, 602-900): a header
 * line, then one line per point (the wall faces and the real cell centres,
 * nx+2 (x ny+2) points, x fastest) with x (y), h+s, the velocity, c, h, s
 * at 17 digits after the point. The fields are padded arrays of n doubles
 * in the case's layout, the face velocities as yz_step_velocity gives them.
 * dimensioned != 0 writes mm, um and m/s with the scales L0 [m], H0 [m],
 * U0 [m/s]. The file is written to path.tmp and renamed. 0 / 2. Host only. */
int yz_write_dat(const char *path, const yz_case *c, const double *h, const double *s, const double *conc,
                 const double *ux, const double *uy, double L0, double H0, double U0, int dimensioned);
/* Reads such a file as an initial condition (C_Restart::ReadAscii_WM1D /
 * ReadAscii_WM2D, Restart.cpp:170-311): point k goes to the padded cell one
 * in from the grid's outer ghost layer, h and s times H_coef, c as read;
 * the other cells of h, s, conc are left as given (a step context made from
 * them forms its ghosts itself). A row must hold the header's count of
 * finite numbers (a seventh / ninth column when the header names PHI), and
 * the file as many points as the grid. 0 / 2. Host only. */
int yz_read_dat(const char *path, const yz_case *c, double H_coef, double *h, double *s, double *conc);
/* yz_write_dat of the context's film, substrate, concentration and face
 * velocities (the reference's field output at the end of a step). */
int yz_step_write_dat(yz_step *st, const char *path, double L0, double H0, double U0, int dimensioned);

/* The VTK structured-grid file (.vts) of a well-mixed film, in the
 * reference's layout (OutputVTK::Write of Field_WM1D_Out / Field_WM2D_Out,
 * Output.cpp:375-637, LubFields.cpp:370-421, 832-880): the points of the
 * .dat (nx+2 (x ny+2)) on two layers, the substrate and the free surface;
 * the header names Conc, Height (h+s), Substrate, Velocity and the terms'
 * V_Ca, V_Ma, V_Bo, V_Pg, V_Sp, and the coordinates (z = s + (h+s - s) k),
 * with the time and Model="WM2D" in a YuzenData tag; the data follow raw,
 * each block a 4-byte length and doubles. ux6 / uy6 are the six velocity
 * fields of yz_step_velocity_terms (uy6 NULL in 1-D). dimensioned != 0
 * writes the time times T0 [s], mm, um and m/s with L0, H0, U0. The driving
 * potential's block (Output.FieldData.WritePotential) is not written.
 * The file is written to path.tmp and renamed. 0 / 2. Host only. */
int yz_write_vts(const char *path, const yz_case *c, double time, const double *h, const double *s,
                 const double *conc, const double *ux6, const double *uy6, double L0, double H0, double U0,
                 double T0, int dimensioned);
/* yz_write_vts of the context's film, substrate, concentration, velocity
 * fields and time. */
int yz_step_write_vts(yz_step *st, const char *path, double L0, double H0, double U0, double T0, int dimensioned);

/* Assembly-level access (parity tests): at film h (n doubles, padded),
 * with the step's dt and the explicit half Ax_old taken at h_old:
 *   ia[13*N] (2-D) / ia[5*N] (1-D): rows of I + theta dt J over the N
 *   interior cells, slot order oM mm om pm Mo mo oo po Po mp op pp oP
 *   (1-D: M m o p P), unscaled;
 *   ax[n]: A(h) h, face by face for the capillary and vertical-gravity
 *   terms (HeightEquationBase::ApplyA), 0 on ghost rows;
 *   rhs[n]: the Newton right-hand side (NewtonKantorovich::MakeRHS),
 *   with - dt B when the case has evaporation or a source: EvapAt of
 *   h_old and of h, the supply at the context's time + dt / 2.
 * Any output pointer may be NULL. */
int yz_step_assemble(yz_step *st, const double *h_old, const double *h,
                     double dt, double *ia, double *ax, double *rhs);

/* ---- the well-mixed concentration (conc_model 1) ------------------------
 * The concentration of the state (n doubles, ghosts included): the field
 * the next step starts from (HeightEquationBase.h:1570-1577). With
 * conc_model 0 it is the frozen input. */
int yz_step_get_c(yz_step *st, double *c);
/* The face volume fluxes of the film h with concentration c (both n
 * doubles, padded), HeightEquationBase::FaceVolumeFluxes
 * (HeightEquationBase.h:852-894): F[a*n + i] is the flux through the face
 * between row i and its neighbour on the high side along axis a, the sum
 * over the enabled terms of their FaceFlux in the order the reference
 * registers them; the low face of a first interior cell is stored at its
 * ghost row; other rows 0. r[n]: the sideways outflow rate of the 1-D
 * centrifugal term (BodyForceKernel.h:153-159), 0 elsewhere. Either output
 * may be NULL. */
int yz_step_face_fluxes(yz_step *st, const double *h, const double *c, double *F, double *r);
/* counters of the concentration since the context was made: [0] solves,
 * [1] their Krylov iterations, [2] solves that needed the multigrid after
 * the Jacobi try, [3] solute moved by RedistributeExcess above 1,
 * [4] below 0 (reference _excessMoved, _deficitMoved, last step) */
int yz_step_conc_stats(yz_step *st, double *out5);

/* ---- learned preconditioner (no reference counterpart) -----------------
 * A per-cell network, 8 -> 64 -> 64 -> 13 with ReLU, gives the 13 weights
 * of row i of an approximate inverse M of the row-scaled IA, times the row
 * scale 1 / (1 + sum_{k != centre} |IA_ik|). Training minimises the probe
 * loss L = (1 / (P N)) sum_p ||v_p - IA M v_p||^2 over Rademacher probes
 * v_p; the network runs on the MFMA units in BF16 with FP32 accumulation.
 * The weights are YZ_LEARN_NPARAM floats, packed W1[64][16] (columns 8..15
 * zero), b1[64], W2[64][64], b2[64], W3[16][64] (rows 13..15 zero),
 * b3[16] (entries 13..15 zero); row o of W holds the weights of output o.
 * Features of cell i, with off = sum_{k != centre} |IA_ik| of the
 * row-scaled row: 0.25 log10(off), the four near neighbours / off (x-,
 * x+, y-, y+), the sum of the four far axial entries / off, h_i and
 * c_i - visc_c0. Valid on a context whose precond is YZ_PC_LEARNED or
 * YZ_PC_LEARNED_AMG. */
enum { YZ_LEARN_NPARAM = 6288, YZ_LEARN_NFEAT = 8, YZ_LEARN_NOUT = 13 };
int yz_learn_get_params(yz_step *st, float *params);
int yz_learn_set_params(yz_step *st, const float *params);
/* Assemble the row-scaled FP32 IA at the context's film h and step, with
 * the explicit half at h (a Newton pass's system at its starting point),
 * and keep it as the training matrix. Any output may be NULL: ia[13 N]
 * (planes, row-scaled), feat[8 N], mscale[N]. */
int yz_learn_sample(yz_step *st, float *ia, float *feat, float *mscale);
/* Probe loss and its gradient (YZ_LEARN_NPARAM floats, padding entries 0)
 * at the current weights on the sampled matrix; probes p = 0..nprobe-1 of
 * the given seed. Entry i of probe p is +1 or -1 by bit 16 of
 * mix(seed * 0x9E3779B1 ^ p * 0x85EBCA77 ^ i * 0xC2B2AE3D), mix(h) =
 * (h ^= h >> 15, h *= 0x2C1B3C6D, h ^= h >> 12, h *= 0x297A2D39,
 * h ^= h >> 15) in 32-bit arithmetic (1: +1). */
int yz_learn_loss_grad(yz_step *st, int nprobe, unsigned seed, double *loss, float *grad);
/* iters Adam steps (beta 0.9 / 0.999, eps 1e-8) of rate lr on the sampled
 * matrix; loss before the first and before the last step */
int yz_learn_train(yz_step *st, int iters, double lr, int nprobe, double *loss_first, double *loss_last);
/* M (13 N planes) for the sampled matrix at the current weights */
int yz_learn_apply(yz_step *st, float *M);
/* counters: [0] Newton passes whose FP32 solve failed and were solved
 * again in FP64, [1] Adam steps taken, [2] training milliseconds,
 * [3] probe loss of the first training step, [4] of the last */
int yz_learn_stats(yz_step *st, double *out5);

/* ---- batched sweep (no reference counterpart) ---------------------------
 * n independent 2-D films advanced one fixed implicit step per call,
 * together on one stream: every Newton round assembles all films still
 * iterating, then solves all their systems in one batched FP32 BiCGSTAB
 * with the learned preconditioner (one launch per kernel for every film),
 * then applies each film's Newton decisions. A film's numbers are the ones
 * it gets in a step context of its own with the same weights (tested).
 * Each case: dim 2, precond YZ_PC_LEARNED, step_control 0, learn_passes 0
 * (the films share one network, set with yz_batch_set_params), the same
 * lin_maxit. h, s, conc:
 * the films' padded arrays, concatenated in film order; dt0[n]. */
typedef struct yz_batch yz_batch;
yz_batch *yz_batch_create(int n, const yz_case *cases, const double *h,
                          const double *s, const double *conc,
                          const double *dt0);
int yz_batch_set_params(yz_batch *b, const float *params);
/* one step for every film: returns 0 when every film's step was accepted,
 * 1 when some film's Newton iteration failed (its status is 1), 2 on a
 * device error; status, nk_iter, lin_iter: n ints each, or NULL */
int yz_batch_advance(yz_batch *b, int *status, int *nk_iter, int *lin_iter);
int yz_batch_get_h(yz_batch *b, int film, double *h);
/* [0] milliseconds of the last yz_batch_advance (HIP events), [1] Newton
 * passes, [2] Krylov iterations, summed over the films, [3] FP64
 * fallbacks so far */
int yz_batch_stats(yz_batch *b, double *out4);
void yz_batch_destroy(yz_batch *b);

/* ---- tile tuner of the fused FP32 iteration (no reference counterpart) --
 * The learned path's Krylov kernels stage 64 x R tiles (R, the tile rows,
 * 8 or 16). yz_tune_tiles solves `nfilms` synthetic systems of nx x ny
 * cells together (one batch, as a sweep solves its films of that size) for
 * `iters` iterations with each R, three times, and makes 8 the process's
 * choice for that size when it is faster by more than 3 % (16 otherwise);
 * rows gets the choice and us2 the microseconds per iteration with R = 16
 * and R = 8. yz_tile_rows returns the choice in force (16 when none was
 * made), yz_set_tile_rows pins one (returns 2 for another R). A film's
 * result depends on its tile rows (the order of the dot products' partial
 * sums), so every context and batch of the process uses the same R for
 * films of one size. Return codes: 0, 2 device error or bad argument. */
int yz_tune_tiles(int nx, int ny, int nfilms, int iters, int *rows, double *us2);
int yz_tile_rows(int nx, int ny);
int yz_set_tile_rows(int nx, int ny, int rows);
/* The row kernel of 2-D films (process-wide): 1, the default, stages each
 * tile of 64 x 4 cells with its halo in LDS and forms every face's
 * quantities once for both of its cells; 0 forms them per cell in
 * registers. Both give the same rows to the bit. Returns the previous
 * setting; values other than 0 and 1 only read it. */
int yz_set_row_kernel(int tiled);
/* Fault injection for tests of the failure paths, the counterpart of the
 * reference's YUZEN_FAULT_AMGCL=k:m (LinearSolverBackend.cpp:628-645): the
 * film's Newton-pass linear solves number k .. k+m-1 from this call on
 * (process-wide, every context) break down, returning 2 before they
 * iterate, so the step takes its failure path (retry from h^n after an
 * extrapolated guess, a cut step under step control, yz_step_advance
 * returning 1 otherwise). k = 0 turns it off. Without a call, the
 * environment variable YZ_FAULT_SOLVE=k:m is read at the first solve.
 * Returns 0. */
int yz_set_fault(long k, long m);

/* ---- linear-solve seam -------------------------------------------------
 * Solve A x = b, A in CRS (rowptr[n+1], col[nnz], val[nnz]; any column
 * order), with BiCGSTAB under settings s. When grid_nx > 0 the rows are
 * the (grid_nx+4)(grid_ny+4) rows of a film-height matrix on that grid
 * (ghost rows included); with grid_nx 0 the grid is read off an interior
 * row's pattern (13-point film-height or 9-point concentration rows in 2-D,
 * 5-point in 1-D). The rows are folded into stencil planes on the device,
 * so AMG and Jacobi both run on the structured grid. A matrix with no such
 * grid, or whose rows do not fit it, is solved as a general CRS on the
 * device: its rows scaled by their diagonals, a product giving each row a
 * group of lanes of a wave, BiCGSTAB preconditioned by SPAI0 (for
 * YZ_PC_AMG: there is no algebraic multigrid on the device) or damped
 * Jacobi; the codes and the contract are the same. */
typedef struct yz_lin_settings {
    int precond;       /* YZ_PC_JACOBI or YZ_PC_AMG */
    int scaling;       /* row scaling by the diagonal */
    int max_iter;
    double epsilon;
    int grid_dim, grid_nx, grid_ny;   /* optional structure, 0: none */
    int bc[4];
    /* LinearSolverSettings::conc (LinearSolverBackend.h:38-51): a
     * concentration system. With YZ_PC_AMG the solve first tries damped
     * Jacobi for at most 300 iterations and, if that does not converge,
     * starts again from the guess with the multigrid, as the reference's
     * backend does for a concentration equation (LinearSolverBackend.cpp:
     * 405-470, 587-626); iter counts both. */
    int conc;
} yz_lin_settings;

typedef struct yz_linctx yz_linctx;
yz_linctx *yz_linsolve_create(void);
void yz_linsolve_destroy(yz_linctx *ctx);
int yz_linsolve(yz_linctx *ctx, int n, int nnz, const int *rowptr,
                const int *col, const double *val, const double *rhs,
                double *x_inout, const yz_lin_settings *s, int *iter,
                double *resid);
/* The same solve with no allocation. No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Every device array
 * of the solve -- the CRS copy, the planes, the Krylov vectors, the
 * multigrid hierarchy -- is carved afresh on each call from a file-scope
 * scratch array of 96 MiB that every GPU the library runs on holds until
 * the process ends; the call makes no hipMalloc or hipFree and works on the
 * current device's default stream. A system larger than the scratch
 * (about 110 000 interior cells) returns 2 with yz_last_error() saying so.
 * Calls that overlap on one GPU race on the scratch. Same arguments, codes
 * and contract as yz_linsolve, without a context. */
int yz_linsolve_noalloc(int n, int nnz, const int *rowptr, const int *col, const double *val,
                        const double *rhs, double *x_inout, const yz_lin_settings *s, int *iter,
                        double *resid);
/* solve-to-solve state for restart (ILinearSolverBackend::Memory /
 * SetMemory): the device backend carries none, so Memory is empty */
int yz_memory_get(yz_linctx *ctx, double *buf, int cap);
int yz_memory_set(yz_linctx *ctx, const double *buf, int len);


/* ------------------------------------------------------------------ */
/* The ConcZ step: a 2-D film whose concentration is resolved through   */
/* its thickness on a zeta grid that stretches with the film. One call  */
/* is one time step: CZ_HeightEquation_2D::TimeAdvance (a predictor and */
/* a midpoint stage of the semi-implicit film, CZ_HeightEquation_2D.cpp */
/* :1949-2020) and AfterTimeStepTrial (:207-386: the conservative       */
/* update, the velocity through the film, the concentration step and    */
/* the coupling passes, RedistributeExcess and RedistributeDeficit),    */
/* with CZ_ConcEquation_2D.cpp for the concentration. The film is 2-D   */
/* (Grid.Dimension "xyh"), the step fixed (no step control); every      */
/* concentration scheme, the limited ones with SettleLimiter's solves   */
/* (CZ_ConcEquation_2D.cpp:945-973); the create call refuses the rest.  */
/* ------------------------------------------------------------------ */
typedef struct yz_cz_case {
    int nx, ny, nz;                 /* real cells: Grid.Nx, Ny, Nz */
    double xmin, xmax, ymin, ymax;  /* nondimensional */
    int bc[4];                      /* YZ_BC_SYMMETRY or YZ_BC_PERIODIC */
    int capillary, vertical_gravity, inplane_gravity, centrifugal, marangoni, evaporation;
    double As, Ma, Bo, Pgx, Pgy, Sp, rot_x, rot_y;
    double sig_c0;                  /* sigma(c) = 1 + Ma (c - sig_c0) */
    double visc_c0, visc_exp;       /* mu(c) = exp(visc_exp (c - visc_c0)) */
    double Df, diff_c0, diff_exp;   /* D(c) = exp(diff_exp (c - diff_c0)) */
    double Ev, evap_coef;           /* E = Ev evap_coef f(c_s) g(h) (Evaporation.cpp:62-103) */
    int evap_func;                  /* YZ_EVAP_RAOULT (default) or YZ_EVAP_POWERN */
    double evap_power, evap_concr;  /* PowerN: max(1 - c - concr, 0)^power */
    double molar_resin, molar_solvent;  /* Raoult: Fluid.MolarMass (2200, 106) */
    double precursor;               /* Initial.SurfaceShape.PrecursorRatio (0: no ramp) */
    int conc_scheme;                /* YZ_CONV_*: Numerics.Convection.Concentration (UTOPIA) */
    int lin_maxit;                  /* LinearSolver.MaxIterations of both (5000) */
    double h_eps, c_eps;            /* LinearSolver.Epsilon of the film and the concentration (1e-13) */
    /* The film's linear solve. YZ_PC_AMG (the default): FP64 BiCGSTAB with
     * the structured SA-AMG, as the reference's backend solves it.
     * YZ_PC_LEARNED (no reference counterpart): FP32 BiCGSTAB right-
     * preconditioned by the learned 13-point approximate inverse (trained
     * in place on the first learn_passes film systems, learn_iters Adam
     * steps each, then frozen), inside FP64 iterative refinement to the same
     * h_eps on the FP64 residual; when a round fails to halve the residual,
     * FP64 AMG finishes the solve from the refined film. */
    int film_precond;
    int learn_passes, learn_iters, learn_probes;
    double learn_lr;
    /* Numerics.TimeStepControl (C_SemiImplicitEquation::TimeStepTrial,
     * SemiImplicit.cpp:100-148; TimeStepAdjustment and
     * LocalTruncationCriterion of the film, ImplicitSolver.cpp:144-190,
     * 271-307): a step whose solve fails or whose film's truncation
     * criterion is not positive is solved again at tsc_change times the
     * step (the run fails below dt_min); after tsc_keep accepted steps with
     * the criterion above 0.9 the next step is dt / tsc_change, at most
     * dt_max. Off (0) by default: a failed step returns its code. */
    int step_control;
    double tsc_change, tsc_criterion, tsc_floor;
    int tsc_keep;
    double dt_min, dt_max;
} yz_cz_case;

void yz_cz_case_defaults(yz_cz_case *c);

/* h, s: the film and the substrate with their ghost columns, (nx+4)(ny+4)
 * rows, row = (nx+4) iy + ix; conc: every cell of the 3-D grid with its
 * ghosts, (nx+4)(ny+4)(nz+4) rows, row = (nx+4)(ny+4) kz + (nx+4) iy + ix
 * (kz = 2 .. nz+1 the real levels, from the substrate up) */
typedef struct yz_cz yz_cz;
yz_cz *yz_cz_create(const yz_cz_case *c, const double *h, const double *s,
                    const double *conc, double dt);
/* one step. Returns 0; 1 the film's solve did not converge, 2 the
 * concentration's, 3 the coupling passes did not settle (the reference
 * stops a run with a fixed step on each: FailedFixedStep), or -1 on a
 * device error (yz_last_error). The state is that of the step's start
 * after a failure. lin_iter: linear iterations of the step (film and
 * concentration), passes: coupling passes taken again, resid: the last
 * film solve's relative residual. */
int yz_cz_advance(yz_cz *st, int *lin_iter, int *passes, double *resid);
/* the step control's clock: [0] the step the next advance takes, [1] the
 * last accepted step (DT().old), [2] the trials it rejected since create */
void yz_cz_clock(yz_cz *st, double *out3);
void yz_cz_get_h(yz_cz *st, double *h);
void yz_cz_get_c(yz_cz *st, double *conc);
/* the volume fluxes of the last concentration solve (BuildVolumeFluxes):
 * qx through x face ix (between ix-1 and ix) of level kz, qy through y face
 * iy, w through the zeta face below cell kz; each (nx+4)(ny+4)(nz+4) */
void yz_cz_volume_fluxes(yz_cz *st, double *qx, double *qy, double *w);
/* counters since create: [0] film solves, [1] their iterations, [2]
 * concentration solves, [3] their iterations, [4] coupling passes taken
 * again, [5] solute moved by RedistributeExcess, [6] moved and [7] made up
 * by RedistributeDeficit (c times cells, summed) */
void yz_cz_stats(yz_cz *st, double *out8);
/* the learned film solve since create: [0] refinement rounds, [1] FP32
 * iterations, [2] solves that FP64 AMG finished, [3] milliseconds of
 * training */
void yz_cz_learn_stats(yz_cz *st, double *out4);
/* the restart state in the reference's field*.state layout and keys
 * (ConcZ_Model_2D::WriteState, ConcZ_Model_2D.cpp:572-600: the clock,
 * height.old, conc.previous and conc.previousDt, solver.height,
 * solver.conc, field.h/s/c); a context created for the same film continues
 * from it to the bit. 0 or 2 (yz_last_error). */
int yz_cz_save_state(yz_cz *st, const char *path);
int yz_cz_load_state(yz_cz *st, const char *path);
void yz_cz_destroy(yz_cz *st);
/* sizeof(yz_cz_case), for bindings to check their mirror */
int yz_sizeof_cz_case(void);

#ifdef __cplusplus
}
#endif
#endif /* YUZEN_GPU_H */
