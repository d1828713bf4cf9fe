// yz_step.hip -- the device-resident implicit film-height step and the
// C-ABI of libyuzen_hip.so (include/yuzen_gpu.h).
//
// Host control mirrors the reference's Newton-Kantorovich driver
// (yuzen/src/numerics/NewtonKantorovich.cpp, paths relative to the
// reference root):
//   TimeAdvance :424-486  -- retry from h^n when the extrapolated start
//                            fails; step control or FailedFixedStep
//   NKLoop      :157-422  -- extrapolated start, the exits (absolute,
//                            relative to the first correction, stagnation
//                            with its three acceptance tests), alpha 1/2
//                            when the correction grew
//   Solve       :74-154   -- assembly, Eisenstat-Walker tolerance,
//                            linear solve, correction zeroed on failure
// and the step control of ImplicitSolver.cpp:144-307. Every vector stays
// on the device; per Newton pass the host reads back three scalars
// (max |Rhs|, the solve's outcome, max |correction|).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/yuzen_gpu.h"
#include "yz_conc.h"
#include "yz_velocity.h"
#include "yz_energy.h"
#include "yz_rows.h"
#include "yz_solver.h"
#include "yz_mlp.h"
#include "yz_state.h"

namespace yz {

static thread_local std::string g_err;
static void set_err(const std::string &s) { g_err = s; }
void set_error(const std::string &s) { g_err = s; }

#define YZ_TRY(x)                                                                           \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) {                                                             \
            set_err(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x);       \
            return 2;                                                                       \
        }                                                                                   \
    } while (0)

static constexpr int TB = 256;
static inline int nb(int n) { return (n + TB - 1) / TB; }

// ghost cells: copies of the cell they mirror or wrap to
// (BoundaryCondition.cpp ApplyEdge / ApplyCorners for Symmetry, Periodic)
__global__ void k_ghosts(Geo g, double *v) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.nxp * g.nyp) return;
    const int ix = p % g.nxp, iy = p / g.nxp;
    const bool inx = ix >= 2 && ix <= g.nxp - 3;
    const bool iny = g.dim == 1 || (iy >= 2 && iy <= g.nyp - 3);
    if (inx && iny) return;
    const int sx = fold_axis(ix - 2, g.nx, g.bc[0] == YZ_BC_PERIODIC ? 1 : 0) + 2;
    const int sy = g.dim == 2 ? fold_axis(iy - 2, g.ny, g.bc[2] == YZ_BC_PERIODIC ? 1 : 0) + 2 : 0;
    v[p] = v[sy * g.nxp + sx];
}

// h += alpha * corr on the interior cells
__global__ void k_update(Geo g, double *h, const double *corr, double alpha) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    const int ix = i % g.nx, iy = i / g.nx;
    const int p = (g.dim == 2 ? (iy + 2) : 0) * g.nxp + ix + 2;
    h[p] += corr[i] * alpha;
}

__global__ void k_absmax(int n, const double *v, unsigned long long *out) {
    double m = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = fmax(m, fabs(v[i]));
    __shared__ double sh[TB / 64];
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mx = 0;
        for (int w = 0; w < TB / 64; ++w) mx = fmax(mx, sh[w]);
        atomicMax(out, (unsigned long long)__double_as_longlong(mx));
    }
}

// extrapolated start g = h_n + r (h_n - h_n-1) over every row; flags a real
// cell that had liquid and would have none (NewtonKantorovich.cpp:178-198)
__global__ void k_extrap(Geo g, const double *hn, const double *hn1, double r, double *out,
                         unsigned long long *move_bits, int *bad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0;
    if (p < g.nxp * g.nyp) {
        const double v = hn[p] + r * (hn[p] - hn1[p]);
        out[p] = v;
        const int ix = p % g.nxp, iy = p / g.nxp;
        const bool inside = ix >= 2 && ix <= g.nxp - 3 && (g.dim == 1 || (iy >= 2 && iy <= g.nyp - 3));
        if (inside && hn[p] > 0 && !(v > 0)) atomicOr(bad, 1);
        m = fabs(v - hn[p]);
    }
    __shared__ double sh[TB / 64];
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mx = 0;
        for (int w = 0; w < TB / 64; ++w) mx = fmax(mx, sh[w]);
        atomicMax(move_bits, (unsigned long long)__double_as_longlong(mx));
    }
}

// interior cells of a padded array -> packed N
__global__ void k_pack(Geo g, const double *pad, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    const int ix = i % g.nx, iy = i / g.nx;
    out[i] = pad[(g.dim == 2 ? iy + 2 : 0) * g.nxp + ix + 2];
}

// the weight floor is ScaleFloor x max |phi_old|, the maximum's bits as
// k_absmax left them: no read-back between the two kernels
__global__ void k_truncation(Geo g, const double *hm, const double *ho, const double *hp, double dt0, double dt,
                             double floor_factor, const unsigned long long *hmaxbits, unsigned long long *out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0;
    const double wfloor = floor_factor * __longlong_as_double((long long)*hmaxbits);
    if (p < g.nxp * g.nyp) {
        double w = fabs(ho[p]);
        if (w < wfloor) w = wfloor;
        const double den = (dt0 + dt) * w;
        // Math::IsZero (MyMath_I.h:20-28): at or below the smallest normal
        e = (fabs(den) <= DBL_MIN) ? 0.0 : fabs(2.0 * (dt / dt0) * (dt0 * hp[p] + dt * hm[p] - (dt0 + dt) * ho[p]) / den);
    }
    __shared__ double sh[TB / 64];
    for (int off = 32; off > 0; off >>= 1) e = fmax(e, __shfl_down(e, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mx = 0;
        for (int w = 0; w < TB / 64; ++w) mx = fmax(mx, sh[w]);
        atomicMax(out, (unsigned long long)__double_as_longlong(mx));
    }
}

__global__ void k_f2d(int n, const float *a, double *b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}

// the nozzle's place and supply at one time, from the host
struct Nozzle {
    int on, ns;
    double xn, yn, flow, radius;
};

// BeforeNKLoop's part of B (HeightEquationBase.h:1330-1341): EvapAt of the
// step's start h^n, and the nozzle's supply at the middle of the step,
// flow x the covered fraction of the cell, found by sampling ns (x ns)
// sub-cells whose centres lie within the radius (WM2D/Source.cpp:29-126;
// WM1D/Source_1D.cpp:26-111); cells farther than the radius plus a cell's
// diagonal (1-D: plus two cells) are not sampled
__global__ void k_bstart(Geo g, BTerm bt, Nozzle nz, const double *hold, const double *c, double dt,
                         double *evold, double *src) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    const int ix = i % g.nx + 2, iy = g.dim == 2 ? i / g.nx + 2 : 0;
    const int p = iy * g.nxp + ix;
    evold[i] = evap_rate(bt, hold[p], c[p], dt);
    double frac = 0.0;
    if (nz.on) {
        const double xc = g.xlo + ix * g.dx + g.dx * 0.5;
        if (g.dim == 2) {
            const double yc = g.ylo + iy * g.dy + g.dy * 0.5;
            const double ex = xc - nz.xn, ey = yc - nz.yn;
            if (sqrt(ex * ex + ey * ey) < nz.radius + sqrt(g.dx * g.dx + g.dy * g.dy)) {
                const double sx = g.dx / (double)nz.ns, sy = g.dy / (double)nz.ns;
                const double x0 = xc - 0.5 * g.dx, y0 = yc - 0.5 * g.dy;
                double hit = 0.0;
                for (int a = 0; a < nz.ns; ++a) {
                    const double u = x0 + ((double)a + 0.5) * sx - nz.xn;
                    for (int b = 0; b < nz.ns; ++b) {
                        const double v = y0 + ((double)b + 0.5) * sy - nz.yn;
                        hit += sqrt(u * u + v * v) <= nz.radius ? 1.0 : 0.0;
                    }
                }
                frac = hit / (double)(nz.ns * nz.ns);
            }
        } else if (fabs(xc - nz.xn) < nz.radius + 2.0 * g.dx) {
            const double sx = g.dx / (double)nz.ns, x0 = xc - 0.5 * g.dx;
            double hit = 0.0;
            for (int a = 0; a < nz.ns; ++a) hit += fabs(x0 + ((double)a + 0.5) * sx - nz.xn) <= nz.radius ? 1.0 : 0.0;
            frac = hit / (double)nz.ns;
        }
    }
    src[i] = nz.flow * frac;
}

static double bits_to_double(unsigned long long b) {
    double d;
    std::memcpy(&d, &b, sizeof(d));
    return d;
}

}  // namespace yz

using namespace yz;

struct yz_step {
    yz_case cs;
    Geo g;
    RowArgs ra;
    hipStream_t st = nullptr;
    bool own_stream = true;
    double *h = nullptr, *hold = nullptr, *hold2 = nullptr, *hnew = nullptr, *s = nullptr, *c = nullptr;
    double *axold = nullptr, *b = nullptr, *corr = nullptr, *scratch = nullptr;
    // the row scale of the last pass that formed IA (JacobianUpdateInterval)
    double *dinv = nullptr;
    double *evold = nullptr, *srcv = nullptr;   // B's parts fixed over a step (N each)
    double *evpv = nullptr;                     // B's evaporation of the last pass (N)
    unsigned long long *dmax = nullptr;   // [0] rhs, [1] corr, [2] move, [3] trunc
    int *dflag = nullptr;
    unsigned long long *hmax = nullptr;   // pinned mirror
    int *hflag = nullptr;
    Solver S;
    // the FP32 inner solve and the learned preconditioner
    SolverT<float> S32;
    float *b32 = nullptr, *corr32 = nullptr, *feat = nullptr, *mscale = nullptr, *M32 = nullptr;
    __half *Yh = nullptr;   // the network's raw rows, FP16 (learned preconditioner alone)
    Mlp mlp;
    long long fallbacks = 0;
    double time = 0, dt_now = 0, dt_old = 0;
    long long steps = 0;
    int tsc_count = 0;
    bool extrap_blocked = false, first_step = true;
    // a step ran since create or load: the last step's evaporated and
    // supplied volumes (evpv, srcv) are this context's own
    bool stepped = false;
    double extrap_move = 0, ew_prev = 0, ew_eta = 0;
    double nk_first = 0, nk_prev = 0;
    int nk_pass = 0;
    long long lin_step = 0;
    // the Newton passes of the last attempt: correction (the larger of the
    // film's and the coupled side's), the solve's tolerance and the
    // relative residual it reached; the stagnation cuts of the run
    // (_stagnationCuts, NewtonKantorovich.cpp:357)
    std::vector<double> plog_corr, plog_eta, plog_lres;
    double last_eps = 0;
    long long stag_cuts = 0;
    // the run's termination tests, reduced on the device after every
    // accepted step and read back with the step's own synchronisation
    int term_on = 0;
    yz_term term{};
    double *dterm = nullptr;    // partials of k_term_cells, then the 8 reduced values
    double *hterm = nullptr;    // their host copy (pinned)
    int term_status = 0, steady_count = 0;
    double term_info[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // the run's time series (yz_step_series_open): where the files go, their
    // scales, and what every step since the opening accumulates (the
    // evaporated, supplied and sideways-outflowed volumes and solute, the
    // energy dissipated), the volume and solute at the opening
    bool ser_on = false;
    std::string ser_dir;
    yz_series_scales ser_sc{};
    double *dser = nullptr;
    double ser_evap = 0, ser_sup = 0, ser_supr = 0, ser_out = 0, ser_outr = 0, ser_vol0 = 0, ser_resin0 = 0;
    double en_F = 0, en_D = 0, en_diss = 0, en_bal = 0;
    bool en_step = false;
    int last_lin_it = 0, last_nk = 0;
    double last_lin_res = 0, last_nk_res = 0;
    hipEvent_t ev[12];
    double tms[6];
    // the well-mixed concentration (conc_model 1): the field's c, the
    // concentration equation's c0 (start of step) and cx (its solve), the
    // last pass's c, the face fluxes at the start (Fc0, rc0) and at the film
    // of the pass (Fc1, rc1), the solute fluxes, the change d (also the next
    // solve's guess) and its own FP64 solver
    bool conc = false;
    double *cfield = nullptr, *c0 = nullptr, *cx = nullptr, *cprev = nullptr, *clast = nullptr;
    // what a limited scheme freezes its limiter at: the latest iterate as the
    // solve gave it, before the cap (_cIterate, HeightEquationBase.h:423-429)
    double *citer = nullptr;
    double *Fc0[2] = {nullptr, nullptr}, *Fc1[2] = {nullptr, nullptr}, *G2[2] = {nullptr, nullptr};
    double *rc0 = nullptr, *rc1 = nullptr;
    double *ex[4] = {nullptr, nullptr, nullptr, nullptr};   // room, gain, f, e
    double *dconc = nullptr, *dsave = nullptr, *bconc = nullptr, *cscratch = nullptr, *csums = nullptr;
    // the step's starting c (padded) and the evaporation StepSources gives
    // the concentration at the film of a pass (N)
    double *cstart = nullptr, *evps = nullptr;
    int *cflags = nullptr, *hcflags = nullptr;
    unsigned long long *cdmax = nullptr, *hcdmax = nullptr;
    double *hcsums = nullptr;
    Solver Sc;
    ConcArgs ca;
    double coupled = 0;
    double cstat[5] = {0, 0, 0, 0, 0};
};

// evaporation or a source: B enters the right-hand side
static bool has_b(const yz_case &c) { return c.evaporation || c.source; }

static int check_case(const yz_case &c) {
    if (c.dim != 1 && c.dim != 2) { set_err("dim must be 1 or 2"); return 1; }
    if (c.nx < 3 || (c.dim == 2 && c.ny < 3)) { set_err("grid too small"); return 1; }
    const int nb_ = c.dim == 2 ? 4 : 2;
    for (int k = 0; k < nb_; ++k)
        if (c.bc[k] != YZ_BC_SYMMETRY && c.bc[k] != YZ_BC_PERIODIC) {
            set_err("only Symmetry and Periodic edges are implemented"); return 1;
        }
    if ((c.bc[0] == YZ_BC_PERIODIC) != (c.bc[1] == YZ_BC_PERIODIC) ||
        (c.dim == 2 && (c.bc[2] == YZ_BC_PERIODIC) != (c.bc[3] == YZ_BC_PERIODIC))) {
        set_err("a Periodic edge needs a Periodic opposite edge"); return 1;
    }
    if (c.jac_interval < 1) { set_err("JacobianUpdateInterval must be at least 1"); return 1; }
    if (c.precond < YZ_PC_JACOBI || c.precond > YZ_PC_LEARNED_AMG) { set_err("unknown preconditioner"); return 1; }
    if (c.dim != 2 && (c.precond == YZ_PC_LEARNED || c.precond == YZ_PC_LEARNED_AMG)) {
        set_err("the learned preconditioner is built for the 2-D film"); return 1;
    }
    if (c.stall_window < 0) { set_err("stall_window must be >= 0"); return 1; }
    if (c.zero_guess < -1 || c.zero_guess > 1) { set_err("zero_guess must be -1, 0 or 1"); return 1; }
    if (c.conc_model != 0 && c.conc_model != 1) { set_err("conc_model must be 0 (frozen) or 1 (well-mixed)"); return 1; }
    if (c.height_scheme < YZ_HSCHEME_KOREN || c.height_scheme > YZ_HSCHEME_KK) {
        set_err("height_scheme: not one of the YZ_HSCHEME_* schemes"); return 1;
    }
    if (c.conc_model == 1) {
        if (c.conc_scheme < YZ_CONV_UPWIND1 || c.conc_scheme > YZ_CONV_VANLEER) {
            set_err("conc_scheme: not one of the YZ_CONV_* schemes"); return 1;
        }
        if (c.conc_lin_maxit < 1 || !(c.conc_lin_eps > 0)) { set_err("conc_lin_maxit >= 1 and conc_lin_eps > 0"); return 1; }
    }
    if (has_b(c)) {
        if (c.evaporation && c.evap_func != YZ_EVAP_RAOULT && c.evap_func != YZ_EVAP_POWERN) {
            set_err("evap_func must be YZ_EVAP_RAOULT or YZ_EVAP_POWERN"); return 1;
        }
        if (c.source && c.src_subregions < 1) { set_err("src_subregions must be >= 1"); return 1; }
    }
    if (c.learn_passes > 0 && (c.learn_iters < 0 || c.learn_probes < 1 || !(c.learn_lr >= 0))) {
        set_err("learn_iters >= 0, learn_probes >= 1 and learn_lr >= 0 are needed to train"); return 1;
    }
    return 0;
}

static void build_geo(const yz_case &c, Geo &g) {
    std::memset(&g, 0, sizeof(g));
    g.dim = c.dim; g.nx = c.nx; g.ny = c.dim == 2 ? c.ny : 1;
    g.nxp = g.nx + 4; g.nyp = c.dim == 2 ? g.ny + 4 : 1;
    g.N = g.nx * g.ny;
    g.ns = c.dim == 2 ? 13 : 5; g.center = c.dim == 2 ? 6 : 2; g.ndir = c.dim == 2 ? 4 : 2;
    const double dx0 = (c.xmax - c.xmin) / (double)c.nx;
    g.xlo = c.xmin - 2 * dx0;
    g.dx = ((c.xmax + 2 * dx0) - g.xlo) / g.nxp;
    if (c.dim == 2) {
        const double dy0 = (c.ymax - c.ymin) / (double)c.ny;
        g.ylo = c.ymin - 2 * dy0;
        g.dy = ((c.ymax + 2 * dy0) - g.ylo) / g.nyp;
    } else { g.ylo = 0; g.dy = 1; }
    for (int k = 0; k < 4; ++k) g.bc[k] = c.bc[k];
}

// capillary C(d) (LaplacePressure_2D.cpp:39-94 / LaplacePressure_1D.cpp:35-77)
static void build_ccap(const Geo &g, double C[4][13]) {
    std::memset(C, 0, sizeof(double) * 52);
    if (g.dim == 2) {
        const double dx2 = g.dx * g.dx, dy2 = g.dy * g.dy, dx4 = dx2 * dx2, dy4 = dy2 * dy2, dxy = dx2 * dy2;
        C[0][S_mm] = +1.0 / dxy; C[0][S_om] = -1.0 / dxy; C[0][S_Mo] = +1.0 / dx4;
        C[0][S_mo] = 1.0 * (-3.0 / dx4 - 2.0 / dxy); C[0][S_oo] = 1.0 * (+3.0 / dx4 + 2.0 / dxy);
        C[0][S_po] = -1.0 / dx4; C[0][S_mp] = +1.0 / dxy; C[0][S_op] = -1.0 / dxy;
        C[1][S_om] = -1.0 / dxy; C[1][S_pm] = +1.0 / dxy; C[1][S_mo] = -1.0 / dx4;
        C[1][S_oo] = 1.0 * (+3.0 / dx4 + 2.0 / dxy); C[1][S_po] = 1.0 * (-3.0 / dx4 - 2.0 / dxy);
        C[1][S_Po] = +1.0 / dx4; C[1][S_op] = -1.0 / dxy; C[1][S_pp] = +1.0 / dxy;
        C[2][S_oM] = +1.0 / dy4; C[2][S_mm] = +1.0 / dxy; C[2][S_om] = 1.0 * (-3.0 / dy4 - 2.0 / dxy);
        C[2][S_pm] = +1.0 / dxy; C[2][S_mo] = -1.0 / dxy; C[2][S_oo] = 1.0 * (+3.0 / dy4 + 2.0 / dxy);
        C[2][S_po] = -1.0 / dxy; C[2][S_op] = -1.0 / dy4;
        C[3][S_om] = -1.0 / dy4; C[3][S_mo] = -1.0 / dxy; C[3][S_oo] = 1.0 * (+3.0 / dy4 + 2.0 / dxy);
        C[3][S_po] = -1.0 / dxy; C[3][S_mp] = +1.0 / dxy; C[3][S_op] = 1.0 * (-3.0 / dy4 - 2.0 / dxy);
        C[3][S_pp] = +1.0 / dxy; C[3][S_oP] = +1.0 / dy4;
    } else {
        const double dx2 = g.dx * g.dx, dx4 = dx2 * dx2;
        C[0][0] = +1.0 / dx4; C[0][1] = 1.0 * (-3.0 / dx4); C[0][2] = 1.0 * (+3.0 / dx4); C[0][3] = -1.0 / dx4;
        C[1][1] = -1.0 / dx4; C[1][2] = 1.0 * (+3.0 / dx4); C[1][3] = 1.0 * (-3.0 / dx4); C[1][4] = +1.0 / dx4;
    }
}

static StencilOp fine_op(const Geo &g) {
    StencilOp op;
    std::memset(&op, 0, sizeof(op));
    op.dim = g.dim; op.nx = g.nx; op.ny = g.ny; op.N = g.N; op.ns = g.ns;
    op.perx = g.bc[0] == YZ_BC_PERIODIC; op.pery = g.dim == 2 && g.bc[2] == YZ_BC_PERIODIC;
    for (int k = 0; k < g.ns; ++k) { op.ox[k] = slot_ox(g.dim, k); op.oy[k] = slot_oy(g.dim, k); }
    return op;
}

// the arm's angle angle0 + omega x (the time the arm has swept at full
// speed: a linear ramp up over Ts, the plateau Tm, a ramp down), the nozzle
// at the arm's end, and the flow rate (1 - exp(-t / tau)) of the full rate
// (SourceBase.cpp:54-112); the run starts at time 0
static Nozzle nozzle_at(const yz_case &c, double tt) {
    Nozzle nz;
    nz.on = c.source; nz.ns = c.src_subregions; nz.radius = c.src_radius;
    nz.xn = nz.yn = nz.flow = 0;
    if (!c.source) return nz;
    const double Ts = c.src_ramp, Tm = c.src_plateau;
    const double t = std::max(0.0, tt - 0.0);
    double swept;
    if (t <= Ts) swept = Ts > 0 ? 0.5 * t * t / Ts : 0.0;
    else if (t <= Ts + Tm) swept = 0.5 * Ts + (t - Ts);
    else if (t <= 2.0 * Ts + Tm) {
        const double u = t - Ts - Tm;
        swept = 0.5 * Ts + Tm + u - 0.5 * u * u / Ts;
    } else swept = Ts + Tm;
    const double ang = c.src_angle0 + c.src_omega * swept;
    nz.xn = c.src_arm_x + c.src_arm_length * std::cos(ang);
    nz.yn = c.src_arm_y + c.src_arm_length * std::sin(ang);
    double rise = 1.0;
    if (!(std::fabs(c.src_tau) <= std::numeric_limits<double>::epsilon())) rise = 1.0 - std::exp(-tt / c.src_tau);
    nz.flow = c.src_rate * rise;
    return nz;
}

// B's parts fixed over the step (evaporation at h^n, the supply at the
// middle of the step t + dt/2), for the step's dt: into the row arguments
static int b_before(yz_step *S, const double *hold, double dt, double t) {
    if (!S->evold) return 0;
    hipLaunchKernelGGL(k_bstart, dim3(nb(S->g.N)), dim3(TB), 0, S->st, S->g, S->ra.bt,
                       nozzle_at(S->cs, t + 0.5 * dt), hold, S->c, dt, S->evold, S->srcv);
    YZ_TRY(hipGetLastError());
    return 0;
}

// the row kernel of 2-D films: the LDS-tiled one (1, the default) or the
// one thread per cell with its faces in registers (0); the same rows
static int g_rows_tiled = 1;

static void rows_launch(yz_step *S, int mode, const double *h) {
    RowArgs a = S->ra;
    a.h = h;
    const dim3 grid(nb(S->g.N)), blk(TB);
    if (S->g.dim == 2 && g_rows_tiled) {
        const dim3 tgrid((S->g.nx + RTX - 1) / RTX, (S->g.ny + RTY - 1) / RTY);
        if (mode == 0) hipLaunchKernelGGL((k_rows_tiled<0>), tgrid, dim3(RTX * RTY), 0, S->st, a);
        else hipLaunchKernelGGL((k_rows_tiled<1>), tgrid, dim3(RTX * RTY), 0, S->st, a);
    } else if (S->g.dim == 2) {
        if (mode == 0) hipLaunchKernelGGL((k_rows<2, 0>), grid, blk, 0, S->st, a);
        else hipLaunchKernelGGL((k_rows<2, 1>), grid, blk, 0, S->st, a);
    } else {
        if (mode == 0) hipLaunchKernelGGL((k_rows<1, 0>), grid, blk, 0, S->st, a);
        else hipLaunchKernelGGL((k_rows<1, 1>), grid, blk, 0, S->st, a);
    }
}

static int sync_read(yz_step *S) {
    YZ_TRY(hipMemcpyAsync(S->hmax, S->dmax, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipMemcpyAsync(S->hflag, S->dflag, sizeof(int), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipStreamSynchronize(S->st));
    return 0;
}

static double ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0;
    return ms;
}


static bool fp32_path(const yz_case &c) {
    return c.krylov_fp32 || c.precond == YZ_PC_LEARNED || c.precond == YZ_PC_LEARNED_AMG;
}

// the FP32 inner solve cannot reach below this relative residual
static constexpr double kEpsFP32 = 1.0e-6;
// EW's cap on the FP32 inner solve (no reference counterpart; the
// reference caps at 1e-4). At 1024^2 the inexact passes of 1e-3 took 15 %
// fewer Krylov iterations per step for 1 % more Newton passes, with the
// film unchanged to 1e-12; 1e-2 gained 1 % more and cost 2 % more passes
static constexpr double kEtaMaxFP32 = 1.0e-3;

// the linear tolerance of this pass (NewtonKantorovich.cpp:108-131):
// Eisenstat-Walker from max |Rhs| (read back into hmax[0]) when on
static double ew_eps(yz_step *S) {
    const yz_case &c = S->cs;
    if (!c.ew) return c.lin_eps;
    const double kEtaMax = fp32_path(c) ? kEtaMaxFP32 : 1.0e-4;
    const double rn = bits_to_double(S->hmax[0]);
    double eta = kEtaMax;
    if (S->nk_pass > 1 && S->ew_prev > 0) {
        const double ratio = rn / S->ew_prev;
        eta = 0.9 * ratio * ratio;
        const double safe = 0.9 * S->ew_eta * S->ew_eta;
        if (safe > 0.1 && safe > eta) eta = safe;
    }
    if (eta > kEtaMax) eta = kEtaMax;
    if (eta < c.lin_eps) eta = c.lin_eps;
    S->ew_eta = eta; S->ew_prev = rn;
    return eta;
}

// the FP32 inner solve's tolerance (no reference counterpart): EW's,
// floored at kEpsFP32, and not tighter than the Newton exit can see. Once
// the previous correction prev is within a few decades of the exit
// threshold (NewtonEpsilon, or NewtonEpsilonRel x the first correction),
// the solve only needs an error below 1 % of that threshold relative to the
// correction sought, so its relative residual may be 0.01 thr / prev (at
// most EW's cap): the last pass of a step no longer runs to 1e-6 for a
// correction the exit test then finds 10^3 times below its threshold (the
// oversolving safeguard of inexact Newton methods)
static double fp32_tol(const yz_case &c, double eps, double first, double prev) {
    double t = eps > kEpsFP32 ? eps : kEpsFP32;
    if (prev > 0) {
        const double thr = std::max(c.nk_eps, c.nk_eps_rel > 0 ? c.nk_eps_rel * first : 0.0);
        t = std::max(t, std::min(kEtaMaxFP32, 0.01 * thr / prev));
    }
    return t;
}

// whether each pass's solve starts from zero rather than from the previous
// pass's correction (zero_guess -1: the reference's guess on the FP64 path,
// zero on the FP32 inner solve)
static bool zero_guess(const yz_case &c) {
    return c.zero_guess == 1 || (c.zero_guess < 0 && fp32_path(c));
}

// the pass again in FP64 with the AMG, after the FP32 solve failed or its
// stall test tripped: FP64 assembly, AMG set-up, FP64 BiCGSTAB from zero.
// Returns the solve's code; iter accumulates.
static int fp64_fallback(yz_step *S, double eps, int *iter, double *resid, int *dev_err) {
    const yz_case &c = S->cs;
    ++S->fallbacks;
    YZ_TRY(hipMemsetAsync(S->corr32, 0, S->g.N * sizeof(float), S->st));
    YZ_TRY(hipMemsetAsync(S->dmax, 0, 4 * sizeof(unsigned long long), S->st));
    rows_launch(S, 1, S->h);
    YZ_TRY(hipGetLastError());
    YZ_TRY(hipMemsetAsync(S->corr, 0, S->g.N * sizeof(double), S->st));
    int it2 = 0;
    double r2 = 0;
    YZ_TRY(hipEventRecord(S->ev[2], S->st));
    if (amg_setup(S->S)) { set_err("preconditioner set-up failed"); *dev_err = 1; return 2; }
    YZ_TRY(hipEventRecord(S->ev[3], S->st));
    const int ch = bicgstab(S->S, S->b, S->corr, YZ_PC_AMG, c.lin_maxit, eps, &it2, &r2);
    YZ_TRY(hipEventRecord(S->ev[4], S->st));
    YZ_TRY(hipEventSynchronize(S->ev[4]));
    S->tms[2] += ev_ms(S->ev[2], S->ev[3]);
    S->tms[3] += ev_ms(S->ev[3], S->ev[4]);
    *iter += it2;
    *resid = r2;
    return ch;
}

// Solve (NewtonKantorovich.cpp:74-154): assemble, EW tolerance, solve.
// With the FP32 inner solve the pass is an inexact-Newton step: the row
// kernel writes FP32 planes, the Krylov vectors are FP32, the tolerance is
// floored at kEpsFP32; the Newton residual, the assembly and the update stay
// FP64. When the FP32 solve fails or its stall test trips, the pass is
// assembled again in FP64 and solved with the FP64 AMG-BiCGSTAB.
// Fault injection for the tests of the failure paths, as the reference's
// YUZEN_FAULT_AMGCL=k:m (LinearSolverBackend.cpp:628-645): the film's
// Newton-pass solves number k .. k+m-1, counted from the setting
// (yz_set_fault, or YZ_FAULT_SOLVE=k:m read at the first solve), break down
// (return 2) before they iterate. Process-wide; k = 0 is off.
static long g_fault_at = 0, g_fault_count = 0, g_fault_seen = 0;
static bool g_fault_read = false;
static bool fault_now() {
    if (!g_fault_read) {
        g_fault_read = true;
        if (const char *e = std::getenv("YZ_FAULT_SOLVE")) {
            char *end = nullptr;
            g_fault_at = std::strtol(e, &end, 10);
            g_fault_count = (end && *end == ':') ? std::strtol(end + 1, nullptr, 10) : 1;
            g_fault_seen = 0;
        }
    }
    return g_fault_at > 0 && ++g_fault_seen >= g_fault_at && g_fault_seen < g_fault_at + g_fault_count;
}

static int nk_solve(yz_step *S, int *iter, double *resid, int *dev_err) {
    *dev_err = 0;
    const yz_case &c = S->cs;
    const bool f32 = fp32_path(c);
    S->ra.dt = S->dt_now;
    S->ra.axold = S->axold;
    S->ra.hold = S->hold;
    RowArgs keep = S->ra;
    // the modified Newton method (NewtonKantorovich.cpp:80-91): IA formed on
    // the passes nk_pass % JacobianUpdateInterval == 0 of the attempt (the
    // first always); the others keep it, with its row scale, its multigrid
    // and its learned inverse, and form only the right-hand side
    const bool jac = S->nk_pass % c.jac_interval == 0;
    S->ra.keep_ia = jac ? 0 : 1;
    S->ra.dinv = S->dinv;
    if (f32) {
        S->ra.ia32 = S->S32.A; S->ra.b32 = S->b32;
        if (c.precond == YZ_PC_LEARNED || c.precond == YZ_PC_LEARNED_AMG) { S->ra.feat = S->feat; S->ra.mscale = S->mscale; }
    }
    YZ_TRY(hipMemsetAsync(S->dmax, 0, 4 * sizeof(unsigned long long), S->st));
    YZ_TRY(hipMemsetAsync(S->dflag, 0, sizeof(int), S->st));
    if (f32) S->ra.nonunit = S->dflag;
    YZ_TRY(hipEventRecord(S->ev[0], S->st));
    rows_launch(S, 1, S->h);
    S->ra = keep;
    YZ_TRY(hipGetLastError());
    YZ_TRY(hipEventRecord(S->ev[1], S->st));
    ++S->nk_pass;
    if (sync_read(S)) { *dev_err = 1; return 2; }
    if (f32 && jac) S->S32.unit_diag = S->hflag[0] == 0;
    S->tms[0] += ev_ms(S->ev[0], S->ev[1]);
    const double eps = ew_eps(S);
    S->last_eps = eps;
    if (fault_now()) {
        // a breakdown made to happen before the solve iterates
        *iter = 0; *resid = 0;
        YZ_TRY(hipMemsetAsync(S->corr, 0, S->g.N * sizeof(double), S->st));
        return 2;
    }
    YZ_TRY(hipEventRecord(S->ev[2], S->st));
    int ch = -1;
    if (f32) {
        int rc = 0;
        if (jac && (c.precond == YZ_PC_LEARNED || c.precond == YZ_PC_LEARNED_AMG)) {
            // in-place training on the first passes' own matrices, then frozen
            if (S->mlp.trained_passes < c.learn_passes && c.learn_iters > 0) {
                YZ_TRY(hipEventRecord(S->ev[8], S->st));
                double l0 = 0, l1 = 0;
                rc = mlp_train(S->mlp, S->S32, S->feat, S->mscale, c.learn_iters, c.learn_lr, c.learn_probes,
                               0x1234u + 104729u * (unsigned)S->mlp.trained_passes, &l0, &l1, S->st);
                YZ_TRY(hipEventRecord(S->ev[9], S->st));
                YZ_TRY(hipEventSynchronize(S->ev[9]));
                S->mlp.train_ms += ev_ms(S->ev[8], S->ev[9]);
                if (S->mlp.trained_passes == 0) S->mlp.loss_first = l0;
                S->mlp.loss_last = l1;
                ++S->mlp.trained_passes;
            }
            if (c.precond == YZ_PC_LEARNED) {
                // the fused iteration reads M as FP16 rows and the FP32 row scale
                if (!rc) rc = mlp_infer_half(S->mlp, S->g.N, S->feat, S->Yh, S->st);
                S->S32.Mh = S->Yh;
                S->S32.Mscale = S->mscale;
            } else {
                if (!rc) rc = mlp_infer(S->mlp, S->g.N, S->feat, S->mscale, S->M32, S->st);
                S->S32.M = S->M32;
            }
        }
        S->S32.stall_window = c.stall_window;
        S->S32.use_graph = c.graphs;
        if (jac && !rc && (c.precond == YZ_PC_AMG || c.precond == YZ_PC_LEARNED_AMG)) rc = amg_setup(S->S32);
        if (jac && !rc && c.precond == YZ_PC_JACOBI) rc = jacobi_setup(S->S32);
        if (rc) { set_err("preconditioner set-up failed"); *dev_err = 1; return 2; }
        YZ_TRY(hipEventRecord(S->ev[3], S->st));
        const double eps32 = fp32_tol(c, eps, S->nk_first, S->nk_prev);
        if (zero_guess(c)) YZ_TRY(hipMemsetAsync(S->corr32, 0, S->g.N * sizeof(float), S->st));
        if (c.precond == YZ_PC_LEARNED)
            ch = bicgstab_stencil(S->S32, (const float *)S->b32, S->corr32, c.lin_maxit, eps32, iter, resid);
        else
            ch = bicgstab(S->S32, (const float *)S->b32, S->corr32, c.precond, c.lin_maxit, eps32, iter, resid);
        YZ_TRY(hipEventRecord(S->ev[4], S->st));
        YZ_TRY(hipEventSynchronize(S->ev[4]));
        S->tms[2] += ev_ms(S->ev[2], S->ev[3]);
        S->tms[3] += ev_ms(S->ev[3], S->ev[4]);
        if (ch == 0) {
            hipLaunchKernelGGL(k_f2d, dim3(nb(S->g.N)), dim3(TB), 0, S->st, S->g.N, (const float *)S->corr32, S->corr);
            return 0;
        }
        ch = fp64_fallback(S, eps, iter, resid, dev_err);
        if (*dev_err) return 2;
    } else {
        int rc = 0;
        if (jac) rc = c.precond == YZ_PC_AMG ? amg_setup(S->S) : jacobi_setup(S->S);
        if (rc) { set_err("preconditioner set-up failed"); *dev_err = 1; return 2; }
        YZ_TRY(hipEventRecord(S->ev[3], S->st));
        if (zero_guess(c)) YZ_TRY(hipMemsetAsync(S->corr, 0, S->g.N * sizeof(double), S->st));
        ch = bicgstab(S->S, S->b, S->corr, c.precond, c.lin_maxit, eps, iter, resid);
        YZ_TRY(hipEventRecord(S->ev[4], S->st));
        YZ_TRY(hipEventSynchronize(S->ev[4]));
        S->tms[2] += ev_ms(S->ev[2], S->ev[3]);
        S->tms[3] += ev_ms(S->ev[3], S->ev[4]);
    }
    if (ch != 0) YZ_TRY(hipMemsetAsync(S->corr, 0, S->g.N * sizeof(double), S->st));
    return ch;
}

// ---- the well-mixed concentration ---------------------------------------

// FaceVolumeFluxes of film h with the working concentration
static void conc_fluxes(yz_step *S, const double *h, double *const *F, double *r) {
    RowArgs a = S->ra;
    a.h = h;
    a.c = S->c;
    launch_face_fluxes(a, F[0], F[1], r, S->st);
}

// whether a real cell of the film S->h holds no liquid (FilmNotPositive)
static int film_not_positive(yz_step *S, bool *no_film) {
    YZ_TRY(hipMemsetAsync(S->cflags, 0, 4 * sizeof(int), S->st));
    launch_film_check(S->g, S->h, S->cflags, S->st);
    YZ_TRY(hipMemcpyAsync(S->hcflags, S->cflags, 4 * sizeof(int), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipStreamSynchronize(S->st));
    *no_film = S->hcflags[0] != 0;
    return 0;
}

// one solve of the concentration (ConcEquationBase.h:261-418): the system
// at the film of the pass (or of the step, commit), the Jacobi try of at
// most 300 iterations and the multigrid after it
// (LinearSolverBackend.cpp:405-470, 587-626), cx = c0 + d and, at the
// step's completion, c from the balance of h c. *ch: the solve's code.
static int conc_solve(yz_step *S, bool commit, int *ch) {
    const yz_case &c = S->cs;
    const Geo &g = S->g;
    ConcArgs a = S->ca;
    a.dt = S->dt_now;
    launch_conc_rows(a, S->st);
    YZ_TRY(hipGetLastError());
    YZ_TRY(hipMemcpyAsync(S->dsave, S->dconc, g.N * sizeof(double), hipMemcpyDeviceToDevice, S->st));
    if (jacobi_setup(S->Sc)) { set_err("concentration preconditioner set-up failed"); return 2; }
    int it = 0, it2 = 0;
    double r = 0;
    int rc = bicgstab(S->Sc, (const double *)S->bconc, S->dconc, YZ_PC_JACOBI, std::min(c.conc_lin_maxit, 300),
                      c.conc_lin_eps, &it, &r);
    S->cstat[0] += 1;
    if (rc != 0) {
        YZ_TRY(hipMemcpyAsync(S->dconc, S->dsave, g.N * sizeof(double), hipMemcpyDeviceToDevice, S->st));
        if (amg_setup(S->Sc)) { set_err("concentration preconditioner set-up failed"); return 2; }
        rc = bicgstab(S->Sc, (const double *)S->bconc, S->dconc, YZ_PC_AMG, c.conc_lin_maxit, c.conc_lin_eps, &it2, &r);
        it += it2;
        S->cstat[2] += 1;
    }
    S->cstat[1] += it;
    if (std::getenv("YZ_DEBUG"))
        std::fprintf(stderr, "[hip] conc solve commit %d: iterations %d (multigrid %d) resid %.3e rc %d\n", (int)commit,
                     it, it2, r, rc);
    // a failed solve starts the next one from zero, and what it left is
    // not used (the pass fails, or the step keeps its last pass)
    if (rc != 0) YZ_TRY(hipMemsetAsync(S->dconc, 0, g.N * sizeof(double), S->st));
    launch_conc_update(a, S->dconc, S->st);
    if (rc == 0 && commit && c.conc_convection) {
        launch_reconstruct(a, S->st);
        launch_ghosts(g, S->cx, S->st);
    }
    YZ_TRY(hipGetLastError());
    *ch = rc;
    return 0;
}

// ---- the run's termination tests (CheckTermination,
// WellMixedModelBase.h:1567-1679) -------------------------------------------
// k_term_cells: per block, over the real cells: the sums of h and c h (the
// fluid volume and the solute, times the cell area), min h, max c, min c,
// whether a c is not a number; over every row: max |h - h_old| and max |h|
// (ChangeRateOfVariable). k_term_final sums the blocks in their order.
// FieldInfo's loop carries a NaN of c only from its last cell (Math::Min/Max
// of a NaN and what follows keeps what follows); here any NaN stops the run.
constexpr int TERM_B = 256, TERM_G = 512, TERM_Q = 8;
constexpr int TERM_OUT = TERM_Q * TERM_G, TERM_MINSTART = TERM_OUT + TERM_Q;
__global__ void __launch_bounds__(TERM_B) k_term_cells(Geo g, const double *h, const double *hold, const double *c,
                                                       double *part) {
    double v[TERM_Q] = {0.0, 0.0, 1e300, -1e300, 1e300, 0.0, 0.0, 0.0};
    const int np = g.nxp * g.nyp;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        const double hh = h[p];
        v[5] = fmax(v[5], fabs(hh - hold[p]));
        v[6] = fmax(v[6], fabs(hh));
        const int ix = p % g.nxp, iy = p / g.nxp;
        const bool real = ix >= 2 && ix < g.nxp - 2 && (g.dim == 1 || (iy >= 2 && iy < g.nyp - 2));
        if (!real) continue;
        const double cc = c[p];
        v[0] += hh;
        v[1] += cc * hh;
        v[2] = fmin(v[2], hh);
        v[3] = fmax(v[3], cc);
        v[4] = fmin(v[4], cc);
        if (!(cc == cc) || !(hh == hh)) v[7] = 1.0;
    }
    __shared__ double sv[TERM_Q][TERM_B];
#pragma unroll
    for (int k = 0; k < TERM_Q; ++k) sv[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = TERM_B / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            sv[0][threadIdx.x] += sv[0][threadIdx.x + off];
            sv[1][threadIdx.x] += sv[1][threadIdx.x + off];
            sv[2][threadIdx.x] = fmin(sv[2][threadIdx.x], sv[2][threadIdx.x + off]);
            sv[3][threadIdx.x] = fmax(sv[3][threadIdx.x], sv[3][threadIdx.x + off]);
            sv[4][threadIdx.x] = fmin(sv[4][threadIdx.x], sv[4][threadIdx.x + off]);
            sv[5][threadIdx.x] = fmax(sv[5][threadIdx.x], sv[5][threadIdx.x + off]);
            sv[6][threadIdx.x] = fmax(sv[6][threadIdx.x], sv[6][threadIdx.x + off]);
            sv[7][threadIdx.x] = fmax(sv[7][threadIdx.x], sv[7][threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < TERM_Q; ++k) part[k * TERM_G + blockIdx.x] = sv[k][0];
}
__global__ void __launch_bounds__(TERM_B) k_term_final(int nblocks, const double *part, double *out) {
    __shared__ double sv[TERM_Q][TERM_B];
    double v[TERM_Q] = {0.0, 0.0, 1e300, -1e300, 1e300, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblocks; b += TERM_B) {
        v[0] += part[0 * TERM_G + b];
        v[1] += part[1 * TERM_G + b];
        v[2] = fmin(v[2], part[2 * TERM_G + b]);
        v[3] = fmax(v[3], part[3 * TERM_G + b]);
        v[4] = fmin(v[4], part[4 * TERM_G + b]);
        v[5] = fmax(v[5], part[5 * TERM_G + b]);
        v[6] = fmax(v[6], part[6 * TERM_G + b]);
        v[7] = fmax(v[7], part[7 * TERM_G + b]);
    }
#pragma unroll
    for (int k = 0; k < TERM_Q; ++k) sv[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = TERM_B / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            sv[0][threadIdx.x] += sv[0][threadIdx.x + off];
            sv[1][threadIdx.x] += sv[1][threadIdx.x + off];
            sv[2][threadIdx.x] = fmin(sv[2][threadIdx.x], sv[2][threadIdx.x + off]);
            sv[3][threadIdx.x] = fmax(sv[3][threadIdx.x], sv[3][threadIdx.x + off]);
            sv[4][threadIdx.x] = fmin(sv[4][threadIdx.x], sv[4][threadIdx.x + off]);
            sv[5][threadIdx.x] = fmax(sv[5][threadIdx.x], sv[5][threadIdx.x + off]);
            sv[6][threadIdx.x] = fmax(sv[6][threadIdx.x], sv[6][threadIdx.x + off]);
            sv[7][threadIdx.x] = fmax(sv[7][threadIdx.x], sv[7][threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < TERM_Q; ++k) out[k] = sv[k][0];
}
// the lowest c of a real cell, from 0 (MinStartConc: a NaN stays)
__global__ void __launch_bounds__(1024) k_min_start(Geo g, const double *c, double *out) {
    __shared__ double sm[1024];
    double m = 0.0;
    for (int i = threadIdx.x; i < g.N; i += blockDim.x) {
        const int p = (g.dim == 2 ? i / g.nx + 2 : 0) * g.nxp + i % g.nx + 2;
        const double v = c[p];
        if (!(v == v) || !(m == m)) m = (m == m) ? v : m;
        else if (v < m) m = v;
    }
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            const double a = sm[threadIdx.x], b = sm[threadIdx.x + off];
            sm[threadIdx.x] = (!(a == a)) ? a : ((!(b == b)) ? b : fmin(a, b));
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sm[0];
}
static int term_launch(yz_step *S) {
    const Geo &g = S->g;
    const int np = g.nxp * g.nyp;
    const int nbk = std::min(TERM_G, (np + TERM_B - 1) / TERM_B);
    hipLaunchKernelGGL(k_term_cells, dim3(nbk), dim3(TERM_B), 0, S->st, g, S->hnew, S->hold,
                       S->conc ? S->cfield : S->c, S->dterm);
    hipLaunchKernelGGL(k_term_final, dim3(1), dim3(TERM_B), 0, S->st, nbk, S->dterm, S->dterm + TERM_OUT);
    YZ_TRY(hipGetLastError());
    YZ_TRY(hipMemcpyAsync(S->hterm, S->dterm + TERM_OUT, (TERM_Q + 1) * sizeof(double), hipMemcpyDeviceToHost, S->st));
    return 0;
}
// the tests, in the reference's order, on the reduced values of the step
// just taken (the clock and the step count already advanced)
static void term_decide(yz_step *S) {
    const yz_term &t = S->term;
    const Geo &g = S->g;
    const double *r = S->hterm;
    const double dA = g.dim == 2 ? g.dx * g.dy : g.dx;
    double chr = 0;
    if (t.steady_eps > 0) {
        double tspan = t.end_time - 0.0;
        if (!(tspan > 0)) tspan = 1.0;
        double rate;
        if (S->steps < 5 || fabs(S->dt_old) <= DBL_MIN) rate = 1;
        else if (fabs(r[6]) <= DBL_MIN) rate = 0;
        else rate = r[5] / (r[6] * S->dt_old);
        chr = rate * tspan;
        if (chr < t.steady_eps) S->steady_count++;
        else S->steady_count = 0;
    }
    const double vol = r[0] * dA, ave = (r[1] * dA) / vol, minH = r[2], maxC = r[3];
    double minC = r[4];
    if (S->conc) minC = std::min(minC, r[TERM_Q]);
    if (r[7] != 0.0 || (S->conc && !(r[TERM_Q] == r[TERM_Q]))) minC = std::nan("");
    double *info = S->term_info;
    info[0] = vol; info[1] = ave; info[2] = minH; info[3] = maxC; info[4] = minC;
    info[5] = chr; info[6] = S->steady_count; info[7] = (double)S->steps;
    int st = YZ_RUN_GOING;
    const bool running = (t.end_time <= 0 || (0.0 <= S->time && S->time <= t.end_time)) &&
                         (t.max_steps <= 0 || S->steps <= t.max_steps - 1);
    if (!std::isfinite(vol) || !std::isfinite(maxC) || !std::isfinite(minC) || !std::isfinite(minH)) st = YZ_RUN_NAN;
    else if (S->conc && (maxC > 1.0 + 0.1 || minC < -0.1)) st = YZ_RUN_CONC_RANGE;
    else if (minH < -1.0e-3) st = YZ_RUN_NEGATIVE_FILM;
    else if (vol < 0.0) st = YZ_RUN_NEGATIVE_VOLUME;
    else if (!running) st = YZ_RUN_END;
    else if (t.dry && (t.end_conc_type == 0 ? ave : (t.end_conc_type == 1 ? maxC : minC)) > t.end_conc_value)
        st = YZ_RUN_DRY;
    else if (t.steady_eps > 0 && S->steady_count > t.steady_steps) st = YZ_RUN_STEADY;
    S->term_status = st;
}

// BeforeNKLoop's concentration part (HeightEquationBase.h:1314-1388)
static int conc_before(yz_step *S) {
    const Geo &g = S->g;
    const size_t np = (size_t)g.nxp * g.nyp;
    // the lowest c of a real cell at the step's start (MinStartConc), for the
    // termination tests
    if (S->term_on) hipLaunchKernelGGL(k_min_start, dim3(1), dim3(1024), 0, S->st, g, S->cfield, S->dterm + TERM_MINSTART);
    YZ_TRY(hipMemcpyAsync(S->c, S->cfield, np * sizeof(double), hipMemcpyDeviceToDevice, S->st));
    YZ_TRY(hipMemcpyAsync(S->cprev, S->cfield, np * sizeof(double), hipMemcpyDeviceToDevice, S->st));
    YZ_TRY(hipMemcpyAsync(S->cstart, S->cfield, np * sizeof(double), hipMemcpyDeviceToDevice, S->st));
    YZ_TRY(hipMemcpyAsync(S->citer, S->cfield, np * sizeof(double), hipMemcpyDeviceToDevice, S->st));
    conc_fluxes(S, S->hold, S->Fc0, S->rc0);
    YZ_TRY(hipMemsetAsync(S->dconc, 0, g.N * sizeof(double), S->st));
    YZ_TRY(hipGetLastError());
    S->coupled = 0;
    return 0;
}

// EveryNKLoop (HeightEquationBase.h:1391-1444): the concentration with the
// fluxes of this pass's film, the working c from it (capped at 1), and how
// far it moved. *ok false: the concentration solve failed (the pass fails).
static int conc_every(yz_step *S, bool *ok) {
    *ok = true;
    bool no_film = false;
    if (film_not_positive(S, &no_film)) return 2;
    if (no_film) { S->coupled = 0; return 0; }
    conc_fluxes(S, S->h, S->Fc1, S->rc1);
    // StepSources (HeightEquationBase.h:936-960): the evaporation at this
    // pass's film with the pass's c
    if (S->cs.evaporation) {
        RowArgs ra = S->ra;
        ra.dt = S->dt_now;
        launch_step_evp(ra, S->h, S->c, S->evps, S->st);
        S->ca.evp = S->evps;
    } else {
        S->ca.evp = nullptr;
    }
    int ch = 0;
    if (conc_solve(S, false, &ch)) return 2;
    if (ch != 0) { *ok = false; return 0; }
    // the limiter of the next pass from the iterate as solved
    const size_t np = (size_t)S->g.nxp * S->g.nyp;
    YZ_TRY(hipMemcpyAsync(S->citer, S->cx, np * sizeof(double), hipMemcpyDeviceToDevice, S->st));
    YZ_TRY(hipMemsetAsync(S->cdmax, 0, sizeof(unsigned long long), S->st));
    launch_conc_take(S->g, S->cx, S->c, S->cprev, S->cdmax, S->st);
    YZ_TRY(hipMemcpyAsync(S->hcdmax, S->cdmax, sizeof(unsigned long long), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipStreamSynchronize(S->st));
    S->coupled = bits_to_double(*S->hcdmax);
    return 0;
}

// RedistributeExcess (HeightEquationBase.h:1128-1273) on cx, with the film
// of the step's end
static int conc_redistribute(yz_step *S) {
    const Geo &g = S->g;
    ExcessArgs e;
    e.g = g; e.h = S->h; e.c = S->cx;
    e.room = S->ex[0]; e.gain = S->ex[1]; e.f = S->ex[2]; e.e = S->ex[3];
    e.part = S->csums; e.sums = S->csums + 256; e.flags = S->cflags; e.lower = 0;
    S->cstat[3] = S->cstat[4] = 0;
    YZ_TRY(hipMemsetAsync(S->cflags, 0, 4 * sizeof(int), S->st));
    launch_excess_flags(e, S->st);
    YZ_TRY(hipMemcpyAsync(S->hcflags, S->cflags, 4 * sizeof(int), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipStreamSynchronize(S->st));
    const bool over = S->hcflags[0] != 0, under = S->hcflags[1] != 0;
    if (!over && !under) return 0;
    for (int lower = 0; lower < 2; ++lower) {
        if ((lower == 0 && !over) || (lower == 1 && !under)) continue;
        e.lower = lower;
        bool any = true;
        for (int pass = 0; pass < 4 && any; ++pass) {
            YZ_TRY(hipMemsetAsync(S->cflags + 2, 0, sizeof(int), S->st));
            launch_excess_pass(e, pass, S->st);
            YZ_TRY(hipGetLastError());
            YZ_TRY(hipMemcpyAsync(S->hcflags, S->cflags, 4 * sizeof(int), hipMemcpyDeviceToHost, S->st));
            YZ_TRY(hipMemcpyAsync(S->hcsums, S->csums + 256, 4 * sizeof(double), hipMemcpyDeviceToHost, S->st));
            YZ_TRY(hipStreamSynchronize(S->st));
            if (pass == 0) S->cstat[3 + lower] = S->hcsums[0];
            any = S->hcflags[2] != 0;
        }
        if (any) launch_excess_spread(e, S->st);
    }
    launch_excess_caps(e, under ? 1 : 0, S->st);
    YZ_TRY(hipGetLastError());
    return 0;
}

// AfterNKLoop's concentration part (HeightEquationBase.h:1447-1577): the
// face fluxes at the film the iteration ended on, made those of the step
// the film took (ConsistentFluxes), the concentration solved once more with
// them and formed from the balance of h c, the excess moved, the step kept
static int conc_after(yz_step *S) {
    const Geo &g = S->g;
    const size_t np = (size_t)g.nxp * g.nyp;
    conc_fluxes(S, S->h, S->Fc1, S->rc1);
    // the evaporation as B took it in the last pass, the one the film's step
    // holds (HeightEquationBase.h:1497-1500)
    S->ca.evp = S->cs.evaporation ? S->evpv : nullptr;
    ConcArgs a = S->ca;
    a.dt = S->dt_now;
    launch_consistent_fluxes(a, S->Fc1[0], S->Fc1[1], S->cscratch, S->st);
    YZ_TRY(hipGetLastError());
    bool no_film = false;
    if (film_not_positive(S, &no_film)) return 2;
    YZ_TRY(hipMemcpyAsync(S->clast, S->cx, np * sizeof(double), hipMemcpyDeviceToDevice, S->st));
    bool solved = true;
    // a limited scheme freezes its limiter at the last iterate: the
    // completing solve is repeated, frozen at its own answer, until that
    // settles (HeightEquationBase.h:1506-1525)
    const bool limited = S->cs.conc_convection && S->ca.lim;
    for (int k = 0; k < (limited ? 8 : 1) && !no_film; ++k) {
        int ch = 0;
        if (conc_solve(S, true, &ch)) return 2;
        if (ch != 0) { solved = false; break; }
        if (!limited) break;
        YZ_TRY(hipMemsetAsync(S->cdmax, 0, sizeof(unsigned long long), S->st));
        launch_conc_settle(g, S->cx, S->citer, S->cdmax, S->st);
        YZ_TRY(hipMemcpyAsync(S->hcdmax, S->cdmax, sizeof(unsigned long long), hipMemcpyDeviceToHost, S->st));
        YZ_TRY(hipStreamSynchronize(S->st));
        if (bits_to_double(*S->hcdmax) <= 1e-12) break;
    }
    if (!solved) YZ_TRY(hipMemcpyAsync(S->cx, S->clast, np * sizeof(double), hipMemcpyDeviceToDevice, S->st));
    if (conc_redistribute(S)) return 2;
    launch_conc_commit(g, S->cx, S->c0, S->c, S->cfield, S->st);
    YZ_TRY(hipGetLastError());
    return 0;
}

// NKLoop (NewtonKantorovich.cpp:157-422)
static int nk_loop(yz_step *S, int *iter_out, double *resid_out, bool *done_out) {
    const yz_case &c = S->cs;
    const Geo &g = S->g;
    const int np = g.nxp * g.nyp;
    S->nk_pass = 0;
    S->lin_step = 0;
    S->extrap_move = 0;
    S->plog_corr.clear(); S->plog_eta.clear(); S->plog_lres.clear();
    bool guessed = false;
    if (c.extrap && !S->extrap_blocked && !S->first_step) {
        const double r = (S->dt_old > 0) ? S->dt_now / S->dt_old : 0.0;
        if (r > 0) {
            YZ_TRY(hipMemsetAsync(S->dmax, 0, 4 * sizeof(unsigned long long), S->st));
            YZ_TRY(hipMemsetAsync(S->dflag, 0, sizeof(int), S->st));
            hipLaunchKernelGGL(k_extrap, dim3(nb(np)), dim3(TB), 0, S->st, g, S->hold, S->hold2, r, S->h,
                               S->dmax + 2, S->dflag);
            if (sync_read(S)) return 2;
            const double move = bits_to_double(S->hmax[2]);
            if (!S->hflag[0] && move > 0) { S->extrap_move = move; guessed = true; }
        }
    }
    if (!guessed) YZ_TRY(hipMemcpyAsync(S->h, S->hold, np * sizeof(double), hipMemcpyDeviceToDevice, S->st));
    // BeforeNKLoop + the explicit half A(h^n) h^n
    if (S->conc && conc_before(S)) return 2;
    if (b_before(S, S->hold, S->dt_now, S->time)) return 2;
    {
        RowArgs a = S->ra;
        a.ax = S->axold;
        RowArgs keep = S->ra;
        S->ra = a;
        rows_launch(S, 0, S->hold);
        S->ra = keep;
        YZ_TRY(hipGetLastError());
    }
    double resid_eq_old = 1e40, resid_first = 0, resid_best = 1e40, resid_eq = 0;
    S->nk_first = S->nk_prev = 0;
    int stall = 0;
    const int stall_limit = 5;
    bool done = false;
    int outer;
    for (outer = 0; outer < c.nk_max; ++outer) {
        if (S->conc) {
            bool ok = true;
            if (conc_every(S, &ok)) return 2;
            if (!ok) { done = false; break; }
        }
        int it = 0, dev_err = 0;
        double lr = 0;
        const int ch = nk_solve(S, &it, &lr, &dev_err);
        if (dev_err) return 2;
        S->lin_step += it;
        YZ_TRY(hipMemsetAsync(S->dmax + 1, 0, sizeof(unsigned long long), S->st));
        hipLaunchKernelGGL(k_absmax, dim3(std::min(nb(g.N), 1024)), dim3(TB), 0, S->st, g.N, S->corr, S->dmax + 1);
        if (sync_read(S)) return 2;
        resid_eq = bits_to_double(S->hmax[1]);
        if (S->conc && S->coupled > resid_eq) resid_eq = S->coupled;
        S->plog_corr.push_back(resid_eq); S->plog_eta.push_back(S->last_eps); S->plog_lres.push_back(lr);
        S->last_lin_it = it; S->last_lin_res = lr;
        if (std::getenv("YZ_DEBUG"))
            std::fprintf(stderr, "[hip] pass %d lin %d resid_lin %.3e corr %.3e rc %d\n", outer, it, lr, resid_eq, ch);
        if (ch != 0) { done = false; break; }
        if (outer == 0) resid_first = S->nk_first = resid_eq;
        S->nk_prev = resid_eq;   // the FP32 inner solve's tolerance reads both
        if (resid_eq < resid_best * 0.9) { resid_best = resid_eq; stall = 0; }
        else stall++;
        if (stall > stall_limit) {
            // floor 1000 eps |h|: the guess's max over every row
            YZ_TRY(hipMemsetAsync(S->dmax + 3, 0, sizeof(unsigned long long), S->st));
            hipLaunchKernelGGL(k_absmax, dim3(std::min(nb(np), 1024)), dim3(TB), 0, S->st, np, S->h, S->dmax + 3);
            if (sync_read(S)) return 2;
            const double floorPhi = 1000.0 * std::numeric_limits<double>::epsilon() * bits_to_double(S->hmax[3]);
            done = (resid_eq < c.nk_eps) || (resid_first > 0 && resid_eq < resid_first * 1.0e-4) || (resid_eq <= floorPhi);
            if (!done) S->stag_cuts++;
            break;
        }
        const double alpha = (resid_eq > resid_eq_old) ? 0.5 : 1.0;
        YZ_TRY(hipEventRecord(S->ev[5], S->st));
        hipLaunchKernelGGL(k_update, dim3(nb(g.N)), dim3(TB), 0, S->st, g, S->h, S->corr, alpha);
        hipLaunchKernelGGL(k_ghosts, dim3(nb(np)), dim3(TB), 0, S->st, g, S->h);
        YZ_TRY(hipEventRecord(S->ev[6], S->st));
        if (outer >= c.nk_min - 1) {
            if (resid_eq < c.nk_eps) { done = true; break; }
            if (c.nk_eps_rel > 0 && resid_first > 0 && resid_eq < resid_first * c.nk_eps_rel) { done = true; break; }
        }
        resid_eq_old = resid_eq;
    }
    *iter_out = outer + 1;
    *resid_out = resid_eq;
    *done_out = done;
    return 0;
}

static yz_step *step_new(const yz_case *c, const double *h, const double *s, const double *conc, double dt0,
                         hipStream_t shared);

extern "C" {

void yz_case_defaults(yz_case *c) {
    std::memset(c, 0, sizeof(*c));
    c->dim = 2; c->nx = c->ny = 16;
    c->xmin = 0; c->xmax = 1; c->ymin = 0; c->ymax = 1;
    c->capillary = 1;
    c->dis_n = 3; c->dis_m = 2; c->hstar = 0.01;
    c->visc_c0 = 0.5; c->sig_c0 = 0.5;
    c->As = 1;
    c->theta = 0.5;
    c->nk_max = 30; c->nk_min = 1; c->jac_interval = 1;
    c->nk_eps = 1e-10; c->nk_eps_rel = 1e-6;
    c->ew = 1; c->extrap = 1;
    c->lin_maxit = 5000; c->lin_eps = 1e-13; c->row_scaling = 1; c->precond = YZ_PC_AMG;
    c->tsc_change = 0.5; c->tsc_criterion = 0.01; c->tsc_floor = 0; c->tsc_keep = 50;
    c->dt_min = 1e-15; c->dt_max = 1e30;
    c->learn_passes = 1; c->learn_iters = 100; c->learn_probes = 2; c->learn_lr = 4e-3;
    c->stall_window = 40;
    c->graphs = 0;
    c->conc_model = 0;
    c->conc_convection = 1; c->conc_diffusion = 1;
    c->conc_scheme = YZ_CONV_UTOPIA;
    c->Df = 1.0; c->diff_c0 = 0.5; c->diff_exp = 0.0;
    c->conc_lin_maxit = 5000; c->conc_lin_eps = 1e-13; c->conc_row_scaling = 1;
    c->zero_guess = -1;
    c->evaporation = 0; c->Ev = 1; c->evap_coef = 1; c->evap_func = YZ_EVAP_RAOULT; c->evap_power = 0.5;
    c->evap_concr = 0; c->molar_resin = 2200; c->molar_solvent = 106; c->precursor = 0;
    c->source = 0; c->src_radius = 1; c->src_rate = 0; c->src_tau = 0;
    c->src_arm_length = 1; c->src_arm_x = 0; c->src_arm_y = 0;
    c->src_omega = 0; c->src_angle0 = 0; c->src_ramp = 0; c->src_plateau = 1e10; c->src_subregions = 1;
}

const char *yz_last_error(void) { return g_err.c_str(); }
int yz_sizeof_case(void) { return (int)sizeof(yz_case); }
int yz_sizeof_lin_settings(void) { return (int)sizeof(yz_lin_settings); }

yz_step *yz_step_create(const yz_case *c, const double *h, const double *s, const double *conc, double dt0) {
    return step_new(c, h, s, conc, dt0, nullptr);
}

}  // extern "C"

// a step context; with a shared stream it is a member of a batch: it runs
// on the batch's stream and has no network of its own (the batch holds the
// shared one)
static yz_step *step_new(const yz_case *c, const double *h, const double *s, const double *conc, double dt0,
                         hipStream_t shared) {
    if (!c || !h || !s || !conc) { set_err("null argument"); return nullptr; }
    if (check_case(*c)) return nullptr;
    yz_step *S = new yz_step();
    S->cs = *c;
    build_geo(*c, S->g);
    const Geo &g = S->g;
    const int np = g.nxp * g.nyp;
    auto fail = [&](const char *what) { set_err(what); yz_step_destroy(S); return (yz_step *)nullptr; };
    if (shared) { S->st = shared; S->own_stream = false; }
    else if (hipStreamCreateWithFlags(&S->st, hipStreamNonBlocking) != hipSuccess) return fail("stream");
    double **pads[] = {&S->h, &S->hold, &S->hold2, &S->hnew, &S->s, &S->c};
    for (double **p : pads) if (hipMalloc((void **)p, np * sizeof(double)) != hipSuccess) return fail("hipMalloc");
    double **vecs[] = {&S->axold, &S->b, &S->corr, &S->scratch, &S->dinv};
    for (double **p : vecs) if (hipMalloc((void **)p, g.N * sizeof(double)) != hipSuccess) return fail("hipMalloc");
    if (has_b(*c))
        for (double **p : {&S->evold, &S->srcv, &S->evpv}) {
            if (hipMalloc((void **)p, g.N * sizeof(double)) != hipSuccess) return fail("hipMalloc");
            if (hipMemset(*p, 0, g.N * sizeof(double)) != hipSuccess) return fail("memset");
        }
    if (hipMalloc((void **)&S->dmax, 4 * sizeof(unsigned long long)) != hipSuccess) return fail("hipMalloc");
    if (hipMalloc((void **)&S->dflag, sizeof(int)) != hipSuccess) return fail("hipMalloc");
    if (hipHostMalloc((void **)&S->hmax, 4 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) return fail("hipHostMalloc");
    if (hipHostMalloc((void **)&S->hflag, sizeof(int), hipHostMallocDefault) != hipSuccess) return fail("hipHostMalloc");
    for (auto &e : S->ev) if (hipEventCreate(&e) != hipSuccess) return fail("event");
    if (hipMemcpy(S->hnew, h, np * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return fail("copy");
    if (hipMemcpy(S->s, s, np * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return fail("copy");
    if (hipMemcpy(S->c, conc, np * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return fail("copy");
    for (double *p : {S->hnew, S->s, S->c}) hipLaunchKernelGGL(k_ghosts, dim3(nb(np)), dim3(TB), 0, S->st, g, p);
    if (hipMemcpyAsync(S->hold, S->hnew, np * sizeof(double), hipMemcpyDeviceToDevice, S->st) != hipSuccess ||
        hipMemcpyAsync(S->hold2, S->hnew, np * sizeof(double), hipMemcpyDeviceToDevice, S->st) != hipSuccess ||
        hipMemcpyAsync(S->h, S->hnew, np * sizeof(double), hipMemcpyDeviceToDevice, S->st) != hipSuccess ||
        hipMemsetAsync(S->corr, 0, g.N * sizeof(double), S->st) != hipSuccess)
        return fail("init");
    if (hipStreamSynchronize(S->st) != hipSuccess) return fail("init");
    if (solver_init(S->S, fine_op(g), S->st)) return fail("solver allocation");
    if (fp32_path(*c)) {
        if (solver_init(S->S32, fine_op(g), S->st)) return fail("solver allocation");
        for (float **p : {&S->b32, &S->corr32, &S->mscale})
            if (hipMalloc((void **)p, g.N * sizeof(float)) != hipSuccess) return fail("hipMalloc");
        if (hipMemset(S->corr32, 0, g.N * sizeof(float)) != hipSuccess) return fail("memset");
        if (c->precond == YZ_PC_LEARNED || c->precond == YZ_PC_LEARNED_AMG) {
            if (hipMalloc((void **)&S->feat, (size_t)8 * g.N * sizeof(float)) != hipSuccess) return fail("hipMalloc");
            if (hipMalloc((void **)&S->Yh, (size_t)g.ns * g.N * sizeof(__half)) != hipSuccess) return fail("hipMalloc");
            if (!shared) {
                if (hipMalloc((void **)&S->M32, (size_t)g.ns * g.N * sizeof(float)) != hipSuccess) return fail("hipMalloc");
                if (mlp_init(S->mlp, g.N, S->st)) return fail("learned preconditioner allocation");
            }
        }
    }
    if (c->conc_model == 1) {
        S->conc = true;
        double **cp[] = {&S->cfield, &S->c0, &S->cx, &S->cprev, &S->clast, &S->Fc0[0], &S->Fc0[1], &S->Fc1[0],
                         &S->Fc1[1], &S->G2[0], &S->G2[1], &S->rc0, &S->rc1, &S->ex[0], &S->ex[1], &S->ex[2],
                         &S->ex[3], &S->cstart, &S->citer};
        for (double **p : cp) {
            if (hipMalloc((void **)p, np * sizeof(double)) != hipSuccess) return fail("hipMalloc");
            if (hipMemset(*p, 0, np * sizeof(double)) != hipSuccess) return fail("memset");
        }
        for (double **p : {&S->dconc, &S->dsave, &S->bconc, &S->evps})
            if (hipMalloc((void **)p, g.N * sizeof(double)) != hipSuccess) return fail("hipMalloc");
        if (hipMemset(S->evps, 0, g.N * sizeof(double)) != hipSuccess) return fail("memset");
        if (hipMemset(S->dconc, 0, g.N * sizeof(double)) != hipSuccess) return fail("memset");
        if (hipMalloc((void **)&S->cscratch, ((size_t)g.N + 2 * g.ny + 256 + 8) * sizeof(double)) != hipSuccess ||
            hipMalloc((void **)&S->csums, (256 + 8) * sizeof(double)) != hipSuccess ||
            hipMalloc((void **)&S->cflags, 4 * sizeof(int)) != hipSuccess ||
            hipMalloc((void **)&S->cdmax, sizeof(unsigned long long)) != hipSuccess)
            return fail("hipMalloc");
        if (hipHostMalloc((void **)&S->hcflags, 4 * sizeof(int), hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&S->hcdmax, sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&S->hcsums, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess)
            return fail("hipHostMalloc");
        for (double *p : {S->cfield, S->c0, S->cx, S->cprev, S->citer})
            if (hipMemcpy(p, S->c, np * sizeof(double), hipMemcpyDeviceToDevice) != hipSuccess) return fail("copy");
        if (solver_init(S->Sc, fine_op(g), S->st)) return fail("solver allocation");
        ConcArgs &k = S->ca;
        std::memset(&k, 0, sizeof(k));
        k.g = g;
        k.theta_h = c->theta;
        k.conv = c->conc_convection; k.diff = c->conc_diffusion;
        k.side = (c->dim == 1 && c->centrifugal) ? 1 : 0;
        // an unlimited scheme's c1 and c2 over the sum of c1
        // (FiniteDifferences.cpp:22-93); a limited one's limiter, frozen at
        // the latest iterate
        k.lim = c->conc_scheme == YZ_CONV_KOREN || c->conc_scheme == YZ_CONV_VANLEER;
        k.lim_vl = c->conc_scheme == YZ_CONV_VANLEER;
        k.climit = S->citer;
        if (!k.lim) {
            double c1[4], c2[4];
            scheme_coefs(c->conc_scheme, c1, c2);
            for (int j = 0; j < 4; ++j) {
                const double a1 = c1[j], a2 = c2[j];
                k.wpos[j] = a1 + 1.0 * a2;
                k.wneg[j] = a1 + (-1.0) * a2;
            }
        }
        k.Df = c->Df; k.diff_c0 = c->diff_c0; k.diff_exp = c->diff_exp;
        k.row_scaling = c->conc_row_scaling;
        k.h = S->h; k.c = S->c; k.hold = S->hold; k.c0 = S->c0;
        for (int ax = 0; ax < 2; ++ax) { k.F0[ax] = S->Fc0[ax]; k.F1[ax] = S->Fc1[ax]; k.G2[ax] = S->G2[ax]; }
        k.r0 = S->rc0; k.r1 = S->rc1;
        k.cx = S->cx;
        k.A = S->Sc.A; k.b = S->bconc;
        // the step's sources (ConcEquationBase.h:310-318; HeightEquationBase.h
        // :1379-1385): the supply at the local c or at the nozzle's
        k.src = c->source ? S->srcv : nullptr;
        k.supplied = (c->source && c->src_conc_supplied) ? 1 : 0;
        k.src_conc = c->src_conc;
        k.evp = nullptr;
    }
    // row kernel arguments
    RowArgs &a = S->ra;
    std::memset(&a, 0, sizeof(a));
    a.g = g;
    a.cap = c->capillary; a.vg = c->vertical_gravity; a.pg = c->inplane_gravity; a.cf = c->centrifugal;
    a.ma = c->marangoni; a.dj = c->disjoining;
    a.Bo = c->Bo; a.Pgx = c->Pgx; a.Pgy = c->Pgy; a.Sp = c->Sp; a.AsMa2 = c->As * c->Ma / 2.0; a.Dp = c->Dp;
    a.hstar = c->hstar; a.pn = c->dis_n; a.pm = c->dis_m; a.rotx = c->rot_x; a.roty = c->rot_y;
    a.visc_c0 = c->visc_c0; a.visc_exp = c->visc_exp; a.sig_c0 = c->sig_c0;
    a.Ma_eff = c->marangoni ? c->Ma : 0.0;
    a.mob.lambda = c->slip > 0 ? c->slip : 0.0; a.mob.a = 3.0 * a.mob.lambda; a.mob.slip = c->slip > 0;
    a.inv_h[0] = 1.0 / g.dx; a.inv_h[1] = 1.0 / g.dy;
    a.inv_h2[0] = 1.0 / (g.dx * g.dx); a.inv_h2[1] = 1.0 / (g.dy * g.dy);
    auto whole = [](double e) { return (e >= 1.0 && e <= 16.0 && e == std::floor(e)) ? (int)e : 0; };
    a.pn_i = whole(c->dis_n); a.pm_i = whole(c->dis_m);
    build_ccap(g, a.Ccap);
    a.theta = c->theta;
    a.row_scaling = c->row_scaling;
    a.hs = height_scheme_of(c->height_scheme);
    a.s = S->s; a.c = S->c;
    a.ia = S->S.A; a.b = S->b; a.rmax = S->dmax;
    a.ax = S->axold;
    a.bt.evap = c->evaporation; a.bt.efunc = c->evap_func; a.bt.Ev = c->Ev; a.bt.coef = c->evap_coef;
    a.bt.power = c->evap_power; a.bt.concr = c->evap_concr; a.bt.mres = c->molar_resin;
    a.bt.msol = c->molar_solvent; a.bt.prec = c->precursor;
    a.evold = S->evold; a.src = S->srcv; a.evp_out = S->evpv;
    // C_TimeSteps::Build (TimeSteps.cpp:38-61) starts DT().old at the
    // initial step: the first step's truncation criterion is taken too
    S->dt_now = dt0; S->dt_old = dt0;
    return S;
}

extern "C" {

void yz_step_destroy(yz_step *S) {
    if (!S) return;
    if (S->dser) (void)hipFree(S->dser);
    solver_free(S->S);
    solver_free(S->S32);
    solver_free(S->Sc);
    for (double *p : {S->cfield, S->c0, S->cx, S->cprev, S->clast, S->citer, S->Fc0[0], S->Fc0[1], S->Fc1[0], S->Fc1[1],
                      S->G2[0], S->G2[1], S->rc0, S->rc1, S->ex[0], S->ex[1], S->ex[2], S->ex[3], S->dconc,
                      S->dsave, S->bconc, S->cscratch, S->csums, S->cstart, S->evps})
        if (p) (void)hipFree(p);
    if (S->cflags) (void)hipFree(S->cflags);
    if (S->cdmax) (void)hipFree(S->cdmax);
    for (void *p : {(void *)S->hcflags, (void *)S->hcdmax, (void *)S->hcsums})
        if (p) (void)hipHostFree(p);
    mlp_free(S->mlp);
    for (float *p : {S->b32, S->corr32, S->feat, S->mscale, S->M32})
        if (p) (void)hipFree(p);
    if (S->Yh) (void)hipFree(S->Yh);
    for (double *p : {S->h, S->hold, S->hold2, S->hnew, S->s, S->c, S->axold, S->b, S->corr, S->scratch, S->dinv,
                      S->evold, S->srcv, S->evpv})
        if (p) (void)hipFree(p);
    if (S->dmax) (void)hipFree(S->dmax);
    if (S->dflag) (void)hipFree(S->dflag);
    if (S->hmax) (void)hipHostFree(S->hmax);
    if (S->hflag) (void)hipHostFree(S->hflag);
    if (S->dterm) (void)hipFree(S->dterm);
    if (S->hterm) (void)hipHostFree(S->hterm);
    for (auto &e : S->ev) if (e) (void)hipEventDestroy(e);
    if (S->st && S->own_stream) (void)hipStreamDestroy(S->st);
    delete S;
}

static int series_step(yz_step *S);

// TimeAdvance (NewtonKantorovich.cpp:424-486) + TimeStepAdjustment
// (ImplicitSolver.cpp:144-190)
int yz_step_advance(yz_step *S, int *nk_iter, int *lin_iter, double *resid) {
    if (!S) { set_err("null step"); return 2; }
    const yz_case &c = S->cs;
    const Geo &g = S->g;
    const int np = g.nxp * g.nyp;
    for (double &t : S->tms) t = 0;
    YZ_TRY(hipEventRecord(S->ev[10], S->st));
    YZ_TRY(hipMemcpyAsync(S->hold2, S->hold, np * sizeof(double), hipMemcpyDeviceToDevice, S->st));
    YZ_TRY(hipMemcpyAsync(S->hold, S->hnew, np * sizeof(double), hipMemcpyDeviceToDevice, S->st));
    bool step_done = false, ok = false;
    double dt_request = S->dt_now;
    long long lin_sum = 0;
    int it = 0;
    double r = 0;
    do {
        const long long cuts0 = S->stag_cuts;
        if (nk_loop(S, &it, &r, &ok)) return 2;
        lin_sum += S->lin_step;
        if (!ok && S->extrap_move > 0) {
            S->stag_cuts = cuts0;
            S->extrap_blocked = true;
            const int e = nk_loop(S, &it, &r, &ok);
            S->extrap_blocked = false;
            if (e) return 2;
            lin_sum += S->lin_step;
        }
        if (c.step_control) {
            double crit = -1.0;
            if (ok) {
                if (fabs(S->dt_old) <= DBL_MIN) crit = 1.0;
                else {
                    double th = c.tsc_criterion;
                    if (fabs(th) <= DBL_MIN) th = 1.0e-2;
                    YZ_TRY(hipMemsetAsync(S->dmax, 0, 4 * sizeof(unsigned long long), S->st));
                    hipLaunchKernelGGL(k_absmax, dim3(std::min(nb(np), 1024)), dim3(TB), 0, S->st, np, S->hold, S->dmax + 2);
                    hipLaunchKernelGGL(k_truncation, dim3(nb(np)), dim3(TB), 0, S->st, g, S->hold2, S->hold, S->h,
                                       S->dt_old, S->dt_now, c.tsc_floor, (const unsigned long long *)(S->dmax + 2),
                                       S->dmax + 3);
                    if (sync_read(S)) return 2;
                    crit = 1.0 - bits_to_double(S->hmax[3]) / th;
                }
            }
            if (crit > 0.0) {
                step_done = true;
                if (crit > 0.90 && S->dt_now < c.dt_max) {
                    if (++S->tsc_count >= c.tsc_keep) {
                        S->tsc_count = 0;
                        dt_request = std::min(S->dt_now / c.tsc_change, c.dt_max);
                    }
                }
            } else {
                S->dt_now *= c.tsc_change;
                dt_request = S->dt_now;
                if (S->dt_now < c.dt_min) {
                    set_err("step fell below MinDt");
                    if (nk_iter) *nk_iter = it;
                    if (lin_iter) *lin_iter = (int)lin_sum;
                    return 1;
                }
            }
        } else {
            if (!ok) {
                set_err("the Newton iteration did not converge at the fixed step");
                if (nk_iter) *nk_iter = it;
                if (lin_iter) *lin_iter = (int)lin_sum;
                if (resid) *resid = r;
                return 1;
            }
            step_done = true;
        }
    } while (!step_done);
    if (S->conc && conc_after(S)) return 2;
    YZ_TRY(hipMemcpyAsync(S->hnew, S->h, np * sizeof(double), hipMemcpyDeviceToDevice, S->st));
    if (S->term_on && term_launch(S)) return 2;
    YZ_TRY(hipEventRecord(S->ev[11], S->st));
    YZ_TRY(hipEventSynchronize(S->ev[11]));
    S->tms[5] = ev_ms(S->ev[10], S->ev[11]);
    S->time += S->dt_now;
    S->dt_old = S->dt_now;
    S->dt_now = dt_request;
    S->first_step = false;
    S->stepped = true;
    ++S->steps;
    S->last_nk = it; S->last_nk_res = r;
    if (S->term_on) term_decide(S);
    if (S->ser_on && series_step(S)) return 2;
    if (nk_iter) *nk_iter = it;
    if (lin_iter) *lin_iter = (int)lin_sum;
    if (resid) *resid = r;
    return 0;
}

int yz_step_get_c(yz_step *S, double *c) {
    if (!S || !c) { set_err("null argument"); return 2; }
    const int np = S->g.nxp * S->g.nyp;
    YZ_TRY(hipMemcpyAsync(c, S->conc ? S->cfield : S->c, np * sizeof(double), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipStreamSynchronize(S->st));
    return 0;
}

int yz_step_face_fluxes(yz_step *S, const double *h, const double *c, double *F, double *r) {
    if (!S || !h || !c) { set_err("null argument"); return 2; }
    const Geo &g = S->g;
    const size_t np = (size_t)g.nxp * g.nyp;
    double *buf = nullptr;
    YZ_TRY(hipMalloc((void **)&buf, 5 * np * sizeof(double)));
    double *dh = buf, *dc = buf + np, *F0 = buf + 2 * np, *F1 = buf + 3 * np, *dr = buf + 4 * np;
    int rc = 0;
    do {
        if (hipMemcpyAsync(dh, h, np * sizeof(double), hipMemcpyHostToDevice, S->st) != hipSuccess ||
            hipMemcpyAsync(dc, c, np * sizeof(double), hipMemcpyHostToDevice, S->st) != hipSuccess ||
            hipMemsetAsync(F0, 0, 3 * np * sizeof(double), S->st) != hipSuccess) { rc = 2; break; }
        hipLaunchKernelGGL(k_ghosts, dim3(nb((int)np)), dim3(TB), 0, S->st, g, dh);
        hipLaunchKernelGGL(k_ghosts, dim3(nb((int)np)), dim3(TB), 0, S->st, g, dc);
        RowArgs a = S->ra;
        a.h = dh;
        a.c = dc;
        launch_face_fluxes(a, F0, F1, dr, S->st);
        if (hipGetLastError() != hipSuccess) { rc = 2; break; }
        if (F && (hipMemcpyAsync(F, F0, np * sizeof(double), hipMemcpyDeviceToHost, S->st) != hipSuccess ||
                  hipMemcpyAsync(F + np, F1, np * sizeof(double), hipMemcpyDeviceToHost, S->st) != hipSuccess)) {
            rc = 2; break;
        }
        if (r && hipMemcpyAsync(r, dr, np * sizeof(double), hipMemcpyDeviceToHost, S->st) != hipSuccess) { rc = 2; break; }
        if (hipStreamSynchronize(S->st) != hipSuccess) rc = 2;
    } while (0);
    (void)hipFree(buf);
    if (rc) set_err("face fluxes: device error");
    return rc;
}

int yz_step_conc_stats(yz_step *S, double *out5) {
    if (!S || !out5) { set_err("null argument"); return 2; }
    for (int k = 0; k < 5; ++k) out5[k] = S->cstat[k];
    return 0;
}

// CalVelocity (HeightEquationBase.h:1583-1601) of the context's film and
// concentration: the face velocities with their boundary rows, padded
int yz_step_velocity(yz_step *S, double *ux, double *uy) {
    if (!S || !ux || (S->g.dim == 2 && !uy)) { set_err("null argument"); return 2; }
    const Geo &g = S->g;
    const size_t np = (size_t)g.nxp * g.nyp;
    double *d = nullptr;
    YZ_TRY(hipMalloc((void **)&d, 2 * np * sizeof(double)));
    double *vx = d, *vy = d + np;
    RowArgs a = S->ra;
    a.h = S->hnew;
    int rc = 0;
    do {
        if (hipMemsetAsync(d, 0, 2 * np * sizeof(double), S->st) != hipSuccess) { rc = 2; break; }
        const int cells = (g.nxp - 3) * (g.dim == 2 ? g.nyp - 3 : 1);
        // the boundary rows twice: once for the terms' fields, once for the
        // mixed field (HeightEquationBase.h:1591-1600)
        if (g.dim == 2) {
            hipLaunchKernelGGL(k_velocity<2>, dim3(nb(cells)), dim3(TB), 0, S->st, a, vx, vy);
            for (int pass = 0; pass < 2; ++pass) {
                hipLaunchKernelGGL(k_velocity_bc_x, dim3(nb(g.nyp)), dim3(TB), 0, S->st, g, vx, vy);
                hipLaunchKernelGGL(k_velocity_bc_y, dim3(nb(g.nxp)), dim3(TB), 0, S->st, g, vx, vy);
            }
        } else {
            hipLaunchKernelGGL(k_velocity<1>, dim3(nb(cells)), dim3(TB), 0, S->st, a, vx, vy);
            for (int pass = 0; pass < 2; ++pass)
                hipLaunchKernelGGL(k_velocity_bc_x, dim3(1), dim3(TB), 0, S->st, g, vx, vy);
        }
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(ux, vx, np * sizeof(double), hipMemcpyDeviceToHost, S->st) != hipSuccess ||
            (g.dim == 2 && hipMemcpyAsync(uy, vy, np * sizeof(double), hipMemcpyDeviceToHost, S->st) != hipSuccess) ||
            hipStreamSynchronize(S->st) != hipSuccess) { rc = 2; break; }
    } while (0);
    (void)hipFree(d);
    if (rc) set_err("device error in yz_step_velocity");
    return rc;
}

// the .vts's velocity fields of the context's film: the mixed one as
// yz_step_velocity gives it, then the capillary, Marangoni, vertical-gravity,
// in-plane gravity and centrifugal terms' own, each with its boundary rows
// once (CalVelocity, HeightEquationBase.h:1589-1597)
int yz_step_velocity_terms(yz_step *S, double *ux6, double *uy6) {
    if (!S || !ux6 || (S->g.dim == 2 && !uy6)) { set_err("null argument"); return 2; }
    const Geo &g = S->g;
    const size_t np = (size_t)g.nxp * g.nyp;
    double *d = nullptr;
    YZ_TRY(hipMalloc((void **)&d, 12 * np * sizeof(double)));
    double *vx = d, *vy = d + 6 * np;
    RowArgs a = S->ra;
    a.h = S->hnew;
    int rc = 0;
    do {
        if (hipMemsetAsync(d, 0, 12 * np * sizeof(double), S->st) != hipSuccess) { rc = 2; break; }
        const int cells = (g.nxp - 3) * (g.dim == 2 ? g.nyp - 3 : 1);
        if (g.dim == 2) hipLaunchKernelGGL(k_velocity_terms<2>, dim3(nb(cells)), dim3(TB), 0, S->st, a, vx, vy, np);
        else hipLaunchKernelGGL(k_velocity_terms<1>, dim3(nb(cells)), dim3(TB), 0, S->st, a, vx, vy, np);
        for (int f = 0; f < 6; ++f)
            for (int pass = 0; pass < (f == 0 ? 2 : 1); ++pass) {
                if (g.dim == 2) {
                    hipLaunchKernelGGL(k_velocity_bc_x, dim3(nb(g.nyp)), dim3(TB), 0, S->st, g, vx + f * np, vy + f * np);
                    hipLaunchKernelGGL(k_velocity_bc_y, dim3(nb(g.nxp)), dim3(TB), 0, S->st, g, vx + f * np, vy + f * np);
                } else {
                    hipLaunchKernelGGL(k_velocity_bc_x, dim3(1), dim3(TB), 0, S->st, g, vx + f * np, vy + f * np);
                }
            }
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(ux6, vx, 6 * np * sizeof(double), hipMemcpyDeviceToHost, S->st) != hipSuccess ||
            (g.dim == 2 && hipMemcpyAsync(uy6, vy, 6 * np * sizeof(double), hipMemcpyDeviceToHost, S->st) != hipSuccess) ||
            hipStreamSynchronize(S->st) != hipSuccess) { rc = 2; break; }
    } while (0);
    (void)hipFree(d);
    if (rc) set_err("device error in yz_step_velocity_terms");
    return rc;
}

// the energy budget of the context's film (Output.Energy, EnergyBudget.h):
// the free energy F and the power D of the face fluxes against the potential
int yz_step_energy(yz_step *S, double *F, double *D) {
    if (!S || !F || !D) { set_err("null argument"); return 2; }
    const Geo &g = S->g;
    const size_t np = (size_t)g.nxp * g.nyp;
    const int nblk = (g.N + EB - 1) / EB;
    double *d = nullptr;
    YZ_TRY(hipMalloc((void **)&d, (3 * np + 2 * (size_t)nblk + 2) * sizeof(double)));
    double *Fx = d, *Fy = d + np, *r = d + 2 * np, *part = d + 3 * np, *out = part + 2 * nblk;
    RowArgs a = S->ra;
    a.h = S->hnew;
    double host[2] = {0, 0};
    int rc = 0;
    do {
        if (hipMemsetAsync(d, 0, 3 * np * sizeof(double), S->st) != hipSuccess) { rc = 2; break; }
        launch_face_fluxes(a, Fx, Fy, r, S->st);
        const double cm = g.dim == 2 ? g.dx * g.dy : g.dx;
        if (g.dim == 2) hipLaunchKernelGGL(k_energy_cells<2>, dim3(nblk), dim3(EB), 0, S->st, a, Fx, Fy, r, part);
        else hipLaunchKernelGGL(k_energy_cells<1>, dim3(nblk), dim3(EB), 0, S->st, a, Fx, Fy, r, part);
        hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(EB), 0, S->st, nblk, part, cm, out);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(host, out, 2 * sizeof(double), hipMemcpyDeviceToHost, S->st) != hipSuccess ||
            hipStreamSynchronize(S->st) != hipSuccess) { rc = 2; break; }
    } while (0);
    (void)hipFree(d);
    if (rc) { set_err("device error in yz_step_energy"); return rc; }
    *F = host[0];
    *D = host[1];
    return 0;
}

// ---- the run's time series (History.dat, CalcInfo.dat, Energy.dat) ------
// FieldInfo, the change rate and the step's balances in one reduction
// (k_series_cells); v[SER_Q] as k_series_cells lays them out
static int series_reduce(yz_step *S, double *v) {
    const Geo &g = S->g;
    if (!S->dser && hipMalloc((void **)&S->dser, (SER_Q * (size_t)SER_G + SER_Q) * sizeof(double)) != hipSuccess) {
        set_err("hipMalloc"); return 2;
    }
    const int np = g.nxp * g.nyp;
    const int nbk = std::min(SER_G, (np + SER_B - 1) / SER_B);
    RowArgs a = S->ra;
    a.h = S->hnew;
    a.c = S->conc ? S->cfield : S->c;
    const bool ran = S->stepped;
    const int side = (ran && g.dim == 1 && S->cs.centrifugal) ? 1 : 0;
    hipLaunchKernelGGL(k_series_cells, dim3(nbk), dim3(SER_B), 0, S->st, a, S->hold, S->conc ? S->cstart : S->c,
                       (ran && S->cs.evaporation) ? S->evpv : nullptr, (ran && S->cs.source) ? S->srcv : nullptr,
                       (S->conc && S->cs.src_conc_supplied) ? 1 : 0, S->cs.src_conc, side, S->dser);
    hipLaunchKernelGGL(k_series_final, dim3(1), dim3(SER_B), 0, S->st, nbk, S->dser, S->dser + SER_Q * SER_G);
    YZ_TRY(hipGetLastError());
    YZ_TRY(hipMemcpyAsync(v, S->dser + SER_Q * SER_G, SER_Q * sizeof(double), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipStreamSynchronize(S->st));
    return 0;
}
static double cell_measure(const Geo &g) { return g.dim == 2 ? g.dx * g.dy : g.dx; }
// what one step adds to the run's sums (WellMixedModelBase.h:1339-1411),
// with the dt the step was solved with
static int series_step(yz_step *S) {
    double v[SER_Q];
    if (series_reduce(S, v)) return 2;
    const double cm = cell_measure(S->g), dt = S->dt_old, th = S->cs.theta;
    if (S->cs.evaporation) S->ser_evap += v[10] * cm * dt;
    if (S->cs.source) {
        S->ser_sup += v[11] * cm * dt;
        S->ser_supr += v[12] * cm * dt;
    }
    if (S->g.dim == 1 && S->cs.centrifugal) {
        S->ser_out += (th * v[13] + (1.0 - th) * v[15]) * cm * dt;
        S->ser_outr += (th * v[14] + (1.0 - th) * v[16]) * cm * dt;
    }
    if (S->ser_sc.energy) {
        const double F0 = S->en_F, D0 = S->en_D;
        if (yz_step_energy(S, &S->en_F, &S->en_D)) return 2;
        S->en_diss += 0.5 * (D0 + S->en_D) * dt;
        S->en_bal = (S->en_F - F0) / dt + 0.5 * (D0 + S->en_D);
        S->en_step = true;
    }
    return 0;
}
// the header line of a time series (OutputTimeline::Build)
static bool series_header(const std::string &path, bool dim, const std::vector<std::string> &labels) {
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    std::string line = dim ? "Time [s]" : "Time ";
    for (const std::string &l : labels) line += "\t" + l;
    line += "\n";
    const bool ok = std::fputs(line.c_str(), f) >= 0;
    return std::fclose(f) == 0 && ok;
}
static bool series_append(const std::string &path, const std::string &row) {
    FILE *f = std::fopen(path.c_str(), "a");
    if (!f) return false;
    const bool ok = std::fputs((row + "\n").c_str(), f) >= 0;
    return std::fclose(f) == 0 && ok;
}
static void put_e(std::string &row, const char *fmt, double v) {
    char b[64];
    std::snprintf(b, sizeof b, fmt, v);
    row += b;
}

int yz_step_field_info(yz_step *S, double *out9) {
    if (!S || !out9) { set_err("null argument"); return 2; }
    double v[SER_Q];
    if (series_reduce(S, v)) return 2;
    out9[0] = v[0]; out9[1] = v[1]; out9[2] = v[1] / v[0];
    out9[3] = v[4]; out9[4] = v[5];   // h + s
    out9[5] = v[2]; out9[6] = v[3];   // h
    out9[7] = v[6]; out9[8] = v[7];   // c
    return 0;
}

int yz_step_series_open(yz_step *S, const char *dir, const yz_series_scales *sc) {
    if (!S || !dir || !sc) { set_err("null argument"); return 2; }
    const bool dim = sc->dimensioned != 0, one = S->g.dim == 1;
    const std::string uV = dim ? (one ? " [m^3/m]" : " [m^3]") : " [-]", uH = dim ? " [m]" : " [-]";
    const std::string d = dir;
    const std::vector<std::string> hist = {
        "Volume" + uV, "ResinVol" + uV, "AveC [-]", "EvaporatedVol" + uV, "OutflowedVol" + uV,
        "OutflowedResin" + uV, "VolErr [-]", "ResinVolErr [-]", "maxHS" + uH, "minHS" + uH, "maxH" + uH,
        "minH" + uH, "maxC [-]", "minC [-]", "SuppliedVol" + uV, "SuppliedResin" + uV};
    const std::vector<std::string> calc = {"NK_iter", "NK_resid", "BCG_iter", "BCG_resid", "deltaT", "ChangeRate"};
    if (!series_header(d + "/History.dat", dim, hist) || !series_header(d + "/CalcInfo.dat", dim, calc)) {
        set_err("yz_step_series_open: cannot write the files"); return 2;
    }
    if (sc->energy) {
        const std::string uE = dim ? (one ? " [J/m]" : " [J]") : " [-]", uP = dim ? (one ? " [W/m]" : " [W]") : " [-]";
        if (!series_header(d + "/Energy.dat", dim, {"FreeEnergy" + uE, "Dissipation" + uP, "Dissipated" + uE,
                                                    "Balance" + uP})) {
            set_err("yz_step_series_open: cannot write the files"); return 2;
        }
    }
    double info[9];
    if (yz_step_field_info(S, info)) return 2;
    S->ser_vol0 = info[0]; S->ser_resin0 = info[1];
    S->ser_evap = S->ser_sup = S->ser_supr = S->ser_out = S->ser_outr = 0;
    S->en_diss = S->en_bal = 0; S->en_step = false;
    if (sc->energy && yz_step_energy(S, &S->en_F, &S->en_D)) return 2;
    S->ser_dir = d;
    S->ser_sc = *sc;
    S->ser_on = true;
    return 0;
}

int yz_step_series_write(yz_step *S) {
    if (!S) { set_err("null step"); return 2; }
    if (!S->ser_on) { set_err("yz_step_series_write: no series open"); return 2; }
    double v[SER_Q];
    if (series_reduce(S, v)) return 2;
    const yz_series_scales &sc = S->ser_sc;
    const bool dim = sc.dimensioned != 0;
    const double ts = dim ? sc.T0 : 1.0;
    double vs = 1.0, hs = 1.0;
    if (dim) {
        vs = sc.L0 * sc.H0;
        if (S->g.dim == 2) vs *= sc.L0;
        hs = sc.H0;
    }
    const double vf = v[0], vr = v[1];
    const double volErr = (vf + S->ser_out + S->ser_evap - S->ser_sup) / S->ser_vol0 - 1.0;
    const double resinErr = (vr + S->ser_outr - S->ser_supr) / S->ser_resin0 - 1.0;
    std::string row;
    put_e(row, "%e", S->time * ts);
    const double cols[16] = {vf * vs, vr * vs, vr / vf, S->ser_evap * vs, S->ser_out * vs, S->ser_outr * vs,
                             volErr, resinErr, v[4] * hs, v[5] * hs, v[2] * hs, v[3] * hs, v[6], v[7],
                             S->ser_sup * vs, S->ser_supr * vs};
    for (double x : cols) put_e(row, "\t%e", x);
    bool ok = series_append(S->ser_dir + "/History.dat", row);
    // ChangeRateOfVariable of the last step
    double rate;
    if (S->steps < 5 || fabs(S->dt_old) <= DBL_MIN) rate = 1;
    else if (fabs(v[9]) <= DBL_MIN) rate = 0;
    else rate = v[8] / (v[9] * S->dt_old);
    row.clear();
    put_e(row, "%e", S->time * ts);
    row += "\t" + std::to_string(S->last_nk);
    put_e(row, "\t%e", S->last_nk_res);
    row += "\t" + std::to_string(S->last_lin_it);
    put_e(row, "\t%e", S->last_lin_res);
    put_e(row, "\t%e", S->dt_now * ts);
    put_e(row, "\t%e", rate / ts);
    ok = ok && series_append(S->ser_dir + "/CalcInfo.dat", row);
    if (sc.energy) {
        double eS = 1, pS = 1;
        if (dim) {
            eS = sc.sigma0 * sc.H0 * sc.H0;
            if (S->g.dim == 1) eS /= sc.L0;
            pS = eS / sc.T0;
        }
        row.clear();
        put_e(row, "%.17e", S->time * ts);
        put_e(row, "\t%.17e", S->en_F * eS);
        put_e(row, "\t%.17e", S->en_D * pS);
        put_e(row, "\t%.17e", S->en_diss * eS);
        put_e(row, "\t%.17e", (S->en_step ? S->en_bal : 0.0) * pS);
        ok = ok && series_append(S->ser_dir + "/Energy.dat", row);
    }
    if (!ok) { set_err("yz_step_series_write: cannot append to the files"); return 2; }
    return 0;
}

// the volume balance of the context's film (History.dat's volume columns,
// WellMixedModelBase.h:1339-1360): the film's volume and solute, and what
// the last step evaporated (EvaporationRate, HeightEquationBase.h:359-369,
// times its dt) and what the nozzle supplied (SuppliedVolume :293-303)
int yz_set_fault(long k, long m) {
    g_fault_read = true;
    g_fault_at = k; g_fault_count = m > 0 ? m : 1; g_fault_seen = 0;
    return 0;
}

int yz_set_row_kernel(int tiled) {
    const int was = g_rows_tiled;
    if (tiled == 0 || tiled == 1) g_rows_tiled = tiled;
    return was;
}

int yz_step_balances(yz_step *S, double *out4) {
    if (!S || !out4) { set_err("null argument"); return 2; }
    const Geo &g = S->g;
    const int nblk = (g.N + EB - 1) / EB;
    double *d = nullptr;
    YZ_TRY(hipMalloc((void **)&d, (4 * (size_t)nblk + 4) * sizeof(double)));
    double host[4] = {0, 0, 0, 0};
    int rc = 0;
    const bool ran = S->stepped;
    hipLaunchKernelGGL(k_balance_cells, dim3(nblk), dim3(EB), 0, S->st, g, S->hnew, S->c,
                       ran ? S->evpv : nullptr, (ran && S->cs.source) ? S->srcv : nullptr, d);
    hipLaunchKernelGGL(k_balance_final, dim3(1), dim3(EB), 0, S->st, nblk, d, d + 4 * nblk);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(host, d + 4 * nblk, 4 * sizeof(double), hipMemcpyDeviceToHost, S->st) != hipSuccess ||
        hipStreamSynchronize(S->st) != hipSuccess) rc = 2;
    (void)hipFree(d);
    if (rc) { set_err("device error in yz_step_balances"); return rc; }
    const double cm = g.dim == 2 ? g.dx * g.dy : g.dx;
    const double dt = ran ? S->dt_old : 0.0;
    out4[0] = host[0] * cm;
    out4[1] = host[1] * cm;
    out4[2] = S->cs.evaporation ? host[2] * cm * dt : 0.0;
    out4[3] = host[3] * cm * dt;
    return 0;
}

int yz_step_write_dat(yz_step *S, const char *path, double L0, double H0, double U0, int dimensioned) {
    if (!S || !path) { set_err("null argument"); return 2; }
    const size_t np = (size_t)S->g.nxp * S->g.nyp;
    std::vector<double> h(np), s(np), c(np), ux(np), uy(np, 0.0);
    if (yz_step_velocity(S, ux.data(), uy.data())) return 2;
    YZ_TRY(hipMemcpyAsync(h.data(), S->hnew, np * sizeof(double), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipMemcpyAsync(s.data(), S->s, np * sizeof(double), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipMemcpyAsync(c.data(), S->c, np * sizeof(double), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipStreamSynchronize(S->st));
    return yz_write_dat(path, &S->cs, h.data(), s.data(), c.data(), ux.data(), uy.data(), L0, H0, U0, dimensioned);
}

int yz_step_write_vts(yz_step *S, const char *path, double L0, double H0, double U0, double T0, int dimensioned) {
    if (!S || !path) { set_err("null argument"); return 2; }
    const size_t np = (size_t)S->g.nxp * S->g.nyp;
    std::vector<double> h(np), s(np), c(np), ux(6 * np), uy(6 * np, 0.0);
    if (yz_step_velocity_terms(S, ux.data(), uy.data())) return 2;
    YZ_TRY(hipMemcpyAsync(h.data(), S->hnew, np * sizeof(double), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipMemcpyAsync(s.data(), S->s, np * sizeof(double), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipMemcpyAsync(c.data(), S->c, np * sizeof(double), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipStreamSynchronize(S->st));
    return yz_write_vts(path, &S->cs, S->time, h.data(), s.data(), c.data(), ux.data(), uy.data(), L0, H0, U0, T0,
                        dimensioned);
}

int yz_step_get_h(yz_step *S, double *h) {
    if (!S || !h) { set_err("null argument"); return 2; }
    const int np = S->g.nxp * S->g.nyp;
    YZ_TRY(hipMemcpyAsync(h, S->hnew, np * sizeof(double), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipStreamSynchronize(S->st));
    return 0;
}

// the restart state (WellMixed_ModelBase::WriteState / BuildRestart,
// WellMixedModelBase.h:896-927, 1196-1221): the clock (time, step, dt now /
// before / of the last step), the step control's count, the film of the
// step before (height.old), the last Newton correction (height.correction,
// which the next step's first solve starts from), the backends' memory
// (none here: empty), the fields h, s, c with their ghosts; and what the
// device step carries besides: the concentration equation's start
// (conc.start, conc_model 1) and the learned preconditioner's weights,
// Adam moments and training counters (learn.*)
static std::vector<double> pad_interior(const Geo &g, const std::vector<double> &in) {
    const int np = g.nxp * g.nyp;
    std::vector<double> out((size_t)np, 0.0);
    auto fold = [](int i, int n, int lo, int hi) {
        const int m = n - 4;
        if (i < 2) return lo == YZ_BC_PERIODIC ? i + m : 3 - i;
        if (i > n - 3) return hi == YZ_BC_PERIODIC ? i - m : 2 * (n - 3) + 1 - i;
        return i;
    };
    for (int iy = 0; iy < g.nyp; ++iy)
        for (int ix = 0; ix < g.nxp; ++ix) {
            const int fx = fold(ix, g.nxp, g.bc[0], g.bc[1]);
            const int fy = g.dim == 2 ? fold(iy, g.nyp, g.bc[2], g.bc[3]) : 0;
            const int q = (g.dim == 2 ? (fy - 2) * g.nx : 0) + (fx - 2);
            out[(size_t)iy * g.nxp + ix] = in[(size_t)q];
        }
    return out;
}

int yz_step_save_state(yz_step *S, const char *path) {
    if (!S || !path) { set_err("null argument"); return 2; }
    const Geo &g = S->g;
    const size_t np = (size_t)g.nxp * g.nyp;
    std::vector<double> hnew(np), hold(np), sv(np), cv(np), corr((size_t)g.N), c0;
    YZ_TRY(hipMemcpyAsync(hnew.data(), S->hnew, np * sizeof(double), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipMemcpyAsync(hold.data(), S->hold, np * sizeof(double), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipMemcpyAsync(sv.data(), S->s, np * sizeof(double), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipMemcpyAsync(cv.data(), S->conc ? S->cfield : S->c, np * sizeof(double), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipMemcpyAsync(corr.data(), S->corr, g.N * sizeof(double), hipMemcpyDeviceToHost, S->st));
    if (S->conc) {
        c0.resize(np);
        YZ_TRY(hipMemcpyAsync(c0.data(), S->c0, np * sizeof(double), hipMemcpyDeviceToHost, S->st));
    }
    const bool learned = S->cs.precond == YZ_PC_LEARNED || S->cs.precond == YZ_PC_LEARNED_AMG;
    std::vector<float> lp, lm, lv;
    if (learned) {
        lp.resize(MLP_NP); lm.resize(MLP_NP); lv.resize(MLP_NP);
        YZ_TRY(hipMemcpyAsync(lp.data(), S->mlp.params, MLP_NP * sizeof(float), hipMemcpyDeviceToHost, S->st));
        YZ_TRY(hipMemcpyAsync(lm.data(), S->mlp.adam_m, MLP_NP * sizeof(float), hipMemcpyDeviceToHost, S->st));
        YZ_TRY(hipMemcpyAsync(lv.data(), S->mlp.adam_v, MLP_NP * sizeof(float), hipMemcpyDeviceToHost, S->st));
    }
    YZ_TRY(hipStreamSynchronize(S->st));
    StateFile st;
    st.set_text("model", g.dim == 2 ? "Well-Mixed-2D" : "Well-Mixed-1D");
    st.set("time", S->time);
    st.set("step", (double)S->steps);
    st.set("dt", std::vector<double>{S->dt_now, S->dt_old, S->dt_old});
    st.set("adjustCount.height", (double)S->tsc_count);
    st.set_vector("height.old", hold);
    st.set_vector("height.correction", pad_interior(g, corr));
    st.set_vector("solver.height", std::vector<double>());
    st.set_vector("solver.conc", std::vector<double>());
    st.set_vector("field.h", hnew);
    st.set_vector("field.s", sv);
    st.set_vector("field.c", cv);
    if (S->conc) st.set_vector("conc.start", c0);
    if (learned) {
        st.set("learn.state", std::vector<double>{(double)S->mlp.trained_passes, (double)S->mlp.adam_t});
        st.set_vector("learn.params", std::vector<double>(lp.begin(), lp.end()));
        st.set_vector("learn.adam_m", std::vector<double>(lm.begin(), lm.end()));
        st.set_vector("learn.adam_v", std::vector<double>(lv.begin(), lv.end()));
    }
    if (!st.write(path)) { set_err(std::string("could not write the state file ") + path); return 2; }
    return 0;
}

int yz_step_load_state(yz_step *S, const char *path) {
    if (!S || !path) { set_err("null argument"); return 2; }
    const Geo &g = S->g;
    const size_t np = (size_t)g.nxp * g.nyp;
    StateFile st;
    if (!st.read(path)) { set_err(std::string("not a readable state file: ") + path); return 2; }
    const std::string model = g.dim == 2 ? "Well-Mixed-2D" : "Well-Mixed-1D";
    if (st.text("model") != model) { set_err("the state file was written by " + st.text("model") + ", not " + model); return 2; }
    const char *need[] = {"field.h", "field.s", "field.c", "height.old"};
    for (const char *k : need)
        if (st.vec(k).size() != np) { set_err(std::string("state vector ") + k + " missing or of another grid"); return 2; }
    const double dt_now = st.get("dt", 0), dt_old = st.get("dt", 1);
    if (!(dt_now > 0) || !(dt_old >= 0) || !std::isfinite(st.get("time"))) { set_err("state file: bad clock"); return 2; }
    const bool learned = S->cs.precond == YZ_PC_LEARNED || S->cs.precond == YZ_PC_LEARNED_AMG;
    YZ_TRY(hipMemcpyAsync(S->hnew, st.vec("field.h").data(), np * sizeof(double), hipMemcpyHostToDevice, S->st));
    YZ_TRY(hipMemcpyAsync(S->h, st.vec("field.h").data(), np * sizeof(double), hipMemcpyHostToDevice, S->st));
    YZ_TRY(hipMemcpyAsync(S->hold, st.vec("height.old").data(), np * sizeof(double), hipMemcpyHostToDevice, S->st));
    YZ_TRY(hipMemcpyAsync(S->s, st.vec("field.s").data(), np * sizeof(double), hipMemcpyHostToDevice, S->st));
    if (S->conc) {
        YZ_TRY(hipMemcpyAsync(S->cfield, st.vec("field.c").data(), np * sizeof(double), hipMemcpyHostToDevice, S->st));
        YZ_TRY(hipMemcpyAsync(S->c, st.vec("field.c").data(), np * sizeof(double), hipMemcpyHostToDevice, S->st));
        const std::vector<double> &c0 = st.vec("conc.start").size() == np ? st.vec("conc.start") : st.vec("field.c");
        YZ_TRY(hipMemcpyAsync(S->c0, c0.data(), np * sizeof(double), hipMemcpyHostToDevice, S->st));
    } else {
        YZ_TRY(hipMemcpyAsync(S->c, st.vec("field.c").data(), np * sizeof(double), hipMemcpyHostToDevice, S->st));
    }
    std::vector<double> corr((size_t)g.N, 0.0);
    if (st.vec("height.correction").size() == np) {
        const std::vector<double> &v = st.vec("height.correction");
        for (int iy = 0; iy < g.ny; ++iy)
            for (int ix = 0; ix < g.nx; ++ix)
                corr[(size_t)iy * g.nx + ix] = v[(size_t)(g.dim == 2 ? iy + 2 : 0) * g.nxp + ix + 2];
    }
    YZ_TRY(hipMemcpyAsync(S->corr, corr.data(), g.N * sizeof(double), hipMemcpyHostToDevice, S->st));
    if (learned && st.vec("learn.params").size() == (size_t)MLP_NP) {
        std::vector<float> p(MLP_NP), m(MLP_NP, 0.0f), v(MLP_NP, 0.0f);
        for (int i = 0; i < MLP_NP; ++i) p[i] = (float)st.vec("learn.params")[i];
        if (st.vec("learn.adam_m").size() == (size_t)MLP_NP)
            for (int i = 0; i < MLP_NP; ++i) m[i] = (float)st.vec("learn.adam_m")[i];
        if (st.vec("learn.adam_v").size() == (size_t)MLP_NP)
            for (int i = 0; i < MLP_NP; ++i) v[i] = (float)st.vec("learn.adam_v")[i];
        YZ_TRY(hipMemcpyAsync(S->mlp.params, p.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice, S->st));
        YZ_TRY(hipMemcpyAsync(S->mlp.adam_m, m.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice, S->st));
        YZ_TRY(hipMemcpyAsync(S->mlp.adam_v, v.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice, S->st));
        S->mlp.trained_passes = (int)st.get("learn.state", 0);
        S->mlp.adam_t = (int)st.get("learn.state", 1);
    }
    YZ_TRY(hipStreamSynchronize(S->st));
    S->time = st.get("time");
    S->steps = (long long)st.get("step");
    S->dt_now = dt_now;
    S->dt_old = dt_old;
    S->tsc_count = (int)st.get("adjustCount.height");
    S->first_step = S->steps == 0 && dt_old == 0;
    S->stepped = false;   // the file holds no last-step evaporation or supply
    return 0;
}

int yz_step_set_h(yz_step *S, const double *h) {
    if (!S || !h) { set_err("null argument"); return 2; }
    const Geo &g = S->g;
    const int np = g.nxp * g.nyp;
    YZ_TRY(hipMemcpyAsync(S->hnew, h, np * sizeof(double), hipMemcpyHostToDevice, S->st));
    hipLaunchKernelGGL(k_ghosts, dim3(nb(np)), dim3(TB), 0, S->st, g, S->hnew);
    YZ_TRY(hipStreamSynchronize(S->st));
    return 0;
}

int yz_step_set_term(yz_step *S, const yz_term *t) {
    if (!S) { set_err("null step"); return 2; }
    if (!t) { S->term_on = 0; return 0; }
    if (!S->dterm) {
        YZ_TRY(hipMalloc((void **)&S->dterm, (TERM_MINSTART + 1) * sizeof(double)));
        YZ_TRY(hipMemset(S->dterm, 0, (TERM_MINSTART + 1) * sizeof(double)));
        YZ_TRY(hipHostMalloc((void **)&S->hterm, (TERM_Q + 1) * sizeof(double), hipHostMallocDefault));
    }
    S->term = *t;
    S->term_on = 1;
    S->term_status = 0;
    return 0;
}

int yz_step_term_status(yz_step *S, int *status, double *info8) {
    if (!S) { set_err("null step"); return 2; }
    if (status) *status = S->term_status;
    if (info8) std::memcpy(info8, S->term_info, sizeof(S->term_info));
    return 0;
}

int yz_step_run(yz_step *S, int nsteps, int *done, int *status) {
    if (!S) { set_err("null step"); return 2; }
    int k = 0, rc = 0;
    for (; k < nsteps; ++k) {
        int nk = 0, lin = 0;
        double r = 0;
        rc = yz_step_advance(S, &nk, &lin, &r);
        if (rc != 0) break;
        if (S->term_on && S->term_status != YZ_RUN_GOING) { ++k; break; }
    }
    if (done) *done = k;
    if (status) *status = S->term_status;
    return rc;
}

int yz_step_pass_log(yz_step *S, double *corr, double *eta, double *lin_resid, int cap, long long *stag_cuts) {
    if (!S) { set_err("null step"); return -1; }
    const int m = (int)S->plog_corr.size();
    for (int i = 0; i < m && i < cap; ++i) {
        if (corr) corr[i] = S->plog_corr[i];
        if (eta) eta[i] = S->plog_eta[i];
        if (lin_resid) lin_resid[i] = S->plog_lres[i];
    }
    if (stag_cuts) *stag_cuts = S->stag_cuts;
    return m;
}

int yz_step_clock(yz_step *S, double *t, double *dt_now, double *dt_old) {
    if (!S) { set_err("null step"); return 2; }
    if (t) *t = S->time;
    if (dt_now) *dt_now = S->dt_now;
    if (dt_old) *dt_old = S->dt_old;
    return 0;
}

int yz_step_kernel_time(yz_step *S, int enable, long long *count, double *total_ms) {
    if (!S) { set_err("null step"); return 2; }
    // the solver in use: the FP32 inner solve when the case runs it
    const bool f32 = fp32_path(S->cs);
    int &on = f32 ? S->S32.ktime : S->S.ktime;
    long long &n = f32 ? S->S32.kcount : S->S.kcount;
    double &ms = f32 ? S->S32.kms : S->S.kms;
    if (count) *count = n;
    if (total_ms) *total_ms = ms;
    if (enable >= 0) { on = enable; n = 0; ms = 0; }
    return 0;
}

int yz_step_timing(yz_step *S, double *ms6) {
    if (!S || !ms6) { set_err("null argument"); return 2; }
    for (int k = 0; k < 6; ++k) ms6[k] = S->tms[k];
    return 0;
}

int yz_step_assemble(yz_step *S, const double *h_old, const double *h, double dt, double *ia, double *ax,
                     double *rhs) {
    if (!S || !h_old || !h) { set_err("null argument"); return 2; }
    const Geo &g = S->g;
    const int np = g.nxp * g.nyp;
    double *dia = nullptr, *dax = nullptr, *drhs = nullptr;
    YZ_TRY(hipMalloc((void **)&dia, (size_t)g.ns * g.N * sizeof(double)));
    YZ_TRY(hipMalloc((void **)&dax, g.N * sizeof(double)));
    YZ_TRY(hipMalloc((void **)&drhs, g.N * sizeof(double)));
    YZ_TRY(hipMemcpyAsync(S->hold, h_old, np * sizeof(double), hipMemcpyHostToDevice, S->st));
    YZ_TRY(hipMemcpyAsync(S->h, h, np * sizeof(double), hipMemcpyHostToDevice, S->st));
    hipLaunchKernelGGL(k_ghosts, dim3(nb(np)), dim3(TB), 0, S->st, g, S->hold);
    hipLaunchKernelGGL(k_ghosts, dim3(nb(np)), dim3(TB), 0, S->st, g, S->h);
    if (b_before(S, S->hold, dt, S->time)) return 2;
    RowArgs keep = S->ra;
    S->ra.ax = S->axold;
    rows_launch(S, 0, S->hold);
    S->ra = keep;
    S->ra.dt = dt; S->ra.axold = S->axold; S->ra.hold = S->hold;
    S->ra.ia_raw = dia; S->ra.ax_raw = dax; S->ra.rhs_raw = drhs;
    YZ_TRY(hipMemsetAsync(S->dmax, 0, 4 * sizeof(unsigned long long), S->st));
    rows_launch(S, 1, S->h);
    S->ra = keep;
    YZ_TRY(hipGetLastError());
    std::vector<double> tax(g.N), trhs(g.N);
    if (ia) YZ_TRY(hipMemcpyAsync(ia, dia, (size_t)g.ns * g.N * sizeof(double), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipMemcpyAsync(tax.data(), dax, g.N * sizeof(double), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipMemcpyAsync(trhs.data(), drhs, g.N * sizeof(double), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipStreamSynchronize(S->st));
    (void)hipFree(dia); (void)hipFree(dax); (void)hipFree(drhs);
    for (int i = 0; i < g.N; ++i) {
        const int ix = i % g.nx, iy = i / g.nx;
        const int p = (g.dim == 2 ? iy + 2 : 0) * g.nxp + ix + 2;
        if (ax) ax[p] = tax[i];
        if (rhs) rhs[p] = trhs[i];
    }
    if (ax || rhs)
        for (int p = 0; p < np; ++p) {
            const int ix = p % g.nxp, iy = p / g.nxp;
            const bool inside = ix >= 2 && ix <= g.nxp - 3 && (g.dim == 1 || (iy >= 2 && iy <= g.nyp - 3));
            if (inside) continue;
            if (ax) ax[p] = 0;
            if (rhs) rhs[p] = 0;
        }
    // the step's own state is left as it was before the call
    YZ_TRY(hipMemcpyAsync(S->hold, S->hnew, np * sizeof(double), hipMemcpyDeviceToDevice, S->st));
    YZ_TRY(hipMemcpyAsync(S->h, S->hnew, np * sizeof(double), hipMemcpyDeviceToDevice, S->st));
    YZ_TRY(hipStreamSynchronize(S->st));
    return 0;
}

// ---- learned preconditioner ------------------------------------------

static int learn_ctx(yz_step *S) {
    if (!S) { set_err("null step"); return 2; }
    if (S->cs.precond != YZ_PC_LEARNED && S->cs.precond != YZ_PC_LEARNED_AMG) {
        set_err("the context was not created with a learned preconditioner"); return 2;
    }
    return 0;
}

int yz_learn_get_params(yz_step *S, float *params) {
    if (int e = learn_ctx(S)) return e;
    if (!params) { set_err("null argument"); return 2; }
    YZ_TRY(hipMemcpyAsync(params, S->mlp.params, MLP_NP * sizeof(float), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipStreamSynchronize(S->st));
    return 0;
}

int yz_learn_set_params(yz_step *S, const float *params) {
    if (int e = learn_ctx(S)) return e;
    if (!params) { set_err("null argument"); return 2; }
    std::vector<float> p(params, params + MLP_NP);
    for (int o = 0; o < 64; ++o)
        for (int i = MLP_IN; i < 16; ++i) p[MLP_W1 + o * 16 + i] = 0.0f;
    for (int o = MLP_OUT; o < 16; ++o) {
        for (int i = 0; i < 64; ++i) p[MLP_W3 + o * 64 + i] = 0.0f;
        p[MLP_B3 + o] = 0.0f;
    }
    for (float v : p)
        if (!std::isfinite(v)) { set_err("non-finite weight"); return 2; }
    YZ_TRY(hipMemcpyAsync(S->mlp.params, p.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice, S->st));
    YZ_TRY(hipStreamSynchronize(S->st));
    return 0;
}

int yz_learn_sample(yz_step *S, float *ia, float *feat, float *mscale) {
    if (int e = learn_ctx(S)) return e;
    const Geo &g = S->g;
    const int np = g.nxp * g.nyp;
    // the Newton system at h = hnew, the step's current film: the explicit
    // half also at h (into scratch), so the state of the step is untouched
    // but for the pass-local iterate h
    YZ_TRY(hipMemcpyAsync(S->h, S->hnew, np * sizeof(double), hipMemcpyDeviceToDevice, S->st));
    if (b_before(S, S->h, S->dt_now, S->time)) return 2;
    RowArgs keep = S->ra;
    S->ra.ax = S->scratch;
    rows_launch(S, 0, S->h);
    S->ra = keep;
    S->ra.dt = S->dt_now; S->ra.axold = S->scratch; S->ra.hold = S->h;
    S->ra.ia32 = S->S32.A; S->ra.b32 = S->b32; S->ra.feat = S->feat; S->ra.mscale = S->mscale;
    YZ_TRY(hipMemsetAsync(S->dmax, 0, 4 * sizeof(unsigned long long), S->st));
    rows_launch(S, 1, S->h);
    S->ra = keep;
    YZ_TRY(hipGetLastError());
    const size_t N = g.N;
    if (ia) YZ_TRY(hipMemcpyAsync(ia, S->S32.A, g.ns * N * sizeof(float), hipMemcpyDeviceToHost, S->st));
    if (feat) YZ_TRY(hipMemcpyAsync(feat, S->feat, MLP_IN * N * sizeof(float), hipMemcpyDeviceToHost, S->st));
    if (mscale) YZ_TRY(hipMemcpyAsync(mscale, S->mscale, N * sizeof(float), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipStreamSynchronize(S->st));
    return 0;
}

int yz_learn_loss_grad(yz_step *S, int nprobe, unsigned seed, double *loss, float *grad) {
    if (int e = learn_ctx(S)) return e;
    if (nprobe < 1) { set_err("nprobe must be positive"); return 2; }
    double l = 0;
    if (mlp_loss_grad(S->mlp, S->S32, S->feat, S->mscale, nprobe, seed, &l, S->st)) {
        set_err("probe loss failed"); return 2;
    }
    if (grad) YZ_TRY(hipMemcpyAsync(grad, S->mlp.grads, MLP_NP * sizeof(float), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipStreamSynchronize(S->st));
    if (loss) *loss = l;
    return 0;
}

int yz_learn_train(yz_step *S, int iters, double lr, int nprobe, double *loss_first, double *loss_last) {
    if (int e = learn_ctx(S)) return e;
    if (iters < 1 || nprobe < 1 || !(lr >= 0)) { set_err("iters, nprobe >= 1 and lr >= 0 are needed"); return 2; }
    YZ_TRY(hipEventRecord(S->ev[8], S->st));
    if (mlp_train(S->mlp, S->S32, S->feat, S->mscale, iters, lr, nprobe, 0x51u + (unsigned)S->mlp.adam_t,
                  loss_first, loss_last, S->st)) {
        set_err("training failed"); return 2;
    }
    YZ_TRY(hipEventRecord(S->ev[9], S->st));
    YZ_TRY(hipEventSynchronize(S->ev[9]));
    S->mlp.train_ms += ev_ms(S->ev[8], S->ev[9]);
    return 0;
}

int yz_learn_apply(yz_step *S, float *M) {
    if (int e = learn_ctx(S)) return e;
    if (!M) { set_err("null argument"); return 2; }
    if (mlp_infer(S->mlp, S->g.N, S->feat, S->mscale, S->M32, S->st)) { set_err("inference failed"); return 2; }
    YZ_TRY(hipMemcpyAsync(M, S->M32, (size_t)S->g.ns * S->g.N * sizeof(float), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipStreamSynchronize(S->st));
    return 0;
}

int yz_learn_stats(yz_step *S, double *out5) {
    if (!S || !out5) { set_err("null argument"); return 2; }
    out5[0] = (double)S->fallbacks;
    out5[1] = (double)S->mlp.adam_t;
    out5[2] = S->mlp.train_ms;
    out5[3] = S->mlp.loss_first;
    out5[4] = S->mlp.loss_last;
    return 0;
}

}  // extern "C"

// ---- batched sweep ------------------------------------------------------
//
// Many independent films advanced one implicit step together. Each film is
// a step context on the batch's one stream, with its own fields, Newton
// state and FP32 solver; the network is the batch's, shared. The Newton
// passes of all films proceed in rounds: every round enqueues the
// assembly of every film still iterating, reads their max |Rhs| with one
// synchronisation, solves all their systems in one batched fused BiCGSTAB
// (one launch per kernel for every film still running), reads their
// corrections' max with one synchronisation, and applies each film's
// Newton decisions (NewtonKantorovich.cpp:157-422, as nk_loop; a film's
// numbers are the ones it gets alone). Rare paths (the stagnation floor,
// the FP64 fallback) run per film.

namespace {

enum { BF_START, BF_PASS, BF_END, BF_DONE };

struct BatchFilm {
    yz_step *S = nullptr;
    int phase = BF_START;
    bool retried = false, ok = false, guessed = false, extrap_pending = false;
    int outer = 0, stall = 0, status = 0, nk_out = 0;
    double resid_eq_old = 0, resid_first = 0, resid_best = 0, resid_eq = 0, eps = 0;
    long long lin_sum = 0;
};

}  // namespace

struct yz_batch {
    hipStream_t st = nullptr;
    std::vector<BatchFilm> f;
    Mlp mlp;                  // the shared network
    FusedTable table;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double ms_last = 0;
    long long passes_last = 0, lin_last = 0;
};

// start (or, with extrapolation blocked, restart) a film's Newton loop:
// the extrapolated guess when the step allows one; its check is read back
// with the round's synchronisation
static int bf_loop_start(BatchFilm &F) {
    yz_step *S = F.S;
    const yz_case &c = S->cs;
    const Geo &g = S->g;
    S->nk_pass = 0;
    S->lin_step = 0;
    S->extrap_move = 0;
    F.guessed = false;
    F.extrap_pending = false;
    F.outer = 0; F.stall = 0;
    F.resid_eq_old = 1e40; F.resid_first = 0; F.resid_best = 1e40; F.resid_eq = 0;
    if (c.extrap && !S->extrap_blocked && !S->first_step) {
        const double r = (S->dt_old > 0) ? S->dt_now / S->dt_old : 0.0;
        if (r > 0) {
            const int np = g.nxp * g.nyp;
            YZ_TRY(hipMemsetAsync(S->dmax, 0, 4 * sizeof(unsigned long long), S->st));
            YZ_TRY(hipMemsetAsync(S->dflag, 0, sizeof(int), S->st));
            hipLaunchKernelGGL(k_extrap, dim3(nb(np)), dim3(TB), 0, S->st, g, S->hold, S->hold2, r, S->h,
                               S->dmax + 2, S->dflag);
            YZ_TRY(hipMemcpyAsync(S->hmax, S->dmax, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, S->st));
            YZ_TRY(hipMemcpyAsync(S->hflag, S->dflag, sizeof(int), hipMemcpyDeviceToHost, S->st));
            F.extrap_pending = true;
        }
    }
    return 0;
}

// after the round's synchronisation: take or drop the guess, then the
// explicit half A(h^n) h^n (BeforeNKLoop)
static int bf_loop_begin(BatchFilm &F) {
    yz_step *S = F.S;
    const int np = S->g.nxp * S->g.nyp;
    if (F.extrap_pending) {
        const double move = bits_to_double(S->hmax[2]);
        if (!S->hflag[0] && move > 0) { S->extrap_move = move; F.guessed = true; }
    }
    if (!F.guessed) YZ_TRY(hipMemcpyAsync(S->h, S->hold, np * sizeof(double), hipMemcpyDeviceToDevice, S->st));
    if (b_before(S, S->hold, S->dt_now, S->time)) return 2;
    RowArgs keep = S->ra;
    S->ra.ax = S->axold;
    rows_launch(S, 0, S->hold);
    S->ra = keep;
    YZ_TRY(hipGetLastError());
    return 0;
}

// the assembly of a pass (FP32 planes, features, row scale), max |Rhs|
// read back with the round's synchronisation
static int bf_assemble(BatchFilm &F) {
    yz_step *S = F.S;
    S->ra.dt = S->dt_now;
    S->ra.axold = S->axold;
    S->ra.hold = S->hold;
    RowArgs keep = S->ra;
    S->ra.ia32 = S->S32.A; S->ra.b32 = S->b32; S->ra.feat = S->feat; S->ra.mscale = S->mscale;
    S->ra.nonunit = S->dflag;
    YZ_TRY(hipMemsetAsync(S->dmax, 0, 4 * sizeof(unsigned long long), S->st));
    YZ_TRY(hipMemsetAsync(S->dflag, 0, sizeof(int), S->st));
    rows_launch(S, 1, S->h);
    S->ra = keep;
    YZ_TRY(hipGetLastError());
    ++S->nk_pass;
    YZ_TRY(hipMemcpyAsync(S->hmax, S->dmax, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, S->st));
    YZ_TRY(hipMemcpyAsync(S->hflag, S->dflag, sizeof(int), hipMemcpyDeviceToHost, S->st));
    return 0;
}

// a film's Newton decisions after its pass (nk_loop's body after the
// solve); corrections' max already in hmax[1]. Sets F.phase.
static int bf_newton(BatchFilm &F, int ch, int lin_it) {
    yz_step *S = F.S;
    const yz_case &c = S->cs;
    const Geo &g = S->g;
    const int np = g.nxp * g.nyp;
    S->lin_step += lin_it;
    const double resid_eq = bits_to_double(S->hmax[1]);
    F.resid_eq = resid_eq;
    bool done = false, end = false;
    if (ch != 0) {
        end = true;
    } else {
        if (F.outer == 0) F.resid_first = resid_eq;
        if (resid_eq < F.resid_best * 0.9) { F.resid_best = resid_eq; F.stall = 0; }
        else F.stall++;
        if (F.stall > 5) {
            YZ_TRY(hipMemsetAsync(S->dmax + 3, 0, sizeof(unsigned long long), S->st));
            hipLaunchKernelGGL(k_absmax, dim3(std::min(nb(np), 1024)), dim3(TB), 0, S->st, np, S->h, S->dmax + 3);
            if (sync_read(S)) return 2;
            const double floorPhi = 1000.0 * std::numeric_limits<double>::epsilon() * bits_to_double(S->hmax[3]);
            done = (resid_eq < c.nk_eps) || (F.resid_first > 0 && resid_eq < F.resid_first * 1.0e-4) ||
                   (resid_eq <= floorPhi);
            end = true;
        } else {
            const double alpha = (resid_eq > F.resid_eq_old) ? 0.5 : 1.0;
            hipLaunchKernelGGL(k_update, dim3(nb(g.N)), dim3(TB), 0, S->st, g, S->h, S->corr, alpha);
            hipLaunchKernelGGL(k_ghosts, dim3(nb(np)), dim3(TB), 0, S->st, g, S->h);
            if (F.outer >= c.nk_min - 1) {
                if (resid_eq < c.nk_eps) { done = true; end = true; }
                else if (c.nk_eps_rel > 0 && F.resid_first > 0 && resid_eq < F.resid_first * c.nk_eps_rel) {
                    done = true; end = true;
                }
            }
            F.resid_eq_old = resid_eq;
        }
    }
    if (!end && F.outer + 1 >= c.nk_max) { end = true; F.outer++; }
    else if (!end) { F.outer++; return F.phase = BF_PASS, 0; }
    F.ok = done;
    F.nk_out = F.outer + 1;
    F.phase = BF_END;
    return 0;
}

// a film's Newton loop ended: retry without the guess, fail, or accept
// (TimeAdvance with a fixed step, NewtonKantorovich.cpp:424-486)
static int bf_loop_end(BatchFilm &F) {
    yz_step *S = F.S;
    const int np = S->g.nxp * S->g.nyp;
    F.lin_sum += S->lin_step;
    if (!F.ok && S->extrap_move > 0 && !F.retried) {
        F.retried = true;
        S->extrap_blocked = true;
        F.phase = BF_START;
        return 0;
    }
    S->extrap_blocked = false;
    if (!F.ok) {
        F.status = 1;
    } else {
        F.status = 0;
        YZ_TRY(hipMemcpyAsync(S->hnew, S->h, np * sizeof(double), hipMemcpyDeviceToDevice, S->st));
        S->time += S->dt_now;
        S->dt_old = S->dt_now;
        S->first_step = false;
    }
    F.phase = BF_DONE;
    return 0;
}

extern "C" {

yz_batch *yz_batch_create(int n, const yz_case *cases, const double *h, const double *s, const double *conc,
                          const double *dt0) {
    if (n < 1 || !cases || !h || !s || !conc || !dt0) { set_err("null argument or empty batch"); return nullptr; }
    for (int i = 0; i < n; ++i) {
        const yz_case &c = cases[i];
        if (c.dim != 2 || c.precond != YZ_PC_LEARNED) {
            set_err("a batch takes 2-D films with the learned preconditioner"); return nullptr;
        }
        if (c.step_control) { set_err("a batch advances fixed steps (step_control 0)"); return nullptr; }
        if (c.learn_passes != 0) { set_err("a batch uses one shared network: learn_passes 0 (yz_batch_set_params)"); return nullptr; }
        if (c.lin_maxit != cases[0].lin_maxit) { set_err("the films of a batch share lin_maxit"); return nullptr; }
        if (c.conc_model != 0) { set_err("a batch takes films with a frozen concentration (conc_model 0)"); return nullptr; }
        if (c.jac_interval != 1) { set_err("a batch forms IA every pass (jac_interval 1)"); return nullptr; }
    }
    yz_batch *B = new yz_batch();
    auto fail = [&](const char *what) { if (what) set_err(what); yz_batch_destroy(B); return (yz_batch *)nullptr; };
    if (hipStreamCreateWithFlags(&B->st, hipStreamNonBlocking) != hipSuccess) return fail("stream");
    for (auto &e : B->ev) if (hipEventCreate(&e) != hipSuccess) return fail("event");
    B->f.resize(n);
    size_t off = 0;
    int nmax = 0;
    for (int i = 0; i < n; ++i) {
        const yz_case &c = cases[i];
        const size_t np = (size_t)(c.nx + 4) * (c.ny + 4);
        B->f[i].S = step_new(&c, h + off, s + off, conc + off, dt0[i], B->st);
        if (!B->f[i].S) return fail(nullptr);
        off += np;
        nmax = std::max(nmax, c.nx * c.ny);
    }
    if (mlp_init(B->mlp, nmax, B->st)) return fail("network allocation");
    return B;
}

int yz_batch_set_params(yz_batch *B, const float *params) {
    if (!B || !params) { set_err("null argument"); return 2; }
    std::vector<float> p(params, params + MLP_NP);
    for (float v : p)
        if (!std::isfinite(v)) { set_err("non-finite weight"); return 2; }
    for (int o = 0; o < 64; ++o)
        for (int i = MLP_IN; i < 16; ++i) p[MLP_W1 + o * 16 + i] = 0.0f;
    for (int o = MLP_OUT; o < 16; ++o) {
        for (int i = 0; i < 64; ++i) p[MLP_W3 + o * 64 + i] = 0.0f;
        p[MLP_B3 + o] = 0.0f;
    }
    YZ_TRY(hipMemcpyAsync(B->mlp.params, p.data(), MLP_NP * sizeof(float), hipMemcpyHostToDevice, B->st));
    YZ_TRY(hipStreamSynchronize(B->st));
    return 0;
}

int yz_batch_advance(yz_batch *B, int *status, int *nk_iter, int *lin_iter) {
    if (!B) { set_err("null batch"); return 2; }
    const int n = (int)B->f.size();
    YZ_TRY(hipEventRecord(B->ev[0], B->st));
    for (BatchFilm &F : B->f) {
        yz_step *S = F.S;
        const int np = S->g.nxp * S->g.nyp;
        for (double &t : S->tms) t = 0;
        YZ_TRY(hipMemcpyAsync(S->hold2, S->hold, np * sizeof(double), hipMemcpyDeviceToDevice, B->st));
        YZ_TRY(hipMemcpyAsync(S->hold, S->hnew, np * sizeof(double), hipMemcpyDeviceToDevice, B->st));
        F.phase = BF_START; F.retried = false; F.ok = false; F.lin_sum = 0; F.status = 0; F.nk_out = 0;
    }
    long long passes = 0;
    std::vector<SolverT<float> *> sv;
    std::vector<const float *> bv;
    std::vector<float *> xv;
    std::vector<double> tol, resid;
    std::vector<int> its, rcs, who;
    for (;;) {
        // Newton loops that start (or restart) in this round
        bool any_start = false;
        for (BatchFilm &F : B->f)
            if (F.phase == BF_START) { if (bf_loop_start(F)) return 2; any_start = true; }
        if (any_start) YZ_TRY(hipStreamSynchronize(B->st));
        for (BatchFilm &F : B->f)
            if (F.phase == BF_START) { if (bf_loop_begin(F)) return 2; F.phase = BF_PASS; }
        // the passes of this round
        who.clear();
        for (int i = 0; i < n; ++i) if (B->f[i].phase == BF_PASS) who.push_back(i);
        if (who.empty()) break;
        for (int i : who) if (bf_assemble(B->f[i])) return 2;
        YZ_TRY(hipStreamSynchronize(B->st));
        const int m = (int)who.size();
        sv.resize(m); bv.resize(m); xv.resize(m); tol.resize(m); resid.resize(m); its.resize(m); rcs.resize(m);
        for (int u = 0; u < m; ++u) {
            BatchFilm &F = B->f[who[u]];
            yz_step *S = F.S;
            F.eps = ew_eps(S);
            S->S32.unit_diag = S->hflag[0] == 0;
            if (mlp_infer_half(B->mlp, S->g.N, S->feat, S->Yh, B->st)) { set_err("inference failed"); return 2; }
            S->S32.Mh = S->Yh;
            S->S32.Mscale = S->mscale;
            S->S32.stall_window = S->cs.stall_window;
            S->S32.use_graph = S->cs.graphs;
            sv[u] = &S->S32; bv[u] = S->b32; xv[u] = S->corr32;
            tol[u] = fp32_tol(S->cs, F.eps, F.resid_first, F.outer > 0 ? F.resid_eq : 0.0);
            if (zero_guess(S->cs)) YZ_TRY(hipMemsetAsync(S->corr32, 0, S->g.N * sizeof(float), B->st));
        }
        if (bicgstab_stencil_batch(B->table, m, sv.data(), bv.data(), xv.data(), B->f[who[0]].S->cs.lin_maxit,
                                   tol.data(), its.data(), resid.data(), rcs.data(), B->st)) {
            set_err("batched solve failed"); return 2;
        }
        passes += m;
        for (int u = 0; u < m; ++u) {
            BatchFilm &F = B->f[who[u]];
            yz_step *S = F.S;
            if (rcs[u] == 0) {
                hipLaunchKernelGGL(k_f2d, dim3(nb(S->g.N)), dim3(TB), 0, B->st, S->g.N, (const float *)S->corr32,
                                   S->corr);
            } else {
                int dev_err = 0;
                rcs[u] = fp64_fallback(S, F.eps, &its[u], &resid[u], &dev_err);
                if (dev_err) return 2;
                if (rcs[u] != 0) YZ_TRY(hipMemsetAsync(S->corr, 0, S->g.N * sizeof(double), B->st));
            }
            YZ_TRY(hipMemsetAsync(S->dmax + 1, 0, sizeof(unsigned long long), B->st));
            hipLaunchKernelGGL(k_absmax, dim3(std::min(nb(S->g.N), 1024)), dim3(TB), 0, B->st, S->g.N, S->corr,
                               S->dmax + 1);
            YZ_TRY(hipMemcpyAsync(S->hmax, S->dmax, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, B->st));
        }
        YZ_TRY(hipStreamSynchronize(B->st));
        for (int u = 0; u < m; ++u) {
            BatchFilm &F = B->f[who[u]];
            if (bf_newton(F, rcs[u], its[u])) return 2;
            if (F.phase == BF_END && bf_loop_end(F)) return 2;
        }
    }
    YZ_TRY(hipEventRecord(B->ev[1], B->st));
    YZ_TRY(hipEventSynchronize(B->ev[1]));
    B->ms_last = ev_ms(B->ev[0], B->ev[1]);
    B->passes_last = passes;
    long long lin = 0;
    int worst = 0;
    for (int i = 0; i < n; ++i) {
        const BatchFilm &F = B->f[i];
        if (status) status[i] = F.status;
        if (nk_iter) nk_iter[i] = F.nk_out;
        if (lin_iter) lin_iter[i] = (int)F.lin_sum;
        lin += F.lin_sum;
        worst = std::max(worst, F.status);
    }
    B->lin_last = lin;
    if (worst) set_err("the Newton iteration did not converge at the fixed step for some film");
    return worst;
}

int yz_batch_get_h(yz_batch *B, int film, double *h) {
    if (!B || !h || film < 0 || film >= (int)B->f.size()) { set_err("bad argument"); return 2; }
    return yz_step_get_h(B->f[film].S, h);
}

int yz_batch_stats(yz_batch *B, double *out4) {
    if (!B || !out4) { set_err("null argument"); return 2; }
    long long fb = 0;
    for (const BatchFilm &F : B->f) fb += F.S->fallbacks;
    out4[0] = B->ms_last;
    out4[1] = (double)B->passes_last;
    out4[2] = (double)B->lin_last;
    out4[3] = (double)fb;
    return 0;
}

void yz_batch_destroy(yz_batch *B) {
    if (!B) return;
    for (BatchFilm &F : B->f) if (F.S) yz_step_destroy(F.S);
    mlp_free(B->mlp);
    fused_table_free(B->table);
    for (auto &e : B->ev) if (e) (void)hipEventDestroy(e);
    if (B->st) (void)hipStreamDestroy(B->st);
    delete B;
}

}  // extern "C"
