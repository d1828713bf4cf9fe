/*
 * yz_oracle.cpp -- CPU FP64 restatement of Yuzen's well-mixed implicit
 * film-height step.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: the parity tests, the
 * smoke check and the CPU-baseline leg of the benchmark load it; the
 * product (yuzen_amd/csrc, libyuzen_hip.so) never includes, links or calls
 * it.
 *
 * It restates, in plain loops over the reference's own row layout (two
 * ghost layers, row = nxp*iy + ix, ghost rows kept in the system as
 * boundary-condition rows), the following reference functions. Each
 * function below cites the file:line it follows; paths are relative to the
 * reference root.
 *
 *   grid, coordinates      yuzen/src/model/Grids.cpp:88-110,
 *                          yuzen/src/grid/RectGrid.h:375-376,462-464
 *   ghost rows (Symmetry,  yuzen/src/model/BoundaryCondition.cpp:
 *   Periodic)              ApplyEdge, ApplyCorners, MirrorXY
 *   mobility               yuzen/src/model/Mobility.h:49-191
 *   limiter, face values   yuzen/src/numerics/FiniteDifferences.cpp:22-228
 *   capillary              LaplacePressureKernel.h:354-579,
 *                          WM2D/LaplacePressure_2D.cpp:39-128,
 *                          WM1D/LaplacePressure_1D.cpp:35-77
 *   vertical gravity       VerticalGravityKernel.h:150-433
 *   disjoining pressure    DisjoiningPressureKernel.h:150-435
 *   in-plane gravity,      BodyForceKernel.h (limited: LimitedFace,
 *   centrifugal            SetMatrixA_Limited, SetJacobiMatrix_Limited),
 *                          ParallelGravityKernel.h, CentrifugalKernel.h
 *   Marangoni              SolutoCapillarityKernel.h (limited forms)
 *   A, J, ApplyA           well_mixed/HeightEquationBase.h:669-827,1061-1091
 *   IA, Rhs, Newton loop   yuzen/src/numerics/NewtonKantorovich.cpp:74-523
 *   step control           yuzen/src/numerics/ImplicitSolver.cpp:144-307
 *   SatisfyBCRows          yuzen/src/numerics/CRS_Matrix.cpp:198-302
 *   ScaleRows              yuzen/src/numerics/LinearSolverBackend.cpp:227-252
 *   BiCGSTAB, SA-AMG,      AMGCL 1.4.4 (un-vendored third-party library,
 *   SPAI0, damped Jacobi   called at LinearSolverBackend.cpp:596-679):
 *                          restated from the published methods (van der
 *                          Vorst's BiCGSTAB with right preconditioning;
 *                          smoothed aggregation of Vanek, Mandel & Brezina
 *                          with constant near-null space, strength threshold
 *                          0.08 halved per level, prolongator smoothed with
 *                          the filtered matrix and omega = 4/3 / rho;
 *                          SPAI0 m_i = a_ii / sum_j a_ij^2; direct solve on
 *                          the coarsest level below 500 rows). The
 *                          aggregation order of AMGCL is not reproduced: the
 *                          hierarchy and the iteration counts are not pinned,
 *                          only the solution (to the solve's tolerance).
 *
 *   face volume fluxes     HeightEquationBase.h:852-894 and each term's
 *                          FaceFlux/SideRate (LaplacePressureKernel.h:480-
 *                          491, VerticalGravityKernel.h:271-287,
 *                          BodyForceKernel.h:118-159,
 *                          SolutoCapillarityKernel.h:104-138,
 *                          DisjoiningPressureKernel.h:160-168)
 *   well-mixed             HeightEquationBase.h:1314-1444 (BeforeNKLoop,
 *   concentration          EveryNKLoop), :1447-1580 (AfterNKLoop), :985-1058
 *                          (ConsistentFluxes), :1128-1273
 *                          (RedistributeExcess, MoveExcess), :1098-1109
 *                          (ClampConc), :1292-1305 (ConcCorrection);
 *                          ConcEquationBase.h:261-418 (Solve), :427-518
 *                          (SoluteFlux, Reconstruct); unlimited face
 *                          schemes FiniteDifferences.cpp:22-93, 152-170;
 *                          the Jacobi-then-multigrid solve of a
 *                          concentration system LinearSolverBackend.cpp:
 *                          405-470, 587-626 (Jacobi in place of SPAI0 for
 *                          the try: iteration counts are not pinned)
 *
 * Only the limited Koren/van Leer height schemes (the default,
 * Numerics.Convection.Height = "Koren") and the Symmetry/Periodic edges are
 * restated. Evaporation and the nozzle source enter B with the concentration
 * frozen (EvapAt HeightEquationBase.h:830-846, SetMatrixA's B :753-787,
 * Evaporation.cpp:62-103, SourceBase.cpp:25-112, WM2D/Source.cpp,
 * WM1D/Source_1D.cpp). The concentration is frozen
 * (Physics.ConcField.Model = "None") with conc_model 0 and solved with the
 * film (Model = "WM", unlimited face schemes, UTOPIA the default) with
 * conc_model 1.
 *
 * How each part is pinned (tests/test_oracle_golden.py,
 * tests/test_oracle_physics.py):
 *   mobility, limiter, Dp, every term's A and J, A(h)h: against the
 *     reference's own term code compiled from its sources
 *     (oracle/build_ref.sh, ref_golden.cpp), to 1e-12;
 *   IA, Rhs, the Newton loop and its exits, step control: against the
 *     exact discrete Crank-Nicolson amplification of linear leveling, the
 *     second-order space convergence and volume conservation; the
 *     reference's Newton loop cannot run here (its linear solve needs
 *     Eigen/AMGCL), so Newton trajectories are parity unpinned against the
 *     reference itself;
 *   BiCGSTAB / SA-AMG / Jacobi: the 1-D Poisson problem of the reference's
 *     backend test and the residual definition; AMG hierarchy and
 *     iteration counts: parity unpinned (only the solution to tolerance);
 *   B (evap_eval, source_at): against the reference's own C_Evaporation
 *     and Source_WM2D / Source_WM1D tables, to the bit; evap_at's caps, the
 *     theta split of B and its place in Rhs: against closed forms of flat
 *     films and the volume the supply adds
 *     (tests/test_oracle_evap_source.py).
 */
#include "../include/yuzen_gpu.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Row loops of the step's hot path (the term assembly, ApplyA, IA, the
 * right-hand side, the matrix-vector products, the multigrid cycle's and
 * BiCGSTAB's vector updates) run over OpenMP threads when the oracle is
 * built with -fopenmp. Every such loop writes only its own row and computes
 * it in the same order as the serial loop, so the results are the same to
 * the bit for any thread count; the sums over rows (dot products, norms,
 * the energy and balance reductions) stay serial, in row order. */
#define YZO_PAR_MIN 4096

namespace yzo {

using std::vector;

/* ------------------------------------------------------------------ */
/* Grid: Grids.cpp:88-110 (two ghost layers, spacing recomputed over the
 * padded extent), RectGrid.h:375-376 (X of a half-shifted site).       */
/* ------------------------------------------------------------------ */
struct Grid {
    int dim = 2, nx = 0, ny = 1, nxp = 0, nyp = 1, n = 0;
    double dx = 1, dy = 1, xlo = 0, ylo = 0;
    int bc[4] = {0, 0, 0, 0};

    void build(const yz_case &c) {
        dim = c.dim; nx = c.nx; ny = (dim == 2) ? c.ny : 1;
        for (int k = 0; k < 4; ++k) bc[k] = c.bc[k];
        const double dx0 = (c.xmax - c.xmin) / (double)nx;
        xlo = c.xmin - 2 * dx0;
        const double xhi = c.xmax + 2 * dx0;
        nxp = nx + 4;
        dx = (xhi - xlo) / nxp;
        if (dim == 2) {
            const double dy0 = (c.ymax - c.ymin) / (double)ny;
            ylo = c.ymin - 2 * dy0;
            const double yhi = c.ymax + 2 * dy0;
            nyp = ny + 4;
            dy = (yhi - ylo) / nyp;
        } else {
            nyp = 1; dy = 1; ylo = 0;
        }
        n = nxp * nyp;
    }
    int row(int ix, int iy) const { return nxp * iy + ix; }
    bool inside(int ix, int iy) const {
        if (ix < 2 || ix > nxp - 3) return false;
        if (dim == 2 && (iy < 2 || iy > nyp - 3)) return false;
        return true;
    }
    /* cell centre (Zface / Xedge: the x bit set) */
    double X(int ix) const { return xlo + ix * dx + dx * 0.5; }
    double Y(int iy) const { return ylo + iy * dy + dy * 0.5; }
    /* face coordinates (HalfXm keeps the index, HalfXp raises it) */
    double Xface(int ix) const { return xlo + ix * dx; }
    double Yface(int iy) const { return ylo + iy * dy; }

    /* the interior point a ghost index maps to along one axis
     * (BoundaryCondition.cpp ApplyEdge: Symmetry EqualVal(L0,L3),
     * EqualVal(L1,L2); Periodic EqualVal(L0,opp3), EqualVal(L1,opp2)) */
    static int fold(int i, int np, int lo_bc, int hi_bc) {
        const int m = np - 4;
        if (i < 2) return lo_bc == YZ_BC_PERIODIC ? i + m : 3 - i;
        if (i > np - 3) return hi_bc == YZ_BC_PERIODIC ? i - m : 2 * (np - 3) + 1 - i;
        return i;
    }
    /* the row a ghost row copies (corners fold along both axes, as
     * MirrorXY does) */
    int src(int ix, int iy) const {
        const int fx = fold(ix, nxp, bc[0], bc[1]);
        const int fy = (dim == 2) ? fold(iy, nyp, bc[2], bc[3]) : iy;
        return row(fx, fy);
    }
};

/* stencil slots: 2-D  oM mm om pm Mo mo oo po Po mp op pp oP  (FullIdx
 * order, GridTraits.h WM2D_Grid::FullIdx); 1-D  M m o p P            */
struct Stencil {
    int ns = 13, center = 6;
    int ox[13], oy[13];
    int adj[4];          /* AdjIdx(d): mo po om op */
    int face[4][4];      /* FaceStencil(d, j): LL L R RR */
    int axis[2][5];      /* AxisStencil(a, j) */
    int ndir = 4, naxis = 2;
    void build(int dim) {
        if (dim == 2) {
            static const int x[13] = {0, -1, 0, 1, -2, -1, 0, 1, 2, -1, 0, 1, 0};
            static const int y[13] = {-2, -1, -1, -1, 0, 0, 0, 0, 0, 1, 1, 1, 2};
            ns = 13; center = 6; ndir = 4; naxis = 2;
            for (int k = 0; k < 13; ++k) { ox[k] = x[k]; oy[k] = y[k]; }
            enum { oM, mm, om, pm, Mo, mo, oo, po, Po, mp, op, pp, oP };
            adj[0] = mo; adj[1] = po; adj[2] = om; adj[3] = op;
            const int f[4][4] = {{Mo, mo, oo, po}, {mo, oo, po, Po},
                                 {oM, om, oo, op}, {om, oo, op, oP}};
            std::memcpy(face, f, sizeof(f));
            const int a[2][5] = {{Mo, mo, oo, po, Po}, {oM, om, oo, op, oP}};
            std::memcpy(axis, a, sizeof(a));
        } else {
            ns = 5; center = 2; ndir = 2; naxis = 1;
            for (int k = 0; k < 5; ++k) { ox[k] = k - 2; oy[k] = 0; }
            adj[0] = 1; adj[1] = 3;
            const int f[2][4] = {{0, 1, 2, 3}, {1, 2, 3, 4}};
            for (int d = 0; d < 2; ++d) for (int j = 0; j < 4; ++j) face[d][j] = f[d][j];
            for (int j = 0; j < 5; ++j) axis[0][j] = j;
        }
    }
    /* StencilDir(k): Ym, Xm, Xp, Yp in 2-D (GridTraits.h WM2D_Grid) */
    int sdir(int k) const { return ndir == 4 ? (k == 0 ? 2 : k == 1 ? 0 : k == 2 ? 1 : 3) : k; }
    int dir_of_adj(int slot) const { for (int d = 0; d < ndir; ++d) if (adj[d] == slot) return d; return -1; }
    bool in_face(int d, int slot) const { for (int j = 0; j < 4; ++j) if (face[d][j] == slot) return true; return false; }
};

/* CRS with the reference's storage rule: the diagonal first
 * (CRS_Matrix.h:26-29, 59-68). Interior rows hold the stencil slots in
 * slot order after the diagonal; a ghost row holds (diag, source). */
struct Csr {
    int n = 0;
    vector<int> ptr, col;
    vector<double> val;
    vector<char> bc;
    vector<int> slotpos;   /* interior row: position of slot k, -1 ghost */
    int ns = 0;
    double &at(int row, int slot) { return val[slotpos[(size_t)row * ns + slot]]; }
    void clear_nonbc() {
        for (int i = 0; i < n; ++i)
            if (!bc[i]) for (int k = ptr[i]; k < ptr[i + 1]; ++k) val[k] = 0;
    }
    void matvec(const vector<double> &x, vector<double> &y) const {
        y.assign(n, 0.0);
#pragma omp parallel for schedule(static) if (n > YZO_PAR_MIN)
        for (int i = 0; i < n; ++i) {
            double s = 0;
            for (int k = ptr[i]; k < ptr[i + 1]; ++k) s += val[k] * x[col[k]];
            y[i] = s;
        }
    }
};

/* ------------------------------------------------------------------ */
/* Mobility (Mobility.h:49-191)                                         */
/* ------------------------------------------------------------------ */
struct Mobility {
    double lambda = 0, a = 0;
    bool slip = false;
    static constexpr double kSeriesRel = 1e-3;
    void set(double l) { lambda = l > 0 ? l : 0.0; a = 3.0 * lambda; slip = lambda > 0; }
    double D(double h) const { const double u = std::fabs(h); return h * u * (u + a); }
    double dD(double h) const { const double u = std::fabs(h); return u * (3.0 * u + 2.0 * a); }
    double DperH(double h) const { const double u = std::fabs(h); return u * (u + a); }
    double dDperH(double h) const { return h < 0 ? -(2.0 * (-h) + a) : 2.0 * h + a; }
    double Face(double h1, double h2) const {
        if (slip) return FaceDK(h1, h2);
        if (!(h1 > 0) || !(h2 > 0)) return 0;
        return 2.0 * (h1 * h1) * (h2 * h2) / (h1 + h2);
    }
    double Face_d1(double h1, double h2) const {
        if (slip) return FaceDK_d1(h1, h2);
        if (!(h1 > 0) || !(h2 > 0)) return 0;
        return 2.0 * (h1 * h2 * h2) * (h1 + h2 + h2) / ((h1 + h2) * (h1 + h2));
    }
    double Upwind(double hf) const { return DperH(hf); }
    double dUpwind(double hf) const { return dDperH(hf); }
    double Shear(double hf) const { return hf + 2.0 * lambda; }
    double G(double h) const {
        const double x = a / h;
        double f;
        if (x < 0.1) {
            double t = x * x, s = 0;
            for (int k = 2; k <= 18; ++k) { s += ((k % 2 == 0) ? t : -t) / k; t *= x; }
            f = s;
        } else f = x - std::log1p(x);
        return -f / (a * a);
    }
    double K(double h) const { const double u = dD(h) / D(h); return 2.0 * u * u - (6.0 * h + 2.0 * a) / D(h); }
    double dK(double h) const {
        const double d = D(h), d1 = dD(h), d2 = 6.0 * h + 2.0 * a, d3 = 6.0;
        const double u = d1 / d, du = d2 / d - u * u;
        return 4.0 * u * du - (d3 * d - d2 * d1) / (d * d);
    }
    double FaceDK(double h1, double h2) const {
        if (!(h1 > 0) || !(h2 > 0)) return 0;
        const double hm = 0.5 * (h1 + h2), dl = h2 - h1;
        if (std::fabs(dl) < kSeriesRel * hm) return D(hm) / (1.0 + dl * dl * K(hm) / 24.0);
        return dl / (G(h2) - G(h1));
    }
    double FaceDK_d1(double h1, double h2) const {
        if (!(h1 > 0) || !(h2 > 0)) return 0;
        const double hm = 0.5 * (h1 + h2), dl = h2 - h1;
        if (std::fabs(dl) < kSeriesRel * hm) {
            const double k = K(hm), q = 1.0 + dl * dl * k / 24.0;
            const double Fhm = (dD(hm) * q - D(hm) * dl * dl * dK(hm) / 24.0) / (q * q);
            const double Fdl = -D(hm) * (dl * k / 12.0) / (q * q);
            return 0.5 * Fhm - Fdl;
        }
        const double dg = G(h2) - G(h1);
        return (-dg + dl / D(h1)) / (dg * dg);
    }
};

/* ------------------------------------------------------------------ */
/* Limited face interpolation (FiniteDifferences.cpp, Koren / VanLeer)  */
/* ------------------------------------------------------------------ */
struct Limiter {
    int vanleer = 0;  /* 0 Koren, 1 van Leer */
    double phi(double a, double b, double *dA = nullptr, double *dB = nullptr) const {
        double p = 0, pa = 0, pb = 0;
        const bool same = (a > 0 && b > 0) || (a < 0 && b < 0);
        if (same) {
            if (!vanleer) {
                const double s = (b > 0) ? 1.0 : -1.0;
                const double aa = std::fabs(a), ab = std::fabs(b);
                const double t1 = 2.0 * aa, t2 = (2.0 * ab + aa) / 3.0, t3 = 2.0 * ab;
                if (t1 <= t2 && t1 <= t3) { p = s * t1; pa = 2.0; pb = 0.0; }
                else if (t2 <= t3) { p = s * t2; pa = 1.0 / 3.0; pb = 2.0 / 3.0; }
                else { p = s * t3; pa = 0.0; pb = 2.0; }
            } else {
                const double sum = a + b, ra = a / sum, rb = b / sum;
                p = 2.0 * a * rb; pa = 2.0 * rb * rb; pb = 2.0 * ra * ra;
            }
        }
        if (dA) *dA = pa;
        if (dB) *dB = pb;
        return p;
    }
    static void pos(double sign, int &uu, int &u, int &d) {
        if (sign >= 0) { uu = 0; u = 1; d = 2; } else { uu = 3; u = 2; d = 1; }
    }
    double value(const double f[4], double sign) const {
        int uu, u, d; pos(sign, uu, u, d);
        return f[u] + 0.5 * phi(f[u] - f[uu], f[d] - f[u]);
    }
    /* downwind form (rowPos = -1 in FaceWeights) */
    void weights(const double f[4], double sign, double w[4]) const { weights_at(f, sign, -1, w); }
    /* FaceWeights (FiniteDifferences.cpp:166-190): the face value as weights
     * of the four points, frozen at f; the row's own cell at position rowpos
     * (LL 0, L 1, R 2, RR 3). Upwind cell of the row: f_U + chi/2 (f_U - f_UU),
     * chi = phi / (f_U - f_UU); otherwise f_U + psi/2 (f_D - f_U),
     * psi = phi / (f_D - f_U). Either keeps the frozen matrix an M-matrix. */
    void weights_at(const double f[4], double sign, int rowpos, double w[4]) const {
        int uu, u, d; pos(sign, uu, u, d);
        const double a = f[u] - f[uu], b = f[d] - f[u];
        const double p = phi(a, b);
        for (int j = 0; j < 4; ++j) w[j] = 0;
        if (rowpos == u) {
            const double chi = (a != 0) ? p / a : 0.0;
            w[u] = 1.0 + 0.5 * chi;
            w[uu] = -0.5 * chi;
        } else {
            const double psi = (b != 0) ? p / b : 0.0;
            w[u] = 1.0 - 0.5 * psi;
            w[d] = 0.5 * psi;
        }
    }
    double derivative(const double f[4], double sign, double g[4]) const {
        int uu, u, d; pos(sign, uu, u, d);
        double pa, pb;
        const double p = phi(f[u] - f[uu], f[d] - f[u], &pa, &pb);
        for (int j = 0; j < 4; ++j) g[j] = 0;
        g[uu] = -0.5 * pa;
        g[u] = 1.0 + 0.5 * (pa - pb);
        g[d] = 0.5 * pb;
        return f[u] + 0.5 * p;
    }
};

static inline double Sign(double v) { return v >= 0.0 ? 1.0 : -1.0; }

/* The fixed coefficients of the unlimited four-point schemes
 * (C_Interpolation_base4::SetCoef, FiniteDifferences.cpp:22-73): c1 and c2
 * of LL L R RR over the sum of c1, for a YZ_CONV_* scheme (UTOPIA, the
 * reference's Upwind_3rd, for any other value) */
static void unlimited_coef(int conv, double c1[4], double c2[4]) {
    double a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
    switch (conv) {
    case YZ_CONV_UPWIND1: a[1] = 1; a[2] = 1; b[1] = 1; b[2] = -1; break;
    case YZ_CONV_CENTRAL2: a[1] = 1; a[2] = 1; break;
    case YZ_CONV_CENTRAL4: a[0] = -1; a[1] = 9; a[2] = 9; a[3] = -1; break;
    case YZ_CONV_UPWIND2: a[0] = -1; a[1] = 3; a[2] = 3; a[3] = -1; b[0] = -1; b[1] = 3; b[2] = -3; b[3] = 1; break;
    case YZ_CONV_QUICK: a[0] = -1; a[1] = 9; a[2] = 9; a[3] = -1; b[0] = -1; b[1] = 3; b[2] = -3; b[3] = 1; break;
    case YZ_CONV_KK: a[0] = -1; a[1] = 7; a[2] = 7; a[3] = -1; b[0] = -3; b[1] = 9; b[2] = -9; b[3] = 3; break;
    default: a[0] = -1; a[1] = 7; a[2] = 7; a[3] = -1; b[0] = -1; b[1] = 3; b[2] = -3; b[3] = 1; break;
    }
    double den = 0;
    for (int j = 0; j < 4; ++j) den += a[j];
    for (int j = 0; j < 4; ++j) { c1[j] = a[j] / den; c2[j] = b[j] / den; }
}
/* the YZ_CONV_* scheme of an unlimited YZ_HSCHEME_* height scheme */
static int height_conv(int hs) {
    switch (hs) {
    case YZ_HSCHEME_UPWIND1: return YZ_CONV_UPWIND1;
    case YZ_HSCHEME_CENTRAL2: return YZ_CONV_CENTRAL2;
    case YZ_HSCHEME_CENTRAL4: return YZ_CONV_CENTRAL4;
    case YZ_HSCHEME_UPWIND2: return YZ_CONV_UPWIND2;
    case YZ_HSCHEME_QUICK: return YZ_CONV_QUICK;
    case YZ_HSCHEME_KK: return YZ_CONV_KK;
    default: return YZ_CONV_UTOPIA;
    }
}

/* ------------------------------------------------------------------ */
/* The film-height equation                                             */
/* ------------------------------------------------------------------ */
struct Equation {
    yz_case cs;
    Grid g;
    Stencil st;
    Mobility mob;
    Limiter lim;
    /* Numerics.Convection.Height: limited (Koren, van Leer: lim) or one of
     * the fixed-coefficient schemes (hc1, hc2 of LL L R RR) */
    bool hlim = true;
    double hc1[4] = {0, 0, 0, 0}, hc2[4] = {0, 0, 0, 0};
    vector<double> h, s, c;          /* the working field (guess) */
    vector<double> Ccap[4];          /* capillary C(d)[slot] */
    vector<double> dFace[4];         /* capillary D of the last SetMatrixA */
    Csr A, Arest, J, IA;

    double visc(double cc) const { return std::exp(cs.visc_exp * (cc - cs.visc_c0)); }
    double sigma(double cc) const { return 1.0 + (cs.marangoni ? cs.Ma : 0.0) * (cc - cs.sig_c0); }
    double spacing(int axis) const { return axis == 0 ? g.dx : g.dy; }
    int nb(int i, int slot) const {
        const int ix = i % g.nxp, iy = i / g.nxp;
        return g.row(ix + st.ox[slot], iy + st.oy[slot]);
    }
    bool ghost_row(int i) const { return !g.inside(i % g.nxp, i / g.nxp); }

    void build_pattern(Csr &m) {
        m.n = g.n; m.ns = st.ns;
        m.ptr.assign(g.n + 1, 0); m.col.clear(); m.val.clear();
        m.bc.assign(g.n, 0); m.slotpos.assign((size_t)g.n * st.ns, -1);
        for (int i = 0; i < g.n; ++i) {
            const int ix = i % g.nxp, iy = i / g.nxp;
            if (!g.inside(ix, iy)) {
                m.bc[i] = 1;
                m.col.push_back(i); m.val.push_back(1.0);
                const int sr = g.src(ix, iy);
                m.col.push_back(sr); m.val.push_back(-1.0);
            } else {
                m.slotpos[(size_t)i * st.ns + st.center] = (int)m.col.size();
                m.col.push_back(i); m.val.push_back(0.0);
                for (int k = 0; k < st.ns; ++k) {
                    if (k == st.center) continue;
                    m.slotpos[(size_t)i * st.ns + k] = (int)m.col.size();
                    m.col.push_back(nb(i, k)); m.val.push_back(0.0);
                }
            }
            m.ptr[i + 1] = (int)m.col.size();
        }
    }

    void build(const yz_case &c_) {
        cs = c_;
        g.build(cs);
        st.build(cs.dim);
        mob.set(cs.slip);
        lim.vanleer = cs.height_scheme == YZ_HSCHEME_VANLEER ? 1 : 0;
        hlim = cs.height_scheme == YZ_HSCHEME_KOREN || cs.height_scheme == YZ_HSCHEME_VANLEER;
        if (!hlim) unlimited_coef(height_conv(cs.height_scheme), hc1, hc2);
        build_pattern(A); Arest = A; J = A; IA = A;
        /* capillary C(d): LaplacePressure_2D.cpp:39-94 / _1D.cpp:35-77 */
        for (int d = 0; d < st.ndir; ++d) Ccap[d].assign(st.ns, 0.0);
        if (cs.dim == 2) {
            enum { oM, mm, om, pm, Mo, mo, oo, po, Po, mp, op, pp, oP };
            const double dx = g.dx, dy = g.dy, dx2 = dx * dx, dy2 = dy * dy;
            const double dx4 = dx2 * dx2, dy4 = dy2 * dy2, dxy = dx2 * dy2;
            vector<double> *C = Ccap;
            C[0][mm] = +1.0 / dxy; C[0][om] = -1.0 / dxy; C[0][Mo] = +1.0 / dx4;
            C[0][mo] = 1.0 * (-3.0 / dx4 - 2.0 / dxy); C[0][oo] = 1.0 * (+3.0 / dx4 + 2.0 / dxy);
            C[0][po] = -1.0 / dx4; C[0][mp] = +1.0 / dxy; C[0][op] = -1.0 / dxy;
            C[1][om] = -1.0 / dxy; C[1][pm] = +1.0 / dxy; C[1][mo] = -1.0 / dx4;
            C[1][oo] = 1.0 * (+3.0 / dx4 + 2.0 / dxy); C[1][po] = 1.0 * (-3.0 / dx4 - 2.0 / dxy);
            C[1][Po] = +1.0 / dx4; C[1][op] = -1.0 / dxy; C[1][pp] = +1.0 / dxy;
            C[2][oM] = +1.0 / dy4; C[2][mm] = +1.0 / dxy; C[2][om] = 1.0 * (-3.0 / dy4 - 2.0 / dxy);
            C[2][pm] = +1.0 / dxy; C[2][mo] = -1.0 / dxy; C[2][oo] = 1.0 * (+3.0 / dy4 + 2.0 / dxy);
            C[2][po] = -1.0 / dxy; C[2][op] = -1.0 / dy4;
            C[3][om] = -1.0 / dy4; C[3][mo] = -1.0 / dxy; C[3][oo] = 1.0 * (+3.0 / dy4 + 2.0 / dxy);
            C[3][po] = -1.0 / dxy; C[3][mp] = +1.0 / dxy; C[3][op] = 1.0 * (-3.0 / dy4 - 2.0 / dxy);
            C[3][pp] = +1.0 / dxy; C[3][oP] = +1.0 / dy4;
        } else {
            const double dx = g.dx, dx2 = dx * dx, dx4 = dx2 * dx2;
            Ccap[0][0] = +1.0 / dx4; Ccap[0][1] = 1.0 * (-3.0 / dx4); Ccap[0][2] = 1.0 * (+3.0 / dx4); Ccap[0][3] = -1.0 / dx4;
            Ccap[1][1] = -1.0 / dx4; Ccap[1][2] = 1.0 * (+3.0 / dx4); Ccap[1][3] = 1.0 * (-3.0 / dx4); Ccap[1][4] = +1.0 / dx4;
        }
        for (int d = 0; d < st.ndir; ++d) dFace[d].assign(g.n, 0.0);
    }

    void fill_ghosts(vector<double> &v) const {
        for (int i = 0; i < g.n; ++i) {
            const int ix = i % g.nxp, iy = i / g.nxp;
            if (!g.inside(ix, iy)) v[i] = v[g.src(ix, iy)];
        }
    }

    /* FaceGradient: LaplacePressure_2D.cpp:105-128, _1D.cpp */
    double face_gradient(const double *v, int d) const {
        if (cs.dim == 2) {
            enum { oM, mm, om, pm, Mo, mo, oo, po, Po, mp, op, pp, oP };
            const double dx2 = g.dx * g.dx, dy2 = g.dy * g.dy;
            const double dn4 = (d < 2) ? dx2 * dx2 : dy2 * dy2, dxy = dx2 * dy2;
            double a1, a, b, b1, am, ap, bm, bp;
            switch (d) {
            case 0: a1 = v[Mo]; a = v[mo]; b = v[oo]; b1 = v[po]; am = v[mm]; ap = v[mp]; bm = v[om]; bp = v[op]; break;
            case 1: a1 = v[mo]; a = v[oo]; b = v[po]; b1 = v[Po]; am = v[om]; ap = v[op]; bm = v[pm]; bp = v[pp]; break;
            case 2: a1 = v[oM]; a = v[om]; b = v[oo]; b1 = v[op]; am = v[mm]; ap = v[pm]; bm = v[mo]; bp = v[po]; break;
            default: a1 = v[om]; a = v[oo]; b = v[op]; b1 = v[oP]; am = v[mo]; ap = v[po]; bm = v[mp]; bp = v[pp]; break;
            }
            const double normal = (((b1 - b) - 2.0 * (b - a)) + (a - a1)) / dn4;
            const double cross = (((bp - b) - (b - bm)) - ((ap - a) - (a - am))) / dxy;
            return normal + cross;
        }
        const double dx2 = g.dx * g.dx, dx4 = dx2 * dx2;
        double a1, a, b, b1;
        if (d == 0) { a1 = v[0]; a = v[1]; b = v[2]; b1 = v[3]; }
        else { a1 = v[1]; a = v[2]; b = v[3]; b1 = v[4]; }
        return (((b1 - b) - 2.0 * (b - a)) + (a - a1)) / dx4;
    }

    /* capillary face mobility, NonlinearTerm (LaplacePressureKernel.h:242-270) */
    double cap_D(int i, int d) const {
        const int a = nb(i, st.adj[d]);
        const double Hoo = h[i], Haj = h[a];
        const double cf = 0.5 * (c[i] + c[a]);
        const double val = mob.slip ? mob.FaceDK(Hoo, Haj) : 2.0 * (Hoo * Hoo) * (Haj * Haj) / (Hoo + Haj);
        return val * sigma(cf) / visc(cf);
    }
    /* NonlinearJacobi (:279-322): derivative at the centre (cen) or the
     * neighbour of face d */
    double cap_dD(int i, int d, bool cen) const {
        const int a = nb(i, st.adj[d]);
        const double H1 = cen ? h[i] : h[a], H2 = cen ? h[a] : h[i];
        const double cf = 0.5 * (c[i] + c[a]);
        const double den = std::pow(H1 + H2, 2.0);
        const double val = mob.slip ? mob.FaceDK_d1(H1, H2) : 2.0 * (H1 * H2 * H2) * (H1 + H2 + H2) / den;
        return val * sigma(cf) / visc(cf);
    }

    /* ---- SetMatrixA of each term ---- */
    void capillary_A(Csr &M) {
        const double Coef = 1.0 / 3.0;
#pragma omp parallel for schedule(static) if (g.n > YZO_PAR_MIN)
        for (int i = 0; i < g.n; ++i) {
            if (M.bc[i]) continue;
            double D[4], Tf[4];
            for (int d = 0; d < st.ndir; ++d) { D[d] = cap_D(i, d); Tf[d] = D[d] / h[i]; dFace[d][i] = D[d]; }
            double topo = 0;
            for (int k = 0; k < st.ns; ++k) {
                double cc = 0;
                for (int d = 0; d < st.ndir; ++d) cc += Ccap[d][k] * Tf[d];
                topo += cc * s[nb(i, k)];
            }
            for (int k = 0; k < st.ns; ++k) {
                double v = 0;
                for (int d = 0; d < st.ndir; ++d) v += Ccap[d][k] * D[d];
                if (k == st.center) { v += topo; M.at(i, k) += Coef * v; }
                else M.at(i, k) += Coef * v;
            }
        }
    }
    /* narrow C(d) of the second-difference terms */
    double narrowC(int d, int slot) const {
        const double hh = spacing(d / 2), h2 = hh * hh;
        if (slot == st.adj[d]) return 1.0 / h2;
        if (slot == st.center) return -1.0 / h2;
        return 0.0;
    }
    double vg_G(int i, int d) const {
        const int a = nb(i, st.adj[d]);
        return mob.Face(h[i], h[a]) / visc(0.5 * (c[i] + c[a]));
    }
    double vg_T(int i, int d) const {
        const int a = nb(i, st.adj[d]);
        const double ho = h[i], ha = h[a];
        const double gg = mob.Face(ho, ha);
        const double mu = visc(0.5 * (c[i] + c[a]));
        return (gg == 0) ? 0.0 : gg / (ho + ha) / mu;
    }
    double vg_dG(int i, int d, bool cen) const {
        const int a = nb(i, st.adj[d]);
        const double ho = h[i], ha = h[a];
        const double mu = visc(0.5 * (c[i] + c[a]));
        return (cen ? mob.Face_d1(ho, ha) : mob.Face_d1(ha, ho)) / mu;
    }
    void vgrav_A(Csr &M) {
        const double Coef = -1.0 * cs.Bo / 3.0;
#pragma omp parallel for schedule(static) if (g.n > YZO_PAR_MIN)
        for (int i = 0; i < g.n; ++i) {
            if (M.bc[i]) continue;
            double G[4], Tface[4], Tsum[4];
            for (int d = 0; d < st.ndir; ++d) { G[d] = vg_G(i, d); Tface[d] = vg_T(i, d); }
            for (int d = 0; d < st.ndir; ++d) {
                Tsum[d] = 0;
                for (int k = 0; k < st.ns; ++k) Tsum[d] += narrowC(d, k) * s[nb(i, k)];
            }
            double Too = 0;
            for (int k = 0; k < st.ndir; ++k) { const int dd = st.sdir(k); Too += Tsum[dd] * Tface[dd]; }
            for (int d = 0; d < st.ndir; ++d) Tsum[d] *= Tface[d];
            /* NarrowIdx order: m o p (1-D) / om mo oo po op (2-D) */
            for (int k = 0; k < st.ns; ++k) {
                const bool isnarrow = (k == st.center) || st.dir_of_adj(k) >= 0;
                if (!isnarrow) continue;
                double v = 0;
                for (int d = 0; d < st.ndir; ++d) v += narrowC(d, k) * G[d];
                if (k == st.center) v += Too; else v += Tsum[st.dir_of_adj(k)];
                M.at(i, k) += Coef * v;
            }
        }
    }
    double Pi(double hh) const {
        if (hh <= 0) return 0;
        const double r = cs.hstar / hh;
        return std::pow(r, cs.dis_n) - std::pow(r, cs.dis_m);
    }
    double dPi(double hh) const {
        if (hh <= 0) return 0;
        const double r = cs.hstar / hh;
        return -(cs.dis_n * std::pow(r, cs.dis_n) - cs.dis_m * std::pow(r, cs.dis_m)) / hh;
    }
    double PiPerH(double hh) const { if (hh <= 0) return 0; return Pi(hh) / hh; }
    void disjoining_A(Csr &M) {
        const double Coef = cs.Dp / 3.0;
#pragma omp parallel for schedule(static) if (g.n > YZO_PAR_MIN)
        for (int i = 0; i < g.n; ++i) {
            if (M.bc[i]) continue;
            double G[4];
            for (int d = 0; d < st.ndir; ++d) G[d] = vg_G(i, d);
            for (int k = 0; k < st.ns; ++k) {
                const bool isnarrow = (k == st.center) || st.dir_of_adj(k) >= 0;
                if (!isnarrow) continue;
                const double php = PiPerH(h[nb(i, k)]);
                double v = 0;
                for (int d = 0; d < st.ndir; ++d) v += narrowC(d, k) * G[d];
                v *= php;
                M.at(i, k) += Coef * v;
            }
        }
    }

    /* wall faces of the upwinded terms (GridTraits.h WallFace) */
    bool wall_face(int i, int d) const {
        const int ix = i % g.nxp, iy = i / g.nxp;
        const int Nx = g.nxp, Ny = g.nyp;
        auto isw = [](int t) { return t == YZ_BC_SYMMETRY || t == YZ_BC_PINNING; };
        if ((d == 0 && ix == 2) || (d == 1 && ix == 1)) return isw(g.bc[0]);
        if ((d == 1 && ix == Nx - 3) || (d == 0 && ix == Nx - 2)) return isw(g.bc[1]);
        if (cs.dim == 2) {
            if ((d == 2 && iy == 2) || (d == 3 && iy == 1)) return isw(g.bc[2]);
            if ((d == 3 && iy == Ny - 3) || (d == 2 && iy == Ny - 2)) return isw(g.bc[3]);
        }
        return false;
    }
    void face_heights(int i, int d, double f[4]) const {
        for (int j = 0; j < 4; ++j) f[j] = h[nb(i, st.face[d][j])];
    }

    /* body forces: mode 0 in-plane gravity, 1 centrifugal */
    double bf_drive_face(int mode, int i, int d) const {
        if (mode == 0) return Sign(d / 2 == 0 ? cs.Pgx : cs.Pgy);
        const int ix = i % g.nxp, iy = i / g.nxp;
        switch (d) {
        case 0: return g.Xface(ix) - cs.rot_x;
        case 1: return g.Xface(ix + 1) - cs.rot_x;
        case 2: return g.Yface(iy) - cs.rot_y;
        default: return g.Yface(iy + 1) - cs.rot_y;
        }
    }
    double bf_weight(int mode, int i, int d) const { return mode == 0 ? 1.0 : bf_drive_face(mode, i, d); }
    double bf_mag(int mode, int axis) const {
        if (mode == 0) return (axis == 0 ? cs.Pgx : cs.Pgy) / 3.0;
        return cs.Sp / 3.0;
    }
    /* LimitedFace (BodyForceKernel.h) */
    void bf_face(int mode, int i, int j, double &hf, double &D, double wt[4], double *gd) const {
        const double dv = bf_drive_face(mode, i, j);
        const double hs = spacing(j / 2);
        double f[4];
        face_heights(i, j, f);
        if (gd) { hf = lim.derivative(f, dv, gd); for (int q = 0; q < 4; ++q) gd[q] /= hs; }
        else hf = lim.value(f, dv);
        lim.weights(f, dv, wt);
        for (int q = 0; q < 4; ++q) wt[q] /= hs;
        const int a = nb(i, st.adj[j]);
        const double cc = 0.5 * (c[i] + c[a]);
        D = (mob.slip ? mob.Upwind(hf) : hf * hf) / visc(cc);
        if (wall_face(i, j)) D = 0;
    }
    /* ---- the fixed-coefficient schemes of the body forces
     * (BodyForceKernel.h:263-545): the face value sum_j (c1_j + Sign c2_j)
     * f_j upwinded by Drive for h_f^2/mu, and the same coefficients per
     * spacing, upwinded by DriveFace, as A's weights ---- */
    /* Drive (ParallelGravityKernel.h:83-88, CentrifugalKernel.h:88-93): the
     * centrifugal arm from the mean of the two cells' coordinates
     * (CellAvgCoord, GridTraits.h:219-220, 488-498) */
    double bf_drive(int mode, int i, int d) const {
        if (mode == 0) return d / 2 == 0 ? cs.Pgx : cs.Pgy;
        if (cs.dim == 1) return bf_drive_face(mode, i, d);
        const int ix = i % g.nxp, iy = i / g.nxp;
        switch (d) {
        case 0: return 0.5 * (g.X(ix - 1) + g.X(ix)) - cs.rot_x;
        case 1: return 0.5 * (g.X(ix + 1) + g.X(ix)) - cs.rot_x;
        case 2: return 0.5 * (g.Y(iy - 1) + g.Y(iy)) - cs.rot_y;
        default: return 0.5 * (g.Y(iy + 1) + g.Y(iy)) - cs.rot_y;
        }
    }
    /* the fixed face value of face d by the sign of pp (FaceStencil order) */
    double bf_face_value(int i, int d, double pp) const {
        double hh = 0;
        for (int j = 0; j < 4; ++j) hh += h[nb(i, st.face[d][j])] * (hc1[j] + Sign(pp) * hc2[j]);
        return hh;
    }
    /* NonlinearTerm (:263-302): h_f^2 / mu(c_f), 0 on a wall face */
    double bf_nonlinear(int mode, int i, int d) const {
        const int a = nb(i, st.adj[d]);
        const double cc = 0.5 * (c[i] + c[a]);
        if (wall_face(i, d)) return 0;
        const double hh = bf_face_value(i, d, bf_drive(mode, i, d));
        const double val = mob.slip ? mob.Upwind(hh) : hh * hh;
        return val / visc(cc);
    }
    /* NonlinearJacobi (:305-336): its derivative by the face's point q */
    double bf_nonlinear_d(int mode, int i, int d, int q) const {
        const int a = nb(i, st.adj[d]);
        const double cc = 0.5 * (c[i] + c[a]);
        if (wall_face(i, d)) return 0;
        const double pp = bf_drive(mode, i, d);
        const double hh = bf_face_value(i, d, pp);
        const double dhdhk = hc1[q] + Sign(pp) * hc2[q];
        const double val = mob.slip ? mob.dUpwind(hh) * dhdhk : 2.0 * hh * dhdhk;
        return val / visc(cc);
    }
    /* C(j)[idx] + Sign(dv) C2(j)[idx] of the face's point q: the
     * coefficients per spacing (SetUpwindCoef(true), UpwindKernelBase.h
     * :122-145) */
    double bf_coef(int j, int q, double dv) const {
        const double hs = spacing(j / 2);
        return hc1[q] / hs + Sign(dv) * (hc2[q] / hs);
    }
    /* SetMatrixA (:344-399) */
    void bodyforce_A_fixed(int mode, Csr &M) {
        const bool side = (mode == 1 && cs.dim == 1);
#pragma omp parallel for schedule(static) if (g.n > YZO_PAR_MIN)
        for (int i = 0; i < g.n; ++i) {
            if (M.bc[i]) continue;
            double dv[4], w[4], D[4], Cf[4][13] = {{0}};
            for (int j = 0; j < st.ndir; ++j) {
                dv[j] = bf_drive_face(mode, i, j);
                w[j] = bf_weight(mode, i, j);
                D[j] = bf_nonlinear(mode, i, j);
                for (int q = 0; q < 4; ++q) Cf[j][st.face[j][q]] = bf_coef(j, q, dv[j]);
            }
            for (int a = 0; a < st.naxis; ++a) {
                const int dm = 2 * a, dp = 2 * a + 1;
                const double Coef = bf_mag(mode, a);
                for (int q = 0; q < 5; ++q) {
                    const int idx = st.axis[a][q];
                    const double v = -w[dm] * Cf[dm][idx] * D[dm] + w[dp] * Cf[dp][idx] * D[dp];
                    M.at(i, idx) += Coef * v;
                }
            }
            if (side) {
                const double hoo = h[i];
                M.at(i, st.center) += mob.slip ? bf_mag(mode, 0) * (mob.DperH(hoo) / visc(c[i]))
                                               : bf_mag(mode, 0) * (hoo * hoo / visc(c[i]));
            }
        }
    }
    /* SetJacobiMatrix (:405-492) */
    void bodyforce_J_fixed(int mode, Csr &M) {
        const bool side = (mode == 1 && cs.dim == 1);
#pragma omp parallel for schedule(static) if (g.n > YZO_PAR_MIN)
        for (int i = 0; i < g.n; ++i) {
            if (M.bc[i]) continue;
            double acc[4], gg[4][13] = {{0}};
            for (int j = 0; j < st.ndir; ++j) {
                const double dv = bf_drive_face(mode, i, j), w = bf_weight(mode, i, j);
                const double sg = (j % 2 == 0) ? -1.0 : +1.0;
                acc[j] = 0;
                for (int q = 0; q < 4; ++q) acc[j] += sg * w * bf_coef(j, q, dv) * h[nb(i, st.face[j][q])];
                for (int q = 0; q < 4; ++q) gg[j][st.face[j][q]] = bf_nonlinear_d(mode, i, j, q);
            }
            for (int a = 0; a < st.naxis; ++a) {
                const int dm = 2 * a, dp = 2 * a + 1;
                const double Coef = bf_mag(mode, a);
                for (int q = 0; q < 5; ++q) {
                    const int idx = st.axis[a][q];
                    const int first = st.dir_of_adj(idx);
                    const bool inM = st.in_face(dm, idx), inP = st.in_face(dp, idx);
                    double v = 0;
                    if (first == dm && inM) v += gg[dm][idx] * acc[dm];
                    if (first == dp && inP) v += gg[dp][idx] * acc[dp];
                    if (first != dm && inM) v += gg[dm][idx] * acc[dm];
                    if (first != dp && inP) v += gg[dp][idx] * acc[dp];
                    M.at(i, idx) += Coef * v;
                }
            }
            if (side) {
                const double hoo = h[i];
                M.at(i, st.center) += mob.slip ? bf_mag(mode, 0) * ((mob.dD(hoo) - mob.DperH(hoo)) / visc(c[i]))
                                               : bf_mag(mode, 0) * (2.0 * hoo * hoo / visc(c[i]));
            }
        }
    }
    void bodyforce_A(int mode, Csr &M) {
        if (!hlim) { bodyforce_A_fixed(mode, M); return; }
        const bool side = (mode == 1 && cs.dim == 1);
#pragma omp parallel for schedule(static) if (g.n > YZO_PAR_MIN)
        for (int i = 0; i < g.n; ++i) {
            if (M.bc[i]) continue;
            double W[4][13] = {{0}}, D[4], w[4];
            for (int j = 0; j < st.ndir; ++j) {
                double hf, wt[4];
                bf_face(mode, i, j, hf, D[j], wt, nullptr);
                w[j] = bf_weight(mode, i, j);
                for (int q = 0; q < 4; ++q) W[j][st.face[j][q]] = wt[q];
            }
            for (int a = 0; a < st.naxis; ++a) {
                const int dm = 2 * a, dp = 2 * a + 1;
                const double Coef = bf_mag(mode, a);
                for (int q = 0; q < 5; ++q) {
                    const int idx = st.axis[a][q];
                    const double v = -w[dm] * W[dm][idx] * D[dm] + w[dp] * W[dp][idx] * D[dp];
                    M.at(i, idx) += Coef * v;
                }
            }
            if (side) {
                const double hoo = h[i];
                M.at(i, st.center) += mob.slip ? bf_mag(mode, 0) * (mob.DperH(hoo) / visc(c[i]))
                                               : bf_mag(mode, 0) * (hoo * hoo / visc(c[i]));
            }
        }
    }
    /* Marangoni (SolutoCapillarityKernel.h, limited) */
    double ma_coef() const { return cs.As * cs.Ma / 2.0; }
    double ma_up(double dc) const { return Sign(ma_coef()) * Sign(dc); }
    void face_conc(int i, int d, double &dc, double &cc) const {
        const int lo = nb(i, st.face[d][1]), hi = nb(i, st.face[d][2]);
        const double hh = spacing(d / 2);
        dc = (c[hi] - c[lo]) / hh;
        cc = (c[hi] + c[lo]) / 2.0;
    }
    /* ---- the fixed-coefficient schemes of the Marangoni term
     * (SolutoCapillarityKernel.h:182-405): S = dc/mu h_f (h_f + 2 Lambda
     * with slip), h_f = sum (C + Up C2) h over the face's points, the
     * coefficients not divided by the spacing (SetUpwindCoef(false)) ---- */
    double ma_coef_q(int q, double up) const { return hc1[q] + up * hc2[q]; }
    double ma_face_value(int i, int d, double up) const {
        double hh = 0;
        for (int q = 0; q < 4; ++q) hh += ma_coef_q(q, up) * h[nb(i, st.face[d][q])];
        return hh;
    }
    /* NonlinearTerm (:182-211) */
    double ma_nonlinear(int i, int d) const {
        if (wall_face(i, d)) return 0;
        double dc, cc;
        face_conc(i, d, dc, cc);
        const double hh = ma_face_value(i, d, ma_up(dc));
        const double Visco = visc(cc);
        if (mob.slip) return dc / Visco * mob.Shear(hh);
        return dc / Visco * hh;
    }
    /* SetMatrixA (:234-320) */
    void marangoni_A_fixed(Csr &M) {
        const double Coef = ma_coef();
#pragma omp parallel for schedule(static) if (g.n > YZO_PAR_MIN)
        for (int i = 0; i < g.n; ++i) {
            if (M.bc[i]) continue;
            double S[4], up[4], Val[13] = {0};
            for (int d = 0; d < st.ndir; ++d) {
                double dc, cc;
                face_conc(i, d, dc, cc);
                up[d] = ma_up(dc);
                S[d] = ma_nonlinear(i, d);
            }
            /* stencil order, then the directions in stencil order */
            for (int k = 0; k < st.ns; ++k)
                for (int kk = 0; kk < st.ndir; ++kk) {
                    const int d = st.sdir(kk);
                    double cq = 0;
                    for (int q = 0; q < 4; ++q) if (st.face[d][q] == k) cq = ma_coef_q(q, up[d]);
                    const double hh = spacing(d / 2);
                    const double sg = (d % 2 == 0) ? -1.0 : +1.0;
                    Val[k] += sg * cq * S[d] / hh;
                }
            for (int k = 0; k < st.ns; ++k) {
                bool wide = (k == st.center);
                for (int d = 0; d < st.ndir; ++d) wide = wide || st.in_face(d, k);
                if (wide) M.at(i, k) += Coef * Val[k];
            }
        }
    }
    /* SetJacobiMatrix (:322-385) with ScatterJacobiWide (UpwindKernelBase.h
     * :160-214) */
    void marangoni_J_fixed(Csr &M) {
        const double Coef = ma_coef();
#pragma omp parallel for schedule(static) if (g.n > YZO_PAR_MIN)
        for (int i = 0; i < g.n; ++i) {
            if (M.bc[i]) continue;
            double a[4], gg[4][13] = {{0}};
            for (int d = 0; d < st.ndir; ++d) {
                double dc, cc;
                face_conc(i, d, dc, cc);
                const double up = ma_up(dc);
                const double hh = spacing(d / 2);
                const double sg = (d % 2 == 0) ? -1.0 : +1.0;
                a[d] = 0;
                for (int q = 0; q < 4; ++q) a[d] += sg * ma_coef_q(q, up) * h[nb(i, st.face[d][q])] / hh;
                const bool none = wall_face(i, d);
                const double Visco = none ? 1.0 : visc(cc);
                for (int q = 0; q < 4; ++q) gg[d][st.face[d][q]] = none ? 0.0 : dc / Visco * ma_coef_q(q, up);
            }
            for (int k = 0; k < st.ns; ++k) {
                bool wide = (k == st.center);
                for (int d = 0; d < st.ndir; ++d) wide = wide || st.in_face(d, k);
                if (!wide) continue;
                double v = 0;
                if (k == st.center) {
                    for (int j = 0; j < st.ndir; ++j) { const int d = st.sdir(j); v += gg[d][k] * a[d]; }
                } else {
                    const int first = st.dir_of_adj(k);
                    if (first >= 0 && st.in_face(first, k)) v += gg[first][k] * a[first];
                    for (int d = 0; d < st.ndir; ++d) {
                        if (d == first || !st.in_face(d, k)) continue;
                        v += gg[d][k] * a[d];
                    }
                }
                M.at(i, k) += Coef * v;
            }
        }
    }
    void marangoni_A(Csr &M) {
        if (!hlim) { marangoni_A_fixed(M); return; }
        const double Coef = ma_coef();
#pragma omp parallel for schedule(static) if (g.n > YZO_PAR_MIN)
        for (int i = 0; i < g.n; ++i) {
            if (M.bc[i]) continue;
            double Val[13] = {0};
            for (int d = 0; d < st.ndir; ++d) {
                double dc, cc, f[4], wt[4];
                face_conc(i, d, dc, cc);
                face_heights(i, d, f);
                const double up = ma_up(dc);
                const double hf = lim.value(f, up);
                lim.weights(f, up, wt);
                const double S = wall_face(i, d) ? 0.0
                               : mob.slip ? dc / visc(cc) * mob.Shear(hf) : dc / visc(cc) * hf;
                const double hh = spacing(d / 2);
                const double sg = (d % 2 == 0) ? -1.0 : +1.0;
                for (int q = 0; q < 4; ++q) Val[st.face[d][q]] += sg * wt[q] * S / hh;
            }
            for (int k = 0; k < st.ns; ++k) {
                bool wide = (k == st.center);
                for (int d = 0; d < st.ndir; ++d) wide = wide || st.in_face(d, k);
                if (wide) M.at(i, k) += Coef * Val[k];
            }
        }
    }

    /* ---- SetJacobiMatrix of each term ---- */
    void capillary_J(Csr &M) {
        const double Coef = 1.0 / 3.0;
#pragma omp parallel for schedule(static) if (g.n > YZO_PAR_MIN)
        for (int i = 0; i < g.n; ++i) {
            if (M.bc[i]) continue;
            const double Hoo = h[i];
            double a[4], b[4], gA[4], gC[4], sA[4], sC[4];
            for (int d = 0; d < st.ndir; ++d) {
                a[d] = b[d] = 0;
                for (int k = 0; k < st.ns; ++k) { a[d] += Ccap[d][k] * h[nb(i, k)]; b[d] += Ccap[d][k] * s[nb(i, k)]; }
                gA[d] = cap_dD(i, d, false);
                gC[d] = cap_dD(i, d, true);
                sA[d] = gA[d] / Hoo;
                sC[d] = gC[d] / Hoo - cap_D(i, d) / (Hoo * Hoo);
            }
            double ga = 0, sb = 0;
            for (int k = 0; k < st.ndir; ++k) { const int dd = st.sdir(k); ga += gC[dd] * a[dd]; }
            for (int d = 0; d < st.ndir; ++d) sb += sC[d] * b[d];
            M.at(i, st.center) += Coef * (ga + sb * Hoo);
            for (int d = 0; d < st.ndir; ++d)
                M.at(i, st.adj[d]) += Coef * (gA[d] * a[d] + sA[d] * b[d] * Hoo);
        }
    }
    void vgrav_J(Csr &M) {
        const double Coef = -1.0 * cs.Bo / 3.0;
#pragma omp parallel for schedule(static) if (g.n > YZO_PAR_MIN)
        for (int i = 0; i < g.n; ++i) {
            if (M.bc[i]) continue;
            double a[4], b[4], gA[4], gC[4], kF[4];
            for (int d = 0; d < st.ndir; ++d) {
                a[d] = b[d] = 0;
                for (int k = 0; k < st.ns; ++k) { a[d] += narrowC(d, k) * h[nb(i, k)]; b[d] += narrowC(d, k) * s[nb(i, k)]; }
                gA[d] = vg_dG(i, d, false);
                gC[d] = vg_dG(i, d, true);
                kF[d] = vg_T(i, d);
            }
            double ga = 0;
            for (int k = 0; k < st.ndir; ++k) { const int dd = st.sdir(k); ga += gC[dd] * (a[dd] + b[dd]) - kF[dd] * b[dd]; }
            M.at(i, st.center) += Coef * ga;
            for (int d = 0; d < st.ndir; ++d)
                M.at(i, st.adj[d]) += Coef * (gA[d] * (a[d] + b[d]) - kF[d] * b[d]);
        }
    }
    void disjoining_J(Csr &M) {
        const double Coef = cs.Dp / 3.0;
#pragma omp parallel for schedule(static) if (g.n > YZO_PAR_MIN)
        for (int i = 0; i < g.n; ++i) {
            if (M.bc[i]) continue;
            double Ad[4], G[4], gA[4], gC[4];
            for (int d = 0; d < st.ndir; ++d) {
                Ad[d] = 0;
                for (int k = 0; k < st.ns; ++k) Ad[d] += narrowC(d, k) * Pi(h[nb(i, k)]);
                G[d] = vg_G(i, d);
                gA[d] = vg_dG(i, d, false);
                gC[d] = vg_dG(i, d, true);
            }
            double ga = 0, gc = 0;
            for (int k = 0; k < st.ndir; ++k) {
                const int dd = st.sdir(k);
                ga += gC[dd] * Ad[dd];
                gc += G[dd] * narrowC(dd, st.center);
            }
            {
                const double hj = h[i];
                const double dpi = dPi(hj) - PiPerH(hj);
                M.at(i, st.center) += Coef * (ga + gc * dpi);
            }
            for (int d = 0; d < st.ndir; ++d) {
                const int idx = st.adj[d];
                const double hj = h[nb(i, idx)];
                const double dpi = dPi(hj) - PiPerH(hj);
                M.at(i, idx) += Coef * (gA[d] * Ad[d] + G[d] * narrowC(d, idx) * dpi);
            }
        }
    }
    void bodyforce_J(int mode, Csr &M) {
        if (!hlim) { bodyforce_J_fixed(mode, M); return; }
        const bool side = (mode == 1 && cs.dim == 1);
#pragma omp parallel for schedule(static) if (g.n > YZO_PAR_MIN)
        for (int i = 0; i < g.n; ++i) {
            if (M.bc[i]) continue;
            double X[4][13] = {{0}};
            for (int j = 0; j < st.ndir; ++j) {
                double hf, D, wt[4], gd[4];
                bf_face(mode, i, j, hf, D, wt, gd);
                const double r3 = !mob.slip ? 3.0 : (mob.Upwind(hf) != 0.0 ? mob.dD(hf) / mob.Upwind(hf) : 0.0);
                const double sg = (j % 2 == 0) ? -1.0 : +1.0;
                const double w = bf_weight(mode, i, j);
                for (int q = 0; q < 4; ++q)
                    X[j][st.face[j][q]] = mob.slip ? sg * w * (r3 * gd[q] - wt[q]) * D : sg * w * (3.0 * gd[q] - wt[q]) * D;
            }
            for (int a = 0; a < st.naxis; ++a) {
                const int dm = 2 * a, dp = 2 * a + 1;
                const double Coef = bf_mag(mode, a);
                for (int q = 0; q < 5; ++q) {
                    const int idx = st.axis[a][q];
                    M.at(i, idx) += Coef * (X[dm][idx] + X[dp][idx]);
                }
            }
            if (side) {
                const double hoo = h[i];
                M.at(i, st.center) += mob.slip ? bf_mag(mode, 0) * ((mob.dD(hoo) - mob.DperH(hoo)) / visc(c[i]))
                                               : bf_mag(mode, 0) * (2.0 * hoo * hoo / visc(c[i]));
            }
        }
    }
    void marangoni_J(Csr &M) {
        if (!hlim) { marangoni_J_fixed(M); return; }
        const double Coef = ma_coef();
#pragma omp parallel for schedule(static) if (g.n > YZO_PAR_MIN)
        for (int i = 0; i < g.n; ++i) {
            if (M.bc[i]) continue;
            double a[4], gg[4][13] = {{0}};
            for (int d = 0; d < st.ndir; ++d) {
                double dc, cc, f[4], wt[4], gd[4];
                face_conc(i, d, dc, cc);
                face_heights(i, d, f);
                const double up = ma_up(dc);
                const double hf = lim.derivative(f, up, gd);
                lim.weights(f, up, wt);
                const double r = wall_face(i, d) ? 0.0 : dc / visc(cc);
                const double hh = spacing(d / 2);
                const double sg = (d % 2 == 0) ? -1.0 : +1.0;
                if (mob.slip) {
                    const double L2 = 2.0 * mob.lambda;
                    a[d] = sg / hh;
                    for (int q = 0; q < 4; ++q) gg[d][st.face[d][q]] = r * ((2.0 * hf + L2) * gd[q] - (hf + L2) * wt[q]);
                    continue;
                }
                a[d] = sg * hf / hh;
                for (int q = 0; q < 4; ++q) gg[d][st.face[d][q]] = r * (2.0 * gd[q] - wt[q]);
            }
            /* ScatterJacobiWide (UpwindKernelBase.h) */
            for (int k = 0; k < st.ns; ++k) {
                bool wide = (k == st.center);
                for (int d = 0; d < st.ndir; ++d) wide = wide || st.in_face(d, k);
                if (!wide) continue;
                double v = 0;
                if (k == st.center) {
                    for (int j = 0; j < st.ndir; ++j) { const int d = st.sdir(j); v += gg[d][k] * a[d]; }
                } else {
                    const int first = st.dir_of_adj(k);
                    if (first >= 0 && st.in_face(first, k)) v += gg[first][k] * a[first];
                    for (int d = 0; d < st.ndir; ++d) {
                        if (d == first || !st.in_face(d, k)) continue;
                        v += gg[d][k] * a[d];
                    }
                }
                M.at(i, k) += Coef * v;
            }
        }
    }

    /* ---- FaceFlux of each term: the flux through face d of cell i
     * along +axis, as the row's A h takes it (PhysicsTermBase.h:160-188) */
    double flux_capillary(int i, int d) const {   /* LaplacePressureKernel.h:480-491 */
        double vh[13], vs[13];
        for (int k = 0; k < st.ns; ++k) { vh[k] = h[nb(i, k)]; vs[k] = s[nb(i, k)]; }
        return (1.0 / 3.0) * (cap_D(i, d) * (face_gradient(vh, d) + face_gradient(vs, d)));
    }
    double flux_vgrav(int i, int d) const {       /* VerticalGravityKernel.h:271-287 */
        const int a = nb(i, st.adj[d]);
        const int lo = (d % 2 == 0) ? a : i, hi = (d % 2 == 0) ? i : a;
        const double dn = spacing(d / 2);
        const double hl = h[lo], hh = h[hi];
        const double mu = visc(0.5 * (c[lo] + c[hi]));
        double gf = mob.Face(hl, hh);
        const double tk = (gf == 0) ? 0.0 : gf / (hl + hh) / mu;
        gf /= mu;
        return (-1.0 * cs.Bo / 3.0) * ((gf * (hh - hl) + (s[hi] - s[lo]) * (tk * (hl + hh))) / (dn * dn));
    }
    double flux_bodyforce(int mode, int i, int j) const {   /* BodyForceKernel.h:118-151 */
        if (!hlim) {
            if (wall_face(i, j)) return 0;
            const double D = bf_nonlinear(mode, i, j);
            const double dv = bf_drive_face(mode, i, j);
            double hs = 0;
            for (int q = 0; q < 4; ++q) hs += bf_coef(j, q, dv) * h[nb(i, st.face[j][q])];
            return bf_mag(mode, j / 2) * (bf_weight(mode, i, j) * hs * D);
        }
        double hf, D, wt[4];
        bf_face(mode, i, j, hf, D, wt, nullptr);
        double hs = 0;
        for (int q = 0; q < 4; ++q) hs += wt[q] * h[nb(i, st.face[j][q])];
        return bf_mag(mode, j / 2) * (bf_weight(mode, i, j) * hs * D);
    }
    double flux_marangoni(int i, int d) const {   /* SolutoCapillarityKernel.h:104-138 */
        double dc, cc, f[4], wt[4];
        face_conc(i, d, dc, cc);
        const double hh = spacing(d / 2);
        if (!hlim) {
            const double hs = ma_face_value(i, d, ma_up(dc));
            const double S = wall_face(i, d) ? 0.0 : dc / visc(cc) * (mob.slip ? mob.Shear(hs) : hs);
            return ma_coef() * (S * hs / hh);
        }
        face_heights(i, d, f);
        const double up = ma_up(dc);
        const double hf = lim.value(f, up);
        lim.weights(f, up, wt);
        const double S = wall_face(i, d) ? 0.0 : dc / visc(cc) * (mob.slip ? mob.Shear(hf) : hf);
        double hs = 0;
        for (int q = 0; q < 4; ++q) hs += wt[q] * f[q];
        return ma_coef() * (S * hs / hh);
    }
    double flux_disjoining(int i, int d) const {  /* DisjoiningPressureKernel.h:160-168 */
        const int a = nb(i, st.adj[d]);
        const double v = narrowC(d, st.adj[d]) * Pi(h[a]) + narrowC(d, st.center) * Pi(h[i]);
        const double f = (cs.Dp / 3.0) * (vg_G(i, d) * v);
        return (d % 2 == 0) ? -f : f;
    }
    double face_flux(int i, int d) const {
        double F = 0;
        if (cs.capillary) F += flux_capillary(i, d);
        if (cs.vertical_gravity) F += flux_vgrav(i, d);
        if (cs.inplane_gravity) F += flux_bodyforce(0, i, d);
        if (cs.marangoni) F += flux_marangoni(i, d);
        if (cs.centrifugal) F += flux_bodyforce(1, i, d);
        if (cs.disjoining) F += flux_disjoining(i, d);
        return F;
    }
    /* SideRate: the 1-D centrifugal term's sideways outflow
     * (BodyForceKernel.h:153-159) */
    double side_rate(int i) const {
        if (!(cs.centrifugal && cs.dim == 1)) return 0;
        const double hoo = h[i], coo = c[i];
        if (mob.slip) return bf_mag(1, 0) * (mob.D(hoo) / visc(coo));
        return bf_mag(1, 0) * (hoo * hoo / visc(coo)) * hoo;
    }
    /* ---- the face velocities (CalVelocity, HeightEquationBase.h:1583-1601):
     * each term's velocity on the Xm face (x) and the Ym face (y) of every
     * cell 2 <= ix <= nxp-2 (2 <= iy <= nyp-2; InsideForVelocity,
     * GridTraits.h:239-244, 515-522), at the working field ---- */
    bool vel_cell(int i) const {
        const int ix = i % g.nxp, iy = i / g.nxp;
        if (ix < 2 || g.nxp - 2 < ix) return false;
        return cs.dim == 1 || !(iy < 2 || g.nyp - 2 < iy);
    }
    /* AxisMask: no y velocity on a grid 2 real cells deep (GridTraits.h:557-558) */
    double axis_mask(int a, double q) const { return (a == 1 && g.nyp == 6) ? 0.0 : q; }
    /* FaceHeight: the mean of the cells on either side of the face */
    double face_height(int i, int a) const { return 0.5 * (h[nb(i, st.adj[2 * a])] + h[i]); }
    /* LaplacePressure_2D.cpp:137-211, LaplacePressure_1D.cpp:84-136 */
    double vel_capillary(int i, int a) const {
        const double Mag = 1.0 / 3.0;
        if (cs.dim == 1) {
            const double dx = g.dx, rdx3 = 1.0 / (dx * dx * dx);
            const double DL = cap_D(i, 0), hL = 0.5 * (h[nb(i, 1)] + h[i]);
            const double hp = h[nb(i, 3)], ho = h[i], hm = h[nb(i, 1)], hM = h[nb(i, 0)];
            const double sp = s[nb(i, 3)], so = s[i], sm = s[nb(i, 1)], sM = s[nb(i, 0)];
            double Qhx = ((hp - 2.0 * ho + hm) - (ho - 2.0 * hm + hM)) * rdx3;
            double Qsx = ((sp - 2.0 * so + sm) - (so - 2.0 * sm + sM)) * rdx3;
            Qhx *= Mag * DL;
            Qsx *= Mag * DL;
            return (Qhx + Qsx) / hL;
        }
        enum { oM, mm, om, pm, Mo, mo, oo, po, Po, mp, op, pp, oP };
        const double dx = g.dx, dy = g.dy;
        const double rdx3 = 1.0 / (dx * dx * dx), rdy3 = 1.0 / (dy * dy * dy);
        const double ratio2x = (dx * dx) / (dy * dy), ratio2y = (dy * dy) / (dx * dx);
        double H[13], S[13];
        for (int k = 0; k < 13; ++k) {
            if (k == Po || k == oP) { H[k] = S[k] = 0; continue; }
            H[k] = h[nb(i, k)]; S[k] = s[nb(i, k)];
        }
        auto qx = [&](const double *v) {
            return (((v[po] - 2.0 * v[oo] + v[mo]) + (v[op] - 2.0 * v[oo] + v[om]) * ratio2x) -
                    ((v[oo] - 2.0 * v[mo] + v[Mo]) + (v[mp] - 2.0 * v[mo] + v[mm]) * ratio2x)) * rdx3;
        };
        auto qy = [&](const double *v) {
            return (((v[po] - 2.0 * v[oo] + v[mo]) * ratio2y + (v[op] - 2.0 * v[oo] + v[om])) -
                    ((v[pm] - 2.0 * v[om] + v[mm]) * ratio2y + (v[oo] - 2.0 * v[om] + v[oM]))) * rdy3;
        };
        if (a == 0) {
            const double DL = cap_D(i, 0), hL = 0.5 * (h[nb(i, mo)] + h[i]);
            double Qhx = qx(H), Qsx = qx(S);
            Qhx *= Mag * DL;
            Qsx *= Mag * DL;
            return (Qhx + Qsx) / hL;
        }
        const double DB = cap_D(i, 2), hB = 0.5 * (h[nb(i, om)] + h[i]);
        double Qhy = qy(H), Qsy = qy(S);
        Qhy *= Mag * DB;
        Qsy *= Mag * DB;
        return (Qhy + Qsy) / hB;
    }
    /* CV(a) = -C(2a) spacing over the face's two cells (VerticalGravityKernel.h
     * :143-161): the difference across the Xm (Ym) face */
    double cv_diff(int i, int a, double vadj, double vcen) const {
        const double hsp = spacing(a), h2 = hsp * hsp;
        const double cadj = -(1.0 / h2) * hsp, ccen = -(-1.0 / h2) * hsp;
        double q = 0;
        q += cadj * vadj;
        q += ccen * vcen;
        (void)i;
        return q;
    }
    /* VerticalGravityKernel.h:435-470 */
    double vel_vgrav(int i, int a) const {
        const int adj = nb(i, st.adj[2 * a]);
        const double Gf = vg_G(i, 2 * a);
        double Q = cv_diff(i, a, h[adj] + s[adj], h[i] + s[i]);
        Q *= (-1.0 * cs.Bo / 3.0) * Gf;
        Q = axis_mask(a, Q);
        return Q / face_height(i, a);
    }
    /* DisjoiningPressureKernel.h:442-475 */
    double vel_disjoining(int i, int a) const {
        const int adj = nb(i, st.adj[2 * a]);
        const double Gf = vg_G(i, 2 * a);
        double Q = cv_diff(i, a, Pi(h[adj]), Pi(h[i]));
        Q *= (cs.Dp / 3.0) * Gf;
        Q = axis_mask(a, Q);
        return Q / face_height(i, a);
    }
    /* SolutoCapillarityKernel.h:387-405 with its NonlinearTerm :182-210
     * (limited face height) */
    double vel_marangoni(int i, int a) const {
        const int d = 2 * a;
        if (!hlim) return axis_mask(a, ma_coef() * ma_nonlinear(i, d));   /* CalVelocity :387-405 */
        double S = 0;
        if (!wall_face(i, d)) {
            double dc, cc, f[4];
            face_conc(i, d, dc, cc);
            face_heights(i, d, f);
            const double hh = lim.value(f, ma_up(dc));
            const double mu = visc(cc);
            S = mob.slip ? dc / mu * mob.Shear(hh) : dc / mu * hh;
        }
        return axis_mask(a, ma_coef() * S);
    }
    /* BodyForceKernel.h:694-720 (limited): -w Mag D, times VelocitySign -1
     * (ParallelGravityKernel.h:112, CentrifugalKernel.h:144) */
    double vel_bodyforce(int mode, int i, int a) const {
        if (!hlim) {
            /* CalVelocity (:494-545): the face value from CV = -C(2a) h and
             * CV2 = -C2(2a) h, and from C, C2 times the spacing */
            const int dm = 2 * a;
            const double D = bf_nonlinear(mode, i, dm);
            const double dv = bf_drive_face(mode, i, dm), wv = bf_weight(mode, i, dm);
            const double hsp = spacing(a);
            double Q = 0, hh = 0;
            for (int q = 0; q < 4; ++q) {
                const double v = h[nb(i, st.face[dm][q])];
                const double cv = -(hc1[q] / hsp) * hsp, cv2 = -(hc2[q] / hsp) * hsp;
                Q += wv * (cv + Sign(dv) * cv2) * v;
                hh += bf_coef(dm, q, dv) * v;
            }
            hh *= hsp;
            Q *= bf_mag(mode, a) * D;
            Q = axis_mask(a, Q);
            if (wall_face(i, dm)) { Q = 0; hh = 1; }
            return -1.0 * Q / hh;
        }
        double hf, D, wt[4];
        bf_face(mode, i, 2 * a, hf, D, wt, nullptr);
        double Q = -bf_weight(mode, i, 2 * a) * bf_mag(mode, a) * D;
        Q = axis_mask(a, Q);
        return -1.0 * Q;
    }
    /* SetVelocityBC (WM2D/HeightEquation_2D.cpp:32-151, WM1D/
     * HeightEquation_1D.cpp:28-86), Symmetry and Periodic edges */
    void velocity_bc(vector<double> &vx, vector<double> &vy) const {
        const int Nx = g.nxp, Ny = g.nyp;
        auto per = [&](int k) { return g.bc[k] == YZ_BC_PERIODIC; };
        if (cs.dim == 1) {
            vx[0] = 0;
            if (per(0)) vx[1] = vx[Nx - 3]; else { vx[1] = 0; vx[2] = 0; }
            if (per(1)) vx[Nx - 1] = vx[3]; else { vx[Nx - 2] = 0.0; vx[Nx - 1] = 0.0; }
            return;
        }
        auto X = [&](int ix, int iy) -> double & { return vx[g.row(ix, iy)]; };
        auto Y = [&](int ix, int iy) -> double & { return vy[g.row(ix, iy)]; };
        for (int jy = 1; jy < Ny - 1; ++jy) {
            X(0, jy) = 0;
            if (per(0)) { Y(1, jy) = Y(Nx - 3, jy); X(1, jy) = X(Nx - 3, jy); }
            else { Y(1, jy) = Y(2, jy); X(1, jy) = 0; X(2, jy) = 0; }
            if (per(1)) { Y(Nx - 2, jy) = Y(2, jy); X(Nx - 1, jy) = X(3, jy); }
            else { Y(Nx - 2, jy) = Y(Nx - 3, jy); X(Nx - 2, jy) = 0.0; X(Nx - 1, jy) = 0.0; }
        }
        for (int ix = 1; ix < Nx - 1; ++ix) {
            Y(ix, 0) = 0;
            if (per(2)) { X(ix, 1) = X(ix, Ny - 3); Y(ix, 1) = Y(ix, Ny - 3); }
            else { X(ix, 1) = X(ix, 2); Y(ix, 1) = 0; Y(ix, 2) = 0; }
            if (per(3)) { X(ix, Ny - 2) = X(ix, 2); Y(ix, Ny - 1) = Y(ix, 3); }
            else { X(ix, Ny - 2) = X(ix, Ny - 3); Y(ix, Ny - 2) = 0; Y(ix, Ny - 1) = 0; }
        }
    }
    /* CalVelocity: every term's field with its boundary rows, their plain sum
     * in Cal_MixedVelocity's order (LubFields.cpp:467-495: capillary,
     * Marangoni, vertical gravity, in-plane gravity, centrifugal,
     * disjoining), the boundary rows once more. bc = 0 leaves out the
     * boundary rows (what the terms' CalVelocity alone write). */
    void velocity(vector<double> &vx, vector<double> &vy, bool bc) const {
        vector<double> t[6][2];
        velocity(vx, vy, bc, t);
    }
    /* the same, with each term's field (its boundary rows once) in t */
    void velocity(vector<double> &vx, vector<double> &vy, bool bc, vector<double> (&t)[6][2]) const {
        const int na = st.naxis;
        const bool on[6] = {cs.capillary != 0, cs.marangoni != 0, cs.vertical_gravity != 0,
                            cs.inplane_gravity != 0, cs.centrifugal != 0, cs.disjoining != 0};
        for (int k = 0; k < 6; ++k)
            for (int a = 0; a < 2; ++a) t[k][a].assign(g.n, 0.0);
        for (int i = 0; i < g.n; ++i) {
            if (!vel_cell(i)) continue;
            for (int a = 0; a < na; ++a) {
                if (on[0]) t[0][a][i] = vel_capillary(i, a);
                if (on[1]) t[1][a][i] = vel_marangoni(i, a);
                if (on[2]) t[2][a][i] = vel_vgrav(i, a);
                if (on[3]) t[3][a][i] = vel_bodyforce(0, i, a);
                if (on[4]) t[4][a][i] = vel_bodyforce(1, i, a);
                if (on[5]) t[5][a][i] = vel_disjoining(i, a);
            }
        }
        if (bc)
            for (int k = 0; k < 6; ++k) velocity_bc(t[k][0], t[k][1]);
        vx.assign(g.n, 0.0); vy.assign(g.n, 0.0);
        for (int i = 0; i < g.n; ++i) {
            vx[i] = t[0][0][i] + t[1][0][i] + t[2][0][i] + t[3][0][i] + t[4][0][i] + t[5][0][i];
            vy[i] = t[0][1][i] + t[1][1][i] + t[2][1][i] + t[3][1][i] + t[4][1][i] + t[5][1][i];
        }
        if (bc) velocity_bc(vx, vy);
    }

    /* ---- the energy budget (EnergyBudget.h:52-166; the parameters of
     * WellMixedModelBase.h:869-880): the free energy F of the working field,
     * the potential phi, and the power of the face fluxes against phi
     * (FluxPower, HeightEquationBase.h:275-291) ---- */
    double en_W(double hh) const {   /* w(H) = -int_{H*}^{H} Pi */
        const double dp = cs.disjoining ? cs.Dp : 0.0;
        if (dp == 0 || !(hh > 0)) return 0;
        const double hs = cs.hstar, n = cs.dis_n, m = cs.dis_m;
        const double in = std::pow(hs, n) * (std::pow(hh, 1.0 - n) - std::pow(hs, 1.0 - n)) / (1.0 - n);
        const double im = std::pow(hs, m) * (std::pow(hh, 1.0 - m) - std::pow(hs, 1.0 - m)) / (1.0 - m);
        return -dp * (in - im);
    }
    double en_cell(double hh, double f) const {
        const double bo = cs.vertical_gravity ? cs.Bo : 0.0;
        return 0.5 * bo * f * f + en_W(hh);
    }
    double en_face(double fi, double fj, double ci, double cj, double dn) const {
        if (!cs.capillary) return 0;
        const double gr = (fj - fi) / dn;
        return 0.25 * (sigma(ci) + sigma(cj)) * gr * gr;
    }
    double free_energy() const {
        auto v = [&](int i) { return h[i] + s[i]; };
        const bool px = g.bc[0] == YZ_BC_PERIODIC, py = cs.dim == 2 && g.bc[2] == YZ_BC_PERIODIC;
        double F = 0;
        for (int i = 0; i < g.n; ++i) {
            if (ghost_row(i)) continue;
            F += en_cell(h[i], v(i));
            const int xp = nb(i, st.adj[1]), xm = nb(i, st.adj[0]);
            F += en_face(v(i), v(xp), c[i], c[xp], g.dx);
            if (cs.dim == 2) {
                const int yp = nb(i, st.adj[3]);
                F += en_face(v(i), v(yp), c[i], c[yp], g.dy);
            }
            if (!px && ghost_row(xm)) F += en_face(v(xm), v(i), c[xm], c[i], g.dx);
            if (cs.dim == 2) {
                const int ym = nb(i, st.adj[2]);
                if (!py && ghost_row(ym)) F += en_face(v(ym), v(i), c[ym], c[i], g.dy);
            }
        }
        return cs.dim == 2 ? F * g.dx * g.dy : F * g.dx;
    }
    /* PotentialAt (LubFields.cpp:300-311) at every real cell, 0 elsewhere */
    void potential(vector<double> &phi) const {
        auto v = [&](int i) { return h[i] + s[i]; };
        phi.assign(g.n, 0.0);
        const double dx2 = g.dx * g.dx, dy2 = g.dy * g.dy;
        const double dp = cs.disjoining ? cs.Dp : 0.0, bo = cs.vertical_gravity ? cs.Bo : 0.0;
        for (int i = 0; i < g.n; ++i) {
            if (ghost_row(i)) continue;
            const double vo = v(i);
            double lap = (v(nb(i, st.adj[1])) - 2.0 * vo + v(nb(i, st.adj[0]))) / dx2;
            if (cs.dim == 2) lap = lap + (v(nb(i, st.adj[3])) - 2.0 * vo + v(nb(i, st.adj[2]))) / dy2;
            double p = 0;
            if (cs.capillary) p += sigma(c[i]) * lap;
            if (dp != 0 && h[i] > 0) {
                const double r = cs.hstar / h[i];
                p += dp * (std::pow(r, cs.dis_n) - std::pow(r, cs.dis_m));
            }
            if (bo != 0) p -= bo * vo;
            phi[i] = p;
        }
    }
    /* FluxPower times the cell measure: the dissipation D */
    double flux_power() const {
        vector<double> phi, F[2], r;
        potential(phi);
        face_volume_fluxes(F, r);
        double p = 0;
        for (int i = 0; i < g.n; ++i) {
            if (ghost_row(i)) continue;
            double div = 0;
            for (int a = 0; a < st.naxis; ++a) div += F[a][i] - F[a][nb(i, st.adj[2 * a])];
            p += phi[i] * (-div - r[i]);
        }
        return p * (cs.dim == 2 ? g.dx * g.dy : g.dx);
    }

    /* FaceVolumeFluxes (HeightEquationBase.h:852-894) of the working field */
    void face_volume_fluxes(vector<double> F[2], vector<double> &r) const {
        for (int a = 0; a < st.naxis; ++a) F[a].assign(g.n, 0.0);
        r.assign(g.n, 0.0);
        for (int i = 0; i < g.n; ++i) {
            if (ghost_row(i)) continue;
            for (int a = 0; a < st.naxis; ++a) {
                const int il = nb(i, st.adj[2 * a]);
                const bool below = ghost_row(il);
                F[a][i] = face_flux(i, 2 * a + 1);
                if (below) F[a][il] = face_flux(i, 2 * a);
            }
            r[i] = side_rate(i);
        }
    }

    /* ---- SetMatrixA (HeightEquationBase.h:669-751), order of
     * RegisterPhysics: capillary, vertical gravity, in-plane gravity,
     * Marangoni, centrifugal, disjoining ---- */
    void set_matrix_A() {
        A.clear_nonbc(); Arest.clear_nonbc();
        if (cs.capillary) capillary_A(A);
        if (cs.vertical_gravity) vgrav_A(A);
        if (cs.inplane_gravity) bodyforce_A(0, Arest);
        if (cs.marangoni) marangoni_A(Arest);
        if (cs.centrifugal) bodyforce_A(1, Arest);
        if (cs.disjoining) disjoining_A(Arest);
#pragma omp parallel for schedule(static) if (g.n > YZO_PAR_MIN)
        for (int i = 0; i < g.n; ++i) {
            if (A.bc[i]) continue;
            for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) A.val[k] += Arest.val[k];
        }
    }
    /* SetMatrixJacobi (:1061-1091) */
    void set_matrix_J() {
#pragma omp parallel for schedule(static) if (g.n > YZO_PAR_MIN)
        for (int i = 0; i < g.n; ++i) {
            if (A.bc[i]) continue;
            for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) J.val[k] = A.val[k];
        }
        if (cs.capillary) capillary_J(J);
        if (cs.vertical_gravity) vgrav_J(J);
        if (cs.inplane_gravity) bodyforce_J(0, J);
        if (cs.marangoni) marangoni_J(J);
        if (cs.centrifugal) bodyforce_J(1, J);
        if (cs.disjoining) disjoining_J(J);
    }
    /* ApplyA (:803-827) with AddConservative of the capillary
     * (LaplacePressureKernel.h:439-477) and vertical gravity
     * (VerticalGravityKernel.h:233-269) */
    void apply_A(const vector<double> &x, vector<double> &y) const {
        const bool any = cs.capillary || cs.vertical_gravity;
        if (!any) { A.matvec(x, y); return; }
        Arest.matvec(x, y);
#pragma omp parallel for schedule(static) if (g.n > YZO_PAR_MIN)
        for (int i = 0; i < g.n; ++i) {
            if (!A.bc[i]) continue;
            double v = 0;
            for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) v += A.val[k] * x[A.col[k]];
            y[i] = v;
        }
        if (cs.capillary) {
            const double Coef = 1.0 / 3.0;
#pragma omp parallel for schedule(static) if (g.n > YZO_PAR_MIN)
            for (int i = 0; i < g.n; ++i) {
                if (A.bc[i]) continue;
                double vh[13], vs[13];
                for (int k = 0; k < st.ns; ++k) { vh[k] = x[nb(i, k)]; vs[k] = s[nb(i, k)]; }
                double r = 0;
                for (int d = 0; d < st.ndir; ++d) {
                    const double f = Coef * (dFace[d][i] * (face_gradient(vh, d) + face_gradient(vs, d)));
                    r += (d % 2 == 0) ? -f : f;
                }
                y[i] += r;
            }
        }
        if (cs.vertical_gravity) {
            const double Coef = -1.0 * cs.Bo / 3.0;
#pragma omp parallel for schedule(static) if (g.n > YZO_PAR_MIN)
            for (int i = 0; i < g.n; ++i) {
                if (A.bc[i]) continue;
                double r = 0;
                for (int d = 0; d < st.ndir; ++d) {
                    const int a = nb(i, st.adj[d]);
                    const int lo = (d % 2 == 0) ? a : i, hi = (d % 2 == 0) ? i : a;
                    const double dn = spacing(d / 2);
                    const double hl = h[lo], hh = h[hi];
                    const double mu = visc(0.5 * (c[lo] + c[hi]));
                    double gf = mob.Face(hl, hh);
                    const double tk = (gf == 0) ? 0.0 : gf / (hl + hh) / mu;
                    gf /= mu;
                    const double xl = x[lo], xh = x[hi];
                    const double f = Coef * ((gf * (xh - xl) + (s[hi] - s[lo]) * (tk * (xl + xh))) / (dn * dn));
                    r += (d % 2 == 0) ? -f : f;
                }
                y[i] += r;
            }
        }
    }
};

/* ------------------------------------------------------------------ */
/* Linear algebra: ScaleRows, BiCGSTAB, SA-AMG, SPAI0, damped Jacobi    */
/* ------------------------------------------------------------------ */
struct SpM {  /* plain CSR, sorted columns not required */
    int n = 0;
    vector<int> ptr, col;
    vector<double> val;
    void mv(const double *x, double *y) const {
#pragma omp parallel for schedule(static) if (n > YZO_PAR_MIN)
        for (int i = 0; i < n; ++i) {
            double s = 0;
            for (int k = ptr[i]; k < ptr[i + 1]; ++k) s += val[k] * x[col[k]];
            y[i] = s;
        }
    }
    double diag(int i) const {
        for (int k = ptr[i]; k < ptr[i + 1]; ++k) if (col[k] == i) return val[k];
        return 0;
    }
};

static double dot(const vector<double> &a, const vector<double> &b) {
    double s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

static SpM transpose(const SpM &A, int ncols) {
    SpM T; T.n = ncols;
    T.ptr.assign(ncols + 1, 0);
    for (int k = 0; k < A.ptr[A.n]; ++k) T.ptr[A.col[k] + 1]++;
    for (int i = 0; i < ncols; ++i) T.ptr[i + 1] += T.ptr[i];
    T.col.resize(A.ptr[A.n]); T.val.resize(A.ptr[A.n]);
    vector<int> pos(T.ptr.begin(), T.ptr.end() - 1);
    for (int i = 0; i < A.n; ++i)
        for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
            const int p = pos[A.col[k]]++;
            T.col[p] = i; T.val[p] = A.val[k];
        }
    return T;
}
/* C = A * B (A: n x m, B: m x p) */
/* C = A B (B with p columns); each row's products are summed in A's then
 * B's storage order and its columns sorted. Over OpenMP threads the rows
 * are cut into contiguous chunks, each built apart and joined in row order:
 * the same matrix as one thread builds. */
static SpM spgemm(const SpM &A, const SpM &B, int p) {
    SpM C; C.n = A.n; C.ptr.assign(A.n + 1, 0);
    int nch = 1;
#ifdef _OPENMP
    if (A.n > YZO_PAR_MIN) nch = omp_get_max_threads();
#endif
    vector<vector<int>> ccol(nch);
    vector<vector<double>> cval(nch);
#pragma omp parallel if (nch > 1)
    {
        int t = 0, T = 1;
#ifdef _OPENMP
        t = omp_get_thread_num(); T = omp_get_num_threads();
#endif
        vector<int> marker(p, -1);
        vector<double> acc(p, 0.0);
        vector<int> cols;
        for (int ch = t; ch < nch; ch += T) {
            const int lo = (int)((long long)A.n * ch / nch), hi = (int)((long long)A.n * (ch + 1) / nch);
            for (int i = lo; i < hi; ++i) {
                cols.clear();
                for (int ka = A.ptr[i]; ka < A.ptr[i + 1]; ++ka) {
                    const int j = A.col[ka]; const double va = A.val[ka];
                    for (int kb = B.ptr[j]; kb < B.ptr[j + 1]; ++kb) {
                        const int c = B.col[kb];
                        if (marker[c] != i) { marker[c] = i; acc[c] = 0; cols.push_back(c); }
                        acc[c] += va * B.val[kb];
                    }
                }
                std::sort(cols.begin(), cols.end());
                for (int c : cols) { ccol[ch].push_back(c); cval[ch].push_back(acc[c]); }
                C.ptr[i + 1] = (int)cols.size();
            }
        }
    }
    for (int i = 0; i < A.n; ++i) C.ptr[i + 1] += C.ptr[i];
    C.col.reserve(C.ptr[A.n]); C.val.reserve(C.ptr[A.n]);
    for (int ch = 0; ch < nch; ++ch) {
        C.col.insert(C.col.end(), ccol[ch].begin(), ccol[ch].end());
        C.val.insert(C.val.end(), cval[ch].begin(), cval[ch].end());
    }
    return C;
}

struct AmgLevel {
    SpM A, P, R;
    vector<double> spai;      /* SPAI0 diagonal */
    vector<double> x, b, r;
    int nc = 0;
};

struct Amg {
    vector<AmgLevel> lv;
    vector<double> lu;        /* dense LU of the coarsest level */
    vector<int> piv;
    int nco = 0;
    int coarse_enough = 500;

    void build(const SpM &A0) {
        lv.clear();
        lv.emplace_back();
        lv[0].A = A0;
        double eps_strong = 0.08;
        while (lv.back().A.n > coarse_enough) {
            AmgLevel &L = lv.back();
            const SpM &A = L.A;
            const int n = A.n;
            vector<double> dia(n);
            for (int i = 0; i < n; ++i) dia[i] = A.diag(i);
            /* strength: a_ij^2 > eps^2 |a_ii a_jj| */
            const double e2 = eps_strong * eps_strong;
            vector<char> strong(A.ptr[n], 0);
            for (int i = 0; i < n; ++i)
                for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
                    const int j = A.col[k];
                    if (j == i) continue;
                    const double v = A.val[k];
                    strong[k] = (v * v > e2 * std::fabs(dia[i] * dia[j])) ? 1 : 0;
                }
            /* aggregation: pass 1 seeds whose strong neighbourhood is free;
             * pass 2 attaches the rest to a neighbouring aggregate; pass 3
             * makes aggregates of what is left */
            const int UND = -1, ISO = -2;
            vector<int> id(n, UND);
            for (int i = 0; i < n; ++i) {
                bool any = false;
                for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) if (strong[k]) { any = true; break; }
                if (!any) id[i] = ISO;
            }
            int na = 0;
            for (int i = 0; i < n; ++i) {
                if (id[i] != UND) continue;
                bool free = true;
                for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
                    if (strong[k] && id[A.col[k]] != UND) { free = false; break; }
                if (!free) continue;
                id[i] = na;
                for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) if (strong[k]) id[A.col[k]] = na;
                ++na;
            }
            vector<int> id2 = id;
            for (int i = 0; i < n; ++i) {
                if (id[i] != UND) continue;
                double best = -1; int ba = UND;
                for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
                    if (strong[k] && id[A.col[k]] >= 0 && std::fabs(A.val[k]) > best) { best = std::fabs(A.val[k]); ba = id[A.col[k]]; }
                if (ba != UND) id2[i] = ba;
            }
            id = id2;
            for (int i = 0; i < n; ++i) {
                if (id[i] != UND) continue;
                id[i] = na;
                for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) if (strong[k] && id[A.col[k]] == UND) id[A.col[k]] = na;
                ++na;
            }
            if (na == 0 || na >= n) break;
            /* tentative prolongator, columns of unit norm */
            vector<int> cnt(na, 0);
            for (int i = 0; i < n; ++i) if (id[i] >= 0) cnt[id[i]]++;
            SpM Pt; Pt.n = n; Pt.ptr.assign(n + 1, 0);
            for (int i = 0; i < n; ++i) {
                if (id[i] >= 0) { Pt.col.push_back(id[i]); Pt.val.push_back(1.0 / std::sqrt((double)cnt[id[i]])); }
                Pt.ptr[i + 1] = (int)Pt.col.size();
            }
            /* filtered matrix: weak connections lumped into the diagonal */
            SpM Af; Af.n = n; Af.ptr.assign(n + 1, 0);
            vector<double> df(n);
            for (int i = 0; i < n; ++i) {
                double dd = 0;
                for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
                    const int j = A.col[k];
                    if (j == i) dd += A.val[k];
                    else if (!strong[k]) dd += A.val[k];
                }
                df[i] = dd;
                Af.col.push_back(i); Af.val.push_back(dd);
                for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
                    if (A.col[k] != i && strong[k]) { Af.col.push_back(A.col[k]); Af.val.push_back(A.val[k]); }
                Af.ptr[i + 1] = (int)Af.col.size();
            }
            /* omega = 4/3 / rho(D^-1 Af), rho by Gershgorin */
            double rho = 0;
            for (int i = 0; i < n; ++i) {
                double sr = 0;
                for (int k = Af.ptr[i]; k < Af.ptr[i + 1]; ++k) sr += std::fabs(Af.val[k]);
                sr /= std::fabs(df[i]);
                rho = std::max(rho, sr);
            }
            const double omega = (4.0 / 3.0) / rho;
            SpM S = Af;
            for (int i = 0; i < n; ++i)
                for (int k = S.ptr[i]; k < S.ptr[i + 1]; ++k)
                    S.val[k] = (S.col[k] == i ? 1.0 : 0.0) - omega * Af.val[k] / df[i];
            L.P = spgemm(S, Pt, na);
            L.R = transpose(L.P, na);
            L.nc = na;
            SpM AP = spgemm(A, L.P, na);
            AmgLevel next;
            next.A = spgemm(L.R, AP, na);
            lv.push_back(std::move(next));
            eps_strong *= 0.5;
        }
        for (auto &L : lv) {
            const SpM &A = L.A;
            L.spai.assign(A.n, 0.0);
            for (int i = 0; i < A.n; ++i) {
                double ss = 0, dd = 0;
                for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) { ss += A.val[k] * A.val[k]; if (A.col[k] == i) dd = A.val[k]; }
                L.spai[i] = (ss > 0) ? dd / ss : 0.0;
            }
            L.x.assign(A.n, 0); L.b.assign(A.n, 0); L.r.assign(A.n, 0);
        }
        /* coarsest: dense LU with partial pivoting */
        const SpM &C = lv.back().A;
        nco = C.n;
        lu.assign((size_t)nco * nco, 0.0);
        for (int i = 0; i < nco; ++i)
            for (int k = C.ptr[i]; k < C.ptr[i + 1]; ++k) lu[(size_t)i * nco + C.col[k]] += C.val[k];
        piv.assign(nco, 0);
        for (int k = 0; k < nco; ++k) {
            int p = k;
            for (int i = k + 1; i < nco; ++i) if (std::fabs(lu[(size_t)i * nco + k]) > std::fabs(lu[(size_t)p * nco + k])) p = i;
            piv[k] = p;
            if (p != k) for (int j = 0; j < nco; ++j) std::swap(lu[(size_t)k * nco + j], lu[(size_t)p * nco + j]);
            const double d = lu[(size_t)k * nco + k];
            if (d == 0) continue;
            for (int i = k + 1; i < nco; ++i) {
                const double m = lu[(size_t)i * nco + k] / d;
                lu[(size_t)i * nco + k] = m;
                for (int j = k + 1; j < nco; ++j) lu[(size_t)i * nco + j] -= m * lu[(size_t)k * nco + j];
            }
        }
    }
    void coarse_solve(const vector<double> &b, vector<double> &x) const {
        x = b;
        for (int k = 0; k < nco; ++k) std::swap(x[k], x[piv[k]]);
        for (int i = 0; i < nco; ++i)
            for (int j = 0; j < i; ++j) x[i] -= lu[(size_t)i * nco + j] * x[j];
        for (int i = nco - 1; i >= 0; --i) {
            for (int j = i + 1; j < nco; ++j) x[i] -= lu[(size_t)i * nco + j] * x[j];
            const double d = lu[(size_t)i * nco + i];
            x[i] = (d != 0) ? x[i] / d : 0.0;
        }
    }
    void cycle(size_t l) {
        AmgLevel &L = lv[l];
        if (l + 1 == lv.size()) { coarse_solve(L.b, L.x); return; }
        const int n = L.A.n;
        /* pre-smoothing from x = 0: x = M b */
        #pragma omp parallel for schedule(static) if (n > YZO_PAR_MIN)
        for (int i = 0; i < n; ++i) L.x[i] = L.spai[i] * L.b[i];
        L.A.mv(L.x.data(), L.r.data());
        #pragma omp parallel for schedule(static) if (n > YZO_PAR_MIN)
        for (int i = 0; i < n; ++i) L.r[i] = L.b[i] - L.r[i];
        AmgLevel &Nx = lv[l + 1];
        L.R.mv(L.r.data(), Nx.b.data());
        cycle(l + 1);
        vector<double> corr(n);
        L.P.mv(Nx.x.data(), corr.data());
        #pragma omp parallel for schedule(static) if (n > YZO_PAR_MIN)
        for (int i = 0; i < n; ++i) L.x[i] += corr[i];
        L.A.mv(L.x.data(), L.r.data());
        #pragma omp parallel for schedule(static) if (n > YZO_PAR_MIN)
        for (int i = 0; i < n; ++i) L.x[i] += L.spai[i] * (L.b[i] - L.r[i]);
    }
    void apply(const vector<double> &rhs, vector<double> &x) {
        lv[0].b = rhs;
        cycle(0);
        x = lv[0].x;
    }
};

/* BiCGSTAB with right preconditioning; returns 0 / 1 / 2 as the backend */
template <class Mat, class Pc>
static int bicgstab(const Mat &A, const vector<double> &b, vector<double> &x, Pc &M,
                    int maxit, double tol, int &iters, double &resid) {
    const int n = A.n;
    vector<double> r(n), rh, p(n, 0), v(n, 0), ph(n), s(n), sh(n), t(n);
    A.mv(x.data(), r.data());
    for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
    rh = r;
    const double nb = std::sqrt(dot(b, b));
    iters = 0;
    if (nb == 0) { std::fill(x.begin(), x.end(), 0.0); resid = 0; return 0; }
    const double eps = tol * nb;
    double res = std::sqrt(dot(r, r));
    if (res <= eps) { resid = res / nb; return 0; }
    double rho1 = 0, rho2 = 1, alpha = 1, omega = 1;
    int it = 0;
    for (; it < maxit; ++it) {
        rho1 = dot(r, rh);
        if (std::fabs(rho1) < 1e-300) { iters = it; resid = res / nb; return 2; }
        if (it == 0) p = r;
        else {
            const double beta = (rho1 * alpha) / (rho2 * omega);
            #pragma omp parallel for schedule(static) if (n > YZO_PAR_MIN)
            for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        }
        M.apply(p, ph);
        A.mv(ph.data(), v.data());
        const double rv = dot(rh, v);
        if (rv == 0) { iters = it; resid = res / nb; return 2; }
        alpha = rho1 / rv;
        #pragma omp parallel for schedule(static) if (n > YZO_PAR_MIN)
        for (int i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
        res = std::sqrt(dot(s, s));
        if (res < eps) {
            for (int i = 0; i < n; ++i) x[i] += alpha * ph[i];
            ++it;
            break;
        }
        M.apply(s, sh);
        A.mv(sh.data(), t.data());
        const double tt = dot(t, t);
        omega = (tt != 0) ? dot(t, s) / tt : 0.0;
        if (omega == 0) { iters = it; resid = res / nb; return 2; }
        #pragma omp parallel for schedule(static) if (n > YZO_PAR_MIN)
        for (int i = 0; i < n; ++i) x[i] += alpha * ph[i] + omega * sh[i];
        #pragma omp parallel for schedule(static) if (n > YZO_PAR_MIN)
        for (int i = 0; i < n; ++i) r[i] = s[i] - omega * t[i];
        res = std::sqrt(dot(r, r));
        rho2 = rho1;
        if (res < eps) { ++it; break; }
        if (!(res == res)) { iters = it; resid = res; return 2; }
    }
    iters = it;
    resid = res / nb;
    if (!(resid == resid)) return 2;
    return resid <= tol ? 0 : 1;
}

struct Jacobi {
    vector<double> dinv;
    void build(const SpM &A) {
        dinv.resize(A.n);
        for (int i = 0; i < A.n; ++i) { const double d = A.diag(i); dinv[i] = d != 0 ? 0.72 / d : 0.0; }
    }
    void apply(const vector<double> &b, vector<double> &x) const {
        x.resize(b.size());
        for (size_t i = 0; i < b.size(); ++i) x[i] = dinv[i] * b[i];
    }
};

/* Fault injection for the tests of the failure paths, as the reference's
 * YUZEN_FAULT_AMGCL=k:m (LinearSolverBackend.cpp:628-645): the film's
 * Newton-pass solves number k .. k+m-1, counted from the setting, break down
 * (return 2) before they iterate. Process-wide; k = 0 is off. */
static long g_fault_at = 0, g_fault_count = 0, g_fault_seen = 0;
static bool fault_now() {
    return g_fault_at > 0 && ++g_fault_seen >= g_fault_at && g_fault_seen < g_fault_at + g_fault_count;
}

/* one solve through the backend contract (LinearSolverBackend.cpp:539-728) */
static int linsolve(const Csr &M, const vector<double> &rhs, vector<double> &x,
                    int precond, int scaling, int maxit, double eps, int &iter, double &resid) {
    SpM A; A.n = M.n; A.ptr = M.ptr; A.col = M.col; A.val = M.val;
    vector<double> b = rhs;
    if (scaling) {
        for (int i = 0; i < A.n; ++i) {
            const double d = M.val[M.ptr[i]];      /* the diagonal is stored first */
            if (std::fabs(d) < 1.0e-30) continue;
            const double inv = 1.0 / d;
            for (int k = A.ptr[i]; k < A.ptr[i + 1]; ++k) A.val[k] *= inv;
            b[i] *= inv;
        }
    }
    if (precond == YZ_PC_AMG) {
        Amg amg; amg.build(A);
        return bicgstab(A, b, x, amg, maxit, eps, iter, resid);
    }
    Jacobi jac; jac.build(A);
    return bicgstab(A, b, x, jac, maxit, eps, iter, resid);
}

}  // namespace yzo
#include "yz_struct_mg.h"
namespace yzo {

/* The same solve through the device's structured multigrid (yz_struct_mg.h):
 * the padded system folded onto its interior cells (ghost rows
 * d x_g - d x_src = r_g give x_g = x_src + r_g / d; an interior row's entries
 * go to the slot of their offset, those of ghost columns moving r_g / d into
 * the right-hand side), scaled by the diagonal, solved, and the ghost entries
 * of x formed again from their rows. Returns 2 for a matrix without that
 * shape. */
static int linsolve_structured(const Grid &g, const Stencil &st, const Csr &M, const vector<double> &rhs,
                               vector<double> &x, int precond, int scaling, int maxit, double eps, int &iter,
                               double &resid) {
    smg::Op op;
    op.dim = g.dim; op.nx = g.nx; op.ny = g.dim == 2 ? g.ny : 1; op.N = op.nx * op.ny; op.ns = st.ns;
    op.perx = g.bc[0] == YZ_BC_PERIODIC; op.pery = g.dim == 2 && g.bc[2] == YZ_BC_PERIODIC;
    for (int k = 0; k < st.ns; ++k) { op.ox[k] = st.ox[k]; op.oy[k] = st.oy[k]; }
    const int n = M.n, N = op.N;
    vector<int> gsrc(n, -1);
    vector<double> gofs(n, 0.0);
    for (int i = 0; i < n; ++i) {
        const int ix = i % g.nxp, iy = i / g.nxp;
        if (g.inside(ix, g.dim == 2 ? iy : 2)) continue;
        double d = 0, o = 0;
        int src = -1, m = 0;
        for (int k = M.ptr[i]; k < M.ptr[i + 1]; ++k) {
            const double v = M.val[k];
            if (v == 0) continue;
            if (M.col[k] == i) d += v;
            else { src = M.col[k]; o += v; ++m; }
        }
        if (m != 1 || d == 0 || o != -d) return 2;
        gsrc[i] = src;
        gofs[i] = rhs[i] / d;
    }
    vector<double> planes((size_t)op.ns * N, 0.0), b(N), xi(N);
    const int c0 = st.center;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad) if (N > YZO_PAR_MIN)
    for (int i = 0; i < N; ++i) {
        const int ix = i % op.nx, iy = i / op.nx;
        const int r = (g.dim == 2 ? iy + 2 : 0) * g.nxp + ix + 2;
        double acc[25] = {0};
        double bi = rhs[r];
        for (int k = M.ptr[r]; k < M.ptr[r + 1]; ++k) {
            const int cc = M.col[k];
            const double v = M.val[k];
            const int dxo = cc % g.nxp - (ix + 2), dyo = cc / g.nxp - (g.dim == 2 ? iy + 2 : 0);
            int slot = -1;
            for (int q = 0; q < op.ns; ++q) if (op.ox[q] == dxo && op.oy[q] == dyo) slot = q;
            if (slot < 0) { if (v != 0) bad = 1; continue; }
            acc[slot] += v;
            if (gsrc[cc] >= 0) bi -= v * gofs[cc];
        }
        const double d = acc[c0];
        const bool sc = scaling && !(std::fabs(d) < 1.0e-30);
        const double inv = sc ? 1.0 / d : 1.0;
        for (int q = 0; q < op.ns; ++q) planes[(size_t)q * N + i] = sc ? acc[q] * inv : acc[q];
        b[i] = sc ? bi * inv : bi;
        xi[i] = x[r];
    }
    if (bad) return 2;
    smg::Mat A;
    A.op = op; A.val = &planes; A.n = N;
    int ch;
    if (precond == YZ_PC_AMG) {
        smg::Hierarchy hier;
        hier.build(op, planes);
        smg::Pc pc;
        pc.h = &hier;
        ch = bicgstab(A, b, xi, pc, maxit, eps, iter, resid);
    } else {
        smg::JacobiPc jac;
        jac.build(op, planes);
        ch = bicgstab(A, b, xi, jac, maxit, eps, iter, resid);
    }
    for (int i = 0; i < N; ++i) {
        const int ix = i % op.nx, iy = i / op.nx;
        x[(g.dim == 2 ? iy + 2 : 0) * g.nxp + ix + 2] = xi[i];
    }
    for (int i = 0; i < n; ++i)
        if (gsrc[i] >= 0) x[i] = x[gsrc[i]] + gofs[i];
    return ch;
}

/* ------------------------------------------------------------------ */
/* Newton-Kantorovich step (NewtonKantorovich.cpp, ImplicitSolver.cpp)  */
/* ------------------------------------------------------------------ */
struct Stepper {
    Equation eq;
    vector<double> phi_old2, phi_old, phi_new, phi_guess, Ax_old, Ax_guess, Rhs, corr;
    double time = 0, dt_now = 0, dt_old = 0;
    int tsc_count = 0;
    bool extrap_blocked = false;
    double extrap_move = 0;
    double ew_prev = 0, ew_eta = 0;
    int nk_pass = 0;
    long long lin_total = 0;
    bool first_step = true;
    long long stag_cuts = 0;           /* _stagnationCuts (NewtonKantorovich.cpp:357) */
    /* the run's termination tests (CheckTermination) */
    long long steps_done = 0;
    int steady_count = 0;
    double min_start_c = 0;            /* _minStart (HeightEquationBase.h:1363-1370) */
    /* the passes of the last attempt: correction, tolerance, residual */
    vector<double> plog_corr, plog_eta, plog_lres;
    double last_eps = 0;
    /* the linear solves: 0 the reference's algebraic multigrid (Amg),
     * 1 the device's structured one (linsolve_structured) */
    int lin_mode = 0;
    int solve_sys(const Csr &M, const vector<double> &rhs, vector<double> &x, int precond, int scaling,
                  int maxit, double eps, int &it, double &r) const {
        if (lin_mode == 1) return linsolve_structured(eq.g, eq.st, M, rhs, x, precond, scaling, maxit, eps, it, r);
        return linsolve(M, rhs, x, precond, scaling, maxit, eps, it, r);
    }

    /* Scripted Newton control (test hook). In place of the assembly and
     * the linear solve, each pass takes its correction (every row alike),
     * the coupled side's correction, the norm of its right-hand side (for
     * the Eisenstat-Walker tolerance) and the solve's return code from
     * scripts, the way the reference's own test drives its Newton loop
     * (yuzen/tests/test_newton_convergence.cpp:38-84: ScriptedNK, Pick
     * by the count of solves so far, the last entry repeated). Everything
     * else -- the exits, the damping, the extrapolation and its retry, the
     * step control -- runs the code below unchanged. */
    struct Script {
        bool on = false;
        vector<double> dH, dC, rn;
        vector<int> rc;
        int call = 0;
        vector<double> eta;            /* the tolerance each solve was given */
        template <class T> T pick(const vector<T> &v) const {
            if (v.empty()) return T(0);
            const size_t k = (size_t)call;
            return k < v.size() ? v[k] : v.back();
        }
    } script;

    /* the well-mixed concentration (conc_model 1) */
    bool conc_on = false;
    vector<double> c_field;            /* the field's c (_field->CC) */
    vector<double> c0, cx;             /* the concentration equation's _phi_old, _phi_new */
    vector<double> c_iter, c_prev;     /* _cIterate, _concPrev */
    vector<double> dguess, dvec, rhsD;
    vector<double> F0[2], F1[2], G2[2], r0, r1;
    Csr CM;                            /* the concentration system (wide stencil, ghost rows) */
    double cw1[4] = {0, 0, 0, 0}, cw2[4] = {0, 0, 0, 0};   /* scheme c1, c2 */
    bool conc_lim = false;             /* a limited scheme (Koren, van Leer) */
    Limiter clim;
    long long conc_settle = 0;         /* the completing solves of a limited scheme, over the run */
    double coupled = 0;
    double cstat[5] = {0, 0, 0, 0, 0};

    /* ---- B: evaporation and the nozzle's supply ---- */
    vector<double> evap_old, src_cell, Bv;
    /* B's evaporation of the last pass (_evapB), and the evaporation and
     * supply the concentration equation takes (StepSources,
     * HeightEquationBase.h:936-960: _evpStep, _srcStep) */
    vector<double> evap_B, evp_step, src_step;
    /* Evaluate (Evaporation.cpp:62-103), uniform in space */
    double evap_eval(double hh, double cc) const {
        const yz_case &c = eq.cs;
        double f_conc;
        if (c.evap_func == YZ_EVAP_POWERN) {
            double v = 1.0 - cc - c.evap_concr;
            v = std::max(v, 0.0);
            f_conc = std::pow(v, c.evap_power);
        } else {
            const double x = (cc < 0.0) ? 0.0 : ((cc > 1.0) ? 1.0 : cc);
            const double chi = (x / c.molar_resin) / (x / c.molar_resin + (1.0 - x) / c.molar_solvent);
            f_conc = 1.0 - chi;
        }
        double f_mod = 1.0;
        if (c.precursor != 0.0) {
            const double v = (hh - c.precursor) / ((20.0 - 1.0) * c.precursor);
            f_mod = std::max(0.0, std::min(v, 1.0));
        }
        return c.evap_coef * f_conc * 1.0 * f_mod;
    }
    /* EvapAt (HeightEquationBase.h:830-846) of row i of the field (H, C) */
    double evap_at(const vector<double> &H, const vector<double> &C, int i, double dt) const {
        double evp = eq.cs.evaporation ? eq.cs.Ev * evap_eval(H[i], C[i]) : eq.cs.Ev * 0.0;
        if (evp <= 0) return evp;
        const double hh = H[i];
        if (evp * dt > 0.5 * hh) evp = (hh > 0) ? 0.5 * hh / dt : 0.0;
        const double cc = (C[i] < 0.0) ? 0.0 : ((C[i] > 1.0) ? 1.0 : C[i]);
        const double solvent = (hh > 0.0) ? hh * (1.0 - cc) : 0.0;
        if (evp * dt > solvent) evp = solvent / dt;
        return evp;
    }
    /* NozzlePos, FlowRateNow (SourceBase.cpp:54-112) at time tt (the run
     * starts at 0) */
    void nozzle(double tt, double &xn, double &yn) const {
        const yz_case &c = eq.cs;
        const double Ts = c.src_ramp, Tm = c.src_plateau;
        const double t = std::max(0.0, tt - 0.0);
        double swept;
        if (t <= Ts) swept = (Ts > 0) ? 0.5 * t * t / Ts : 0.0;
        else if (t <= Ts + Tm) swept = 0.5 * Ts + (t - Ts);
        else if (t <= 2.0 * Ts + Tm) {
            const double tau = t - Ts - Tm;
            swept = 0.5 * Ts + Tm + tau - 0.5 * tau * tau / Ts;
        } else swept = Ts + Tm;
        const double theta = c.src_angle0 + c.src_omega * swept;
        xn = c.src_arm_x + c.src_arm_length * std::cos(theta);
        yn = c.src_arm_y + c.src_arm_length * std::sin(theta);
    }
    double flow_now(double tt) const {
        const double tau = eq.cs.src_tau;
        double coeff = 1.0;
        if (!(std::fabs(tau) <= std::numeric_limits<double>::epsilon())) coeff = 1.0 - std::exp(-tt / tau);
        return eq.cs.src_rate * coeff;
    }
    /* the supply of every row at time tt: updateFraction + operator()
     * (WM2D/Source.cpp:29-126, WM1D/Source_1D.cpp:26-111) */
    void source_at(double tt) {
        const yz_case &c = eq.cs;
        const Grid &g = eq.g;
        src_cell.assign(g.n, 0.0);
        if (!c.source) return;
        double xn, yn;
        nozzle(tt, xn, yn);
        const double phi = flow_now(tt);
        const int ns = c.src_subregions;
        for (int iy = 0; iy < g.nyp; ++iy)
            for (int ix = 0; ix < g.nxp; ++ix) {
                const double xc = g.X(ix);
                double frac = 0.0;
                if (g.dim == 2) {
                    const double yc = g.Y(iy);
                    const double rr = std::sqrt(std::pow(xc - xn, 2) + std::pow(yc - yn, 2));
                    if (rr < c.src_radius + std::sqrt(g.dx * g.dx + g.dy * g.dy)) {
                        const double dxs = g.dx / (double)ns, dys = g.dy / (double)ns;
                        const double xL = xc - 0.5 * g.dx, yB = yc - 0.5 * g.dy;
                        double alpha = 0.0;
                        for (int a = 0; a < ns; a++) {
                            const double xx = xL + ((double)a + 0.5) * dxs;
                            for (int b = 0; b < ns; b++) {
                                const double yy = yB + ((double)b + 0.5) * dys;
                                const double r2 = std::sqrt(std::pow(xx - xn, 2) + std::pow(yy - yn, 2));
                                alpha += (r2 <= c.src_radius ? 1.0 : 0.0);
                            }
                        }
                        frac = alpha / (double)(ns * ns);
                    }
                } else {
                    const double rr = std::fabs(xc - xn);
                    if (rr < c.src_radius + 2.0 * g.dx) {
                        const double dxs = g.dx / (double)ns, xL = xc - 0.5 * g.dx;
                        double alpha = 0.0;
                        for (int a = 0; a < ns; a++) {
                            const double xx = xL + ((double)a + 0.5) * dxs;
                            alpha += (std::fabs(xx - xn) <= c.src_radius ? 1.0 : 0.0);
                        }
                        frac = alpha / (double)ns;
                    }
                }
                src_cell[g.row(ix, iy)] = phi * frac;
            }
    }
    /* BeforeNKLoop's part of B (:1330-1341): the evaporation at the start of
     * the step, the supply at its middle */
    void b_before() {
        const int n = eq.g.n;
        evap_old.assign(n, 0.0);
        for (int i = 0; i < n; ++i)
            if (!eq.A.bc[i]) evap_old[i] = evap_at(phi_old, eq.c, i, dt_now);
        source_at(time + 0.5 * dt_now);
    }
    /* B of the pass (SetMatrixA :765-787) at the guess */
    void b_pass() {
        const int n = eq.g.n;
        const double th = eq.cs.theta;
        Bv.assign(n, 0.0);
        evap_B.assign(n, 0.0);
        if (!eq.cs.evaporation && !eq.cs.source) return;
#pragma omp parallel for schedule(static) if (n > YZO_PAR_MIN)
        for (int i = 0; i < n; ++i) {
            if (eq.A.bc[i]) continue;
            const double evp = (1.0 - th) * evap_old[i] + th * evap_at(phi_guess, eq.c, i, dt_now);
            evap_B[i] = evp;
            Bv[i] = evp - src_cell[i];
        }
    }
    /* StepSources (HeightEquationBase.h:936-960): the evaporation at the
     * film of the pass, Crank-Nicolson with the step's start, and the
     * supply; both 0 on the boundary rows and without the terms */
    void step_sources() {
        const int n = eq.g.n;
        const double th = eq.cs.theta;
        evp_step.assign(n, 0.0);
        src_step.assign(n, 0.0);
        if (!eq.cs.evaporation && !eq.cs.source) return;
        for (int i = 0; i < n; ++i) {
            if (eq.A.bc[i]) continue;
            evp_step[i] = (1.0 - th) * evap_old[i] + th * evap_at(phi_guess, eq.c, i, dt_now);
            src_step[i] = src_cell[i];
        }
    }

    void init(const yz_case &c, const double *h, const double *s, const double *conc, double dt0) {
        eq.build(c);
        const int n = eq.g.n;
        eq.h.assign(h, h + n); eq.s.assign(s, s + n); eq.c.assign(conc, conc + n);
        eq.fill_ghosts(eq.h); eq.fill_ghosts(eq.s); eq.fill_ghosts(eq.c);
        conc_on = c.conc_model == 1;
        c_field = eq.c;
        if (conc_on) conc_init();
        phi_new = eq.h; phi_old = eq.h; phi_old2 = eq.h;
        Ax_old.assign(n, 0); Ax_guess.assign(n, 0); Rhs.assign(n, 0); corr.assign(n, 0);
        /* C_TimeSteps::Build (TimeSteps.cpp:38-61) starts DT().old at the
         * initial step, so the first step's truncation criterion is taken,
         * with phi^{n-1} = phi^n */
        dt_now = dt0; dt_old = dt0;
    }
    /* a scripted stepper over n rows, none of them a boundary row */
    void init_script(const yz_case &c, int n, const double *phi0, double dt0) {
        eq.cs = c;
        eq.g.n = n;
        eq.A.n = n;
        eq.A.bc.assign(n, 0);
        phi_new.assign(phi0, phi0 + n); phi_old = phi_new; phi_old2 = phi_new;
        Ax_old.assign(n, 0); Ax_guess.assign(n, 0); Rhs.assign(n, 0); corr.assign(n, 0);
        script.on = true;
        dt_now = dt0; dt_old = dt0;
    }
    static double absmax(const vector<double> &v) {
        double m = 0;
        for (double x : v) { const double a = std::fabs(x); if (a > m) m = a; }
        return m;
    }
    void make_IA() {
        const double th = eq.cs.theta, dt = dt_now;
        Csr &J = eq.J, &IA = eq.IA;
#pragma omp parallel for schedule(static) if (J.n > YZO_PAR_MIN)
        for (int i = 0; i < J.n; ++i) {
            const int p0 = J.ptr[i], p1 = J.ptr[i + 1];
            if (J.bc[i]) { for (int k = p0; k < p1; ++k) IA.val[k] = J.val[k]; continue; }
            IA.val[p0] = 1.0 + th * dt * J.val[p0];
            for (int k = p0 + 1; k < p1; ++k) IA.val[k] = th * dt * J.val[k];
        }
    }
    void make_rhs() {
        const double th = eq.cs.theta, dt = dt_now;
#pragma omp parallel for schedule(static) if (eq.g.n > YZO_PAR_MIN)
        for (int i = 0; i < eq.g.n; ++i) {
            if (eq.A.bc[i]) { Rhs[i] = 0; continue; }
            Rhs[i] = phi_old[i] - phi_guess[i] - (dt * th * Ax_guess[i]) - (dt * (1.0 - th) * Ax_old[i]) - dt * Bv[i];
        }
    }
    /* SatisfyBCRows for single-row sets */
    void satisfy_bc(vector<double> &x) const {
        const Csr &M = eq.IA;
        for (int i = 0; i < M.n; ++i) {
            if (!M.bc[i]) continue;
            double sum = 0;
            for (int k = M.ptr[i] + 1; k < M.ptr[i + 1]; ++k) sum += M.val[k] * x[M.col[k]];
            x[i] = (Rhs[i] - sum) / M.val[M.ptr[i]];
        }
    }
    /* ---------------- the well-mixed concentration ---------------- */
    /* the scheme's coefficients (FiniteDifferences.cpp:22-93): c1 and c2
     * of LL L R RR, over the sum of c1 */
    void conc_scheme() {
        conc_lim = eq.cs.conc_scheme == YZ_CONV_KOREN || eq.cs.conc_scheme == YZ_CONV_VANLEER;
        clim.vanleer = eq.cs.conc_scheme == YZ_CONV_VANLEER;
        if (!conc_lim) unlimited_coef(eq.cs.conc_scheme, cw1, cw2);
    }
    /* FaceWeights / FaceValue (FiniteDifferences.cpp:152-190): the fixed
     * coefficients of an unlimited scheme, or a limited one's weights frozen
     * at f with the row's cell at rowpos */
    void conc_weights(const double f[4], double sign, int rowpos, double w[4]) const {
        if (conc_lim) { clim.weights_at(f, sign, rowpos, w); return; }
        for (int j = 0; j < 4; ++j) w[j] = cw1[j] + Sign(sign) * cw2[j];
    }
    double conc_face(const double f[4], double sign) const {
        if (conc_lim) return clim.value(f, sign);
        double v = 0;
        for (int j = 0; j < 4; ++j) v += (cw1[j] + Sign(sign) * cw2[j]) * f[j];
        return v;
    }
    /* KFace (ConcEquationBase.h:146-151) */
    double kface(double ha, double hb, double ca, double cb, double dn) const {
        if (!eq.cs.conc_diffusion) return 0;
        return eq.cs.Df * std::exp(eq.cs.diff_exp * (0.5 * (ca + cb) - eq.cs.diff_c0)) * (0.5 * (ha + hb)) / (dn * dn);
    }
    bool wide_slot(int k) const {
        if (k == eq.st.center) return true;
        for (int d = 0; d < eq.st.ndir; ++d) if (eq.st.in_face(d, k)) return true;
        return false;
    }
    void conc_init() {
        conc_scheme();
        const int n = eq.g.n, ns = eq.st.ns;
        /* the pattern: ghost rows (1, -1 to the row they copy) and the
         * wide stencil of every real cell, the diagonal first
         * (ConcEquationBase.h:242-258) */
        Csr &m = CM;
        m.n = n; m.ns = ns;
        m.ptr.assign(n + 1, 0); m.col.clear(); m.val.clear();
        m.bc.assign(n, 0); m.slotpos.assign((size_t)n * ns, -1);
        for (int i = 0; i < n; ++i) {
            const int ix = i % eq.g.nxp, iy = i / eq.g.nxp;
            if (!eq.g.inside(ix, iy)) {
                m.bc[i] = 1;
                m.col.push_back(i); m.val.push_back(1.0);
                m.col.push_back(eq.g.src(ix, iy)); m.val.push_back(-1.0);
            } else {
                m.slotpos[(size_t)i * ns + eq.st.center] = (int)m.col.size();
                m.col.push_back(i); m.val.push_back(0.0);
                for (int k = 0; k < ns; ++k) {
                    if (k == eq.st.center || !wide_slot(k)) continue;
                    m.slotpos[(size_t)i * ns + k] = (int)m.col.size();
                    m.col.push_back(eq.nb(i, k)); m.val.push_back(0.0);
                }
            }
            m.ptr[i + 1] = (int)m.col.size();
        }
        c0 = c_field; cx = c_field; c_iter = c_field; c_prev = c_field;
        dguess.assign(n, 0.0); dvec.assign(n, 0.0); rhsD.assign(n, 0.0);
    }
    bool film_not_positive(const vector<double> &hh) const {
        for (int i = 0; i < eq.g.n; ++i)
            if (!eq.ghost_row(i) && !(hh[i] > 0.0)) return true;
        return false;
    }
    /* the ghost rows from the rows they copy (SatisfyBCRows on the
     * concentration's single-row sets, b = rhs) */
    void conc_bc(vector<double> &x, const vector<double> *rhs) const {
        for (int i = 0; i < eq.g.n; ++i) {
            if (!CM.bc[i]) continue;
            const int p = CM.ptr[i];
            const double b = rhs ? (*rhs)[i] : 0.0;
            x[i] = (b - CM.val[p + 1] * x[CM.col[p + 1]]) / CM.val[p];
        }
    }
    /* SoluteFlux (ConcEquationBase.h:427-457) of face d of real cell i */
    double solute_flux(int i, int d) const {
        const double th = 0.5;
        const int a = d / 2;
        const bool high = (d % 2 == 1);
        const int ia = eq.nb(i, eq.st.adj[d]);
        const int lo = high ? i : ia, hi = high ? ia : i;
        const double f1v = F1[a][lo], f0v = F0[a][lo];
        double f1[4], f0[4];
        for (int q = 0; q < 4; ++q) {
            const int j = eq.nb(i, eq.st.face[d][q]);
            f1[q] = cx[j]; f0[q] = c0[j];
        }
        double G1 = f1v * conc_face(f1, f1v);
        double G0 = f0v * conc_face(f0, f0v);
        if (eq.cs.conc_diffusion) {
            const double dn = eq.spacing(a);
            const double K1 = kface(eq.h[lo], eq.h[hi], eq.c[lo], eq.c[hi], dn);
            const double K0 = kface(phi_old[lo], phi_old[hi], c0[lo], c0[hi], dn);
            G1 -= K1 * (cx[hi] - cx[lo]);
            G0 -= K0 * (c0[hi] - c0[lo]);
        }
        return th * G1 + (1.0 - th) * G0;
    }
    /* Reconstruct (ConcEquationBase.h:459-518): c1 from the balance of
     * h c with the solute fluxes formed once per face from c1 */
    void reconstruct() {
        const double th = 0.5, dt = dt_now;
        const int n = eq.g.n;
        for (int a = 0; a < eq.st.naxis; ++a) G2[a].assign(n, 0.0);
        for (int i = 0; i < n; ++i) {
            if (CM.bc[i]) continue;
            for (int a = 0; a < eq.st.naxis; ++a) {
                G2[a][i] = solute_flux(i, 2 * a + 1);
                const int il = eq.nb(i, eq.st.adj[2 * a]);
                if (CM.bc[il]) G2[a][il] = solute_flux(i, 2 * a);
            }
        }
        const vector<double> c1 = cx;
        for (int i = 0; i < n; ++i) {
            if (CM.bc[i]) continue;
            const double h1 = phi_guess[i], h0 = phi_old[i];
            if (!(h1 > 0.0)) continue;
            double div = 0;
            for (int a = 0; a < eq.st.naxis; ++a) div += G2[a][i] - G2[a][eq.nb(i, eq.st.adj[2 * a])];
            const double cc0 = c0[i];
            double q = h0 * cc0 - dt * div;
            if (!src_step.empty())
                q += dt * src_step[i] * (eq.cs.src_conc_supplied ? eq.cs.src_conc : th * c1[i] + (1.0 - th) * cc0);
            if (eq.cs.centrifugal && eq.cs.dim == 1) q -= dt * (th * r1[i] * c1[i] + (1.0 - th) * r0[i] * cc0);
            cx[i] = q / h1;
        }
    }
    /* one solve of the concentration (ConcEquationBase.h:261-418): the
     * change d = c1 - c0 with the film of the pass (commit false) or of
     * the step (commit true); the linear solve tries damped Jacobi for at
     * most 300 iterations and falls back to the multigrid
     * (LinearSolverBackend.cpp:405-470, 587-626) */
    int conc_solve(bool commit, int &iter, double &resid) {
        const double dt = dt_now, th = 0.5;
        const int n = eq.g.n;
        for (int i = 0; i < n; ++i) {
            if (!CM.bc[i]) continue;
            double v = 0;
            for (int k = CM.ptr[i]; k < CM.ptr[i + 1]; ++k) v += CM.val[k] * c0[CM.col[k]];
            rhsD[i] = 0.0 - v;
        }
        for (int i = 0; i < n; ++i) {
            if (CM.bc[i]) continue;
            double M[13];
            for (int k = 0; k < eq.st.ns; ++k) M[k] = 0;
            const double h1 = phi_guess[i], h0 = phi_old[i];
            const double rr1 = (eq.cs.centrifugal && eq.cs.dim == 1) ? r1[i] : 0.0;
            /* the supply: at the local c, theta d of it is the new level's;
             * at the nozzle's c, all of it is known (ConcEquationBase.h
             * :310-318) */
            const bool supplied = eq.cs.source && eq.cs.src_conc_supplied;
            const double S = src_step.empty() ? 0.0 : src_step[i];
            double diag = h1 + th * dt * (rr1 - (supplied ? 0.0 : S));
            double R = dt * (evp_step.empty() ? 0.0 : evp_step[i]) * c0[i];
            if (supplied) R += dt * S * (eq.cs.src_conc - c0[i]);
            for (int d = 0; d < eq.st.ndir; ++d) {
                const int a = d / 2;
                const bool high = (d % 2 == 1);
                const double sg = high ? 1.0 : -1.0;
                const int adj = eq.st.adj[d];
                const int ia = eq.nb(i, adj);
                const int lo = high ? i : ia;
                const double f1v = F1[a][lo], f0v = F0[a][lo];
                if (eq.cs.conc_convection) {
                    /* a limited scheme's weights frozen at the latest iterate
                     * (LimiterC, ConcEquationBase.h:157-161, 336-355) */
                    double f0[4], fl[4], w1[4], w0[4];
                    int col[4];
                    for (int q = 0; q < 4; ++q) {
                        col[q] = eq.st.face[d][q];
                        f0[q] = c0[eq.nb(i, col[q])];
                        fl[q] = c_iter[eq.nb(i, col[q])];
                    }
                    /* the row's cell is L of its high face, R of its low one */
                    const int rowpos = high ? 1 : 2;
                    conc_weights(fl, f1v, rowpos, w1);
                    conc_weights(f0, f0v, rowpos, w0);
                    double e1 = 0, e0 = 0;
                    for (int q = 0; q < 4; ++q) {
                        M[col[q]] += th * dt * sg * f1v * w1[q];
                        e1 += w1[q] * (f0[q] - c0[i]);
                        e0 += w0[q] * (f0[q] - c0[i]);
                    }
                    R -= dt * sg * (th * f1v * e1 + (1.0 - th) * f0v * e0);
                } else {
                    diag += th * dt * sg * f1v;
                }
                if (eq.cs.conc_diffusion) {
                    const double dn = eq.spacing(a);
                    const double K1 = kface(eq.h[i], eq.h[ia], eq.c[i], eq.c[ia], dn);
                    const double K0 = kface(h0, phi_old[ia], c0[i], c0[ia], dn);
                    diag += th * dt * K1;
                    M[adj] -= th * dt * K1;
                    const double dc0 = high ? (c0[ia] - c0[i]) : (c0[i] - c0[ia]);
                    R += dt * sg * (th * K1 * dc0 + (1.0 - th) * K0 * dc0);
                }
            }
            M[eq.st.center] += diag;
            for (int k = 0; k < eq.st.ns; ++k)
                if (wide_slot(k)) CM.at(i, k) = M[k];
            rhsD[i] = R;
        }
        dvec = dguess;
        const vector<double> x0 = dvec;
        const int maxit = eq.cs.conc_lin_maxit;
        int it = 0;
        double r = 0;
        int ch = solve_sys(CM, rhsD, dvec, YZ_PC_JACOBI, eq.cs.conc_row_scaling, std::min(maxit, 300),
                          eq.cs.conc_lin_eps, it, r);
        cstat[0] += 1;
        if (ch != 0) {
            const int it_try = it;
            dvec = x0;
            ch = solve_sys(CM, rhsD, dvec, YZ_PC_AMG, eq.cs.conc_row_scaling, maxit, eq.cs.conc_lin_eps, it, r);
            it += it_try;
            cstat[2] += 1;
        }
        cstat[1] += it;
        if (std::getenv("YZO_DEBUG"))
            std::fprintf(stderr, "[oracle] conc solve commit %d: iterations %d resid %.3e rc %d\n", (int)commit, it, r, ch);
        if (std::getenv("YZO_DEBUG") && ch != 0) {
            double dmin = 1e300, worst = 1e300, hmin = 1e300, fmax = 0;
            int ndd = 0;
            for (int i = 0; i < n; ++i) {
                if (CM.bc[i]) continue;
                const double dg = CM.val[CM.ptr[i]];
                double off = 0;
                for (int k = CM.ptr[i] + 1; k < CM.ptr[i + 1]; ++k) off += std::fabs(CM.val[k]);
                dmin = std::min(dmin, dg);
                worst = std::min(worst, (std::fabs(dg) - off) / std::fabs(dg));
                if (std::fabs(dg) < off) ++ndd;
                hmin = std::min(hmin, phi_guess[i]);
                for (int a = 0; a < eq.st.naxis; ++a) fmax = std::max(fmax, std::fabs(F1[a][i]));
            }
            std::fprintf(stderr, "[oracle]   failed system: min diag %.3e, rows not diagonally dominant %d, worst margin %.3e, "
                         "min h %.3e, max |F1| %.3e\n", dmin, ndd, worst, hmin, fmax);
        }
        iter = it; resid = r;
        if (ch == 0) {
            conc_bc(dvec, &rhsD);
            dguess = dvec;
        } else {
            bool finite = true;
            for (int i = 0; i < n && finite; ++i) finite = std::isfinite(dvec[i]);
            if (!finite) std::fill(dvec.begin(), dvec.end(), 0.0);
            std::fill(dguess.begin(), dguess.end(), 0.0);
        }
        for (int i = 0; i < n; ++i) cx[i] = c0[i] + dvec[i];
        if (ch == 0 && commit && eq.cs.conc_convection) reconstruct();
        conc_bc(cx, nullptr);
        return ch;
    }
    /* BeforeNKLoop's concentration part (HeightEquationBase.h:1314-1388) */
    void conc_before() {
        eq.c = c_field;
        c_iter = c_field;
        c_prev = c_field;
        eq.h = phi_old;
        eq.face_volume_fluxes(F0, r0);
        std::fill(dguess.begin(), dguess.end(), 0.0);
        coupled = 0;
    }
    void clamp_conc() {
        for (double &v : eq.c) if (v > 1.0) v = 1.0;
    }
    /* EveryNKLoop (HeightEquationBase.h:1391-1444) */
    bool conc_every() {
        eq.h = phi_guess;
        if (film_not_positive(phi_guess)) { coupled = 0; return true; }
        eq.face_volume_fluxes(F1, r1);
        step_sources();
        int it = 0; double r = 0;
        if (conc_solve(false, it, r) != 0) return false;
        eq.c = cx;
        c_iter = cx;
        clamp_conc();
        double dmax = 0;
        for (int i = 0; i < eq.g.n; ++i) {
            const double d = std::fabs(eq.c[i] - c_prev[i]);
            if (d > dmax) dmax = d;
            c_prev[i] = eq.c[i];
        }
        coupled = dmax;
        return true;
    }
    /* ConsistentFluxes (HeightEquationBase.h:985-1058) */
    void consistent_fluxes(double dt) {
        const int n = eq.g.n;
        const double th = eq.cs.theta;
        vector<double> rho(n, 0.0);
        double sum = 0;
        int cnt = 0;
        for (int i = 0; i < n; ++i) {
            if (CM.bc[i]) continue;
            double div1 = 0, div0 = 0;
            for (int a = 0; a < eq.st.naxis; ++a) {
                const int il = eq.nb(i, eq.st.adj[2 * a]);
                div1 += F1[a][i] - F1[a][il];
                div0 += F0[a][i] - F0[a][il];
            }
            const double es = evp_step.empty() ? 0.0 : evp_step[i], ss = src_step.empty() ? 0.0 : src_step[i];
            const double r = (phi_old[i] - phi_guess[i]) / dt
                           - (th * (div1 + r1[i]) + (1.0 - th) * (div0 + r0[i]) + (es - ss));
            rho[i] = r;
            sum += r;
            ++cnt;
        }
        if (cnt == 0) return;
        const double mean = sum / cnt;
        const int nx = eq.g.nxp, ny = eq.g.nyp;
        const double w = 1.0 / th;
        if (eq.cs.dim == 1) {
            double acc = 0;
            for (int ix = 2; ix < nx - 3; ix++) { acc += rho[ix] - mean; F1[0][ix] += w * acc; }
            return;
        }
        const int m = nx - 4;
        vector<double> rowSum(ny, 0.0);
        for (int jy = 2; jy <= ny - 3; jy++) {
            double r = 0;
            for (int ix = 2; ix <= nx - 3; ix++) r += rho[jy * nx + ix] - mean;
            rowSum[jy] = r;
        }
        double accY = 0;
        for (int jy = 2; jy < ny - 3; jy++) {
            accY += rowSum[jy] / m;
            for (int ix = 2; ix <= nx - 3; ix++) F1[1][jy * nx + ix] += w * accY;
        }
        for (int jy = 2; jy <= ny - 3; jy++) {
            const double share = rowSum[jy] / m;
            double acc = 0;
            for (int ix = 2; ix < nx - 3; ix++) {
                acc += (rho[jy * nx + ix] - mean) - share;
                F1[0][jy * nx + ix] += w * acc;
            }
        }
    }
    /* MoveExcess (HeightEquationBase.h:1172-1273) on the film phi_guess */
    void move_excess(vector<double> &c, bool lower) {
        const vector<double> &hh = phi_guess;
        const int n = eq.g.n;
        const double bound = lower ? 0.0 : 1.0;
        auto outside = [lower](double v) { return lower ? (v < 0.0) : (v > 1.0); };
        auto beyond = [lower](double h, double v) { return lower ? -h * v : h * (v - 1.0); };
        auto roomOf = [lower](double h, double v) { return lower ? h * v : h * (1.0 - v); };
        double &moved = cstat[lower ? 4 : 3];
        vector<double> room(n, 0.0), gain(n, 0.0);
        bool any = true;
        const int ndir = eq.st.ndir;
        for (int pass = 0; pass < 4 && any; ++pass) {
            for (int i = 0; i < n; ++i) {
                const double h = hh[i];
                const bool inside = lower ? (c[i] > 0.0) : (c[i] < 1.0);
                room[i] = (!CM.bc[i] && h > 0.0 && inside) ? roomOf(h, c[i]) : 0.0;
                gain[i] = 0.0;
            }
            any = false;
            for (int i = 0; i < n; ++i) {
                if (CM.bc[i] || !outside(c[i])) continue;
                const double h = hh[i];
                if (!(h > 0.0)) { c[i] = bound; continue; }
                const double e = beyond(h, c[i]);
                if (pass == 0) moved += e;
                double r = 0.0;
                for (int d = 0; d < ndir; ++d) r += room[eq.nb(i, eq.st.adj[d])];
                if (!(r > 0.0)) { any = true; continue; }
                const double f = (e < r) ? e / r : 1.0;
                for (int d = 0; d < ndir; ++d) {
                    const int j = eq.nb(i, eq.st.adj[d]);
                    gain[j] += f * room[j];
                }
                if (f < 1.0) c[i] = bound;
                else { c[i] = lower ? -(e - r) / h : 1.0 + (e - r) / h; any = true; }
            }
            for (int i = 0; i < n; ++i) {
                if (gain[i] > 0.0) {
                    const double dc = gain[i] / hh[i];
                    c[i] += lower ? -dc : dc;
                    if (outside(c[i])) any = true;
                }
            }
        }
        if (!any) return;
        double left = 0.0, total = 0.0;
        for (int i = 0; i < n; ++i) {
            if (CM.bc[i]) continue;
            const double h = hh[i];
            if (outside(c[i])) { left += beyond(h, c[i]); c[i] = bound; }
            else if (h > 0.0) total += roomOf(h, c[i]);
        }
        const double f = (total > 0.0) ? std::min(1.0, left / total) : 0.0;
        if (f > 0.0)
            for (int i = 0; i < n; ++i) {
                if (CM.bc[i] || !(hh[i] > 0.0)) continue;
                if (lower) { if (!(c[i] > 0.0)) continue; c[i] -= f * c[i]; if (c[i] < 0.0) c[i] = 0.0; }
                else { if (!(c[i] < 1.0)) continue; c[i] += f * (1.0 - c[i]); if (c[i] > 1.0) c[i] = 1.0; }
            }
    }
    /* RedistributeExcess (HeightEquationBase.h:1128-1170) */
    void redistribute(vector<double> &c) {
        cstat[3] = cstat[4] = 0;
        bool over = false, under = false;
        for (int i = 0; i < eq.g.n; ++i) {
            if (CM.bc[i]) continue;
            if (c[i] > 1.0) over = true;
            if (c[i] < 0.0) under = true;
        }
        if (over) move_excess(c, false);
        if (under) move_excess(c, true);
        if (over || under) conc_bc(c, nullptr);
        for (double &v : c) {
            if (v > 1.0) v = 1.0;
            if (under && v < 0.0) v = 0.0;
        }
    }
    /* AfterNKLoop's concentration part (HeightEquationBase.h:1447-1577) */
    void conc_after() {
        eq.h = phi_guess;
        eq.face_volume_fluxes(F1, r1);
        step_sources();
        /* the evaporation as B took it in the last pass, the one the film's
         * step holds (HeightEquationBase.h:1497-1500) */
        if (eq.cs.evaporation && evap_B.size() == (size_t)eq.g.n) evp_step = evap_B;
        consistent_fluxes(dt_now);
        const bool noFilm = film_not_positive(phi_guess);
        const vector<double> lastPass = cx;
        bool solved = true;
        /* a limited scheme freezes its limiter at the last iterate: the
         * completing solve is repeated, frozen at its own answer, until that
         * settles (HeightEquationBase.h:1506-1525) */
        const bool limited = eq.cs.conc_convection && conc_lim;
        for (int k = 0; k < (limited ? 8 : 1) && !noFilm; ++k) {
            int it = 0; double r = 0;
            if (conc_solve(true, it, r) != 0) { solved = false; break; }
            if (!limited) break;
            double dmax = 0;
            for (size_t i = 0; i < cx.size(); ++i) dmax = std::max(dmax, std::fabs(cx[i] - c_iter[i]));
            c_iter = cx;
            ++conc_settle;
            if (dmax <= 1e-12) break;
        }
        if (!solved) cx = lastPass;
        redistribute(cx);
        c0 = cx;
        eq.c = cx;
        c_iter = cx;
        clamp_conc();
        c_field = eq.c;
    }

    /* Solve (NewtonKantorovich.cpp:74-154) */
    int solve(int &iter, double &resid) {
        if (script.on) return solve_scripted(iter, resid);
        eq.h = phi_guess;
        eq.set_matrix_A();
        if (nk_pass % eq.cs.jac_interval == 0) { eq.set_matrix_J(); make_IA(); }
        ++nk_pass;
        eq.apply_A(phi_guess, Ax_guess);
        b_pass();
        make_rhs();
        double eps = eq.cs.lin_eps;
        if (eq.cs.ew) eps = ew_tolerance(absmax(Rhs));
        last_eps = eps;
        /* the previous pass's correction as the guess (the reference's), or
         * zero (zero_guess 1; the oracle is the FP64 path, so -1 is the
         * reference's) */
        if (eq.cs.zero_guess == 1) std::fill(corr.begin(), corr.end(), 0.0);
        if (fault_now()) {
            /* a breakdown made to happen before the solve iterates */
            iter = 0; resid = 0;
            std::fill(corr.begin(), corr.end(), 0.0);
            return 2;
        }
        const int ch = solve_sys(eq.IA, Rhs, corr, eq.cs.precond == YZ_PC_JACOBI ? YZ_PC_JACOBI : YZ_PC_AMG,
                                eq.cs.row_scaling, eq.cs.lin_maxit, eps, iter, resid);
        if (ch == 0) satisfy_bc(corr);
        if (ch != 0) std::fill(corr.begin(), corr.end(), 0.0);
        return ch;
    }
    /* Eisenstat-Walker choice 2 (NewtonKantorovich.cpp:108-126) on the
     * norm rn of this pass's right-hand side; pass counts from 1 */
    double ew_tolerance(double rn) {
        const double kEtaMax = 1.0e-4;
        double eta = kEtaMax;
        if (nk_pass > 1 && ew_prev > 0) {
            const double ratio = rn / ew_prev;
            eta = 0.9 * ratio * ratio;
            const double safe = 0.9 * ew_eta * ew_eta;
            if (safe > 0.1 && safe > eta) eta = safe;
        }
        if (eta > kEtaMax) eta = kEtaMax;
        if (eta < eq.cs.lin_eps) eta = eq.cs.lin_eps;
        ew_eta = eta; ew_prev = rn;
        return eta;
    }
    /* Solve of the scripted stepper: the tolerance is formed as in solve(),
     * the correction is the script's; a failed solve leaves no correction
     * (NewtonKantorovich.cpp:137-152) */
    int solve_scripted(int &iter, double &resid) {
        ++nk_pass;
        double eps = eq.cs.lin_eps;
        if (eq.cs.ew) eps = ew_tolerance(script.pick(script.rn));
        script.eta.push_back(eps);
        last_eps = eps;
        const double d = script.pick(script.dH);
        const int ch = script.pick(script.rc);
        std::fill(corr.begin(), corr.end(), ch == 0 ? d : 0.0);
        iter = 1;
        resid = 0.0;
        ++script.call;
        return ch;
    }
    /* NKLoop (:157-422) */
    bool nk_loop(int &iter, double &resid_eq) {
        nk_pass = 0;
        int lin_step = 0;
        plog_corr.clear(); plog_eta.clear(); plog_lres.clear();
        phi_guess = phi_old;
        extrap_move = 0;
        if (eq.cs.extrap && !extrap_blocked && !first_step) {
            const double r = (dt_old > 0) ? dt_now / dt_old : 0.0;
            bool ok = (r > 0);
            double move = 0;
            for (size_t i = 0; i < phi_old.size() && ok; ++i) {
                const double gg = phi_old[i] + r * (phi_old[i] - phi_old2[i]);
                if (!eq.A.bc[i] && phi_old[i] > 0 && !(gg > 0)) ok = false;
                const double m = std::fabs(gg - phi_old[i]);
                if (m > move) move = m;
            }
            if (ok && move > 0) {
                for (size_t i = 0; i < phi_old.size(); ++i) phi_guess[i] = phi_old[i] + r * (phi_old[i] - phi_old2[i]);
                extrap_move = move;
            }
        }
        coupled = 0;
        if (!script.on) {
            /* BeforeNKLoop: the working field from the state of the step */
            if (conc_on) conc_before(); else eq.c = c_field;
            if (conc_on) {
                /* the lowest c of a real cell at the start of the step */
                min_start_c = 0;
                for (int i = 0; i < eq.g.n; ++i) {
                    const double cv = c_field[i];
                    if (!eq.A.bc[i] && (cv < min_start_c || !(cv == cv))) min_start_c = cv;
                }
            }
            eq.h = phi_old;
            eq.set_matrix_A();
            eq.apply_A(phi_old, Ax_old);
            b_before();
        }

        double resid_eq_old = 1e40, resid_first = 0, resid_best = 1e40;
        int stall = 0;
        const int stall_limit = 5;
        bool done = false;
        int outer;
        resid_eq = 0;
        for (outer = 0; outer < eq.cs.nk_max; ++outer) {
            if (script.on) coupled = script.pick(script.dC);
            else if (conc_on && !conc_every()) { done = false; break; }
            int it = 0; double lr = 0;
            const int ch = solve(it, lr);
            lin_step += it;
            resid_eq = absmax(corr);
            if (coupled > resid_eq) resid_eq = coupled;
            plog_corr.push_back(resid_eq); plog_eta.push_back(last_eps); plog_lres.push_back(lr);
            if (std::getenv("YZO_DEBUG"))
                std::fprintf(stderr, "[oracle] pass %d lin %d resid_lin %.3e corr %.3e rc %d\n",
                             outer, it, lr, resid_eq, ch);
            if (ch != 0) { done = false; break; }
            if (outer == 0) resid_first = resid_eq;
            if (resid_eq < resid_best * 0.9) { resid_best = resid_eq; stall = 0; }
            else stall++;
            if (stall > stall_limit) {
                const double floorPhi = 1000.0 * std::numeric_limits<double>::epsilon() * absmax(phi_guess);
                done = (resid_eq < eq.cs.nk_eps) || (resid_first > 0 && resid_eq < resid_first * 1.0e-4)
                    || (resid_eq <= floorPhi);
                if (!done) stag_cuts++;
                break;
            }
            const double alpha = (resid_eq > resid_eq_old) ? 0.5 : 1.0;
            for (size_t i = 0; i < phi_guess.size(); ++i) phi_guess[i] += corr[i] * alpha;
            if (outer >= eq.cs.nk_min - 1) {
                if (resid_eq < eq.cs.nk_eps) { done = true; break; }
                if (eq.cs.nk_eps_rel > 0 && resid_first > 0 && resid_eq < resid_first * eq.cs.nk_eps_rel) { done = true; break; }
            }
            resid_eq_old = resid_eq;
        }
        iter = outer + 1;
        lin_total = lin_step;
        return done;
    }
    /* Math::IsZero (MyMath_I.h:20-28): at or below the smallest normal */
    static bool is_zero(double v) { return std::fabs(v) <= std::numeric_limits<double>::min(); }
    /* LocalTruncationCriterion (ImplicitSolver.cpp:271-307) */
    double truncation_criterion() const {
        double th = eq.cs.tsc_criterion;
        if (is_zero(th)) th = 1.0e-2;
        if (is_zero(dt_old)) return 1.0;
        const double wfloor = eq.cs.tsc_floor * absmax(phi_old);
        double ek = 0;
        for (size_t i = 0; i < phi_guess.size(); ++i) {
            const double hm = phi_old2[i], ho = phi_old[i], hp = phi_guess[i];
            double w = std::fabs(ho);
            if (w < wfloor) w = wfloor;
            const double den = (dt_old + dt_now) * w;
            const double e = is_zero(den) ? 0.0 : 2.0 * (dt_now / dt_old) * (dt_old * hp + dt_now * hm - (dt_old + dt_now) * ho) / den;
            ek = std::max(ek, std::fabs(e));
        }
        return 1.0 - ek / th;
    }
    /* TimeAdvance (:424-486) with TimeStepAdjustment (ImplicitSolver.cpp:144-190) */
    int advance(int &iter, int &lin, double &resid) {
        phi_old2 = phi_old;
        phi_old = phi_new;
        bool step_done = false;
        bool ok = false;
        double dt_request = dt_now;
        long long lin_sum = 0;
        do {
            const long long cuts0 = stag_cuts;
            ok = nk_loop(iter, resid);
            lin_sum += lin_total;
            if (!ok && extrap_move > 0) {
                stag_cuts = cuts0;
                extrap_blocked = true;
                ok = nk_loop(iter, resid);
                lin_sum += lin_total;
                extrap_blocked = false;
            }
            if (eq.cs.step_control) {
                const double crit = ok ? truncation_criterion() : -1.0;
                if (crit > 0.0) {
                    step_done = true;
                    if (crit > 0.90 && dt_now < eq.cs.dt_max) {
                        if (++tsc_count >= eq.cs.tsc_keep) {
                            tsc_count = 0;
                            dt_request = std::min(dt_now / eq.cs.tsc_change, eq.cs.dt_max);
                        }
                    }
                } else {
                    dt_now *= eq.cs.tsc_change;
                    dt_request = dt_now;
                    if (dt_now < eq.cs.dt_min) { lin = (int)lin_sum; return 1; }
                }
            } else {
                if (!ok) { lin = (int)lin_sum; return 1; }
                step_done = true;
            }
        } while (!step_done);
        phi_new = phi_guess;
        if (conc_on) conc_after();
        if (!script.on) eq.h = phi_new;
        time += dt_now;
        dt_old = dt_now;
        dt_now = dt_request;
        first_step = false;
        ++steps_done;
        lin = (int)lin_sum;
        return 0;
    }
    /* ChangeRateOfVariable (ImplicitSolver.cpp:225-240): the relative rate
     * of change per unit time of the step just taken, over every row */
    double change_rate() const {
        if (steps_done < 5) return 1;
        const double dt = dt_old;                  /* LastStepDT */
        if (is_zero(dt)) return 1;
        double dmax = 0;
        for (size_t i = 0; i < phi_new.size(); ++i) dmax = std::max(dmax, std::fabs(phi_new[i] - phi_old[i]));
        const double scale = absmax(phi_new);
        if (is_zero(scale)) return 0;
        return dmax / (scale * dt);
    }
    /* CheckTermination (WellMixedModelBase.h:1567-1679) after a step, with
     * FieldInfo over the real cells (FieldInfo.cpp:52-136); info8 as
     * yz_step_term_status */
    int check_termination(const yz_term &t, double *info) {
        const Grid &g = eq.g;
        double chr = 0;
        if (t.steady_eps > 0) {
            double tspan = t.end_time - 0.0;
            if (!(tspan > 0)) tspan = 1.0;
            chr = change_rate() * tspan;
            if (chr < t.steady_eps) steady_count++;
            else steady_count = 0;
        }
        const double dA = g.dim == 2 ? g.dx * g.dy : g.dx;
        double vol = 0, solute = 0, minH = 1e40, maxC = -1e40, minC = 1e40;
        for (int i = 0; i < g.n; ++i) {
            if (eq.A.bc[i]) continue;
            const double hh = phi_new[i], cc = c_field[i];
            vol += hh * dA;
            solute += cc * hh * dA;
            minH = std::min(minH, hh);
            maxC = std::max(maxC, cc);
            minC = std::min(minC, cc);
        }
        const double ave = solute / vol;
        if (conc_on) minC = std::min(minC, min_start_c);
        if (info) {
            info[0] = vol; info[1] = ave; info[2] = minH; info[3] = maxC; info[4] = minC;
            info[5] = chr; info[6] = steady_count; info[7] = (double)steps_done;
        }
        if (!std::isfinite(vol) || !std::isfinite(maxC) || !std::isfinite(minC) || !std::isfinite(minH)) return YZ_RUN_NAN;
        if (conc_on && (maxC > 1.0 + 0.1 || minC < -0.1)) return YZ_RUN_CONC_RANGE;
        if (minH < -1.0e-3) return YZ_RUN_NEGATIVE_FILM;
        if (vol < 0.0) return YZ_RUN_NEGATIVE_VOLUME;
        const bool running = (t.end_time <= 0 || (0.0 <= time && time <= t.end_time)) &&
                             (t.max_steps <= 0 || steps_done <= t.max_steps - 1);
        if (!running) return YZ_RUN_END;
        if (t.dry) {
            const double v = t.end_conc_type == 0 ? ave : (t.end_conc_type == 1 ? maxC : minC);
            if (v > t.end_conc_value) return YZ_RUN_DRY;
        }
        if (t.steady_eps > 0 && steady_count > t.steady_steps) return YZ_RUN_STEADY;
        return YZ_RUN_GOING;
    }
};

static thread_local std::string g_err;

}  // namespace yzo

/* ------------------------------------------------------------------ */
/* C API of the oracle (ctypes)                                         */
/* ------------------------------------------------------------------ */
extern "C" {

/* the fault injection of the film's Newton-pass solves (fault_now):
 * solves k .. k + m - 1 from now break down; k = 0 turns it off */
void yzo_set_fault(long k, long m) {
    yzo::g_fault_at = k; yzo::g_fault_count = m > 0 ? m : 1; yzo::g_fault_seen = 0;
}

/* the threads of the row loops: n > 0 sets them; returns the number in
 * force (1 when built without OpenMP) */
int yzo_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

void yzo_case_defaults(yz_case *c) {
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
    c->learn_passes = 2; c->learn_iters = 100; c->learn_probes = 4; c->learn_lr = 2e-3;
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

void *yzo_create(const yz_case *c, const double *h, const double *s, const double *conc, double dt0) {
    auto *st = new yzo::Stepper();
    st->init(*c, h, s, conc, dt0);
    return st;
}
/* a stepper of the scripted Newton control (test hook, see Stepper::Script):
 * n rows starting at phi0; the case gives the Newton and step-control
 * numerics */
void *yzo_script_create(const yz_case *c, int n, const double *phi0, double dt0,
                        const double *dH, int ndH, const double *dC, int ndC,
                        const double *rn, int nrn, const int *rc, int nrc) {
    auto *st = new yzo::Stepper();
    st->init_script(*c, n, phi0, dt0);
    st->script.dH.assign(dH, dH + ndH);
    st->script.dC.assign(dC, dC + ndC);
    st->script.rn.assign(rn, rn + nrn);
    st->script.rc.assign(rc, rc + nrc);
    return st;
}
/* solves so far, stagnation cuts, and the tolerances the solves were given
 * (up to `cap` of them); returns how many there were */
int yzo_script_info(void *p, int *calls, long long *cuts, double *eta, int cap) {
    auto *st = static_cast<yzo::Stepper *>(p);
    if (calls) *calls = st->script.call;
    if (cuts) *cuts = st->stag_cuts;
    const int m = (int)st->script.eta.size();
    for (int i = 0; i < m && i < cap; ++i) eta[i] = st->script.eta[i];
    return m;
}
/* the termination tests after the last step (yz_step_term_status) */
int yzo_check(void *p, const yz_term *t, double *info8) {
    return static_cast<yzo::Stepper *>(p)->check_termination(*t, info8);
}
/* the passes of the last step's accepted attempt (yz_step_pass_log) */
int yzo_pass_log(void *p, double *corr, double *eta, double *lres, int cap) {
    auto *st = static_cast<yzo::Stepper *>(p);
    const int m = (int)st->plog_corr.size();
    for (int i = 0; i < m && i < cap; ++i) {
        if (corr) corr[i] = st->plog_corr[i];
        if (eta) eta[i] = st->plog_eta[i];
        if (lres) lres[i] = st->plog_lres[i];
    }
    return m;
}
void yzo_destroy(void *p) { delete static_cast<yzo::Stepper *>(p); }
/* the stepper's linear solves: 0 the reference's algebraic multigrid,
 * 1 the device's structured multigrid */
void yzo_set_lin_mode(void *p, int mode) { static_cast<yzo::Stepper *>(p)->lin_mode = mode; }
int yzo_n(void *p) { return static_cast<yzo::Stepper *>(p)->eq.g.n; }
int yzo_advance(void *p, int *nk, int *lin, double *resid) {
    auto *st = static_cast<yzo::Stepper *>(p);
    int it = 0, li = 0; double r = 0;
    const int rc = st->advance(it, li, r);
    if (nk) *nk = it;
    if (lin) *lin = li;
    if (resid) *resid = r;
    return rc;
}
void yzo_get_h(void *p, double *h) {
    auto *st = static_cast<yzo::Stepper *>(p);
    std::copy(st->phi_new.begin(), st->phi_new.end(), h);
}
void yzo_set_h(void *p, const double *h) {
    auto *st = static_cast<yzo::Stepper *>(p);
    st->phi_new.assign(h, h + st->eq.g.n);
    st->eq.fill_ghosts(st->phi_new);
}
void yzo_get_c(void *p, double *c) {
    auto *st = static_cast<yzo::Stepper *>(p);
    std::copy(st->c_field.begin(), st->c_field.end(), c);
}
/* FaceVolumeFluxes of film h with concentration c: F[2 n] (axis-major), r[n] */
void yzo_face_fluxes(void *p, const double *h, const double *c, double *F, double *r) {
    auto *st = static_cast<yzo::Stepper *>(p);
    const int n = st->eq.g.n;
    const std::vector<double> keep_h = st->eq.h, keep_c = st->eq.c;
    st->eq.h.assign(h, h + n); st->eq.c.assign(c, c + n);
    st->eq.fill_ghosts(st->eq.h); st->eq.fill_ghosts(st->eq.c);
    std::vector<double> FF[2], rr;
    st->eq.face_volume_fluxes(FF, rr);
    for (int a = 0; a < st->eq.st.naxis; ++a) std::copy(FF[a].begin(), FF[a].end(), F + (size_t)a * n);
    if (r) std::copy(rr.begin(), rr.end(), r);
    st->eq.h = keep_h; st->eq.c = keep_c;
}
/* the concentration system of the last solve (unscaled CRS, the diagonal
 * first; ghost rows (1, -1)) and its right-hand side */
int yzo_conc_nnz(void *p) { return (int)static_cast<yzo::Stepper *>(p)->CM.col.size(); }
void yzo_conc_system(void *p, int *rowptr, int *col, double *val, double *rhs) {
    auto *st = static_cast<yzo::Stepper *>(p);
    const yzo::Csr &M = st->CM;
    std::copy(M.ptr.begin(), M.ptr.end(), rowptr);
    std::copy(M.col.begin(), M.col.end(), col);
    std::copy(M.val.begin(), M.val.end(), val);
    std::copy(st->rhsD.begin(), st->rhsD.end(), rhs);
}
/* an unlimited face scheme (conc_scheme): out = FaceValue, FaceWeights[4] */
void yzo_conc_face(int scheme, const double *f, double sign, int rowpos, double *out) {
    yzo::Stepper st;
    st.eq.cs.conc_scheme = scheme;
    st.conc_scheme();
    out[0] = st.conc_face(f, sign);
    st.conc_weights(f, sign, rowpos, out + 1);
}
/* [0] concentration solves, [1] their iterations, [2] solves that needed
 * the multigrid, [3] solute moved above 1, [4] below 0 (last step) */
void yzo_conc_stats(void *p, double *out5) {
    auto *st = static_cast<yzo::Stepper *>(p);
    for (int k = 0; k < 5; ++k) out5[k] = st->cstat[k];
}
/* the solves that completed the steps of a limited concentration scheme,
 * over the run (each step at least one, at most 8) */
long long yzo_conc_settle(void *p) { return static_cast<yzo::Stepper *>(p)->conc_settle; }
void yzo_clock(void *p, double *t, double *dt_now, double *dt_old) {
    auto *st = static_cast<yzo::Stepper *>(p);
    if (t) *t = st->time;
    if (dt_now) *dt_now = st->dt_now;
    if (dt_old) *dt_old = st->dt_old;
}
int yzo_nnz(void *p) { return (int)static_cast<yzo::Stepper *>(p)->eq.A.col.size(); }
/* CRS structure (rowptr n+1, col nnz) and the slot position table */
void yzo_pattern(void *p, int *rowptr, int *col, char *bc) {
    auto *st = static_cast<yzo::Stepper *>(p);
    const yzo::Csr &A = st->eq.A;
    std::copy(A.ptr.begin(), A.ptr.end(), rowptr);
    std::copy(A.col.begin(), A.col.end(), col);
    if (bc) std::copy(A.bc.begin(), A.bc.end(), bc);
}
/* Assemble at film h with the explicit half at h_old: A, J, IA values
 * (nnz each), ApplyA(h) (n), Rhs (n). Any pointer may be NULL. */
void yzo_assemble(void *p, const double *h_old, const double *h, double dt,
                  double *A, double *J, double *IA, double *ax, double *rhs) {
    auto *st = static_cast<yzo::Stepper *>(p);
    const int n = st->eq.g.n;
    std::vector<double> ho(h_old, h_old + n), hg(h, h + n);
    st->eq.fill_ghosts(ho); st->eq.fill_ghosts(hg);
    const double save_dt = st->dt_now;
    st->dt_now = dt;
    st->phi_old = ho;
    st->eq.h = ho; st->eq.set_matrix_A(); st->eq.apply_A(ho, st->Ax_old);
    st->phi_guess = hg;
    st->b_before();
    st->b_pass();
    st->eq.h = hg; st->eq.set_matrix_A(); st->eq.set_matrix_J(); st->make_IA();
    st->eq.apply_A(hg, st->Ax_guess);
    st->make_rhs();
    const int nnz = (int)st->eq.A.val.size();
    if (A) std::copy(st->eq.A.val.begin(), st->eq.A.val.begin() + nnz, A);
    if (J) std::copy(st->eq.J.val.begin(), st->eq.J.val.begin() + nnz, J);
    if (IA) std::copy(st->eq.IA.val.begin(), st->eq.IA.val.begin() + nnz, IA);
    if (ax) std::copy(st->Ax_guess.begin(), st->Ax_guess.end(), ax);
    if (rhs) std::copy(st->Rhs.begin(), st->Rhs.end(), rhs);
    st->dt_now = save_dt;
}
/* the face velocities of the field (h, c) (padded): vx on the Xm face and
 * vy on the Ym face of every cell; bc = 0 without the boundary rows */
void yzo_velocity(void *p, const double *h, const double *c, int bc, double *vx, double *vy) {
    auto *st = static_cast<yzo::Stepper *>(p);
    auto &eq = st->eq;
    const int n = eq.g.n;
    const std::vector<double> hs = eq.h, cs = eq.c;
    eq.h.assign(h, h + n); eq.c.assign(c, c + n);
    eq.fill_ghosts(eq.h); eq.fill_ghosts(eq.c);
    std::vector<double> ux, uy;
    eq.velocity(ux, uy, bc != 0);
    std::copy(ux.begin(), ux.end(), vx);
    if (vy) std::copy(uy.begin(), uy.end(), vy);
    eq.h = hs; eq.c = cs;
}
/* the velocity fields of the .vts (Field_WM1D_Out / Field_WM2D_Out,
 * LubFields.cpp:370-383, 832-846): vx6, vy6 six padded fields each, the
 * mixed one, then the capillary, Marangoni, vertical-gravity, in-plane
 * gravity and centrifugal terms' own, each with its boundary rows */
void yzo_velocity_terms(void *p, const double *h, const double *c, double *vx6, double *vy6) {
    auto *st = static_cast<yzo::Stepper *>(p);
    auto &eq = st->eq;
    const int n = eq.g.n;
    const std::vector<double> hs = eq.h, cs = eq.c;
    eq.h.assign(h, h + n); eq.c.assign(c, c + n);
    eq.fill_ghosts(eq.h); eq.fill_ghosts(eq.c);
    std::vector<double> ux, uy, t[6][2];
    eq.velocity(ux, uy, true, t);
    std::copy(ux.begin(), ux.end(), vx6);
    if (vy6) std::copy(uy.begin(), uy.end(), vy6);
    for (int k = 0; k < 5; ++k) {
        std::copy(t[k][0].begin(), t[k][0].end(), vx6 + (size_t)(k + 1) * n);
        if (vy6) std::copy(t[k][1].begin(), t[k][1].end(), vy6 + (size_t)(k + 1) * n);
    }
    eq.h = hs; eq.c = cs;
}
/* the energy budget of the field (h, c): out3 = F, D, and phi into phi
 * (padded, may be NULL) */
void yzo_energy(void *p, const double *h, const double *c, double *out2, double *phi) {
    auto *st = static_cast<yzo::Stepper *>(p);
    auto &eq = st->eq;
    const int n = eq.g.n;
    const std::vector<double> hs = eq.h, cs = eq.c;
    eq.h.assign(h, h + n); eq.c.assign(c, c + n);
    eq.fill_ghosts(eq.h); eq.fill_ghosts(eq.c);
    out2[0] = eq.free_energy();
    out2[1] = eq.flux_power();
    if (phi) {
        std::vector<double> ph;
        eq.potential(ph);
        std::copy(ph.begin(), ph.end(), phi);
    }
    eq.h = hs; eq.c = cs;
}
/* B's pieces (for the tests against the reference's tables): out2 =
 * Ev E(h, c) of the law alone (Evaporation.cpp:62-103 times Ev), EvapAt's
 * capped rate at step dt (HeightEquationBase.h:830-846) */
void yzo_evap(void *p, double h, double c, double dt, double *out2) {
    auto *st = static_cast<yzo::Stepper *>(p);
    const std::vector<double> H(1, h), Cv(1, c);
    out2[0] = st->eq.cs.evaporation ? st->eq.cs.Ev * st->evap_eval(h, c) : 0.0;
    out2[1] = st->evap_at(H, Cv, 0, dt);
}
/* the nozzle's supply of every cell of the padded grid at time t */
void yzo_source(void *p, double t, double *out) {
    auto *st = static_cast<yzo::Stepper *>(p);
    st->source_at(t);
    std::copy(st->src_cell.begin(), st->src_cell.end(), out);
}
/* Mobility (for the tests against the reference's tables): out =
 * Face(h1,h2), Face_d1(h1,h2), Face_d1(h2,h1), D(h1), dD(h1), DperH(h1) */
void yzo_mobility(double lambda, double h1, double h2, double *out) {
    yzo::Mobility m;
    m.set(lambda);
    out[0] = m.Face(h1, h2); out[1] = m.Face_d1(h1, h2); out[2] = m.Face_d1(h2, h1);
    out[3] = m.D(h1); out[4] = m.dD(h1); out[5] = m.DperH(h1);
}
/* Koren limiter: out = value, weights[4], derivative value, g[4] */
void yzo_limiter(const double *f, double sign, int vanleer, double *out) {
    yzo::Limiter k;
    k.vanleer = vanleer;
    out[0] = k.value(f, sign);
    k.weights(f, sign, out + 1);
    out[5] = k.derivative(f, sign, out + 6);
}
/* the backend contract on a caller's CRS */
int yzo_linsolve(int n, int nnz, const int *rowptr, const int *col, const double *val,
                 const double *rhs, double *x, int precond, int scaling, int maxit, double eps,
                 int *iter, double *resid) {
    yzo::Csr M;
    M.n = n; M.ptr.assign(rowptr, rowptr + n + 1); M.col.assign(col, col + nnz);
    M.val.assign(val, val + nnz); M.bc.assign(n, 0);
    /* the diagonal first, as CRS_Matrix stores it */
    for (int i = 0; i < n; ++i) {
        for (int k = M.ptr[i]; k < M.ptr[i + 1]; ++k)
            if (M.col[k] == i && k != M.ptr[i]) { std::swap(M.col[k], M.col[M.ptr[i]]); std::swap(M.val[k], M.val[M.ptr[i]]); }
    }
    std::vector<double> b(rhs, rhs + n), xx(x, x + n);
    int it = 0; double r = 0;
    const int rc = yzo::linsolve(M, b, xx, precond, scaling, maxit, eps, it, r);
    std::copy(xx.begin(), xx.end(), x);
    if (iter) *iter = it;
    if (resid) *resid = r;
    return rc;
}

}  /* extern "C" */

/* the ConcZ step, restated in its own file on the same CRS, BiCGSTAB and
 * multigrid */
#include "yz_cz_oracle.cpp"
