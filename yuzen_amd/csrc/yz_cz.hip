// yz_cz.hip -- the ConcZ step on the device: a 2-D film whose
// concentration is resolved through its thickness on a zeta grid that
// stretches with the film.
//
// One yz_cz_advance is one time step of the reference's ConcZ model:
//   CZ_HeightEquation_2D::Solve (CZ_HeightEquation_2D.cpp:1988-2020): the
//     semi-implicit film solved twice, with the coefficients of the start
//     of the step and again with those of the midpoint (SolveStage
//     :1949-1963, SetMatrixA :750-1081, the ghost rows :1083-1142);
//   AfterTimeStepTrial (:207-386): the conservative update through the
//     face fluxes (StoreFaceFluxParts / ConservativeUpdate :1161-1301),
//     the velocity through the film (CalVelocity :1381-1749, the in-plane
//     components), the concentration step on the moving zeta grid
//     (CZ_ConcEquation_2D.cpp: BuildVolumeFluxes :144-263, SetMatrixA
//     :290-570, SetBoundaryCondition :588-732, SolveStage :755-893,
//     ExplicitPart :895-913), the coupling passes with Aitken relaxation
//     when the Marangoni term is on, RedistributeExcess and
//     RedistributeDeficit (:388-570), and the ghosts of the end state
//     (RefreshGhosts, ApplyBoundaryRows :985-1005).
//
// Layout in HBM (FP64 throughout): the film grid is (nx+4) x (ny+4) rows
// with two ghost layers per side, row = Nx iy + ix; the concentration grid
// adds (nz+4) levels, row = Nx Ny kz + Nx iy + ix, so every level is a
// contiguous plane and a column is strided by Nx Ny. The linear systems are
// held as planes over the real cells only (ghost rows are exact copies in
// the increments solved for, and are folded into the stencil): 13 planes
// for the film (the structured multigrid solver of yz_solver.hip), 21 for
// the concentration, which is solved here by BiCGSTAB preconditioned with
// the exact solve of each column's zeta coupling (a banded LU per column:
// the diffusion through the film is the stiff direction).
//
// The kernels are HBM-bound stencil sweeps: one thread per cell, or per
// column where the reference's recursion runs up the column (the Simpson
// integral of the velocity, the zeta volume fluxes, the column solves and
// the redistribution), with the level planes read coalesced across the
// threads of a wave.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/yuzen_gpu.h"
#include "yz_limiter.h"
#include "yz_mlp.h"
#include "yz_schemes.h"
#include "yz_solver.h"
#include "yz_state.h"

namespace yz {
void set_error(const std::string &s);
}

using namespace yz;

namespace {

constexpr int TBK = 256;
constexpr int RB = 4096;         // reduction blocks
inline int nblk(long n) { return (int)std::min<long>((n + TBK - 1) / TBK, 65535L * 16); }

#define CZ_CHECK(x)                                                            \
    do {                                                                       \
        hipError_t e_ = (x);                                                   \
        if (e_ != hipSuccess) {                                                \
            yz::set_error(std::string("HIP: ") + hipGetErrorString(e_));       \
            return -1;                                                         \
        }                                                                      \
    } while (0)

struct Geo3 {
    int Nx, Ny, Nz, nx, ny, nz, n2, n3, nint;
    double dx, dy, dz, xlo, ylo;
    int wall[4];
    int perx, pery;
    int off[21];      // the slots' offsets among the real cells (no edge within reach)
};

__host__ __device__ inline int I2(const Geo3 &g, int i, int j) { return j * g.Nx + i; }
__host__ __device__ inline int I3(const Geo3 &g, int i, int j, int k) { return k * g.n2 + j * g.Nx + i; }
// the real index of a padded cell (ix, jy, kz all real)
__host__ __device__ inline int IR(const Geo3 &g, int i, int j, int k) {
    return ((k - 2) * g.ny + (j - 2)) * g.nx + (i - 2);
}
// the index a ghost index maps to along one axis (the film's ghost rows,
// CZ_HeightEquation_2D.cpp:1083-1142; the same rule on the 3-D grid)
__host__ __device__ inline int fold_pad(int i, int np, int per) {
    const int m = np - 4;
    if (i < 2) return per ? i + m : 3 - i;
    if (i > np - 3) return per ? i - m : 2 * (np - 3) + 1 - i;
    return i;
}

// the 21 slots of a concentration row: the zeta column (the band the
// column solves factor), the x line, the y line, then the x-zeta and
// y-zeta corners; slot 2 is the diagonal
#define YZ_S21X {0, 0, 0, 0, 0, -2, -1, 1, 2, 0, 0, 0, 0, -1, 1, -1, 1, 0, 0, 0, 0}
#define YZ_S21Y {0, 0, 0, 0, 0, 0, 0, 0, 0, -2, -1, 1, 2, 0, 0, 0, 0, -1, 1, -1, 1}
#define YZ_S21Z {-2, -1, 0, 1, 2, 0, 0, 0, 0, 0, 0, 0, 0, -1, -1, 1, 1, -1, -1, 1, 1}
__constant__ int c_s21x[21] = YZ_S21X;
__constant__ int c_s21y[21] = YZ_S21Y;
__constant__ int c_s21z[21] = YZ_S21Z;
constexpr int kDiag21 = 2;
__host__ __device__ constexpr int slot21(int ox, int oy, int oz) {
    return (ox == 0 && oy == 0) ? 2 + oz
         : (oy == 0 && oz == 0) ? (ox < 0 ? 7 + ox : 6 + ox)
         : (ox == 0 && oz == 0) ? (oy < 0 ? 11 + oy : 10 + oy)
         : (oy == 0) ? 13 + (ox > 0) + 2 * (oz > 0)
         : 17 + (oy > 0) + 2 * (oz > 0);
}

struct Args {
    Geo3 g;
    yz_cz_case cs;
    double dt, theta;
    double w1[4], w2[4];       // face weights of the scheme: c1, c2
    // a limited scheme (Koren, or van Leer: lim_vl): the weights of a face
    // frozen at the field fC, upwind form for the upwind cell, downwind form
    // for the other (AdvWeights, CZ_ConcEquation_2D.cpp:575-593)
    int lim, lim_vl;
    const double *S;
    double *hO, *hN, *fH, *tH, *fC, *tC;
    double *Dx, *Dy, *sgx, *sgy, *Mx, *My;
    double *qCapX, *qExpX, *qCapY, *qExpY, *qX, *qY, *colSrc, *dtdh;
    double *Qx, *Qy, *W, *zr;  // zr: the z ghost right-hand sides, 4 x n2
    double *hP, *hb, *hx;      // film system (solver planes), interior
    double *cA, *react;        // concentration A (21 planes) and the surface term, interior
    double *cP, *cb;           // row-scaled I + theta dt A, its right-hand side, interior
    int *flags;                // [0] right-hand side not zero, [1] below 0, [2] columns, [3] left
};

__device__ inline double visc(const yz_cz_case &c, double x) { return 1.0 * exp(c.visc_exp * (x - c.visc_c0)); }
__device__ inline double diffu(const yz_cz_case &c, double x) { return 1.0 * exp(c.diff_exp * (x - c.diff_c0)); }
__device__ inline double sigm(const yz_cz_case &c, double x) { return 1.0 + c.Ma * (x - c.sig_c0); }
__device__ inline double sgn(double v) { return v >= 0.0 ? 1.0 : -1.0; }
// fourth-order first derivatives: at the face between f0 and f1 (from
// f_-1 .. f2), and at the centre of a cell (from its +-1, +-2 neighbours)
__device__ inline double dface4(double fm1, double f0, double f1, double f2, double ds) {
    return (fm1 - 27.0 * f0 + 27.0 * f1 - f2) / (24.0 * ds);
}
__device__ inline double dcentre4(double fm2, double fm1, double f1, double f2, double ds) {
    return (fm2 - 8.0 * fm1 + 8.0 * f1 - f2) / (12.0 * ds);
}

// Evaluate (Evaporation.cpp:62-103, 148-220), uniform in space
__device__ inline double evap_eval(const yz_cz_case &c, double hh, double cc) {
    double f_conc;
    if (c.evap_func == YZ_EVAP_POWERN) {
        double v = 1.0 - cc - c.evap_concr;
        v = fmax(v, 0.0);
        f_conc = pow(v, c.evap_power);
    } else {
        const double x = (cc < 0.0) ? 0.0 : ((cc > 1.0) ? 1.0 : cc);
        const double chi = (x / c.molar_resin) / (x / c.molar_resin + (1.0 - x) / c.molar_solvent);
        f_conc = 1.0 - chi;
    }
    double f_mod = 1.0;
    if (c.precursor != 0.0) {
        const double v = (hh - c.precursor) / ((20.0 - 1.0) * c.precursor);
        f_mod = fmax(0.0, fmin(v, 1.0));
    }
    return c.evap_coef * f_conc * 1.0 * f_mod;
}
__device__ inline double solvent_limit(double evp, double h, double c, double dt) {
    if (!(evp > 0.0)) return evp;
    const double cc = (c < 0.0) ? 0.0 : ((c > 1.0) ? 1.0 : c);
    const double solvent = (h > 0.0) ? h * (1.0 - cc) : 0.0;
    if (evp * dt > solvent) evp = solvent / dt;
    return evp;
}
__device__ inline double column_mean(const Geo3 &g, const double *C, int ix, int jy) {
    double s = 0.0;
    for (int kz = 2; kz < g.Nz - 2; kz++) s += C[I3(g, ix, jy, kz)];
    return s / (double)(g.Nz - 4);
}
__device__ inline double evap_rate(const Args &a, const double *H, const double *C, int ix, int jy) {
    if (!a.cs.evaporation) return 0.0;
    const Geo3 &g = a.g;
    const double cc = 0.5 * (C[I3(g, ix, jy, g.Nz - 3)] + C[I3(g, ix, jy, g.Nz - 2)]);
    return evap_eval(a.cs, H[I2(g, ix, jy)], cc);
}
__device__ inline bool wall_face(const Args &a, int ix, int jy, int dir) {
    const Geo3 &g = a.g;
    if (dir == 0) return (ix == 2 && g.wall[0]) || (ix == g.Nx - 2 && g.wall[1]);
    return (jy == 2 && g.wall[2]) || (jy == g.Ny - 2 && g.wall[3]);
}

// ---------------------------------------------------------------------
// film: the face quantities of the stage (FaceMobility with its depth
// mean of 1/mu, the surface tension of the face, MarangoniFlux) and the
// face flux parts (StoreFaceFluxParts), one thread per cell for the
// x face on its low side and the y face below it
// ---------------------------------------------------------------------
__device__ void face_values(const Args &a, int ix, int jy, int dir, double &D, double &sg, double &M) {
    const Geo3 &g = a.g;
    const int ax = dir == 0, ay = dir == 1;
    const int lo2 = I2(g, ix - ax, jy - ay), hi2 = I2(g, ix, jy);
    // both depth means of 1/mu in one sweep (the sums in the same order)
    double n2 = 0, d2 = 0, n1 = 0, d1 = 0;
    for (int kz = 2; kz < g.Nz - 2; kz++) {
        const double c = 0.5 * (a.fC[kz * g.n2 + lo2] + a.fC[kz * g.n2 + hi2]);
        const double r = 1.0 - g.dz * (-1.5 + (double)kz);
        const double mu = visc(a.cs, c);
        n2 += (r * r) / mu; d2 += r * r;
        n1 += r / mu; d1 += r;
    }
    const double hl = a.fH[lo2], hh = a.fH[hi2];
    const double cl_t = a.fC[(g.Nz - 2) * g.n2 + lo2], cl_m = a.fC[(g.Nz - 3) * g.n2 + lo2];
    const double ch_t = a.fC[(g.Nz - 2) * g.n2 + hi2], ch_m = a.fC[(g.Nz - 3) * g.n2 + hi2];
    sg = sigm(a.cs, 0.25 * (cl_t + cl_m + ch_t + ch_m));
    const bool w = wall_face(a, ix, jy, dir);
    D = 0.0;
    M = 0.0;
    if (!w) {
        const double denom = hl + hh;
        if (!(denom <= 0.0)) {
            const double mob = 2.0 * (hl * hl) * (hh * hh) / denom;
            D = mob / 3.0 * (n2 / d2);
        }
        if (a.cs.marangoni && !(hl + hh <= 0.0)) {
            const double ds = dir == 0 ? g.dx : g.dy;
            const double cs1 = 0.5 * (cl_t + cl_m), cs2 = 0.5 * (ch_t + ch_m);
            const double dcds = (cs2 - cs1) / ds;
            M = 0.5 * (hl * hh) * a.cs.As * a.cs.Ma * dcds * (n1 / d1);
        }
    }
}

__global__ void k_faces(Args a) {
    const Geo3 &g = a.g;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.n2) return;
    const int ix = p % g.Nx, jy = p / g.Nx;
    const double rdx = 1.0 / g.dx, rdx2 = rdx * rdx, rdy = 1.0 / g.dy, rdy2 = rdy * rdy;
    auto Sf = [&](int i, int j) { return a.S[I2(g, i, j)]; };
    auto Hf = [&](int i, int j) { return a.fH[I2(g, i, j)]; };
    if (ix >= 2 && ix <= g.Nx - 2 && jy >= 2 && jy < g.Ny - 2) {
        double D, sg, M;
        face_values(a, ix, jy, 0, D, sg, M);
        a.Dx[p] = D; a.sgx[p] = sg; a.Mx[p] = M;
        double cap = 0.0, expl = 0.0;
        if (a.cs.capillary) {
            cap = D * sg;
            double d3s = (Sf(ix + 1, jy) - 3.0 * Sf(ix, jy) + 3.0 * Sf(ix - 1, jy) - Sf(ix - 2, jy)) * rdx * rdx2;
            d3s += ((Sf(ix, jy + 1) - 2.0 * Sf(ix, jy) + Sf(ix, jy - 1))
                    - (Sf(ix - 1, jy + 1) - 2.0 * Sf(ix - 1, jy) + Sf(ix - 1, jy - 1))) * rdx * rdy2;
            expl += cap * d3s;
        }
        if (a.cs.vertical_gravity)
            expl -= D * a.cs.Bo * ((Hf(ix, jy) + Sf(ix, jy)) - (Hf(ix - 1, jy) + Sf(ix - 1, jy))) * rdx;
        if (a.cs.inplane_gravity) expl += D * a.cs.Pgx;
        if (a.cs.centrifugal) expl += D * a.cs.Sp * ((g.xlo + ix * g.dx) - a.cs.rot_x);
        if (a.cs.marangoni) expl += M;
        a.qCapX[p] = cap;
        a.qExpX[p] = expl;
    }
    if (ix >= 2 && ix < g.Nx - 2 && jy >= 2 && jy <= g.Ny - 2) {
        double D, sg, M;
        face_values(a, ix, jy, 1, D, sg, M);
        a.Dy[p] = D; a.sgy[p] = sg; a.My[p] = M;
        double cap = 0.0, expl = 0.0;
        if (a.cs.capillary) {
            cap = D * sg;
            double d3s = (Sf(ix, jy + 1) - 3.0 * Sf(ix, jy) + 3.0 * Sf(ix, jy - 1) - Sf(ix, jy - 2)) * rdy * rdy2;
            d3s += ((Sf(ix - 1, jy) - 2.0 * Sf(ix, jy) + Sf(ix + 1, jy))
                    - (Sf(ix - 1, jy - 1) - 2.0 * Sf(ix, jy - 1) + Sf(ix + 1, jy - 1))) * rdy * rdx2;
            expl += cap * d3s;
        }
        if (a.cs.vertical_gravity)
            expl -= D * a.cs.Bo * ((Hf(ix, jy) + Sf(ix, jy)) - (Hf(ix, jy - 1) + Sf(ix, jy - 1))) * rdy;
        if (a.cs.inplane_gravity) expl += D * a.cs.Pgy;
        if (a.cs.centrifugal) expl += D * a.cs.Sp * ((g.ylo + jy * g.dy) - a.cs.rot_y);
        if (a.cs.marangoni) expl += M;
        a.qCapY[p] = cap;
        a.qExpY[p] = expl;
    }
}

// The film's row as the divergence of its four face fluxes
// (CZ_HeightEquation_2D.cpp:750-1081 gives the same operator entry by
// entry). The capillary flux through a face is
//   kappa_f (d^3/dn^3 + d/dn d^2/dt^2)(h + s),   kappa_f = D_f sigma_f
// (k_faces), with n the face normal and t the tangent: a fixed set of taps
// around the face, below. A row of A is then (F_high - F_low) / ds over
// both axes, so each face's taps enter the rows on either side of it with
// opposite signs and the operator is conservative by construction. The
// explicit fluxes of the stage (substrate relief, gravity, spin, Marangoni)
// are k_faces' qExp, differenced the same way. The row of
// I + theta dt A, its right-hand side h_old - (1 - theta) dt A h_old - dt B
// (SemiImplicit.cpp:158-190) and the scaling by the diagonal follow.
struct FaceTap {
    int n, t;      // cell offset along the normal (from the face's high cell) and the tangent
    double w;      // weight
    bool mixed;    // the d/dn d^2/dt^2 part (else the normal third difference)
};
constexpr FaceTap kFaceTaps[10] = {
    {+1, 0, 1.0, false}, {0, 0, -3.0, false}, {-1, 0, 3.0, false}, {-2, 0, -1.0, false},
    {0, +1, 1.0, true},  {0, 0, -2.0, true},  {0, -1, 1.0, true},
    {-1, +1, -1.0, true}, {-1, 0, 2.0, true}, {-1, -1, -1.0, true}};
// the solver's slot of a cell offset (oM mm om pm Mo mo oo po Po mp op pp oP)
__host__ __device__ constexpr int slot13(int ox, int oy) {
    return oy == -2 ? 0 : oy == 2 ? 12 : oy == -1 ? 2 + ox : oy == 1 ? 10 + ox : 6 + ox;
}
constexpr int kSlotX[13] = {0, -1, 0, 1, -2, -1, 0, 1, 2, -1, 0, 1, 0};
constexpr int kSlotY[13] = {-2, -1, -1, -1, 0, 0, 0, 0, 0, 1, 1, 1, 2};

__global__ void k_hrows(Args a) {
    const Geo3 &g = a.g;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = g.nx * g.ny;
    if (q >= N) return;
    const int ix = q % g.nx + 2, jy = q / g.nx + 2;
    const int p = I2(g, ix, jy);
    const yz_cz_case &cs = a.cs;
    const double rdx = 1.0 / g.dx, rdy = 1.0 / g.dy;
    // per axis: the normal part's scale 1/ds^4 and the mixed part's 1/(dx^2 dy^2)
    const double rnx = (rdx * rdx) * (rdx * rdx), rny = (rdy * rdy) * (rdy * rdy);
    const double rmix = (rdx * rdx) * (rdy * rdy);
    double A[13];
#pragma unroll
    for (int k = 0; k < 13; ++k) A[k] = 0.0;
    if (cs.capillary) {
        // kappa of the low and high face of each axis
        const double kx[2] = {a.Dx[p] * a.sgx[p], a.Dx[p + 1] * a.sgx[p + 1]};
        const double ky[2] = {a.Dy[p] * a.sgy[p], a.Dy[p + g.Nx] * a.sgy[p + g.Nx]};
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const double sg = side ? 1.0 : -1.0;
#pragma unroll
            for (int t = 0; t < 10; ++t) {
                const FaceTap &f = kFaceTaps[t];
                const double wx = sg * f.w * kx[side] * (f.mixed ? rmix : rnx);
                const double wy = sg * f.w * ky[side] * (f.mixed ? rmix : rny);
                A[slot13(side + f.n, f.t)] += wx;
                A[slot13(f.t, side + f.n)] += wy;
            }
        }
    }
    // the explicit fluxes' divergence, evaporation with its solvent limits
    double Evap = cs.Ev * evap_rate(a, a.fH, a.fC, ix, jy);
    if (Evap > 0.0) {
        Evap = solvent_limit(Evap, a.fH[p], column_mean(g, a.fC, ix, jy), a.dt);
        Evap = solvent_limit(Evap, a.tH[p], column_mean(g, a.tC, ix, jy), a.dt);
    }
    const double divq = (a.qExpX[p + 1] - a.qExpX[p]) * rdx + (a.qExpY[p + g.Nx] - a.qExpY[p]) * rdy;
    a.colSrc[p] = -Evap;
    const double B = divq + Evap;      // dh/dt = -A h - B
    double Ax = 0.0;
#pragma unroll
    for (int k = 0; k < 13; ++k) Ax += A[k] * a.hO[I2(g, ix + kSlotX[k], jy + kSlotY[k])];
    const double th = a.theta, dt = a.dt;
    const double rhs = a.hO[p] - dt * (1.0 - th) * Ax - dt * B;
    const double d = 1.0 + th * dt * A[6];
    const bool scale = !(fabs(d) < 1.0e-30);
    const double inv = scale ? 1.0 / d : 1.0;
#pragma unroll
    for (int k = 0; k < 13; ++k) {
        const double v = (k == 6) ? d : th * dt * A[k];
        a.hP[(size_t)k * N + q] = scale ? v * inv : v;
    }
    a.hb[q] = scale ? rhs * inv : rhs;
    a.hx[q] = a.hN[p];
}

// the film's ghosts from the real rows they copy (CopyGhosts: a corner
// reads an edge, so both axes fold)
__global__ void k_hghosts(Geo3 g, double *h) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.n2) return;
    const int ix = p % g.Nx, jy = p / g.Nx;
    const bool gx = ix < 2 || ix > g.Nx - 3, gy = jy < 2 || jy > g.Ny - 3;
    if (!gx && !gy) return;
    const int sx = fold_pad(ix, g.Nx, g.perx), sy = fold_pad(jy, g.Ny, g.pery);
    h[p] = h[I2(g, sx, sy)];
}
__global__ void k_hscatter(Geo3 g, const double *x, double *h) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nx * g.ny) return;
    h[I2(g, q % g.nx + 2, q / g.nx + 2)] = x[q];
}

// ConservativeUpdate (:1232-1301): the face fluxes with the
// Crank-Nicolson h of the solve, then the end of the step through them
__device__ inline double hcv(const Args &a, int i, int j) {
    const int p = I2(a.g, i, j);
    const double u = a.hN[p] * a.theta;
    const double v = a.hO[p] * (1.0 - a.theta);
    return u + v;
}
__global__ void k_hflux(Args a) {
    const Geo3 &g = a.g;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.n2) return;
    const int ix = p % g.Nx, jy = p / g.Nx;
    const double rdx = 1.0 / g.dx, rdx2 = rdx * rdx, rdy = 1.0 / g.dy, rdy2 = rdy * rdy;
    if (ix >= 2 && ix <= g.Nx - 2 && jy >= 2 && jy < g.Ny - 2) {
        double d3h = (hcv(a, ix + 1, jy) - 3.0 * hcv(a, ix, jy) + 3.0 * hcv(a, ix - 1, jy) - hcv(a, ix - 2, jy)) * rdx * rdx2;
        d3h += ((hcv(a, ix, jy + 1) - 2.0 * hcv(a, ix, jy) + hcv(a, ix, jy - 1))
                - (hcv(a, ix - 1, jy + 1) - 2.0 * hcv(a, ix - 1, jy) + hcv(a, ix - 1, jy - 1))) * rdx * rdy2;
        a.qX[p] = a.qCapX[p] * d3h + a.qExpX[p];
    }
    if (ix >= 2 && ix < g.Nx - 2 && jy >= 2 && jy <= g.Ny - 2) {
        double d3h = (hcv(a, ix, jy + 1) - 3.0 * hcv(a, ix, jy) + 3.0 * hcv(a, ix, jy - 1) - hcv(a, ix, jy - 2)) * rdy * rdy2;
        d3h += ((hcv(a, ix - 1, jy) - 2.0 * hcv(a, ix, jy) + hcv(a, ix + 1, jy))
                - (hcv(a, ix - 1, jy - 1) - 2.0 * hcv(a, ix, jy - 1) + hcv(a, ix + 1, jy - 1))) * rdy * rdx2;
        a.qY[p] = a.qCapY[p] * d3h + a.qExpY[p];
    }
}
__global__ void k_hupdate(Args a) {
    const Geo3 &g = a.g;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nx * g.ny) return;
    const int ix = q % g.nx + 2, jy = q / g.nx + 2;
    const int p = I2(g, ix, jy);
    double div = (a.qX[p + 1] - a.qX[p]) * (1.0 / g.dx);
    div += (a.qY[p + g.Nx] - a.qY[p]) * (1.0 / g.dy);
    a.hN[p] = a.hO[p] + a.dt * (a.colSrc[p] - div);
}
// the midpoint film and dh/dt of the step (CalDtDh)
__global__ void k_hmid(Args a, double *mid, int with_rate) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.g.n2) return;
    mid[p] = (a.hO[p] + a.hN[p]) * 0.5;
    if (with_rate) {
        const double d = a.hN[p] - a.hO[p];
        a.dtdh[p] = d / a.dt;
    }
}

// ---------------------------------------------------------------------
// The velocity through the film at the faces and the layer volume fluxes
// (CalVelocity, CZ_HeightEquation_2D.cpp:1381-1530; BuildVolumeFluxes,
// CZ_ConcEquation_2D.cpp:144-263). In the lubrication limit the velocity
// along a face normal at height z above the substrate is
//   u(z) = G int_s^z (s + h - z') / mu(z') dz' + tau int_s^z dz' / mu(z'),
// G the driving gradient of the face (capillary pressure, gravity along
// the slope, in-plane gravity, spin) and tau the Marangoni shear at the
// surface; the integrals run by Simpson's rule between the levels' centres
// (the first from the substrate, half a level long), mu at the centres and
// at the interfaces from the level values of c averaged over the face's two
// cells (at the substrate, over the first level and its ghost). A layer's
// volume flux is the face film height times the mean of u over the layer
// (the quadratic through the neighbouring centres; u = 0 at the
// substrate), shifted evenly through the depth so that the layers add up to
// the film's own flux through the face.
//
// One block works on a tile of TW consecutive faces of one row and one
// direction; the layout of its LDS is [node][face]:
//   phase 1  all threads: 1/mu at the 2 nz + 3 nodes of every face of the
//            tile (the exponentials spread over the whole block);
//   phase 2  one thread per face: the running Simpson sums up the column,
//            u at the nz + 1 centres (the last one the ghost above the
//            surface);
//   phase 3  all threads: the layer means, into the LDS the nodes left;
//   phase 4  one thread per face: the layers' total, in level order;
//   phase 5  all threads: Q of every layer with the shift, coalesced along
//            the row.
// ---------------------------------------------------------------------
constexpr int kVelThreads = 256;
// the widest tile (64, 32, 16 or 8 faces) whose LDS fits in 64 KiB
inline int vel_tile(int nz) {
    for (int tw = 64; tw > 8; tw >>= 1)
        if ((size_t)(3 * nz + 4) * tw * sizeof(double) <= 65536) return tw;
    return 8;
}
inline size_t vel_lds(int nz, int tw) { return (size_t)(3 * nz + 4) * tw * sizeof(double) + 8 * tw * sizeof(double); }

__global__ void __launch_bounds__(kVelThreads) k_face_velocity(Args a, int TW) {
    extern __shared__ double lds[];
    const Geo3 &g = a.g;
    const yz_cz_case &cs = a.cs;
    const int nz = g.nz, nodes = 2 * nz + 3;
    const int dir = blockIdx.z, ax = dir == 0, ay = dir == 1;
    double *node = lds;                               // [nodes][TW]: 1/mu; later the layer fluxes
    double *uc = lds + (size_t)nodes * TW;            // [nz + 1][TW]: u at the centres
    double *fp = uc + (size_t)(nz + 1) * TW;          // [8][TW]: per-face scalars
    double *f_drive = fp, *f_shear = fp + TW, *f_top = fp + 2 * TW, *f_sub = fp + 3 * TW;
    double *f_dzz = fp + 4 * TW, *f_hf = fp + 5 * TW, *f_shift = fp + 6 * TW, *f_ok = fp + 7 * TW;
    const int jy = 2 + blockIdx.y;
    const int ix0 = 2 + blockIdx.x * TW;
    // the faces of this direction: x faces ix in [2, Nx-2] on rows [2, Ny-3],
    // y faces ix in [2, Nx-3] on rows [2, Ny-2]
    const int ixmax = dir == 0 ? g.Nx - 2 : g.Nx - 3, jymax = dir == 0 ? g.Ny - 3 : g.Ny - 2;
    if (jy > jymax) return;
    const int t = threadIdx.x;
    auto cell = [&](int f, int k) { return I2(g, ix0 + f + k * ax, jy + k * ay); };   // k along the normal
    auto cval = [&](int f, int k, int kz) { return a.fC[kz * g.n2 + cell(f, k)]; };

    // phase 0: the face's own scalars
    if (t < TW) {
        const int ix = ix0 + t;
        const bool ok = ix <= ixmax;
        f_ok[t] = ok ? 1.0 : 0.0;
        if (ok) {
            const double ds = dir == 0 ? g.dx : g.dy, dn = dir == 0 ? g.dy : g.dx;
            const int tx = ay, ty = ax;                   // the tangent
            auto hs = [&](int k, int m) {
                const int p = I2(g, ix + k * ax + m * tx, jy + k * ay + m * ty);
                return a.fH[p] + a.S[p];
            };
            const double hlo = a.fH[cell(t, -1)], hhi = a.fH[cell(t, 0)];
            const double slo = a.S[cell(t, -1)], shi = a.S[cell(t, 0)];
            const double hface = 0.5 * (hlo + hhi), sface = 0.5 * (slo + shi);
            const int top = g.Nz - 2, below = g.Nz - 3;
            double drive = 0.0, shear = 0.0;
            if (cs.capillary) {
                const double sig = sigm(cs, 0.25 * (cval(t, 0, top) + cval(t, 0, below) + cval(t, -1, top) + cval(t, -1, below)));
                const double d3n = (hs(1, 0) - 3.0 * hs(0, 0) + 3.0 * hs(-1, 0) - hs(-2, 0)) / (ds * ds * ds);
                const double dnt2 = ((hs(0, 1) - 2.0 * hs(0, 0) + hs(0, -1)) - (hs(-1, 1) - 2.0 * hs(-1, 0) + hs(-1, -1))) /
                                    (ds * dn * dn);
                drive += sig * (d3n + dnt2);
            }
            if (cs.vertical_gravity) drive -= cs.Bo * dface4(hs(-2, 0), hs(-1, 0), hs(0, 0), hs(1, 0), ds);
            if (cs.inplane_gravity) drive += dir == 0 ? cs.Pgx : cs.Pgy;
            if (cs.centrifugal)
                drive += cs.Sp * (dir == 0 ? (g.xlo + ix * g.dx) - cs.rot_x : (g.ylo + jy * g.dy) - cs.rot_y);
            if (cs.marangoni) {
                const double gt = dface4(cval(t, -2, top), cval(t, -1, top), cval(t, 0, top), cval(t, 1, top), ds);
                const double gb = dface4(cval(t, -2, below), cval(t, -1, below), cval(t, 0, below), cval(t, 1, below), ds);
                shear = cs.As * cs.Ma * 0.5 * (gt + gb);
            }
            const bool wall = wall_face(a, ix, jy, dir);
            f_drive[t] = wall ? 0.0 : drive;
            f_shear[t] = wall ? 0.0 : shear;
            f_top[t] = hface + sface;
            f_sub[t] = sface;
            f_dzz[t] = g.dz * hface;
            f_hf[t] = hface;
            f_shift[t] = dir == 0 ? a.qX[cell(t, 0)] : a.qY[cell(t, 0)];
        }
    }
    // phase 1: 1/mu at the nodes: 0 the substrate, 2j + 1 the lower end's
    // midpoint of interval j, 2j + 2 the centre of level j + 2
    for (int e = t; e < nodes * TW; e += kVelThreads) {
        const int m = e / TW, f = e % TW;
        if (ix0 + f > ixmax) continue;
        double c;
        if (m == 0) {
            c = 0.25 * (cval(f, 0, 2) + cval(f, 0, 1) + cval(f, -1, 2) + cval(f, -1, 1));
        } else if (m == 1) {
            c = 0.375 * (cval(f, 0, 2) + cval(f, -1, 2)) + 0.125 * (cval(f, 0, 1) + cval(f, -1, 1));
        } else if (m % 2 == 0) {
            const int kz = m / 2 + 1;
            c = 0.5 * (cval(f, 0, kz) + cval(f, -1, kz));
        } else {
            const int kz = (m + 1) / 2 + 1;
            c = 0.25 * (cval(f, 0, kz) + cval(f, 0, kz - 1) + cval(f, -1, kz) + cval(f, -1, kz - 1));
        }
        node[e] = 1.0 / visc(cs, c);
    }
    __syncthreads();
    // phase 2: u at the centres, from the running Simpson sums
    if (t < TW && f_ok[t] != 0.0) {
        const double dzz = f_dzz[t], zs = f_sub[t], ztop = f_top[t], G = f_drive[t], tau = f_shear[t];
        double i1 = 0.0, iz = 0.0;                 // int dz/mu, int z dz/mu
        double zlo = zs, wlo = node[t];
        for (int j = 0; j <= nz; ++j) {
            const double zmid = (j == 0) ? zs + 0.25 * dzz : zs + dzz * (double)j;
            const double zc = zs + dzz * ((double)j + 0.5);
            const double len = (j == 0) ? 0.5 * dzz : dzz;
            const double wm = node[(2 * j + 1) * TW + t], wc = node[(2 * j + 2) * TW + t];
            i1 += (wlo + 4.0 * wm + wc) * (len / 6.0);
            iz += (zlo * wlo + 4.0 * zmid * wm + zc * wc) * (len / 6.0);
            uc[j * TW + t] = G * (ztop * i1 - iz) + tau * i1;
            zlo = zc;
            wlo = wc;
        }
    }
    __syncthreads();
    // phase 3: layer fluxes h_f <u>_layer into the node storage
    double *lay = node;                                // [nz][TW]
    for (int e = t; e < nz * TW; e += kVelThreads) {
        const int j = e / TW, f = e % TW;
        if (f_ok[f] == 0.0) continue;
        const double u0 = uc[j * TW + f], u1 = uc[(j + 1) * TW + f];
        const double mean = (j == 0) ? u0 + (u1 - 3.0 * u0) / 18.0 : u0 + (u1 - 2.0 * u0 + uc[(j - 1) * TW + f]) / 24.0;
        lay[e] = f_hf[f] * mean;
    }
    __syncthreads();
    // phase 4: what the layers carry against the film's own flux
    if (t < TW && f_ok[t] != 0.0) {
        double tot = 0.0;
        for (int j = 0; j < nz; ++j) tot += lay[j * TW + t] * g.dz;
        f_shift[t] -= tot;
    }
    __syncthreads();
    // phase 5: the layers' fluxes, shifted evenly
    double *Q = dir == 0 ? a.Qx : a.Qy;
    for (int e = t; e < nz * TW; e += kVelThreads) {
        const int j = e / TW, f = e % TW;
        if (f_ok[f] == 0.0) continue;
        Q[(j + 2) * g.n2 + cell(f, 0)] = lay[e] + f_shift[f];
    }
}
// the zeta volume fluxes of a column from the substrate up (the discrete
// geometric conservation law), the free surface from the film's sources
__global__ void k_w(Args a) {
    const Geo3 &g = a.g;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nx * g.ny) return;
    const int ix = q % g.nx + 2, jy = q / g.nx + 2;
    const int p = I2(g, ix, jy);
    const double ht = a.dtdh[p];
    double w = 0.0;
    a.W[2 * g.n2 + p] = 0.0;
    for (int kz = 2; kz < g.Nz - 2; kz++) {
        const int r = kz * g.n2 + p;
        double div = ht + (a.Qx[r + 1] - a.Qx[r]) / g.dx;
        div += (a.Qy[r + g.Nx] - a.Qy[r]) / g.dy;
        w -= g.dz * div;
        a.W[r + g.n2] = w;
    }
    a.W[(g.Nz - 2) * g.n2 + p] = -(a.colSrc[p] + 0.0);
}

// ---------------------------------------------------------------------
// the concentration's ghost rows: the right-hand sides of the z ghosts of
// every real column (SetBoundaryCondition), and the ghosts of a vector
// from the rows (ApplyBoundaryRows: the z ghosts from the real cells,
// then the y ghost rows, then the x ghost columns, which is the fixed
// point the reference's sweeps reach)
// ---------------------------------------------------------------------
__global__ void k_cbc(Args a, int atEnd) {
    const Geo3 &g = a.g;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nx * g.ny) return;
    const int ix = q % g.nx + 2, jy = q / g.nx + 2;
    const int p = I2(g, ix, jy);
    const yz_cz_case &cs = a.cs;
    auto C = [&](int i, int j, int k) { return a.fC[I3(g, i, j, k)]; };
    auto Hh = [&](int i, int j) { return a.fH[I2(g, i, j)]; };
    auto Ss = [&](int i, int j) { return a.S[I2(g, i, j)]; };
    const double hh = Hh(ix, jy);
    const double hx0 = (Hh(ix - 2, jy) - 8.0 * Hh(ix - 1, jy) + 8.0 * Hh(ix + 1, jy) - Hh(ix + 2, jy)) / (12.0 * g.dx);
    const double sx0 = (Ss(ix - 2, jy) - 8.0 * Ss(ix - 1, jy) + 8.0 * Ss(ix + 1, jy) - Ss(ix + 2, jy)) / (12.0 * g.dx);
    const double hy0 = (Hh(ix, jy - 2) - 8.0 * Hh(ix, jy - 1) + 8.0 * Hh(ix, jy + 1) - Hh(ix, jy + 2)) / (12.0 * g.dy);
    const double sy0 = (Ss(ix, jy - 2) - 8.0 * Ss(ix, jy - 1) + 8.0 * Ss(ix, jy + 1) - Ss(ix, jy + 2)) / (12.0 * g.dy);
    {
        const int k = 2;
        const double cb = 0.5 * (C(ix, jy, k - 1) + C(ix, jy, k));
        const double D = cs.Df * diffu(cs, cb);
        const double cx = (C(ix + 1, jy, k) - C(ix - 1, jy, k)) / (2.0 * g.dx);
        const double cy = (C(ix, jy + 1, k) - C(ix, jy - 1, k)) / (2.0 * g.dy);
        const double hA = sx0, hAy = sy0;
        const double cz = hh * (0.0 / D + hA * cx + hAy * cy) / (cs.As + hA * hA + hAy * hAy);
        a.zr[0 * g.n2 + p] = -3.0 * g.dz * cz;
        a.zr[1 * g.n2 + p] = -g.dz * cz;
    }
    {
        const int k = g.Nz - 3;
        const double csf = 0.5 * (C(ix, jy, k) + C(ix, jy, k + 1));
        const double D = cs.Df * diffu(cs, csf);
        const double cx = (C(ix + 1, jy, k) - C(ix - 1, jy, k)) / (2.0 * g.dx);
        const double cy = (C(ix, jy + 1, k) - C(ix, jy - 1, k)) / (2.0 * g.dy);
        const double hA = hx0 + sx0, hAy = hy0 + sy0;
        double WT = a.W[(g.Nz - 2) * g.n2 + p];
        if (atEnd) {
            WT = cs.Ev * evap_rate(a, a.fH, a.fC, ix, jy);
            double cm = 0.0;
            for (int kk = 2; kk < g.Nz - 2; kk++) cm += C(ix, jy, kk);
            cm /= (double)(g.Nz - 4);
            WT = solvent_limit(WT, hh, cm, a.dt);
        }
        const double G = WT * csf;
        const double cz = hh * (G / D + hA * cx + hAy * cy) / (cs.As + hA * hA + hAy * hAy);
        const double grad = g.dz * cz;
        a.zr[2 * g.n2 + p] = grad;
        a.zr[3 * g.n2 + p] = 3.0 * grad;
    }
}
// rows c(0) = c(3) + r0, c(1) = c(2) + r1, c(Nz-2) = c(Nz-3) + r2,
// c(Nz-1) = c(Nz-4) + r3 of every real column
__global__ void k_cghost_z(Geo3 g, const double *zr, double *x) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nx * g.ny) return;
    const int p = I2(g, q % g.nx + 2, q / g.nx + 2);
    double s;
    s = zr[0 * g.n2 + p]; s -= (-1.0) * x[3 * g.n2 + p]; x[0 * g.n2 + p] = s / 1.0;
    s = zr[1 * g.n2 + p]; s -= (-1.0) * x[2 * g.n2 + p]; x[1 * g.n2 + p] = s / 1.0;
    s = zr[2 * g.n2 + p]; s -= (-1.0) * x[(g.Nz - 3) * g.n2 + p]; x[(g.Nz - 2) * g.n2 + p] = s / 1.0;
    s = zr[3 * g.n2 + p]; s -= (-1.0) * x[(g.Nz - 4) * g.n2 + p]; x[(g.Nz - 1) * g.n2 + p] = s / 1.0;
}
// the y ghost rows of the real x columns, every level
__global__ void k_cghost_y(Geo3 g, double *x) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = g.nx * 4;
    if (t >= per * g.Nz) return;
    const int kz = t / per, r = t % per;
    const int ix = r % g.nx + 2, gj = r / g.nx;
    const int jy = gj < 2 ? gj : g.Ny - 4 + gj;
    const int sy = fold_pad(jy, g.Ny, g.pery);
    x[I3(g, ix, jy, kz)] = 0.0 + x[I3(g, ix, sy, kz)];
}
// the x ghost columns, every row and level
__global__ void k_cghost_x(Geo3 g, double *x) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = 4 * g.Ny;
    if (t >= per * g.Nz) return;
    const int kz = t / per, r = t % per;
    const int jy = r / 4, gi = r % 4;
    const int ix = gi < 2 ? gi : g.Nx - 4 + gi;
    const int sx = fold_pad(ix, g.Nx, g.perx);
    x[I3(g, ix, jy, kz)] = 0.0 + x[I3(g, sx, jy, kz)];
}

// ---------------------------------------------------------------------
// The concentration row of one real cell as a balance of face fluxes.
// Through each face runs a flux P_f per unit cell volume, along the axis:
//   advection   Q_f c_f / ds, c_f the scheme's face value over the four
//               cells along the axis (weights c1 + sign(Q_f) c2);
//   diffusion   -G_f, the diffusive flux in the stretched coordinates:
//     lateral face  G = (h_f D_f / ds) [ (c_hi - c_lo) / ds
//                        - (zeta h_s + s_s)_f / h_f (c_{k+1} - c_{k-1},
//                          summed over both cells) / (4 dzeta) ],
//     zeta face     G = (D_f / dzeta) [ (As + |grad(zeta h + s)|^2) / h
//                        (c_hi - c_lo) / dzeta - (zeta h_x + s_x)
//                        (c_{i+1} - c_{i-1}, summed over both levels)
//                        / (4 dx) - (the same along y) ],
//     none through a wall;
// the cell's inflow is P_low - P_high on every axis. The substrate carries
// nothing; the free surface carries the evaporation flux W_T c of the cell
// itself, a reaction term. A row of A is minus the inflow over h; its
// diagonal is minus the sum of its off-diagonal entries plus the reaction
// term, so that a uniform c is kept exactly. This is the operator of
// CZ_ConcEquation_2D::SetMatrixA (CZ_ConcEquation_2D.cpp:301-573). The
// stage's right-hand side is -dt (A c_old) in the balance form
// react c_r + sum_k A_k (c_k - c_r) (ExplicitPart, :895-913).
// ---------------------------------------------------------------------
// react c_r + sum over the off-diagonal slots of A_k (c_k - c_r)
__device__ inline double balance_form(const Geo3 &g, const double (&A)[21], double react, const double *c, int ix,
                                      int jy, int kz) {
    const double cr = c[I3(g, ix, jy, kz)];
    double s = react * cr;
#pragma unroll
    for (int k = 0; k < 21; ++k)
        if (k != kDiag21) s += A[k] * (c[I3(g, ix + c_s21x[k], jy + c_s21y[k], kz + c_s21z[k])] - cr);
    return s;
}

__global__ void k_crows(Args a, const double *cold) {
    const Geo3 &g = a.g;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nint) return;
    const int ix = q % g.nx + 2, jy = (q / g.nx) % g.ny + 2, kz = q / (g.nx * g.ny) + 2;
    const yz_cz_case &cs = a.cs;
    const double dx = g.dx, dy = g.dy, dz = g.dz;
    auto cc = [&](int i, int j, int k) { return a.fC[I3(g, i, j, k)]; };
    auto hh = [&](int i, int j) { return a.fH[I2(g, i, j)]; };
    auto ss = [&](int i, int j) { return a.S[I2(g, i, j)]; };
    const int r3 = I3(g, ix, jy, kz);
    const double zeta = dz * ((double)kz - 1.5);
    const bool bottom = kz == 2, top = kz == g.Nz - 3;
    double v[21];
#pragma unroll
    for (int k = 0; k < 21; ++k) v[k] = 0.0;
    auto add = [&](int ox, int oy, int oz, double w) { v[slot21(ox, oy, oz)] += w; };

    // lateral faces; side 0 is the low face (inflow +P), side 1 the high (-P)
#pragma unroll
    for (int axis = 0; axis < 2; ++axis) {
        const int ex = axis == 0, ey = axis == 1;
        const double ds = axis == 0 ? dx : dy;
        const double *Qf = axis == 0 ? a.Qx : a.Qy;
        const int fstride = axis == 0 ? 1 : g.Nx;
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const double sg = side ? -1.0 : 1.0;
            const int hi = side, lo = side - 1;       // the face's cells, offsets along the axis
            const double Q = Qf[r3 + side * fstride];
            const double sq = sgn(Q);
            double w[4];
            if (a.lim) {
                // the cell is R of its low face, L of its high one
                double f[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) f[t] = cc(ix + (lo - 1 + t) * ex, jy + (lo - 1 + t) * ey, kz);
                limited_weights(f, Q, side ? 1 : 2, a.lim_vl, w);
            } else {
#pragma unroll
                for (int t = 0; t < 4; ++t) w[t] = a.w1[t] + sq * a.w2[t];
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) add((lo - 1 + t) * ex, (lo - 1 + t) * ey, 0, sg * Q * w[t] / ds);
            if (wall_face(a, ix + hi * ex, jy + hi * ey, axis)) continue;
            const int li = ix + lo * ex, lj = jy + lo * ey, hi_i = ix + hi * ex, hi_j = jy + hi * ey;
            const double hf = 0.5 * (hh(li, lj) + hh(hi_i, hi_j));
            const double Df = cs.Df * diffu(cs, 0.5 * (cc(li, lj, kz) + cc(hi_i, hi_j, kz)));
            const double hs = dface4(hh(li - ex, lj - ey), hh(li, lj), hh(hi_i, hi_j), hh(hi_i + ex, hi_j + ey), ds);
            const double sd = dface4(ss(li - ex, lj - ey), ss(li, lj), ss(hi_i, hi_j), ss(hi_i + ex, hi_j + ey), ds);
            const double ka = hf * Df / ds;
            const double km = ka * ((hs * zeta + sd) / hf) / (4.0 * dz);
            add(hi * ex, hi * ey, 0, -sg * ka / ds);
            add(lo * ex, lo * ey, 0, +sg * ka / ds);
            add(lo * ex, lo * ey, +1, +sg * km);
            add(hi * ex, hi * ey, +1, +sg * km);
            add(lo * ex, lo * ey, -1, -sg * km);
            add(hi * ex, hi * ey, -1, -sg * km);
        }
    }
    // zeta faces
    double rct = 0.0;
    {
        const double hc = hh(ix, jy);
        const double hx0 = dcentre4(hh(ix - 2, jy), hh(ix - 1, jy), hh(ix + 1, jy), hh(ix + 2, jy), dx);
        const double sx0 = dcentre4(ss(ix - 2, jy), ss(ix - 1, jy), ss(ix + 1, jy), ss(ix + 2, jy), dx);
        const double hy0 = dcentre4(hh(ix, jy - 2), hh(ix, jy - 1), hh(ix, jy + 1), hh(ix, jy + 2), dy);
        const double sy0 = dcentre4(ss(ix, jy - 2), ss(ix, jy - 1), ss(ix, jy + 1), ss(ix, jy + 2), dy);
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const double sg = side ? -1.0 : 1.0;
            const int hi = side, lo = side - 1;
            if (side == 0 && bottom) continue;
            const double W = a.W[r3 + side * g.n2];
            if (side == 1 && top) { rct += W / dz; continue; }
            const double sw = sgn(W);
            double w[4];
            if (a.lim) {
                double f[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) f[t] = cc(ix, jy, kz + lo - 1 + t);
                limited_weights(f, W, side ? 1 : 2, a.lim_vl, w);
            } else {
#pragma unroll
                for (int t = 0; t < 4; ++t) w[t] = a.w1[t] + sw * a.w2[t];
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) add(0, 0, lo - 1 + t, sg * W * w[t] / dz);
            const double zf = zeta + (side ? 0.5 : -0.5) * dz;
            const double Dz = cs.Df * diffu(cs, 0.5 * (cc(ix, jy, kz + lo) + cc(ix, jy, kz + hi)));
            const double gx = hx0 * zf + sx0, gy = hy0 * zf + sy0;
            const double K = Dz * ((cs.As + gx * gx + gy * gy) / hc) / (dz * dz);
            const double kxw = Dz * gx / (4.0 * dx) / dz, kyw = Dz * gy / (4.0 * dy) / dz;
            add(0, 0, hi, -sg * K);
            add(0, 0, lo, +sg * K);
            add(+1, 0, lo, +sg * kxw); add(+1, 0, hi, +sg * kxw);
            add(-1, 0, lo, -sg * kxw); add(-1, 0, hi, -sg * kxw);
            add(0, +1, lo, +sg * kyw); add(0, +1, hi, +sg * kyw);
            add(0, -1, lo, -sg * kyw); add(0, -1, hi, -sg * kyw);
        }
    }
    const double rh = 1.0 / hh(ix, jy);
    double A[21];
    double offsum = 0.0;
#pragma unroll
    for (int k = 0; k < 21; ++k) {
        if (k == kDiag21) continue;
        A[k] = -v[k] * rh;
        offsum += A[k];
    }
    const double react = -rct * rh;
    A[kDiag21] = -offsum + react;
    const int n = g.nint;
#pragma unroll
    for (int k = 0; k < 21; ++k) a.cA[(size_t)k * n + q] = A[k];
    a.react[q] = react;
    const double rhs = -a.dt * balance_form(g, A, react, cold, ix, jy, kz);
    if (rhs != 0.0) atomicOr(&a.flags[0], 1);
    const double th = a.theta, dt = a.dt;
    const double d = 1.0 + th * dt * A[kDiag21];
    const bool sc = !(fabs(d) < 1.0e-30);
    const double inv = sc ? 1.0 / d : 1.0;
#pragma unroll
    for (int k = 0; k < 21; ++k) {
        const double val = (k == kDiag21) ? d : th * dt * A[k];
        a.cP[(size_t)k * n + q] = sc ? val * inv : val;
    }
    a.cb[q] = sc ? rhs * inv : rhs;
}

// ---------------------------------------------------------------------
// the concentration system on the real cells: 21 planes, the neighbours
// folded at the edges (x and y: wrapped or mirrored; zeta: mirrored, the
// increments of the ghost rows being copies)
// ---------------------------------------------------------------------
__device__ inline int fold_in(int j, int n, int per) {
    if (j >= 0 && j < n) return j;
    if (per) { j %= n; if (j < 0) j += n; return j; }
    if (j < 0) return -1 - j;
    return 2 * n - 1 - j;
}
__device__ inline int col3(const Geo3 &g, int i, int j, int k, int s) {
    const int ii = fold_in(i + c_s21x[s], g.nx, g.perx);
    const int jj = fold_in(j + c_s21y[s], g.ny, g.pery);
    const int kk = fold_in(k + c_s21z[s], g.nz, 0);
    return (kk * g.ny + jj) * g.nx + ii;
}
// the 21-slot product of row q: plain offsets where no slot reaches an edge
__device__ inline double prod21(const Geo3 &g, const double *P, const double *x, int q) {
    const int n = g.nint;
    const int i = q % g.nx, j = (q / g.nx) % g.ny, k = q / (g.nx * g.ny);
    double s = 0.0;
    if (i >= 2 && i < g.nx - 2 && j >= 2 && j < g.ny - 2 && k >= 2 && k < g.nz - 2) {
#pragma unroll
        for (int t = 0; t < 21; ++t) s += P[(size_t)t * n + q] * x[q + g.off[t]];
    } else {
        for (int t = 0; t < 21; ++t) s += P[(size_t)t * n + q] * x[col3(g, i, j, k, t)];
    }
    return s;
}

// block partial sums of up to 3 products, reduced in a fixed order
template <int K>
__device__ inline void block_partials(double (&v)[K], double *part) {
    __shared__ double red[K][TBK];
    for (int k = 0; k < K; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = TBK / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < K; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < K; ++k) part[k * RB + blockIdx.x] = red[k][0];
}
// y = P x, and partial (w, y)
__global__ void k3_spmv(Geo3 g, const double *P, const double *x, double *y, const double *w, double *part) {
    double acc[1] = {0.0};
    const int n = g.nint;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const double s = prod21(g, P, x, q);
        y[q] = s;
        if (w) acc[0] += w[q] * s;
    }
    if (part) block_partials<1>(acc, part);
}
// the column preconditioner: the zeta bands of each column, folded, by a
// banded LU without pivoting (forward elimination streamed up the column)
__global__ void k3_lu(Geo3 g, const double *P, double *L1, double *L2, double *U0, double *U1, double *U2) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = g.nx * g.ny;
    if (c >= nc) return;
    const int n = g.nint, nz = g.nz;
    // the band row of level l: entries at offsets -2 .. 2
    auto band = [&](int l, double r[5]) {
        for (int o = 0; o < 5; ++o) r[o] = 0.0;
        const int q = l * nc + c;
        for (int m = 0; m < 5; ++m) {      // slots 0..4: levels l-2 .. l+2
            int t = l + m - 2;
            if (t < 0) t = -1 - t;
            if (t >= nz) t = 2 * nz - 1 - t;
            r[t - l + 2] += P[(size_t)m * n + q];
        }
    };
    double r0[5], r1[5], r2[5];
    band(0, r0);
    if (nz > 1) band(1, r1);
    for (int k = 0; k < nz; ++k) {
        if (k + 2 < nz) band(k + 2, r2);
        const int q = k * nc + c;
        const double piv = r0[2];
        const double ip = 1.0 / piv;
        U0[q] = ip; U1[q] = r0[3]; U2[q] = r0[4];
        if (k + 1 < nz) {
            const double f1 = r1[1] * ip;
            L1[q + nc] = f1;
            r1[2] -= f1 * r0[3];
            r1[3] -= f1 * r0[4];
        }
        if (k + 2 < nz) {
            const double f2 = r2[0] * ip;
            L2[q + 2 * nc] = f2;
            r2[1] -= f2 * r0[3];
            r2[2] -= f2 * r0[4];
        }
        for (int o = 0; o < 5; ++o) { r0[o] = r1[o]; r1[o] = r2[o]; }
    }
}
__global__ void k3_prec(Geo3 g, const double *L1, const double *L2, const double *U0, const double *U1,
                        const double *U2, const double *b, double *x) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = g.nx * g.ny;
    if (c >= nc) return;
    const int nz = g.nz;
    double y1 = 0.0, y2 = 0.0;   // y[k-1], y[k-2]
    for (int k = 0; k < nz; ++k) {
        const int q = k * nc + c;
        double y = b[q];
        if (k >= 1) y -= L1[q] * y1;
        if (k >= 2) y -= L2[q] * y2;
        x[q] = y;
        y2 = y1; y1 = y;
    }
    double x1 = 0.0, x2 = 0.0;   // x[k+1], x[k+2]
    for (int k = nz - 1; k >= 0; --k) {
        const int q = k * nc + c;
        double v = x[q];
        if (k + 1 < nz) v -= U1[q] * x1;
        if (k + 2 < nz) v -= U2[q] * x2;
        v *= U0[q];
        x[q] = v;
        x2 = x1; x1 = v;
    }
}
// one block sums the partials of K quantities in order
__global__ void k_final(const double *part, int K, double *out) {
    __shared__ double red[TBK];
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
        for (int b = threadIdx.x; b < RB; b += TBK) s += part[k * RB + b];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int t = TBK / 2; t > 0; t >>= 1) {
            if ((int)threadIdx.x < t) red[threadIdx.x] += red[threadIdx.x + t];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[k] = red[0];
        __syncthreads();
    }
}
// r = b - A x (x = 0 here: r = b), rh = r; partials (b, b), (r, r)
__global__ void k3_init(int n, const double *b, double *r, double *rh, double *p, double *v, double *part) {
    double acc[2] = {0.0, 0.0};
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const double v0 = b[q];
        r[q] = v0; rh[q] = v0; p[q] = 0.0; v[q] = 0.0;
        acc[0] += v0 * v0;
        acc[1] += v0 * v0;
    }
    block_partials<2>(acc, part);
}
__global__ void k3_p(int n, double beta, double omega, const double *r, double *p, const double *v) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
        p[q] = r[q] + beta * (p[q] - omega * v[q]);
}
// s = r - alpha v; partial (s, s)
__global__ void k3_s(int n, double alpha, const double *r, const double *v, double *s, double *part) {
    double acc[1] = {0.0};
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const double t = r[q] - alpha * v[q];
        s[q] = t;
        acc[0] += t * t;
    }
    block_partials<1>(acc, part);
}
// t = A sh; partials (t, t), (t, s)
__global__ void k3_t(Geo3 g, const double *P, const double *sh, double *t, const double *s, double *part) {
    double acc[2] = {0.0, 0.0};
    const int n = g.nint;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const double sum = prod21(g, P, sh, q);
        t[q] = sum;
        acc[0] += sum * sum;
        acc[1] += sum * s[q];
    }
    block_partials<2>(acc, part);
}
// x += alpha ph + omega sh; r = s - omega t; partials (r, r), (r, rh)
__global__ void k3_xr(int n, double alpha, double omega, double *x, const double *ph, const double *sh, double *r,
                      const double *s, const double *t, const double *rh, double *part) {
    double acc[2] = {0.0, 0.0};
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        x[q] += alpha * ph[q] + omega * sh[q];
        const double v = s[q] - omega * t[q];
        r[q] = v;
        acc[0] += v * v;
        acc[1] += v * rh[q];
    }
    block_partials<2>(acc, part);
}
__global__ void k3_xa(int n, double alpha, double *x, const double *ph) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) x[q] += alpha * ph[q];
}

// ---------------------------------------------------------------------
// after the solve: the increment with its ghosts from the rows, the
// solute the solver's residual leaves put back evenly, the end state
// ---------------------------------------------------------------------
// cn = c_old + d on the real cells (the ghosts follow from the rows)
__global__ void k_cplus(Geo3 g, const double *cold, const double *dint, double *cn) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.nint) return;
    const int ix = q % g.nx + 2, jy = (q / g.nx) % g.ny + 2, kz = q / (g.nx * g.ny) + 2;
    const int r = I3(g, ix, jy, kz);
    cn[r] = cold[r] + dint[q];
}
// d = cn - c_old and ct = d theta + c_old, every row
__global__ void k_cdiff(int n3, double theta, const double *cold, const double *cn, double *d, double *ct) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n3) return;
    const double v = cn[r] - cold[r];
    d[r] = v;
    const double t = v * theta;
    ct[r] = t + cold[r];
}
// partials of sum h (-dt (A ct + B) - d) and sum h (1 + theta dt react)
__global__ void k_csolute(Args a, const double *ct, const double *d, double *part) {
    const Geo3 &g = a.g;
    double acc[2] = {0.0, 0.0};
    const int n = g.nint;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const int ix = q % g.nx + 2, jy = (q / g.nx) % g.ny + 2, kz = q / (g.nx * g.ny) + 2;
        const int r3 = I3(g, ix, jy, kz);
        double A[21];
#pragma unroll
        for (int k = 0; k < 21; ++k) A[k] = a.cA[(size_t)k * n + q];
        const double s = balance_form(g, A, a.react[q], ct, ix, jy, kz);
        const double h = a.fH[I2(g, ix, jy)];
        acc[0] += h * (-a.dt * (s + 0.0) - d[r3]);
        acc[1] += h * (1.0 + a.theta * a.dt * a.react[q]);
    }
    block_partials<2>(acc, part);
}
__global__ void k_cnew(int n3, double alpha, const double *cold, const double *d, double *cnew) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n3) return;
    cnew[r] = cold[r] + (d[r] + alpha);
}

// ---------------------------------------------------------------------
// coupling passes: r = (c0 + c) / 2 - cm with max |r|; Aitken sums
// ---------------------------------------------------------------------
__global__ void k_resid(int n3, const double *c0, const double *c, const double *cm, double *r,
                        unsigned long long *emax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (i < n3) {
        const double v = (c0[i] + c[i]) * 0.5 - cm[i];
        r[i] = v;
        e = fabs(v);
    }
    for (int o = 32; o > 0; o >>= 1) e = fmax(e, __shfl_xor(e, o));
    if ((threadIdx.x & 63) == 0 && e > 0.0) atomicMax(emax, (unsigned long long)__double_as_longlong(e));
}
// SettleLimiter's test: emax[0] = max |c - prev|, emax[1] = max |c - c_old|
// (bits of non-negative doubles, atomicMax); then prev = c
__global__ void k_settle_max(int n3, const double *c, double *prev, const double *cold, unsigned long long *emax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0, f = 0.0;
    if (i < n3) {
        const double v = c[i];
        e = fabs(v - prev[i]);
        f = fabs(v - cold[i]);
        prev[i] = v;
    }
    for (int o = 32; o > 0; o >>= 1) { e = fmax(e, __shfl_xor(e, o)); f = fmax(f, __shfl_xor(f, o)); }
    if ((threadIdx.x & 63) == 0) {
        if (e > 0.0) atomicMax(emax, (unsigned long long)__double_as_longlong(e));
        if (f > 0.0) atomicMax(emax + 1, (unsigned long long)__double_as_longlong(f));
    }
}
__global__ void k_aitken(int n3, const double *r, const double *rp, double *part) {
    double acc[2] = {0.0, 0.0};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += gridDim.x * blockDim.x) {
        const double d = r[i] - rp[i];
        acc[0] += rp[i] * d;
        acc[1] += d * d;
    }
    block_partials<2>(acc, part);
}
__global__ void k_relax(int n3, double omega, const double *r, double *cm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) cm[i] += omega * r[i];
}
__global__ void k_avg(int n3, const double *x, const double *y, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) out[i] = (x[i] + y[i]) * 0.5;
}
__global__ void k_extrap(int n3, double rr, const double *c, const double *prev, double *cm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) { const double v = c[i]; cm[i] = v + rr * (v - prev[i]); }
}

// ---------------------------------------------------------------------
// RedistributeExcess and RedistributeDeficit (CZ_ConcEquation_2D.cpp
// :388-570), one wave per real column (kColPerBlock columns to a block);
// a column's levels are taken 64 at a time, one to a lane.
//
// Excess: what a level holds above 1 runs down the column, level to level;
// with out_k the excess leaving level k downward,
//   out_k = max(c_k - 1 + out_{k+1}, 0),
// a recurrence of maps x -> max(a + x, b), which compose associatively as
// (a, b) o (a', b') = (a + a', max(a + b', b)): a scan over the lanes.
// What reaches the bottom level above 1 fills the room (1 - c) of the
// levels from the bottom up (a prefix sum of the room), and what no level
// has room for is spread evenly through the column. The solute moved is
// every excess that left a level plus what filled room.
//
// Deficit: within a column, the solute below 0 is taken from the levels
// above 0 in proportion to what they hold, or, when the column holds too
// little, the column is made uniform at its (negative) mean; between
// columns, a column whose mean is below 0 takes from the room (h times the
// mean) of its four neighbours, which give up the fractions taken, for up
// to four passes; what is left is taken from every column in proportion.
// ---------------------------------------------------------------------
constexpr int kColPerBlock = 4;
inline int col_blocks(int ncol) { return (ncol + kColPerBlock - 1) / kColPerBlock; }
__device__ inline int col_index() { return blockIdx.x * kColPerBlock + (int)(threadIdx.x >> 6); }
__device__ inline int lane64() { return (int)(threadIdx.x & 63); }
// the sum over the 64 lanes, the same in every lane (butterfly, fixed order)
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// inclusive prefix sum over the lanes
__device__ inline double wave_scan(double v, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}
// the sum of a column's real levels, in chunk order
__device__ inline double col_total(const Geo3 &g, const double *c, int p, int lane) {
    double s = 0.0;
    for (int k0 = 0; k0 < g.nz; k0 += 64) {
        const int k = k0 + lane;
        s += wave_sum(k < g.nz ? c[(k + 2) * g.n2 + p] : 0.0);
    }
    return s;
}
__device__ inline int col_pad(const Geo3 &g, int q) { return I2(g, q % g.nx + 2, q / g.nx + 2); }

__global__ void __launch_bounds__(64 * kColPerBlock) k_excess(Geo3 g, double *c, double *moved) {
    const int q = col_index();
    if (q >= g.nx * g.ny) return;
    const int lane = lane64(), p = col_pad(g, q), nz = g.nz;
    const double NEG = -1.0e300;
    // every lane keeps to its own levels: i = 0 the top (level nz + 1), the
    // chunks of 64 from the top down
    auto row = [&](int i) { return (nz + 1 - i) * g.n2 + p; };
    // down the column
    double carry = 0.0, mv = 0.0, ybot = 0.0;   // what enters the chunk from above; moved; the bottom's value
    for (int i0 = 0; i0 < nz; i0 += 64) {
        const int i = i0 + lane;
        const bool lv = i < nz, bottom = i == nz - 1;
        const double ck = lv ? c[row(i)] : 0.0;
        double A = (lv && !bottom) ? ck - 1.0 : 0.0, B = (lv && !bottom) ? 0.0 : NEG;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const double a2 = __shfl_up(A, o, 64), b2 = __shfl_up(B, o, 64);
            if (lane >= o) { B = fmax(A + b2, B); A = A + a2; }
        }
        const double out = fmax(A + carry, B);                 // leaves this level downward
        double above = __shfl_up(out, 1, 64);                   // entered it from above
        if (lane == 0) above = carry;
        if (lv && !bottom) {
            c[row(i)] = (out > 0.0) ? 1.0 : ck + above;
            mv += out;
        }
        if (bottom) ybot = ck + above;
        carry = __shfl(out, 63, 64);
    }
    mv = wave_sum(mv);
    const int lb = (nz - 1) & 63;
    const double e = __shfl(ybot, lb, 64) - 1.0;
    if (!(e > 0.0)) {
        if (lane == lb) c[row(nz - 1)] = ybot;
        if (lane == 0) moved[q] = mv;
        return;
    }
    if (lane == lb) c[row(nz - 1)] = 1.0;
    // the bottom's excess fills the room of the levels above it, bottom up:
    // chunks from the bottom one, lanes from the highest
    double below = 0.0, filled = 0.0;           // room in the chunks below; solute taken
    for (int i0 = ((nz - 1) / 64) * 64; i0 >= 0; i0 -= 64) {
        const int i = i0 + lane;
        const bool lv = i < nz - 1;
        const double ck = lv ? c[row(i)] : 1.0;
        const double room = fmax(1.0 - ck, 0.0);
        double upto = room;                     // room of this level and every level below it in the chunk
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const double u = __shfl_down(upto, o, 64);
            if (lane + o < 64) upto += u;
        }
        upto += below;
        const double take = fmin(fmax(e - (upto - room), 0.0), room);
        if (lv && take > 0.0) c[row(i)] = ck + take;
        filled += take;
        below = __shfl(upto, 0, 64);
    }
    filled = wave_sum(filled);
    mv += filled;
    const double left = e - filled;
    if (left > 0.0)
        for (int i = lane; i < nz; i += 64) c[row(i)] += left / (double)nz;
    if (lane == 0) moved[q] = mv;
}
// whether any real cell is below 0
__global__ void k_below(Geo3 g, const double *c, int *flags) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nint) return;
    const int ix = t % g.nx + 2, jy = (t / g.nx) % g.ny + 2, kz = t / (g.nx * g.ny) + 2;
    if (c[I3(g, ix, jy, kz)] < 0.0) atomicOr(&flags[1], 1);
}
// within each column: the solute below 0 from its levels above 0
__global__ void __launch_bounds__(64 * kColPerBlock) k_deficit_col(Geo3 g, const double *H, double *c, double *moved,
                                                                   int *flags) {
    const int q = col_index();
    if (q >= g.nx * g.ny) return;
    const int lane = lane64(), p = col_pad(g, q), nz = g.nz;
    double lack = 0.0, hold = 0.0;
    for (int k0 = 0; k0 < nz; k0 += 64) {
        const int k = k0 + lane;
        const double ck = k < nz ? c[(k + 2) * g.n2 + p] : 0.0;
        lack += wave_sum(ck < 0.0 ? -ck : 0.0);
        hold += wave_sum(ck < 0.0 ? 0.0 : ck);
    }
    if (lane == 0) moved[q] = lack * H[p] / (double)nz;
    if (lack == 0.0) return;
    if (hold >= lack) {
        const double f = lack / hold;
        for (int k = lane; k < nz; k += 64) {
            const int r = (k + 2) * g.n2 + p;
            const double ck = c[r];
            c[r] = (ck < 0.0) ? 0.0 : ck - f * ck;
        }
    } else {
        const double m = (hold - lack) / (double)nz;
        for (int k = lane; k < nz; k += 64) c[(k + 2) * g.n2 + p] = m;
        if (lane == 0) atomicOr(&flags[2], 1);
    }
}
__device__ inline void col_fill(const Geo3 &g, double *c, int p, int lane, double v) {
    for (int k = lane; k < g.nz; k += 64) c[(k + 2) * g.n2 + p] = v;
}
__device__ inline void col_scale(const Geo3 &g, double *c, int p, int lane, double f) {
    for (int k = lane; k < g.nz; k += 64) c[(k + 2) * g.n2 + p] *= f;
}
// a pass between columns: the room (h times the mean) of every column
__global__ void __launch_bounds__(64 * kColPerBlock) k_def_room(Geo3 g, const double *H, const double *c, double *room,
                                                                double *fcol) {
    const int q = col_index();
    if (q >= g.nx * g.ny) return;
    const int lane = lane64(), p = col_pad(g, q);
    const double m = col_total(g, c, p, lane) / (double)g.nz, h = H[p];
    if (lane == 0) {
        room[q] = (m > 0.0 && h > 0.0) ? h * m : 0.0;
        fcol[q] = 0.0;
    }
}
// the columns below 0: the fraction f of their neighbours' room they take
// (all of it when that is not enough) and their new mean; added: what a
// column without film makes up
__global__ void __launch_bounds__(64 * kColPerBlock) k_def_take(Geo3 g, const double *H, double *c, const double *room,
                                                                double *fcol, double *added, int *flags) {
    const int q = col_index();
    if (q >= g.nx * g.ny) return;
    const int lane = lane64(), p = col_pad(g, q);
    const int i = q % g.nx, j = q / g.nx;
    const double m = col_total(g, c, p, lane) / (double)g.nz, h = H[p];
    if (lane == 0) added[q] = 0.0;
    if (!(m < 0.0)) return;
    if (!(h > 0.0)) {
        if (lane == 0) added[q] = -(h * m);
        col_scale(g, c, p, lane, 0.0);
        return;
    }
    const double need = -h * m;
    double avail = 0.0;
    if (i > 0) avail += room[q - 1];
    if (i < g.nx - 1) avail += room[q + 1];
    if (j > 0) avail += room[q - g.nx];
    if (j < g.ny - 1) avail += room[q + g.nx];
    if (!(avail > 0.0)) {
        if (lane == 0) atomicOr(&flags[3], 1);
        return;
    }
    const double f = (need < avail) ? need / avail : 1.0;
    if (lane == 0) {
        fcol[q] = f;
        if (f >= 1.0) atomicOr(&flags[3], 1);
    }
    col_fill(g, c, p, lane, (f < 1.0) ? 0.0 : -(need - avail) / h);
}
// the neighbours give up what was taken from them; the fractions are added
// west, south, north, east
__global__ void __launch_bounds__(64 * kColPerBlock) k_def_give(Geo3 g, const double *fcol, double *c, int *flags) {
    const int q = col_index();
    if (q >= g.nx * g.ny) return;
    const int lane = lane64(), p = col_pad(g, q);
    const int i = q % g.nx, j = q / g.nx;
    double tk = 0.0;
    bool got = false;
    const int nb[4] = {i > 0 ? q - 1 : -1, j > 0 ? q - g.nx : -1, j < g.ny - 1 ? q + g.nx : -1, i < g.nx - 1 ? q + 1 : -1};
    for (int u = 0; u < 4; ++u)
        if (nb[u] >= 0 && fcol[nb[u]] > 0.0) { tk += fcol[nb[u]]; got = true; }
    if (!(got && tk > 0.0)) return;
    col_scale(g, c, p, lane, 1.0 - tk);
    if (col_total(g, c, p, lane) / (double)g.nz < 0.0 && lane == 0) atomicOr(&flags[3], 1);
}
// what is still lacking, and the solute every other column holds
__global__ void __launch_bounds__(64 * kColPerBlock) k_def_left(Geo3 g, const double *H, double *c, double *lack,
                                                                double *total) {
    const int q = col_index();
    if (q >= g.nx * g.ny) return;
    const int lane = lane64(), p = col_pad(g, q);
    const double m = col_total(g, c, p, lane) / (double)g.nz, h = H[p];
    if (lane == 0) { lack[q] = 0.0; total[q] = 0.0; }
    if (m < 0.0) {
        if (lane == 0) lack[q] = -(h * m);
        col_scale(g, c, p, lane, 0.0);
    } else if (h > 0.0 && lane == 0) {
        total[q] = h * m;
    }
}
__global__ void __launch_bounds__(64 * kColPerBlock) k_def_scale(Geo3 g, const double *H, double f, double *c) {
    const int q = col_index();
    if (q >= g.nx * g.ny) return;
    const int lane = lane64(), p = col_pad(g, q);
    if (H[p] > 0.0 && col_total(g, c, p, lane) / (double)g.nz > 0.0) col_scale(g, c, p, lane, 1.0 - f);
}
// partial sums of a per-column array in a fixed order (reduced by k_final)
__global__ void k_colsum(int n, const double *v, double *part) {
    double acc[1] = {0.0};
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) acc[0] += v[q];
    block_partials<1>(acc, part);
}


// ---- the learned film solve ----------------------------------------------

// The film's FP32 system and the learned preconditioner's inputs from the
// FP64 row-scaled planes P (13 x N) and the film h at the solve's start:
// the FP32 planes (a diagonal the scaling made 1 within an ulp stored as
// exactly 1, else *nonunit set), and per row the features the network takes
// (the stiffness and shape of the scaled row, the film) and the row scale
// 1 / (1 + sum of |off-diagonal|), as the well-mixed film's row kernel forms
// them
__global__ void k_cz_feat(int N, const double *__restrict__ P, const double *__restrict__ h, float *__restrict__ A32,
                          float *__restrict__ feat, float *__restrict__ mscale, int *__restrict__ nonunit) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double IA[13];
#pragma unroll
    for (int k = 0; k < 13; ++k) IA[k] = P[(size_t)k * N + i];
#pragma unroll
    for (int k = 0; k < 13; ++k) A32[(size_t)k * N + i] = (float)IA[k];
    if (fabs(IA[6] - 1.0) <= 1.0e-12) A32[(size_t)6 * N + i] = 1.0f;
    else atomicOr(nonunit, 1);
    double off = 0;
#pragma unroll
    for (int k = 0; k < 13; ++k) if (k != 6) off += fabs(IA[k]);
    // the far points of the four faces: Mo, Po, oM, oP
    const double far = IA[4] + IA[8] + IA[0] + IA[12];
    const double so = off > 1e-30 ? off : 1e-30;
    feat[0 * (size_t)N + i] = (float)(0.25 * log10(so));
    feat[1 * (size_t)N + i] = (float)(IA[5] / so);
    feat[2 * (size_t)N + i] = (float)(IA[7] / so);
    feat[3 * (size_t)N + i] = (float)(IA[2] / so);
    feat[4 * (size_t)N + i] = (float)(IA[10] / so);
    feat[5 * (size_t)N + i] = (float)(far / so);
    feat[6 * (size_t)N + i] = (float)h[i];
    feat[7 * (size_t)N + i] = 0.0f;
    mscale[i] = (float)(1.0 / (1.0 + off));
}
// b32 = r / s (the residual at a scale FP32 holds well)
__global__ void k_cz_to32(int N, const double *__restrict__ r, double inv_s, float *__restrict__ b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) b[i] = (float)(r[i] * inv_s);
}
// x += s d
__global__ void k_cz_refine(int N, double *__restrict__ x, double s, const float *__restrict__ d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) x[i] += s * (double)d[i];
}

}  // namespace

// ---------------------------------------------------------------------
// host
// ---------------------------------------------------------------------
struct yz_cz {
    Geo3 g;
    yz_cz_case cs;
    double dt = 0;
    hipStream_t st = nullptr;
    Args a;
    Solver Sh;                              // the film's solver
    // device state
    double *S = nullptr, *hO = nullptr, *hN = nullptr, *fH = nullptr, *tH = nullptr, *mid = nullptr;
    double *hO_keep = nullptr, *hN_keep = nullptr;
    double *fC = nullptr, *tC = nullptr, *cO = nullptr, *cN = nullptr, *c0 = nullptr, *cm = nullptr;
    double *cPrev = nullptr, *keepO = nullptr, *keepN = nullptr, *prevKeep = nullptr, *r = nullptr, *rp = nullptr;
    double *cn = nullptr, *d = nullptr, *ct = nullptr;
    double *cSettle = nullptr;              // SettleLimiter's previous answer
    // the step control: the start of the step before (_phi_old2), DT().old,
    // the count of accepted steps toward lengthening, the rejected trials
    double *hO2 = nullptr;
    double dt_old = 0;
    int tsc_count = 0;
    long long cuts = 0;
    double *f2[16] = {nullptr};
    double *Qx = nullptr, *Qy = nullptr, *W = nullptr, *zr = nullptr;
    double *cA = nullptr, *react = nullptr, *cP = nullptr, *cb = nullptr;
    double *kv[9] = {nullptr};              // x r rh p v s t ph sh (real cells)
    double *lu[5] = {nullptr};
    double *part = nullptr, *dsc = nullptr, *hsc = nullptr;
    double *colv[4] = {nullptr};
    int *flags = nullptr, *hflags = nullptr;
    unsigned long long *emax = nullptr;
    double dtPrev = 1.0;
    bool prevSet = false;
    double time = 0;
    long long steps = 0;
    double stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // the learned film solve (film_precond YZ_PC_LEARNED): the FP32 system,
    // the network's inputs and rows, the FP64 residual and the guess kept
    SolverT<float> S32;
    float *b32 = nullptr, *x32 = nullptr, *feat = nullptr, *mscale = nullptr;
    __half *Yh = nullptr;
    double *r64 = nullptr, *xkeep = nullptr;
    int *nonunit = nullptr, *hnonunit = nullptr;
    Mlp mlp;
    double lstats[4] = {0, 0, 0, 0};
};

namespace {

template <typename T>
int dal(T **p, size_t n) {
    if (hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
        yz::set_error("ConcZ: device allocation failed");
        return -1;
    }
    return hipMemset(*p, 0, std::max<size_t>(n, 1) * sizeof(T)) == hipSuccess ? 0 : -1;
}
int cp_d2d(yz_cz *z, double *dst, const double *src, size_t n) {
    CZ_CHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, z->st));
    return 0;
}
int sync(yz_cz *z) {
    CZ_CHECK(hipStreamSynchronize(z->st));
    CZ_CHECK(hipGetLastError());
    return 0;
}
// read K reduced partials
int read_sums(yz_cz *z, int K, double *out) {
    hipLaunchKernelGGL(k_final, dim3(1), dim3(TBK), 0, z->st, z->part, K, z->dsc);
    CZ_CHECK(hipMemcpyAsync(z->hsc, z->dsc, K * sizeof(double), hipMemcpyDeviceToHost, z->st));
    if (sync(z)) return -1;
    for (int k = 0; k < K; ++k) out[k] = z->hsc[k];
    return 0;
}
int read_flags(yz_cz *z) {
    CZ_CHECK(hipMemcpyAsync(z->hflags, z->flags, 4 * sizeof(int), hipMemcpyDeviceToHost, z->st));
    return sync(z);
}
int zero_flags(yz_cz *z) {
    CZ_CHECK(hipMemsetAsync(z->flags, 0, 4 * sizeof(int), z->st));
    return 0;
}
double colsum(yz_cz *z, const double *v, int *err) {
    hipLaunchKernelGGL(k_colsum, dim3(RB), dim3(TBK), 0, z->st, z->g.nx * z->g.ny, v, z->part);
    double out = 0.0;
    if (read_sums(z, 1, &out)) *err = 1;
    return out;
}

int apply_rows(yz_cz *z, double *x) {
    const Geo3 &g = z->g;
    const int nc = g.nx * g.ny;
    hipLaunchKernelGGL(k_cghost_z, dim3(nblk(nc)), dim3(TBK), 0, z->st, g, z->zr, x);
    hipLaunchKernelGGL(k_cghost_y, dim3(nblk((long)g.nx * 4 * g.Nz)), dim3(TBK), 0, z->st, g, x);
    hipLaunchKernelGGL(k_cghost_x, dim3(nblk((long)4 * g.Ny * g.Nz)), dim3(TBK), 0, z->st, g, x);
    return 0;
}

// The film's solve with the learned preconditioner: FP32 BiCGSTAB with the
// network's approximate inverse on the scaled FP64 residual, the correction
// added in FP64, round after round until the FP64 residual meets h_eps;
// FP64 AMG finishes from the refined film when a round fails to halve the
// residual. Returns the solve's code (-1 on a device error).
constexpr int kRefineRounds = 10;
constexpr double kInnerTol = 1e-5;
constexpr int kInnerMaxit = 600;
int h_solve_learned(yz_cz *z, int *iters, double *resid) {
    Args &a = z->a;
    const int N = z->Sh.op0.N;
    const yz_cz_case &c = z->cs;
    const dim3 gb(nblk(N)), tb(TBK);
    CZ_CHECK(hipMemsetAsync(z->nonunit, 0, sizeof(int), z->st));
    hipLaunchKernelGGL(k_cz_feat, gb, tb, 0, z->st, N, (const double *)a.hP, (const double *)a.hx, z->S32.A, z->feat,
                       z->mscale, z->nonunit);
    CZ_CHECK(hipMemcpyAsync(z->hnonunit, z->nonunit, sizeof(int), hipMemcpyDeviceToHost, z->st));
    if (sync(z)) return -1;
    z->S32.unit_diag = z->hnonunit[0] == 0;
    if (z->mlp.trained_passes < c.learn_passes && c.learn_iters > 0) {
        hipEvent_t e0, e1;
        CZ_CHECK(hipEventCreate(&e0));
        CZ_CHECK(hipEventCreate(&e1));
        CZ_CHECK(hipEventRecord(e0, z->st));
        double l0 = 0, l1 = 0;
        const int rc = mlp_train(z->mlp, z->S32, z->feat, z->mscale, c.learn_iters, c.learn_lr, c.learn_probes,
                                 0x2345u + 104729u * (unsigned)z->mlp.trained_passes, &l0, &l1, z->st);
        CZ_CHECK(hipEventRecord(e1, z->st));
        CZ_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc) { yz::set_error("ConcZ: training of the learned preconditioner failed"); return -1; }
        z->lstats[3] += ms;
        ++z->mlp.trained_passes;
    }
    if (mlp_infer_half(z->mlp, N, z->feat, z->Yh, z->st)) { yz::set_error("ConcZ: inference failed"); return -1; }
    z->S32.Mh = z->Yh;
    z->S32.Mscale = z->mscale;
    double bb = 0, rr = 0;
    if (residual_norms(z->Sh, (const double *)a.hx, (const double *)a.hb, z->r64, &bb, &rr)) return -1;
    const double nb = std::sqrt(bb), eps = c.h_eps * nb;
    double res = std::sqrt(rr);
    if (nb == 0) {
        CZ_CHECK(hipMemsetAsync(a.hx, 0, N * sizeof(double), z->st));
        *resid = 0;
        return 0;
    }
    for (int round = 0; round < kRefineRounds && res > eps; ++round) {
        const double sc = res / std::sqrt((double)N);
        hipLaunchKernelGGL(k_cz_to32, gb, tb, 0, z->st, N, (const double *)z->r64, 1.0 / sc, z->b32);
        CZ_CHECK(hipMemsetAsync(z->x32, 0, N * sizeof(float), z->st));
        int it = 0;
        double r32 = 0;
        const int ch = bicgstab_stencil(z->S32, (const float *)z->b32, z->x32, kInnerMaxit, kInnerTol, &it, &r32);
        if (ch < 0) return -1;
        *iters += it;
        z->lstats[1] += it;
        z->lstats[0] += 1;
        if (std::getenv("YZ_DEBUG"))
            std::fprintf(stderr, "[cz] learned round %d: |r| %.3e of %.3e, FP32 iterations %d (rc %d, %.2e)\n", round,
                         res, nb, it, ch, r32);
        if (ch == 2) break;                 // a breakdown: the correction is not kept
        CZ_CHECK(hipMemcpyAsync(z->xkeep, a.hx, N * sizeof(double), hipMemcpyDeviceToDevice, z->st));
        hipLaunchKernelGGL(k_cz_refine, gb, tb, 0, z->st, N, a.hx, sc, (const float *)z->x32);
        if (residual_norms(z->Sh, (const double *)a.hx, (const double *)a.hb, z->r64, &bb, &rr)) return -1;
        const double next = std::sqrt(rr);
        if (!(next <= 0.5 * res)) {
            // the round did not halve the residual: its correction is taken
            // back, and the residual of the film before it restored
            CZ_CHECK(hipMemcpyAsync(a.hx, z->xkeep, N * sizeof(double), hipMemcpyDeviceToDevice, z->st));
            if (residual_norms(z->Sh, (const double *)a.hx, (const double *)a.hb, z->r64, &bb, &rr)) return -1;
            res = std::sqrt(rr);
            break;
        }
        res = next;
    }
    if (res <= eps) {
        *resid = res / nb;
        return 0;
    }
    // FP64 AMG from the refined film
    z->lstats[2] += 1;
    if (amg_setup(z->Sh)) { yz::set_error("ConcZ: film preconditioner set-up failed"); return -1; }
    int it = 0;
    const int ch = bicgstab(z->Sh, (const double *)a.hb, a.hx, YZ_PC_AMG, c.lin_maxit, c.h_eps, &it, resid);
    *iters += it;
    return ch;
}

// SolveStage of the film: assemble at the field, solve into hN
int h_stage(yz_cz *z, int &iter, double &resid) {
    const Geo3 &g = z->g;
    Args &a = z->a;
    hipLaunchKernelGGL(k_faces, dim3(nblk(g.n2)), dim3(TBK), 0, z->st, a);
    hipLaunchKernelGGL(k_hrows, dim3(nblk(g.nx * g.ny)), dim3(TBK), 0, z->st, a);
    if (sync(z)) return -1;
    int it = 0;
    double rs = 0;
    int ch;
    if (z->cs.film_precond == YZ_PC_LEARNED) {
        ch = h_solve_learned(z, &it, &rs);
        if (ch < 0) return -1;
    } else {
        if (amg_setup(z->Sh)) { yz::set_error("ConcZ: film preconditioner set-up failed"); return -1; }
        ch = bicgstab(z->Sh, (const double *)a.hb, a.hx, YZ_PC_AMG, z->cs.lin_maxit, z->cs.h_eps, &it, &rs);
    }
    iter += it;
    resid = rs;
    z->stats[0] += 1;
    z->stats[1] += it;
    if (ch != 0) return 1;
    hipLaunchKernelGGL(k_hscatter, dim3(nblk(g.nx * g.ny)), dim3(TBK), 0, z->st, g, (const double *)a.hx, z->hN);
    hipLaunchKernelGGL(k_hghosts, dim3(nblk(g.n2)), dim3(TBK), 0, z->st, g, z->hN);
    return 0;
}
void conservative_update(yz_cz *z) {
    const Geo3 &g = z->g;
    hipLaunchKernelGGL(k_hghosts, dim3(nblk(g.n2)), dim3(TBK), 0, z->st, g, z->hN);
    hipLaunchKernelGGL(k_hflux, dim3(nblk(g.n2)), dim3(TBK), 0, z->st, z->a);
    hipLaunchKernelGGL(k_hupdate, dim3(nblk(g.nx * g.ny)), dim3(TBK), 0, z->st, z->a);
    hipLaunchKernelGGL(k_hghosts, dim3(nblk(g.n2)), dim3(TBK), 0, z->st, g, z->hN);
}

// BiCGSTAB on the concentration system (the oracle's and the reference
// backend's contract: ||r|| <= eps ||b||, 0 / 1 / 2), x = 0 to start
int c_krylov(yz_cz *z, int &iters, double &resid) {
    const Geo3 &g = z->g;
    const int n = g.nint;
    double *x = z->kv[0], *r = z->kv[1], *rh = z->kv[2], *p = z->kv[3], *v = z->kv[4], *s = z->kv[5], *t = z->kv[6],
           *ph = z->kv[7], *sh = z->kv[8];
    const double *P = z->cP;
    const int nc = g.nx * g.ny;
    hipLaunchKernelGGL(k3_lu, dim3(nblk(nc)), dim3(TBK), 0, z->st, g, P, z->lu[0], z->lu[1], z->lu[2], z->lu[3], z->lu[4]);
    CZ_CHECK(hipMemsetAsync(x, 0, n * sizeof(double), z->st));
    hipLaunchKernelGGL(k3_init, dim3(RB), dim3(TBK), 0, z->st, n, (const double *)z->cb, r, rh, p, v, z->part);
    double sc[3];
    if (read_sums(z, 2, sc)) return -1;
    const double nb = std::sqrt(sc[0]);
    iters = 0;
    resid = 0;
    if (nb == 0) return 0;
    const double eps = z->cs.c_eps * nb;
    double res = std::sqrt(sc[1]);
    if (res <= eps) { resid = res / nb; return 0; }
    double rho1 = sc[1], rho2 = 1, alpha = 1, omega = 1;   // (r, rh) = (r, r) at the start
    int it = 0;
    const int maxit = z->cs.lin_maxit;
    for (; it < maxit; ++it) {
        if (std::fabs(rho1) < 1e-300) { iters = it; resid = res / nb; return 2; }
        const double beta = it == 0 ? 0.0 : (rho1 * alpha) / (rho2 * omega);
        hipLaunchKernelGGL(k3_p, dim3(RB), dim3(TBK), 0, z->st, n, beta, it == 0 ? 0.0 : omega, (const double *)r, p,
                           (const double *)v);
        hipLaunchKernelGGL(k3_prec, dim3(nblk(nc)), dim3(TBK), 0, z->st, g, z->lu[0], z->lu[1], z->lu[2], z->lu[3],
                           z->lu[4], (const double *)p, ph);
        hipLaunchKernelGGL(k3_spmv, dim3(RB), dim3(TBK), 0, z->st, g, P, (const double *)ph, v, (const double *)rh,
                           z->part);
        if (read_sums(z, 1, sc)) return -1;
        const double rv = sc[0];
        if (rv == 0) { iters = it; resid = res / nb; return 2; }
        alpha = rho1 / rv;
        hipLaunchKernelGGL(k3_s, dim3(RB), dim3(TBK), 0, z->st, n, alpha, (const double *)r, (const double *)v, s, z->part);
        if (read_sums(z, 1, sc)) return -1;
        res = std::sqrt(sc[0]);
        if (res < eps) {
            hipLaunchKernelGGL(k3_xa, dim3(RB), dim3(TBK), 0, z->st, n, alpha, x, (const double *)ph);
            ++it;
            break;
        }
        hipLaunchKernelGGL(k3_prec, dim3(nblk(nc)), dim3(TBK), 0, z->st, g, z->lu[0], z->lu[1], z->lu[2], z->lu[3],
                           z->lu[4], (const double *)s, sh);
        hipLaunchKernelGGL(k3_t, dim3(RB), dim3(TBK), 0, z->st, g, P, (const double *)sh, t, (const double *)s, z->part);
        if (read_sums(z, 2, sc)) return -1;
        const double tt = sc[0];
        omega = (tt != 0) ? sc[1] / tt : 0.0;
        if (omega == 0) { iters = it; resid = res / nb; return 2; }
        hipLaunchKernelGGL(k3_xr, dim3(RB), dim3(TBK), 0, z->st, n, alpha, omega, x, (const double *)ph,
                           (const double *)sh, r, (const double *)s, (const double *)t, (const double *)rh, z->part);
        if (read_sums(z, 2, sc)) return -1;
        res = std::sqrt(sc[0]);
        rho2 = rho1;
        rho1 = sc[1];
        if (res < eps) { ++it; break; }
        if (!(res == res)) { iters = it; resid = res; return 2; }
    }
    iters = it;
    resid = res / nb;
    if (!(resid == resid)) return 2;
    return resid <= z->cs.c_eps ? 0 : 1;
}

// SolveStage of the concentration at the field (fC, fH), from cO into cN
int c_stage(yz_cz *z, int &iter, double &resid) {
    const Geo3 &g = z->g;
    Args &a = z->a;
    hipLaunchKernelGGL(k_cbc, dim3(nblk(g.nx * g.ny)), dim3(TBK), 0, z->st, a, 0);
    apply_rows(z, z->cO);
    if (zero_flags(z)) return -1;
    hipLaunchKernelGGL(k_crows, dim3(nblk(g.nint)), dim3(TBK), 0, z->st, a, (const double *)z->cO);
    if (read_flags(z)) return -1;
    const bool any = z->hflags[0] != 0;
    int ch = 0, it = 0;
    double rs = 0;
    if (any) {
        ch = c_krylov(z, it, rs);
        if (ch < 0) return -1;
        z->stats[2] += 1;
        z->stats[3] += it;
    } else {
        CZ_CHECK(hipMemsetAsync(z->kv[0], 0, g.nint * sizeof(double), z->st));
    }
    iter += it;
    resid = rs;
    if (ch != 0) return 2;
    // the increment with its ghosts from the rows, then the solute put back
    CZ_CHECK(hipMemsetAsync(z->cn, 0, g.n3 * sizeof(double), z->st));
    hipLaunchKernelGGL(k_cplus, dim3(nblk(g.nint)), dim3(TBK), 0, z->st, g, (const double *)z->cO,
                       (const double *)z->kv[0], z->cn);
    apply_rows(z, z->cn);
    hipLaunchKernelGGL(k_cdiff, dim3(nblk(g.n3)), dim3(TBK), 0, z->st, g.n3, z->a.theta, (const double *)z->cO,
                       (const double *)z->cn, z->d, z->ct);
    double alpha = 0.0;
    if (any) {
        hipLaunchKernelGGL(k_csolute, dim3(RB), dim3(TBK), 0, z->st, a, (const double *)z->ct, (const double *)z->d,
                           z->part);
        double sc[2];
        if (read_sums(z, 2, sc)) return -1;
        alpha = (sc[1] > 0.0) ? sc[0] / sc[1] : 0.0;
    }
    hipLaunchKernelGGL(k_cnew, dim3(nblk(g.n3)), dim3(TBK), 0, z->st, g.n3, alpha, (const double *)z->cO,
                       (const double *)z->d, z->cN);
    apply_rows(z, z->cN);
    return 0;
}

// the velocity through the film and the volume fluxes of the pass, at the
// field (fH, fC); BuildVolumeFluxes reads nothing the concentration's
// stages change, so both stages share them
void volume_fluxes(yz_cz *z) {
    const Geo3 &g = z->g;
    const int tw = vel_tile(g.nz);
    hipLaunchKernelGGL(k_face_velocity, dim3((g.nx + tw) / tw, g.ny + 1, 2), dim3(kVelThreads), vel_lds(g.nz, tw),
                       z->st, z->a, tw);
    hipLaunchKernelGGL(k_w, dim3(nblk(g.nx * g.ny)), dim3(TBK), 0, z->st, z->a);
}

// SettleLimiter (CZ_ConcEquation_2D.cpp:945-973): a limited scheme's step
// solved again with its limiter frozen at the midpoint of what came out,
// until that settles to 1e-12 of the step's change (8 passes at most)
int c_settle(yz_cz *z, int &iter, double &resid) {
    if (!z->a.lim) return 0;
    const Geo3 &g = z->g;
    for (int pass = 0; pass < 8; ++pass) {
        hipLaunchKernelGGL(k_avg, dim3(nblk(g.n3)), dim3(TBK), 0, z->st, g.n3, (const double *)z->cO,
                           (const double *)z->cN, z->fC);
        if (cp_d2d(z, z->cSettle, z->cN, g.n3)) return -1;
        int it = 0;
        double rs = 0;
        const int ch = c_stage(z, it, rs);
        iter += it;
        resid = rs;
        if (ch != 0) return ch;
        CZ_CHECK(hipMemsetAsync(z->emax, 0, 2 * sizeof(unsigned long long), z->st));
        hipLaunchKernelGGL(k_settle_max, dim3(nblk(g.n3)), dim3(TBK), 0, z->st, g.n3, (const double *)z->cN,
                           z->cSettle, (const double *)z->cO, z->emax);
        unsigned long long eb[2];
        CZ_CHECK(hipMemcpyAsync(eb, z->emax, sizeof(eb), hipMemcpyDeviceToHost, z->st));
        CZ_CHECK(hipStreamSynchronize(z->st));
        double dmax, cmax;
        std::memcpy(&dmax, &eb[0], sizeof(dmax));
        std::memcpy(&cmax, &eb[1], sizeof(cmax));
        if (!(dmax > 1e-12 * cmax)) break;
    }
    return 0;
}

int c_solve_at_mid(yz_cz *z, const double *cmid, int &iter, double &resid) {
    const size_t n3 = z->g.n3;
    if (cp_d2d(z, z->fC, cmid, n3)) return -1;
    int ch = c_stage(z, iter, resid);
    if (ch == 0) ch = c_settle(z, iter, resid);
    if (cp_d2d(z, z->fC, z->cO, n3)) return -1;
    return ch;
}
int c_solve(yz_cz *z, int &iter, double &resid) {
    const Geo3 &g = z->g;
    int ch = c_stage(z, iter, resid);
    if (ch != 0) return ch;
    // the midpoint of the prediction
    hipLaunchKernelGGL(k_avg, dim3(nblk(g.n3)), dim3(TBK), 0, z->st, g.n3, (const double *)z->cO,
                       (const double *)z->cN, z->fC);
    ch = c_stage(z, iter, resid);
    if (ch == 0) ch = c_settle(z, iter, resid);
    if (cp_d2d(z, z->fC, z->cO, g.n3)) return -1;
    return ch;
}

int redistribute_deficit(yz_cz *z, double &moved, double &added) {
    const Geo3 &g = z->g;
    const int nc = g.nx * g.ny;
    moved = added = 0;
    int err = 0;
    if (zero_flags(z)) return -1;
    hipLaunchKernelGGL(k_below, dim3(nblk(g.nint)), dim3(TBK), 0, z->st, g, (const double *)z->cN, z->flags);
    if (read_flags(z)) return -1;
    if (!z->hflags[1]) return 0;
    hipLaunchKernelGGL(k_deficit_col, dim3(col_blocks(nc)), dim3(64 * kColPerBlock), 0, z->st, g, (const double *)z->fH, z->cN,
                       z->colv[0], z->flags);
    moved = colsum(z, z->colv[0], &err);
    if (err || read_flags(z)) return -1;
    if (!z->hflags[2]) return 0;
    bool left = true;
    for (int pass = 0; pass < 4 && left; ++pass) {
        hipLaunchKernelGGL(k_def_room, dim3(col_blocks(nc)), dim3(64 * kColPerBlock), 0, z->st, g, (const double *)z->fH,
                           (const double *)z->cN, z->colv[0], z->colv[1]);
        if (zero_flags(z)) return -1;
        hipLaunchKernelGGL(k_def_take, dim3(col_blocks(nc)), dim3(64 * kColPerBlock), 0, z->st, g, (const double *)z->fH, z->cN,
                           (const double *)z->colv[0], z->colv[1], z->colv[2], z->flags);
        added += colsum(z, z->colv[2], &err);
        hipLaunchKernelGGL(k_def_give, dim3(col_blocks(nc)), dim3(64 * kColPerBlock), 0, z->st, g, (const double *)z->colv[1], z->cN,
                           z->flags);
        if (err || read_flags(z)) return -1;
        left = z->hflags[3] != 0;
    }
    if (!left) return 0;
    hipLaunchKernelGGL(k_def_left, dim3(col_blocks(nc)), dim3(64 * kColPerBlock), 0, z->st, g, (const double *)z->fH, z->cN, z->colv[0],
                       z->colv[1]);
    const double lack = colsum(z, z->colv[0], &err);
    const double total = colsum(z, z->colv[1], &err);
    if (err) return -1;
    const double f = (total > 0.0) ? std::min(1.0, lack / total) : 0.0;
    if (f > 0.0) hipLaunchKernelGGL(k_def_scale, dim3(col_blocks(nc)), dim3(64 * kColPerBlock), 0, z->st, g, (const double *)z->fH, f, z->cN);
    if (total < lack) added += lack - total;
    return 0;
}

// one step (TimeAdvance, Solve, AfterTimeStepTrial)
// LocalTruncationCriterion (ImplicitSolver.cpp:271-307) of the film the
// trial's solve gave: max over every row of |2 (dt/dt0) (dt0 h+ + dt h- -
// (dt0 + dt) h) / ((dt0 + dt) w)|, w = |h| floored at tsc_floor max |h|
__global__ void k_cz_absmax(int n, const double *v, unsigned long long *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = i < n ? fabs(v[i]) : 0.0;
    for (int o = 32; o > 0; o >>= 1) e = fmax(e, __shfl_xor(e, o));
    if ((threadIdx.x & 63) == 0 && e > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(e));
}
__global__ void k_cz_trunc(int n, const double *hm, const double *ho, const double *hp, double dt0, double dt,
                           double floor_factor, const unsigned long long *hmaxbits, unsigned long long *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (i < n) {
        const double wfloor = floor_factor * __longlong_as_double((long long)*hmaxbits);
        double w = fabs(ho[i]);
        if (w < wfloor) w = wfloor;
        const double den = (dt0 + dt) * w;
        e = (fabs(den) <= DBL_MIN) ? 0.0 : fabs(2.0 * (dt / dt0) * (dt0 * hp[i] + dt * hm[i] - (dt0 + dt) * ho[i]) / den);
    }
    for (int o = 32; o > 0; o >>= 1) e = fmax(e, __shfl_xor(e, o));
    if ((threadIdx.x & 63) == 0 && e > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(e));
}
int cz_criterion(yz_cz *z, double *crit) {
    const yz_cz_case &c = z->cs;
    double th = c.tsc_criterion;
    if (fabs(th) <= DBL_MIN) th = 1.0e-2;
    if (fabs(z->dt_old) <= DBL_MIN) { *crit = 1.0; return 0; }
    const int n = (int)z->g.n2;
    CZ_CHECK(hipMemsetAsync(z->emax, 0, 2 * sizeof(unsigned long long), z->st));
    hipLaunchKernelGGL(k_cz_absmax, dim3(nblk(n)), dim3(TBK), 0, z->st, n, (const double *)z->hO, z->emax);
    hipLaunchKernelGGL(k_cz_trunc, dim3(nblk(n)), dim3(TBK), 0, z->st, n, (const double *)z->hO2,
                       (const double *)z->hO, (const double *)z->hN, z->dt_old, z->dt, c.tsc_floor,
                       (const unsigned long long *)z->emax, z->emax + 1);
    unsigned long long eb[2];
    CZ_CHECK(hipMemcpyAsync(eb, z->emax, sizeof(eb), hipMemcpyDeviceToHost, z->st));
    CZ_CHECK(hipStreamSynchronize(z->st));
    double e;
    std::memcpy(&e, &eb[1], sizeof(e));
    *crit = 1.0 - e / th;
    return 0;
}

// one trial of the step (TimeAdvance, Solve, AfterTimeStepTrial); with
// crit, the film's truncation criterion after its solve, and a film it
// rejects is not carried on (4, the step's start kept)
int cz_advance(yz_cz *z, int &lin, int &passes_out, double &resid_out, double *crit) {
    const Geo3 &g = z->g;
    const size_t n2 = g.n2, n3 = g.n3;
    lin = 0; passes_out = 0; resid_out = 0;
    auto restore = [&]() -> int {
        if (cp_d2d(z, z->hO, z->hO_keep, n2) || cp_d2d(z, z->hN, z->hN_keep, n2) || cp_d2d(z, z->fH, z->tH, n2) ||
            cp_d2d(z, z->fC, z->tC, n3))
            return -1;
        return sync(z);
    };
    if (cp_d2d(z, z->hO_keep, z->hO, n2) || cp_d2d(z, z->hN_keep, z->hN, n2)) return -1;
    if (cp_d2d(z, z->hO, z->hN, n2) || cp_d2d(z, z->fH, z->tH, n2) || cp_d2d(z, z->fC, z->tC, n3)) return -1;
    int iter = 0;
    double resid = 0;
    int ch = h_stage(z, iter, resid);
    if (ch < 0) return -1;
    if (ch != 0) { if (restore()) return -1; return 1; }
    hipLaunchKernelGGL(k_hmid, dim3(nblk(g.n2)), dim3(TBK), 0, z->st, z->a, z->fH, 0);
    const bool extrap = z->prevSet;
    if (extrap) {
        const double rr = 0.5 * z->dt / z->dtPrev;
        hipLaunchKernelGGL(k_extrap, dim3(nblk(g.n3)), dim3(TBK), 0, z->st, g.n3, rr, (const double *)z->cN,
                           (const double *)z->cPrev, z->cm);
        if (cp_d2d(z, z->fC, z->cm, n3)) return -1;
    }
    ch = h_stage(z, iter, resid);
    if (ch < 0) return -1;
    if (cp_d2d(z, z->fH, z->hO, n2)) return -1;
    if (extrap && cp_d2d(z, z->fC, z->cN, n3)) return -1;
    if (ch != 0) { if (restore()) return -1; return 1; }
    if (crit) {
        if (cz_criterion(z, crit)) return -1;
        if (!(*crit > 0.0)) { lin = iter; if (restore()) return -1; return 4; }
    }

    conservative_update(z);
    if (cp_d2d(z, z->fH, z->hN, n2)) return -1;
    if (cp_d2d(z, z->c0, z->cN, n3)) return -1;
    if (!extrap && cp_d2d(z, z->cm, z->c0, n3)) return -1;
    if (cp_d2d(z, z->keepO, z->cO, n3) || cp_d2d(z, z->keepN, z->cN, n3) || cp_d2d(z, z->prevKeep, z->cPrev, n3))
        return -1;
    const double dtPrevKeep = z->dtPrev;
    const bool prevSetKeep = z->prevSet;
    if (cp_d2d(z, z->cPrev, z->c0, n3)) return -1;
    z->dtPrev = z->dt;
    z->prevSet = true;
    const int maxPass = z->cs.marangoni ? 1 + 8 : 1;
    const double floor = std::max(1e-10, 10.0 * z->cs.c_eps);
    double e0 = 0, e = 0, omega = 1.0;
    int passes = 0, fail = 0, pass;
    for (pass = 0; pass < maxPass; pass++) {
        hipLaunchKernelGGL(k_hmid, dim3(nblk(g.n2)), dim3(TBK), 0, z->st, z->a, z->mid, 0);
        if (pass > 0) {
            if (cp_d2d(z, z->fH, z->mid, n2) || cp_d2d(z, z->fC, z->cm, n3)) return -1;
            int it2 = 0;
            double r2 = 0;
            ch = h_stage(z, it2, r2);
            if (ch < 0) return -1;
            iter += it2;
            resid = r2;
            if (cp_d2d(z, z->fC, z->c0, n3)) return -1;
            if (ch != 0) { fail = 1; break; }
            conservative_update(z);
        }
        hipLaunchKernelGGL(k_hmid, dim3(nblk(g.n2)), dim3(TBK), 0, z->st, z->a, z->mid, 1);
        if (cp_d2d(z, z->fH, z->mid, n2) || cp_d2d(z, z->fC, z->cm, n3)) return -1;
        volume_fluxes(z);
        if (cp_d2d(z, z->fC, z->c0, n3)) return -1;
        int cit = 0;
        double crs = 0;
        int cch;
        if (pass > 0) cch = c_solve_at_mid(z, z->cm, cit, crs);
        else if (maxPass > 1) {
            if (cp_d2d(z, z->cO, z->cN, n3)) return -1;
            cch = c_solve_at_mid(z, z->cm, cit, crs);
        } else {
            if (cp_d2d(z, z->cO, z->cN, n3)) return -1;
            cch = c_solve(z, cit, crs);
        }
        if (cch < 0) return -1;
        iter += cit;
        if (cch != 0) { fail = 2; break; }
        if (maxPass == 1) break;
        passes = pass;
        CZ_CHECK(hipMemsetAsync(z->emax, 0, sizeof(unsigned long long), z->st));
        hipLaunchKernelGGL(k_resid, dim3(nblk(g.n3)), dim3(TBK), 0, z->st, g.n3, (const double *)z->c0,
                           (const double *)z->cN, (const double *)z->cm, z->r, z->emax);
        unsigned long long eb = 0;
        CZ_CHECK(hipMemcpyAsync(&eb, z->emax, sizeof(eb), hipMemcpyDeviceToHost, z->st));
        if (sync(z)) return -1;
        std::memcpy(&e, &eb, sizeof(e));
        if (pass == 0) e0 = e;
        const bool settled = !(e > 0.1 * e0) || e <= floor;
        if (pass > 0 && !settled) {
            hipLaunchKernelGGL(k_aitken, dim3(RB), dim3(TBK), 0, z->st, g.n3, (const double *)z->r,
                               (const double *)z->rp, z->part);
            double sc[2];
            if (read_sums(z, 2, sc)) return -1;
            if (sc[1] > 0) omega = -omega * sc[0] / sc[1];
        }
        if (settled) break;
        hipLaunchKernelGGL(k_relax, dim3(nblk(g.n3)), dim3(TBK), 0, z->st, g.n3, omega, (const double *)z->r, z->cm);
        if (cp_d2d(z, z->rp, z->r, n3)) return -1;
    }
    if (!fail && maxPass > 1 && pass == maxPass) fail = 3;
    if (cp_d2d(z, z->fH, z->hN, n2)) return -1;
    lin = iter;
    resid_out = resid;
    if (fail) {
        if (cp_d2d(z, z->cO, z->keepO, n3) || cp_d2d(z, z->cN, z->keepN, n3) || cp_d2d(z, z->cPrev, z->prevKeep, n3))
            return -1;
        z->dtPrev = dtPrevKeep;
        z->prevSet = prevSetKeep;
        if (restore()) return -1;
        return fail;
    }
    z->stats[4] += passes;
    passes_out = passes;
    // RedistributeExcess, RedistributeDeficit
    const int nc = g.nx * g.ny;
    hipLaunchKernelGGL(k_excess, dim3(col_blocks(nc)), dim3(64 * kColPerBlock), 0, z->st, g, z->cN, z->colv[3]);
    int err = 0;
    z->stats[5] += colsum(z, z->colv[3], &err);
    if (err) return -1;
    double mv = 0, ad = 0;
    if (redistribute_deficit(z, mv, ad)) return -1;
    z->stats[6] += mv;
    z->stats[7] += ad;
    // the ghosts of the end state from its own conditions (RefreshGhosts)
    if (cp_d2d(z, z->fC, z->cN, n3)) return -1;
    hipLaunchKernelGGL(k_cbc, dim3(nblk(nc)), dim3(TBK), 0, z->st, z->a, 1);
    apply_rows(z, z->cN);
    if (cp_d2d(z, z->fC, z->cN, n3)) return -1;
    if (cp_d2d(z, z->tH, z->fH, n2) || cp_d2d(z, z->tC, z->fC, n3)) return -1;
    z->time += z->dt;
    ++z->steps;
    return sync(z);
}

}  // namespace

extern "C" {

void yz_cz_case_defaults(yz_cz_case *c) {
    std::memset(c, 0, sizeof(*c));
    c->nx = c->ny = 16; c->nz = 8;
    c->xmin = 0; c->xmax = 1; c->ymin = 0; c->ymax = 1;
    c->capillary = 1;
    c->As = 1; c->sig_c0 = 0.5; c->visc_c0 = 0.5; c->Df = 1; c->diff_c0 = 0.5;
    c->Ev = 1; c->evap_coef = 1; c->evap_func = YZ_EVAP_RAOULT; c->evap_power = 0.5;
    c->molar_resin = 2200; c->molar_solvent = 106;
    c->conc_scheme = YZ_CONV_UTOPIA;
    c->lin_maxit = 5000; c->h_eps = 1e-13; c->c_eps = 1e-13;
    c->film_precond = YZ_PC_AMG;
    c->learn_passes = 3; c->learn_iters = 150; c->learn_probes = 4; c->learn_lr = 2e-3;
    c->step_control = 0; c->tsc_change = 0.5; c->tsc_criterion = 0.01; c->tsc_floor = 0; c->tsc_keep = 50;
    c->dt_min = 1e-15; c->dt_max = 1e30;
}

int yz_sizeof_cz_case(void) { return (int)sizeof(yz_cz_case); }

yz_cz *yz_cz_create(const yz_cz_case *c, const double *h, const double *s, const double *conc, double dt) {
    if (!c || !h || !s || !conc) { yz::set_error("ConcZ: null argument"); return nullptr; }
    if (c->nx < 4 || c->ny < 4 || c->nz < 2) {
        yz::set_error("ConcZ: the film is 2-D here (nx, ny >= 4, nz >= 2); a 1-D film (Ny = 1) is not taken");
        return nullptr;
    }
    for (int k = 0; k < 4; ++k)
        if (c->bc[k] != YZ_BC_SYMMETRY && c->bc[k] != YZ_BC_PERIODIC) {
            yz::set_error("ConcZ: only Symmetry and Periodic edges are taken");
            return nullptr;
        }
    if ((c->bc[0] == YZ_BC_PERIODIC) != (c->bc[1] == YZ_BC_PERIODIC) ||
        (c->bc[2] == YZ_BC_PERIODIC) != (c->bc[3] == YZ_BC_PERIODIC)) {
        yz::set_error("ConcZ: a periodic edge needs its opposite edge periodic");
        return nullptr;
    }
    if (c->film_precond != YZ_PC_AMG && c->film_precond != YZ_PC_LEARNED) {
        yz::set_error("ConcZ: film_precond is YZ_PC_AMG or YZ_PC_LEARNED");
        return nullptr;
    }
    if (c->step_control && (!(c->tsc_change > 0 && c->tsc_change < 1) || c->tsc_keep < 1 || !(c->dt_min > 0) ||
                            !(c->dt_max >= dt))) {
        yz::set_error("ConcZ: step control needs 0 < tsc_change < 1, tsc_keep >= 1, dt_min > 0, dt_max >= dt");
        return nullptr;
    }
    if (c->conc_scheme < YZ_CONV_UPWIND1 || c->conc_scheme > YZ_CONV_VANLEER) {
        yz::set_error("ConcZ: unknown concentration scheme");
        return nullptr;
    }
    if (!(dt > 0) || !(c->xmax > c->xmin) || !(c->ymax > c->ymin) || c->lin_maxit < 1) {
        yz::set_error("ConcZ: bad dt, extent or iteration cap");
        return nullptr;
    }
    yz_cz *z = new yz_cz();
    z->cs = *c;
    z->dt = dt;
    z->dt_old = dt;   // DT().old starts at the initial step
    Geo3 &g = z->g;
    g.nx = c->nx; g.ny = c->ny; g.nz = c->nz;
    g.Nx = c->nx + 4; g.Ny = c->ny + 4; g.Nz = c->nz + 4;
    g.n2 = g.Nx * g.Ny; g.n3 = g.n2 * g.Nz; g.nint = g.nx * g.ny * g.nz;
    // the grid with its ghosts (Grids.cpp:155-184)
    const double dx0 = (c->xmax - c->xmin) / (double)c->nx, dy0 = (c->ymax - c->ymin) / (double)c->ny;
    g.xlo = c->xmin - 2 * dx0; g.ylo = c->ymin - 2 * dy0;
    g.dx = ((c->xmax + 2 * dx0) - g.xlo) / (double)g.Nx;
    g.dy = ((c->ymax + 2 * dy0) - g.ylo) / (double)g.Ny;
    const double dz0 = (1.0 - 0.0) / (double)c->nz;
    g.dz = ((1.0 + 2 * dz0) - (0.0 - 2 * dz0)) / (double)g.Nz;
    for (int k = 0; k < 4; ++k) g.wall[k] = c->bc[k] != YZ_BC_PERIODIC;
    g.perx = !g.wall[0]; g.pery = !g.wall[2];
    {
        const int cox[21] = YZ_S21X, coy[21] = YZ_S21Y, coz[21] = YZ_S21Z;
        for (int t = 0; t < 21; ++t) g.off[t] = (coz[t] * g.ny + coy[t]) * g.nx + cox[t];
    }
    bool ok = hipStreamCreateWithFlags(&z->st, hipStreamNonBlocking) == hipSuccess;
    const size_t n2 = g.n2, n3 = g.n3, ni = g.nint;
    double **a2[] = {&z->S, &z->hO, &z->hN, &z->fH, &z->tH, &z->mid, &z->hO_keep, &z->hN_keep, &z->hO2};
    for (double **p : a2) ok = ok && !dal(p, n2);
    for (int k = 0; k < 16; ++k) ok = ok && !dal(&z->f2[k], n2);
    double **a3[] = {&z->fC, &z->tC, &z->cO, &z->cN, &z->c0, &z->cm, &z->cPrev, &z->keepO, &z->keepN,
                     &z->prevKeep, &z->r, &z->rp, &z->cn, &z->d, &z->ct, &z->Qx, &z->Qy, &z->W, &z->cSettle};
    for (double **p : a3) ok = ok && !dal(p, n3);
    ok = ok && !dal(&z->zr, 4 * n2);
    ok = ok && !dal(&z->cA, 21 * ni) && !dal(&z->cP, 21 * ni) && !dal(&z->react, ni) && !dal(&z->cb, ni);
    for (int k = 0; k < 9; ++k) ok = ok && !dal(&z->kv[k], ni);
    for (int k = 0; k < 5; ++k) ok = ok && !dal(&z->lu[k], ni);
    for (int k = 0; k < 4; ++k) ok = ok && !dal(&z->colv[k], (size_t)g.nx * g.ny);
    ok = ok && !dal(&z->part, 4 * RB) && !dal(&z->dsc, 8) && !dal(&z->flags, 4) && !dal(&z->emax, 2);
    ok = ok && hipHostMalloc((void **)&z->hsc, 8 * sizeof(double), hipHostMallocDefault) == hipSuccess;
    ok = ok && hipHostMalloc((void **)&z->hflags, 4 * sizeof(int), hipHostMallocDefault) == hipSuccess;
    // the film's solver over the real cells
    StencilOp op;
    std::memset(&op, 0, sizeof(op));
    op.dim = 2; op.nx = g.nx; op.ny = g.ny; op.N = g.nx * g.ny; op.ns = 13;
    op.perx = g.perx; op.pery = g.pery;
    {
        const int ox[13] = {0, -1, 0, 1, -2, -1, 0, 1, 2, -1, 0, 1, 0};
        const int oy[13] = {-2, -1, -1, -1, 0, 0, 0, 0, 0, 1, 1, 1, 2};
        for (int k = 0; k < 13; ++k) { op.ox[k] = ox[k]; op.oy[k] = oy[k]; }
    }
    double *hb = nullptr, *hx = nullptr;
    ok = ok && solver_init(z->Sh, op, z->st) == 0;
    ok = ok && !dal(&hb, (size_t)op.N) && !dal(&hx, (size_t)op.N);
    if (ok && c->film_precond == YZ_PC_LEARNED) {
        ok = solver_init(z->S32, op, z->st) == 0 && !dal(&z->b32, (size_t)op.N) && !dal(&z->x32, (size_t)op.N) &&
             !dal(&z->feat, (size_t)8 * op.N) && !dal(&z->mscale, (size_t)op.N) &&
             !dal(&z->Yh, (size_t)13 * op.N) && !dal(&z->r64, (size_t)op.N) && !dal(&z->xkeep, (size_t)op.N) &&
             !dal(&z->nonunit, (size_t)1) &&
             hipHostMalloc((void **)&z->hnonunit, sizeof(int), hipHostMallocDefault) == hipSuccess &&
             mlp_init(z->mlp, op.N, z->st) == 0;
        z->S32.use_graph = 1;
        // a diverging FP32 solve stops once its residual has not halved over
        // 40 iterations
        z->S32.stall_window = 40;
        z->S32.stall_factor = 0.5;
    }
    if (!ok) {
        if (yz_last_error()[0] == 0) yz::set_error("ConcZ: device set-up failed");
        yz_cz_destroy(z);
        return nullptr;
    }
    ok = hipMemcpy(z->S, s, n2 * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(z->hN, h, n2 * sizeof(double), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(z->cN, conc, n3 * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(z->hO, z->hN, n2 * sizeof(double), hipMemcpyDeviceToDevice) == hipSuccess &&
         hipMemcpy(z->fH, z->hN, n2 * sizeof(double), hipMemcpyDeviceToDevice) == hipSuccess &&
         hipMemcpy(z->tH, z->hN, n2 * sizeof(double), hipMemcpyDeviceToDevice) == hipSuccess &&
         hipMemcpy(z->cO, z->cN, n3 * sizeof(double), hipMemcpyDeviceToDevice) == hipSuccess &&
         hipMemcpy(z->fC, z->cN, n3 * sizeof(double), hipMemcpyDeviceToDevice) == hipSuccess &&
         hipMemcpy(z->tC, z->cN, n3 * sizeof(double), hipMemcpyDeviceToDevice) == hipSuccess;
    if (!ok) { yz::set_error("ConcZ: upload failed"); yz_cz_destroy(z); return nullptr; }
    Args &a = z->a;
    std::memset(&a, 0, sizeof(a));
    a.g = g; a.cs = *c; a.dt = dt; a.theta = 0.5;
    a.lim = c->conc_scheme == YZ_CONV_KOREN || c->conc_scheme == YZ_CONV_VANLEER;
    a.lim_vl = c->conc_scheme == YZ_CONV_VANLEER;
    if (!a.lim) scheme_coefs(c->conc_scheme, a.w1, a.w2);
    a.S = z->S; a.hO = z->hO; a.hN = z->hN; a.fH = z->fH; a.tH = z->tH; a.fC = z->fC; a.tC = z->tC;
    a.Dx = z->f2[0]; a.Dy = z->f2[1]; a.sgx = z->f2[2]; a.sgy = z->f2[3]; a.Mx = z->f2[4]; a.My = z->f2[5];
    a.qCapX = z->f2[6]; a.qExpX = z->f2[7]; a.qCapY = z->f2[8]; a.qExpY = z->f2[9]; a.qX = z->f2[10];
    a.qY = z->f2[11]; a.colSrc = z->f2[12]; a.dtdh = z->f2[13];
    a.Qx = z->Qx; a.Qy = z->Qy; a.W = z->W; a.zr = z->zr;
    a.hP = z->Sh.A; a.hb = hb; a.hx = hx;
    a.cA = z->cA; a.react = z->react; a.cP = z->cP; a.cb = z->cb; a.flags = z->flags;
    return z;
}

int yz_cz_advance(yz_cz *z, int *lin_iter, int *passes, double *resid) {
    if (!z) { yz::set_error("ConcZ: null context"); return -1; }
    int l = 0, p = 0;
    double r = 0;
    int rc;
    if (!z->cs.step_control) {
        rc = cz_advance(z, l, p, r, nullptr);
    } else {
        // TimeStepTrial (SemiImplicit.cpp:100-148) with TimeStepAdjustment
        // (ImplicitSolver.cpp:144-190): the trial repeated at a shorter step
        // until one is accepted; a step accepted with the criterion above
        // 0.9 counts toward lengthening the next one
        const size_t n2 = z->g.n2;
        if (cp_d2d(z, z->hO2, z->hO, n2)) return -1;
        int lsum = 0;
        for (;;) {
            double crit = -1.0;
            rc = cz_advance(z, l, p, r, &crit);
            if (rc < 0) return -1;
            lsum += l;
            if (rc == 0 && crit > 0.0) {
                double dt_next = z->dt;
                if (crit > 0.90 && z->dt < z->cs.dt_max && ++z->tsc_count >= z->cs.tsc_keep) {
                    z->tsc_count = 0;
                    dt_next = std::min(z->dt / z->cs.tsc_change, z->cs.dt_max);
                }
                z->dt_old = z->dt;
                z->dt = dt_next;
                z->a.dt = dt_next;
                break;
            }
            ++z->cuts;
            z->dt *= z->cs.tsc_change;
            z->a.dt = z->dt;
            if (z->dt < z->cs.dt_min) { yz::set_error("ConcZ: the step fell below dt_min"); rc = 1; break; }
        }
        l = lsum;
    }
    if (lin_iter) *lin_iter = l;
    if (passes) *passes = p;
    if (resid) *resid = r;
    return rc;
}

void yz_cz_clock(yz_cz *z, double *out3) {
    if (!z || !out3) return;
    out3[0] = z->dt; out3[1] = z->dt_old; out3[2] = (double)z->cuts;
}

void yz_cz_get_h(yz_cz *z, double *h) {
    if (!z || !h) return;
    (void)hipStreamSynchronize(z->st);
    (void)hipMemcpy(h, z->hN, z->g.n2 * sizeof(double), hipMemcpyDeviceToHost);
}
void yz_cz_get_c(yz_cz *z, double *c) {
    if (!z || !c) return;
    (void)hipStreamSynchronize(z->st);
    (void)hipMemcpy(c, z->cN, z->g.n3 * sizeof(double), hipMemcpyDeviceToHost);
}
void yz_cz_volume_fluxes(yz_cz *z, double *qx, double *qy, double *w) {
    if (!z) return;
    (void)hipStreamSynchronize(z->st);
    const size_t b = z->g.n3 * sizeof(double);
    if (qx) (void)hipMemcpy(qx, z->Qx, b, hipMemcpyDeviceToHost);
    if (qy) (void)hipMemcpy(qy, z->Qy, b, hipMemcpyDeviceToHost);
    if (w) (void)hipMemcpy(w, z->W, b, hipMemcpyDeviceToHost);
}
void yz_cz_stats(yz_cz *z, double *out8) {
    if (!z || !out8) return;
    for (int k = 0; k < 8; ++k) out8[k] = z->stats[k];
}
void yz_cz_learn_stats(yz_cz *z, double *out4) {
    if (!z || !out4) return;
    for (int k = 0; k < 4; ++k) out4[k] = z->lstats[k];
}
// the restart state (ConcZ_Model_2D::WriteState / BuildRestart,
// ConcZ_Model_2D.cpp:572-600, 664-690): the clock, the step control's count
// (0: the step is fixed here), the film at the start of the last step, the
// concentration at the start of the last step and that step's dt (the
// corrector extrapolates c to the next midpoint from them), the backends'
// memory (none), the fields with their ghosts
int yz_cz_save_state(yz_cz *z, const char *path) {
    if (!z || !path) { yz::set_error("ConcZ: null argument"); return 2; }
    const size_t n2 = z->g.n2, n3 = z->g.n3;
    std::vector<double> h(n2), hold(n2), sv(n2), c(n3), cp;
    if (hipStreamSynchronize(z->st) != hipSuccess ||
        hipMemcpy(h.data(), z->hN, n2 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hold.data(), z->hO, n2 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(sv.data(), z->S, n2 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(c.data(), z->cN, n3 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        yz::set_error("ConcZ: download failed");
        return 2;
    }
    if (z->prevSet) {
        cp.resize(n3);
        if (hipMemcpy(cp.data(), z->cPrev, n3 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
            yz::set_error("ConcZ: download failed");
            return 2;
        }
    }
    StateFile st;
    st.set_text("model", "ConcZ-3D");
    st.set("time", z->time);
    st.set("step", (double)z->steps);
    // the clock (PutClock, RestartState.cpp:168-183): the step about to be
    // taken, the last accepted one, the one the last step was solved with
    st.set("dt", std::vector<double>{z->dt, z->dt_old, z->dt_old});
    st.set("adjustCount.height", (double)z->tsc_count);
    st.set_vector("height.old", hold);
    if (z->prevSet) {
        st.set("conc.previousDt", z->dtPrev);
        st.set_vector("conc.previous", cp);
    }
    st.set_vector("solver.height", std::vector<double>());
    st.set_vector("solver.conc", std::vector<double>());
    st.set_vector("field.h", h);
    st.set_vector("field.s", sv);
    st.set_vector("field.c", c);
    if (!st.write(path)) { yz::set_error(std::string("ConcZ: could not write ") + path); return 2; }
    return 0;
}

int yz_cz_load_state(yz_cz *z, const char *path) {
    if (!z || !path) { yz::set_error("ConcZ: null argument"); return 2; }
    const size_t n2 = z->g.n2, n3 = z->g.n3;
    StateFile st;
    if (!st.read(path)) { yz::set_error(std::string("ConcZ: not a readable state file: ") + path); return 2; }
    if (st.text("model") != "ConcZ-3D") { yz::set_error("ConcZ: the state file was written by " + st.text("model")); return 2; }
    if (st.vec("field.h").size() != n2 || st.vec("field.s").size() != n2 || st.vec("height.old").size() != n2 ||
        st.vec("field.c").size() != n3) {
        yz::set_error("ConcZ: state vectors missing or of another grid");
        return 2;
    }
    const bool prev = st.has("conc.previousDt") && st.vec("conc.previous").size() == n3;
    bool ok = hipStreamSynchronize(z->st) == hipSuccess;
    for (double *d : {z->hN, z->fH, z->tH})
        ok = ok && hipMemcpy(d, st.vec("field.h").data(), n2 * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(z->hO, st.vec("height.old").data(), n2 * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(z->S, st.vec("field.s").data(), n2 * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    for (double *d : {z->cN, z->cO, z->fC, z->tC})
        ok = ok && hipMemcpy(d, st.vec("field.c").data(), n3 * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    if (prev) ok = ok && hipMemcpy(z->cPrev, st.vec("conc.previous").data(), n3 * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) { yz::set_error("ConcZ: upload failed"); return 2; }
    z->prevSet = prev;
    z->dtPrev = prev ? st.get("conc.previousDt") : 1.0;
    z->time = st.get("time");
    z->steps = (long long)st.get("step");
    // under the step control the clock's steps and the lengthening count
    // continue from the file (TakeClock with restoreDt)
    if (z->cs.step_control) {
        z->dt = st.get("dt", 0, z->dt);
        z->dt_old = st.get("dt", 1, z->dt);
        z->a.dt = z->dt;
        z->tsc_count = (int)st.get("adjustCount.height");
    }
    return 0;
}

void yz_cz_destroy(yz_cz *z) {
    if (!z) return;
    if (z->st) (void)hipStreamSynchronize(z->st);
    solver_free(z->Sh);
    solver_free(z->S32);
    mlp_free(z->mlp);
    for (float *p : {z->b32, z->x32, z->feat, z->mscale}) if (p) (void)hipFree(p);
    if (z->Yh) (void)hipFree(z->Yh);
    for (double *p : {z->r64, z->xkeep}) if (p) (void)hipFree(p);
    if (z->nonunit) (void)hipFree(z->nonunit);
    if (z->hnonunit) (void)hipHostFree(z->hnonunit);
    double *ptrs[] = {z->S, z->hO, z->hN, z->fH, z->tH, z->mid, z->hO_keep, z->hN_keep, z->hO2, z->fC, z->tC, z->cO,
                      z->cN, z->c0, z->cm, z->cPrev, z->keepO, z->keepN, z->prevKeep, z->r, z->rp, z->cn, z->d,
                      z->ct, z->Qx, z->Qy, z->W, z->zr, z->cA, z->react, z->cP, z->cb, z->part, z->dsc, z->cSettle,
                      z->a.hb, z->a.hx};
    for (double *p : ptrs) if (p) (void)hipFree(p);
    for (double *p : z->f2) if (p) (void)hipFree(p);
    for (double *p : z->kv) if (p) (void)hipFree(p);
    for (double *p : z->lu) if (p) (void)hipFree(p);
    for (double *p : z->colv) if (p) (void)hipFree(p);
    if (z->flags) (void)hipFree(z->flags);
    if (z->emax) (void)hipFree(z->emax);
    if (z->hsc) (void)hipHostFree(z->hsc);
    if (z->hflags) (void)hipHostFree(z->hflags);
    if (z->st) (void)hipStreamDestroy(z->st);
    delete z;
}

}  // extern "C"
