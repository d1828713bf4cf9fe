// yz_velocity.h -- the film's face velocities (CalVelocity,
// HeightEquationBase.h:1583-1601), for output and diagnostics.
//
// One thread per cell 2 <= ix <= nxp-2 (2 <= iy <= nyp-2; InsideForVelocity,
// GridTraits.h:239-244, 515-522) forms the x velocity on its Xm face and the
// y velocity on its Ym face as the plain sum of the terms' velocities in
// Cal_MixedVelocity's order (LubFields.cpp:467-495): capillary
// (LaplacePressure_2D.cpp:137-211, LaplacePressure_1D.cpp:84-136),
// Marangoni (SolutoCapillarityKernel.h:182-210, 387-405), vertical gravity
// (VerticalGravityKernel.h:435-470), in-plane gravity and centrifugal force
// (the limited BodyForceKernel.h:694-720), disjoining pressure
// (DisjoiningPressureKernel.h:442-475). Two more kernels then write the
// boundary rows (SetVelocityBC, WM2D/HeightEquation_2D.cpp:32-151,
// WM1D/HeightEquation_1D.cpp:28-86): the side columns row by row, then the
// bottom and top rows column by column, in that order, as the reference
// does. The reference writes the boundary rows into each term's field,
// sums them and writes them once more. Every boundary row is a copy or a
// zero, so the sum of the terms' rows is the rows of the sum, to the bit,
// and the two passes over the sum give the reference's field. Two passes
// are not one: on a Periodic film the first pass fills x on the ghost row
// iy = 1 after the side columns have copied from it, so only the second
// gives the outer corner face (nxp-1, 1) its value.
#pragma once
#include "yz_rows.h"

namespace yz {

template <int DIM>
struct VelCell {
    static constexpr int NS = DIM == 2 ? 13 : 5;
    static constexpr int C0 = DIM == 2 ? 6 : 2;
    double H[NS], S[NS], Cc[NS];
    int ixp, iyp;
};

// the wall faces a cell's Xm / Ym face can be (GridTraits.h WallFace): the
// first real cell's, and the first ghost's on the far side
template <int DIM>
__device__ inline bool vel_wall(const Geo &g, const VelCell<DIM> &q, int a) {
    if (a == 0) {
        if (q.ixp == 2) return is_wall(g.bc[0]);
        if (q.ixp == g.nxp - 2) return is_wall(g.bc[1]);
        return false;
    }
    if (q.iyp == 2) return is_wall(g.bc[2]);
    if (q.iyp == g.nyp - 2) return is_wall(g.bc[3]);
    return false;
}

// AxisMask: no y velocity on a grid two real cells deep
__device__ inline double vel_mask(const Geo &g, int a, double v) { return (a == 1 && g.nyp == 6) ? 0.0 : v; }

// each term's velocity on the face (a = 0: Xm, 1: Ym) into t: capillary,
// Marangoni, vertical gravity, in-plane gravity, centrifugal, disjoining;
// a term the case leaves out is 0
template <int DIM>
__device__ void face_terms(const RowArgs &A, const VelCell<DIM> &q, int a, double t[6]) {
    constexpr int C0 = VelCell<DIM>::C0;
    const Geo &g = A.g;
    const Mob &mob = A.mob;
    const int d = 2 * a;
    const int as = adj_slot(DIM, d);
    const double hsp = a == 0 ? g.dx : g.dy;
    const double hface = 0.5 * (q.H[as] + q.H[C0]);
    const double cface = 0.5 * (q.Cc[C0] + q.Cc[as]);
    // CV(a): the difference across the face, -C(2a) times the spacing
    const double h2 = hsp * hsp;
    const double cadj = -(1.0 / h2) * hsp, ccen = -(-1.0 / h2) * hsp;
#pragma unroll
    for (int k = 0; k < 6; ++k) t[k] = 0.0;
    if (A.cap) {
        const double Hoo = q.H[C0], Haj = q.H[as];
        const double val = mob.slip ? mob.FaceDK(Hoo, Haj) : 2.0 * (Hoo * Hoo) * (Haj * Haj) / (Hoo + Haj);
        const double Df = val * sig_of(A, cface) / visc_of(A, cface);
        const double Mag = 1.0 / 3.0;
        double Qh, Qs;
        if (DIM == 1) {
            const double rdx3 = 1.0 / (g.dx * g.dx * g.dx);
            const double *h = q.H, *s = q.S;
            Qh = ((h[3] - 2.0 * h[2] + h[1]) - (h[2] - 2.0 * h[1] + h[0])) * rdx3;
            Qs = ((s[3] - 2.0 * s[2] + s[1]) - (s[2] - 2.0 * s[1] + s[0])) * rdx3;
        } else {
            const double dx = g.dx, dy = g.dy;
            const double *h = q.H, *s = q.S;
            if (a == 0) {
                const double rdx3 = 1.0 / (dx * dx * dx), r2 = (dx * dx) / (dy * dy);
                Qh = (((h[S_po] - 2.0 * h[S_oo] + h[S_mo]) + (h[S_op] - 2.0 * h[S_oo] + h[S_om]) * r2) -
                      ((h[S_oo] - 2.0 * h[S_mo] + h[S_Mo]) + (h[S_mp] - 2.0 * h[S_mo] + h[S_mm]) * r2)) * rdx3;
                Qs = (((s[S_po] - 2.0 * s[S_oo] + s[S_mo]) + (s[S_op] - 2.0 * s[S_oo] + s[S_om]) * r2) -
                      ((s[S_oo] - 2.0 * s[S_mo] + s[S_Mo]) + (s[S_mp] - 2.0 * s[S_mo] + s[S_mm]) * r2)) * rdx3;
            } else {
                const double rdy3 = 1.0 / (dy * dy * dy), r2 = (dy * dy) / (dx * dx);
                Qh = (((h[S_po] - 2.0 * h[S_oo] + h[S_mo]) * r2 + (h[S_op] - 2.0 * h[S_oo] + h[S_om])) -
                      ((h[S_pm] - 2.0 * h[S_om] + h[S_mm]) * r2 + (h[S_oo] - 2.0 * h[S_om] + h[S_oM]))) * rdy3;
                Qs = (((s[S_po] - 2.0 * s[S_oo] + s[S_mo]) * r2 + (s[S_op] - 2.0 * s[S_oo] + s[S_om])) -
                      ((s[S_pm] - 2.0 * s[S_om] + s[S_mm]) * r2 + (s[S_oo] - 2.0 * s[S_om] + s[S_oM]))) * rdy3;
            }
        }
        Qh *= Mag * Df;
        Qs *= Mag * Df;
        t[0] = (Qh + Qs) / hface;
    }
    if (A.ma) {
        double S = 0;
        if (!vel_wall<DIM>(g, q, a)) {
            const int lo = face_slot(DIM, d, 1), hi = face_slot(DIM, d, 2);
            const double dc = (q.Cc[hi] - q.Cc[lo]) / hsp;
            const double cc = (q.Cc[hi] + q.Cc[lo]) / 2.0;
            double f[4], wt[4], gd[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) f[k] = q.H[face_slot(DIM, d, k)];
            const double hf = face_interp(f, sgn(A.AsMa2) * sgn(dc), wt, gd, A.hs);
            const double mu = visc_of(A, cc);
            S = mob.slip ? dc / mu * (hf + 2.0 * mob.lambda) : dc / mu * hf;
        }
        t[1] = vel_mask(g, a, A.AsMa2 * S);
    }
    if (A.vg) {
        const double Gf = mob.Face(q.H[C0], q.H[as]) / visc_of(A, cface);
        double Q = 0;
        Q += cadj * (q.H[as] + q.S[as]);
        Q += ccen * (q.H[C0] + q.S[C0]);
        Q *= (-1.0 * A.Bo / 3.0) * Gf;
        t[2] = vel_mask(g, a, Q) / hface;
    }
#pragma unroll
    for (int mode = 0; mode < 2; ++mode) {
        if (mode == 0 && !A.pg) continue;
        if (mode == 1 && !A.cf) continue;
        double dv;
        if (mode == 0) dv = sgn(a == 0 ? A.Pgx : A.Pgy);
        else if (a == 0) dv = (g.xlo + q.ixp * g.dx) - A.rotx;
        else dv = (g.ylo + q.iyp * g.dy) - A.roty;
        const double w = mode == 0 ? 1.0 : dv;
        const double Mag = mode == 0 ? ((a == 0 ? A.Pgx : A.Pgy) / 3.0) : A.Sp / 3.0;
        double f[4], wt[4], gd[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) f[k] = q.H[face_slot(DIM, d, k)];
        const double hf = face_interp(f, dv, wt, gd, A.hs);
        double D = (mob.slip ? mob.DperH(hf) : hf * hf) / visc_of(A, cface);
        if (vel_wall<DIM>(g, q, a)) D = 0;
        const double Q = -w * Mag * D;
        t[3 + mode] = -1.0 * vel_mask(g, a, Q);
    }
    if (A.dj) {
        const double Gf = mob.Face(q.H[C0], q.H[as]) / visc_of(A, cface);
        auto Pi = [&](double hh) {
            if (hh <= 0) return 0.0;
            const double r = A.hstar / hh;
            return pow(r, A.pn) - pow(r, A.pm);
        };
        double Q = 0;
        Q += cadj * Pi(q.H[as]);
        Q += ccen * Pi(q.H[C0]);
        Q *= (A.Dp / 3.0) * Gf;
        t[5] = vel_mask(g, a, Q) / hface;
    }
}

// the mixed velocity: the plain sum of the terms the case has, in order
template <int DIM>
__device__ double face_velocity(const RowArgs &A, const VelCell<DIM> &q, int a) {
    double t[6];
    face_terms<DIM>(A, q, a, t);
    const bool on[6] = {A.cap != 0, A.ma != 0, A.vg != 0, A.pg != 0, A.cf != 0, A.dj != 0};
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k)
        if (on[k]) v += t[k];
    return v;
}

template <int DIM>
__global__ void k_velocity(RowArgs A, double *vx, double *vy) {
    constexpr int NS = VelCell<DIM>::NS;
    const Geo &g = A.g;
    const int nxv = g.nxp - 3, nyv = DIM == 2 ? g.nyp - 3 : 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nxv * nyv) return;
    VelCell<DIM> q;
    q.ixp = t % nxv + 2;
    q.iyp = DIM == 2 ? t / nxv + 2 : 0;
    const int p = q.iyp * g.nxp + q.ixp;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        // the points two cells on (Po, oP; 1-D P) leave the grid at the
        // last cells and no face velocity uses them
        const bool far = DIM == 2 ? (k == S_Po || k == S_oP) : (k == 4);
        if (far) { q.H[k] = q.S[k] = q.Cc[k] = 0.0; continue; }
        const int r = p + slot_ox(DIM, k) + slot_oy(DIM, k) * g.nxp;
        q.H[k] = A.h[r]; q.S[k] = A.s[r]; q.Cc[k] = A.c[r];
    }
    vx[p] = face_velocity<DIM>(A, q, 0);
    if (DIM == 2) vy[p] = face_velocity<DIM>(A, q, 1);
}

// the mixed field and the first five terms' own fields (the .vts's velocity
// arrays, LubFields.cpp:378-383, 840-845): field f of six at vx + f np,
// vy + f np; the mixed one as k_velocity writes it. A term's own value is
// stored as the term formed it (a term the case leaves out is 0)
template <int DIM>
__global__ void k_velocity_terms(RowArgs A, double *vx, double *vy, size_t np) {
    constexpr int NS = VelCell<DIM>::NS;
    const Geo &g = A.g;
    const int nxv = g.nxp - 3, nyv = DIM == 2 ? g.nyp - 3 : 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nxv * nyv) return;
    VelCell<DIM> q;
    q.ixp = t % nxv + 2;
    q.iyp = DIM == 2 ? t / nxv + 2 : 0;
    const int p = q.iyp * g.nxp + q.ixp;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const bool far = DIM == 2 ? (k == S_Po || k == S_oP) : (k == 4);
        if (far) { q.H[k] = q.S[k] = q.Cc[k] = 0.0; continue; }
        const int r = p + slot_ox(DIM, k) + slot_oy(DIM, k) * g.nxp;
        q.H[k] = A.h[r]; q.S[k] = A.s[r]; q.Cc[k] = A.c[r];
    }
    const bool on[6] = {A.cap != 0, A.ma != 0, A.vg != 0, A.pg != 0, A.cf != 0, A.dj != 0};
    for (int a = 0; a < DIM; ++a) {
        double tv[6];
        face_terms<DIM>(A, q, a, tv);
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k)
            if (on[k]) v += tv[k];
        double *out = a == 0 ? vx : vy;
        out[p] = v;
#pragma unroll
        for (int k = 0; k < 5; ++k) out[(k + 1) * np + p] = tv[k];
    }
}

// the side columns of the boundary rows, one thread per row 1 <= jy <= nyp-2
// (1-D: one thread)
__global__ void k_velocity_bc_x(Geo g, double *vx, double *vy) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int Nx = g.nxp;
    const bool px = g.bc[0] == YZ_BC_PERIODIC, qx = g.bc[1] == YZ_BC_PERIODIC;
    if (g.dim == 1) {
        if (t != 0) return;
        vx[0] = 0;
        if (px) vx[1] = vx[Nx - 3]; else { vx[1] = 0; vx[2] = 0; }
        if (qx) vx[Nx - 1] = vx[3]; else { vx[Nx - 2] = 0; vx[Nx - 1] = 0; }
        return;
    }
    const int jy = t + 1;
    if (jy > g.nyp - 2) return;
    double *X = vx + (size_t)jy * Nx, *Y = vy + (size_t)jy * Nx;
    X[0] = 0;
    if (px) { Y[1] = Y[Nx - 3]; X[1] = X[Nx - 3]; }
    else { Y[1] = Y[2]; X[1] = 0; X[2] = 0; }
    if (qx) { Y[Nx - 2] = Y[2]; X[Nx - 1] = X[3]; }
    else { Y[Nx - 2] = Y[Nx - 3]; X[Nx - 2] = 0; X[Nx - 1] = 0; }
}

// then the bottom and top rows, one thread per column 1 <= ix <= nxp-2
__global__ void k_velocity_bc_y(Geo g, double *vx, double *vy) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (ix > g.nxp - 2) return;
    const int Nx = g.nxp, Ny = g.nyp;
    const bool py = g.bc[2] == YZ_BC_PERIODIC, qy = g.bc[3] == YZ_BC_PERIODIC;
    auto X = [&](int iy) -> double & { return vx[(size_t)iy * Nx + ix]; };
    auto Y = [&](int iy) -> double & { return vy[(size_t)iy * Nx + ix]; };
    Y(0) = 0;
    if (py) { X(1) = X(Ny - 3); Y(1) = Y(Ny - 3); }
    else { X(1) = X(2); Y(1) = 0; Y(2) = 0; }
    if (qy) { X(Ny - 2) = X(2); Y(Ny - 1) = Y(3); }
    else { X(Ny - 2) = X(Ny - 3); Y(Ny - 2) = 0; Y(Ny - 1) = 0; }
}

}  // namespace yz
