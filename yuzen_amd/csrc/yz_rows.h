// yz_rows.h -- the fused row kernel of a Newton pass.
//
// One thread per interior cell. From the stencil of the film h, the
// substrate s and the frozen concentration c it forms, in registers,
//   * A(h)h as the reference's ApplyA forms it (HeightEquationBase.h:
//     803-827): the capillary and vertical-gravity terms face by face
//     (LaplacePressureKernel.h:439-477, VerticalGravityKernel.h:233-269),
//     the other terms as the product of their entries of A with h;
//   * the row of the Jacobian J = A + sum of the terms' SetJacobiMatrix
//     (HeightEquationBase.h:1061-1091) and of IA = I + theta dt J
//     (NewtonKantorovich.cpp:488-510);
//   * the Newton right-hand side Rhs = h_old - h - dt (theta A(h)h +
//     (1-theta) A(h_old)h_old) - dt B (NewtonKantorovich.cpp:512-523), B
//     the evaporation and the nozzle's supply when the case has them;
// and writes the row-scaled IA planes and right-hand side
// (LinearSolverBackend.cpp:227-252, ScaleRows: divided by the diagonal).
// A and J are never stored: the pass reads h, s, c once and writes the
// 13 planes of IA and one vector.
#pragma once
#include "yz_physics.h"

namespace yz {

// evaporation in B: the rate law (Evaporation.cpp:62-103, uniform in
// space) and its caps (EvapAt, HeightEquationBase.h:830-846)
struct BTerm {
    int evap, efunc;
    double Ev, coef, power, concr, mres, msol, prec;
};

// Ev E(h, c), at most half the film and at most its solvent per step
__device__ inline double evap_rate(const BTerm &b, double hh, double cc, double dt) {
    double evp = b.Ev * 0.0;
    if (b.evap) {
        double f_conc;
        if (b.efunc == YZ_EVAP_POWERN) {
            const double v = fmax(1.0 - cc - b.concr, 0.0);
            f_conc = pow(v, b.power);
        } else {
            // Raoult: 1 - the resin's mole fraction of the clamped c
            const double x = cc < 0.0 ? 0.0 : (cc > 1.0 ? 1.0 : cc);
            const double chi = (x / b.mres) / (x / b.mres + (1.0 - x) / b.msol);
            f_conc = 1.0 - chi;
        }
        double f_mod = 1.0;
        if (b.prec != 0.0) {
            // the precursor ramp: 0 at h*, 1 from 20 h*
            const double v = (hh - b.prec) / ((20.0 - 1.0) * b.prec);
            f_mod = fmax(0.0, fmin(v, 1.0));
        }
        evp = b.Ev * (b.coef * f_conc * 1.0 * f_mod);
    }
    if (evp <= 0) return evp;
    if (evp * dt > 0.5 * hh) evp = hh > 0 ? 0.5 * hh / dt : 0.0;
    const double x = cc < 0.0 ? 0.0 : (cc > 1.0 ? 1.0 : cc);
    const double solvent = hh > 0.0 ? hh * (1.0 - x) : 0.0;
    if (evp * dt > solvent) evp = solvent / dt;
    return evp;
}

struct RowArgs {
    Geo g;
    int cap, vg, pg, cf, ma, dj;
    double Bo, Pgx, Pgy, Sp, AsMa2, Dp, hstar, pn, pm, rotx, roty;
    double visc_c0, visc_exp, sig_c0, Ma_eff;
    // 1/dx, 1/dy, 1/dx^2, 1/dy^2 (formed once on the host), and the
    // disjoining exponents n, m as whole numbers when they are (0: not)
    double inv_h[2], inv_h2[2];
    int pn_i, pm_i;
    Mob mob;
    double Ccap[4][13];
    double theta, dt;
    int row_scaling;
    HScheme hs;                // the upwinded terms' face scheme of the film height
    const double *h, *s, *c;   // padded
    const double *hold;        // padded, h at the start of the step
    const double *axold;       // N: A(h_old) h_old
    double *ax;                // mode 0: A(h)h (N)
    double *ia;                // mode 1: ns planes x N, row-scaled
    double *b;                 // mode 1: row-scaled Rhs (N)
    double *ia_raw;            // optional: unscaled IA planes
    double *rhs_raw;           // optional: unscaled Rhs (N)
    double *ax_raw;            // optional: A(h)h (N) in mode 1
    float *ia32;               // mode 1, FP32 inner solve: row-scaled IA planes (replaces ia)
    float *b32;                // mode 1, FP32 inner solve: row-scaled Rhs (replaces b)
    float *feat;               // optional: 8 x N features of the learned preconditioner
    float *mscale;             // optional: 1 / (1 + sum_k!=c |IA_k|) of the scaled row
    unsigned long long *rmax;  // mode 1: max |Rhs| (bits of a non-negative double)
    int *nonunit;              // FP32 planes: set when a row's stored diagonal is not exactly 1
    // the modified Newton method (JacobianUpdateInterval > 1): a pass that
    // keeps the last IA (keep_ia) writes only the right-hand side, scaled by
    // the row scale dinv that the last IA's pass stored (mode 1)
    int keep_ia;
    double *dinv;
    // optional, mode 1: B = (1 - theta) evold + theta EvapAt(h, c) - src
    // enters Rhs as - dt B (SetMatrixA, HeightEquationBase.h:765-787); J
    // carries no derivative of it
    BTerm bt;
    const double *evold;       // N: EvapAt(h_old, c) of the step's start
    const double *src;         // N: the nozzle's supply at the middle of the step
    double *evp_out;           // optional, N: the evaporation part of B of the pass
};

__device__ inline double visc_of(const RowArgs &a, double cc) { return exp(a.visc_exp * (cc - a.visc_c0)); }
// r^e: repeated products when e is a small whole number k (the usual
// disjoining exponents 3 and 2), pow otherwise
__device__ inline double pow_k(double r, double e, int k) {
    if (k <= 0) return pow(r, e);
    double p = r;
    for (int j = 1; j < k; ++j) p *= r;
    return p;
}
__device__ inline double sig_of(const RowArgs &a, double cc) { return 1.0 + a.Ma_eff * (cc - a.sig_c0); }

// the third difference of v across face d (LaplacePressure_2D.cpp:105-128)
template <int DIM>
__device__ inline double face_grad(const RowArgs &A, const double *v, int d) {
    if (DIM == 2) {
        const double dx2 = A.g.dx * A.g.dx, dy2 = A.g.dy * A.g.dy;
        const double dn4 = (d < 2) ? dx2 * dx2 : dy2 * dy2, dxy = dx2 * dy2;
        double a1, a, b, b1, am, ap, bm, bp;
        switch (d) {
        case 0: a1 = v[S_Mo]; a = v[S_mo]; b = v[S_oo]; b1 = v[S_po]; am = v[S_mm]; ap = v[S_mp]; bm = v[S_om]; bp = v[S_op]; break;
        case 1: a1 = v[S_mo]; a = v[S_oo]; b = v[S_po]; b1 = v[S_Po]; am = v[S_om]; ap = v[S_op]; bm = v[S_pm]; bp = v[S_pp]; break;
        case 2: a1 = v[S_oM]; a = v[S_om]; b = v[S_oo]; b1 = v[S_op]; am = v[S_mm]; ap = v[S_pm]; bm = v[S_mo]; bp = v[S_po]; break;
        default: a1 = v[S_om]; a = v[S_oo]; b = v[S_op]; b1 = v[S_oP]; am = v[S_mo]; ap = v[S_po]; bm = v[S_mp]; bp = v[S_pp]; break;
        }
        const double normal = (((b1 - b) - 2.0 * (b - a)) + (a - a1)) / dn4;
        const double cross = (((bp - b) - (b - bm)) - ((ap - a) - (a - am))) / dxy;
        return normal + cross;
    }
    const double dx2 = A.g.dx * A.g.dx, dx4 = dx2 * dx2;
    double a1, a, b, b1;
    if (d == 0) { a1 = v[0]; a = v[1]; b = v[2]; b1 = v[3]; }
    else { a1 = v[1]; a = v[2]; b = v[3]; b1 = v[4]; }
    return (((b1 - b) - 2.0 * (b - a)) + (a - a1)) / dx4;
}

__device__ inline bool is_wall(int t) { return t == YZ_BC_SYMMETRY || t == YZ_BC_PINNING; }

// What the row of a cell takes from one of its faces, oriented for that
// cell (c: the cell's own film, a: the neighbour's across the face)
struct FaceQ {
    double mu;               // viscosity at the face's mean concentration
    double D, gC, gA;        // capillary: M sigma / mu and its derivatives in h_c, h_a (mode 1)
    double G, T, gCf, gAf;   // potential terms: M / mu, M / (h_c + h_a) / mu, dM/dh_c / mu, dM/dh_a / mu
};

// The quantities of a face from the films hc, ha and concentrations cc, ca
// of its two cells (Fm, Fm1, Fm2: the slip mobility and its derivatives in
// hc and ha, slip only). Every expression is symmetric in the two cells:
// with hc and ha (cc and ca, Fm1 and Fm2) exchanged the values are the same
// to the bit and gC, gA (gCf, gAf) trade places, so a face can be formed
// once for both of its cells.
template <int MODE>
__device__ inline void face_q(const RowArgs &A, double hc, double ha, double cc, double ca, double Fm, double Fm1,
                              double Fm2, FaceQ &q) {
    const Mob &mob = A.mob;
    q.D = q.gC = q.gA = q.G = q.T = q.gCf = q.gAf = 0;
    const double cf = 0.5 * (cc + ca);
    q.mu = visc_of(A, cf);
    if (A.cap) {
        const double Hoo = hc, Haj = ha;
        const double so = sig_of(A, cf) / q.mu;
        const double val = mob.slip ? Fm : 2.0 * (Hoo * Hoo) * (Haj * Haj) / (Hoo + Haj);
        q.D = val * so;
        if (MODE == 1) {
            const double den = (Hoo + Haj) * (Hoo + Haj);
            const double vc = mob.slip ? Fm1 : 2.0 * (Hoo * Haj * Haj) * (Hoo + Haj + Haj) / den;
            const double va = mob.slip ? Fm2 : 2.0 * (Haj * Hoo * Hoo) * (Haj + Hoo + Hoo) / den;
            q.gC = vc * so;
            q.gA = va * so;
        }
    }
    if (A.vg || A.dj) {
        const double ho = hc, mu = q.mu;
        const double gf = mob.slip ? Fm : ((ho > 0 && ha > 0) ? 2.0 * (ho * ho) * (ha * ha) / (ho + ha) : 0.0);
        q.G = gf / mu;
        q.T = (gf == 0) ? 0.0 : gf / (ho + ha) / mu;
        if (MODE == 1) {
            if (mob.slip) { q.gCf = Fm1 / mu; q.gAf = Fm2 / mu; }
            else if (ho > 0 && ha > 0) {
                const double den = (ho + ha) * (ho + ha);
                q.gCf = 2.0 * (ho * ha * ha) * (ho + ha + ha) / den / mu;
                q.gAf = 2.0 * (ha * ho * ho) * (ha + ho + ho) / den / mu;
            }
        }
    }
}

// the disjoining pressure at a point: Pi = (h*/h)^n - (h*/h)^m, dPi/dh and
// Pi / h (0 where the film is not positive)
__device__ inline void pi_point(const RowArgs &A, double hh, double &pis, double &dpis, double &pih) {
    pis = dpis = pih = 0;
    if (hh > 0) {
        const double ih = 1.0 / hh;
        const double r = A.hstar * ih;
        const double rn = pow_k(r, A.pn, A.pn_i), rm = pow_k(r, A.pm, A.pm_i);
        pis = rn - rm;
        dpis = -(A.pn * rn - A.pm * rm) * ih;
        pih = pis * ih;
    }
}

// The row of cell i (padded position p, padded coordinates ixp, iyp) from
// its stencil H, S, Cc, its faces fq and the disjoining values of its
// centre and neighbours: A(h)h (mode 0), or the residual, IA and what goes
// with them (mode 1; rabs: |Rhs|)
template <int DIM, int MODE>
__device__ __forceinline__ void row_body(const RowArgs &A, int i, int ixp, int iyp, int p, const double (&H)[DIM == 2 ? 13 : 5],
                                         const double (&S)[DIM == 2 ? 13 : 5], const double (&Cc)[DIM == 2 ? 13 : 5],
                                         const FaceQ (&fq)[DIM == 2 ? 4 : 2], const double (&PiS)[DIM == 2 ? 13 : 5],
                                         const double (&dPiS)[DIM == 2 ? 13 : 5], const double (&PiH)[DIM == 2 ? 13 : 5],
                                         double &rabs) {
    constexpr int NS = DIM == 2 ? 13 : 5;
    constexpr int ND = DIM == 2 ? 4 : 2;
    constexpr int C0 = DIM == 2 ? 6 : 2;
    const Geo &g = A.g;
    {
        double Ar[NS], Jc[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) { Ar[k] = 0; Jc[k] = 0; }
        double ax = 0;
        const Mob &mob = A.mob;
        auto spacing = [&](int axis) { return axis == 0 ? g.dx : g.dy; };
        auto wall = [&](int d) {
            if (d == 0) return ixp == 2 && is_wall(g.bc[0]);
            if (d == 1) return ixp == g.nxp - 3 && is_wall(g.bc[1]);
            if (d == 2) return iyp == 2 && is_wall(g.bc[2]);
            return iyp == g.nyp - 3 && is_wall(g.bc[3]);
        };

        // ---- capillary pressure (conservative; coefficient 1/3) ----
        if (A.cap) {
            const double Coef = 1.0 / 3.0;
            const double Hoo = H[C0];
            double D[ND], gA[ND], gC[ND];
#pragma unroll
            for (int d = 0; d < ND; ++d) { D[d] = fq[d].D; gC[d] = fq[d].gC; gA[d] = fq[d].gA; }
            // residual, face by face
            double r = 0;
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                const double f = Coef * (D[d] * (face_grad<DIM>(A, H, d) + face_grad<DIM>(A, S, d)));
                r += (d % 2 == 0) ? -f : f;
            }
            ax += r;
            if (MODE == 1) {
                double topo = 0;
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    double cc = 0;
#pragma unroll
                    for (int d = 0; d < ND; ++d) cc += A.Ccap[d][k] * (D[d] / Hoo);
                    topo += cc * S[k];
                }
#pragma unroll
                for (int k = 0; k < NS; ++k) {
                    double v = 0;
#pragma unroll
                    for (int d = 0; d < ND; ++d) v += A.Ccap[d][k] * D[d];
                    if (k == C0) v += topo;
                    Jc[k] += Coef * v;
                }
                double av[ND], bv[ND];
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    av[d] = bv[d] = 0;
#pragma unroll
                    for (int k = 0; k < NS; ++k) { av[d] += A.Ccap[d][k] * H[k]; bv[d] += A.Ccap[d][k] * S[k]; }
                }
                double ga = 0, sb = 0;
#pragma unroll
                for (int k = 0; k < ND; ++k) { const int dd = sdir(ND, k); ga += gC[dd] * av[dd]; }
#pragma unroll
                for (int d = 0; d < ND; ++d) sb += (gC[d] / Hoo - D[d] / (Hoo * Hoo)) * bv[d];
                Jc[C0] += Coef * (ga + sb * Hoo);
#pragma unroll
                for (int d = 0; d < ND; ++d)
                    Jc[adj_slot(DIM, d)] += Coef * (gA[d] * av[d] + (gA[d] / Hoo) * bv[d] * Hoo);
            }
        }

        // face mobility of the potential terms (vertical gravity, disjoining)
        double Gf[ND], Tk[ND], gAf[ND], gCf[ND], inv2[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            Gf[d] = fq[d].G; Tk[d] = fq[d].T; gCf[d] = fq[d].gCf; gAf[d] = fq[d].gAf;
            inv2[d] = A.inv_h2[d / 2];
        }

        // ---- vertical gravity (conservative; -Bo/3) ----
        if (A.vg) {
            const double Coef = -1.0 * A.Bo / 3.0;
            double r = 0;
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                const int a = adj_slot(DIM, d);
                const int lo = (d % 2 == 0) ? a : C0, hi = (d % 2 == 0) ? C0 : a;
                const double dn = spacing(d / 2);
                const double f = Coef * ((Gf[d] * (H[hi] - H[lo]) + (S[hi] - S[lo]) * (Tk[d] * (H[lo] + H[hi]))) / (dn * dn));
                r += (d % 2 == 0) ? -f : f;
            }
            ax += r;
            if (MODE == 1) {
                double Tsum[ND];
#pragma unroll
                for (int d = 0; d < ND; ++d) Tsum[d] = inv2[d] * S[adj_slot(DIM, d)] + (-inv2[d]) * S[C0];
                double Too = 0, cen = 0;
#pragma unroll
                for (int k = 0; k < ND; ++k) { const int dd = sdir(ND, k); Too += Tsum[dd] * Tk[dd]; }
#pragma unroll
                for (int d = 0; d < ND; ++d) cen += (-inv2[d]) * Gf[d];
                Jc[C0] += Coef * (cen + Too);
#pragma unroll
                for (int d = 0; d < ND; ++d) Jc[adj_slot(DIM, d)] += Coef * (inv2[d] * Gf[d] + Tsum[d] * Tk[d]);
                double gs = 0;
#pragma unroll
                for (int k = 0; k < ND; ++k) {
                    const int dd = sdir(ND, k);
                    const double ad = inv2[dd] * H[adj_slot(DIM, dd)] + (-inv2[dd]) * H[C0];
                    gs += gCf[dd] * (ad + Tsum[dd]) - Tk[dd] * Tsum[dd];
                }
                Jc[C0] += Coef * gs;
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const double ad = inv2[d] * H[adj_slot(DIM, d)] + (-inv2[d]) * H[C0];
                    Jc[adj_slot(DIM, d)] += Coef * (gAf[d] * (ad + Tsum[d]) - Tk[d] * Tsum[d]);
                }
            }
        }

        // ---- in-plane gravity (Pg/3) and centrifugal (Sp/3): limited upwind ----
#pragma unroll
        for (int mode = 0; mode < 2; ++mode) {
            if (mode == 0 && !A.pg) continue;
            if (mode == 1 && !A.cf) continue;
#pragma unroll
            for (int j = 0; j < ND; ++j) {
                double dv;
                if (mode == 0) dv = sgn(j < 2 ? A.Pgx : A.Pgy);
                else if (j == 0) dv = (g.xlo + ixp * g.dx) - A.rotx;
                else if (j == 1) dv = (g.xlo + (ixp + 1) * g.dx) - A.rotx;
                else if (j == 2) dv = (g.ylo + iyp * g.dy) - A.roty;
                else dv = (g.ylo + (iyp + 1) * g.dy) - A.roty;
                const double w = (mode == 0) ? 1.0 : dv;
                const double Coef = (mode == 0) ? ((j < 2 ? A.Pgx : A.Pgy) / 3.0) : A.Sp / 3.0;
                const double ihs = A.inv_h[j / 2];
                double f[4], wt[4], gd[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) f[q] = H[face_slot(DIM, j, q)];
                const double hf = face_interp(f, dv, wt, gd, A.hs);
#pragma unroll
                for (int q = 0; q < 4; ++q) { wt[q] *= ihs; gd[q] *= ihs; }
                double D = (mob.slip ? mob.DperH(hf) : hf * hf) / fq[j].mu;
                if (wall(j)) D = 0;
                const double sg = (j % 2 == 0) ? -1.0 : 1.0;
#pragma unroll
                for (int q = 0; q < 4; ++q) Ar[face_slot(DIM, j, q)] += Coef * (sg * w * wt[q] * D);
                if (MODE == 1) {
                    const double r3 = !mob.slip ? 3.0 : (mob.DperH(hf) != 0.0 ? mob.dD(hf) / mob.DperH(hf) : 0.0);
#pragma unroll
                    for (int q = 0; q < 4; ++q) Jc[face_slot(DIM, j, q)] += Coef * (sg * w * (r3 * gd[q] - wt[q]) * D);
                }
            }
            if (mode == 1 && DIM == 1) {
                // sideways outflow of the 1-D centrifugal term (BodyForceKernel.h SideOutflow)
                const double hoo = H[C0], mu = visc_of(A, Cc[C0]), Mg = A.Sp / 3.0;
                Ar[C0] += mob.slip ? Mg * (mob.DperH(hoo) / mu) : Mg * (hoo * hoo / mu);
                if (MODE == 1)
                    Jc[C0] += mob.slip ? Mg * ((mob.dD(hoo) - mob.DperH(hoo)) / mu) : Mg * (2.0 * hoo * hoo / mu);
            }
        }

        // ---- Marangoni (As Ma / 2): limited upwind by the sign of the stress ----
        if (A.ma) {
            const double Coef = A.AsMa2;
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                const int lo = face_slot(DIM, d, 1), hi = face_slot(DIM, d, 2);
                const double ihh = A.inv_h[d / 2];
                const double dc = (Cc[hi] - Cc[lo]) * ihh;
                const double up = sgn(Coef) * sgn(dc);
                double f[4], wt[4], gd[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) f[q] = H[face_slot(DIM, d, q)];
                const double hf = face_interp(f, up, wt, gd, A.hs);
                const bool wl = wall(d);
                const double mu = fq[d].mu;
                const double Sf = wl ? 0.0 : (mob.slip ? dc / mu * (hf + 2.0 * mob.lambda) : dc / mu * hf);
                const double sg = (d % 2 == 0) ? -1.0 : 1.0;
                const double Sfh = Sf * ihh;
#pragma unroll
                for (int q = 0; q < 4; ++q) Ar[face_slot(DIM, d, q)] += Coef * (sg * wt[q] * Sfh);
                if (MODE == 1) {
                    const double r = wl ? 0.0 : dc / mu;
                    if (mob.slip) {
                        const double L2 = 2.0 * mob.lambda, ad = sg * ihh;
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            Jc[face_slot(DIM, d, q)] += Coef * (r * ((2.0 * hf + L2) * gd[q] - (hf + L2) * wt[q]) * ad);
                    } else {
                        const double ad = sg * hf * ihh;
#pragma unroll
                        for (int q = 0; q < 4; ++q) Jc[face_slot(DIM, d, q)] += Coef * (r * (2.0 * gd[q] - wt[q]) * ad);
                    }
                }
            }
        }

        // ---- disjoining pressure (Dp/3), Pi = (h*/h)^n - (h*/h)^m ----
        if (A.dj) {
            const double Coef = A.Dp / 3.0;
            // (h*/h)^n and (h*/h)^m per point (pi_point), from the caller
            auto PiPerH = [&](int k) { return PiH[k]; };
            double cen = 0;
#pragma unroll
            for (int d = 0; d < ND; ++d) cen += (-inv2[d]) * Gf[d];
            Ar[C0] += Coef * (cen * PiPerH(C0));
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                const int a = adj_slot(DIM, d);
                Ar[a] += Coef * ((inv2[d] * Gf[d]) * PiPerH(a));
            }
            if (MODE == 1) {
                double Ad[ND];
#pragma unroll
                for (int d = 0; d < ND; ++d) Ad[d] = inv2[d] * PiS[adj_slot(DIM, d)] + (-inv2[d]) * PiS[C0];
                double ga = 0, gc = 0;
#pragma unroll
                for (int k = 0; k < ND; ++k) {
                    const int dd = sdir(ND, k);
                    ga += gCf[dd] * Ad[dd];
                    gc += Gf[dd] * (-inv2[dd]);
                }
                const double dpc = dPiS[C0] - PiPerH(C0);
                Jc[C0] += Coef * (ga + gc * dpc);
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    const int a = adj_slot(DIM, d);
                    const double dpa = dPiS[a] - PiPerH(a);
                    Jc[a] += Coef * (gAf[d] * Ad[d] + Gf[d] * inv2[d] * dpa);
                }
            }
        }

        // A(h)h of the terms held as entries
#pragma unroll
        for (int k = 0; k < NS; ++k) ax += Ar[k] * H[k];

        if (MODE == 0) {
            A.ax[i] = ax;
        } else {
            double R = A.hold[p] - H[C0] - A.dt * A.theta * ax - A.dt * (1.0 - A.theta) * A.axold[i];
            if (A.evold) {
                const double evp = (1.0 - A.theta) * A.evold[i] + A.theta * evap_rate(A.bt, H[C0], Cc[C0], A.dt);
                R = R - A.dt * (evp - A.src[i]);
                if (A.evp_out) A.evp_out[i] = evp;
            }
            rabs = fabs(R);
            if (A.rhs_raw) A.rhs_raw[i] = R;
            if (A.ax_raw) A.ax_raw[i] = ax;
            const double td = A.theta * A.dt;
            double IA[NS];
#pragma unroll
            for (int k = 0; k < NS; ++k) IA[k] = td * (Jc[k] + Ar[k]);
            IA[C0] = 1.0 + IA[C0];
            if (A.keep_ia) {
                const double inv = A.dinv[i];
                if (A.b32) A.b32[i] = (float)(A.row_scaling ? R * inv : R);
                else A.b[i] = A.row_scaling ? R * inv : R;
                return;
            }
            if (A.ia_raw) {
#pragma unroll
                for (int k = 0; k < NS; ++k) A.ia_raw[(size_t)k * g.N + i] = IA[k];
            }
            const double dg = IA[C0];
            const double inv = (A.row_scaling && fabs(dg) >= 1.0e-30) ? 1.0 / dg : 1.0;
            if (A.dinv) A.dinv[i] = inv;
            if (A.ia32) {
#pragma unroll
                for (int k = 0; k < NS; ++k) A.ia32[(size_t)k * g.N + i] = (float)(A.row_scaling ? IA[k] * inv : IA[k]);
                // a finite diagonal scaled by its own inverse is 1 within an
                // FP64 ulp, so exactly 1 in FP32: the fused FP32 iteration may
                // then skip that plane (the flag tells it when it may not)
                const bool unit = A.row_scaling && fabs(dg) >= 1.0e-30 && isfinite(dg);
                if (unit) A.ia32[(size_t)C0 * g.N + i] = 1.0f;
                else if (A.nonunit) atomicOr(A.nonunit, 1);
                A.b32[i] = (float)(A.row_scaling ? R * inv : R);
            } else {
#pragma unroll
                for (int k = 0; k < NS; ++k) A.ia[(size_t)k * g.N + i] = A.row_scaling ? IA[k] * inv : IA[k];
                A.b[i] = A.row_scaling ? R * inv : R;
            }
            if (A.feat) {
                // the learned preconditioner's inputs: the stiffness and the
                // shape of the scaled row (the face mobilities and spacings
                // it carries), the film height and the concentration
                double off = 0, far = 0;
#pragma unroll
                for (int k = 0; k < NS; ++k) if (k != C0) off += fabs(IA[k] * inv);
#pragma unroll
                for (int d = 0; d < ND; ++d) far += IA[face_slot(DIM, d, d % 2 == 0 ? 0 : 3)] * inv;
                const double so = off > 1e-30 ? off : 1e-30;
                const size_t N = g.N;
                A.feat[0 * N + i] = (float)(0.25 * log10(so));
                A.feat[1 * N + i] = (float)(IA[adj_slot(DIM, 0)] * inv / so);
                A.feat[2 * N + i] = (float)(IA[adj_slot(DIM, 1)] * inv / so);
                A.feat[3 * N + i] = (float)(DIM == 2 ? IA[adj_slot(DIM, 2)] * inv / so : 0.0);
                A.feat[4 * N + i] = (float)(DIM == 2 ? IA[adj_slot(DIM, 3)] * inv / so : 0.0);
                A.feat[5 * N + i] = (float)(far / so);
                A.feat[6 * N + i] = (float)H[C0];
                A.feat[7 * N + i] = (float)(Cc[C0] - A.visc_c0);
                A.mscale[i] = (float)(1.0 / (1.0 + off));
            }
        }
    }
}

// the largest |Rhs| of the block into A.rmax (mode 1)
__device__ inline void rows_rmax(const RowArgs &A, double rabs) {
    __shared__ double sh[4];
    for (int off = 32; off > 0; off >>= 1) rabs = fmax(rabs, __shfl_down(rabs, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = rabs;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, sh[w]);
        atomicMax(A.rmax, (unsigned long long)__double_as_longlong(m));
    }
}

// One thread per interior cell, its stencil and faces formed in registers
// (1-D films; 2-D films where the tiled kernel below is switched off).
// mode 0: A(h)h only (the explicit half); mode 1: the whole pass
template <int DIM, int MODE>
__global__ void __launch_bounds__(256) k_rows(RowArgs A) {
    constexpr int NS = DIM == 2 ? 13 : 5;
    constexpr int ND = DIM == 2 ? 4 : 2;
    constexpr int C0 = DIM == 2 ? 6 : 2;
    const Geo &g = A.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double rabs = 0;
    if (i < g.N) {
        const int ix = i % g.nx, iy = i / g.nx;
        const int ixp = ix + 2, iyp = DIM == 2 ? iy + 2 : 0;
        const int p = iyp * g.nxp + ixp;
        double H[NS], S[NS], Cc[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int q = p + slot_ox(DIM, k) + slot_oy(DIM, k) * g.nxp;
            H[k] = A.h[q]; S[k] = A.s[q]; Cc[k] = A.c[q];
        }
        FaceQ fq[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            const int a = adj_slot(DIM, d);
            double Fm = 0, Fm1 = 0, Fm2 = 0;
            if (A.mob.slip) A.mob.FaceDK3(H[C0], H[a], Fm, Fm1, Fm2);
            face_q<MODE>(A, H[C0], H[a], Cc[C0], Cc[a], Fm, Fm1, Fm2, fq[d]);
        }
        double PiS[NS], dPiS[NS], PiH[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) { PiS[k] = 0; dPiS[k] = 0; PiH[k] = 0; }
        if (A.dj) {
#pragma unroll
            for (int d = -1; d < ND; ++d) {
                const int k = d < 0 ? C0 : adj_slot(DIM, d);
                pi_point(A, H[k], PiS[k], dPiS[k], PiH[k]);
            }
        }
        row_body<DIM, MODE>(A, i, ixp, iyp, p, H, S, Cc, fq, PiS, dPiS, PiH, rabs);
    }
    if (MODE == 1) rows_rmax(A, rabs);
}

// The same rows of a 2-D film from tiles of 64 x 4 cells (256 threads).
// The tile's film, substrate and concentration with their 2-cell halo are
// staged in LDS once; g(h) of the slip mobility and the disjoining powers
// are formed once per point; every face's quantities (face_q) once per face
// for both of its cells. A cell's row then reads its stencil and its four
// faces from LDS: the same values as k_rows forms for itself, so the same
// rows to the bit, with about half the transcendental work per cell.
constexpr int RTX = 64, RTY = 4;
constexpr int RLX = RTX + 4, RLY = RTY + 4;                        // with the halo
constexpr int RNXF = (RTX + 1) * RTY, RNYF = RTX * (RTY + 1);      // x-faces, y-faces of a tile

template <int MODE>
__global__ void __launch_bounds__(256) k_rows_tiled(RowArgs A) {
    constexpr int NS = 13, ND = 4, C0 = 6;
    __shared__ double sH[RLY][RLX], sS[RLY][RLX], sC[RLY][RLX];
    __shared__ double sG[RLY - 2][RLX - 2];              // g(h), points of the 1-cell halo
    __shared__ double sP[3][RLY - 2][RLX - 2];           // Pi, dPi/dh, Pi/h there
    __shared__ double sF[8][RNXF + RNYF];                // face_q of every face, for its lower cell
    const Geo &g = A.g;
    const int bx = blockIdx.x * RTX, by = blockIdx.y * RTY;   // the tile's first interior cell
    const int tid = threadIdx.x;
    // the tile and its halo (LDS point (lx, ly) is padded point (bx + lx, by + ly))
    for (int q = tid; q < RLY * RLX; q += 256) {
        const int ly = q / RLX, lx = q - ly * RLX;
        const int px = bx + lx, py = by + ly;
        double h = 0, s = 0, c = 0;
        if (px < g.nxp && py < g.nyp) {
            const int j = py * g.nxp + px;
            h = A.h[j]; s = A.s[j]; c = A.c[j];
        }
        sH[ly][lx] = h; sS[ly][lx] = s; sC[ly][lx] = c;
    }
    __syncthreads();
    const bool slip = A.mob.slip;
    if (slip || A.dj) {
        for (int q = tid; q < (RLY - 2) * (RLX - 2); q += 256) {
            const int ly = q / (RLX - 2), lx = q - ly * (RLX - 2);
            const double h = sH[ly + 1][lx + 1];
            if (slip) sG[ly][lx] = h > 0 ? A.mob.Gint(h) : 0.0;
            if (A.dj) pi_point(A, h, sP[0][ly][lx], sP[1][ly][lx], sP[2][ly][lx]);
        }
        __syncthreads();
    }
    // faces: x-face (fx, fy) lies between tile cells (fx - 1, fy) and
    // (fx, fy), fx = 0..RTX; y-face (fx, fy) between (fx, fy - 1) and (fx, fy)
    for (int f = tid; f < RNXF + RNYF; f += 256) {
        int llx, lly, ulx, uly;   // LDS points of the lower and upper cell
        if (f < RNXF) {
            const int fy = f / (RTX + 1), fx = f - fy * (RTX + 1);
            llx = fx + 1; lly = fy + 2; ulx = fx + 2; uly = fy + 2;
        } else {
            const int e = f - RNXF, fy = e / RTX, fx = e - fy * RTX;
            llx = fx + 2; lly = fy + 1; ulx = fx + 2; uly = fy + 2;
        }
        const double hl = sH[lly][llx], hu = sH[uly][ulx];
        double Fm = 0, Fm1 = 0, Fm2 = 0;
        if (slip) A.mob.FaceDK3T<true>(hl, hu, sG[lly - 1][llx - 1], sG[uly - 1][ulx - 1], Fm, Fm1, Fm2);
        FaceQ q;
        face_q<MODE>(A, hl, hu, sC[lly][llx], sC[uly][ulx], Fm, Fm1, Fm2, q);
        sF[0][f] = q.mu; sF[1][f] = q.D; sF[2][f] = q.gC; sF[3][f] = q.gA;
        sF[4][f] = q.G; sF[5][f] = q.T; sF[6][f] = q.gCf; sF[7][f] = q.gAf;
    }
    __syncthreads();
    const int tx = tid & (RTX - 1), ty = tid / RTX;
    const int ix = bx + tx, iy = by + ty;
    double rabs = 0;
    if (ix < g.nx && iy < g.ny) {
        const int i = iy * g.nx + ix;
        const int ixp = ix + 2, iyp = iy + 2;
        const int p = iyp * g.nxp + ixp;
        double H[NS], S[NS], Cc[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int ly = ty + 2 + slot_oy(2, k), lx = tx + 2 + slot_ox(2, k);
            H[k] = sH[ly][lx]; S[k] = sS[ly][lx]; Cc[k] = sC[ly][lx];
        }
        // faces 0 (x-) and 2 (y-): the cell is their upper cell, so the
        // derivatives trade places; faces 1 (x+) and 3 (y+): the lower
        FaceQ fq[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            const int f = d == 0 ? ty * (RTX + 1) + tx : d == 1 ? ty * (RTX + 1) + tx + 1
                        : d == 2 ? RNXF + ty * RTX + tx : RNXF + (ty + 1) * RTX + tx;
            const bool upper = (d % 2) == 0;
            fq[d].mu = sF[0][f]; fq[d].D = sF[1][f];
            fq[d].gC = upper ? sF[3][f] : sF[2][f];
            fq[d].gA = upper ? sF[2][f] : sF[3][f];
            fq[d].G = sF[4][f]; fq[d].T = sF[5][f];
            fq[d].gCf = upper ? sF[7][f] : sF[6][f];
            fq[d].gAf = upper ? sF[6][f] : sF[7][f];
        }
        double PiS[NS], dPiS[NS], PiH[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) { PiS[k] = 0; dPiS[k] = 0; PiH[k] = 0; }
        if (A.dj) {
#pragma unroll
            for (int d = -1; d < ND; ++d) {
                const int k = d < 0 ? C0 : adj_slot(2, d);
                const int ly = ty + 1 + slot_oy(2, k), lx = tx + 1 + slot_ox(2, k);
                PiS[k] = sP[0][ly][lx]; dPiS[k] = sP[1][ly][lx]; PiH[k] = sP[2][ly][lx];
            }
        }
        row_body<2, MODE>(A, i, ixp, iyp, p, H, S, Cc, fq, PiS, dPiS, PiH, rabs);
    }
    if (MODE == 1) rows_rmax(A, rabs);
}

}  // namespace yz
