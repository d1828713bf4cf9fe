// yz_energy.h -- the film's energy budget (Output.Energy: EnergyBudget.h
// :52-166 with the parameters of WellMixedModelBase.h:869-892), for
// diagnostics.
//
//   F = sum over the real cells of Bo/2 f^2 + Dp w(h) and of
//       sigma_f/2 ((f_j - f_i)/dn)^2 over the cell's high faces (and its low
//       face against a ghost on an edge that is not Periodic), times the
//       cell measure; f = h + s, w(h) = -int_{h*}^{h} Pi;
//   D = sum over the real cells of phi_i (-div_i - r_i), times the cell
//       measure (FluxPower, HeightEquationBase.h:275-291): phi the potential
//       sigma lap f + Dp Pi(h) - Bo f (LubFields.cpp:300-311), div the
//       difference of the cell's face volume fluxes, r its sideways outflow.
//
// One thread per real cell forms its part of F and of D; each block sums its
// threads' parts in a fixed tree, and one block then sums the blocks' in a
// fixed order, so the result does not depend on scheduling.
#pragma once
#include "yz_rows.h"

namespace yz {

static constexpr int EB = 256;

template <int DIM>
__global__ void __launch_bounds__(EB) k_energy_cells(RowArgs A, const double *Fx, const double *Fy,
                                                     const double *rr, double *part) {
    const Geo &g = A.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0, pw = 0.0;
    if (i < g.N) {
        const int ix = i % g.nx + 2, iy = DIM == 2 ? i / g.nx + 2 : 0;
        const int p = iy * g.nxp + ix;
        const double bo = A.vg ? A.Bo : 0.0, dp = A.dj ? A.Dp : 0.0;
        auto f = [&](int q) { return A.h[q] + A.s[q]; };
        auto face = [&](double fi, double fj, double ci, double cj, double dn) {
            if (!A.cap) return 0.0;
            const double gr = (fj - fi) / dn;
            return 0.25 * (sig_of(A, ci) + sig_of(A, cj)) * gr * gr;
        };
        const double hh = A.h[p], fo = f(p);
        double w = 0.0;
        if (dp != 0 && hh > 0) {
            const double hs = A.hstar, n = A.pn, m = A.pm;
            const double in = pow(hs, n) * (pow(hh, 1.0 - n) - pow(hs, 1.0 - n)) / (1.0 - n);
            const double im = pow(hs, m) * (pow(hh, 1.0 - m) - pow(hs, 1.0 - m)) / (1.0 - m);
            w = -dp * (in - im);
        }
        e = 0.5 * bo * fo * fo + w;
        const int xp = p + 1, xm = p - 1;
        e += face(fo, f(xp), A.c[p], A.c[xp], g.dx);
        if (DIM == 2) e += face(fo, f(p + g.nxp), A.c[p], A.c[p + g.nxp], g.dy);
        if (g.bc[0] != YZ_BC_PERIODIC && ix == 2) e += face(f(xm), fo, A.c[xm], A.c[p], g.dx);
        if (DIM == 2 && g.bc[2] != YZ_BC_PERIODIC && iy == 2)
            e += face(f(p - g.nxp), fo, A.c[p - g.nxp], A.c[p], g.dy);
        // the potential and its power against the face fluxes
        double lap = (f(xp) - 2.0 * fo + f(xm)) / (g.dx * g.dx);
        if (DIM == 2) lap = lap + (f(p + g.nxp) - 2.0 * fo + f(p - g.nxp)) / (g.dy * g.dy);
        double phi = 0.0;
        if (A.cap) phi += sig_of(A, A.c[p]) * lap;
        if (dp != 0 && hh > 0) {
            const double r = A.hstar / hh;
            phi += dp * (pow(r, A.pn) - pow(r, A.pm));
        }
        if (bo != 0) phi -= bo * fo;
        double div = 0.0;
        div += Fx[p] - Fx[p - 1];
        if (DIM == 2) div += Fy[p] - Fy[p - g.nxp];
        pw = phi * (-div - rr[p]);
    }
    __shared__ double se[EB], sp[EB];
    se[threadIdx.x] = e;
    sp[threadIdx.x] = pw;
    __syncthreads();
    for (int off = EB / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            se[threadIdx.x] += se[threadIdx.x + off];
            sp[threadIdx.x] += sp[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = se[0];
        part[2 * blockIdx.x + 1] = sp[0];
    }
}

// the blocks' parts in a fixed order, times the cell measure
__global__ void __launch_bounds__(EB) k_energy_final(int nblocks, const double *part, double cm, double *out) {
    __shared__ double se[EB], sp[EB];
    double e = 0.0, p = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += EB) {
        e += part[2 * b];
        p += part[2 * b + 1];
    }
    se[threadIdx.x] = e;
    sp[threadIdx.x] = p;
    __syncthreads();
    for (int off = EB / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            se[threadIdx.x] += se[threadIdx.x + off];
            sp[threadIdx.x] += sp[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = se[0] * cm;
        out[1] = sp[0] * cm;
    }
}

// the step's volume balance: per real cell h, h c, and (when the case has
// B) the evaporation and the supply of the last pass, summed per block in a
// fixed tree
__global__ void __launch_bounds__(EB) k_balance_cells(Geo g, const double *h, const double *c, const double *evp,
                                                      const double *src, double *part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (i < g.N) {
        const int p = (g.dim == 2 ? i / g.nx + 2 : 0) * g.nxp + i % g.nx + 2;
        v[0] = h[p];
        v[1] = h[p] * c[p];
        if (evp) v[2] = evp[i];
        if (src) v[3] = src[i];
    }
    __shared__ double sv[4][EB];
#pragma unroll
    for (int k = 0; k < 4; ++k) sv[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = EB / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
#pragma unroll
            for (int k = 0; k < 4; ++k) sv[k][threadIdx.x] += sv[k][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k) part[4 * blockIdx.x + k] = sv[k][0];
}

__global__ void __launch_bounds__(EB) k_balance_final(int nblocks, const double *part, double *out) {
    __shared__ double sv[4][EB];
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblocks; b += EB)
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] += part[4 * b + k];
#pragma unroll
    for (int k = 0; k < 4; ++k) sv[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = EB / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
#pragma unroll
            for (int k = 0; k < 4; ++k) sv[k][threadIdx.x] += sv[k][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k) out[k] = sv[k][0];
}

// ---- the run's time series (History.dat, CalcInfo.dat) ----------------
// k_series_cells: per block, over the real cells, FieldInfo's quantities
// (FieldInfo.cpp:52-136): [0] sum h dA, [1] sum c h dA, [2] max h, [3] min h,
// [4] max h+s, [5] min h+s, [6] max c, [7] min c; over every row the pieces
// of ChangeRateOfVariable (ImplicitSolver.cpp:225-240): [8] max |h - h_old|,
// [9] max |h|; and the step's balances over the real cells: [10] B's
// evaporation of the last pass, [11] the nozzle's supply, [12] the supply
// times the concentration it arrives at (SuppliedResin,
// HeightEquationBase.h:325-351: the nozzle's, or the mean of the step's
// start and end c), [13] the 1-D centrifugal term's sideways
// rate at h and [14] times c, [15] the same at h_old and [16] times the
// step's starting c (SideOutflowRate, BodyForceKernel.h:85-109).
// k_series_final reduces the blocks in their order.
constexpr int SER_Q = 17, SER_B = 256, SER_G = 512;
__device__ inline void series_init(double v[SER_Q]) {
#pragma unroll
    for (int k = 0; k < SER_Q; ++k) v[k] = 0.0;
    v[2] = v[4] = v[6] = -1e40;
    v[3] = v[5] = v[7] = 1e40;
}
__device__ inline void series_merge(double *a, const double *b) {
    a[0] += b[0]; a[1] += b[1];
    a[2] = fmax(a[2], b[2]); a[3] = fmin(a[3], b[3]);
    a[4] = fmax(a[4], b[4]); a[5] = fmin(a[5], b[5]);
    a[6] = fmax(a[6], b[6]); a[7] = fmin(a[7], b[7]);
    a[8] = fmax(a[8], b[8]); a[9] = fmax(a[9], b[9]);
#pragma unroll
    for (int k = 10; k < SER_Q; ++k) a[k] += b[k];
}
__global__ void __launch_bounds__(SER_B) k_series_cells(RowArgs A, const double *hold, const double *cstart,
                                                        const double *evp, const double *src, int supplied,
                                                        double src_conc, int side, double *part) {
    const Geo &g = A.g;
    double v[SER_Q];
    series_init(v);
    const int np = g.nxp * g.nyp;
    const double Mg = A.Sp / 3.0;
    auto rate = [&](double hh, double cc) {
        const double mu = visc_of(A, cc);
        return A.mob.slip ? Mg * (A.mob.D(hh) / mu) : Mg * (hh * hh / mu) * hh;
    };
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < np; p += gridDim.x * blockDim.x) {
        const double hh = A.h[p];
        v[8] = fmax(v[8], fabs(hh - hold[p]));
        v[9] = fmax(v[9], fabs(hh));
        const int ix = p % g.nxp, iy = p / g.nxp;
        const bool real = ix >= 2 && ix < g.nxp - 2 && (g.dim == 1 || (iy >= 2 && iy < g.nyp - 2));
        if (!real) continue;
        const int i = (g.dim == 2 ? (iy - 2) * g.nx : 0) + ix - 2;
        const double cc = A.c[p], hs = hh + A.s[p];
        // each cell's volume and solute times the cell measure, in
        // FieldInfo's products (h dx dy, c h dx dy)
        if (g.dim == 2) {
            v[0] += hh * g.dx * g.dy;
            v[1] += cc * hh * g.dx * g.dy;
        } else {
            v[0] += hh * g.dx;
            v[1] += cc * hh * g.dx;
        }
        v[2] = fmax(v[2], hh); v[3] = fmin(v[3], hh);
        v[4] = fmax(v[4], hs); v[5] = fmin(v[5], hs);
        v[6] = fmax(v[6], cc); v[7] = fmin(v[7], cc);
        if (evp) v[10] += evp[i];
        if (src) {
            v[11] += src[i];
            v[12] += src[i] * (supplied ? src_conc : 0.5 * (cstart[p] + cc));
        }
        if (side) {
            const double r1 = rate(hh, cc), c0 = cstart[p], r0 = rate(hold[p], c0);
            v[13] += r1; v[14] += r1 * cc;
            v[15] += r0; v[16] += r0 * c0;
        }
    }
    __shared__ double sv[SER_Q][SER_B];
#pragma unroll
    for (int k = 0; k < SER_Q; ++k) sv[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = SER_B / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            double a[SER_Q], b[SER_Q];
#pragma unroll
            for (int k = 0; k < SER_Q; ++k) { a[k] = sv[k][threadIdx.x]; b[k] = sv[k][threadIdx.x + off]; }
            series_merge(a, b);
#pragma unroll
            for (int k = 0; k < SER_Q; ++k) sv[k][threadIdx.x] = a[k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < SER_Q; ++k) part[k * SER_G + blockIdx.x] = sv[k][0];
}
__global__ void __launch_bounds__(SER_B) k_series_final(int nblocks, const double *part, double *out) {
    double v[SER_Q];
    series_init(v);
    for (int b = threadIdx.x; b < nblocks; b += SER_B) {
        double w[SER_Q];
#pragma unroll
        for (int k = 0; k < SER_Q; ++k) w[k] = part[k * SER_G + b];
        series_merge(v, w);
    }
    __shared__ double sv[SER_Q][SER_B];
#pragma unroll
    for (int k = 0; k < SER_Q; ++k) sv[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = SER_B / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            double a[SER_Q], b[SER_Q];
#pragma unroll
            for (int k = 0; k < SER_Q; ++k) { a[k] = sv[k][threadIdx.x]; b[k] = sv[k][threadIdx.x + off]; }
            series_merge(a, b);
#pragma unroll
            for (int k = 0; k < SER_Q; ++k) sv[k][threadIdx.x] = a[k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < SER_Q; ++k) out[k] = sv[k][0];
}

}  // namespace yz
