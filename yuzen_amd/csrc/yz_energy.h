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

}  // namespace yz
