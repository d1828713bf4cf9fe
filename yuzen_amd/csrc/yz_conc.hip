// yz_conc.hip -- the well-mixed concentration on the device (gfx950). One
// thread per cell for every per-cell loop of the reference; the sums the
// reference forms in cell order (the mean residual of ConsistentFluxes, the
// excess left after MoveExcess) are formed here in a fixed order too (block
// partials, then one block), so that every run gives the same bits.
#include <hip/hip_runtime.h>

#include "yz_conc.h"

namespace yz {

namespace {

constexpr int CB = 256;         // threads per block
constexpr int RB = 256;         // blocks of a fixed-order reduction
inline int cblocks(int n) { return (n + CB - 1) / CB; }

__device__ inline int pad_index(const Geo &g, int i) {
    const int ix = i % g.nx, iy = i / g.nx;
    return (g.dim == 2 ? iy + 2 : 0) * g.nxp + ix + 2;
}
__device__ inline bool ghost_row(const Geo &g, int p) {
    const int ix = p % g.nxp, iy = p / g.nxp;
    if (ix < 2 || ix > g.nxp - 3) return true;
    return g.dim == 2 && (iy < 2 || iy > g.nyp - 3);
}
// the row of stencil slot k from padded row p
template <int DIM>
__device__ inline int slot_row(const Geo &g, int p, int k) {
    return p + slot_ox(DIM, k) + slot_oy(DIM, k) * g.nxp;
}

// ---- FaceFlux of each term (the oracle's flux_*; reference
// LaplacePressureKernel.h:480-491, VerticalGravityKernel.h:271-287,
// BodyForceKernel.h:118-159, SolutoCapillarityKernel.h:104-138,
// DisjoiningPressureKernel.h:160-168) -----------------------------------
template <int DIM>
struct Cell {
    static constexpr int NS = DIM == 2 ? 13 : 5;
    static constexpr int C0 = DIM == 2 ? 6 : 2;
    double H[NS], S[NS], Cc[NS];
    int ixp, iyp;
};

template <int DIM>
__device__ inline double Pi_of(const RowArgs &A, double hh) {
    if (hh <= 0) return 0;
    const double r = A.hstar / hh;
    return pow(r, A.pn) - pow(r, A.pm);
}

template <int DIM>
__device__ inline bool wall_of(const RowArgs &A, const Cell<DIM> &q, int d) {
    const Geo &g = A.g;
    if (d == 0) return q.ixp == 2 && is_wall(g.bc[0]);
    if (d == 1) return q.ixp == g.nxp - 3 && is_wall(g.bc[1]);
    if (d == 2) return q.iyp == 2 && is_wall(g.bc[2]);
    return q.iyp == g.nyp - 3 && is_wall(g.bc[3]);
}

template <int DIM>
__device__ double face_flux(const RowArgs &A, const Cell<DIM> &q, int d) {
    constexpr int C0 = Cell<DIM>::C0;
    const Geo &g = A.g;
    const Mob &mob = A.mob;
    const int as = adj_slot(DIM, d);
    const double hsp = (d / 2 == 0) ? g.dx : g.dy;
    double F = 0;
    if (A.cap) {
        const double Hoo = q.H[C0], Haj = q.H[as];
        const double cf = 0.5 * (q.Cc[C0] + q.Cc[as]);
        const double val = mob.slip ? mob.FaceDK(Hoo, Haj) : 2.0 * (Hoo * Hoo) * (Haj * Haj) / (Hoo + Haj);
        const double D = val * sig_of(A, cf) / visc_of(A, cf);
        F += (1.0 / 3.0) * (D * (face_grad<DIM>(A, q.H, d) + face_grad<DIM>(A, q.S, d)));
    }
    if (A.vg) {
        const int lo = (d % 2 == 0) ? as : C0, hi = (d % 2 == 0) ? C0 : as;
        const double hl = q.H[lo], hh = q.H[hi];
        const double mu = visc_of(A, 0.5 * (q.Cc[lo] + q.Cc[hi]));
        double gf = mob.Face(hl, hh);
        const double tk = (gf == 0) ? 0.0 : gf / (hl + hh) / mu;
        gf /= mu;
        F += (-1.0 * A.Bo / 3.0) * ((gf * (hh - hl) + (q.S[hi] - q.S[lo]) * (tk * (hl + hh))) / (hsp * hsp));
    }
#pragma unroll
    for (int mode = 0; mode < 2; ++mode) {
        // in-plane gravity is summed before the Marangoni term, the
        // centrifugal term after it (the reference's registration order)
        if (mode == 1 && A.ma) {
            double f[4], wt[4], gd[4];
            const int lo = face_slot(DIM, d, 1), hi = face_slot(DIM, d, 2);
            const double dc = (q.Cc[hi] - q.Cc[lo]) / hsp;
            const double cc = (q.Cc[hi] + q.Cc[lo]) / 2.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) f[k] = q.H[face_slot(DIM, d, k)];
            const double up = sgn(A.AsMa2) * sgn(dc);
            const double hf = face_interp(f, up, wt, gd, A.hs);
            const double Sf = wall_of<DIM>(A, q, d) ? 0.0 : dc / visc_of(A, cc) * (mob.slip ? hf + 2.0 * mob.lambda : hf);
            double hs = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) hs += wt[k] * f[k];
            F += A.AsMa2 * (Sf * hs / hsp);
        }
        if (mode == 0 && !A.pg) continue;
        if (mode == 1 && !A.cf) continue;
        double dv;
        if (mode == 0) dv = sgn(d < 2 ? A.Pgx : A.Pgy);
        else if (d == 0) dv = (g.xlo + q.ixp * g.dx) - A.rotx;
        else if (d == 1) dv = (g.xlo + (q.ixp + 1) * g.dx) - A.rotx;
        else if (d == 2) dv = (g.ylo + q.iyp * g.dy) - A.roty;
        else dv = (g.ylo + (q.iyp + 1) * g.dy) - A.roty;
        const double w = (mode == 0) ? 1.0 : dv;
        const double Mag = (mode == 0) ? ((d < 2 ? A.Pgx : A.Pgy) / 3.0) : A.Sp / 3.0;
        double f[4], wt[4], gd[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) f[k] = q.H[face_slot(DIM, d, k)];
        const double hf = face_interp(f, dv, wt, gd, A.hs);
#pragma unroll
        for (int k = 0; k < 4; ++k) wt[k] /= hsp;
        const double cc = 0.5 * (q.Cc[C0] + q.Cc[as]);
        double D = (mob.slip ? mob.DperH(hf) : hf * hf) / visc_of(A, cc);
        if (wall_of<DIM>(A, q, d)) D = 0;
        double hs = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) hs += wt[k] * f[k];
        F += Mag * (w * hs * D);
    }
    if (A.dj) {
        const double v = (1.0 / (hsp * hsp)) * Pi_of<DIM>(A, q.H[as]) + (-1.0 / (hsp * hsp)) * Pi_of<DIM>(A, q.H[C0]);
        const double G = mob.Face(q.H[C0], q.H[as]) / visc_of(A, 0.5 * (q.Cc[C0] + q.Cc[as]));
        const double f = (A.Dp / 3.0) * (G * v);
        F += (d % 2 == 0) ? -f : f;
    }
    return F;
}

template <int DIM>
__global__ void __launch_bounds__(CB) k_face_flux(RowArgs A, double *Fx, double *Fy, double *r) {
    constexpr int NS = Cell<DIM>::NS;
    constexpr int C0 = Cell<DIM>::C0;
    const Geo &g = A.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    Cell<DIM> q;
    const int ix = i % g.nx, iy = i / g.nx;
    q.ixp = ix + 2;
    q.iyp = DIM == 2 ? iy + 2 : 0;
    const int p = q.iyp * g.nxp + q.ixp;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int t = slot_row<DIM>(g, p, k);
        q.H[k] = A.h[t]; q.S[k] = A.s[t]; q.Cc[k] = A.c[t];
    }
    Fx[p] = face_flux<DIM>(A, q, 1);
    if (ix == 0) Fx[p - 1] = face_flux<DIM>(A, q, 0);
    if (DIM == 2) {
        Fy[p] = face_flux<DIM>(A, q, 3);
        if (iy == 0) Fy[p - g.nxp] = face_flux<DIM>(A, q, 2);
    }
    double rr = 0;
    if (DIM == 1 && A.cf) {
        const double hoo = q.H[C0], coo = q.Cc[C0], Mg = A.Sp / 3.0;
        rr = A.mob.slip ? Mg * (A.mob.D(hoo) / visc_of(A, coo)) : Mg * (hoo * hoo / visc_of(A, coo)) * hoo;
    }
    r[p] = rr;
}

// ---- the concentration system (ConcEquationBase.h:261-385) -------------
__device__ inline double kface(const ConcArgs &a, double ha, double hb, double ca, double cb, double dn) {
    if (!a.diff) return 0;
    return a.Df * exp(a.diff_exp * (0.5 * (ca + cb) - a.diff_c0)) * (0.5 * (ha + hb)) / (dn * dn);
}

template <int DIM>
__global__ void __launch_bounds__(CB) k_conc_rows(ConcArgs a) {
    constexpr int NS = DIM == 2 ? 13 : 5;
    constexpr int ND = DIM == 2 ? 4 : 2;
    constexpr int C0 = DIM == 2 ? 6 : 2;
    const Geo &g = a.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    const int p = pad_index(g, i);
    const double th = 0.5, dt = a.dt;
    double M[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) M[k] = 0;
    const double h1 = a.h[p], h0 = a.hold[p];
    const double c0p = a.c0[p];
    const double rr1 = a.side ? a.r1[p] : 0.0;
    // the supply: at the local c, theta d of it is the new level's; at the
    // nozzle's c, all of it is known (ConcEquationBase.h:310-318)
    const double S = a.src ? a.src[i] : 0.0;
    double diag = h1 + th * dt * (rr1 - (a.supplied ? 0.0 : S));
    double R = dt * (a.evp ? a.evp[i] : 0.0) * c0p;
    if (a.supplied) R += dt * S * (a.src_conc - c0p);
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        const int ax = d / 2;
        const bool high = (d % 2 == 1);
        const double sg = high ? 1.0 : -1.0;
        const int as = adj_slot(DIM, d);
        const int ia = slot_row<DIM>(g, p, as);
        const int lo = high ? p : ia;
        const double f1v = a.F1[ax][lo], f0v = a.F0[ax][lo];
        if (a.conv) {
            double w1[4], w0[4], f0[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) f0[q] = a.c0[slot_row<DIM>(g, p, face_slot(DIM, d, q))];
            if (a.lim) {
                // the row's cell is L of its high face, R of its low one
                double fl[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) fl[q] = a.climit[slot_row<DIM>(g, p, face_slot(DIM, d, q))];
                const int rowpos = high ? 1 : 2;
                limited_weights(fl, f1v, rowpos, a.lim_vl, w1);
                limited_weights(f0, f0v, rowpos, a.lim_vl, w0);
            } else {
                const double *u1 = f1v >= 0.0 ? a.wpos : a.wneg;
                const double *u0 = f0v >= 0.0 ? a.wpos : a.wneg;
#pragma unroll
                for (int q = 0; q < 4; ++q) { w1[q] = u1[q]; w0[q] = u0[q]; }
            }
            double e1 = 0, e0 = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int col = face_slot(DIM, d, q);
                M[col] += th * dt * sg * f1v * w1[q];
                e1 += w1[q] * (f0[q] - c0p);
                e0 += w0[q] * (f0[q] - c0p);
            }
            R -= dt * sg * (th * f1v * e1 + (1.0 - th) * f0v * e0);
        } else {
            diag += th * dt * sg * f1v;
        }
        if (a.diff) {
            const double dn = ax == 0 ? g.dx : g.dy;
            const double K1 = kface(a, a.h[p], a.h[ia], a.c[p], a.c[ia], dn);
            const double K0 = kface(a, h0, a.hold[ia], c0p, a.c0[ia], dn);
            diag += th * dt * K1;
            M[as] -= th * dt * K1;
            const double dc0 = high ? (a.c0[ia] - c0p) : (c0p - a.c0[ia]);
            R += dt * sg * (th * K1 * dc0 + (1.0 - th) * K0 * dc0);
        }
    }
    M[C0] += diag;
    const double dg = M[C0];
    const double inv = (a.row_scaling && fabs(dg) >= 1.0e-30) ? 1.0 / dg : 1.0;
#pragma unroll
    for (int k = 0; k < NS; ++k) a.A[(size_t)k * g.N + i] = a.row_scaling ? M[k] * inv : M[k];
    a.b[i] = a.row_scaling ? R * inv : R;
}

__global__ void k_conc_update(Geo g, const double *c0, const double *d, double *cx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    const int p = pad_index(g, i);
    cx[p] = c0[p] + d[i];
}

__global__ void k_ghosts_c(Geo g, double *v) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.nxp * g.nyp) return;
    if (!ghost_row(g, p)) return;
    const int ix = p % g.nxp, iy = p / g.nxp;
    const int sx = fold_axis(ix - 2, g.nx, g.bc[0] == YZ_BC_PERIODIC ? 1 : 0) + 2;
    const int sy = g.dim == 2 ? fold_axis(iy - 2, g.ny, g.bc[2] == YZ_BC_PERIODIC ? 1 : 0) + 2 : 0;
    v[p] = v[sy * g.nxp + sx];
}

// SoluteFlux (ConcEquationBase.h:427-457) of face d of the real cell at p
template <int DIM>
__device__ double solute_flux(const ConcArgs &a, int p, int d) {
    const Geo &g = a.g;
    const double th = 0.5;
    const int ax = d / 2;
    const bool high = (d % 2 == 1);
    const int ia = slot_row<DIM>(g, p, adj_slot(DIM, d));
    const int lo = high ? p : ia, hi = high ? ia : p;
    const double f1v = a.F1[ax][lo], f0v = a.F0[ax][lo];
    double v1 = 0, v0 = 0;
    if (a.lim) {
        // a limited scheme's face value evaluated at that very data
        double f1[4], f0[4], w[4], gd[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = slot_row<DIM>(g, p, face_slot(DIM, d, q));
            f1[q] = a.cx[j];
            f0[q] = a.c0[j];
        }
        v1 = limited_face(f1, f1v, w, gd, a.lim_vl);
        v0 = limited_face(f0, f0v, w, gd, a.lim_vl);
    } else {
        const double *w1 = f1v >= 0.0 ? a.wpos : a.wneg;
        const double *w0 = f0v >= 0.0 ? a.wpos : a.wneg;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = slot_row<DIM>(g, p, face_slot(DIM, d, q));
            v1 += w1[q] * a.cx[j];
            v0 += w0[q] * a.c0[j];
        }
    }
    double G1 = f1v * v1;
    double G0 = f0v * v0;
    if (a.diff) {
        const double dn = ax == 0 ? g.dx : g.dy;
        const double K1 = kface(a, a.h[lo], a.h[hi], a.c[lo], a.c[hi], dn);
        const double K0 = kface(a, a.hold[lo], a.hold[hi], a.c0[lo], a.c0[hi], dn);
        G1 -= K1 * (a.cx[hi] - a.cx[lo]);
        G0 -= K0 * (a.c0[hi] - a.c0[lo]);
    }
    return th * G1 + (1.0 - th) * G0;
}

template <int DIM>
__global__ void __launch_bounds__(CB) k_solute_flux(ConcArgs a) {
    const Geo &g = a.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    const int ix = i % g.nx, iy = i / g.nx;
    const int p = pad_index(g, i);
    a.G2[0][p] = solute_flux<DIM>(a, p, 1);
    if (ix == 0) a.G2[0][p - 1] = solute_flux<DIM>(a, p, 0);
    if (DIM == 2) {
        a.G2[1][p] = solute_flux<DIM>(a, p, 3);
        if (iy == 0) a.G2[1][p - g.nxp] = solute_flux<DIM>(a, p, 2);
    }
}

// Reconstruct's second loop (ConcEquationBase.h:490-517)
template <int DIM>
__global__ void __launch_bounds__(CB) k_reconstruct(ConcArgs a) {
    const Geo &g = a.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    const int p = pad_index(g, i);
    const double th = 0.5;
    const double h1 = a.h[p], h0 = a.hold[p];
    if (!(h1 > 0.0)) return;
    double div = 0;
    div += a.G2[0][p] - a.G2[0][p - 1];
    if (DIM == 2) div += a.G2[1][p] - a.G2[1][p - g.nxp];
    const double cc0 = a.c0[p];
    double q = h0 * cc0 - a.dt * div;
    if (a.src) q += a.dt * a.src[i] * (a.supplied ? a.src_conc : th * a.cx[p] + (1.0 - th) * cc0);
    if (a.side) q -= a.dt * (th * a.r1[p] * a.cx[p] + (1.0 - th) * a.r0[p] * cc0);
    a.cx[p] = q / h1;
}

__global__ void k_conc_take(Geo g, const double *cx, double *c, double *cprev, unsigned long long *dmax) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0;
    if (p < g.nxp * g.nyp) {
        double v = cx[p];
        if (v > 1.0) v = 1.0;
        c[p] = v;
        m = fabs(v - cprev[p]);
        cprev[p] = v;
    }
    __shared__ double sh[CB / 64];
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mx = 0;
        for (int w = 0; w < CB / 64; ++w) mx = fmax(mx, sh[w]);
        atomicMax(dmax, (unsigned long long)__double_as_longlong(mx));
    }
}

__global__ void k_conc_settle(int n, const double *cx, double *climit, unsigned long long *dmax) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0;
    if (p < n) {
        const double v = cx[p];
        m = fabs(v - climit[p]);
        climit[p] = v;
    }
    __shared__ double sh[CB / 64];
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mx = 0;
        for (int w = 0; w < CB / 64; ++w) mx = fmax(mx, sh[w]);
        atomicMax(dmax, (unsigned long long)__double_as_longlong(mx));
    }
}

__global__ void k_film_check(Geo g, const double *h, int *flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    if (!(h[pad_index(g, i)] > 0.0)) atomicOr(flag, 1);
}

__global__ void k_conc_commit(int n, const double *cx, double *c0, double *c, double *cfield) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double v = cx[p];
    c0[p] = v;
    const double w = v > 1.0 ? 1.0 : v;
    c[p] = w;
    cfield[p] = w;
}

__global__ void k_nonfinite(int n, const double *d, int *flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !isfinite(d[i])) atomicOr(flag, 1);
}
__global__ void k_zero_if(int n, double *d, const int *flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && *flag) d[i] = 0.0;
}

// ---- fixed-order sums -----------------------------------------------------
__device__ inline double block_sum(double v, double *sh) {
    // a tree of fixed shape over the block's threads
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = CB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}
// part[b] = sum over the grid-stride slice of block b of v (RB blocks)
__global__ void __launch_bounds__(CB) k_sum_part(int n, const double *v, double *part) {
    __shared__ double sh[CB];
    double s = 0;
    for (int i = blockIdx.x * CB + threadIdx.x; i < n; i += RB * CB) s += v[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}
__global__ void __launch_bounds__(CB) k_sum_final(const double *part, double *out) {
    __shared__ double sh[CB];
    const double t = block_sum(part[threadIdx.x], sh);
    if (threadIdx.x == 0) *out = t;
}
void fixed_sum(int n, const double *v, double *part, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_sum_part, dim3(RB), dim3(CB), 0, st, n, v, part);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(CB), 0, st, (const double *)part, out);
}

// ---- ConsistentFluxes (HeightEquationBase.h:985-1058) ------------------
__global__ void k_cf_rho(ConcArgs a, const double *Fx, const double *Fy, double *rho) {
    const Geo &g = a.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    const int p = pad_index(g, i);
    const double th = a.theta_h;
    double div1 = 0, div0 = 0;
    div1 += Fx[p] - Fx[p - 1];
    div0 += a.F0[0][p] - a.F0[0][p - 1];
    if (g.dim == 2) {
        div1 += Fy[p] - Fy[p - g.nxp];
        div0 += a.F0[1][p] - a.F0[1][p - g.nxp];
    }
    const double es = a.evp ? a.evp[i] : 0.0, ss = a.src ? a.src[i] : 0.0;
    rho[i] = (a.hold[p] - a.h[p]) / a.dt - (th * (div1 + a.r1[p]) + (1.0 - th) * (div0 + a.r0[p]) + (es - ss));
}
// rowsum[jy] = sum over the row of (rho - mean), one block per row
__global__ void __launch_bounds__(CB) k_cf_rowsum(Geo g, const double *rho, const double *mean_sum,
                                                  double *rowsum) {
    __shared__ double sh[CB];
    const int jy = blockIdx.x;
    const double mean = *mean_sum / (double)g.N;
    double s = 0;
    for (int ix = threadIdx.x; ix < g.nx; ix += CB) s += rho[jy * g.nx + ix] - mean;
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) rowsum[jy] = t;
}
// along y: the running sum of the rows' shares (one thread; ny values)
__global__ void k_cf_accy(Geo g, const double *rowsum, double *accy) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double acc = 0;
    for (int jy = 0; jy + 1 < g.ny; ++jy) {
        acc += rowsum[jy] / (double)g.nx;
        accy[jy] = acc;
    }
}
__global__ void k_cf_apply_y(Geo g, double w, const double *accy, double *Fy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    const int iy = i / g.nx;
    if (iy + 1 >= g.ny) return;
    Fy[pad_index(g, i)] += w * accy[iy];
}
// along x within each row: the running sum of (rho - mean) - share over the
// faces between real cells, one block per row (chunks per thread, then an
// exclusive scan of the chunk sums)
__global__ void __launch_bounds__(CB) k_cf_scan_x(Geo g, double w, const double *rho, const double *mean_sum,
                                                  const double *rowsum, double *Fx) {
    __shared__ double sh[CB];
    const int jy = blockIdx.x;
    const double mean = *mean_sum / (double)g.N;
    const double share = (g.dim == 2) ? rowsum[jy] / (double)g.nx : 0.0;
    const int m = g.nx - 1;   // faces between real cells of the row
    const int chunk = (m + CB - 1) / CB;
    const int b = threadIdx.x * chunk, e = min(b + chunk, m);
    double s = 0;
    for (int ix = b; ix < e; ++ix) s += (rho[jy * g.nx + ix] - mean) - share;
    sh[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double acc = 0;
        for (int t = 0; t < CB; ++t) { const double v = sh[t]; sh[t] = acc; acc += v; }
    }
    __syncthreads();
    double acc = sh[threadIdx.x];
    const int row = (g.dim == 2 ? jy + 2 : 0) * g.nxp + 2;
    for (int ix = b; ix < e; ++ix) {
        acc += (rho[jy * g.nx + ix] - mean) - share;
        Fx[row + ix] += w * acc;
    }
}

// ---- RedistributeExcess (HeightEquationBase.h:1128-1273) ---------------
__global__ void k_ex_flags(ExcessArgs a) {
    const Geo &g = a.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    const double v = a.c[pad_index(g, i)];
    if (v > 1.0) atomicOr(a.flags + 0, 1);
    if (v < 0.0) atomicOr(a.flags + 1, 1);
}
__device__ inline bool ex_outside(int lower, double v) { return lower ? (v < 0.0) : (v > 1.0); }
__device__ inline double ex_beyond(int lower, double h, double v) { return lower ? -h * v : h * (v - 1.0); }
__device__ inline double ex_room(int lower, double h, double v) { return lower ? h * v : h * (1.0 - v); }

__global__ void k_ex_room(ExcessArgs a) {
    const Geo &g = a.g;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.nxp * g.nyp) return;
    const double h = a.h[p], v = a.c[p];
    const bool inside = a.lower ? (v > 0.0) : (v < 1.0);
    a.room[p] = (!ghost_row(g, p) && h > 0.0 && inside) ? ex_room(a.lower, h, v) : 0.0;
    a.gain[p] = 0.0;
    a.f[p] = -1.0;
    a.e[p] = 0.0;
}
// the cells outside the bound: what they hold beyond it goes to the room of
// their face neighbours (f[p] the fraction taken, -1 when none)
__global__ void k_ex_move(ExcessArgs a, int pass) {
    const Geo &g = a.g;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.nxp * g.nyp || ghost_row(g, p)) return;
    const double v = a.c[p];
    if (!ex_outside(a.lower, v)) return;
    const double bound = a.lower ? 0.0 : 1.0;
    const double h = a.h[p];
    if (!(h > 0.0)) { a.c[p] = bound; return; }
    const double e = ex_beyond(a.lower, h, v);
    if (pass == 0) a.e[p] = e;
    const int nbr[4] = {p - 1, p + 1, p - g.nxp, p + g.nxp};
    const int nd = g.dim == 2 ? 4 : 2;
    double r = 0.0;
    for (int d = 0; d < nd; ++d) r += a.room[nbr[d]];
    if (!(r > 0.0)) { atomicOr(a.flags + 2, 1); return; }
    const double f = (e < r) ? e / r : 1.0;
    a.f[p] = f;
    if (f < 1.0) a.c[p] = bound;
    else { a.c[p] = a.lower ? -(e - r) / h : 1.0 + (e - r) / h; atomicOr(a.flags + 2, 1); }
}
// the gains, gathered in the order the reference's cell loop adds them
// (the giving neighbours by increasing row), then applied
__global__ void k_ex_gain(ExcessArgs a) {
    const Geo &g = a.g;
    const int n = g.nxp * g.nyp;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double rj = a.room[j];
    double gain = 0.0;
    const int cand2[4] = {j - g.nxp, j - 1, j + 1, j + g.nxp};
    const int cand1[2] = {j - 1, j + 1};
    const int nc = g.dim == 2 ? 4 : 2;
    for (int k = 0; k < nc; ++k) {
        const int i = g.dim == 2 ? cand2[k] : cand1[k];
        if (i < 0 || i >= n || ghost_row(g, i)) continue;
        const double fi = a.f[i];
        if (fi >= 0.0) gain += fi * rj;
    }
    a.gain[j] = gain;
}
__global__ void k_ex_apply(ExcessArgs a) {
    const Geo &g = a.g;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= g.nxp * g.nyp) return;
    const double gn = a.gain[j];
    if (gn > 0.0) {
        const double dc = gn / a.h[j];
        const double v = a.c[j] + (a.lower ? -dc : dc);
        a.c[j] = v;
        if (ex_outside(a.lower, v)) atomicOr(a.flags + 2, 1);
    }
}
// what is left outside after the passes, and the room there is
__global__ void k_ex_left(ExcessArgs a) {
    const Geo &g = a.g;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.nxp * g.nyp) return;
    double left = 0.0, total = 0.0;
    if (!ghost_row(g, p)) {
        const double h = a.h[p], v = a.c[p];
        if (ex_outside(a.lower, v)) { left = ex_beyond(a.lower, h, v); a.c[p] = a.lower ? 0.0 : 1.0; }
        else if (h > 0.0) total = ex_room(a.lower, h, v);
    }
    a.room[p] = left;
    a.gain[p] = total;
}
__global__ void k_ex_spread(ExcessArgs a) {
    const Geo &g = a.g;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.nxp * g.nyp || ghost_row(g, p)) return;
    const double left = a.sums[1], total = a.sums[2];
    const double f = (total > 0.0) ? fmin(1.0, left / total) : 0.0;
    if (!(f > 0.0) || !(a.h[p] > 0.0)) return;
    double v = a.c[p];
    if (a.lower) {
        if (!(v > 0.0)) return;
        v -= f * v;
        if (v < 0.0) v = 0.0;
    } else {
        if (!(v < 1.0)) return;
        v += f * (1.0 - v);
        if (v > 1.0) v = 1.0;
    }
    a.c[p] = v;
}
__global__ void k_ex_caps(Geo g, double *c, int under) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.nxp * g.nyp) return;
    double v = c[p];
    if (v > 1.0) v = 1.0;
    if (under && v < 0.0) v = 0.0;
    c[p] = v;
}

}  // namespace

__global__ void k_step_evp(RowArgs A, const double *h, const double *c, double *evp) {
    const Geo &g = A.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    const int p = pad_index(g, i);
    evp[i] = (1.0 - A.theta) * A.evold[i] + A.theta * evap_rate(A.bt, h[p], c[p], A.dt);
}
void launch_step_evp(const RowArgs &a, const double *h, const double *c, double *evp, hipStream_t st) {
    hipLaunchKernelGGL(k_step_evp, dim3((a.g.N + CB - 1) / CB), dim3(CB), 0, st, a, h, c, evp);
}

void launch_face_fluxes(const RowArgs &a, double *Fx, double *Fy, double *r, hipStream_t st) {
    if (a.g.dim == 2) hipLaunchKernelGGL(k_face_flux<2>, dim3(cblocks(a.g.N)), dim3(CB), 0, st, a, Fx, Fy, r);
    else hipLaunchKernelGGL(k_face_flux<1>, dim3(cblocks(a.g.N)), dim3(CB), 0, st, a, Fx, Fy, r);
}
void launch_conc_rows(const ConcArgs &a, hipStream_t st) {
    if (a.g.dim == 2) hipLaunchKernelGGL(k_conc_rows<2>, dim3(cblocks(a.g.N)), dim3(CB), 0, st, a);
    else hipLaunchKernelGGL(k_conc_rows<1>, dim3(cblocks(a.g.N)), dim3(CB), 0, st, a);
}
void launch_ghosts(const Geo &g, double *v, hipStream_t st) {
    hipLaunchKernelGGL(k_ghosts_c, dim3(cblocks(g.nxp * g.nyp)), dim3(CB), 0, st, g, v);
}
void launch_conc_update(const ConcArgs &a, const double *d, hipStream_t st) {
    hipLaunchKernelGGL(k_conc_update, dim3(cblocks(a.g.N)), dim3(CB), 0, st, a.g, a.c0, d, a.cx);
    launch_ghosts(a.g, a.cx, st);
}
void launch_reconstruct(const ConcArgs &a, hipStream_t st) {
    if (a.g.dim == 2) {
        hipLaunchKernelGGL(k_solute_flux<2>, dim3(cblocks(a.g.N)), dim3(CB), 0, st, a);
        hipLaunchKernelGGL(k_reconstruct<2>, dim3(cblocks(a.g.N)), dim3(CB), 0, st, a);
    } else {
        hipLaunchKernelGGL(k_solute_flux<1>, dim3(cblocks(a.g.N)), dim3(CB), 0, st, a);
        hipLaunchKernelGGL(k_reconstruct<1>, dim3(cblocks(a.g.N)), dim3(CB), 0, st, a);
    }
}
void launch_conc_take(const Geo &g, const double *cx, double *c, double *cprev, unsigned long long *dmax,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_conc_take, dim3(cblocks(g.nxp * g.nyp)), dim3(CB), 0, st, g, cx, c, cprev, dmax);
}
void launch_conc_settle(const Geo &g, const double *cx, double *climit, unsigned long long *dmax, hipStream_t st) {
    const int n = g.nxp * g.nyp;
    hipLaunchKernelGGL(k_conc_settle, dim3(cblocks(n)), dim3(CB), 0, st, n, cx, climit, dmax);
}
void launch_film_check(const Geo &g, const double *h, int *flag, hipStream_t st) {
    hipLaunchKernelGGL(k_film_check, dim3(cblocks(g.N)), dim3(CB), 0, st, g, h, flag);
}
void launch_conc_commit(const Geo &g, const double *cx, double *c0, double *c, double *cfield, hipStream_t st) {
    const int n = g.nxp * g.nyp;
    hipLaunchKernelGGL(k_conc_commit, dim3(cblocks(n)), dim3(CB), 0, st, n, cx, c0, c, cfield);
}
void launch_zero_if_nonfinite(int n, double *d, int *flag, hipStream_t st) {
    hipLaunchKernelGGL(k_nonfinite, dim3(cblocks(n)), dim3(CB), 0, st, n, (const double *)d, flag);
    hipLaunchKernelGGL(k_zero_if, dim3(cblocks(n)), dim3(CB), 0, st, n, d, (const int *)flag);
}

void launch_consistent_fluxes(const ConcArgs &a, double *Fx, double *Fy, double *scratch, hipStream_t st) {
    const Geo &g = a.g;
    double *rho = scratch;                  // N
    double *rowsum = rho + g.N;             // ny
    double *accy = rowsum + g.ny;           // ny
    double *part = accy + g.ny;             // RB
    double *sum = part + RB;                // 1
    const double w = 1.0 / a.theta_h;
    hipLaunchKernelGGL(k_cf_rho, dim3(cblocks(g.N)), dim3(CB), 0, st, a, (const double *)Fx, (const double *)Fy, rho);
    fixed_sum(g.N, rho, part, sum, st);
    if (g.dim == 2) {
        hipLaunchKernelGGL(k_cf_rowsum, dim3(g.ny), dim3(CB), 0, st, g, (const double *)rho, (const double *)sum,
                           rowsum);
        hipLaunchKernelGGL(k_cf_accy, dim3(1), dim3(64), 0, st, g, (const double *)rowsum, accy);
        hipLaunchKernelGGL(k_cf_apply_y, dim3(cblocks(g.N)), dim3(CB), 0, st, g, w, (const double *)accy, Fy);
    }
    hipLaunchKernelGGL(k_cf_scan_x, dim3(g.dim == 2 ? g.ny : 1), dim3(CB), 0, st, g, w, (const double *)rho,
                       (const double *)sum, (const double *)rowsum, Fx);
}

void launch_excess_flags(const ExcessArgs &a, hipStream_t st) {
    hipLaunchKernelGGL(k_ex_flags, dim3(cblocks(a.g.N)), dim3(CB), 0, st, a);
}
void launch_excess_pass(const ExcessArgs &a, int pass, hipStream_t st) {
    const int n = a.g.nxp * a.g.nyp;
    hipLaunchKernelGGL(k_ex_room, dim3(cblocks(n)), dim3(CB), 0, st, a);
    hipLaunchKernelGGL(k_ex_move, dim3(cblocks(n)), dim3(CB), 0, st, a, pass);
    hipLaunchKernelGGL(k_ex_gain, dim3(cblocks(n)), dim3(CB), 0, st, a);
    hipLaunchKernelGGL(k_ex_apply, dim3(cblocks(n)), dim3(CB), 0, st, a);
    if (pass == 0) fixed_sum(n, a.e, a.part, a.sums + 0, st);
}
void launch_excess_spread(const ExcessArgs &a, hipStream_t st) {
    const int n = a.g.nxp * a.g.nyp;
    hipLaunchKernelGGL(k_ex_left, dim3(cblocks(n)), dim3(CB), 0, st, a);
    fixed_sum(n, a.room, a.part, a.sums + 1, st);
    fixed_sum(n, a.gain, a.part, a.sums + 2, st);
    hipLaunchKernelGGL(k_ex_spread, dim3(cblocks(n)), dim3(CB), 0, st, a);
}
void launch_excess_caps(const ExcessArgs &a, int under, hipStream_t st) {
    launch_ghosts(a.g, a.c, st);
    hipLaunchKernelGGL(k_ex_caps, dim3(cblocks(a.g.nxp * a.g.nyp)), dim3(CB), 0, st, a.g, a.c, under);
}

}  // namespace yz
