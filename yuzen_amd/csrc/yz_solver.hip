// yz_solver.hip -- BiCGSTAB and structured smoothed-aggregation multigrid
// over stencil operators (yz_stencil.h), FP64, for gfx950.
//
// BiCGSTAB follows the method of van der Vorst (1992) with right
// preconditioning, as the reference's backends run it (Eigen's BiCGSTAB,
// AMGCL's bicgstab: reference yuzen/src/numerics/LinearSolverBackend.cpp:
// 87-111, 662-679): convergence on ||r|| <= tol ||b|| of the row-scaled
// system; breakdown (rho, (rh,v) or omega zero, NaN) returns 2.
//
// The multigrid is smoothed aggregation (Vanek, Mandel & Brezina 1996), the
// method of the reference's AMGCL preconditioner (smoothed_aggregation +
// spai0), laid out for the structured grid: aggregates are the 3x3 blocks
// of cells (3 in 1-D), the tentative prolongator the normalised constant
// on each, smoothed once with damped Jacobi, P = (I - omega D^-1 A) P_t,
// omega = 4/3 / rho(D^-1 A) with rho bounded by Gershgorin; the coarse
// operator is the Galerkin product P^T A P, a 5x5 box stencil on the grid
// of aggregates; SPAI0 (m_i = a_ii / sum_j a_ij^2) smooths before and after
// the coarse correction; the coarsest level (<= 128 cells) is solved with
// its dense inverse. Every level is a set of value planes: no column
// indices, no hashing, no atomics in the set-up.
//
// Reductions are deterministic: fixed grid, per-block partial sums in LDS
// (wave64 shuffles), one finishing block summing the partials in a fixed
// order.
#include "yz_solver.h"

#include <cmath>
#include <cstdio>
#include <cstring>

#include "../../include/yuzen_gpu.h"

namespace yz {

#define YZ_CHECK(x)                                                                 \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_),  \
                         __FILE__, __LINE__);                                       \
            return -1;                                                              \
        }                                                                           \
    } while (0)

static constexpr int BLK = 256;
static constexpr int RED_BLOCKS = 1024;   // partial sums per reduction
static constexpr int COARSE_MAX = 128;   // coarsest dense matrix fits in LDS

static inline int nblk(int n) { return (n + BLK - 1) / BLK; }

// ---------------------------------------------------------------------
// block reduction of m values (wave64)
// ---------------------------------------------------------------------
template <int M>
__device__ inline void block_sum_store(double (&v)[M], double *part, int stride) {
    __shared__ double sh[M][BLK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        double x = v[m];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) sh[m][wave] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int m = 0; m < M; ++m) {
            double s = 0;
            for (int w = 0; w < BLK / 64; ++w) s += sh[m][w];
            part[m * stride + blockIdx.x] = s;
        }
    }
}

enum { FIN_INIT, FIN_ALPHA, FIN_SS, FIN_OMEGA, FIN_NEXT };
__device__ inline void apply_scalars(int op, double *sc, const double *sums);

// sum the partials of quantity m (one block, fixed order) and apply the
// BiCGSTAB scalar step op to them
template <int M>
__global__ void k_finish(const double *part, double *sc, int op) {
    __shared__ double sh[M][BLK];
    for (int m = 0; m < M; ++m) {
        double s = 0;
        for (int b = threadIdx.x; b < RED_BLOCKS; b += BLK) s += part[m * RED_BLOCKS + b];
        sh[m][threadIdx.x] = s;
    }
    __syncthreads();
    for (int w = BLK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int m = 0; m < M; ++m) sh[m][threadIdx.x] += sh[m][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double sums[M];
        for (int m = 0; m < M; ++m) sums[m] = sh[m][0];
        apply_scalars(op, sc, sums);
    }
}

// ---------------------------------------------------------------------
// stencil products
// ---------------------------------------------------------------------
template <int NS, typename T>
__device__ inline double row_dot(const StencilOp &op, const T *__restrict__ val,
                                 const T *__restrict__ x, int i, int ix, int iy) {
    double s = 0;
    const int N = op.N;
    const int rx = op.dim == 2 ? op.nx : 0;
    bool inside = true;
#pragma unroll
    for (int k = 0; k < NS; ++k) inside = inside && stencil_inside(op, ix, iy, k);
    if (inside) {
#pragma unroll
        for (int k = 0; k < NS; ++k) s += val[(size_t)k * N + i] * x[i + op.ox[k] + op.oy[k] * rx];
    } else {
#pragma unroll
        for (int k = 0; k < NS; ++k) s += val[(size_t)k * N + i] * x[stencil_col(op, ix, iy, k)];
    }
    return s;
}

template <int NS, typename T>
__global__ void __launch_bounds__(BLK) k_spmv(StencilOp op, const T *__restrict__ val,
                                              const T *__restrict__ x, T *__restrict__ y) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= op.N) return;
    const int ix = i % op.nx, iy = i / op.nx;
    y[i] = row_dot<NS>(op, val, x, i, ix, iy);
}

template <class F>
static void dispatch_ns(int ns, F f) {
    switch (ns) {
    case 5: f(std::integral_constant<int, 5>()); break;
    case 13: f(std::integral_constant<int, 13>()); break;
    default: f(std::integral_constant<int, 25>()); break;
    }
}


// ---------------------------------------------------------------------
// BiCGSTAB kernels. Device scalars sc[]:
// 0 rho1, 1 rho2, 2 alpha, 3 omega, 4 beta, 5 (rh,v), 6 (s,s), 7 (t,t),
// 8 (t,s), 9 (r,r), 10 (b,b), 11 (r,rh) next
// ---------------------------------------------------------------------
enum { SC_RHO1, SC_RHO2, SC_ALPHA, SC_OMEGA, SC_BETA, SC_RV, SC_SS, SC_TT, SC_TS, SC_RR, SC_BB,
       SC_RRH, SC_N };

template <int NS, typename T>
__global__ void __launch_bounds__(BLK) k_init_residual(StencilOp op, const T *__restrict__ val,
                                                       const T *__restrict__ x, const T *__restrict__ b,
                                                       T *__restrict__ r, T *__restrict__ rh,
                                                       double *__restrict__ part) {
    double acc[2] = {0, 0};
    for (int i = blockIdx.x * BLK + threadIdx.x; i < op.N; i += RED_BLOCKS * BLK) {
        const int ix = i % op.nx, iy = i / op.nx;
        const double ri = b[i] - row_dot<NS>(op, val, x, i, ix, iy);
        r[i] = ri; rh[i] = ri;
        acc[0] += b[i] * b[i];
        acc[1] += ri * ri;
    }
    block_sum_store<2>(acc, part, RED_BLOCKS);
}

template <typename T>
__global__ void k_p_update(int N, int first, const T *__restrict__ r, T *__restrict__ p,
                           const T *__restrict__ v, const double *__restrict__ sc) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= N) return;
    if (first) { p[i] = r[i]; return; }
    const double beta = sc[SC_BETA], omega = sc[SC_OMEGA];
    p[i] = r[i] + beta * (p[i] - omega * v[i]);
}

template <int NS, typename T>
__global__ void __launch_bounds__(BLK) k_spmv_dot(StencilOp op, const T *__restrict__ val,
                                                  const T *__restrict__ x, T *__restrict__ y,
                                                  const T *__restrict__ w, double *__restrict__ part) {
    double acc[1] = {0};
    for (int i = blockIdx.x * BLK + threadIdx.x; i < op.N; i += RED_BLOCKS * BLK) {
        const int ix = i % op.nx, iy = i / op.nx;
        const double yi = row_dot<NS>(op, val, x, i, ix, iy);
        y[i] = yi;
        acc[0] += w[i] * yi;
    }
    block_sum_store<1>(acc, part, RED_BLOCKS);
}

template <typename T>
__global__ void k_s_update(int N, const T *__restrict__ r, const T *__restrict__ v,
                           T *__restrict__ s, const double *__restrict__ sc, double *__restrict__ part) {
    double acc[1] = {0};
    const double alpha = sc[SC_ALPHA];
    for (int i = blockIdx.x * BLK + threadIdx.x; i < N; i += RED_BLOCKS * BLK) {
        const double si = r[i] - alpha * v[i];
        s[i] = si;
        acc[0] += si * si;
    }
    block_sum_store<1>(acc, part, RED_BLOCKS);
}

template <int NS, typename T>
__global__ void __launch_bounds__(BLK) k_spmv_dot2(StencilOp op, const T *__restrict__ val,
                                                   const T *__restrict__ x, T *__restrict__ t,
                                                   const T *__restrict__ s, double *__restrict__ part) {
    double acc[2] = {0, 0};
    for (int i = blockIdx.x * BLK + threadIdx.x; i < op.N; i += RED_BLOCKS * BLK) {
        const int ix = i % op.nx, iy = i / op.nx;
        const double ti = row_dot<NS>(op, val, x, i, ix, iy);
        t[i] = ti;
        acc[0] += ti * ti;
        acc[1] += ti * s[i];
    }
    block_sum_store<2>(acc, part, RED_BLOCKS);
}

template <typename T>
__global__ void k_xr_update(int N, T *__restrict__ x, const T *__restrict__ ph,
                            const T *__restrict__ sh, T *__restrict__ r,
                            const T *__restrict__ s, const T *__restrict__ t,
                            const T *__restrict__ rh, const double *__restrict__ sc,
                            double *__restrict__ part) {
    double acc[2] = {0, 0};
    const double alpha = sc[SC_ALPHA], omega = sc[SC_OMEGA];
    for (int i = blockIdx.x * BLK + threadIdx.x; i < N; i += RED_BLOCKS * BLK) {
        x[i] += alpha * ph[i] + omega * sh[i];
        const double ri = s[i] - omega * t[i];
        r[i] = ri;
        acc[0] += ri * ri;
        acc[1] += ri * rh[i];
    }
    block_sum_store<2>(acc, part, RED_BLOCKS);
}

template <typename T>
__global__ void k_x_alpha(int N, T *__restrict__ x, const T *__restrict__ ph,
                          const double *__restrict__ sc) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= N) return;
    x[i] += sc[SC_ALPHA] * ph[i];
}

// scalar steps of the iteration (thread 0 of the finishing block)
__device__ inline void apply_scalars(int op, double *sc, const double *sums) {
    switch (op) {
    case FIN_INIT:
        sc[SC_BB] = sums[0]; sc[SC_RR] = sums[1]; sc[SC_RHO1] = sums[1];
        sc[SC_RHO2] = 1; sc[SC_ALPHA] = 1; sc[SC_OMEGA] = 1; sc[SC_BETA] = 0;
        break;
    case FIN_ALPHA:
        sc[SC_RV] = sums[0];
        sc[SC_ALPHA] = (sums[0] != 0) ? sc[SC_RHO1] / sums[0] : 0.0;
        break;
    case FIN_SS:
        sc[SC_SS] = sums[0];
        break;
    case FIN_OMEGA:
        sc[SC_TT] = sums[0]; sc[SC_TS] = sums[1];
        sc[SC_OMEGA] = (sums[0] != 0) ? sums[1] / sums[0] : 0.0;
        break;
    default:
        sc[SC_RR] = sums[0];
        sc[SC_RHO2] = sc[SC_RHO1];
        sc[SC_RHO1] = sums[1];
        sc[SC_BETA] = (sc[SC_RHO1] * sc[SC_ALPHA]) / (sc[SC_RHO2] * sc[SC_OMEGA]);
        break;
    }
}

// ---------------------------------------------------------------------
// Jacobi (the reference's damped_jacobi preconditioner: 0.72 D^-1)
// ---------------------------------------------------------------------
template <int NS, typename T>
__global__ void k_diag(StencilOp op, const T *__restrict__ val, T *__restrict__ dg) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= op.N) return;
    const int ix = i % op.nx, iy = i / op.nx;
    double d = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k)
        if (stencil_col(op, ix, iy, k) == i) d += val[(size_t)k * op.N + i];
    dg[i] = d;
}
template <typename T>
__global__ void k_jacobi_inv(int N, T *dg) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= N) return;
    dg[i] = dg[i] != 0 ? 0.72 / dg[i] : 0.0;
}
template <typename T>
__global__ void k_scale(int N, const T *__restrict__ d, const T *__restrict__ b, T *__restrict__ x) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= N) return;
    x[i] = d[i] * b[i];
}

// ---------------------------------------------------------------------
// multigrid set-up
// ---------------------------------------------------------------------
// merge the slots of row i that fold onto the same column; returns count
template <int NS, typename T>
__device__ inline int merged_row(const StencilOp &op, const T *val, int i, int ix, int iy,
                                 int *cols, double *vals) {
    int m = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const double v = val[(size_t)k * op.N + i];
        const int c = stencil_col(op, ix, iy, k);
        int q = 0;
        for (; q < m; ++q) if (cols[q] == c) break;
        if (q == m) { cols[m] = c; vals[m] = v; ++m; }
        else vals[q] += v;
    }
    return m;
}

// SPAI0 diagonal, true diagonal and the Gershgorin bound of D^-1 A
template <int NS, typename T>
__global__ void __launch_bounds__(BLK) k_level_diag(StencilOp op, const T *__restrict__ val, T *__restrict__ minv,
                             T *__restrict__ dg, unsigned long long *__restrict__ rho_bits) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    double g = 0;
    if (i < op.N) {
        const int ix = i % op.nx, iy = i / op.nx;
        double d = 0, ss = 0, off = 0;
        bool inside = true;
#pragma unroll
        for (int k = 0; k < NS; ++k) inside = inside && stencil_inside(op, ix, iy, k);
        if (inside) {
            // every slot its own column (the offsets are distinct)
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                const double v = val[(size_t)k * op.N + i];
                ss += v * v;
                if (op.ox[k] == 0 && op.oy[k] == 0) d = v; else off += fabs(v);
            }
        } else {
            int cols[NS];
            double vals[NS];
            const int m = merged_row<NS>(op, val, i, ix, iy, cols, vals);
            for (int q = 0; q < m; ++q) {
                ss += vals[q] * vals[q];
                if (cols[q] == i) d = vals[q]; else off += fabs(vals[q]);
            }
        }
        minv[i] = (ss > 0) ? d / ss : 0.0;
        dg[i] = d;
        g = (d != 0) ? (fabs(d) + off) / fabs(d) : 0.0;
    }
    // block max, then one atomic per block (non-negative doubles order as
    // their bit patterns)
    __shared__ double sh[BLK / 64];
    for (int off = 32; off > 0; off >>= 1) g = fmax(g, __shfl_down(g, off, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = g;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mx = 0;
        for (int w = 0; w < BLK / 64; ++w) mx = fmax(mx, sh[w]);
        atomicMax(rho_bits, (unsigned long long)__double_as_longlong(mx));
    }
}

__device__ inline int agg_size(int I, int n) {
    const int lo = 3 * I, hi = min(3 * I + 3, n);
    return hi - lo;
}
// coarse offset between aggregates along an axis, wrapped when periodic
__device__ inline int coarse_diff(int J, int I, int nc, int periodic) {
    int d = J - I;
    if (periodic) {
        if (d > nc / 2) d -= nc;
        if (d < -(nc / 2)) d += nc;
    }
    return d;
}

// the offset u = i - 3J of fine cell i from the first cell of aggregate J
// along one axis, wrapped into [-3, 5] when the axis is periodic
__device__ inline int agg_offset(int i, int J, int n, int periodic) {
    int u = i - 3 * J;
    if (periodic) {
        if (u > 5) u -= n;
        if (u < -3) u += n;
    }
    return u;
}

// smoothed prolongator, P(i, agg(i) + e), e in [-1,1]^DIM, written twice:
// per fine cell (P[e][i], for the prolongation) and per coarse cell
// (R[u][J] = P(i, J), u the fine offset in [-3,5]^DIM from aggregate J,
// for the restriction)
template <int NS, int DIM, typename T>
__global__ void __launch_bounds__(BLK) k_build_P(StencilOp op, const T *__restrict__ val, const T *__restrict__ dg,
                          const unsigned long long *__restrict__ rho_bits, int ncx, int ncy,
                          T *__restrict__ P, T *__restrict__ R) {
    constexpr int NP = DIM == 2 ? 9 : 3;
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= op.N) return;
    const int ix = i % op.nx, iy = i / op.nx;
    const double rho = __longlong_as_double((long long)*rho_bits);
    const double omega = (rho > 0) ? (4.0 / 3.0) / rho : 0.0;
    const int Ix = ix / 3, Iy = iy / 3;
    const double w = (dg[i] != 0) ? omega / dg[i] : 0.0;
    int ek[NS];
    double vk[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int c = stencil_col(op, ix, iy, k);
        const int jx = c % op.nx, jy = c / op.nx;
        const int Jx = jx / 3, Jy = jy / 3;
        const int ex = coarse_diff(Jx, Ix, ncx, op.perx);
        const int ey = DIM == 2 ? coarse_diff(Jy, Iy, ncy, op.pery) : 0;
        const double tJ = 1.0 / sqrt((double)(agg_size(Jx, op.nx) * (DIM == 2 ? agg_size(Jy, op.ny) : 1)));
        ek[k] = DIM == 2 ? (ey + 1) * 3 + (ex + 1) : ex + 1;
        vk[k] = w * val[(size_t)k * op.N + i] * tJ;
    }
    const double tI = 1.0 / sqrt((double)(agg_size(Ix, op.nx) * (DIM == 2 ? agg_size(Iy, op.ny) : 1)));
    const int ncN = ncx * (DIM == 2 ? ncy : 1);
#pragma unroll
    for (int e = 0; e < NP; ++e) {
        double pv = (e == NP / 2) ? tI : 0.0;
#pragma unroll
        for (int k = 0; k < NS; ++k) if (ek[k] == e) pv -= vk[k];
        P[(size_t)e * op.N + i] = pv;
        // the coarse cell J = agg(i) + e, and i's offset from it
        const int ex = DIM == 2 ? e % 3 - 1 : e - 1;
        const int ey = DIM == 2 ? e / 3 - 1 : 0;
        int Jx = Ix + ex, Jy = Iy + ey;
        if (op.perx) Jx = (Jx + ncx) % ncx;
        else if (Jx < 0 || Jx >= ncx) continue;
        if (DIM == 2) {
            if (op.pery) Jy = (Jy + ncy) % ncy;
            else if (Jy < 0 || Jy >= ncy) continue;
        }
        const int u = agg_offset(ix, Jx, op.nx, op.perx) + 3;
        const int v = DIM == 2 ? agg_offset(iy, Jy, op.ny, op.pery) + 3 : 0;
        R[(size_t)(v * 9 + u) * ncN + Jy * ncx + Jx] = pv;
    }
}

// A P (i, agg(i) + f), f in [-2,2]^DIM
template <int NS, int DIM, typename T>
__global__ void __launch_bounds__(BLK) k_build_AP(StencilOp op, const T *__restrict__ val, const T *__restrict__ P,
                           int ncx, int ncy, T *__restrict__ AP) {
    constexpr int NF = DIM == 2 ? 25 : 5;
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= op.N) return;
    const int ix = i % op.nx, iy = i / op.nx;
    const int Ix = ix / 3, Iy = iy / 3;
    int dxk[NS], dyk[NS], ck[NS];
    double ak[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int c = stencil_col(op, ix, iy, k);
        ck[k] = c;
        ak[k] = val[(size_t)k * op.N + i];
        dxk[k] = coarse_diff((c % op.nx) / 3, Ix, ncx, op.perx);
        dyk[k] = DIM == 2 ? coarse_diff((c / op.nx) / 3, Iy, ncy, op.pery) : 0;
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const int fx = DIM == 2 ? f % 5 - 2 : f - 2;
        const int fy = DIM == 2 ? f / 5 - 2 : 0;
        double acc = 0;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            const int ex = fx - dxk[k], ey = fy - dyk[k];
            if (ex < -1 || ex > 1 || ey < -1 || ey > 1) continue;
            const int e = DIM == 2 ? (ey + 1) * 3 + (ex + 1) : ex + 1;
            acc += ak[k] * P[(size_t)e * op.N + ck[k]];
        }
        AP[(size_t)f * op.N + i] = acc;
    }
}

// Galerkin product on the coarse grid: Ac(I, I + q), q in [-2,2]^DIM
template <int DIM, typename T>
__global__ void __launch_bounds__(BLK) k_build_RAP(StencilOp fop, const T *__restrict__ P, const T *__restrict__ AP,
                            StencilOp cop, T *__restrict__ Ac) {
    constexpr int NQ = DIM == 2 ? 25 : 5;
    const int I = blockIdx.x * BLK + threadIdx.x;
    if (I >= cop.N) return;
    const int Ix = I % cop.nx, Iy = I / cop.nx;
    double ac[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) ac[q] = 0;
#pragma unroll
    for (int ey = (DIM == 2 ? -1 : 0); ey <= (DIM == 2 ? 1 : 0); ++ey) {
#pragma unroll
        for (int ex = -1; ex <= 1; ++ex) {
            int Kx = Ix + ex, Ky = Iy + ey;
            bool ok = true;
            if (cop.perx) Kx = (Kx + cop.nx) % cop.nx;
            else ok = ok && Kx >= 0 && Kx < cop.nx;
            if (DIM == 2) {
                if (cop.pery) Ky = (Ky + cop.ny) % cop.ny;
                else ok = ok && Ky >= 0 && Ky < cop.ny;
            }
            if (!ok) continue;
            // fine cells i of aggregate K; P(i, I) is the slot I - agg(i) = -e
            const int pe = DIM == 2 ? (1 - ey) * 3 + (1 - ex) : (1 - ex);
            const int x0 = 3 * Kx, x1 = min(3 * Kx + 3, fop.nx);
            const int y0 = DIM == 2 ? 3 * Ky : 0, y1 = DIM == 2 ? min(3 * Ky + 3, fop.ny) : 1;
            for (int fy = y0; fy < y1; ++fy)
                for (int fx = x0; fx < x1; ++fx) {
                    const int i = fy * fop.nx + fx;
                    const double pv = P[(size_t)pe * fop.N + i];
                    // AP(i, K + g) -> Ac(I, I + e + g)
#pragma unroll
                    for (int gy = (DIM == 2 ? -2 : 0); gy <= (DIM == 2 ? 2 : 0); ++gy) {
#pragma unroll
                        for (int gx = -2; gx <= 2; ++gx) {
                            const int qx = ex + gx, qy = ey + gy;
                            if (qx < -2 || qx > 2 || qy < -2 || qy > 2) continue;
                            const int g = DIM == 2 ? (gy + 2) * 5 + (gx + 2) : gx + 2;
                            const int q = DIM == 2 ? (qy + 2) * 5 + (qx + 2) : qx + 2;
                            ac[q] += pv * AP[(size_t)g * fop.N + i];
                        }
                    }
                }
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) Ac[(size_t)q * cop.N + I] = ac[q];
}

// dense coarsest matrix
template <int NS, typename T>
__global__ void k_dense(StencilOp op, const T *__restrict__ val, double *__restrict__ Dn) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= op.N) return;
    const int ix = i % op.nx, iy = i / op.nx;
    for (int j = 0; j < op.N; ++j) Dn[(size_t)i * op.N + j] = 0;
    for (int k = 0; k < NS; ++k) Dn[(size_t)i * op.N + stencil_col(op, ix, iy, k)] += val[(size_t)k * op.N + i];
}

// In-place Gauss-Jordan inversion with row pivoting, one block of 1024
// threads. M (n x n, row-major) is the working matrix: in LDS when
// n <= 128 (128 KB), else the global scratch it is given. The row swaps of
// the elimination are undone as column swaps at the end. The inverse is
// written transposed (invT[j n + i] = inv(i, j)) so that the coarse
// product reads it coalesced.
template <bool SHARED>
__device__ void gj_body(int n, double *M, int *perm, double *red_v, int *red_i, double *invT) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int k = 0; k < n; ++k) {
        double best = -1; int br = k;
        for (int r = k + tid; r < n; r += nt) {
            const double a = fabs(M[r * n + k]);
            if (a > best) { best = a; br = r; }
        }
        red_v[tid] = best; red_i[tid] = br;
        __syncthreads();
        for (int w = nt / 2; w > 0; w >>= 1) {
            if (tid < w && (red_v[tid + w] > red_v[tid] || (red_v[tid + w] == red_v[tid] && red_i[tid + w] < red_i[tid]))) {
                red_v[tid] = red_v[tid + w]; red_i[tid] = red_i[tid + w];
            }
            __syncthreads();
        }
        const int pr = red_i[0];
        if (tid == 0) perm[k] = pr;
        if (pr != k)
            for (int j = tid; j < n; j += nt) { const double t = M[k * n + j]; M[k * n + j] = M[pr * n + j]; M[pr * n + j] = t; }
        __syncthreads();
        const double piv = M[k * n + k];
        const double pinv = (piv != 0) ? 1.0 / piv : 0.0;
        __syncthreads();
        for (int j = tid; j < n; j += nt) M[k * n + j] = (j == k) ? pinv : M[k * n + j] * pinv;
        __syncthreads();
        // eliminate column k from the other rows; f is read before any write
        for (int t = tid; t < n * n; t += nt) {
            const int r = t / n, j = t % n;
            if (r == k || j == k) continue;
            M[r * n + j] -= M[r * n + k] * M[k * n + j];
        }
        __syncthreads();
        for (int r = tid; r < n; r += nt)
            if (r != k) M[r * n + k] = -M[r * n + k] * M[k * n + k];
        __syncthreads();
    }
    // undo the pivoting as column swaps, last first
    for (int k = n - 1; k >= 0; --k) {
        const int pr = perm[k];
        if (pr != k)
            for (int r = tid; r < n; r += nt) { const double t = M[r * n + k]; M[r * n + k] = M[r * n + pr]; M[r * n + pr] = t; }
        __syncthreads();
    }
    for (int t = tid; t < n * n; t += nt) {
        const int i = t / n, j = t % n;
        invT[j * n + i] = M[i * n + j];
    }
}

__global__ void __launch_bounds__(1024) k_gauss_jordan(int n, double *Ag, double *invT) {
    __shared__ double red_v[1024];
    __shared__ int red_i[1024];
    __shared__ int perm[1024];
    if (n <= 128) {
        __shared__ double Ms[128 * 128];
        for (int t = threadIdx.x; t < n * n; t += blockDim.x) Ms[t] = Ag[t];
        __syncthreads();
        gj_body<true>(n, Ms, perm, red_v, red_i, invT);
    } else {
        gj_body<false>(n, Ag, perm, red_v, red_i, invT);
    }
}

// ---------------------------------------------------------------------
// V-cycle kernels
// ---------------------------------------------------------------------
// pre-smoothing from zero and the residual: x = M b, r = b - A x
template <int NS, typename T>
__global__ void __launch_bounds__(BLK) k_pre(StencilOp op, const T *__restrict__ val,
                                             const T *__restrict__ m, const T *__restrict__ b,
                                             T *__restrict__ x, T *__restrict__ r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= op.N) return;
    const int ix = i % op.nx, iy = i / op.nx;
    double s = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int c = stencil_col(op, ix, iy, k);
        s += val[(size_t)k * op.N + i] * (m[c] * b[c]);
    }
    x[i] = m[i] * b[i];
    r[i] = b[i] - s;
}

// restriction b_c = P^T r from the per-coarse-cell weights R: a block of
// 16 x 16 coarse cells (256 in 1-D) stages the fine residual it needs
// (a 56 x 56 tile; 776 cells in 1-D) in LDS, coalesced; each thread then
// sums its 81 (9) weights, read coalesced, against the tile
// TC x TC coarse cells per block in 2-D: 16 x 16 on the large levels, 8 x 8
// (one wave) on the small ones, so that a coarse level of a few thousand
// cells still spreads over a hundred CUs
template <int DIM, typename T, int TC = 16>
__global__ void __launch_bounds__(256) k_restrict(StencilOp fop, const T *__restrict__ R,
                                                  const T *__restrict__ r, StencilOp cop,
                                                  T *__restrict__ bc) {
    constexpr int TX = DIM == 2 ? TC : 256, TY = DIM == 2 ? TC : 1;
    constexpr int FX = 3 * TX + 8, FY = DIM == 2 ? 3 * TY + 8 : 1;
    __shared__ double tile[FY * FX];
    const int X0 = blockIdx.x * TX, Y0 = DIM == 2 ? blockIdx.y * TY : 0;
    const int fx0 = 3 * X0 - 3, fy0 = DIM == 2 ? 3 * Y0 - 3 : 0;
    for (int t = threadIdx.x; t < FX * FY; t += TX * TY) {
        const int lx = t % FX, ly = t / FX;
        int gx = fx0 + lx, gy = fy0 + ly;
        bool ok = true;
        if (fop.perx) gx = ((gx % fop.nx) + fop.nx) % fop.nx;
        else ok = ok && gx >= 0 && gx < fop.nx;
        if (DIM == 2) {
            if (fop.pery) gy = ((gy % fop.ny) + fop.ny) % fop.ny;
            else ok = ok && gy >= 0 && gy < fop.ny;
        }
        tile[t] = ok ? r[gy * fop.nx + gx] : 0.0;
    }
    __syncthreads();
    const int tx = DIM == 2 ? threadIdx.x % TX : threadIdx.x, ty = DIM == 2 ? threadIdx.x / TX : 0;
    const int Ix = X0 + tx, Iy = Y0 + ty;
    if (Ix >= cop.nx || (DIM == 2 && Iy >= cop.ny)) return;
    const int I = Iy * cop.nx + Ix;
    double s = 0;
    if (DIM == 2) {
#pragma unroll
        for (int v = 0; v < 9; ++v)
#pragma unroll
            for (int u = 0; u < 9; ++u)
                s += R[(size_t)(v * 9 + u) * cop.N + I] * tile[(3 * ty + v) * FX + 3 * tx + u];
    } else {
#pragma unroll
        for (int u = 0; u < 9; ++u) s += R[(size_t)u * cop.N + I] * tile[3 * tx + u];
    }
    bc[I] = s;
}

// prolongation x += P x_c
template <typename T>
__global__ void k_prolong(StencilOp fop, const T *__restrict__ P, StencilOp cop,
                          const T *__restrict__ xc, T *__restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fop.N) return;
    const bool two = fop.dim == 2;
    const int ix = i % fop.nx, iy = i / fop.nx;
    const int Ix = ix / 3, Iy = iy / 3;
    double s = 0;
    const int ylo = two ? -1 : 0, yhi = two ? 1 : 0;
    for (int ey = ylo; ey <= yhi; ++ey) {
        int Jy = Iy + ey;
        if (two) {
            if (cop.pery) Jy = (Jy + cop.ny) % cop.ny;
            else if (Jy < 0 || Jy >= cop.ny) continue;
        }
        for (int ex = -1; ex <= 1; ++ex) {
            int Jx = Ix + ex;
            if (cop.perx) Jx = (Jx + cop.nx) % cop.nx;
            else if (Jx < 0 || Jx >= cop.nx) continue;
            const int e = two ? (ey + 1) * 3 + (ex + 1) : ex + 1;
            s += P[(size_t)e * fop.N + i] * xc[Jy * cop.nx + Jx];
        }
    }
    x[i] += s;
}

// post-smoothing, out of place: y = x + M (b - A x)
template <int NS, typename T>
__global__ void __launch_bounds__(BLK) k_post(StencilOp op, const T *__restrict__ val,
                                              const T *__restrict__ m, const T *__restrict__ b,
                                              const T *__restrict__ x, T *__restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= op.N) return;
    const int ix = i % op.nx, iy = i / op.nx;
    const double ax = row_dot<NS>(op, val, x, i, ix, iy);
    y[i] = x[i] + m[i] * (b[i] - ax);
}

// The coarse levels of a 2-D hierarchy in two kernels per visit instead of
// four. A level of a few thousand cells costs each kernel its launch ramp
// and a memory round trip or two whatever its size, so the down-sweep's
// smoothing, residual and restriction run as one kernel per tile of coarse
// cells and the up-sweep's prolongation and smoothing as one per tile of
// fine cells, each recomputing the few halo values it needs from its
// neighbours' inputs. Every value is formed by the same arithmetic, in the
// same order, as the separate kernels (k_pre, k_restrict, k_prolong,
// k_post), so the cycle's output is the same to the bit.
//
// Down-sweep of a TC x TC tile of coarse cells: x = M b on the fine window
// of the tile's restriction plus the stencil's 2-cell reach (staged in LDS
// at unfolded positions, each holding the folded cell's value, so the
// stencil reads them without folding), the residual r = b - A x on the
// restriction's window (3 TC + 6 fine cells per axis; zero off a
// non-periodic grid), b_c = R r, and x written for the tile's own fine cells
// (the up-sweep smooths from it). On these levels a kernel's time is its
// chain of memory round trips, so each thread issues all the loads of a
// phase (K1 window values, G2 residual rows at a time) before it uses any.
template <int NS, typename T, int TC, int NT>
__global__ void __launch_bounds__(NT) k_down2(StencilOp op, const T *__restrict__ val, const T *__restrict__ m,
                                              const T *__restrict__ b, T *__restrict__ x,
                                              const T *__restrict__ R, StencilOp cop, T *__restrict__ bc) {
    constexpr int FW = 3 * TC + 6, XW = FW + 4;
    constexpr int K1 = (XW * XW + NT - 1) / NT, K2 = (FW * FW + NT - 1) / NT, G2 = 2;
    // values as the separate kernels store them: x and r in T
    __shared__ T xs[XW * XW];
    __shared__ double rs[FW * FW];
    const int X0 = blockIdx.x * TC, Y0 = blockIdx.y * TC;
    const int fx0 = 3 * X0 - 3, fy0 = 3 * Y0 - 3;
    {
        T mv[K1], bv[K1];
#pragma unroll
        for (int q = 0; q < K1; ++q) {
            const int t = threadIdx.x + q * NT;
            mv[q] = 0; bv[q] = 0;
            if (t < XW * XW) {
                const int gx = fold_axis(fx0 - 2 + t % XW, op.nx, op.perx);
                const int gy = fold_axis(fy0 - 2 + t / XW, op.ny, op.pery);
                const int c = gy * op.nx + gx;
                mv[q] = m[c]; bv[q] = b[c];
            }
        }
#pragma unroll
        for (int q = 0; q < K1; ++q) {
            const int t = threadIdx.x + q * NT;
            if (t < XW * XW) xs[t] = mv[q] * bv[q];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q0 = 0; q0 < K2; q0 += G2) {
        T av[G2][NS], bi[G2];
        int ii[G2];
#pragma unroll
        for (int g = 0; g < G2; ++g) {
            const int t = threadIdx.x + (q0 + g) * NT;
            ii[g] = -1;
            if (q0 + g < K2 && t < FW * FW) {
                int gx = fx0 + t % FW, gy = fy0 + t / FW;
                bool ok = true;
                if (op.perx) gx = ((gx % op.nx) + op.nx) % op.nx;
                else ok = ok && gx >= 0 && gx < op.nx;
                if (op.pery) gy = ((gy % op.ny) + op.ny) % op.ny;
                else ok = ok && gy >= 0 && gy < op.ny;
                if (ok) {
                    const int i = gy * op.nx + gx;
                    ii[g] = i;
#pragma unroll
                    for (int k = 0; k < NS; ++k) av[g][k] = val[(size_t)k * op.N + i];
                    bi[g] = b[i];
                }
            }
        }
#pragma unroll
        for (int g = 0; g < G2; ++g) {
            const int t = threadIdx.x + (q0 + g) * NT;
            if (q0 + g < K2 && t < FW * FW) {
                const int lx = t % FW, ly = t / FW;
                double rv = 0.0;
                if (ii[g] >= 0) {
                    double s = 0;
#pragma unroll
                    for (int k = 0; k < NS; ++k) s += av[g][k] * xs[(ly + 2 + op.oy[k]) * XW + lx + 2 + op.ox[k]];
                    rv = (T)(bi[g] - s);
                }
                rs[t] = rv;
            }
        }
    }
    for (int t = threadIdx.x; t < 9 * TC * TC; t += NT) {
        const int ux = t % (3 * TC), uy = t / (3 * TC);
        const int gx = 3 * X0 + ux, gy = 3 * Y0 + uy;
        if (gx < op.nx && gy < op.ny) x[gy * op.nx + gx] = xs[(uy + 5) * XW + ux + 5];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < TC * TC; t += NT) {
        const int tx = t % TC, ty = t / TC;
        const int Ix = X0 + tx, Iy = Y0 + ty;
        if (Ix >= cop.nx || Iy >= cop.ny) continue;
        const int I = Iy * cop.nx + Ix;
        double s = 0;
#pragma unroll
        for (int v = 0; v < 9; ++v)
#pragma unroll
            for (int u = 0; u < 9; ++u) s += R[(size_t)(v * 9 + u) * cop.N + I] * rs[(3 * ty + v) * FW + 3 * tx + u];
        bc[I] = s;
    }
}

// Up-sweep of a TF x TF tile of fine cells: x' = x + P x_c on the tile and
// its 2-cell halo (staged in LDS as above, every load of a thread's window
// values issued before their sums), then y = x' + M (b - A x') for the
// tile's cells.
template <int NS, typename T, int TF, int NT>
__global__ void __launch_bounds__(NT) k_up2(StencilOp op, const T *__restrict__ val, const T *__restrict__ m,
                                            const T *__restrict__ b, const T *__restrict__ x,
                                            const T *__restrict__ P, StencilOp cop, const T *__restrict__ xc,
                                            T *__restrict__ y) {
    constexpr int W = TF + 4;
    constexpr int K1 = (W * W + NT - 1) / NT;
    __shared__ T xs[W * W];   // x' as k_prolong stores it, in T
    const int X0 = blockIdx.x * TF, Y0 = blockIdx.y * TF;
    {
        T pv[K1][9], cv[K1][9], xv[K1];
#pragma unroll
        for (int q = 0; q < K1; ++q) {
            const int t = threadIdx.x + q * NT;
            xv[q] = 0;
#pragma unroll
            for (int e = 0; e < 9; ++e) { pv[q][e] = 0; cv[q][e] = 0; }
            if (t < W * W) {
                const int gx = fold_axis(X0 - 2 + t % W, op.nx, op.perx);
                const int gy = fold_axis(Y0 - 2 + t / W, op.ny, op.pery);
                const int i = gy * op.nx + gx;
                const int Ix = gx / 3, Iy = gy / 3;
                xv[q] = x[i];
#pragma unroll
                for (int e = 0; e < 9; ++e) {
                    int Jx = Ix + e % 3 - 1, Jy = Iy + e / 3 - 1;
                    bool ok = true;
                    if (cop.perx) Jx = (Jx + cop.nx) % cop.nx;
                    else ok = ok && Jx >= 0 && Jx < cop.nx;
                    if (cop.pery) Jy = (Jy + cop.ny) % cop.ny;
                    else ok = ok && Jy >= 0 && Jy < cop.ny;
                    if (ok) {
                        pv[q][e] = P[(size_t)e * op.N + i];
                        cv[q][e] = xc[Jy * cop.nx + Jx];
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < K1; ++q) {
            const int t = threadIdx.x + q * NT;
            if (t < W * W) {
                // the coarse neighbours off a non-periodic grid hold zeros,
                // and P x_c sums the others in k_prolong's order
                double s = 0;
                const int gx = fold_axis(X0 - 2 + t % W, op.nx, op.perx);
                const int gy = fold_axis(Y0 - 2 + t / W, op.ny, op.pery);
                const int Ix = gx / 3, Iy = gy / 3;
#pragma unroll
                for (int e = 0; e < 9; ++e) {
                    const int Jx = Ix + e % 3 - 1, Jy = Iy + e / 3 - 1;
                    const bool in = (cop.perx || (Jx >= 0 && Jx < cop.nx)) && (cop.pery || (Jy >= 0 && Jy < cop.ny));
                    if (in) s += pv[q][e] * cv[q][e];
                }
                xs[t] = (T)(xv[q] + s);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x >= TF * TF) return;
    const int tx = threadIdx.x % TF, ty = threadIdx.x / TF;
    const int ix = X0 + tx, iy = Y0 + ty;
    if (ix >= op.nx || iy >= op.ny) return;
    const int i = iy * op.nx + ix;
    double ax = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) ax += val[(size_t)k * op.N + i] * xs[(ty + 2 + op.oy[k]) * W + tx + 2 + op.ox[k]];
    const double xi = xs[(ty + 2) * W + tx + 2];
    y[i] = xi + m[i] * (b[i] - ax);
}

// x = A^-1 b with the inverse stored transposed (coalesced over i): 64 rows
// per block, the block's DMV_W waves each take every DMV_W-th column, and
// their partial sums are added in wave order (a fixed order, so the result
// does not depend on scheduling). The coarsest level has a few hundred rows:
// one thread per row walking all columns left most of the chip idle.
static constexpr int DMV_W = 16;
template <typename T>
__global__ void __launch_bounds__(64 * DMV_W) k_dense_mv(int n, const double *__restrict__ AinvT,
                                                         const T *__restrict__ b, T *__restrict__ x) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    double s = 0;
    if (i < n)
        for (int j = wave; j < n; j += DMV_W) s += AinvT[(size_t)j * n + i] * b[j];
    __shared__ double part[DMV_W][64];
    part[wave][lane] = s;
    __syncthreads();
    if (wave == 0 && i < n) {
        double t = 0;
        for (int w = 0; w < DMV_W; ++w) t += part[w][lane];
        x[i] = t;
    }
}
static inline int dmv_blocks(int n) { return (n + 63) / 64; }
// the block size of a level's smoothing and prolongation: one wave on a
// level of fewer than 128 x 128 cells, so that its few thousand cells
// spread over more CUs than 256-thread blocks would reach
static inline int lvl_blk(int n) { return n < 128 * 128 ? 64 : BLK; }
static inline int lvl_grid(int n) { return (n + lvl_blk(n) - 1) / lvl_blk(n); }

// residual r = b - A x, and the explicit-stencil smoothing step
// out = xbase + M t (the learned approximate inverse as the finest smoother)
template <int NS, typename T>
__global__ void __launch_bounds__(BLK) k_resid(StencilOp op, const T *__restrict__ val, const T *__restrict__ x,
                                               const T *__restrict__ b, T *__restrict__ r) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= op.N) return;
    const int ix = i % op.nx, iy = i / op.nx;
    r[i] = b[i] - row_dot<NS>(op, val, x, i, ix, iy);
}
template <int NS, typename T>
__global__ void __launch_bounds__(BLK) k_spmv_add(StencilOp op, const T *__restrict__ val, const T *__restrict__ t,
                                                  const T *__restrict__ xbase, T *__restrict__ out) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= op.N) return;
    const int ix = i % op.nx, iy = i / op.nx;
    out[i] = xbase[i] + row_dot<NS>(op, val, t, i, ix, iy);
}

// ---------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------
static DevArena *g_arena = nullptr;
void solver_use_arena(DevArena *a) { g_arena = a; }
void *arena_carve(size_t n) {
    if (!g_arena) return nullptr;
    const size_t off = (g_arena->used + 255) & ~(size_t)255;
    if (off > g_arena->cap || n > g_arena->cap - off) { g_arena->overflow = true; return nullptr; }
    g_arena->used = off + n;
    return g_arena->base + off;
}
// the host copy of the scalars for a solver in an arena (no pinned
// allocation; the device-to-host copies of read_scalars then go through
// the runtime's staging)
static double g_arena_host_sc[64];

template <typename T>
static int dalloc(T **p, size_t n) {
    if (*p) return 0;
    if (g_arena) {
        *p = (T *)arena_carve(n * sizeof(T));
        if (!*p) {
            return -1;
        }
        return 0;
    }
    YZ_CHECK(hipMalloc((void **)p, n * sizeof(T)));
    return 0;
}
static bool same_op(const StencilOp &a, const StencilOp &b) {
    if (a.dim != b.dim || a.nx != b.nx || a.ny != b.ny || a.N != b.N || a.ns != b.ns || a.perx != b.perx ||
        a.pery != b.pery)
        return false;
    for (int k = 0; k < a.ns; ++k)
        if (a.ox[k] != b.ox[k] || a.oy[k] != b.oy[k]) return false;
    return true;
}

static StencilOp box_op(int dim, int nx, int ny, int perx, int pery) {
    StencilOp op;
    std::memset(&op, 0, sizeof(op));
    op.dim = dim; op.nx = nx; op.ny = dim == 2 ? ny : 1; op.N = op.nx * op.ny;
    op.perx = perx; op.pery = pery;
    if (dim == 2) {
        op.ns = 25;
        for (int q = 0; q < 25; ++q) { op.ox[q] = q % 5 - 2; op.oy[q] = q / 5 - 2; }
    } else {
        op.ns = 5;
        for (int q = 0; q < 5; ++q) { op.ox[q] = q - 2; op.oy[q] = 0; }
    }
    return op;
}

template <typename T>
void launch_spmv(const StencilOp &op, const T *val, const T *x, T *y, hipStream_t st) {
    dispatch_ns(op.ns, [&](auto c) {
        constexpr int NS = decltype(c)::value;
        hipLaunchKernelGGL((k_spmv<NS, T>), dim3(nblk(op.N)), dim3(BLK), 0, st, op, val, x, y);
    });
}

template <typename T>
int solver_init(SolverT<T> &S, const StencilOp &op0, hipStream_t stream) {
    // the hierarchy's shapes hang on the operator's grid, edges and slots:
    // reused only for the same ones (a solver in an arena is made afresh)
    if (S.kr.N == op0.N && !S.in_arena && !g_arena && same_op(S.op0, op0)) { S.stream = stream; return 0; }
    solver_free(S);
    S.op0 = op0;
    S.stream = stream;
    S.in_arena = g_arena != nullptr;
    const int N = op0.N;
    if (dalloc(&S.A, (size_t)op0.ns * N)) return -1;
    Krylov<T> &k = S.kr;
    k.N = N;
    T **vs[] = {&k.r, &k.rh, &k.p, &k.v, &k.s, &k.t, &k.ph, &k.sh};
    for (T **v : vs) if (dalloc(v, N)) return -1;
    if (dalloc(&k.part, 4 * RED_BLOCKS)) return -1;
    if (dalloc(&k.sc, 64)) return -1;
    YZ_CHECK(hipMemset(k.sc, 0, 64 * sizeof(double)));
    if (S.in_arena) k.host_sc = g_arena_host_sc;
    else YZ_CHECK(hipHostMalloc((void **)&k.host_sc, 64 * sizeof(double), hipHostMallocDefault));
    // multigrid hierarchy shapes (fixed for the grid)
    Amg<T> &g = S.amg;
    g.lv.clear();
    AmgLevel<T> L0;
    L0.op = op0;
    g.lv.push_back(L0);
    while ((int)g.lv.size() < S.max_levels) {
        AmgLevel<T> &L = g.lv.back();
        if (L.op.N <= COARSE_MAX) break;
        const int ncx = (L.op.nx + 2) / 3;
        const int ncy = L.op.dim == 2 ? (L.op.ny + 2) / 3 : 1;
        if (L.op.perx && ncx < 5) break;
        if (L.op.dim == 2 && L.op.pery && ncy < 5) break;
        if (ncx * ncy >= L.op.N) break;
        L.ncx = ncx; L.ncy = ncy;
        AmgLevel<T> C;
        C.op = box_op(L.op.dim, ncx, ncy, L.op.perx, L.op.pery);
        g.lv.push_back(C);
    }
    for (size_t l = 0; l < g.lv.size(); ++l) {
        AmgLevel<T> &L = g.lv[l];
        const int n = L.op.N;
        if (l > 0 && dalloc(&L.val, (size_t)L.op.ns * n)) return -1;
        if (dalloc(&L.minv, n) || dalloc(&L.x, n) || dalloc(&L.b, n) || dalloc(&L.r, n) || dalloc(&L.t, n)) return -1;
        if (l + 1 < g.lv.size()) {
            const int np = L.op.dim == 2 ? 9 : 3, nf = L.op.dim == 2 ? 25 : 5;
            const int nr = L.op.dim == 2 ? 81 : 9;
            if (dalloc(&L.P, (size_t)np * n) || dalloc(&L.AP, (size_t)nf * n)) return -1;
            const size_t rn = (size_t)nr * g.lv[l + 1].op.N;
            if (dalloc(&L.R, rn)) return -1;
            // weights of fine cells outside the grid are never written
            YZ_CHECK(hipMemset(L.R, 0, rn * sizeof(T)));
        }
    }
    g.lv[0].val = S.A;
    g.nco = g.lv.back().op.N;
    if (dalloc(&g.coarse_inv, (size_t)g.nco * g.nco) || dalloc(&g.coarse_dense, (size_t)g.nco * g.nco)) return -1;
    if (dalloc(&g.scal, 2 * S.max_levels)) return -1;
    if (dalloc(&S.dinv_jac, N)) return -1;
    return 0;
}

template <typename T>
void solver_free(SolverT<T> &S) {
    const bool own = !S.in_arena;
    auto f = [own](auto *&p) { if (p && own) (void)hipFree((void *)p); p = nullptr; };
    for (size_t l = 0; l < S.amg.lv.size(); ++l) {
        AmgLevel<T> &L = S.amg.lv[l];
        if (l > 0) f(L.val);
        f(L.minv); f(L.P); f(L.R); f(L.AP); f(L.x); f(L.b); f(L.r); f(L.t);
    }
    S.amg.lv.clear();
    f(S.amg.coarse_inv); f(S.amg.coarse_dense); f(S.amg.scal);
    f(S.A); f(S.dinv_jac);
    Krylov<T> &k = S.kr;
    f(k.r); f(k.rh); f(k.p); f(k.v); f(k.s); f(k.t); f(k.ph); f(k.sh); f(k.part); f(k.sc);
    f(k.p2); f(k.fpart);
    k.ftiles = 0;
    fused_table_free(k.table);
    for (auto &e : S.kev) if (e) { (void)hipEventDestroy(e); e = nullptr; }
    if (S.kts) { (void)hipFree(S.kts); S.kts = nullptr; }
    if (k.host_sc && own) (void)hipHostFree(k.host_sc);
    k.host_sc = nullptr;
    k.N = 0;
    S.in_arena = 0;
}

template <typename T>
int amg_setup(SolverT<T> &S) {
    Amg<T> &g = S.amg;
    hipStream_t st = S.stream;
    g.lv[0].val = S.A;
    YZ_CHECK(hipMemsetAsync(g.scal, 0, 2 * S.max_levels * sizeof(double), st));
    for (size_t l = 0; l < g.lv.size(); ++l) {
        AmgLevel<T> &L = g.lv[l];
        unsigned long long *rho = reinterpret_cast<unsigned long long *>(g.scal) + l;
        dispatch_ns(L.op.ns, [&](auto c) {
            constexpr int NS = decltype(c)::value;
            // L.r holds the true diagonal during the set-up
            hipLaunchKernelGGL((k_level_diag<NS, T>), dim3(nblk(L.op.N)), dim3(BLK), 0, st, L.op, L.val, L.minv,
                               L.r, rho);
        });
        if (l + 1 == g.lv.size()) break;
        AmgLevel<T> &C = g.lv[l + 1];
        dispatch_ns(L.op.ns, [&](auto c) {
            constexpr int NS = decltype(c)::value;
            if (L.op.dim == 2) {
                hipLaunchKernelGGL((k_build_P<NS, 2, T>), dim3(nblk(L.op.N)), dim3(BLK), 0, st, L.op, L.val, L.r,
                                   rho, L.ncx, L.ncy, L.P, L.R);
                hipLaunchKernelGGL((k_build_AP<NS, 2, T>), dim3(nblk(L.op.N)), dim3(BLK), 0, st, L.op, L.val, L.P,
                                   L.ncx, L.ncy, L.AP);
            } else {
                hipLaunchKernelGGL((k_build_P<NS, 1, T>), dim3(nblk(L.op.N)), dim3(BLK), 0, st, L.op, L.val, L.r,
                                   rho, L.ncx, L.ncy, L.P, L.R);
                hipLaunchKernelGGL((k_build_AP<NS, 1, T>), dim3(nblk(L.op.N)), dim3(BLK), 0, st, L.op, L.val, L.P,
                                   L.ncx, L.ncy, L.AP);
            }
        });
        if (L.op.dim == 2)
            hipLaunchKernelGGL((k_build_RAP<2, T>), dim3(nblk(C.op.N)), dim3(BLK), 0, st, L.op, L.P, L.AP, C.op,
                               C.val);
        else
            hipLaunchKernelGGL((k_build_RAP<1, T>), dim3(nblk(C.op.N)), dim3(BLK), 0, st, L.op, L.P, L.AP, C.op,
                               C.val);
    }
    AmgLevel<T> &Lc = g.lv.back();
    dispatch_ns(Lc.op.ns, [&](auto c) {
        constexpr int NS = decltype(c)::value;
        hipLaunchKernelGGL((k_dense<NS, T>), dim3(nblk(Lc.op.N)), dim3(BLK), 0, st, Lc.op, Lc.val, g.coarse_dense);
    });
    hipLaunchKernelGGL(k_gauss_jordan, dim3(1), dim3(1024), 0, st, g.nco, g.coarse_dense, g.coarse_inv);
    YZ_CHECK(hipGetLastError());
    g.built = 1;
    return 0;
}

template <typename T>
int jacobi_setup(SolverT<T> &S) {
    dispatch_ns(S.op0.ns, [&](auto c) {
        constexpr int NS = decltype(c)::value;
        hipLaunchKernelGGL((k_diag<NS, T>), dim3(nblk(S.op0.N)), dim3(BLK), 0, S.stream, S.op0, S.A, S.dinv_jac);
    });
    hipLaunchKernelGGL((k_jacobi_inv<T>), dim3(nblk(S.op0.N)), dim3(BLK), 0, S.stream, S.op0.N, S.dinv_jac);
    YZ_CHECK(hipGetLastError());
    return 0;
}

// whether level l of the cycle runs as the two fused kernels (k_down2,
// k_up2): the coarse levels of a 2-D hierarchy. The finest level keeps the
// separate kernels: its halo recomputation would re-read about a quarter
// more of the largest operator for one launch saved.
template <typename T>
static inline bool fused_level(const AmgLevel<T> &L, int l) {
    return l >= 1 && L.op.dim == 2;
}

// one V-cycle: out = M^-1 in; with smoother != null the finest level is
// smoothed with that explicit stencil instead of SPAI0
template <typename T>
static void vcycle(SolverT<T> &S, const T *in, T *out, const T *smoother) {
    Amg<T> &g = S.amg;
    hipStream_t st = S.stream;
    const int nl = (int)g.lv.size();
    if (nl == 1) {
        hipLaunchKernelGGL((k_dense_mv<T>), dim3(dmv_blocks(g.nco)), dim3(64 * DMV_W), 0, st, g.nco, g.coarse_inv, in, out);
        return;
    }
    // downward
    const T *b = in;
    for (int l = 0; l + 1 < nl; ++l) {
        AmgLevel<T> &L = g.lv[l];
        AmgLevel<T> &C = g.lv[l + 1];
        if (fused_level(L, l)) {
            dispatch_ns(L.op.ns, [&](auto c) {
                constexpr int NS = decltype(c)::value;
                if (C.op.N < 128 * 128)
                    hipLaunchKernelGGL((k_down2<NS, T, 8, 512>), dim3((C.op.nx + 7) / 8, (C.op.ny + 7) / 8), dim3(512),
                                       0, st, L.op, L.val, L.minv, b, L.x, L.R, C.op, C.b);
                else
                    hipLaunchKernelGGL((k_down2<NS, T, 16, 512>), dim3((C.op.nx + 15) / 16, (C.op.ny + 15) / 16),
                                       dim3(512), 0, st, L.op, L.val, L.minv, b, L.x, L.R, C.op, C.b);
            });
            b = C.b;
            continue;
        }
        if (l == 0 && smoother) {
            launch_spmv(L.op, smoother, b, L.x, st);
            dispatch_ns(L.op.ns, [&](auto c) {
                constexpr int NS = decltype(c)::value;
                hipLaunchKernelGGL((k_resid<NS, T>), dim3(nblk(L.op.N)), dim3(BLK), 0, st, L.op, L.val, L.x, b, L.r);
            });
        } else {
            dispatch_ns(L.op.ns, [&](auto c) {
                constexpr int NS = decltype(c)::value;
                hipLaunchKernelGGL((k_pre<NS, T>), dim3(lvl_grid(L.op.N)), dim3(lvl_blk(L.op.N)), 0, st, L.op, L.val, L.minv, b, L.x,
                                   L.r);
            });
        }
        if (L.op.dim == 2 && C.op.N < 128 * 128)
            hipLaunchKernelGGL((k_restrict<2, T, 8>), dim3((C.op.nx + 7) / 8, (C.op.ny + 7) / 8), dim3(64), 0, st,
                               L.op, L.R, L.r, C.op, C.b);
        else if (L.op.dim == 2)
            hipLaunchKernelGGL((k_restrict<2, T>), dim3((C.op.nx + 15) / 16, (C.op.ny + 15) / 16), dim3(256), 0, st,
                               L.op, L.R, L.r, C.op, C.b);
        else
            hipLaunchKernelGGL((k_restrict<1, T>), dim3((C.op.nx + 255) / 256), dim3(256), 0, st, L.op, L.R, L.r,
                               C.op, C.b);
        b = C.b;
    }
    AmgLevel<T> &Lc = g.lv[nl - 1];
    hipLaunchKernelGGL((k_dense_mv<T>), dim3(dmv_blocks(g.nco)), dim3(64 * DMV_W), 0, st, g.nco, g.coarse_inv, Lc.b, Lc.x);
    // upward
    for (int l = nl - 2; l >= 0; --l) {
        AmgLevel<T> &L = g.lv[l];
        AmgLevel<T> &C = g.lv[l + 1];
        const T *bl = (l == 0) ? in : L.b;
        T *dst = (l == 0) ? out : L.r;   // r is free here
        if (fused_level(L, l)) {
            dispatch_ns(L.op.ns, [&](auto c) {
                constexpr int NS = decltype(c)::value;
                if (L.op.N < 128 * 128)
                    hipLaunchKernelGGL((k_up2<NS, T, 8, 256>), dim3((L.op.nx + 7) / 8, (L.op.ny + 7) / 8), dim3(256),
                                       0, st, L.op, L.val, L.minv, bl, L.x, L.P, C.op, C.x, dst);
                else
                    hipLaunchKernelGGL((k_up2<NS, T, 16, 256>), dim3((L.op.nx + 15) / 16, (L.op.ny + 15) / 16),
                                       dim3(256), 0, st, L.op, L.val, L.minv, bl, L.x, L.P, C.op, C.x, dst);
            });
            if (l > 0) std::swap(L.x, L.r);
            continue;
        }
        hipLaunchKernelGGL((k_prolong<T>), dim3(lvl_grid(L.op.N)), dim3(lvl_blk(L.op.N)), 0, st, L.op, L.P, C.op, C.x, L.x);
        if (l == 0 && smoother) {
            dispatch_ns(L.op.ns, [&](auto c) {
                constexpr int NS = decltype(c)::value;
                hipLaunchKernelGGL((k_resid<NS, T>), dim3(nblk(L.op.N)), dim3(BLK), 0, st, L.op, L.val, L.x, bl, L.t);
                hipLaunchKernelGGL((k_spmv_add<NS, T>), dim3(nblk(L.op.N)), dim3(BLK), 0, st, L.op, smoother, L.t,
                                   L.x, dst);
            });
        } else {
            dispatch_ns(L.op.ns, [&](auto c) {
                constexpr int NS = decltype(c)::value;
                hipLaunchKernelGGL((k_post<NS, T>), dim3(lvl_grid(L.op.N)), dim3(lvl_blk(L.op.N)), 0, st, L.op, L.val, L.minv, bl,
                                   L.x, dst);
            });
        }
        if (l > 0) std::swap(L.x, L.r);
    }
}

template <typename T>
static void precond_apply(SolverT<T> &S, int precond, const T *in, T *out) {
    switch (precond) {
    case YZ_PC_AMG: vcycle(S, in, out, (const T *)nullptr); break;
    case YZ_PC_LEARNED_AMG: vcycle(S, in, out, S.M); break;
    case YZ_PC_LEARNED: launch_spmv(S.op0, S.M, in, out, S.stream); break;
    default:
        hipLaunchKernelGGL((k_scale<T>), dim3(nblk(S.op0.N)), dim3(BLK), 0, S.stream, S.op0.N, S.dinv_jac, in, out);
        break;
    }
}

template <typename T>
static int read_scalars(SolverT<T> &S) {
    YZ_CHECK(hipMemcpyAsync(S.kr.host_sc, S.kr.sc, SC_N * sizeof(double), hipMemcpyDeviceToHost, S.stream));
    YZ_CHECK(hipStreamSynchronize(S.stream));
    return 0;
}

template <typename T>
int bicgstab(SolverT<T> &S, const T *b, T *x, int precond, int maxit, double tol, int *iters, double *resid) {
    Krylov<T> &k = S.kr;
    const StencilOp &op = S.op0;
    const int N = op.N;
    hipStream_t st = S.stream;
    *iters = 0;
    *resid = 0;
    dispatch_ns(op.ns, [&](auto c) {
        constexpr int NS = decltype(c)::value;
        hipLaunchKernelGGL((k_init_residual<NS, T>), dim3(RED_BLOCKS), dim3(BLK), 0, st, op, S.A, x, b, k.r, k.rh,
                           k.part);
    });
    hipLaunchKernelGGL(k_finish<2>, dim3(1), dim3(BLK), 0, st, k.part, k.sc, (int)FIN_INIT);
    if (read_scalars(S)) return 2;
    const double *h = k.host_sc;
    const double nb = std::sqrt(h[SC_BB]);
    if (nb == 0) {
        YZ_CHECK(hipMemsetAsync(x, 0, N * sizeof(T), st));
        return 0;
    }
    const double eps = tol * nb;
    double res = std::sqrt(h[SC_RR]);
    if (res <= eps) { *resid = res / nb; return 0; }
    double res_mark = res;
    int it = 0;
    for (; it < maxit; ++it) {
        if (it > 0 && std::fabs(h[SC_RHO1]) < 1e-300) { *iters = it; *resid = res / nb; return 2; }
        hipLaunchKernelGGL((k_p_update<T>), dim3(nblk(N)), dim3(BLK), 0, st, N, it == 0 ? 1 : 0, k.r, k.p, k.v, k.sc);
        precond_apply(S, precond, (const T *)k.p, k.ph);
        if (S.ktime) {
            if (!S.kev[0]) { (void)hipEventCreate(&S.kev[0]); (void)hipEventCreate(&S.kev[1]); }
            (void)hipEventRecord(S.kev[0], st);
        }
        dispatch_ns(op.ns, [&](auto c) {
            constexpr int NS = decltype(c)::value;
            hipLaunchKernelGGL((k_spmv_dot<NS, T>), dim3(RED_BLOCKS), dim3(BLK), 0, st, op, S.A, k.ph, k.v, k.rh,
                               k.part);
        });
        if (S.ktime) (void)hipEventRecord(S.kev[1], st);
        hipLaunchKernelGGL(k_finish<1>, dim3(1), dim3(BLK), 0, st, k.part, k.sc, (int)FIN_ALPHA);
        hipLaunchKernelGGL((k_s_update<T>), dim3(RED_BLOCKS), dim3(BLK), 0, st, N, k.r, k.v, k.s, k.sc, k.part);
        hipLaunchKernelGGL(k_finish<1>, dim3(1), dim3(BLK), 0, st, k.part, k.sc, (int)FIN_SS);
        if (read_scalars(S)) return 2;
        if (S.ktime) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, S.kev[0], S.kev[1]) == hipSuccess) { S.kms += ms; S.kcount++; }
        }
        if (h[SC_RV] == 0) { *iters = it; *resid = res / nb; return 2; }
        res = std::sqrt(h[SC_SS]);
        if (res < eps) {
            hipLaunchKernelGGL((k_x_alpha<T>), dim3(nblk(N)), dim3(BLK), 0, st, N, x, k.ph, k.sc);
            ++it;
            break;
        }
        precond_apply(S, precond, (const T *)k.s, k.sh);
        dispatch_ns(op.ns, [&](auto c) {
            constexpr int NS = decltype(c)::value;
            hipLaunchKernelGGL((k_spmv_dot2<NS, T>), dim3(RED_BLOCKS), dim3(BLK), 0, st, op, S.A, k.sh, k.t, k.s,
                               k.part);
        });
        hipLaunchKernelGGL(k_finish<2>, dim3(1), dim3(BLK), 0, st, k.part, k.sc, (int)FIN_OMEGA);
        hipLaunchKernelGGL((k_xr_update<T>), dim3(RED_BLOCKS), dim3(BLK), 0, st, N, x, k.ph, k.sh, k.r, k.s, k.t,
                           k.rh, k.sc, k.part);
        hipLaunchKernelGGL(k_finish<2>, dim3(1), dim3(BLK), 0, st, k.part, k.sc, (int)FIN_NEXT);
        if (read_scalars(S)) return 2;
        if (h[SC_OMEGA] == 0) { *iters = it; *resid = res / nb; return 2; }
        res = std::sqrt(h[SC_RR]);
        if (res < eps) { ++it; break; }
        if (!(res == res)) { *iters = it; *resid = res; return 2; }
        // stall test: the residual must fall by stall_factor every window
        if (S.stall_window > 0 && (it + 1) % S.stall_window == 0) {
            if (res > S.stall_factor * res_mark) { *iters = it + 1; *resid = res / nb; return 3; }
            res_mark = res;
        }
    }
    *iters = it;
    *resid = res / nb;
    if (!(*resid == *resid)) return 2;
    return (*resid <= tol) ? 0 : 1;
}

// the two value types
template int solver_init<double>(SolverT<double> &, const StencilOp &, hipStream_t);
template int solver_init<float>(SolverT<float> &, const StencilOp &, hipStream_t);
template void solver_free<double>(SolverT<double> &);
template void solver_free<float>(SolverT<float> &);
template int amg_setup<double>(SolverT<double> &);
template int amg_setup<float>(SolverT<float> &);
template int jacobi_setup<double>(SolverT<double> &);
template int jacobi_setup<float>(SolverT<float> &);
template <typename T>
int residual_norms(SolverT<T> &S, const T *x, const T *b, T *r, double *bb, double *rr) {
    const StencilOp &op = S.op0;
    Krylov<T> &k = S.kr;
    dispatch_ns(op.ns, [&](auto c) {
        constexpr int NS = decltype(c)::value;
        hipLaunchKernelGGL((k_init_residual<NS, T>), dim3(RED_BLOCKS), dim3(BLK), 0, S.stream, op, S.A, x, b, r, k.rh,
                           k.part);
    });
    hipLaunchKernelGGL(k_finish<2>, dim3(1), dim3(BLK), 0, S.stream, k.part, k.sc, (int)FIN_INIT);
    if (read_scalars(S)) return 2;
    *bb = k.host_sc[SC_BB];
    *rr = k.host_sc[SC_RR];
    return 0;
}
template int residual_norms<double>(SolverT<double> &, const double *, const double *, double *, double *, double *);
template int bicgstab<double>(SolverT<double> &, const double *, double *, int, int, double, int *, double *);
template int bicgstab<float>(SolverT<float> &, const float *, float *, int, int, double, int *, double *);
template void launch_spmv<double>(const StencilOp &, const double *, const double *, double *, hipStream_t);
template void launch_spmv<float>(const StencilOp &, const float *, const float *, float *, hipStream_t);

}  // namespace yz
