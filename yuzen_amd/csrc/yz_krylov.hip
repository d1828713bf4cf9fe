// yz_krylov.hip -- BiCGSTAB right-preconditioned by an explicit 13-point
// stencil M (the learned approximate inverse), as one fused device
// iteration for the 2-D film.
//
// The method is the one of bicgstab() in yz_solver.hip (van der Vorst 1992,
// right preconditioning, convergence on ||r|| <= tol ||b||, the half-step
// exit on ||s||, breakdown on (rh, v) = 0, omega = 0, rho = 0 or NaN, the
// stall test), rearranged for the device:
//
//   K1  p_new = r + beta (p - omega v);  ph = M p_new    (p_new recomputed
//       at the 13 columns from r, p, v: no separate update pass; p is
//       double-buffered so no cell reads a neighbour's new value)
//   K2  v = A ph;  (rh, v)                 -> alpha
//   K3  s = r - alpha v;  sh = M s;  (s, s) -> half-step exit
//   K4  t = A sh;  (t, t), (t, s)          -> omega
//   K5  x += alpha ph + omega sh;  r = s - omega t;  (r, r), (r, rh)
//                                          -> beta, exits, stall test
//
// Every dot product is reduced in FP64 in a fixed order with no atomics and
// no finishing kernel: each of the G blocks writes its partial, and every
// block of the next kernel sums the G partials in the same order, so all
// blocks derive the same scalars and the same exit decision (converged,
// half-step exit, breakdown, stall, iteration cap). Block 0 records the
// scalars for the kernel after it; scalars and partials are double-buffered
// by the parity of the kernel's sequence number, so no block reads what
// another block of the same kernel writes. Only stream order makes the
// values visible (a last-arriving-block finish needs device-scope atomics,
// which serialise, or release fences, which write the L2 back, on this
// many-XCD chip). After an exit every kernel returns at once; the host
// enqueues a chunk of iterations and reads the state once per chunk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>

#include "../../include/yuzen_gpu.h"
#include "yz_solver.h"

namespace yz {

#define KR_CHECK(x)                                                                 \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_),  \
                         __FILE__, __LINE__);                                       \
            return -1;                                                              \
        }                                                                           \
    } while (0)

namespace {

constexpr int TX = 64, TY = 4, TB = TX * TY;
constexpr int G = 512;     // blocks of every kernel (grid-stride over 64 x 4 tiles): partials per reduction
constexpr int CHUNK = 8;   // iterations enqueued between two reads of the state

// slots of the 2-D 13-point stencil: oM mm om pm Mo mo oo po Po mp op pp oP
constexpr int OX[13] = {0, -1, 0, 1, -2, -1, 0, 1, 2, -1, 0, 1, 0};
constexpr int OY[13] = {-2, -1, -1, -1, 0, 0, 0, 0, 0, 1, 1, 1, 2};

// the scalars of the iteration; two copies, by the parity of the kernel's
// sequence number q: kernel q reads copy (q-1)&1 and block 0 writes copy q&1
enum { S_RHO, S_ALPHA, S_OMEGA, S_BETA, S_RES, S_EPS, S_NB, S_MARK, S_STATE, S_IT, S_RR, S_N = 16 };
enum { ST_RUN = 0, ST_DONE = 1, ST_HALF = 2, ST_BREAK = 3, ST_STALL = 4, ST_CAP = 5, ST_ZERO = 7 };

struct Ctl {
    double *sc;      // 2 x S_N
    double *part;    // 2 (parity) x 2 (quantities) x G
    int ntx, nty;    // tiles
    int maxit, window;
    double factor, tol;
};

struct Seq {
    const double *in;    // scalars of kernel q - 1
    double *out;         // scalars of kernel q
    const double *pin;   // partials of kernel q - 1
    double *pout;        // partials of kernel q
};
__device__ inline Seq seq(const Ctl &c, int q) {
    Seq s;
    s.in = c.sc + ((q - 1) & 1) * S_N;
    s.out = c.sc + (q & 1) * S_N;
    s.pin = c.part + ((q - 1) & 1) * 2 * G;
    s.pout = c.part + (q & 1) * 2 * G;
    return s;
}

// every block sums the previous kernel's G partials of K quantities, in
// the same fixed order, so all blocks hold the same totals
template <int K>
__device__ inline void sum_partials(const double *__restrict__ pin, double (&tot)[K]) {
    __shared__ double sh[K][TB / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int m = 0; m < K; ++m) {
        double x = 0;
        for (int t = threadIdx.x; t < G; t += TB) x += pin[m * G + t];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) sh[m][wave] = x;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < K; ++m) {
        double s = 0;
        for (int w = 0; w < TB / 64; ++w) s += sh[m][w];
        tot[m] = s;
    }
    __syncthreads();
}

// this block's partials of K quantities
template <int K>
__device__ inline void store_partials(double (&v)[K], double *__restrict__ pout) {
    __shared__ double sh[K][TB / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int m = 0; m < K; ++m) {
        double x = v[m];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) sh[m][wave] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int m = 0; m < K; ++m) {
            double s = 0;
            for (int w = 0; w < TB / 64; ++w) s += sh[m][w];
            pout[m * G + blockIdx.x] = s;
        }
    }
}

__device__ inline void load_sc(const double *__restrict__ in, double (&c)[S_N]) {
#pragma unroll
    for (int k = 0; k < S_N; ++k) c[k] = in[k];
}
__device__ inline void commit_sc(const double (&c)[S_N], double *__restrict__ out) {
    if (blockIdx.x == 0 && threadIdx.x < S_N) out[threadIdx.x] = c[threadIdx.x];
}

// the columns of the 13 slots of cell (ix, iy): direct offsets inside,
// folded at the edges
__device__ inline void columns(const StencilOp &op, int ix, int iy, int (&j)[13]) {
    const int i = iy * op.nx + ix;
    if (ix >= 2 && ix < op.nx - 2 && iy >= 2 && iy < op.ny - 2) {
#pragma unroll
        for (int k = 0; k < 13; ++k) j[k] = i + OX[k] + OY[k] * op.nx;
    } else {
#pragma unroll
        for (int k = 0; k < 13; ++k) j[k] = stencil_col(op, ix, iy, k);
    }
}

// cell of tile t of this thread
__device__ inline bool tile_cell(const StencilOp &op, const Ctl &c, int t, int &ix, int &iy) {
    ix = (t % c.ntx) * TX + (threadIdx.x & (TX - 1));
    iy = (t / c.ntx) * TY + (threadIdx.x / TX);
    return ix < op.nx && iy < op.ny;
}

// K0 (q = 0): r = b - A x, rh = r, p = 0, v = 0; partials (b, b), (r, r)
__global__ void __launch_bounds__(TB) k_f_init(StencilOp op, const float *__restrict__ A,
                                               const float *__restrict__ b, const float *__restrict__ x,
                                               float *__restrict__ r, float *__restrict__ rh,
                                               float *__restrict__ p, float *__restrict__ v, Ctl c) {
    const Seq sq = seq(c, 0);
    double acc[2] = {0, 0};
    for (int t = blockIdx.x; t < c.ntx * c.nty; t += G) {
        int ix, iy;
        if (!tile_cell(op, c, t, ix, iy)) continue;
        const int i = iy * op.nx + ix;
        int j[13];
        columns(op, ix, iy, j);
        float ax = 0;
#pragma unroll
        for (int k = 0; k < 13; ++k) ax = fmaf(A[k * op.N + i], x[j[k]], ax);
        const float bi = b[i];
        const float ri = bi - ax;
        r[i] = ri; rh[i] = ri; p[i] = 0.0f; v[i] = 0.0f;
        acc[0] += (double)bi * bi;
        acc[1] += (double)ri * ri;
    }
    store_partials<2>(acc, sq.pout);
}

// K1 (q = 5k + 1): the scalar step closing iteration k - 1 (or opening the
// solve, k = 0) from K5's (K0's) partials, then p_new = r + beta (p_old -
// omega v), ph = M p_new. With work = 0 only the scalar step runs (one
// block, at the end of a chunk, so the host can read the state; the K1
// that follows recomputes the same scalars from the same inputs).
__global__ void __launch_bounds__(TB) k_f_pm(StencilOp op, const float *__restrict__ M,
                                             const float *__restrict__ r, const float *__restrict__ pold,
                                             const float *__restrict__ v, float *__restrict__ pnew,
                                             float *__restrict__ ph, Ctl c, int q, int k, int work) {
    const Seq sq = seq(c, q);
    double s[S_N];
    load_sc(sq.in, s);
    if (k == 0 || s[S_STATE] == ST_RUN) {
        double tot[2];
        sum_partials<2>(sq.pin, tot);
        if (k == 0) {
            const double nb = sqrt(tot[0]), res = sqrt(tot[1]);
            s[S_NB] = nb; s[S_EPS] = c.tol * nb; s[S_RR] = tot[1]; s[S_RES] = res; s[S_MARK] = res;
            s[S_RHO] = tot[1]; s[S_ALPHA] = 1; s[S_OMEGA] = 1; s[S_BETA] = 0; s[S_IT] = 0;
            s[S_STATE] = (nb == 0) ? ST_ZERO : (res <= s[S_EPS] ? ST_DONE : ST_RUN);
        } else {
            const double it = s[S_IT] + 1;
            const double res = sqrt(tot[0]);
            s[S_IT] = it; s[S_RR] = tot[0]; s[S_RES] = res;
            s[S_BETA] = (tot[1] / s[S_RHO]) * (s[S_ALPHA] / s[S_OMEGA]);
            s[S_RHO] = tot[1];
            int st = ST_RUN;
            if (res < s[S_EPS]) st = ST_DONE;
            else if (!(res == res)) st = ST_BREAK;
            else if (fabs(s[S_RHO]) < 1e-300) st = ST_BREAK;
            else if (c.window > 0 && ((long long)it % c.window) == 0) {
                if (res > c.factor * s[S_MARK]) st = ST_STALL;
                else s[S_MARK] = res;
            }
            if (st == ST_RUN && it >= c.maxit) st = ST_CAP;
            s[S_STATE] = st;
        }
    }
    commit_sc(s, sq.out);
    if (!work || s[S_STATE] != ST_RUN) return;
    const float beta = (float)s[S_BETA], bw = (float)(s[S_BETA] * s[S_OMEGA]);
    for (int t = blockIdx.x; t < c.ntx * c.nty; t += G) {
        int ix, iy;
        if (!tile_cell(op, c, t, ix, iy)) continue;
        const int i = iy * op.nx + ix;
        int j[13];
        columns(op, ix, iy, j);
        float a = 0;
#pragma unroll
        for (int kk = 0; kk < 13; ++kk) {
            const int qq = j[kk];
            a = fmaf(M[kk * op.N + i], r[qq] + beta * pold[qq] - bw * v[qq], a);
        }
        pnew[i] = r[i] + beta * pold[i] - bw * v[i];
        ph[i] = a;
    }
}

// K2 (q = 5k + 2): v = A ph; partial (rh, v)
__global__ void __launch_bounds__(TB) k_f_av(StencilOp op, const float *__restrict__ A,
                                             const float *__restrict__ ph, const float *__restrict__ rh,
                                             float *__restrict__ v, Ctl c, int q) {
    const Seq sq = seq(c, q);
    double s[S_N];
    load_sc(sq.in, s);
    commit_sc(s, sq.out);
    if (s[S_STATE] != ST_RUN) return;
    double acc[1] = {0};
    for (int t = blockIdx.x; t < c.ntx * c.nty; t += G) {
        int ix, iy;
        if (!tile_cell(op, c, t, ix, iy)) continue;
        const int i = iy * op.nx + ix;
        int j[13];
        columns(op, ix, iy, j);
        float a = 0;
#pragma unroll
        for (int k = 0; k < 13; ++k) a = fmaf(A[k * op.N + i], ph[j[k]], a);
        v[i] = a;
        acc[0] += (double)rh[i] * a;
    }
    store_partials<1>(acc, sq.pout);
}

// K3 (q = 5k + 3): alpha = rho / (rh, v); s = r - alpha v; sh = M s;
// partial (s, s)
__global__ void __launch_bounds__(TB) k_f_sm(StencilOp op, const float *__restrict__ M,
                                             const float *__restrict__ r, const float *__restrict__ v,
                                             float *__restrict__ sv, float *__restrict__ sh, Ctl c, int q) {
    const Seq sq = seq(c, q);
    double s[S_N];
    load_sc(sq.in, s);
    if (s[S_STATE] == ST_RUN) {
        double tot[1];
        sum_partials<1>(sq.pin, tot);
        if (tot[0] == 0 || !(tot[0] == tot[0])) s[S_STATE] = ST_BREAK;
        else s[S_ALPHA] = s[S_RHO] / tot[0];
    }
    commit_sc(s, sq.out);
    if (s[S_STATE] != ST_RUN) return;
    const float alpha = (float)s[S_ALPHA];
    double acc[1] = {0};
    for (int t = blockIdx.x; t < c.ntx * c.nty; t += G) {
        int ix, iy;
        if (!tile_cell(op, c, t, ix, iy)) continue;
        const int i = iy * op.nx + ix;
        int j[13];
        columns(op, ix, iy, j);
        float a = 0;
#pragma unroll
        for (int k = 0; k < 13; ++k) {
            const int qq = j[k];
            a = fmaf(M[k * op.N + i], r[qq] - alpha * v[qq], a);
        }
        const float si = r[i] - alpha * v[i];
        sv[i] = si;
        sh[i] = a;
        acc[0] += (double)si * si;
    }
    store_partials<1>(acc, sq.pout);
}

// K4 (q = 5k + 4): the half-step exit on ||s||; t = A sh; partials (t, t),
// (t, s)
__global__ void __launch_bounds__(TB) k_f_at(StencilOp op, const float *__restrict__ A,
                                             const float *__restrict__ sh, const float *__restrict__ sv,
                                             float *__restrict__ tv, Ctl c, int q) {
    const Seq sq = seq(c, q);
    double s[S_N];
    load_sc(sq.in, s);
    if (s[S_STATE] == ST_RUN) {
        double tot[1];
        sum_partials<1>(sq.pin, tot);
        const double res = sqrt(tot[0]);
        if (!(res == res)) s[S_STATE] = ST_BREAK;
        else if (res < s[S_EPS]) { s[S_RES] = res; s[S_STATE] = ST_HALF; }
    }
    commit_sc(s, sq.out);
    if (s[S_STATE] != ST_RUN) return;
    double acc[2] = {0, 0};
    for (int t = blockIdx.x; t < c.ntx * c.nty; t += G) {
        int ix, iy;
        if (!tile_cell(op, c, t, ix, iy)) continue;
        const int i = iy * op.nx + ix;
        int j[13];
        columns(op, ix, iy, j);
        float a = 0;
#pragma unroll
        for (int k = 0; k < 13; ++k) a = fmaf(A[k * op.N + i], sh[j[k]], a);
        tv[i] = a;
        acc[0] += (double)a * a;
        acc[1] += (double)a * sv[i];
    }
    store_partials<2>(acc, sq.pout);
}

// K5 (q = 5k + 5): omega = (t, s) / (t, t); x += alpha ph + omega sh;
// r = s - omega t; partials (r, r), (r, rh). After the half-step exit:
// x += alpha ph, and the solve ends.
__global__ void __launch_bounds__(TB) k_f_xr(int nx, int ny, float *__restrict__ x, const float *__restrict__ ph,
                                             const float *__restrict__ sh, const float *__restrict__ sv,
                                             const float *__restrict__ tv, float *__restrict__ r,
                                             const float *__restrict__ rh, Ctl c, int q) {
    const Seq sq = seq(c, q);
    double s[S_N];
    load_sc(sq.in, s);
    const bool half = s[S_STATE] == ST_HALF;
    if (half) {
        s[S_IT] += 1;
        s[S_STATE] = ST_DONE;
    } else if (s[S_STATE] == ST_RUN) {
        double tot[2];
        sum_partials<2>(sq.pin, tot);
        const double omega = tot[0] != 0 ? tot[1] / tot[0] : 0.0;
        if (omega == 0 || !(omega == omega)) s[S_STATE] = ST_BREAK;
        else s[S_OMEGA] = omega;
    }
    commit_sc(s, sq.out);
    if (!half && s[S_STATE] != ST_RUN) return;
    const float alpha = (float)s[S_ALPHA], omega = (float)s[S_OMEGA];
    double acc[2] = {0, 0};
    for (int t = blockIdx.x; t < c.ntx * c.nty; t += G) {
        const int ix = (t % c.ntx) * TX + (threadIdx.x & (TX - 1));
        const int iy = (t / c.ntx) * TY + (threadIdx.x / TX);
        if (ix >= nx || iy >= ny) continue;
        const int i = iy * nx + ix;
        if (half) {
            x[i] += alpha * ph[i];
        } else {
            x[i] += alpha * ph[i] + omega * sh[i];
            const float ri = sv[i] - omega * tv[i];
            r[i] = ri;
            acc[0] += (double)ri * ri;
            acc[1] += (double)ri * rh[i];
        }
    }
    if (!half) store_partials<2>(acc, sq.pout);
}

}  // namespace

int bicgstab_stencil(SolverT<float> &S, const float *b, float *x, int maxit, double tol, int *iters,
                     double *resid) {
    const StencilOp &op = S.op0;
    *iters = 0;
    *resid = 0;
    bool ok = op.dim == 2 && op.ns == 13 && S.M;
    for (int k = 0; ok && k < 13; ++k) ok = op.ox[k] == OX[k] && op.oy[k] == OY[k];
    if (!ok) return bicgstab(S, b, x, YZ_PC_LEARNED, maxit, tol, iters, resid);
    Krylov<float> &kr = S.kr;
    hipStream_t st = S.stream;
    const int ntx = (op.nx + TX - 1) / TX, nty = (op.ny + TY - 1) / TY;
    if (kr.ftiles != 1) {
        if (kr.p2) (void)hipFree(kr.p2);
        if (kr.fpart) (void)hipFree(kr.fpart);
        kr.p2 = nullptr; kr.fpart = nullptr; kr.ftiles = 0;
        KR_CHECK(hipMalloc((void **)&kr.p2, (size_t)op.N * sizeof(float)));
        KR_CHECK(hipMalloc((void **)&kr.fpart, (size_t)2 * 2 * G * sizeof(double)));
        kr.ftiles = 1;
    }
    Ctl c;
    c.sc = kr.sc + 32;   // two copies of S_N scalars (the solver's array holds 64)
    c.part = kr.fpart; c.ntx = ntx; c.nty = nty;
    c.maxit = maxit; c.window = S.stall_window; c.factor = S.stall_factor; c.tol = tol;
    const dim3 grid(G), blk(TB);
    const double *h = kr.host_sc;
    float *P[2] = {kr.p, kr.p2};
    // the scalars as K1 of iteration k would see them
    auto check = [&](int k) -> int {
        const int q = 5 * k + 1;
        hipLaunchKernelGGL(k_f_pm, dim3(1), blk, 0, st, op, S.M, (const float *)kr.r, (const float *)P[k & 1],
                           (const float *)kr.v, P[(k + 1) & 1], kr.ph, c, q, k, 0);
        KR_CHECK(hipGetLastError());
        KR_CHECK(hipMemcpyAsync(kr.host_sc, c.sc + (q & 1) * S_N, S_N * sizeof(double), hipMemcpyDeviceToHost, st));
        KR_CHECK(hipStreamSynchronize(st));
        return 0;
    };
    hipLaunchKernelGGL(k_f_init, grid, blk, 0, st, op, (const float *)S.A, b, (const float *)x, kr.r, kr.rh, P[0],
                       kr.v, c);
    if (check(0)) return 2;
    int state = (int)h[S_STATE];
    if (state == ST_ZERO) {
        KR_CHECK(hipMemsetAsync(x, 0, op.N * sizeof(float), st));
        return 0;
    }
    int k = 0;
    while (state == ST_RUN && k < maxit) {
        const int n = std::min(CHUNK, maxit - k);
        for (int u = 0; u < n; ++u, ++k) {
            const int q = 5 * k;
            hipLaunchKernelGGL(k_f_pm, grid, blk, 0, st, op, S.M, (const float *)kr.r, (const float *)P[k & 1],
                               (const float *)kr.v, P[(k + 1) & 1], kr.ph, c, q + 1, k, 1);
            const bool timed = S.ktime && u == 0;
            if (timed) {
                if (!S.kev[0]) { (void)hipEventCreate(&S.kev[0]); (void)hipEventCreate(&S.kev[1]); }
                (void)hipEventRecord(S.kev[0], st);
            }
            hipLaunchKernelGGL(k_f_av, grid, blk, 0, st, op, (const float *)S.A, (const float *)kr.ph,
                               (const float *)kr.rh, kr.v, c, q + 2);
            if (timed) (void)hipEventRecord(S.kev[1], st);
            hipLaunchKernelGGL(k_f_sm, grid, blk, 0, st, op, S.M, (const float *)kr.r, (const float *)kr.v, kr.s,
                               kr.sh, c, q + 3);
            hipLaunchKernelGGL(k_f_at, grid, blk, 0, st, op, (const float *)S.A, (const float *)kr.sh,
                               (const float *)kr.s, kr.t, c, q + 4);
            hipLaunchKernelGGL(k_f_xr, grid, blk, 0, st, op.nx, op.ny, x, (const float *)kr.ph,
                               (const float *)kr.sh, (const float *)kr.s, (const float *)kr.t, kr.r,
                               (const float *)kr.rh, c, q + 5);
        }
        KR_CHECK(hipGetLastError());
        if (check(k)) return 2;
        if (S.ktime) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, S.kev[0], S.kev[1]) == hipSuccess) { S.kms += ms; S.kcount++; }
        }
        state = (int)h[S_STATE];
    }
    *iters = (int)h[S_IT];
    const double nb = h[S_NB];
    *resid = nb > 0 ? h[S_RES] / nb : 0.0;
    switch (state) {
    case ST_DONE: return 0;
    case ST_STALL: return 3;
    case ST_CAP:
    case ST_RUN: return (*resid <= tol) ? 0 : 1;
    default: return 2;
    }
}

}  // namespace yz
