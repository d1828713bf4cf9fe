// yz_krylov.hip -- BiCGSTAB right-preconditioned by an explicit 13-point
// stencil M (the learned approximate inverse), as one fused device
// iteration for the 2-D film.
//
// The method is the one of bicgstab() in yz_solver.hip (van der Vorst 1992,
// right preconditioning, convergence on ||r|| <= tol ||b||, the half-step
// exit on ||s||, breakdown on (rh, v) = 0, omega = 0, rho = 0 or NaN, the
// stall test), rearranged for the device into four kernels per iteration k:
//
//   K1  x += alpha ph + omega sh (iteration k-1);  r = s - omega t;
//       p = r + beta (p_old - omega v);  ph = M p;  partial (r, r)
//       (r and p are formed again at each of the 13 columns from s, t,
//       p_old, v, so no separate update pass runs; p is double-buffered)
//   K2  the exits on ||r|| (converged, breakdown, stall, cap);
//       v = A ph;  partial (rh, v)
//   K3  alpha = rho / (rh, v);  s = r - alpha v;  sh = M s;
//       partials (s, s), (s, rh)
//   K4  the half-step exit on ||s||;  t = A sh;
//       partials (t, t), (t, s), (t, rh)
// and K1 of the next iteration takes omega = (t, s) / (t, t) and
// rho_new = (r_new, rh) = (s, rh) - omega (t, rh), beta = (rho_new / rho)
// (alpha / omega) from those partials.
//
// M is read as FP16 raw rows and an FP32 row scale (M = diag(m) Y): the
// preconditioner is any fixed linear operator, so its rounding changes the
// iteration count at most, never the answer.
//
// Every dot product is reduced in FP64 in a fixed order with no atomics and
// no finishing kernel: each of the G blocks writes its partial, and every
// block of the next kernel sums the G partials in the same order, so all
// blocks derive the same scalars and the same exit decision. Block 0
// records the scalars for the kernel after it; scalars and partials are
// double-buffered by the parity of the kernel's sequence number, so no
// block reads what another block of the same kernel writes. Only stream
// order makes the values visible (a last-arriving-block finish needs
// device-scope atomics, which serialise at the memory side, or release
// fences, which write the L2 back, on this many-XCD chip). After an exit
// every kernel returns at once; the host enqueues a chunk of iterations
// and reads the state once per chunk.
#include <hip/hip_fp16.h>
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

constexpr int TX = 64, TY = 16, TB = 256;  // 64 x 16 tiles, 256 threads, four rows per thread
constexpr int RPT = TY / (TB / TX);         // rows per thread
constexpr int LW = TX + 4, LH = TY + 4;     // LDS tile with the 2-cell halo
// blocks of every kernel (grid-stride over the tiles), so partials per
// reduction: measured at 1024^2 against 256, 512, 1536, 2048 and 4096
constexpr int G = 1024;
constexpr int CHUNK = 8;   // iterations enqueued between two reads of the state (even)

// slots of the 2-D 13-point stencil: oM mm om pm Mo mo oo po Po mp op pp oP
constexpr int OX[13] = {0, -1, 0, 1, -2, -1, 0, 1, 2, -1, 0, 1, 0};
constexpr int OY[13] = {-2, -1, -1, -1, 0, 0, 0, 0, 0, 1, 1, 1, 2};

// the scalars of the iteration; two copies, by the parity of the kernel's
// sequence number q: kernel q reads copy (q-1)&1 and block 0 writes copy q&1
// (S_CHK: the partial (r, r) in flight closes an iteration, so K2 applies
// the exits to it)
enum { S_RHO, S_ALPHA, S_OMEGA, S_BETA, S_RES, S_EPS, S_NB, S_MARK, S_STATE, S_IT, S_SRH, S_CHK, S_N = 16 };
enum { ST_RUN = 0, ST_DONE = 1, ST_HALF = 2, ST_BREAK = 3, ST_STALL = 4, ST_CAP = 5, ST_ZERO = 7 };
constexpr int NQ = 3;      // quantities per partial set

struct Ctl {
    double *sc;      // 2 x S_N
    double *part;    // 2 (parity) x NQ x G
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
    s.pin = c.part + ((q - 1) & 1) * NQ * G;
    s.pout = c.part + (q & 1) * NQ * G;
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

// stage f(j) over tile (x0, y0) and its 2-cell halo into L (folded at the
// edges like the stencil's columns): coalesced row loads instead of 13
// scattered loads per cell and vector
template <class F>
__device__ inline void stage(const StencilOp &op, int x0, int y0, float (&L)[LH][LW], F f) {
    const bool inside = x0 >= 2 && x0 + TX + 2 <= op.nx && y0 >= 2 && y0 + TY + 2 <= op.ny;
    for (int q = threadIdx.x; q < LH * LW; q += TB) {
        const int ly = q / LW, lx = q - ly * LW;
        int gx = x0 + lx - 2, gy = y0 + ly - 2;
        if (!inside) { gx = fold_axis(gx, op.nx, op.perx); gy = fold_axis(gy, op.ny, op.pery); }
        L[ly][lx] = f(gy * op.nx + gx);
    }
}

// the 13-point product of planes W (row i) with the staged tile at local
// cell (lx, ly) (tile coordinates, without the halo)
template <class T, class Cv>
__device__ inline float apply13(const T *__restrict__ W, int N, int i, const float (&L)[LH][LW], int lx, int ly, Cv cv) {
    float a = 0;
#pragma unroll
    for (int k = 0; k < 13; ++k) a = fmaf(cv(W[k * N + i]), L[ly + 2 + OY[k]][lx + 2 + OX[k]], a);
    return a;
}
__device__ inline float idf(float v) { return v; }
__device__ inline float h2f(__half v) { return __half2float(v); }

struct Vecs {
    float *x, *r, *rh, *p0, *p1, *ph, *v, *s, *sh, *t;
};

// the cells of tile (x0, y0) this thread owns: column tx, rows ty, ty + 4,
// ty + 8, ty + 12
template <class F>
__device__ inline void for_own(const StencilOp &op, int x0, int y0, F f) {
    const int lx = threadIdx.x & (TX - 1);
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int ly = threadIdx.x / TX + r * (TB / TX);
        if (x0 + lx < op.nx && y0 + ly < op.ny) f(lx, ly, (y0 + ly) * op.nx + x0 + lx);
    }
}

// K0 (q = 0): r = b - A x, rh = r; partials (b, b), (r, r)
__global__ void __launch_bounds__(TB) k_f_init(StencilOp op, const float *__restrict__ A,
                                               const float *__restrict__ b, Vecs V, Ctl c) {
    __shared__ float L[LH][LW];
    const Seq sq = seq(c, 0);
    double acc[2] = {0, 0};
    const float *__restrict__ x = V.x;
    for (int t = blockIdx.x; t < c.ntx * c.nty; t += G) {
        const int x0 = (t % c.ntx) * TX, y0 = (t / c.ntx) * TY;
        __syncthreads();
        stage(op, x0, y0, L, [&](int j) { return x[j]; });
        __syncthreads();
        for_own(op, x0, y0, [&](int lx, int ly, int i) {
            const float bi = b[i];
            const float ri = bi - apply13(A, op.N, i, L, lx, ly, idf);
            V.r[i] = ri; V.rh[i] = ri;
            acc[0] += (double)bi * bi;
            acc[1] += (double)ri * ri;
        });
    }
    store_partials<2>(acc, sq.pout);
}

// K1 (q = 4k + 1; first: k = 0; par: k & 1). k = 0: opens the solve from
// K0's partials, p = r. k >= 1: omega, rho, beta from K4's partials; the x
// update of iteration k - 1 (or, after the half-step exit, x += alpha ph
// and the end); r = s - omega t; p = r + beta (p_old - omega v); ph = M p;
// partial (r, r).
__global__ void __launch_bounds__(TB) k_f_pm(StencilOp op, const __half *__restrict__ Y,
                                             const float *__restrict__ ms, Vecs V, Ctl c, int q, int first,
                                             int par) {
    __shared__ float L[LH][LW];
    const Seq sq = seq(c, q);
    double s[S_N];
    load_sc(sq.in, s);
    const bool half = !first && s[S_STATE] == ST_HALF;
    const double alpha_prev = s[S_ALPHA];
    if (first) {
        double tot[2];
        sum_partials<2>(sq.pin, tot);
        const double nb = sqrt(tot[0]), res = sqrt(tot[1]);
        s[S_NB] = nb; s[S_EPS] = c.tol * nb; s[S_RES] = res; s[S_MARK] = res;
        s[S_RHO] = tot[1]; s[S_ALPHA] = 1; s[S_OMEGA] = 1; s[S_BETA] = 0; s[S_IT] = 0; s[S_SRH] = 0;
        s[S_CHK] = 0;
        s[S_STATE] = (nb == 0) ? ST_ZERO : (res <= s[S_EPS] ? ST_DONE : ST_RUN);
    } else if (half) {
        s[S_IT] += 1;
        s[S_STATE] = ST_DONE;
    } else if (s[S_STATE] == ST_RUN) {
        double tot[3];
        sum_partials<3>(sq.pin, tot);   // (t, t), (t, s), (t, rh)
        const double omega = tot[0] != 0 ? tot[1] / tot[0] : 0.0;
        if (omega == 0 || !(omega == omega)) {
            s[S_STATE] = ST_BREAK;
        } else {
            const double rho = s[S_SRH] - omega * tot[2];
            s[S_BETA] = (rho / s[S_RHO]) * (s[S_ALPHA] / omega);
            s[S_RHO] = rho;
            s[S_OMEGA] = omega;
            s[S_CHK] = 1;
        }
    }
    commit_sc(s, sq.out);
    if (!half && s[S_STATE] != ST_RUN) return;
    if (half) {
        const float a = (float)alpha_prev;
        for (int t = blockIdx.x; t < c.ntx * c.nty; t += G) {
            const int x0 = (t % c.ntx) * TX, y0 = (t / c.ntx) * TY;
            for_own(op, x0, y0, [&](int, int, int i) { V.x[i] += a * V.ph[i]; });
        }
        return;
    }
    const float *__restrict__ pold = par ? V.p0 : V.p1;
    float *__restrict__ pnew = par ? V.p1 : V.p0;
    const float alpha = (float)s[S_ALPHA], omega = (float)s[S_OMEGA];
    const float beta = (float)s[S_BETA], bw = (float)(s[S_BETA] * s[S_OMEGA]);
    const float *__restrict__ sv = V.s, *__restrict__ tv = V.t, *__restrict__ vv = V.v, *__restrict__ rv = V.r;
    double acc[1] = {0};
    for (int t = blockIdx.x; t < c.ntx * c.nty; t += G) {
        const int x0 = (t % c.ntx) * TX, y0 = (t / c.ntx) * TY;
        __syncthreads();
        if (first) stage(op, x0, y0, L, [&](int j) { return rv[j]; });
        else stage(op, x0, y0, L, [&](int j) { return (sv[j] - omega * tv[j]) + beta * pold[j] - bw * vv[j]; });
        __syncthreads();
        for_own(op, x0, y0, [&](int lx, int ly, int i) {
            const float pi = L[ly + 2][lx + 2];
            const float ri = first ? pi : sv[i] - omega * tv[i];
            if (!first) {
                V.x[i] += alpha * V.ph[i] + omega * V.sh[i];
                V.r[i] = ri;
            }
            pnew[i] = pi;
            V.ph[i] = ms[i] * apply13(Y, op.N, i, L, lx, ly, h2f);
            acc[0] += (double)ri * ri;
        });
    }
    store_partials<1>(acc, sq.pout);
}

// K2 (q = 4k + 2): for k >= 1 the exits on ||r_k||; v = A ph; partial (rh, v)
__global__ void __launch_bounds__(TB) k_f_av(StencilOp op, const float *__restrict__ A, Vecs V, Ctl c, int q,
                                             int work) {
    __shared__ float L[LH][LW];
    const Seq sq = seq(c, q);
    double s[S_N];
    load_sc(sq.in, s);
    if (s[S_CHK] != 0 && s[S_STATE] == ST_RUN) {
        double tot[1];
        sum_partials<1>(sq.pin, tot);
        const double it = s[S_IT] + 1;
        s[S_CHK] = 0;
        const double res = sqrt(tot[0]);
        s[S_IT] = it; s[S_RES] = res;
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
    commit_sc(s, sq.out);
    if (!work || s[S_STATE] != ST_RUN) return;
    const float *__restrict__ ph = V.ph;
    double acc[1] = {0};
    for (int t = blockIdx.x; t < c.ntx * c.nty; t += G) {
        const int x0 = (t % c.ntx) * TX, y0 = (t / c.ntx) * TY;
        __syncthreads();
        stage(op, x0, y0, L, [&](int j) { return ph[j]; });
        __syncthreads();
        for_own(op, x0, y0, [&](int lx, int ly, int i) {
            const float a = apply13(A, op.N, i, L, lx, ly, idf);
            V.v[i] = a;
            acc[0] += (double)V.rh[i] * a;
        });
    }
    store_partials<1>(acc, sq.pout);
}

// K3 (q = 4k + 3): alpha = rho / (rh, v); s = r - alpha v; sh = M s;
// partials (s, s), (s, rh)
__global__ void __launch_bounds__(TB) k_f_sm(StencilOp op, const __half *__restrict__ Y,
                                             const float *__restrict__ ms, Vecs V, Ctl c, int q) {
    __shared__ float L[LH][LW];
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
    const float *__restrict__ rv = V.r, *__restrict__ vv = V.v;
    double acc[2] = {0, 0};
    for (int t = blockIdx.x; t < c.ntx * c.nty; t += G) {
        const int x0 = (t % c.ntx) * TX, y0 = (t / c.ntx) * TY;
        __syncthreads();
        stage(op, x0, y0, L, [&](int j) { return rv[j] - alpha * vv[j]; });
        __syncthreads();
        for_own(op, x0, y0, [&](int lx, int ly, int i) {
            const float si = L[ly + 2][lx + 2];
            V.s[i] = si;
            V.sh[i] = ms[i] * apply13(Y, op.N, i, L, lx, ly, h2f);
            acc[0] += (double)si * si;
            acc[1] += (double)si * V.rh[i];
        });
    }
    store_partials<2>(acc, sq.pout);
}

// K4 (q = 4k + 4): the half-step exit on ||s||; t = A sh; partials (t, t),
// (t, s), (t, rh)
__global__ void __launch_bounds__(TB) k_f_at(StencilOp op, const float *__restrict__ A, Vecs V, Ctl c, int q) {
    __shared__ float L[LH][LW];
    const Seq sq = seq(c, q);
    double s[S_N];
    load_sc(sq.in, s);
    if (s[S_STATE] == ST_RUN) {
        double tot[2];
        sum_partials<2>(sq.pin, tot);
        const double res = sqrt(tot[0]);
        s[S_SRH] = tot[1];
        if (!(res == res)) s[S_STATE] = ST_BREAK;
        else if (res < s[S_EPS]) { s[S_RES] = res; s[S_STATE] = ST_HALF; }
    }
    commit_sc(s, sq.out);
    if (s[S_STATE] != ST_RUN) return;
    const float *__restrict__ sh = V.sh;
    double acc[3] = {0, 0, 0};
    for (int t = blockIdx.x; t < c.ntx * c.nty; t += G) {
        const int x0 = (t % c.ntx) * TX, y0 = (t / c.ntx) * TY;
        __syncthreads();
        stage(op, x0, y0, L, [&](int j) { return sh[j]; });
        __syncthreads();
        for_own(op, x0, y0, [&](int lx, int ly, int i) {
            const float a = apply13(A, op.N, i, L, lx, ly, idf);
            V.t[i] = a;
            acc[0] += (double)a * a;
            acc[1] += (double)a * V.s[i];
            acc[2] += (double)a * V.rh[i];
        });
    }
    store_partials<3>(acc, sq.pout);
}

}  // namespace

int bicgstab_stencil(SolverT<float> &S, const float *b, float *x, int maxit, double tol, int *iters,
                     double *resid) {
    const StencilOp &op = S.op0;
    *iters = 0;
    *resid = 0;
    bool ok = op.dim == 2 && op.ns == 13 && S.Mh && S.Mscale;
    for (int k = 0; ok && k < 13; ++k) ok = op.ox[k] == OX[k] && op.oy[k] == OY[k];
    if (!ok) {
        if (!S.M) return 2;
        return bicgstab(S, b, x, YZ_PC_LEARNED, maxit, tol, iters, resid);
    }
    Krylov<float> &kr = S.kr;
    hipStream_t st = S.stream;
    const int ntx = (op.nx + TX - 1) / TX, nty = (op.ny + TY - 1) / TY;
    if (!kr.ftiles) {
        KR_CHECK(hipMalloc((void **)&kr.p2, (size_t)op.N * sizeof(float)));
        KR_CHECK(hipMalloc((void **)&kr.fpart, (size_t)2 * NQ * G * sizeof(double)));
        kr.ftiles = 1;
    }
    Ctl c;
    c.sc = kr.sc + 32;   // two copies of S_N scalars (the solver's array holds 64)
    c.part = kr.fpart; c.ntx = ntx; c.nty = nty;
    c.maxit = maxit; c.window = S.stall_window; c.factor = S.stall_factor; c.tol = tol;
    Vecs V;
    V.x = x; V.r = kr.r; V.rh = kr.rh; V.p0 = kr.p; V.p1 = kr.p2; V.ph = kr.ph; V.v = kr.v; V.s = kr.s;
    V.sh = kr.sh; V.t = kr.t;
    const dim3 grid(G), blk(TB);
    const double *h = kr.host_sc;
    // the scalars as K2 would see them next (one block, no work; the K2
    // that follows derives the same values from the same inputs)
    auto check = [&]() -> int {
        hipLaunchKernelGGL(k_f_av, dim3(1), blk, 0, st, op, (const float *)S.A, V, c, 2, 0);
        KR_CHECK(hipGetLastError());
        KR_CHECK(hipMemcpyAsync(kr.host_sc, c.sc, S_N * sizeof(double), hipMemcpyDeviceToHost, st));
        KR_CHECK(hipStreamSynchronize(st));
        return 0;
    };
    // two iterations, k even then odd: every kernel argument is fixed by
    // its position (sequence parities 2, 3, 4, 5; p buffers by k's parity),
    // so the pair is captured once as a HIP graph and replayed
    auto pair = [&](bool timed) {
        for (int u = 0; u < 2; ++u) {
            if (timed && u == 0) (void)hipEventRecord(S.kev[0], st);
            hipLaunchKernelGGL(k_f_av, grid, blk, 0, st, op, (const float *)S.A, V, c, 2, 1);
            if (timed && u == 0) (void)hipEventRecord(S.kev[1], st);
            hipLaunchKernelGGL(k_f_sm, grid, blk, 0, st, op, S.Mh, S.Mscale, V, c, 3);
            hipLaunchKernelGGL(k_f_at, grid, blk, 0, st, op, (const float *)S.A, V, c, 4);
            hipLaunchKernelGGL(k_f_pm, grid, blk, 0, st, op, S.Mh, S.Mscale, V, c, 5, 0, u == 0 ? 1 : 0);
        }
    };
    if (S.use_graph && (!kr.fgexec || kr.fg_x != x || kr.fg_maxit != maxit || kr.fg_window != c.window ||
                        kr.fg_factor != c.factor)) {
        if (kr.fgexec) { (void)hipGraphExecDestroy(kr.fgexec); kr.fgexec = nullptr; }
        hipGraph_t g = nullptr;
        KR_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        pair(false);
        KR_CHECK(hipStreamEndCapture(st, &g));
        KR_CHECK(hipGraphInstantiate(&kr.fgexec, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
        kr.fg_x = x; kr.fg_maxit = maxit; kr.fg_window = c.window; kr.fg_factor = c.factor;
    }
    if (S.ktime && !S.kev[0]) { (void)hipEventCreate(&S.kev[0]); (void)hipEventCreate(&S.kev[1]); }
    hipLaunchKernelGGL(k_f_init, grid, blk, 0, st, op, (const float *)S.A, b, V, c);
    hipLaunchKernelGGL(k_f_pm, grid, blk, 0, st, op, S.Mh, S.Mscale, V, c, 1, 1, 0);
    if (check()) return 2;
    int state = (int)h[S_STATE];
    if (state == ST_ZERO) {
        KR_CHECK(hipMemsetAsync(x, 0, op.N * sizeof(float), st));
        return 0;
    }
    // chunks of CHUNK iterations (an even number: every chunk starts at an
    // even k); iterations past an exit, or past maxit (K2 sets the cap),
    // return at once
    for (int k = 0; state == ST_RUN && k < maxit + CHUNK; k += CHUNK) {
        for (int u = 0; u < CHUNK / 2; ++u) {
            const bool timed = S.ktime && u == 0;
            if (S.use_graph && !timed) KR_CHECK(hipGraphLaunch(kr.fgexec, st));
            else pair(timed);
        }
        KR_CHECK(hipGetLastError());
        if (check()) return 2;
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
