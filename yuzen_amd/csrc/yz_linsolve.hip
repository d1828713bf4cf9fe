// yz_linsolve.hip -- the linear-solve seam: one CRS system through the
// device BiCGSTAB, with the contract of ILinearSolverBackend::Solve
// (reference yuzen/src/numerics/LinearSolverBackend.h:56-103; the AMGCL
// backend's Solve, LinearSolverBackend.cpp:539-728): x is the initial guess
// and the result, the matrix and right-hand side are not modified, the
// codes are 0 converged / 1 iteration cap / 2 breakdown or a matrix the
// seam does not take, and nothing throws across the boundary.
//
// The CRS is a film-height or well-mixed concentration matrix of a
// structured grid in the reference's row layout: two ghost layers per side
// whose rows are copies x_g - x_src = r_g (BoundaryCondition::EqualVal:
// Symmetry and Periodic edges), interior rows within the 13-point stencil.
// The grid comes from the settings or is read off the matrix itself (the
// column offsets of an interior row of the 13-point film-height pattern or
// of the 9-point wide concentration pattern, ConcEquationBase.h:242-258).
//
// The CRS goes to the device as it is; one thread per ghost row checks that
// it is a copy row and records its source and offset, one thread per
// interior row folds its entries into the stencil planes (ghost columns
// onto the cells they copy, their offsets into the right-hand side) and
// scales it; the system is solved on the device, and the ghost entries of
// x are formed again from their rows (as CRS_Matrix::SatisfyBCRows makes
// them) before x comes back.
//
// yz_linsolve_noalloc is the same solve on device memory the library holds
// from its load: a file-scope array (kScratchBytes) that every call carves
// afresh, so the call allocates nothing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/yuzen_gpu.h"
#include "yz_physics.h"
#include "yz_solver.h"

using namespace yz;

namespace yz {
void set_error(const std::string &s);
}

namespace {

constexpr int LB = 256;
inline int lblk(long n) { return (int)((n + LB - 1) / LB); }

// the grid of a system and how its ghost rows copy
struct LinGrid {
    int n, nnz, dim, nx, ny, nxp, nyp;
    int bc[4];
};

__device__ inline bool lin_inside(const LinGrid &G, int i) {
    const int ix = i % G.nxp, iy = i / G.nxp;
    return ix >= 2 && ix <= G.nxp - 3 && (G.dim == 1 || (iy >= 2 && iy <= G.nyp - 3));
}

// ghost rows: d x_g + o x_src = r_g with o = -d, so x_g = x_src + r_g / d
__global__ void k_lin_ghosts(LinGrid G, const int *rowptr, const int *col, const double *val, const double *rhs,
                             int *gsrc, double *gofs, int *bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G.n) return;
    if (lin_inside(G, i)) { gsrc[i] = -1; gofs[i] = 0.0; return; }
    double d = 0, o = 0;
    int src = -1, m = 0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
        const double v = val[k];
        if (v == 0) continue;
        if (col[k] == i) d += v;
        else { src = col[k]; o += v; ++m; }
    }
    if (m != 1 || d == 0 || o != -d || src < 0 || src >= G.n || !lin_inside(G, src)) {
        atomicOr(bad, 1);
        gsrc[i] = -1;
        gofs[i] = 0.0;
        return;
    }
    gsrc[i] = src;
    gofs[i] = rhs[i] / d;
}

// an interior row into the planes: its entries by slot (ghost columns fold
// onto the cell they copy, their offsets into b), scaled by its diagonal
__global__ void k_lin_rows(LinGrid G, StencilOp op, int scaling, const int *rowptr, const int *col,
                           const double *val, const double *rhs, const double *xin, const int *gsrc,
                           const double *gofs, double *planes, double *b, double *x, int *bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= op.N) return;
    const int ix = i % op.nx, iy = i / op.nx;
    const int r = (G.dim == 2 ? iy + 2 : 0) * G.nxp + ix + 2;
    double acc[13];
#pragma unroll
    for (int q = 0; q < 13; ++q) acc[q] = 0.0;
    double bi = rhs[r];
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
        const int cc = col[k];
        const double v = val[k];
        const int dxo = cc % G.nxp - (ix + 2), dyo = cc / G.nxp - (G.dim == 2 ? iy + 2 : 0);
        int slot = -1;
        for (int q = 0; q < op.ns; ++q)
            if (op.ox[q] == dxo && op.oy[q] == dyo) slot = q;
        if (slot < 0 || cc < 0 || cc >= G.n) {
            if (v != 0) atomicOr(bad, 2);
            continue;
        }
        acc[slot] += v;
        if (gsrc[cc] >= 0) bi -= v * gofs[cc];
    }
    const int c0 = G.dim == 2 ? 6 : 2;
    const double d = acc[c0];
    const bool sc = scaling && !(fabs(d) < 1.0e-30);
    const double inv = sc ? 1.0 / d : 1.0;
    for (int q = 0; q < op.ns; ++q) planes[(size_t)q * op.N + i] = sc ? acc[q] * inv : acc[q];
    b[i] = sc ? bi * inv : bi;
    x[i] = xin[r];
}

// the solution back into the padded rows: interior cells, then ghosts
__global__ void k_lin_out_interior(LinGrid G, StencilOp op, const double *x, double *xout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= op.N) return;
    const int ix = i % op.nx, iy = i / op.nx;
    xout[(G.dim == 2 ? iy + 2 : 0) * G.nxp + ix + 2] = x[i];
}
__global__ void k_lin_out_ghosts(LinGrid G, const int *gsrc, const double *gofs, double *xout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G.n) return;
    if (gsrc[i] >= 0) xout[i] = xout[gsrc[i]] + gofs[i];
}

// the device arrays of one solve besides the solver's own
struct LinBufs {
    int *rowptr = nullptr, *col = nullptr, *gsrc = nullptr, *flag = nullptr;
    double *val = nullptr, *rhs = nullptr, *xin = nullptr, *gofs = nullptr, *xout = nullptr;
    double *b = nullptr, *x = nullptr;
};

// the grid of a system: given by the settings, or read off an interior
// row's pattern (the largest forward offset of the 13-point film-height or
// 9-point concentration row is two padded rows, of the 5-point 1-D row 2)
int lin_grid(int n, int nnz, const int *rowptr, const int *col, const yz_lin_settings *s, LinGrid &G) {
    G.n = n; G.nnz = nnz;
    G.dim = s->grid_dim; G.nx = s->grid_nx; G.ny = s->grid_ny;
    for (int k = 0; k < 4; ++k) G.bc[k] = s->bc[k];
    if (G.nx <= 0) {
        int maxoff = 0, kind = 0;
        for (int i = 0; i < n && kind == 0; ++i) {
            const int m = rowptr[i + 1] - rowptr[i];
            if (m != 13 && m != 9 && m != 5) continue;
            int mo = 0;
            for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) mo = std::max(mo, col[k] - i);
            if ((m == 13 || m == 9) && mo >= 4 && mo % 2 == 0) { maxoff = mo; kind = 13; }
            else if (m == 5 && mo == 2) { maxoff = mo; kind = 5; }
        }
        if (kind == 13) {
            const int nxp = maxoff / 2;
            if (n % nxp != 0) return 2;
            G.dim = 2; G.nx = nxp - 4; G.ny = n / nxp - 4;
        } else if (kind == 5) {
            G.dim = 1; G.nx = n - 4; G.ny = 1;
        } else {
            return 2;
        }
        G.bc[0] = G.bc[1] = G.bc[2] = G.bc[3] = -1;
    }
    G.nxp = G.nx + 4;
    G.nyp = G.dim == 2 ? G.ny + 4 : 1;
    if (G.nx < 3 || (G.dim == 2 && G.ny < 3) || G.nxp * G.nyp != n) return 2;
    return 0;
}

StencilOp lin_op(const LinGrid &G) {
    StencilOp op;
    std::memset(&op, 0, sizeof(op));
    op.dim = G.dim; op.nx = G.nx; op.ny = G.dim == 2 ? G.ny : 1; op.N = op.nx * op.ny;
    op.ns = G.dim == 2 ? 13 : 5;
    op.perx = G.bc[0] == YZ_BC_PERIODIC; op.pery = G.dim == 2 && G.bc[2] == YZ_BC_PERIODIC;
    for (int k = 0; k < op.ns; ++k) { op.ox[k] = slot_ox(G.dim, k); op.oy[k] = slot_oy(G.dim, k); }
    return op;
}

constexpr int kNotGrid = 99;   // the rows do not fold onto a structured grid

#define LIN_TRY(x)                                                        \
    do {                                                                  \
        if ((x) != hipSuccess) { yz::set_error("HIP error in the linear-solve seam"); return 2; } \
    } while (0)

// the solve on device buffers B (sized for the system) and solver S
int lin_solve(Solver &S, LinBufs &B, LinGrid G, const int *rowptr, const int *col, const double *val,
              const double *rhs, double *x_inout, const yz_lin_settings *s, int *iter, double *resid,
              hipStream_t st) {
    const int n = G.n, nnz = G.nnz;
    LIN_TRY(hipMemcpyAsync(B.rowptr, rowptr, (size_t)(n + 1) * sizeof(int), hipMemcpyHostToDevice, st));
    LIN_TRY(hipMemcpyAsync(B.col, col, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice, st));
    LIN_TRY(hipMemcpyAsync(B.val, val, (size_t)nnz * sizeof(double), hipMemcpyHostToDevice, st));
    LIN_TRY(hipMemcpyAsync(B.rhs, rhs, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
    LIN_TRY(hipMemcpyAsync(B.xin, x_inout, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
    int zero2[2] = {0, 0};
    LIN_TRY(hipMemcpyAsync(B.flag, zero2, sizeof(zero2), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_lin_ghosts, dim3(lblk(n)), dim3(LB), 0, st, G, B.rowptr, B.col, B.val, B.rhs, B.gsrc,
                       B.gofs, B.flag);
    // the edges' kinds from where the first ghosts copy from, when not given
    if (G.bc[0] < 0) {
        int src[2] = {-1, -1};
        const int r0 = (G.dim == 2 ? 2 : 0) * G.nxp;      // first ghost of the first real row
        LIN_TRY(hipMemcpyAsync(&src[0], B.gsrc + r0, sizeof(int), hipMemcpyDeviceToHost, st));
        if (G.dim == 2) LIN_TRY(hipMemcpyAsync(&src[1], B.gsrc + 2, sizeof(int), hipMemcpyDeviceToHost, st));
        LIN_TRY(hipStreamSynchronize(st));
        G.bc[0] = G.bc[1] = (src[0] >= 0 && src[0] % G.nxp == G.nxp - 4) ? YZ_BC_PERIODIC : YZ_BC_SYMMETRY;
        if (G.dim == 2) G.bc[2] = G.bc[3] = (src[1] >= 0 && src[1] / G.nxp == G.nyp - 4) ? YZ_BC_PERIODIC : YZ_BC_SYMMETRY;
        else G.bc[2] = G.bc[3] = YZ_BC_SYMMETRY;
    }
    const StencilOp op = lin_op(G);
    if (solver_init(S, op, st)) { yz::set_error("solver set-up failed"); return 2; }
    hipLaunchKernelGGL(k_lin_rows, dim3(lblk(op.N)), dim3(LB), 0, st, G, op, s->scaling, B.rowptr, B.col, B.val,
                       B.rhs, B.xin, B.gsrc, B.gofs, S.A, B.b, B.x, B.flag);
    int bad = 0;
    LIN_TRY(hipMemcpyAsync(&bad, B.flag, sizeof(int), hipMemcpyDeviceToHost, st));
    LIN_TRY(hipStreamSynchronize(st));
    if (bad) return kNotGrid;
    const int pc = s->precond == YZ_PC_AMG ? YZ_PC_AMG : YZ_PC_JACOBI;
    int it = 0, tried = 0;
    double rr = 0;
    int rc = -1;
    if (s->conc && pc == YZ_PC_AMG) {
        // a concentration system: damped Jacobi first, at most 300
        // iterations; from the guess again with the multigrid if it fails
        if (jacobi_setup(S) != 0) return 2;
        rc = bicgstab(S, B.b, B.x, YZ_PC_JACOBI, std::min(s->max_iter, 300), s->epsilon, &it, &rr);
        if (rc != 0) {
            tried = it;
            hipLaunchKernelGGL(k_lin_rows, dim3(lblk(op.N)), dim3(LB), 0, st, G, op, s->scaling, B.rowptr, B.col,
                               B.val, B.rhs, B.xin, B.gsrc, B.gofs, S.A, B.b, B.x, B.flag);
        }
    }
    if (rc != 0) {
        if ((pc == YZ_PC_AMG ? amg_setup(S) : jacobi_setup(S)) != 0) return 2;
        rc = bicgstab(S, B.b, B.x, pc, s->max_iter, s->epsilon, &it, &rr);
        it += tried;
    }
    LIN_TRY(hipMemcpyAsync(B.xout, B.xin, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_lin_out_interior, dim3(lblk(op.N)), dim3(LB), 0, st, G, op, B.x, B.xout);
    hipLaunchKernelGGL(k_lin_out_ghosts, dim3(lblk(n)), dim3(LB), 0, st, G, B.gsrc, B.gofs, B.xout);
    LIN_TRY(hipMemcpyAsync(x_inout, B.xout, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
    LIN_TRY(hipStreamSynchronize(st));
    if (iter) *iter = it;
    if (resid) *resid = rr;
    return rc;
}

// the arrays of a system of n rows and nnz entries, in a row, 256-aligned:
// from the context's allocation or from an arena
template <class Carve>
bool lin_bufs(LinBufs &B, int n, int nnz, int N, Carve carve) {
    B.rowptr = (int *)carve((size_t)(n + 1) * sizeof(int));
    B.col = (int *)carve((size_t)nnz * sizeof(int));
    B.gsrc = (int *)carve((size_t)n * sizeof(int));
    B.flag = (int *)carve(2 * sizeof(int));
    B.val = (double *)carve((size_t)nnz * sizeof(double));
    B.rhs = (double *)carve((size_t)n * sizeof(double));
    B.xin = (double *)carve((size_t)n * sizeof(double));
    B.gofs = (double *)carve((size_t)n * sizeof(double));
    B.xout = (double *)carve((size_t)n * sizeof(double));
    B.b = (double *)carve((size_t)N * sizeof(double));
    B.x = (double *)carve((size_t)N * sizeof(double));
    return B.rowptr && B.col && B.gsrc && B.flag && B.val && B.rhs && B.xin && B.gofs && B.xout && B.b && B.x;
}
size_t lin_bufs_bytes(int n, int nnz, int N) {
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return up((size_t)(n + 1) * 4) + up((size_t)nnz * 4) + up((size_t)n * 4) + up(8) + up((size_t)nnz * 8) +
           5 * up((size_t)n * 8) + 2 * up((size_t)N * 8);
}

// ---- the scratch of the no-allocation entry point ---------------------
// Every GPU the library runs on holds this array from the module's load
// until the process ends. Host code must not use its name as an address:
// yz_linsolve_noalloc takes the device address of this GPU's copy with
// hipGetSymbolAddress, once per device, and uses that.
constexpr size_t kScratchBytes = (size_t)96 << 20;
__device__ __align__(256) char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
char *g_scratch[kMaxDevices] = {};

// ---------------------------------------------------------------------
// A CRS with no structured grid: the system as it is on the device, its
// rows scaled by their diagonals (ScaleRows), and BiCGSTAB with a
// point-wise preconditioner: SPAI0 (m_i = a_ii / sum_j a_ij^2, AMGCL's
// relaxation; the multigrid setting, for want of an algebraic multigrid on
// the device) or damped Jacobi 0.72 / a_ii (the Jacobi and None settings,
// LinearSolverBackend.cpp:662-679). A product y = A x gives each row a
// group of G lanes of a wave (G from the mean row length, up to the whole
// wave for long rows); the lanes stride over the row's entries and sum in a
// fixed shuffle tree. Dot products: fixed block partials summed in order on
// the host, so a solve gives the same bits every run.
// ---------------------------------------------------------------------
constexpr int CRB = 256;          // reduction blocks
template <int G>
__global__ void __launch_bounds__(256) k_crs_spmv(int n, const int *__restrict__ rp, const int *__restrict__ col,
                                                  const double *__restrict__ val, const double *__restrict__ x,
                                                  double *__restrict__ y) {
    const int lane = threadIdx.x % G;
    const int row = (blockIdx.x * 256 + threadIdx.x) / G;
    double s = 0.0;
    if (row < n)
        for (int k = rp[row] + lane; k < rp[row + 1]; k += G) s += val[k] * x[col[k]];
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) s += __shfl_down(s, o, G);
    if (row < n && lane == 0) y[row] = s;
}
// the scaled copy of the system and the preconditioner's diagonal
__global__ void k_crs_prep(int n, int scaling, int spai, const int *__restrict__ rp, const int *__restrict__ col,
                           const double *__restrict__ val, const double *__restrict__ rhs, double *__restrict__ vs,
                           double *__restrict__ b, double *__restrict__ m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) if (col[k] == i) d += val[k];
    const bool sc = scaling && !(fabs(d) < 1.0e-30);
    const double inv = sc ? 1.0 / d : 1.0;
    double dd = 0.0, ss = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
        const double v = sc ? val[k] * inv : val[k];
        vs[k] = v;
        ss += v * v;
        if (col[k] == i) dd += v;
    }
    b[i] = sc ? rhs[i] * inv : rhs[i];
    m[i] = spai ? (ss > 0 ? dd / ss : 0.0) : (dd != 0 ? 0.72 / dd : 0.0);
}
// partial sums of a.b and c.d over fixed blocks
__global__ void __launch_bounds__(256) k_crs_dot2(int n, const double *a, const double *b, const double *c,
                                                  const double *d, double *part) {
    __shared__ double sh[2][256];
    double u = 0, v = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += CRB * 256) {
        u += a[i] * b[i];
        if (c) v += c[i] * d[i];
    }
    sh[0][threadIdx.x] = u; sh[1][threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) { sh[0][threadIdx.x] += sh[0][threadIdx.x + w]; sh[1][threadIdx.x] += sh[1][threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[blockIdx.x] = sh[0][0]; part[CRB + blockIdx.x] = sh[1][0]; }
}
// y = a x + b y + c z (what BiCGSTAB's updates need), and y = m x
__global__ void k_crs_lin(int n, double a, const double *x, double b, double *y, double c, const double *z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = a * x[i];
    if (b != 0.0) v += b * y[i];
    if (z) v += c * z[i];
    y[i] = v;
}
__global__ void k_crs_diag(int n, const double *m, const double *x, double *y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = m[i] * x[i];
}

struct CrsSolve {
    int n, G;
    const int *rp, *col;
    const double *val;
    double *part, *hpart;
    hipStream_t st;
    void spmv(const double *x, double *y) const {
        const int blocks = (int)(((long)n * G + 255) / 256);
        switch (G) {
        case 2: hipLaunchKernelGGL(k_crs_spmv<2>, dim3(blocks), dim3(256), 0, st, n, rp, col, val, x, y); break;
        case 4: hipLaunchKernelGGL(k_crs_spmv<4>, dim3(blocks), dim3(256), 0, st, n, rp, col, val, x, y); break;
        case 8: hipLaunchKernelGGL(k_crs_spmv<8>, dim3(blocks), dim3(256), 0, st, n, rp, col, val, x, y); break;
        case 16: hipLaunchKernelGGL(k_crs_spmv<16>, dim3(blocks), dim3(256), 0, st, n, rp, col, val, x, y); break;
        case 32: hipLaunchKernelGGL(k_crs_spmv<32>, dim3(blocks), dim3(256), 0, st, n, rp, col, val, x, y); break;
        default: hipLaunchKernelGGL(k_crs_spmv<64>, dim3(blocks), dim3(256), 0, st, n, rp, col, val, x, y); break;
        }
    }
    int dot2(const double *a, const double *b, const double *c, const double *d, double &u, double &v) const {
        hipLaunchKernelGGL(k_crs_dot2, dim3(CRB), dim3(256), 0, st, n, a, b, c, d, part);
        if (hipMemcpyAsync(hpart, part, 2 * CRB * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return -1;
        u = 0; v = 0;
        for (int k = 0; k < CRB; ++k) { u += hpart[k]; v += hpart[CRB + k]; }
        return 0;
    }
    void lin(double a, const double *x, double b, double *y, double c = 0, const double *z = nullptr) const {
        hipLaunchKernelGGL(k_crs_lin, dim3(lblk(n)), dim3(LB), 0, st, n, a, x, b, y, c, z);
    }
};

// BiCGSTAB with right preconditioning by the diagonal m: the method and the
// tests of the stencil solver (0 converged, 1 iteration cap, 2 breakdown)
int crs_bicgstab(const CrsSolve &K, const double *m, const double *b, double *x, double **w, int maxit, double tol,
                 int *iters, double *resid) {
    const int n = K.n;
    double *r = w[0], *rh = w[1], *p = w[2], *v = w[3], *sv = w[4], *t = w[5], *ph = w[6], *sh = w[7];
    *iters = 0; *resid = 0;
    K.spmv(x, r);
    K.lin(1.0, b, -1.0, r);                     // r = b - A x
    if (hipMemcpyAsync(rh, r, n * sizeof(double), hipMemcpyDeviceToDevice, K.st) != hipSuccess) return 2;
    double bb, rr;
    if (K.dot2(b, b, r, r, bb, rr)) return 2;
    const double nb = std::sqrt(bb);
    if (nb == 0) { (void)hipMemsetAsync(x, 0, n * sizeof(double), K.st); return 0; }
    const double eps = tol * nb;
    double res = std::sqrt(rr);
    if (res <= eps) { *resid = res / nb; return 0; }
    double rho1 = rr, rho2 = 1, alpha = 1, omega = 1, dum;
    int it = 0;
    for (; it < maxit; ++it) {
        if (it > 0 && std::fabs(rho1) < 1e-300) { *iters = it; *resid = res / nb; return 2; }
        if (it == 0) K.lin(1.0, r, 0.0, p);
        else {
            const double beta = (rho1 * alpha) / (rho2 * omega);
            K.lin(1.0, r, beta, p, -beta * omega, v);   // p = r + beta (p - omega v)
        }
        hipLaunchKernelGGL(k_crs_diag, dim3(lblk(n)), dim3(LB), 0, K.st, n, m, p, ph);
        K.spmv(ph, v);
        double rv;
        if (K.dot2(rh, v, nullptr, nullptr, rv, dum)) return 2;
        if (rv == 0) { *iters = it; *resid = res / nb; return 2; }
        alpha = rho1 / rv;
        K.lin(1.0, r, 0.0, sv, -alpha, v);          // s = r - alpha v
        double ss;
        if (K.dot2(sv, sv, nullptr, nullptr, ss, dum)) return 2;
        res = std::sqrt(ss);
        if (res < eps) { K.lin(alpha, ph, 1.0, x); ++it; break; }
        hipLaunchKernelGGL(k_crs_diag, dim3(lblk(n)), dim3(LB), 0, K.st, n, m, sv, sh);
        K.spmv(sh, t);
        double tt, ts;
        if (K.dot2(t, t, t, sv, tt, ts)) return 2;
        omega = tt != 0 ? ts / tt : 0.0;
        if (omega == 0) { *iters = it; *resid = res / nb; return 2; }
        K.lin(alpha, ph, 1.0, x, omega, sh);       // x += alpha ph + omega sh
        K.lin(1.0, sv, 0.0, r, -omega, t);          // r = s - omega t
        double rrh;
        if (K.dot2(r, r, r, rh, rr, rrh)) return 2;
        rho2 = rho1;
        rho1 = rrh;
        res = std::sqrt(rr);
        if (res < eps) { ++it; break; }
        if (!(res == res)) { *iters = it; *resid = res; return 2; }
    }
    *iters = it;
    *resid = res / nb;
    if (!(*resid == *resid)) return 2;
    return *resid <= tol ? 0 : 1;
}

}  // namespace

struct yz_linctx {
    Solver S;
    hipStream_t st = nullptr;
    LinBufs B;
    char *mem = nullptr;
    size_t mem_bytes = 0;
    char *crs_mem = nullptr;       // the general CRS path's arrays
    size_t crs_bytes = 0;
    double *crs_host = nullptr;    // pinned block partials
};

namespace {
// the general CRS path (no structured grid): the system, its scaled copy,
// the preconditioner's diagonal and the Krylov vectors on the device
int lin_solve_crs(yz_linctx *ctx, int n, int nnz, const int *rowptr, const int *col, const double *val,
                  const double *rhs, double *x_inout, const yz_lin_settings *s, int *iter, double *resid) {
    hipStream_t st = ctx->st;
    const size_t need = (size_t)(n + 1) * 4 + (size_t)nnz * 4 + 2 * (size_t)nnz * 8 + 12 * (size_t)n * 8 +
                        2 * CRB * 8 + 16 * 256;
    if (need > ctx->crs_bytes) {
        if (ctx->crs_mem) (void)hipFree(ctx->crs_mem);
        ctx->crs_mem = nullptr;
        ctx->crs_bytes = 0;
        if (hipMalloc((void **)&ctx->crs_mem, need) != hipSuccess) { yz::set_error("device allocation failed"); return 2; }
        ctx->crs_bytes = need;
    }
    if (!ctx->crs_host && hipHostMalloc((void **)&ctx->crs_host, 2 * CRB * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        yz::set_error("host allocation failed");
        return 2;
    }
    size_t used = 0;
    auto carve = [&](size_t b) -> void * {
        const size_t off = (used + 255) & ~(size_t)255;
        used = off + b;
        return ctx->crs_mem + off;
    };
    int *drp = (int *)carve((size_t)(n + 1) * 4), *dcol = (int *)carve((size_t)nnz * 4);
    double *dval = (double *)carve((size_t)nnz * 8), *vs = (double *)carve((size_t)nnz * 8);
    double *drhs = (double *)carve((size_t)n * 8), *b = (double *)carve((size_t)n * 8);
    double *x = (double *)carve((size_t)n * 8), *m = (double *)carve((size_t)n * 8);
    double *w[8];
    for (double *&v : w) v = (double *)carve((size_t)n * 8);
    double *part = (double *)carve(2 * CRB * 8);
    LIN_TRY(hipMemcpyAsync(drp, rowptr, (size_t)(n + 1) * sizeof(int), hipMemcpyHostToDevice, st));
    LIN_TRY(hipMemcpyAsync(dcol, col, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice, st));
    LIN_TRY(hipMemcpyAsync(dval, val, (size_t)nnz * sizeof(double), hipMemcpyHostToDevice, st));
    LIN_TRY(hipMemcpyAsync(drhs, rhs, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
    LIN_TRY(hipMemcpyAsync(x, x_inout, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
    const int spai = s->precond == YZ_PC_AMG ? 1 : 0;
    hipLaunchKernelGGL(k_crs_prep, dim3(lblk(n)), dim3(LB), 0, st, n, s->scaling, spai, drp, dcol, dval, drhs, vs, b, m);
    // lanes per row: the power of two nearest half the mean row length, 2 .. 64
    int G = 2;
    while (G < 64 && 2 * G < (nnz + n - 1) / n) G *= 2;
    CrsSolve K{n, G, drp, dcol, vs, part, ctx->crs_host, st};
    int it = 0;
    double rr = 0;
    const int rc = crs_bicgstab(K, m, b, x, w, s->max_iter, s->epsilon, &it, &rr);
    LIN_TRY(hipMemcpyAsync(x_inout, x, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
    LIN_TRY(hipStreamSynchronize(st));
    if (iter) *iter = it;
    if (resid) *resid = rr;
    return rc;
}
}  // namespace

extern "C" {

yz_linctx *yz_linsolve_create(void) {
    yz_linctx *c = new yz_linctx();
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) { delete c; return nullptr; }
    return c;
}

void yz_linsolve_destroy(yz_linctx *c) {
    if (!c) return;
    solver_free(c->S);
    if (c->mem) (void)hipFree(c->mem);
    if (c->crs_mem) (void)hipFree(c->crs_mem);
    if (c->crs_host) (void)hipHostFree(c->crs_host);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
}

int yz_memory_get(yz_linctx *ctx, double *buf, int cap) {
    (void)ctx; (void)buf; (void)cap;
    return 0;   // no solve-to-solve state
}
int yz_memory_set(yz_linctx *ctx, const double *buf, int len) {
    (void)ctx; (void)buf;
    return len == 0 ? 0 : 1;
}

int yz_linsolve(yz_linctx *ctx, int n, int nnz, const int *rowptr, const int *col, const double *val,
                const double *rhs, double *x_inout, const yz_lin_settings *s, int *iter, double *resid) {
    if (iter) *iter = 0;
    if (resid) *resid = 0;
    if (!ctx || !rowptr || !col || !val || !rhs || !x_inout || !s || n <= 0 || nnz < n) return 2;
    LinGrid G;
    // a matrix with no film-height or concentration grid takes the general path
    if (lin_grid(n, nnz, rowptr, col, s, G))
        return lin_solve_crs(ctx, n, nnz, rowptr, col, val, rhs, x_inout, s, iter, resid);
    const int N = G.nx * (G.dim == 2 ? G.ny : 1);
    const size_t need = lin_bufs_bytes(n, nnz, N);
    if (need > ctx->mem_bytes) {
        if (ctx->mem) (void)hipFree(ctx->mem);
        ctx->mem = nullptr;
        ctx->mem_bytes = 0;
        if (hipMalloc((void **)&ctx->mem, need) != hipSuccess) { yz::set_error("device allocation failed"); return 2; }
        ctx->mem_bytes = need;
    }
    size_t used = 0;
    char *base = ctx->mem;
    auto carve = [&](size_t b) -> void * {
        const size_t off = (used + 255) & ~(size_t)255;
        used = off + b;
        return base + off;
    };
    lin_bufs(ctx->B, n, nnz, N, carve);
    const int rc = lin_solve(ctx->S, ctx->B, G, rowptr, col, val, rhs, x_inout, s, iter, resid, ctx->st);
    if (rc == kNotGrid) return lin_solve_crs(ctx, n, nnz, rowptr, col, val, rhs, x_inout, s, iter, resid);
    return rc;
}

int yz_linsolve_noalloc(int n, int nnz, const int *rowptr, const int *col, const double *val, const double *rhs,
                        double *x_inout, const yz_lin_settings *s, int *iter, double *resid) {
    if (iter) *iter = 0;
    if (resid) *resid = 0;
    if (!rowptr || !col || !val || !rhs || !x_inout || !s || n <= 0 || nnz < n) return 2;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) { yz::set_error("no device"); return 2; }
    if (!g_scratch[dev]) {
        // this device's copy of the scratch array, with the device current
        char *p = nullptr;
        if (hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            yz::set_error("the scratch array's device address is not available");
            return 2;
        }
        g_scratch[dev] = p;
    }
    LinGrid G;
    if (lin_grid(n, nnz, rowptr, col, s, G)) { yz::set_error("no structured grid in the matrix"); return 2; }
    const int N = G.nx * (G.dim == 2 ? G.ny : 1);
    DevArena arena;
    arena.base = g_scratch[dev];
    arena.cap = kScratchBytes;
    solver_use_arena(&arena);
    LinBufs B;
    auto too_big = [&]() {
        yz::set_error("the system exceeds the no-allocation scratch (" + std::to_string(kScratchBytes >> 20) +
                      " MiB, about 110 000 interior cells)");
        return 2;
    };
    lin_bufs(B, n, nnz, N, [](size_t b) { return arena_carve(b); });
    if (arena.overflow) { solver_use_arena(nullptr); return too_big(); }
    // a solver made afresh in the arena on every call (the default stream)
    static Solver S;
    solver_free(S);
    int rc = lin_solve(S, B, G, rowptr, col, val, rhs, x_inout, s, iter, resid, nullptr);
    solver_use_arena(nullptr);
    solver_free(S);
    if (arena.overflow) rc = too_big();
    if (rc == kNotGrid) {
        yz::set_error("the matrix is not a structured film or concentration system (the general CRS path "
                      "allocates: yz_linsolve)");
        rc = 2;
    }
    return rc;
}

}  // extern "C"
