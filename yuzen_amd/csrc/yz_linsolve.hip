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
    if (bad) { yz::set_error("the matrix is not a structured film or concentration system"); return 2; }
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

}  // namespace

struct yz_linctx {
    Solver S;
    hipStream_t st = nullptr;
    LinBufs B;
    char *mem = nullptr;
    size_t mem_bytes = 0;
};

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
    if (lin_grid(n, nnz, rowptr, col, s, G)) { yz::set_error("no structured grid in the matrix"); return 2; }
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
    return lin_solve(ctx->S, ctx->B, G, rowptr, col, val, rhs, x_inout, s, iter, resid, ctx->st);
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
    return rc;
}

}  // extern "C"
