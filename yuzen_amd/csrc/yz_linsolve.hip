// yz_linsolve.hip -- the linear-solve seam: one CRS system through the
// device BiCGSTAB, with the contract of ILinearSolverBackend::Solve
// (reference yuzen/src/numerics/LinearSolverBackend.h:56-103; the AMGCL
// backend's Solve, LinearSolverBackend.cpp:539-728): x is the initial guess
// and the result, the matrix and right-hand side are not modified, the
// codes are 0 converged / 1 iteration cap / 2 breakdown, and nothing
// throws across the boundary.
//
// The CRS is a film-height or well-mixed concentration matrix of a
// structured grid in the reference's row layout: two ghost layers per side
// whose rows are copies x_g - x_src = r_g (BoundaryCondition::EqualVal:
// Symmetry and Periodic edges), interior rows within the 13-point stencil.
// The grid comes from the settings or is read off the matrix itself (the
// column offsets of an interior row of the 13-point film-height pattern or
// of the 9-point wide concentration pattern, ConcEquationBase.h:242-258). The ghost rows are
// eliminated, the interior rows become stencil planes, and the system is
// solved on the device; the ghost entries of x are formed again from their
// rows afterwards, as CRS_Matrix::SatisfyBCRows makes them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/yuzen_gpu.h"
#include "yz_physics.h"
#include "yz_solver.h"

using namespace yz;

namespace {
thread_local std::string g_lin_err;
}

struct yz_linctx {
    Solver S;
    hipStream_t st = nullptr;
    double *b = nullptr, *x = nullptr;
    int N = 0;
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
    if (c->b) (void)hipFree(c->b);
    if (c->x) (void)hipFree(c->x);
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
    // ---- the grid: given, or read off the pattern ----
    int dim = s->grid_dim, nx = s->grid_nx, ny = s->grid_ny;
    int bc[4] = {s->bc[0], s->bc[1], s->bc[2], s->bc[3]};
    if (nx <= 0) {
        int maxoff = 0, cnt = 0;
        for (int i = 0; i < n && cnt == 0; ++i) {
            const int m = rowptr[i + 1] - rowptr[i];
            if (m != 13 && m != 9 && m != 5) continue;
            int mo = 0;
            for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) mo = std::max(mo, col[k] - i);
            if ((m == 13 || m == 9) && mo >= 4 && mo % 2 == 0) { maxoff = mo; cnt = 13; }
            else if (m == 5 && mo == 2) { maxoff = mo; cnt = 5; }
        }
        if (cnt == 13) {
            const int nxp = maxoff / 2;
            if (n % nxp != 0) return 2;
            dim = 2; nx = nxp - 4; ny = n / nxp - 4;
        } else if (cnt == 5) {
            dim = 1; nx = n - 4; ny = 1;
        } else return 2;
        bc[0] = bc[1] = bc[2] = bc[3] = -1;
    }
    const int nxp = nx + 4, nyp = dim == 2 ? ny + 4 : 1;
    if (nx < 3 || (dim == 2 && ny < 3) || nxp * nyp != n) return 2;
    auto inside = [&](int ix, int iy) {
        return ix >= 2 && ix <= nxp - 3 && (dim == 1 || (iy >= 2 && iy <= nyp - 3));
    };
    // ghost rows: x_g = x_src + r_g / d (diag d, the other entry -d)
    std::vector<int> gsrc(n, -1);
    std::vector<double> gofs(n, 0.0);
    for (int i = 0; i < n; ++i) {
        const int ix = i % nxp, iy = i / nxp;
        if (inside(ix, iy)) continue;
        double d = 0, o = 0;
        int src = -1, m = 0;
        for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
            if (val[k] == 0) continue;
            if (col[k] == i) d += val[k];
            else { src = col[k]; o += val[k]; ++m; }
        }
        if (m != 1 || d == 0 || o != -d) return 2;
        const int sx = src % nxp, sy = src / nxp;
        if (!inside(sx, sy)) return 2;
        gsrc[i] = src;
        gofs[i] = rhs[i] / d;
    }
    // boundary kinds from the ghost sources, when not given
    if (bc[0] < 0) {
        const int iy0 = dim == 2 ? 2 : 0;
        bc[0] = (gsrc[iy0 * nxp + 0] % nxp == nxp - 4) ? YZ_BC_PERIODIC : YZ_BC_SYMMETRY;
        bc[1] = bc[0];
        if (dim == 2) bc[2] = bc[3] = ((gsrc[0 * nxp + 2] / nxp) == nyp - 4) ? YZ_BC_PERIODIC : YZ_BC_SYMMETRY;
        else bc[2] = bc[3] = YZ_BC_SYMMETRY;
    }
    // ---- interior rows -> stencil planes ----
    StencilOp op;
    std::memset(&op, 0, sizeof(op));
    op.dim = dim; op.nx = nx; op.ny = dim == 2 ? ny : 1; op.N = op.nx * op.ny;
    op.ns = dim == 2 ? 13 : 5;
    op.perx = bc[0] == YZ_BC_PERIODIC; op.pery = dim == 2 && bc[2] == YZ_BC_PERIODIC;
    for (int k = 0; k < op.ns; ++k) { op.ox[k] = slot_ox(dim, k); op.oy[k] = slot_oy(dim, k); }
    const int N = op.N;
    std::vector<double> planes((size_t)op.ns * N, 0.0), b(N), x(N);
    for (int iy = 0; iy < op.ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            const int i = iy * nx + ix;
            const int r = (dim == 2 ? iy + 2 : 0) * nxp + ix + 2;
            double bi = rhs[r];
            for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) {
                const int cc = col[k];
                const int dxo = cc % nxp - (ix + 2), dyo = cc / nxp - (dim == 2 ? iy + 2 : 0);
                int slot = -1;
                for (int q = 0; q < op.ns; ++q)
                    if (op.ox[q] == dxo && op.oy[q] == dyo) { slot = q; break; }
                if (slot < 0) { if (val[k] != 0) return 2; continue; }
                planes[(size_t)slot * N + i] += val[k];
                if (gsrc[cc] >= 0) bi -= val[k] * gofs[cc];
            }
            b[i] = bi;
            x[i] = x_inout[r];
        }
    // row scaling by the stored diagonal
    if (s->scaling) {
        const int c0 = dim == 2 ? 6 : 2;
        for (int i = 0; i < N; ++i) {
            const double d = planes[(size_t)c0 * N + i];
            if (std::fabs(d) < 1.0e-30) continue;
            const double inv = 1.0 / d;
            for (int q = 0; q < op.ns; ++q) planes[(size_t)q * N + i] *= inv;
            b[i] *= inv;
        }
    }
    Solver &S = ctx->S;
    if (solver_init(S, op, ctx->st)) return 2;
    if (ctx->N != N) {
        if (ctx->b) (void)hipFree(ctx->b);
        if (ctx->x) (void)hipFree(ctx->x);
        ctx->b = ctx->x = nullptr;
        if (hipMalloc((void **)&ctx->b, N * sizeof(double)) != hipSuccess) return 2;
        if (hipMalloc((void **)&ctx->x, N * sizeof(double)) != hipSuccess) return 2;
        ctx->N = N;
    }
    if (hipMemcpyAsync(S.A, planes.data(), planes.size() * sizeof(double), hipMemcpyHostToDevice, ctx->st) != hipSuccess) return 2;
    if (hipMemcpyAsync(ctx->b, b.data(), N * sizeof(double), hipMemcpyHostToDevice, ctx->st) != hipSuccess) return 2;
    if (hipMemcpyAsync(ctx->x, x.data(), N * sizeof(double), hipMemcpyHostToDevice, ctx->st) != hipSuccess) return 2;
    const int pc = s->precond == YZ_PC_AMG ? YZ_PC_AMG : YZ_PC_JACOBI;
    int it = 0, tried = 0;
    double rr = 0;
    int rc = -1;
    if (s->conc && pc == YZ_PC_AMG) {
        // a concentration system: damped Jacobi first, at most 300
        // iterations; from the guess again with the multigrid if it fails
        if (jacobi_setup(S) != 0) return 2;
        rc = bicgstab(S, ctx->b, ctx->x, YZ_PC_JACOBI, std::min(s->max_iter, 300), s->epsilon, &it, &rr);
        if (rc != 0) {
            tried = it;
            if (hipMemcpyAsync(ctx->x, x.data(), N * sizeof(double), hipMemcpyHostToDevice, ctx->st) != hipSuccess) return 2;
        }
    }
    if (rc != 0) {
        if ((pc == YZ_PC_AMG ? amg_setup(S) : jacobi_setup(S)) != 0) return 2;
        rc = bicgstab(S, ctx->b, ctx->x, pc, s->max_iter, s->epsilon, &it, &rr);
        it += tried;
    }
    if (hipMemcpyAsync(x.data(), ctx->x, N * sizeof(double), hipMemcpyDeviceToHost, ctx->st) != hipSuccess) return 2;
    if (hipStreamSynchronize(ctx->st) != hipSuccess) return 2;
    for (int iy = 0; iy < op.ny; ++iy)
        for (int ix = 0; ix < nx; ++ix)
            x_inout[(dim == 2 ? iy + 2 : 0) * nxp + ix + 2] = x[iy * nx + ix];
    for (int i = 0; i < n; ++i)
        if (gsrc[i] >= 0) x_inout[i] = x_inout[gsrc[i]] + gofs[i];
    if (iter) *iter = it;
    if (resid) *resid = rr;
    return rc;
}

}  // extern "C"
