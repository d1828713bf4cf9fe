// yz_solver.h -- the device linear solver of the film-height step:
// BiCGSTAB with right preconditioning over a structured stencil operator,
// preconditioned by a structured smoothed-aggregation multigrid (SPAI0
// smoothing, dense inverse on the coarsest level) or damped Jacobi.
#pragma once
#include <hip/hip_runtime.h>
#include <vector>

#include "yz_stencil.h"

namespace yz {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// one multigrid level
struct AmgLevel {
    StencilOp op;
    double *val = nullptr;     // ns planes x N
    double *minv = nullptr;    // SPAI0 diagonal
    double *P = nullptr;       // 9 planes x N (1-D: 3): prolongator to the next level
    double *R = nullptr;       // 81 planes x Nc (1-D: 9): P^T by coarse cell, for the restriction
    double *AP = nullptr;      // 25 planes x N (1-D: 5): A P, set-up scratch
    double *x = nullptr, *b = nullptr, *r = nullptr;
    int ncx = 0, ncy = 0;      // next level's grid
};

struct Amg {
    std::vector<AmgLevel> lv;  // lv[0].val aliases the fine operator
    double *coarse_inv = nullptr;   // dense inverse of the coarsest level
    double *coarse_dense = nullptr; // dense matrix scratch
    int nco = 0;
    double *scal = nullptr;    // device scalars: [0] rho bound (as bits)
    int built = 0;
};

struct Krylov {
    int N = 0;
    double *r = nullptr, *rh = nullptr, *p = nullptr, *v = nullptr, *s = nullptr,
           *t = nullptr, *ph = nullptr, *sh = nullptr;
    double *part = nullptr;    // reduction partials
    double *sc = nullptr;      // device scalars
    double *host_sc = nullptr; // pinned mirror
};

struct Solver {
    StencilOp op0;             // the fine operator
    double *A = nullptr;       // fine planes (row-scaled)
    double *dinv_jac = nullptr;
    Amg amg;
    Krylov kr;
    hipStream_t stream = nullptr;
    int max_levels = 12;
    // launch timing of the fine-level stencil product of the iteration
    // (k_spmv_dot), with HIP events on the solver's stream
    int ktime = 0;
    hipEvent_t kev[2] = {nullptr, nullptr};
    long long kcount = 0;
    double kms = 0;
};

// allocate (idempotent for the same grid)
int solver_init(Solver &S, const StencilOp &op0, hipStream_t stream);
void solver_free(Solver &S);
// multigrid set-up on the current fine planes S.A
int amg_setup(Solver &S);
// Jacobi set-up
int jacobi_setup(Solver &S);
// BiCGSTAB: x (device, N) is the initial guess and the result; b (device);
// precond YZ_PC_AMG / YZ_PC_JACOBI. Returns 0 / 1 / 2 as the backend
// contract; iters and the relative residual out.
int bicgstab(Solver &S, const double *b, double *x, int precond, int maxit,
             double tol, int *iters, double *resid);

// device helpers used by the step
void launch_spmv(const StencilOp &op, const double *val, const double *x, double *y,
                 hipStream_t st);

}  // namespace yz
