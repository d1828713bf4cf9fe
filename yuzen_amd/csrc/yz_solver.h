// yz_solver.h -- the device linear solver of the film-height step:
// BiCGSTAB with right preconditioning over a structured stencil operator,
// preconditioned by a structured smoothed-aggregation multigrid, damped
// Jacobi, or an explicit 13-point stencil M (the learned approximate
// inverse), alone or as the multigrid's finest smoother. Templated on the
// value type: FP64 for the plain method, FP32 for the inexact-Newton inner
// solve (dot products and the Krylov scalars stay FP64, and so does the
// coarsest level's dense solve).
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <vector>

#include "yz_stencil.h"

namespace yz {

template <typename T>
struct AmgLevel {
    StencilOp op;
    T *val = nullptr;     // ns planes x N
    T *minv = nullptr;    // SPAI0 diagonal
    T *P = nullptr;       // 9 planes x N (1-D: 3): prolongator to the next level
    T *R = nullptr;       // 81 planes x Nc (1-D: 9): P^T by coarse cell, for the restriction
    T *AP = nullptr;      // 25 planes x N (1-D: 5): A P, set-up scratch
    T *x = nullptr, *b = nullptr, *r = nullptr, *t = nullptr;
    int ncx = 0, ncy = 0; // next level's grid
};

template <typename T>
struct Amg {
    std::vector<AmgLevel<T>> lv;   // lv[0].val aliases the fine operator
    double *coarse_inv = nullptr;  // dense inverse of the coarsest level (transposed), FP64
    double *coarse_dense = nullptr;
    int nco = 0;
    double *scal = nullptr;        // per level: Gershgorin bound of D^-1 A (bits)
    int built = 0;
};

// device and pinned tables of one launch of the fused iteration (the
// films, the block-to-film map, the state read) and, for a single film,
// its captured HIP graph with what the capture baked in
struct FusedTable {
    void *d_films = nullptr, *h_films = nullptr;
    void *d_blk[2] = {nullptr, nullptr}, *h_blk[2] = {nullptr, nullptr};
    double *d_gather = nullptr, *h_gather = nullptr;
    int cap_films = 0, cap_blk = 0;
    hipGraphExec_t gexec = nullptr;
    std::vector<long long> g_key;   // what the captured graph was made for
};
void fused_table_free(FusedTable &T);
// the fused iteration's tile rows for a film of nx x ny cells (the tile
// tuner's choice, 16 when it has none), and setting that choice (8 or 16)
int fused_tile_rows(int nx, int ny);
int fused_tile_set(int nx, int ny, int rows);

template <typename T>
struct Krylov {
    int N = 0;
    T *r = nullptr, *rh = nullptr, *p = nullptr, *v = nullptr, *s = nullptr, *t = nullptr,
      *ph = nullptr, *sh = nullptr;
    double *part = nullptr;    // reduction partials
    double *sc = nullptr;      // device scalars
    double *host_sc = nullptr; // pinned mirror
    // the fused iteration of the explicit-stencil preconditioner
    // (bicgstab_stencil): second search direction, per-block partials
    T *p2 = nullptr;
    double *fpart = nullptr;
    int ftiles = 0;
    FusedTable table;   // the launch tables when this solver runs alone
};

template <typename T>
struct SolverT {
    StencilOp op0;             // the fine operator
    T *A = nullptr;            // fine planes (row-scaled)
    T *dinv_jac = nullptr;
    const T *M = nullptr;      // explicit 13-point preconditioner / finest smoother (planes), or null
    // the same preconditioner as FP16 raw rows and an FP32 row scale,
    // M = diag(Mscale) Mh (the fused FP32 iteration reads these)
    const __half *Mh = nullptr;
    const float *Mscale = nullptr;
    Amg<T> amg;
    Krylov<T> kr;
    hipStream_t stream = nullptr;
    int max_levels = 12;
    // launch timing of the fine-level stencil product of the iteration
    // (k_spmv_dot), with HIP events on the solver's stream
    int ktime = 0;
    hipEvent_t kev[2] = {nullptr, nullptr};
    long long kcount = 0;
    double kms = 0;
    // the fused FP32 iteration's K1 timing slots (device) and the device
    // clock's rate (kHz)
    unsigned long long *kts = nullptr;
    double kclock_khz = 0;
    // the stall test of the learned preconditioner: give up when the
    // residual has not fallen by stall_factor over stall_window iterations
    int stall_window = 0;
    double stall_factor = 0.5;
    // replay the fused iteration from a captured HIP graph
    int use_graph = 0;
    // every row's diagonal of A is stored as exactly 1 (row-scaled FP32
    // planes): the fused iteration does not read that plane
    int unit_diag = 0;
    // the device arrays were carved from a caller's arena (solver_free
    // then frees nothing)
    int in_arena = 0;
};

// device memory given by a caller: while one is in use (solver_use_arena),
// solver_init carves the solver's arrays from it, 256-byte aligned, and
// calls no allocator; null ends it
struct DevArena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    bool overflow = false;   // a request did not fit
};
void solver_use_arena(DevArena *a);
// the next block of n bytes of the arena in use, or null when none is in
// use or it is full
void *arena_carve(size_t n);

typedef SolverT<double> Solver;

// allocate (idempotent for the same grid)
template <typename T> int solver_init(SolverT<T> &S, const StencilOp &op0, hipStream_t stream);
template <typename T> void solver_free(SolverT<T> &S);
// multigrid set-up on the current fine planes S.A
template <typename T> int amg_setup(SolverT<T> &S);
// Jacobi set-up
template <typename T> int jacobi_setup(SolverT<T> &S);
// BiCGSTAB: x (device, N) is the initial guess and the result; b (device);
// precond YZ_PC_*. Returns 0 / 1 / 2 as the backend contract, 3 when the
// stall test stopped it; iters and the relative residual out.
template <typename T>
int bicgstab(SolverT<T> &S, const T *b, T *x, int precond, int maxit, double tol, int *iters,
             double *resid);
// r = b - A x on the fine planes S.A (the solver's own row products), with
// (b, b) and (r, r) in fixed-order sums; synchronises the solver's stream
template <typename T>
int residual_norms(SolverT<T> &S, const T *x, const T *b, T *r, double *bb, double *rr);

// BiCGSTAB right-preconditioned by the explicit 13-point stencil S.M, with
// the iteration fused into five kernels per step and no host round trip
// inside a chunk of iterations (2-D, 13-point operators; others go to
// bicgstab). FP32. Same contract as bicgstab.
int bicgstab_stencil(SolverT<float> &S, const float *b, float *x, int maxit, double tol, int *iters,
                     double *resid);
// the same for nf films at once on stream st, each with its own solver,
// right-hand side and guess/result: one launch per kernel covers every film
// still running, and a film's arithmetic is the same as alone. rc, iters,
// resid per film (bicgstab's codes); returns nonzero on a device error or
// an operator it cannot take.
int bicgstab_stencil_batch(FusedTable &T, int nf, SolverT<float> *const *S, const float *const *b,
                           float *const *x, int maxit, const double *tol, int *iters, double *resid, int *rc,
                           hipStream_t st);

template <typename T>
void launch_spmv(const StencilOp &op, const T *val, const T *x, T *y, hipStream_t st);

}  // namespace yz
