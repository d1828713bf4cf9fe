// yz_stencil.h -- structured stencil operators on one GPU.
//
// A level of the linear system lives on a structured grid of nx x ny
// cells (interior cells only: the reference's ghost rows are exact copies
// of interior rows, x_ghost = x_source, and are eliminated). An operator
// is held as ns planes of values, val[k][i], one plane per stencil slot,
// so every slot is a coalesced stream and there are no column indices to
// read: the column of slot k of row i is the cell at offset (ox[k], oy[k])
// folded back into the grid along each axis -- mirrored at a Symmetry edge
// (ghost -1 is cell 0, -2 is cell 1), wrapped at a Periodic one. Several
// slots of a row near an edge may fold onto the same column; their values
// simply add in a product.
//
// Fine level: the 13-point film-height stencil (5-point in 1-D). Coarse
// levels of the multigrid: the 5x5 box (5 in 1-D).
#pragma once
#include <hip/hip_runtime.h>

namespace yz {

enum { MAXS = 25 };

struct StencilOp {
    int dim;          // 1 or 2
    int nx, ny, N;    // grid, N = nx*ny
    int ns;           // slots
    int perx, pery;   // 1: periodic axis
    int ox[MAXS], oy[MAXS];
};

__host__ __device__ inline int fold_axis(int j, int n, int periodic) {
    if (j >= 0 && j < n) return j;
    if (periodic) { j %= n; if (j < 0) j += n; return j; }
    if (j < 0) j = -1 - j;
    if (j >= n) j = 2 * n - 1 - j;
    if (j < 0) j = 0;
    if (j >= n) j = n - 1;
    return j;
}

__host__ __device__ inline int stencil_col(const StencilOp &op, int ix, int iy, int k) {
    const int jx = fold_axis(ix + op.ox[k], op.nx, op.perx);
    const int jy = (op.dim == 2) ? fold_axis(iy + op.oy[k], op.ny, op.pery) : 0;
    return jy * op.nx + jx;
}

// the column of slot k when the unfolded neighbour lies inside the grid
// (fast path), -1 otherwise
__host__ __device__ inline bool stencil_inside(const StencilOp &op, int ix, int iy, int k) {
    const int jx = ix + op.ox[k];
    if (jx < 0 || jx >= op.nx) return false;
    if (op.dim == 2) {
        const int jy = iy + op.oy[k];
        if (jy < 0 || jy >= op.ny) return false;
    }
    return true;
}

}  // namespace yz
