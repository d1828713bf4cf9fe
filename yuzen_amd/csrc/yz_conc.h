// yz_conc.h -- the well-mixed concentration on the device: the face volume
// fluxes of the film, the concentration system of a Newton pass or of the
// step's completion, the solute fluxes and the reconstruction of c from the
// balance of h c, the fluxes made consistent with the film's step, and the
// redistribution of what a front over- or undershoots.
//
// Reference (paths relative to yuzen/src/model/well_mixed/):
//   FaceVolumeFluxes            HeightEquationBase.h:852-894, each term's
//                               FaceFlux / SideRate (PhysicsTermBase.h:160-188)
//   the concentration system    ConcEquationBase.h:261-418 (Solve)
//   SoluteFlux, Reconstruct     ConcEquationBase.h:427-518
//   ConsistentFluxes            HeightEquationBase.h:985-1058
//   ClampConc, ConcCorrection   HeightEquationBase.h:1098-1109, 1292-1305
//   RedistributeExcess          HeightEquationBase.h:1128-1273
#pragma once
#include "yz_rows.h"
#include "yz_stencil.h"

namespace yz {

struct ConcArgs {
    Geo g;
    double dt;
    double theta_h;          // the film's theta (ConsistentFluxes)
    int conv, diff;          // EnableConvection, EnableDiffusion
    int side;                // the 1-D centrifugal term's sideways outflow is on
    double wpos[4], wneg[4]; // face weights c1 + Sign(F) c2 for F >= 0 / F < 0
    // a limited scheme (Koren, or van Leer: lim_vl): its weights frozen at
    // climit, the latest iterate (padded), and its face values of the
    // solution formed by the limiter itself
    int lim, lim_vl;
    const double *climit;
    double Df, diff_c0, diff_exp;
    int row_scaling;
    const double *h;         // the working film (the guess), padded
    const double *c;         // the working concentration (clamped), padded
    const double *hold;      // the film at the start of the step, padded
    const double *c0;        // the concentration at the start of the step, padded
    const double *F0[2], *F1[2], *r0, *r1;   // face fluxes, padded
    double *cx;              // the concentration of the solve, padded
    double *G2[2];           // solute fluxes, padded
    double *A, *b;           // the row-scaled system: ns planes x N, N
    // the step's sources (StepSources, HeightEquationBase.h:936-960), N each
    // or null: the evaporation rate and the nozzle's supply rate; supplied:
    // the supply arrives at src_conc, else at the local c
    const double *evp, *src;
    int supplied;
    double src_conc;
};

// FaceVolumeFluxes of the film a.h with concentration a.c: F[a][p] the flux
// through the high face of real cell p along axis a, and of the low face of
// a first cell at its ghost row; r[p] the sideways outflow
void launch_face_fluxes(const RowArgs &a, double *Fx, double *Fy, double *r, hipStream_t st);
// rows of the concentration system and its right-hand side
void launch_conc_rows(const ConcArgs &a, hipStream_t st);
// cx = c0 + d on the real cells, then the ghosts
void launch_conc_update(const ConcArgs &a, const double *d, hipStream_t st);
// Reconstruct: the solute fluxes of the solution, then c from h c
void launch_reconstruct(const ConcArgs &a, hipStream_t st);
// the ghosts of a padded array from the cells they copy
void launch_ghosts(const Geo &g, double *v, hipStream_t st);
// EveryNKLoop's tail over every row: c = min(cx, 1), dmax = max |c - cprev|,
// cprev = c (dmax: bits of a non-negative double, atomicMax)
void launch_conc_take(const Geo &g, const double *cx, double *c, double *cprev, unsigned long long *dmax,
                      hipStream_t st);
// the frozen limiter state settling at the step's completion, over every
// row: dmax = max |cx - climit| (bits of a non-negative double, atomicMax),
// then climit = cx
void launch_conc_settle(const Geo &g, const double *cx, double *climit, unsigned long long *dmax, hipStream_t st);
// any real cell with !(h > 0): *flag |= 1
void launch_film_check(const Geo &g, const double *h, int *flag, hipStream_t st);
// the end of the step over every row: c0 = cx, c = cfield = min(cx, 1)
void launch_conc_commit(const Geo &g, const double *cx, double *c0, double *c, double *cfield, hipStream_t st);
// zero out d when it holds a value that is not finite (a failed solve)
void launch_zero_if_nonfinite(int n, double *d, int *flag, hipStream_t st);
// StepSources' evaporation at the film of the pass: evp = (1 - theta) evold
// + theta EvapAt(h, c) over the real cells (N)
void launch_step_evp(const RowArgs &a, const double *h, const double *c, double *evp, hipStream_t st);

// ConsistentFluxes: F1 gains the flux whose divergence is what the step's
// divergence misses of the film's step. scratch: at least N + 2 ny + 1024 + 8
// doubles.
void launch_consistent_fluxes(const ConcArgs &a, double *Fx, double *Fy, double *scratch, hipStream_t st);

// RedistributeExcess, on the device: the flags of c over and under the
// bounds (flags[0], flags[1]), one pass of MoveExcess (room, move, gain),
// the spread of what is left, the final caps. The host drives the passes.
struct ExcessArgs {
    Geo g;
    const double *h;     // the film of the step's end, padded
    double *c;           // the concentration, padded
    double *room, *gain, *f, *e;   // padded scratch
    double *part;        // 256 block partials
    double *sums;        // [0] moved, [1] left, [2] total, [3] cut
    int *flags;          // [0] over, [1] under, [2] any
    int lower;
};
void launch_excess_flags(const ExcessArgs &a, hipStream_t st);
void launch_excess_pass(const ExcessArgs &a, int pass, hipStream_t st);
// what is left after the passes: left and total summed in a fixed order,
// then spread over the cells with room
void launch_excess_spread(const ExcessArgs &a, hipStream_t st);
void launch_excess_caps(const ExcessArgs &a, int under, hipStream_t st);

}  // namespace yz
