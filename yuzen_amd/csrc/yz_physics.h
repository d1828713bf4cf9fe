// yz_physics.h -- per-cell physics of the well-mixed film-height equation,
// device side (gfx950). Everything a cell's row needs is computed here
// from the 13-point (2-D) / 5-point (1-D) stencil of the film h, the
// substrate s and the frozen concentration c, held in registers:
//
//   * the row of A(h) (all six terms) and of the Jacobian J = d(A(h)h)/dh,
//   * A(h)h in the form the reference's ApplyA forms it: the capillary and
//     vertical-gravity terms face by face (differences of differences), the
//     upwinded and disjoining terms as the product of their entries with h.
//
// Formulas follow the reference kernels (paths relative to the reference
// root, yuzen/src/model/well_mixed/):
//   capillary          LaplacePressureKernel.h:242-579,
//                      WM2D/LaplacePressure_2D.cpp:39-128, WM1D/...:35-77
//   vertical gravity   VerticalGravityKernel.h:150-433
//   disjoining         DisjoiningPressureKernel.h:150-435
//   body forces        BodyForceKernel.h (LimitedFace, *_Limited; the
//                      fixed-coefficient forms :263-545), ParallelGravity-
//                      Kernel.h, CentrifugalKernel.h
//   Marangoni          SolutoCapillarityKernel.h (limited and fixed forms)
//   mobility           ../Mobility.h:49-191
//   limiter            ../../numerics/FiniteDifferences.cpp:94-228
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

#include "../../include/yuzen_gpu.h"
#include "yz_limiter.h"
#include "yz_schemes.h"

namespace yz {

// Stencil slots. 2-D: oM mm om pm Mo mo oo po Po mp op pp oP (the
// reference's FullIdx order); 1-D: M m o p P.
enum { S_oM, S_mm, S_om, S_pm, S_Mo, S_mo, S_oo, S_po, S_Po, S_mp, S_op, S_pp, S_oP };

struct Geo {
    int dim, nx, ny;        // interior cells
    int nxp, nyp;           // padded extents (nyp = 1 in 1-D)
    int N;                  // interior unknowns nx*ny
    int ns, center, ndir;   // stencil size, centre slot, faces
    double dx, dy, xlo, ylo;
    int bc[4];
};

__host__ __device__ inline int slot_ox(int dim, int k) {
    if (dim == 1) return k - 2;
    const int t[13] = {0, -1, 0, 1, -2, -1, 0, 1, 2, -1, 0, 1, 0};
    return t[k];
}
__host__ __device__ inline int slot_oy(int dim, int k) {
    if (dim == 1) return 0;
    const int t[13] = {-2, -1, -1, -1, 0, 0, 0, 0, 0, 1, 1, 1, 2};
    return t[k];
}
// AdjIdx(d): the neighbour across face d (Xm, Xp, Ym, Yp)
__host__ __device__ inline int adj_slot(int dim, int d) {
    if (dim == 1) return d == 0 ? 1 : 3;
    const int t[4] = {S_mo, S_po, S_om, S_op};
    return t[d];
}
// FaceStencil(d, j): LL L R RR of face d
__host__ __device__ inline int face_slot(int dim, int d, int j) {
    if (dim == 1) return d == 0 ? j : j + 1;
    const int t[4][4] = {{S_Mo, S_mo, S_oo, S_po}, {S_mo, S_oo, S_po, S_Po},
                         {S_oM, S_om, S_oo, S_op}, {S_om, S_oo, S_op, S_oP}};
    return t[d][j];
}
// AxisStencil(a, q)
__host__ __device__ inline int axis_slot(int dim, int a, int q) {
    if (dim == 1) return q;
    const int t[2][5] = {{S_Mo, S_mo, S_oo, S_po, S_Po}, {S_oM, S_om, S_oo, S_op, S_oP}};
    return t[a][q];
}
// StencilDir(k): Ym Xm Xp Yp in 2-D
__host__ __device__ inline int sdir(int ndir, int k) {
    if (ndir == 2) return k;
    const int t[4] = {2, 0, 1, 3};
    return t[k];
}

__device__ inline double sgn(double v) { return v >= 0.0 ? 1.0 : -1.0; }

// ---- mobility (Mobility.h) -------------------------------------------
struct Mob {
    double lambda, a;
    bool slip;
    __device__ double D(double h) const { const double u = fabs(h); return h * u * (u + a); }
    __device__ double dD(double h) const { const double u = fabs(h); return u * (3.0 * u + 2.0 * a); }
    __device__ double DperH(double h) const { const double u = fabs(h); return u * (u + a); }
    __device__ double Gint(double h) const {
        const double x = a / h;
        double f;
        if (x < 0.1) {
            double t = x * x, s = 0;
            for (int k = 2; k <= 18; ++k) { s += ((k % 2 == 0) ? t : -t) / k; t *= x; }
            f = s;
        } else f = x - log1p(x);
        return -f / (a * a);
    }
    __device__ double K(double h) const { const double u = dD(h) / D(h); return 2.0 * u * u - (6.0 * h + 2.0 * a) / D(h); }
    __device__ double dK(double h) const {
        const double d = D(h), d1 = dD(h), d2 = 6.0 * h + 2.0 * a, d3 = 6.0;
        const double u = d1 / d, du = d2 / d - u * u;
        return 4.0 * u * du - (d3 * d - d2 * d1) / (d * d);
    }
    __device__ double FaceDK(double h1, double h2) const {
        if (!(h1 > 0) || !(h2 > 0)) return 0;
        const double hm = 0.5 * (h1 + h2), dl = h2 - h1;
        if (fabs(dl) < 1e-3 * hm) return D(hm) / (1.0 + dl * dl * K(hm) / 24.0);
        return dl / (Gint(h2) - Gint(h1));
    }
    __device__ double FaceDK_d1(double h1, double h2) const {
        if (!(h1 > 0) || !(h2 > 0)) return 0;
        const double hm = 0.5 * (h1 + h2), dl = h2 - h1;
        if (fabs(dl) < 1e-3 * hm) {
            const double k = K(hm), q = 1.0 + dl * dl * k / 24.0;
            const double Fhm = (dD(hm) * q - D(hm) * dl * dl * dK(hm) / 24.0) / (q * q);
            const double Fdl = -D(hm) * (dl * k / 12.0) / (q * q);
            return 0.5 * Fhm - Fdl;
        }
        const double dg = Gint(h2) - Gint(h1);
        return (-dg + dl / D(h1)) / (dg * dg);
    }
    // FaceDK(h1, h2) with both derivatives, dF/dh1 and dF/dh2, from one
    // evaluation of g at each height (the same expressions as FaceDK,
    // FaceDK_d1(h1, h2) and FaceDK_d1(h2, h1)). CACHED: g(h1), g(h2) come
    // from the caller (G1, G2, each Gint of its height) instead of being
    // formed here; the values are the same. Swapping h1 and h2 gives the
    // same F and swaps F1 and F2, to the bit (hm and dl * dl do not change,
    // dl and the differences of g change sign exactly).
    template <bool CACHED>
    __device__ void FaceDK3T(double h1, double h2, double G1, double G2, double &F, double &F1, double &F2) const {
        if (!(h1 > 0) || !(h2 > 0)) { F = F1 = F2 = 0; return; }
        const double hm = 0.5 * (h1 + h2), dl = h2 - h1;
        if (fabs(dl) < 1e-3 * hm) {
            const double Dm = D(hm), dDm = dD(hm), k = K(hm), dk = dK(hm);
            const double q = 1.0 + dl * dl * k / 24.0;
            F = Dm / q;
            const double Fhm = (dDm * q - Dm * dl * dl * dk / 24.0) / (q * q);
            const double Fdl = -Dm * (dl * k / 12.0) / (q * q);
            F1 = 0.5 * Fhm - Fdl;
            // the same at (h2, h1): hm unchanged, dl -> -dl
            const double Fdl2 = -Dm * ((-dl) * k / 12.0) / (q * q);
            F2 = 0.5 * Fhm - Fdl2;
            return;
        }
        const double g1 = CACHED ? G1 : Gint(h1), g2 = CACHED ? G2 : Gint(h2);
        const double dg = g2 - g1;
        F = dl / dg;
        F1 = (-dg + dl / D(h1)) / (dg * dg);
        const double dg2 = g1 - g2, dl2 = h1 - h2;
        F2 = (-dg2 + dl2 / D(h2)) / (dg2 * dg2);
    }
    __device__ void FaceDK3(double h1, double h2, double &F, double &F1, double &F2) const {
        FaceDK3T<false>(h1, h2, 0.0, 0.0, F, F1, F2);
    }
    // Mobility::Face / Face_d1 (0 when a side is dry)
    __device__ double Face(double h1, double h2) const {
        if (slip) return FaceDK(h1, h2);
        if (!(h1 > 0) || !(h2 > 0)) return 0;
        return 2.0 * (h1 * h1) * (h2 * h2) / (h1 + h2);
    }
    __device__ double Face_d1(double h1, double h2) const {
        if (slip) return FaceDK_d1(h1, h2);
        if (!(h1 > 0) || !(h2 > 0)) return 0;
        return 2.0 * (h1 * h2 * h2) * (h1 + h2 + h2) / ((h1 + h2) * (h1 + h2));
    }
};

// The face scheme of the upwinded terms' film height (Numerics.Convection.
// Height): kind 0 Koren, 1 van Leer (limited), 2 a fixed-coefficient scheme
// whose coefficients of LL L R RR are c1 + c2 with the flow in +, c1 - c2
// against it (FiniteDifferences.cpp:152-170).
struct HScheme {
    int kind;
    double c1[4], c2[4];
};
// The face value of f by the scheme, with A's weights w (sum w f = the face
// value) and its derivative g by f: for a fixed scheme both are the
// coefficients c1 + Sign(sign) c2.
__device__ inline double face_interp(const double f[4], double sign, double w[4], double g[4], const HScheme &S) {
    if (S.kind < 2) return limited_face(f, sign, w, g, S.kind);
    const double sg = sgn(sign);
    double v = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        w[j] = S.c1[j] + sg * S.c2[j];
        g[j] = w[j];
        v += f[j] * w[j];
    }
    return v;
}

// The height scheme of the case (YZ_HSCHEME_*) as the device applies it.
// Host code.
inline HScheme height_scheme_of(int hs) {
    HScheme S{};
    if (hs == YZ_HSCHEME_KOREN || hs == YZ_HSCHEME_VANLEER) {
        S.kind = hs == YZ_HSCHEME_VANLEER ? 1 : 0;
        return S;
    }
    S.kind = 2;
    int conv = YZ_CONV_UTOPIA;
    switch (hs) {
    case YZ_HSCHEME_UPWIND1: conv = YZ_CONV_UPWIND1; break;
    case YZ_HSCHEME_CENTRAL2: conv = YZ_CONV_CENTRAL2; break;
    case YZ_HSCHEME_CENTRAL4: conv = YZ_CONV_CENTRAL4; break;
    case YZ_HSCHEME_UPWIND2: conv = YZ_CONV_UPWIND2; break;
    case YZ_HSCHEME_QUICK: conv = YZ_CONV_QUICK; break;
    case YZ_HSCHEME_KK: conv = YZ_CONV_KK; break;
    default: break;
    }
    scheme_coefs(conv, S.c1, S.c2);
    return S;
}

}  // namespace yz
