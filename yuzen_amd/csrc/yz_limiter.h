// yz_limiter.h -- the limited face interpolation (reference
// yuzen/src/numerics/FiniteDifferences.cpp:94-190): Koren's and van Leer's
// limiters with their derivatives, the limited face value with its
// downwind-form weights and derivative, and the weights frozen at a field
// with the row's cell upwind or downwind. Device code, shared by the film
// height's upwinded terms and both concentration models.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace yz {

// ---- Koren limiter (FiniteDifferences.cpp Limiter/FaceValue/...) -------
__device__ inline double koren(double a, double b, double *pa, double *pb) {
    double p = 0, qa = 0, qb = 0;
    if ((a > 0 && b > 0) || (a < 0 && b < 0)) {
        const double s = (b > 0) ? 1.0 : -1.0;
        const double aa = fabs(a), ab = fabs(b);
        const double t1 = 2.0 * aa, t2 = (2.0 * ab + aa) / 3.0, t3 = 2.0 * ab;
        if (t1 <= t2 && t1 <= t3) { p = s * t1; qa = 2.0; qb = 0.0; }
        else if (t2 <= t3) { p = s * t2; qa = 1.0 / 3.0; qb = 2.0 / 3.0; }
        else { p = s * t3; qa = 0.0; qb = 2.0; }
    }
    *pa = qa; *pb = qb;
    return p;
}
// van Leer: 2ab / (a + b) where a and b share a strict sign, written with
// the ratios a / (a + b) and b / (a + b), both in (0, 1), so that it
// neither underflows nor overflows for tiny or huge differences
// (FiniteDifferences.cpp:112-124); 0 otherwise. pa, pb: its derivatives.
__device__ inline double vanleer(double a, double b, double *pa, double *pb) {
    double p = 0, qa = 0, qb = 0;
    if ((a > 0 && b > 0) || (a < 0 && b < 0)) {
        const double sum = a + b, ra = a / sum, rb = b / sum;
        p = 2.0 * a * rb;
        qa = 2.0 * rb * rb;
        qb = 2.0 * ra * ra;
    }
    *pa = qa; *pb = qb;
    return p;
}
// The limited face value hf, its downwind-form weights w (sum w f = hf)
// and its derivative g = d hf / d f; Koren, or van Leer (vl).
__device__ inline double limited_face(const double f[4], double sign, double w[4], double g[4], int vl = 0) {
    int uu, u, d;
    if (sign >= 0) { uu = 0; u = 1; d = 2; } else { uu = 3; u = 2; d = 1; }
    const double a = f[u] - f[uu], b = f[d] - f[u];
    double pa, pb;
    const double p = vl ? vanleer(a, b, &pa, &pb) : koren(a, b, &pa, &pb);
#pragma unroll
    for (int j = 0; j < 4; ++j) { w[j] = 0; g[j] = 0; }
    const double psi = (b != 0) ? p / b : 0.0;
    w[u] = 1.0 - 0.5 * psi;
    w[d] = 0.5 * psi;
    g[uu] = -0.5 * pa;
    g[u] = 1.0 + 0.5 * (pa - pb);
    g[d] = 0.5 * pb;
    return f[u] + 0.5 * p;
}

// FaceWeights of a limited scheme (FiniteDifferences.cpp:166-190): the face
// value as weights of the four points frozen at f, with the row's own cell
// at rowpos (LL 0, L 1, R 2, RR 3). The row's cell upwind: f_U + chi/2
// (f_U - f_UU), chi = phi / (f_U - f_UU); otherwise f_U + psi/2 (f_D - f_U),
// psi = phi / (f_D - f_U); either keeps the frozen system an M-matrix.
__device__ inline void limited_weights(const double f[4], double sign, int rowpos, int vl, double w[4]) {
    int uu, u, d;
    if (sign >= 0) { uu = 0; u = 1; d = 2; } else { uu = 3; u = 2; d = 1; }
    const double a = f[u] - f[uu], b = f[d] - f[u];
    double pa, pb;
    const double p = vl ? vanleer(a, b, &pa, &pb) : koren(a, b, &pa, &pb);
#pragma unroll
    for (int j = 0; j < 4; ++j) w[j] = 0;
    if (rowpos == u) {
        const double chi = (a != 0) ? p / a : 0.0;
        w[u] = 1.0 + 0.5 * chi;
        w[uu] = -0.5 * chi;
    } else {
        const double psi = (b != 0) ? p / b : 0.0;
        w[u] = 1.0 - 0.5 * psi;
        w[d] = 0.5 * psi;
    }
}

}  // namespace yz
