// yz_schemes.h -- the fixed coefficients of the unlimited four-point face
// schemes (C_Interpolation_base4::SetCoef, FiniteDifferences.cpp:22-73),
// shared by the film height's upwinded terms and the two concentration
// models. Host code.
#pragma once
#include "../../include/yuzen_gpu.h"

namespace yz {

// The fixed coefficients of the unlimited four-point schemes
// (FiniteDifferences.cpp:22-73): c1 and c2 of LL L R RR over the sum of c1,
// for a YZ_CONV_* scheme (UTOPIA for any other value). Host code.
inline void scheme_coefs(int conv, double c1[4], double c2[4]) {
    double a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
    switch (conv) {
    case YZ_CONV_UPWIND1: a[1] = 1; a[2] = 1; b[1] = 1; b[2] = -1; break;
    case YZ_CONV_CENTRAL2: a[1] = 1; a[2] = 1; break;
    case YZ_CONV_CENTRAL4: a[0] = -1; a[1] = 9; a[2] = 9; a[3] = -1; break;
    case YZ_CONV_UPWIND2: a[0] = -1; a[1] = 3; a[2] = 3; a[3] = -1; b[0] = -1; b[1] = 3; b[2] = -3; b[3] = 1; break;
    case YZ_CONV_QUICK: a[0] = -1; a[1] = 9; a[2] = 9; a[3] = -1; b[0] = -1; b[1] = 3; b[2] = -3; b[3] = 1; break;
    case YZ_CONV_KK: a[0] = -1; a[1] = 7; a[2] = 7; a[3] = -1; b[0] = -3; b[1] = 9; b[2] = -9; b[3] = 3; break;
    default: a[0] = -1; a[1] = 7; a[2] = 7; a[3] = -1; b[0] = -1; b[1] = 3; b[2] = -3; b[3] = 1; break;
    }
    double den = 0;
    for (int j = 0; j < 4; ++j) den += a[j];
    for (int j = 0; j < 4; ++j) { c1[j] = a[j] / den; c2[j] = b[j] / den; }
}

}  // namespace yz
