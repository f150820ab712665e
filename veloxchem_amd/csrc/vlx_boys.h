// Boys-function table geometry (host and device).
#pragma once

// T_k = k * BOYS_DT for k < BOYS_NGRID, Taylor order BOYS_ORDER.  With
// |T - T_k| <= BOYS_DT / 2 = 0.125 the truncation error is below
// 0.125^11 / 11! ~ 6e-18 relative.  T >= BOYS_TMAX uses the asymptotic form.
// A table row of the table for L holds the BOYS_ORDER + 1 Taylor coefficients
// of F_L at T_k, then those of F_{L-1} (zeros for L = 0): the kernels evaluate
// both and take e^-T = (2L - 1) F_{L-1} - 2T F_L for the downward recursion.
constexpr int BOYS_ORDER = 10;
constexpr double BOYS_DT = 0.25;
constexpr double BOYS_INV_DT = 4.0;
constexpr int BOYS_NGRID = 145;
constexpr double BOYS_TMAX = 36.0;
constexpr int BOYS_ROW = 2 * (BOYS_ORDER + 1);
constexpr int BOYS_TAB_SIZE = BOYS_NGRID * BOYS_ROW;
constexpr int VLX_MAX_L = 3;
constexpr int VLX_MAX_LTOT = 4 * VLX_MAX_L;

// Fills tab[(L * BOYS_NGRID + k) * BOYS_ROW + j] = F_{L+j}(T_k) (-1)^j / j! and
// tab[(L * BOYS_NGRID + k) * BOYS_ROW + BOYS_ORDER + 1 + j] = F_{L-1+j}(T_k) (-1)^j / j!
// (j <= BOYS_ORDER; the second block is zero for L = 0), for L = 0..VLX_MAX_LTOT.
void vlx_build_boys_table(double* tab);
