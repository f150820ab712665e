// Host generation of the Boys-function Taylor table used by the kernels.
//
// F_n(T) = integral_0^1 t^(2n) exp(-T t^2) dt is evaluated in long double
// (64-bit mantissa) from the convergent series
//   F_n(T) = exp(-T) sum_{k>=0} (2T)^k / ((2n+1)(2n+3)...(2n+2k+1)),
// whose terms are all positive (no cancellation), so each table entry is
// accurate to well below double rounding.  The reference tabulates the same
// function (src/math/BoysFunc.hpp:24806-24888); this table is generated here
// from the definition, with its own grid spacing and order.
#include <cmath>

#include "vlx_boys.h"

static long double boys_series(int n, long double t)
{
    long double term = 1.0L / (2 * n + 1);
    long double sum = term;
    for (int k = 1; k < 2000; k++) {
        term *= 2.0L * t / (2 * n + 2 * k + 1);
        sum += term;
        if (term < 1.0e-22L * sum) break;
    }
    return expl(-t) * sum;
}

void vlx_build_boys_table(double* tab)
{
    const int nmax = VLX_MAX_LTOT + BOYS_ORDER;
    static long double f[BOYS_NGRID][VLX_MAX_LTOT + BOYS_ORDER + 1];
    for (int k = 0; k < BOYS_NGRID; k++)
        for (int n = 0; n <= nmax; n++) f[k][n] = boys_series(n, (long double)k * BOYS_DT);
    for (int L = 0; L <= VLX_MAX_LTOT; L++)
        for (int k = 0; k < BOYS_NGRID; k++) {
            long double fact = 1.0L;
            double* row = tab + (size_t)(L * BOYS_NGRID + k) * BOYS_ROW;
            for (int j = 0; j <= BOYS_ORDER; j++) {
                if (j > 0) fact *= j;
                const long double sgn = (j & 1) ? -1.0L : 1.0L;
                row[j] = (double)(sgn * f[k][L + j] / fact);
                row[BOYS_ORDER + 1 + j] = (L > 0) ? (double)(sgn * f[k][L - 1 + j] / fact) : 0.0;
            }
        }
}
