/*
 * CPU FP64 oracle for the two-electron Fock build -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker.  The product path (the HIP
 * library under veloxchem_amd/) never calls into it.
 *
 * It restates VeloxChem's CPU Fock build semantics with an independent
 * integral algorithm:
 *   - electron-repulsion integrals by McMurchie-Davidson (Hermite expansion
 *     E coefficients + R_tuv auxiliary integrals), not the Obara-Saika/HGP
 *     scheme the GPU kernels use, so one cannot hide the other's mistakes;
 *   - Boys function by series + downward recursion (small T) or erf +
 *     upward recursion (large T), no tabulation;
 *   - Cartesian component order and cart->sph factors as
 *     src/general/SphericalMomentum.hpp:18-134;
 *   - Schwarz screening as src/orbdata/BlockedGtoPairBlock.cpp:32-42 (bin =
 *     -floor(log10(sqrt(max))), dropped when sqrt <= 1e-15), the maximum taken
 *     over all spherical components of (ab|ab) starting from zero
 *     (src/t4c_func/T4CUtils.cpp:513-519) and quartets kept iff
 *     bin_bra + bin_ket <= ithreshold (src/general/OpenMPFunc.cpp:171,234);
 *   - the erf-attenuated operator erf(omega r)/r of the range-separated labels
 *     (j_rs, k_rs, kx_rs: src/t4c_func/T4CMatrixDistributor.cpp:51-54) by the
 *     substitution F_m(T) -> s^(m+1/2) F_m(s T), s = omega^2 / (omega^2 + rho)
 *     (src/t4c_func/T4CUtils.cpp:422-465, src/math/BoysFunc.hpp:24890-24940),
 *     which in the Hermite form is rho -> s rho plus a factor sqrt(s);
 *   - J/K digestion per AO quartet with the component-symmetry skips and the
 *     1/2 factors of src/t4c_func/T4CLocalDistributor.cpp:130-280 (2jk),
 *     :293-560 (2jkx), :568-766 (j), :769-1005 (k), :1013-1245 (kx), then
 *     G = F' + F'^T as accumulate + symmetrize do
 *     (src/t4c_func/T4CMatrixDistributor.cpp:140-320, src/math/SubMatrix.cpp:167-178).
 *
 * Parity pins: tests/test_oracle.py checks this file against the reference's
 * own fixtures (STO-3G water ERI tensor, J/K of the def2-SVPD and def2-TZVP
 * water dimers), the reference screener's known (ab|ab) maxima, and against
 * closed forms (Boys F_m(0) = 1/(2m+1), 1F1 representation, (ss|ss)).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OMAXL 3
#define OMAXL4 (4 * OMAXL)

typedef struct {
    int nshells;
    const int *l;
    const int *nprim;
    const int *prim_offset;
    const double *exps;
    const double *norms;
    const double *centers; /* nshells x 3, bohr */
    const int *ao_base;
    const int *ao_stride;
    int nao;
} oracle_basis_t;

enum { LBL_2JK = 0, LBL_2JKX = 1, LBL_J = 2, LBL_K = 3, LBL_KX = 4 };

/* ---------------------------------------------------------------- Boys */

void oracle_boys(int mmax, double t, double *f)
{
    if (t < 40.0) {
        /* F_m(T) = e^{-T} sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)) */
        const double et = exp(-t);
        double term = 1.0 / (2 * mmax + 1);
        double sum = term;
        for (int k = 1; k < 400; k++) {
            term *= 2.0 * t / (2 * mmax + 2 * k + 1);
            sum += term;
            if (term < 1.0e-18 * sum) break;
        }
        f[mmax] = et * sum;
        for (int m = mmax - 1; m >= 0; m--) f[m] = (2.0 * t * f[m + 1] + et) / (2 * m + 1);
    } else {
        const double et = exp(-t);
        f[0] = 0.5 * sqrt(M_PI / t) * erf(sqrt(t));
        for (int m = 0; m < mmax; m++) f[m + 1] = ((2 * m + 1) * f[m] - et) / (2.0 * t);
    }
}

/* ------------------------------------------------ Cartesian components */

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }
static int nsph(int l) { return 2 * l + 1; }

/* component c of shell l -> exponents (i,j,k); order xx,xy,xz,yy,yz,zz,... */
static void cart_exps(int l, int c, int *ex)
{
    int n = 0;
    for (int i = l; i >= 0; i--)
        for (int j = l - i; j >= 0; j--) {
            if (n == c) {
                ex[0] = i;
                ex[1] = j;
                ex[2] = l - i - j;
                return;
            }
            n++;
        }
}

/* cart->sph factors, spherical component m as a sum over Cartesian ones.
 * Values restate src/general/SphericalMomentum.hpp:18-134. */
static int sph_terms(int l, int m, int *cidx, double *coef)
{
    if (l == 0) { cidx[0] = 0; coef[0] = 1.0; return 1; }
    if (l == 1) {
        static const int map[3] = {1, 2, 0};
        cidx[0] = map[m]; coef[0] = 1.0; return 1;
    }
    if (l == 2) {
        const double f3 = 2.0 * sqrt(3.0);
        switch (m) {
        case 0: cidx[0] = 1; coef[0] = f3; return 1;
        case 1: cidx[0] = 4; coef[0] = f3; return 1;
        case 2: cidx[0] = 0; coef[0] = -1.0; cidx[1] = 3; coef[1] = -1.0; cidx[2] = 5; coef[2] = 2.0; return 3;
        case 3: cidx[0] = 2; coef[0] = f3; return 1;
        default: cidx[0] = 0; coef[0] = 0.5 * f3; cidx[1] = 3; coef[1] = -0.5 * f3; return 2;
        }
    }
    /* l == 3 */
    const double f5 = sqrt(2.5), f15 = 2.0 * sqrt(15.0), f3 = sqrt(1.5);
    switch (m) {
    case 0: cidx[0] = 1; coef[0] = 3.0 * f5; cidx[1] = 6; coef[1] = -f5; return 2;
    case 1: cidx[0] = 4; coef[0] = f15; return 1;
    case 2: cidx[0] = 8; coef[0] = 4.0 * f3; cidx[1] = 1; coef[1] = -f3; cidx[2] = 6; coef[2] = -f3; return 3;
    case 3: cidx[0] = 9; coef[0] = 2.0; cidx[1] = 2; coef[1] = -3.0; cidx[2] = 7; coef[2] = -3.0; return 3;
    case 4: cidx[0] = 5; coef[0] = 4.0 * f3; cidx[1] = 0; coef[1] = -f3; cidx[2] = 3; coef[2] = -f3; return 3;
    case 5: cidx[0] = 2; coef[0] = 0.5 * f15; cidx[1] = 7; coef[1] = -0.5 * f15; return 2;
    default: cidx[0] = 0; coef[0] = f5; cidx[1] = 3; coef[1] = -3.0 * f5; return 2;
    }
}

/* ---------------------------------------- McMurchie-Davidson integrals */

/* 1D Hermite expansion coefficients E[i][j][t], i<=la, j<=lb, t<=i+j,
 * without the exp(-mu X^2) factor (applied once in 3D). */
#define EDIM (OMAXL + 1)
#define TDIM (2 * OMAXL + 1)
static void hermite_e(int la, int lb, double p, double xpa, double xpb, double e[EDIM][EDIM][TDIM])
{
    memset(e, 0, sizeof(double) * EDIM * EDIM * TDIM);
    const double h = 0.5 / p;
    e[0][0][0] = 1.0;
    for (int i = 0; i <= la; i++) {
        for (int j = 0; j <= lb; j++) {
            if (i == 0 && j == 0) continue;
            for (int t = 0; t <= i + j; t++) {
                double v;
                if (j == 0) { /* raise i from (i-1, j) */
                    v = xpa * e[i - 1][j][t];
                    if (t > 0) v += h * e[i - 1][j][t - 1];
                    if (t + 1 <= i - 1 + j) v += (t + 1) * e[i - 1][j][t + 1];
                } else { /* raise j from (i, j-1) */
                    v = xpb * e[i][j - 1][t];
                    if (t > 0) v += h * e[i][j - 1][t - 1];
                    if (t + 1 <= i + j - 1) v += (t + 1) * e[i][j - 1][t + 1];
                }
                e[i][j][t] = v;
            }
        }
    }
}

#define RDIM (OMAXL4 + 1)
/* R[t][u][v] (level n = 0) for t+u+v <= L */
static void hermite_r(int L, double alpha, const double *pq, double *r0 /* RDIM^3 */)
{
    static __thread double rn[RDIM + 1][RDIM][RDIM][RDIM];
    double f[RDIM + 1];
    const double t2 = alpha * (pq[0] * pq[0] + pq[1] * pq[1] + pq[2] * pq[2]);
    oracle_boys(L, t2, f);
    double fac = 1.0;
    for (int n = 0; n <= L; n++) {
        rn[n][0][0][0] = fac * f[n];
        fac *= -2.0 * alpha;
    }
    for (int n = L - 1; n >= 0; n--) {
        const int top = L - n;
        for (int s = 1; s <= top; s++)
            for (int t = 0; t <= s; t++)
                for (int u = 0; u <= s - t; u++) {
                    const int v = s - t - u;
                    double val;
                    if (t > 0) {
                        val = pq[0] * rn[n + 1][t - 1][u][v];
                        if (t > 1) val += (t - 1) * rn[n + 1][t - 2][u][v];
                    } else if (u > 0) {
                        val = pq[1] * rn[n + 1][t][u - 1][v];
                        if (u > 1) val += (u - 1) * rn[n + 1][t][u - 2][v];
                    } else {
                        val = pq[2] * rn[n + 1][t][u][v - 1];
                        if (v > 1) val += (v - 1) * rn[n + 1][t][u][v - 2];
                    }
                    rn[n][t][u][v] = val;
                }
    }
    for (int t = 0; t <= L; t++)
        for (int u = 0; u <= L - t; u++)
            for (int v = 0; v <= L - t - u; v++) r0[(t * RDIM + u) * RDIM + v] = rn[0][t][u][v];
}

/* Cartesian (ab|cd) block, index ((a*nb + b)*nc + c)*nd + d */
static void eri_cart(const oracle_basis_t *bs, int sa, int sb, int sc, int sd, double omega, double *out)
{
    const int la = bs->l[sa], lb = bs->l[sb], lc = bs->l[sc], ld = bs->l[sd];
    const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
    const int L = la + lb + lc + ld;
    const double *A = bs->centers + 3 * sa, *B = bs->centers + 3 * sb;
    const double *C = bs->centers + 3 * sc, *D = bs->centers + 3 * sd;
    double *r0 = malloc(sizeof(double) * RDIM * RDIM * RDIM);
    int ea[15][3], eb[15][3], ec[15][3], ed[15][3];
    for (int i = 0; i < na; i++) cart_exps(la, i, ea[i]);
    for (int i = 0; i < nb; i++) cart_exps(lb, i, eb[i]);
    for (int i = 0; i < nc; i++) cart_exps(lc, i, ec[i]);
    for (int i = 0; i < nd; i++) cart_exps(ld, i, ed[i]);
    memset(out, 0, sizeof(double) * na * nb * nc * nd);
    double eab[3][EDIM][EDIM][TDIM], ecd[3][EDIM][EDIM][TDIM];
    const double ab2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]);
    const double cd2 = (C[0] - D[0]) * (C[0] - D[0]) + (C[1] - D[1]) * (C[1] - D[1]) + (C[2] - D[2]) * (C[2] - D[2]);
    for (int ia = 0; ia < bs->nprim[sa]; ia++)
        for (int ib = 0; ib < bs->nprim[sb]; ib++) {
            const double a = bs->exps[bs->prim_offset[sa] + ia], b = bs->exps[bs->prim_offset[sb] + ib];
            const double p = a + b;
            double P[3];
            for (int x = 0; x < 3; x++) P[x] = (a * A[x] + b * B[x]) / p;
            for (int x = 0; x < 3; x++) hermite_e(la, lb, p, P[x] - A[x], P[x] - B[x], eab[x]);
            const double kab = exp(-a * b / p * ab2) * bs->norms[bs->prim_offset[sa] + ia] * bs->norms[bs->prim_offset[sb] + ib];
            for (int ic = 0; ic < bs->nprim[sc]; ic++)
                for (int id = 0; id < bs->nprim[sd]; id++) {
                    const double c = bs->exps[bs->prim_offset[sc] + ic], d = bs->exps[bs->prim_offset[sd] + id];
                    const double q = c + d;
                    double Q[3], pq[3];
                    for (int x = 0; x < 3; x++) Q[x] = (c * C[x] + d * D[x]) / q;
                    for (int x = 0; x < 3; x++) hermite_e(lc, ld, q, Q[x] - C[x], Q[x] - D[x], ecd[x]);
                    const double kcd = exp(-c * d / q * cd2) * bs->norms[bs->prim_offset[sc] + ic] * bs->norms[bs->prim_offset[sd] + id];
                    for (int x = 0; x < 3; x++) pq[x] = P[x] - Q[x];
                    const double alpha = p * q / (p + q);
                    const double s = (omega != 0.0) ? omega * omega / (omega * omega + alpha) : 1.0;
                    hermite_r(L, alpha * s, pq, r0);
                    const double pref = 2.0 * pow(M_PI, 2.5) / (p * q * sqrt(p + q)) * kab * kcd * sqrt(s);
                    for (int i = 0; i < na; i++)
                        for (int j = 0; j < nb; j++) {
                            const int *xa = ea[i], *xb = eb[j];
                            for (int k = 0; k < nc; k++)
                                for (int l = 0; l < nd; l++) {
                                    const int *xc = ec[k], *xd = ed[l];
                                    double sum = 0.0;
                                    for (int t = 0; t <= xa[0] + xb[0]; t++)
                                        for (int u = 0; u <= xa[1] + xb[1]; u++)
                                            for (int v = 0; v <= xa[2] + xb[2]; v++) {
                                                const double e1 = eab[0][xa[0]][xb[0]][t] * eab[1][xa[1]][xb[1]][u] * eab[2][xa[2]][xb[2]][v];
                                                if (e1 == 0.0) continue;
                                                double inner = 0.0;
                                                for (int tt = 0; tt <= xc[0] + xd[0]; tt++)
                                                    for (int uu = 0; uu <= xc[1] + xd[1]; uu++)
                                                        for (int vv = 0; vv <= xc[2] + xd[2]; vv++) {
                                                            const double e2 = ecd[0][xc[0]][xd[0]][tt] * ecd[1][xc[1]][xd[1]][uu] * ecd[2][xc[2]][xd[2]][vv];
                                                            const double sgn = ((tt + uu + vv) & 1) ? -1.0 : 1.0;
                                                            inner += sgn * e2 * r0[((t + tt) * RDIM + (u + uu)) * RDIM + (v + vv)];
                                                        }
                                                sum += e1 * inner;
                                            }
                                    out[((i * nb + j) * nc + k) * nd + l] += pref * sum;
                                }
                        }
                }
        }
    free(r0);
}

/* spherical (ab|cd) block, index ((a*nb + b)*nc + c)*nd + d over 2l+1 comps */
static int eri_sph_omega(const oracle_basis_t *bs, int sa, int sb, int sc, int sd, double omega, double *out);

int oracle_eri_sph(const oracle_basis_t *bs, int sa, int sb, int sc, int sd, double *out)
{
    return eri_sph_omega(bs, sa, sb, sc, sd, 0.0, out);
}

/* erf-attenuated (ab|erf(omega r)/r|cd) block, same layout */
int oracle_eri_sph_erf(const oracle_basis_t *bs, int sa, int sb, int sc, int sd, double omega, double *out)
{
    return eri_sph_omega(bs, sa, sb, sc, sd, omega, out);
}

static int eri_sph_omega(const oracle_basis_t *bs, int sa, int sb, int sc, int sd, double omega, double *out)
{
    const int ls[4] = {bs->l[sa], bs->l[sb], bs->l[sc], bs->l[sd]};
    for (int i = 0; i < 4; i++)
        if (ls[i] < 0 || ls[i] > OMAXL) return -1;
    int nc_[4], ns_[4];
    for (int i = 0; i < 4; i++) { nc_[i] = ncart(ls[i]); ns_[i] = nsph(ls[i]); }
    const int ncart4 = nc_[0] * nc_[1] * nc_[2] * nc_[3];
    double *cur = malloc(sizeof(double) * ncart4);
    eri_cart(bs, sa, sb, sc, sd, omega, cur);
    /* transform one index at a time; dims[] tracks the current shape */
    int dims[4] = {nc_[0], nc_[1], nc_[2], nc_[3]};
    for (int ax = 0; ax < 4; ax++) {
        int ndims[4] = {dims[0], dims[1], dims[2], dims[3]};
        ndims[ax] = ns_[ax];
        const int ntot = ndims[0] * ndims[1] * ndims[2] * ndims[3];
        double *nxt = calloc(ntot, sizeof(double));
        int idx[4];
        for (idx[0] = 0; idx[0] < ndims[0]; idx[0]++)
            for (idx[1] = 0; idx[1] < ndims[1]; idx[1]++)
                for (idx[2] = 0; idx[2] < ndims[2]; idx[2]++)
                    for (idx[3] = 0; idx[3] < ndims[3]; idx[3]++) {
                        int cidx[6];
                        double coef[6];
                        const int nt = sph_terms(ls[ax], idx[ax], cidx, coef);
                        double v = 0.0;
                        for (int t = 0; t < nt; t++) {
                            int src[4] = {idx[0], idx[1], idx[2], idx[3]};
                            src[ax] = cidx[t];
                            v += coef[t] * cur[((src[0] * dims[1] + src[1]) * dims[2] + src[2]) * dims[3] + src[3]];
                        }
                        nxt[((idx[0] * ndims[1] + idx[1]) * ndims[2] + idx[2]) * ndims[3] + idx[3]] = v;
                    }
        free(cur);
        cur = nxt;
        for (int i = 0; i < 4; i++) dims[i] = ndims[i];
    }
    memcpy(out, cur, sizeof(double) * ns_[0] * ns_[1] * ns_[2] * ns_[3]);
    free(cur);
    return 0;
}

/* max over all spherical components of (ab|ab), starting from zero */
int oracle_pair_max(const oracle_basis_t *bs, int sa, int sb, double *out)
{
    const int n = nsph(bs->l[sa]) * nsph(bs->l[sb]);
    double *buf = malloc(sizeof(double) * n * n);
    int rc = oracle_eri_sph(bs, sa, sb, sa, sb, buf);
    double mx = 0.0;
    for (int i = 0; i < n * n; i++)
        if (mx < buf[i]) mx = buf[i];
    free(buf);
    *out = mx;
    return rc;
}

static int pair_bin(double maxval)
{
    const double fact = sqrt(maxval);
    if (fact > 1.0e-15) return (fact < 1.0) ? (int)(-floor(log10(fact))) : 0;
    return -1;
}

/* Full AO ERI tensor (nao^4), no screening */
int oracle_eri_tensor(const oracle_basis_t *bs, double *out)
{
    const int n = bs->nao;
    for (int sa = 0; sa < bs->nshells; sa++)
        for (int sb = 0; sb < bs->nshells; sb++)
            for (int sc = 0; sc < bs->nshells; sc++)
                for (int sd = 0; sd < bs->nshells; sd++) {
                    const int na = nsph(bs->l[sa]), nb = nsph(bs->l[sb]), nc = nsph(bs->l[sc]), nd = nsph(bs->l[sd]);
                    double *buf = malloc(sizeof(double) * na * nb * nc * nd);
                    if (oracle_eri_sph(bs, sa, sb, sc, sd, buf)) { free(buf); return -1; }
                    for (int i = 0; i < na; i++)
                        for (int j = 0; j < nb; j++)
                            for (int k = 0; k < nc; k++)
                                for (int l = 0; l < nd; l++) {
                                    const size_t p = bs->ao_base[sa] + (size_t)i * bs->ao_stride[sa];
                                    const size_t q = bs->ao_base[sb] + (size_t)j * bs->ao_stride[sb];
                                    const size_t r = bs->ao_base[sc] + (size_t)k * bs->ao_stride[sc];
                                    const size_t s = bs->ao_base[sd] + (size_t)l * bs->ao_stride[sd];
                                    out[((p * n + q) * n + r) * n + s] = buf[((i * nb + j) * nc + k) * nd + l];
                                }
                    free(buf);
                }
    return 0;
}

/* per-pair maxima and bins for all unordered shell pairs (sa <= sb) */
int oracle_pair_bins(const oracle_basis_t *bs, double *maxvals, int *bins)
{
    const int ns = bs->nshells;
    int rc = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : rc)
    for (int sa = 0; sa < ns; sa++)
        for (int sb = sa; sb < ns; sb++) {
            const int idx = sa * ns + sb;
            rc |= oracle_pair_max(bs, sa, sb, &maxvals[idx]);
            bins[idx] = pair_bin(maxvals[idx]);
        }
    return rc;
}

/* Digest one shell quartet into F' following the AO-level skips/factors of
 * src/t4c_func/T4CLocalDistributor.cpp:130-280 and siblings. */
static void digest(const oracle_basis_t *bs, int sa, int sb, int sc, int sd, const double *buf, const double *dm,
                   int label, double xf, double *fp)
{
    const int n = bs->nao;
    const int na = nsph(bs->l[sa]), nb = nsph(bs->l[sb]), nc = nsph(bs->l[sc]), nd = nsph(bs->l[sd]);
    const int refp = bs->ao_base[sa], refq = bs->ao_base[sb], refr = bs->ao_base[sc], refs = bs->ao_base[sd];
    const double fj = (label == LBL_J) ? 2.0 : 4.0;
    double fk = 0.0;
    if (label == LBL_2JK) fk = -1.0;
    else if (label == LBL_2JKX) fk = -xf;
    else if (label == LBL_K) fk = 1.0;
    else if (label == LBL_KX) fk = xf;
    const int do_j = (label == LBL_2JK || label == LBL_2JKX || label == LBL_J);
    const int do_k = (label != LBL_J);
    for (int i = 0; i < na; i++) {
        const int p = i * bs->ao_stride[sa] + refp;
        for (int j = 0; j < nb; j++) {
            if (refp == refq && j < i) continue;
            const int q = j * bs->ao_stride[sb] + refq;
            for (int k = 0; k < nc; k++)
                for (int l = 0; l < nd; l++) {
                    if (refr == refs && l < k) continue;
                    if (refp == refr && refq == refs) {
                        if (k < i) continue;
                        if (i == k && l < j) continue;
                    }
                    const int r = k * bs->ao_stride[sc] + refr;
                    const int s = l * bs->ao_stride[sd] + refs;
                    double fval = buf[((i * nb + j) * nc + k) * nd + l];
                    if (p == q) fval *= 0.5;
                    if (r == s) fval *= 0.5;
                    if (p == r && q == s) fval *= 0.5;
                    if (do_j) {
                        fp[(size_t)p * n + q] += fj * fval * dm[(size_t)r * n + s];
                        fp[(size_t)r * n + s] += fj * fval * dm[(size_t)p * n + q];
                    }
                    if (do_k) {
                        fp[(size_t)p * n + r] += fk * fval * dm[(size_t)q * n + s];
                        fp[(size_t)p * n + s] += fk * fval * dm[(size_t)q * n + r];
                        fp[(size_t)q * n + r] += fk * fval * dm[(size_t)p * n + s];
                        fp[(size_t)q * n + s] += fk * fval * dm[(size_t)p * n + r];
                    }
                }
        }
    }
}

/* internal: build with an optional bra-pair stride (1 = everything); antisym: G = F' - F'^T */
static long fock_impl(const oracle_basis_t *bs, const double *dm, int label, double xf, double omega, int ithreshold,
                      int nthreads, int stride, const int *bins_in, double *gout, long *total, int antisym);

/*
 * Bounded CPU-baseline sample: only bra pairs ip with ip % stride == 0 are
 * computed.  Returns the number of shell quartets computed; *total receives
 * the number a full build computes.  gout holds the partial result.
 */
long oracle_fock_sampled(const oracle_basis_t *bs, const double *dm, int label, double xf, int ithreshold,
                         int nthreads, int stride, const int *bins, double *gout, long *total)
{
    return fock_impl(bs, dm, label, xf, 0.0, ithreshold, nthreads, stride, bins, gout, total, 0);
}

/*
 * G = F' + F'^T for a symmetric density dm (nao x nao, reference AO order).
 * ithreshold < 0 computes every quartet (the reference's unscreened
 * _compute_fock_omp(basis, molecule, ...) path); otherwise the quartet set is
 * the reference's binned Schwarz selection.  Returns the number of shell
 * quartets computed (>= 0) or -1 on bad input.
 */
long oracle_fock(const oracle_basis_t *bs, const double *dm, int label, double xf, int ithreshold, int nthreads,
                 double *gout)
{
    long total = 0;
    return fock_impl(bs, dm, label, xf, 0.0, ithreshold, nthreads, 1, NULL, gout, &total, 0);
}

/*
 * Range-separated labels: label j/k/kx digestion of the erf(omega r)/r
 * integrals (the reference's j_rs, k_rs, kx_rs).  Screening uses the
 * Coulomb pair maxima, as the reference's screener does.
 */
long oracle_fock_erf(const oracle_basis_t *bs, const double *dm, int label, double xf, double omega, int ithreshold,
                     int nthreads, double *gout)
{
    long total = 0;
    if (label != LBL_J && label != LBL_K && label != LBL_KX) return -1;
    return fock_impl(bs, dm, label, xf, omega, ithreshold, nthreads, 1, NULL, gout, &total, 0);
}

/*
 * General (non-symmetric) density, as the reference digests mat_t.general
 * densities (its tests: tests/test_fock_driver.py:320-360, K compared with
 * the transposed k.gen fixture).  With S = (D + D^T)/2 and A = (D - D^T)/2:
 * J(D) = J(S) because (ab|cd) = (ab|dc), and K(D) = K(S) + K(A): with the
 * reference's K_ac = sum_bd (ab|cd) D_bd and the unique-quartet digestion
 * F'_ac += (ab|cd) A_bd of A, K(A) = F'_A - F'_A^T (antisymmetric, zero
 * diagonal).  Returns the quartets of the S build.
 */
long oracle_fock_general(const oracle_basis_t *bs, const double *dm, int label, double xf, int ithreshold,
                         int nthreads, double *gout)
{
    const int n = bs->nao;
    const size_t nn = (size_t)n * n;
    double kc;
    switch (label) {
    case LBL_2JK: kc = -1.0; break;
    case LBL_2JKX: kc = -xf; break;
    case LBL_J: kc = 0.0; break;
    case LBL_K: kc = 1.0; break;
    case LBL_KX: kc = xf; break;
    default: return -1;
    }
    double *s = calloc(nn, sizeof(double)), *a = calloc(nn, sizeof(double));
    for (int p = 0; p < n; p++)
        for (int q = 0; q < n; q++) {
            const double x = dm[(size_t)p * n + q], y = dm[(size_t)q * n + p];
            s[(size_t)p * n + q] = 0.5 * (x + y);
            a[(size_t)p * n + q] = 0.5 * (x - y);
        }
    long total = 0;
    long nq = fock_impl(bs, s, label, xf, 0.0, ithreshold, nthreads, 1, NULL, gout, &total, 0);
    if (nq >= 0 && kc != 0.0) {
        double *ga = malloc(sizeof(double) * nn);
        const long na = fock_impl(bs, a, LBL_K, 1.0, 0.0, ithreshold, nthreads, 1, NULL, ga, &total, 1);
        if (na < 0) nq = -1;
        else
            for (size_t i = 0; i < nn; i++) gout[i] += kc * ga[i];  /* ga = F'_A - F'_A^T */
        free(ga);
    }
    free(s);
    free(a);
    return nq;
}

/* bins_in: optional nshells x nshells pair bins from oracle_pair_bins */
static long fock_impl(const oracle_basis_t *bs, const double *dm, int label, double xf, double omega, int ithreshold,
                      int nthreads, int stride, const int *bins_in, double *gout, long *total, int antisym)
{
    const int ns = bs->nshells, n = bs->nao;
    if (stride < 1) return -1;
    if (label < 0 || label > 4) return -1;
    double *maxv = malloc(sizeof(double) * ns * ns);
    int *bins = malloc(sizeof(int) * ns * ns);
    if (ithreshold >= 0 && bins_in) {
        memcpy(bins, bins_in, sizeof(int) * ns * ns);
    } else if (ithreshold >= 0) {
        if (oracle_pair_bins(bs, maxv, bins)) { free(maxv); free(bins); return -1; }
    } else {
        for (int i = 0; i < ns * ns; i++) bins[i] = 0;
    }
    /* list of kept pairs (sa <= sb) */
    int npairs = 0;
    int *pa = malloc(sizeof(int) * ns * (ns + 1) / 2), *pb = malloc(sizeof(int) * ns * (ns + 1) / 2);
    for (int sa = 0; sa < ns; sa++)
        for (int sb = sa; sb < ns; sb++)
            if (bins[sa * ns + sb] >= 0) { pa[npairs] = sa; pb[npairs] = sb; npairs++; }
    long nquart = 0;
    long ntotal = 0;
    for (int ip = 0; ip < npairs; ip++) {
        const int bi = bins[pa[ip] * ns + pb[ip]];
        for (int jp = ip; jp < npairs; jp++)
            if (ithreshold < 0 || bi + bins[pa[jp] * ns + pb[jp]] <= ithreshold) ntotal++;
    }
    *total = ntotal;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    memset(gout, 0, sizeof(double) * n * n);
#pragma omp parallel reduction(+ : nquart)
    {
        double *fp = calloc((size_t)n * n, sizeof(double));
        double *buf = malloc(sizeof(double) * 7 * 7 * 7 * 7);
#pragma omp for schedule(dynamic)
        for (int ip = 0; ip < npairs; ip++) {
            if (ip % stride) continue;
            const int bi = bins[pa[ip] * ns + pb[ip]];
            for (int jp = ip; jp < npairs; jp++) {
                const int bj = bins[pa[jp] * ns + pb[jp]];
                if (ithreshold >= 0 && bi + bj > ithreshold) continue;
                eri_sph_omega(bs, pa[ip], pb[ip], pa[jp], pb[jp], omega, buf);
                digest(bs, pa[ip], pb[ip], pa[jp], pb[jp], buf, dm, label, xf, fp);
                nquart++;
            }
        }
#pragma omp critical
        for (size_t i = 0; i < (size_t)n * n; i++) gout[i] += fp[i];
        free(fp);
        free(buf);
    }
    /* symmetrize: G = F' + F'^T (antisym: G = F' - F'^T) */
    for (int p = 0; p < n; p++)
        for (int q = p; q < n; q++) {
            const double x = gout[(size_t)p * n + q], y = gout[(size_t)q * n + p];
            gout[(size_t)p * n + q] = antisym ? x - y : x + y;
            gout[(size_t)q * n + p] = antisym ? y - x : x + y;
        }
    free(maxv);
    free(bins);
    free(pa);
    free(pb);
    return nquart;
}
