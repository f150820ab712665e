// One-electron integrals S, T and V in the spherical AO basis (host C++).
//
// Inputs of the SCF loop around the two-electron Fock build: the reference
// computes them with its COverlapDriver / CKineticEnergyDriver /
// CNuclearPotentialDriver (src/t2c_overlap, src/t2c_kinetic_energy,
// src/t2c_nuclear_potential).  Here they are McMurchie-Davidson integrals over
// the same normalised contracted shells, transformed to real solid harmonics
// with the factors of src/general/SphericalMomentum.hpp and placed in the
// reference AO order (ao_base + m * ao_stride).  O(N^2) work, threaded over
// shell pairs.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <vector>

#include "../../include/vlxfock.h"

namespace {

constexpr int LMAX = 4;
constexpr int DI = LMAX + 1;          // i index
constexpr int DJ = LMAX + 3;          // j index (kinetic needs j + 2)
constexpr int DT = 2 * LMAX + 3;      // t index

int ncart(int l) { return (l + 1) * (l + 2) / 2; }

void cart_exps(int l, int c, int* e)
{
    int n = 0;
    for (int i = l; i >= 0; i--)
        for (int j = l - i; j >= 0; j--, n++)
            if (n == c) {
                e[0] = i;
                e[1] = j;
                e[2] = l - i - j;
                return;
            }
}

// spherical component m of shell l as Cartesian (index, factor) terms
int sph_terms(int l, int m, int* ci, double* cf)
{
    const double f3 = 2.0 * std::sqrt(3.0), g5 = std::sqrt(2.5), g15 = 2.0 * std::sqrt(15.0), g3 = std::sqrt(1.5);
    const double h35 = 4.0 * std::sqrt(35.0), h17 = 4.0 * std::sqrt(17.5), h5 = 4.0 * std::sqrt(5.0),
                 h2 = 4.0 * std::sqrt(2.5);
    static const int pmap[3] = {1, 2, 0};
    switch (l) {
    case 0: ci[0] = 0; cf[0] = 1.0; return 1;
    case 1: ci[0] = pmap[m]; cf[0] = 1.0; return 1;
    case 2:
        switch (m) {
        case 0: ci[0] = 1; cf[0] = f3; return 1;
        case 1: ci[0] = 4; cf[0] = f3; return 1;
        case 2: ci[0] = 0; cf[0] = -1.0; ci[1] = 3; cf[1] = -1.0; ci[2] = 5; cf[2] = 2.0; return 3;
        case 3: ci[0] = 2; cf[0] = f3; return 1;
        default: ci[0] = 0; cf[0] = 0.5 * f3; ci[1] = 3; cf[1] = -0.5 * f3; return 2;
        }
    case 3:
        switch (m) {
        case 0: ci[0] = 1; cf[0] = 3.0 * g5; ci[1] = 6; cf[1] = -g5; return 2;
        case 1: ci[0] = 4; cf[0] = g15; return 1;
        case 2: ci[0] = 8; cf[0] = 4.0 * g3; ci[1] = 1; cf[1] = -g3; ci[2] = 6; cf[2] = -g3; return 3;
        case 3: ci[0] = 9; cf[0] = 2.0; ci[1] = 2; cf[1] = -3.0; ci[2] = 7; cf[2] = -3.0; return 3;
        case 4: ci[0] = 5; cf[0] = 4.0 * g3; ci[1] = 0; cf[1] = -g3; ci[2] = 3; cf[2] = -g3; return 3;
        case 5: ci[0] = 2; cf[0] = 0.5 * g15; ci[1] = 7; cf[1] = -0.5 * g15; return 2;
        default: ci[0] = 0; cf[0] = g5; ci[1] = 3; cf[1] = -3.0 * g5; return 2;
        }
    default: {
        // g: x^a y^b z^c in the order xxxx, xxxy, xxxz, xxyy, xxyz, xxzz, xyyy, xyyz, xyzz, xzzz,
        // yyyy, yyyz, yyzz, yzzz, zzzz (indices 0..14)
        struct T { int n; int i[6]; double f[6]; };
        const T g[9] = {
            {2, {1, 6}, {h35, -h35}},
            {2, {4, 11}, {3.0 * h17, -h17}},
            {3, {8, 1, 6}, {6.0 * h5, -h5, -h5}},
            {3, {13, 4, 11}, {4.0 * h2, -3.0 * h2, -3.0 * h2}},
            {6, {14, 0, 10, 3, 5, 12}, {8.0, 3.0, 3.0, 6.0, -24.0, -24.0}},
            {3, {9, 2, 7}, {4.0 * h2, -3.0 * h2, -3.0 * h2}},
            {4, {5, 12, 0, 10}, {3.0 * h5, -3.0 * h5, -0.5 * h5, 0.5 * h5}},
            {2, {2, 7}, {h17, -3.0 * h17}},
            {3, {0, 10, 3}, {0.25 * h35, 0.25 * h35, -1.5 * h35}},
        };
        for (int k = 0; k < g[m].n; k++) {
            ci[k] = g[m].i[k];
            cf[k] = g[m].f[k];
        }
        return g[m].n;
    }
    }
}

// Boys F_0..F_m(t): series + downward recursion below t = 35, erf + upward above.
void boys(int m, double t, double* f)
{
    const double et = std::exp(-t);
    if (t < 35.0) {
        double term = 1.0 / (2 * m + 1), sum = term;
        for (int k = 1; k < 500; k++) {
            term *= 2.0 * t / (2 * m + 2 * k + 1);
            sum += term;
            if (term < 1e-17 * sum) break;
        }
        f[m] = et * sum;
        for (int n = m - 1; n >= 0; n--) f[n] = (2.0 * t * f[n + 1] + et) / (2 * n + 1);
    } else {
        f[0] = 0.5 * std::sqrt(M_PI / t) * std::erf(std::sqrt(t));
        for (int n = 0; n < m; n++) f[n + 1] = ((2 * n + 1) * f[n] - et) / (2.0 * t);
    }
}

// 1D Hermite coefficients E[i][j][t]
void hermite(int imax, int jmax, double p, double xpa, double xpb, double e[DI][DJ][DT])
{
    for (int i = 0; i < DI; i++)
        for (int j = 0; j < DJ; j++)
            for (int t = 0; t < DT; t++) e[i][j][t] = 0.0;
    e[0][0][0] = 1.0;
    const double h = 0.5 / p;
    for (int i = 0; i <= imax; i++)
        for (int j = 0; j <= jmax; j++) {
            if (i == 0 && j == 0) continue;
            for (int t = 0; t <= i + j; t++) {
                double v;
                if (j == 0) {
                    v = xpa * e[i - 1][0][t];
                    if (t > 0) v += h * e[i - 1][0][t - 1];
                    if (t + 1 <= i - 1) v += (t + 1) * e[i - 1][0][t + 1];
                } else {
                    v = xpb * e[i][j - 1][t];
                    if (t > 0) v += h * e[i][j - 1][t - 1];
                    if (t + 1 <= i + j - 1) v += (t + 1) * e[i][j - 1][t + 1];
                }
                e[i][j][t] = v;
            }
        }
}

struct Shells {
    int n, nao;
    const int *l, *nprim, *poff, *base, *stride;
    const double *exps, *norms, *xyz;
    int natoms;
    const double *charges, *atoms;
};

// Cartesian blocks of S, T, V for shells (sa, sb)
void cart_block(const Shells& s, int sa, int sb, double* cs, double* ct, double* cv)
{
    const int la = s.l[sa], lb = s.l[sb];
    const int na = ncart(la), nb = ncart(lb);
    const double* A = s.xyz + 3 * sa;
    const double* B = s.xyz + 3 * sb;
    int ea[15][3], eb[15][3];
    for (int i = 0; i < na; i++) cart_exps(la, i, ea[i]);
    for (int i = 0; i < nb; i++) cart_exps(lb, i, eb[i]);
    std::fill(cs, cs + na * nb, 0.0);
    std::fill(ct, ct + na * nb, 0.0);
    std::fill(cv, cv + na * nb, 0.0);
    const double ab2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]);
    static thread_local double e[3][DI][DJ][DT];
    const int L = la + lb;
    static thread_local double rt[2 * LMAX + 1][2 * LMAX + 1][2 * LMAX + 1][2 * LMAX + 1];
    for (int i = 0; i < s.nprim[sa]; i++)
        for (int j = 0; j < s.nprim[sb]; j++) {
            const double a = s.exps[s.poff[sa] + i], b = s.exps[s.poff[sb] + j];
            const double p = a + b, mu = a * b / p;
            // every S, T and V term of the primitive pair carries e^(-mu |AB|^2): below e^-60
            // (~1e-26) the pair adds nothing at double precision to the matrices it feeds
            if (mu * ab2 > 60.0) continue;
            const double k = s.norms[s.poff[sa] + i] * s.norms[s.poff[sb] + j] * std::exp(-mu * ab2);
            double P[3];
            for (int x = 0; x < 3; x++) P[x] = (a * A[x] + b * B[x]) / p;
            for (int x = 0; x < 3; x++) hermite(la, lb + 2, p, P[x] - A[x], P[x] - B[x], e[x]);
            const double s1 = std::sqrt(M_PI / p);
            // 1D overlap and kinetic factors
            auto s1d = [&](int x, int ii, int jj) { return (jj < 0) ? 0.0 : e[x][ii][jj][0] * s1; };
            auto t1d = [&](int x, int ii, int jj) {
                return -0.5 * (jj * (jj - 1) * s1d(x, ii, jj - 2) - 2.0 * b * (2 * jj + 1) * s1d(x, ii, jj) +
                               4.0 * b * b * s1d(x, ii, jj + 2));
            };
            for (int ia = 0; ia < na; ia++)
                for (int ib = 0; ib < nb; ib++) {
                    const int *xa = ea[ia], *xb = eb[ib];
                    const double sx = s1d(0, xa[0], xb[0]), sy = s1d(1, xa[1], xb[1]), sz = s1d(2, xa[2], xb[2]);
                    cs[ia * nb + ib] += k * sx * sy * sz;
                    ct[ia * nb + ib] += k * (t1d(0, xa[0], xb[0]) * sy * sz + sx * t1d(1, xa[1], xb[1]) * sz +
                                             sx * sy * t1d(2, xa[2], xb[2]));
                }
            // nuclear attraction
            for (int c = 0; c < s.natoms; c++) {
                const double* C = s.atoms + 3 * c;
                const double pc[3] = {P[0] - C[0], P[1] - C[1], P[2] - C[2]};
                double f[2 * LMAX + 1];
                boys(L, p * (pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]), f);
                // R^n_tuv, level by level
                double fac = 1.0;
                for (int nn = 0; nn <= L; nn++) {
                    rt[nn][0][0][0] = fac * f[nn];
                    fac *= -2.0 * p;
                }
                for (int nn = L - 1; nn >= 0; nn--)
                    for (int sum = 1; sum <= L - nn; sum++)
                        for (int t = 0; t <= sum; t++)
                            for (int u = 0; u <= sum - t; u++) {
                                const int v = sum - t - u;
                                double val;
                                if (t > 0) {
                                    val = pc[0] * rt[nn + 1][t - 1][u][v];
                                    if (t > 1) val += (t - 1) * rt[nn + 1][t - 2][u][v];
                                } else if (u > 0) {
                                    val = pc[1] * rt[nn + 1][t][u - 1][v];
                                    if (u > 1) val += (u - 1) * rt[nn + 1][t][u - 2][v];
                                } else {
                                    val = pc[2] * rt[nn + 1][t][u][v - 1];
                                    if (v > 1) val += (v - 1) * rt[nn + 1][t][u][v - 2];
                                }
                                rt[nn][t][u][v] = val;
                            }
                const double pref = -s.charges[c] * 2.0 * M_PI / p * k;
                for (int ia = 0; ia < na; ia++)
                    for (int ib = 0; ib < nb; ib++) {
                        const int *xa = ea[ia], *xb = eb[ib];
                        double acc = 0.0;
                        for (int t = 0; t <= xa[0] + xb[0]; t++)
                            for (int u = 0; u <= xa[1] + xb[1]; u++)
                                for (int v = 0; v <= xa[2] + xb[2]; v++)
                                    acc += e[0][xa[0]][xb[0]][t] * e[1][xa[1]][xb[1]][u] * e[2][xa[2]][xb[2]][v] *
                                           rt[0][t][u][v];
                        cv[ia * nb + ib] += pref * acc;
                    }
            }
        }
}

}  // namespace

extern "C" int vlx_one_electron(int nshells, const int* l, const int* nprim, const int* prim_offset,
                                const double* exps, const double* norms, const double* centers, const int* ao_base,
                                const int* ao_stride, int nao, int natoms, const double* charges,
                                const double* atom_xyz, double* S, double* T, double* V)
{
    if (nshells <= 0 || nao <= 0 || !l || !nprim || !prim_offset || !exps || !norms || !centers || !ao_base ||
        !ao_stride || !S || !T || !V || (natoms > 0 && (!charges || !atom_xyz)))
        return VLX_ERR_ARG;
    for (int i = 0; i < nshells; i++)
        if (l[i] < 0 || l[i] > LMAX) return VLX_ERR_UNSUPPORTED;
    Shells s{nshells, nao, l, nprim, prim_offset, ao_base, ao_stride, exps, norms, centers, natoms, charges, atom_xyz};
    std::atomic<int> next{0};
    auto work = [&]() {
        std::vector<double> cs(225), ct(225), cv(225);
        for (;;) {
            const int sa = next.fetch_add(1);
            if (sa >= nshells) break;
            for (int sb = 0; sb <= sa; sb++) {
                cart_block(s, sa, sb, cs.data(), ct.data(), cv.data());
                const int la = l[sa], lb = l[sb], nb = ncart(lb);
                for (int ma = 0; ma < 2 * la + 1; ma++) {
                    int cia[6], cib[6];
                    double cfa[6], cfb[6];
                    const int nta = sph_terms(la, ma, cia, cfa);
                    for (int mb = 0; mb < 2 * lb + 1; mb++) {
                        const int ntb = sph_terms(lb, mb, cib, cfb);
                        double vs = 0.0, vt = 0.0, vv = 0.0;
                        for (int x = 0; x < nta; x++)
                            for (int y = 0; y < ntb; y++) {
                                const double c = cfa[x] * cfb[y];
                                const int idx = cia[x] * nb + cib[y];
                                vs += c * cs[idx];
                                vt += c * ct[idx];
                                vv += c * cv[idx];
                            }
                        const size_t p = ao_base[sa] + (size_t)ma * ao_stride[sa];
                        const size_t q = ao_base[sb] + (size_t)mb * ao_stride[sb];
                        S[p * nao + q] = S[q * nao + p] = vs;
                        T[p * nao + q] = T[q * nao + p] = vt;
                        V[p * nao + q] = V[q * nao + p] = vv;
                    }
                }
            }
        }
    };
    const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nth; t++) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
    return VLX_OK;
}
