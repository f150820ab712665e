// Wavefront-cooperative electron-repulsion kernels for the largest classes
// (f shells), where straight-line recursions would neither fit the register
// file nor compile in reasonable time.
//
// One shell quartet at a time per wavefront (work-group = one wavefront),
// McMurchie-Davidson in matrix form, everything staged in LDS:
//   U_AB[ab][p]  = k_ab sum_{cart a,b} S_a S_b E^{ab,x}_t E^{ab,y}_u E^{ab,z}_v
//                  (Hermite expansion of a bra primitive pair projected on the
//                   real solid harmonics of src/general/SphericalMomentum.hpp)
//   R[n][tuv]    = Hermite Coulomb integrals, by the recursion
//                  R^n_{t+1,u,v} = t R^{n+1}_{t-1,u,v} + X_PQ R^{n+1}_{t,u,v}
//                  from R^n_000 = (-2 rho)^n F_n(rho |PQ|^2)
//   X[p][cd]     = sum_q R[p+q] (-1)^{|q|} U_CD[cd][q]
//   out[ab][cd] += sum_p U_AB[ab][p] X[p][cd] / sqrt(zeta + eta)
// with lanes parallel over matrix entries.  The spherical block in LDS is
// then digested (Fock), stored (explicit quartets) or reduced (screener).
#pragma once

#include <algorithm>

#include "eri_kernels.h"

namespace md {

constexpr int nherm(int l) { return (l + 1) * (l + 2) * (l + 3) / 6; }
constexpr int D1 = VLX_MAX_L + 1;
constexpr int DT = 2 * VLX_MAX_L + 1;
constexpr int E3 = D1 * D1 * DT;  // 1D coefficients per axis; an E array holds 3 axes

// linear index of Hermite (t,u,v): degree-major, then (t descending, u descending)
__device__ __forceinline__ int hidx(int t, int u, int v)
{
    const int n = t + u + v, s = u + v;
    return n * (n + 1) * (n + 2) / 6 + s * (s + 1) / 2 + v;
}

__device__ __forceinline__ void hdecode(int idx, int& t, int& u, int& v)
{
    int n = 0;
    while (nherm(n) <= idx) n++;
    const int rem = idx - (n > 0 ? nherm(n - 1) : 0);
    int s = 0;
    while ((s + 1) * (s + 2) / 2 <= rem) s++;
    v = rem - s * (s + 1) / 2;
    u = s - v;
    t = n - s;
}

// Cartesian exponents of component c of shell l (xx, xy, xz, yy, yz, zz, ...)
__device__ __forceinline__ void cart_exp(int l, int c, int& x, int& y, int& z)
{
    int i = 0;
    for (int xx = l; xx >= 0; xx--)
        for (int yy = l - xx; yy >= 0; yy--, i++)
            if (i == c) {
                x = xx;
                y = yy;
                z = l - xx - yy;
                return;
            }
    x = y = z = 0;
}

// Spherical component m of shell l as Cartesian terms (index, factor); the
// factors and orders are those of src/general/SphericalMomentum.hpp:18-134.
__device__ __forceinline__ int sph_terms(int l, int m, int* ci, double* cf)
{
    const double f3 = 3.46410161513775458705;   // 2 sqrt(3)
    const double g5 = 1.58113883008418966599;   // sqrt(5/2)
    const double g15 = 7.74596669241483377036;  // 2 sqrt(15)
    const double g3 = 1.22474487139158904910;   // sqrt(3/2)
    switch (l * 8 + m) {
    case 0: ci[0] = 0; cf[0] = 1.0; return 1;
    case 8: ci[0] = 1; cf[0] = 1.0; return 1;
    case 9: ci[0] = 2; cf[0] = 1.0; return 1;
    case 10: ci[0] = 0; cf[0] = 1.0; return 1;
    case 16: ci[0] = 1; cf[0] = f3; return 1;
    case 17: ci[0] = 4; cf[0] = f3; return 1;
    case 18: ci[0] = 0; cf[0] = -1.0; ci[1] = 3; cf[1] = -1.0; ci[2] = 5; cf[2] = 2.0; return 3;
    case 19: ci[0] = 2; cf[0] = f3; return 1;
    case 20: ci[0] = 0; cf[0] = 0.5 * f3; ci[1] = 3; cf[1] = -0.5 * f3; return 2;
    case 24: ci[0] = 1; cf[0] = 3.0 * g5; ci[1] = 6; cf[1] = -g5; return 2;
    case 25: ci[0] = 4; cf[0] = g15; return 1;
    case 26: ci[0] = 8; cf[0] = 4.0 * g3; ci[1] = 1; cf[1] = -g3; ci[2] = 6; cf[2] = -g3; return 3;
    case 27: ci[0] = 9; cf[0] = 2.0; ci[1] = 2; cf[1] = -3.0; ci[2] = 7; cf[2] = -3.0; return 3;
    case 28: ci[0] = 5; cf[0] = 4.0 * g3; ci[1] = 0; cf[1] = -g3; ci[2] = 3; cf[2] = -g3; return 3;
    case 29: ci[0] = 2; cf[0] = 0.5 * g15; ci[1] = 7; cf[1] = -0.5 * g15; return 2;
    case 30: ci[0] = 0; cf[0] = g5; ci[1] = 3; cf[1] = -3.0 * g5; return 2;
    default: return 0;
    }
}

template <int la, int lb, int lc, int ld>
struct Sizes {
    static constexpr int L = la + lb + lc + ld;
    static constexpr int nhab = nherm(la + lb), nhcd = nherm(lc + ld), nr = nherm(L);
    static constexpr int nsa = 2 * la + 1, nsb = 2 * lb + 1, nsc = 2 * lc + 1, nsd = 2 * ld + 1;
    static constexpr int nsab = nsa * nsb, nscd = nsc * nsd;
    // LDS doubles: E (bra, ket), U_AB, U_CD, two R levels, X, out, scalars
    static constexpr int lds = 6 * E3 + nsab * nhab + nhcd * nscd + 2 * nr + nhab * nscd + nsab * nscd + 32;
    // the Boys table is staged in LDS when it fits beside the work arrays, else read from HBM/L2
    static constexpr bool tab_in_lds = (lds + nsab * nscd + BOYS_TAB_SIZE) * 8 <= 160 * 1024;
};

// 1D Hermite coefficients E^{ij}_t (i <= la, j <= lb) of one primitive pair for
// axis x = threadIdx.x < 3; pa = P - A, pb = P - B.
__device__ __forceinline__ void hermite_1d(int la, int lb, double p, const double* pa, const double* pb,
                                           double* __restrict__ e)
{
    const int x = threadIdx.x;
    if (x >= 3) return;
    double* ex = e + x * E3;
    for (int k = 0; k < E3; k++) ex[k] = 0.0;
    const double h = 0.5 / p;
    ex[0] = 1.0;
    for (int i = 0; i <= la; i++)
        for (int j = 0; j <= lb; j++) {
            if (i == 0 && j == 0) continue;
            for (int t = 0; t <= i + j; t++) {
                double v;
                if (j == 0) {
                    const double* src = ex + ((i - 1) * D1 + j) * DT;
                    v = pa[x] * src[t];
                    if (t > 0) v += h * src[t - 1];
                    if (t + 1 <= i - 1 + j) v += (t + 1) * src[t + 1];
                } else {
                    const double* src = ex + (i * D1 + (j - 1)) * DT;
                    v = pb[x] * src[t];
                    if (t > 0) v += h * src[t - 1];
                    if (t + 1 <= i + j - 1) v += (t + 1) * src[t + 1];
                }
                ex[(i * D1 + j) * DT + t] = v;
            }
        }
}

// Bra: u[ab * nh + p]; ket (sign (-1)^{t+u+v}): u[p * ns + cd]; scaled by k.
__device__ __forceinline__ void build_u(int la, int lb, const double* __restrict__ e, double k, bool ket,
                                        double* __restrict__ u)
{
    const int nsb = 2 * lb + 1, ns = (2 * la + 1) * nsb, nh = nherm(la + lb);
    for (int idx = threadIdx.x; idx < ns * nh; idx += blockDim.x) {
        const int ab = idx / nh, ph = idx % nh;
        int t, uu, v;
        hdecode(ph, t, uu, v);
        int cia[3], cib[3];
        double cfa[3], cfb[3];
        const int nta = sph_terms(la, ab / nsb, cia, cfa);
        const int ntb = sph_terms(lb, ab % nsb, cib, cfb);
        double acc = 0.0;
        for (int ia = 0; ia < nta; ia++) {
            int ax, ay, az;
            cart_exp(la, cia[ia], ax, ay, az);
            for (int ib = 0; ib < ntb; ib++) {
                int bx, by, bz;
                cart_exp(lb, cib[ib], bx, by, bz);
                if (t > ax + bx || uu > ay + by || v > az + bz) continue;
                acc += cfa[ia] * cfb[ib] * e[(ax * D1 + bx) * DT + t] * e[E3 + (ay * D1 + by) * DT + uu] *
                       e[2 * E3 + (az * D1 + bz) * DT + v];
            }
        }
        if (ket) {
            if ((t + uu + v) & 1) acc = -acc;
            u[ph * ns + ab] = k * acc;
        } else {
            u[ab * nh + ph] = k * acc;
        }
    }
}

// Cooperative spherical block (ab|cd) -> out[ab * nscd + cd] in LDS.
// ab_vec = A - B, cd_vec = C - D (P - B = (P - A) + (A - B)).
template <int la, int lb, int lc, int ld>
__device__ void quartet(const PrimPair* __restrict__ bra_prims, int bra_np, const double* ab_vec,
                        const PrimPair* __restrict__ ket_prims, int ket_np, const double* cd_vec,
                        const double* __restrict__ tab, const double omega2, double* __restrict__ lds,
                        double* __restrict__ out)
{
    using S = Sizes<la, lb, lc, ld>;
    constexpr int L = S::L;
    double* ebra = lds;
    double* eket = ebra + 3 * E3;
    double* uab = eket + 3 * E3;
    double* ucd = uab + S::nsab * S::nhab;
    double* ra = ucd + S::nhcd * S::nscd;
    double* rb = ra + S::nr;
    double* x = rb + S::nr;
    double* scal = x + S::nhab * S::nscd;  // Boys values F_0..F_L
    for (int k = threadIdx.x; k < S::nsab * S::nscd; k += blockDim.x) out[k] = 0.0;
    for (int ib = 0; ib < bra_np; ib++) {
        const PrimPair bp = bra_prims[ib];
        {
            const double pa[3] = {bp.pax, bp.pay, bp.paz};
            const double pb[3] = {bp.pax + ab_vec[0], bp.pay + ab_vec[1], bp.paz + ab_vec[2]};
            __syncthreads();
            hermite_1d(la, lb, bp.zeta, pa, pb, ebra);
            __syncthreads();
            build_u(la, lb, ebra, bp.k, false, uab);
        }
        for (int ik = 0; ik < ket_np; ik++) {
            const PrimPair kp = ket_prims[ik];
            const double qc[3] = {kp.pax, kp.pay, kp.paz};
            const double qd[3] = {kp.pax + cd_vec[0], kp.pay + cd_vec[1], kp.paz + cd_vec[2]};
            __syncthreads();
            hermite_1d(lc, ld, kp.zeta, qc, qd, eket);
            const double s = bp.zeta + kp.zeta;
            // erf(omega r)/r: rho -> sfac rho and a factor sqrt(sfac), sfac = omega^2 / (omega^2 + rho)
            const double rho_c = bp.zeta * kp.zeta / s;
            const double sfac = (omega2 > 0.0) ? omega2 / (omega2 + rho_c) : 1.0;
            const double rho = (omega2 > 0.0) ? rho_c * sfac : rho_c;
            const double pq[3] = {bp.px - kp.px, bp.py - kp.py, bp.pz - kp.pz};
            if (threadIdx.x == 3) {
                double F[L + 1];
                boys_eval<L, double>(rho * (pq[0] * pq[0] + pq[1] * pq[1] + pq[2] * pq[2]), tab, F);
                for (int m = 0; m <= L; m++) scal[m] = F[m];
            }
            __syncthreads();
            build_u(lc, ld, eket, kp.k, true, ucd);
            // R levels n = L .. 0 (ping-pong between ra and rb)
            double* cur = (L & 1) ? rb : ra;
            if (threadIdx.x == 0) {
                double f = 1.0;
                for (int m = 0; m < L; m++) f *= -2.0 * rho;
                cur[0] = f * scal[L];
            }
            for (int n = L - 1; n >= 0; n--) {
                double* prev = cur;
                cur = (cur == ra) ? rb : ra;
                __syncthreads();
                double fn = 1.0;
                for (int m = 0; m < n; m++) fn *= -2.0 * rho;
                for (int idx = threadIdx.x; idx < nherm(L - n); idx += blockDim.x) {
                    int t, u, v;
                    hdecode(idx, t, u, v);
                    double val;
                    if (t > 0) {
                        val = pq[0] * prev[hidx(t - 1, u, v)];
                        if (t > 1) val += (t - 1) * prev[hidx(t - 2, u, v)];
                    } else if (u > 0) {
                        val = pq[1] * prev[hidx(t, u - 1, v)];
                        if (u > 1) val += (u - 1) * prev[hidx(t, u - 2, v)];
                    } else if (v > 0) {
                        val = pq[2] * prev[hidx(t, u, v - 1)];
                        if (v > 1) val += (v - 1) * prev[hidx(t, u, v - 2)];
                    } else {
                        val = fn * scal[n];
                    }
                    cur[idx] = val;
                }
            }
            __syncthreads();
            // X[p][cd] = sum_q R[p+q] U_CD[q][cd]
            for (int idx = threadIdx.x; idx < S::nhab * S::nscd; idx += blockDim.x) {
                const int p = idx / S::nscd, cd = idx % S::nscd;
                int t, u, v;
                hdecode(p, t, u, v);
                double acc = 0.0;
                for (int q = 0; q < S::nhcd; q++) {
                    int tq, uq, vq;
                    hdecode(q, tq, uq, vq);
                    acc = fma(cur[hidx(t + tq, u + uq, v + vq)], ucd[q * S::nscd + cd], acc);
                }
                x[idx] = acc;
            }
            __syncthreads();
            const double w = (omega2 > 0.0) ? sqrt(sfac) / sqrt(s) : 1.0 / sqrt(s);
            for (int idx = threadIdx.x; idx < S::nsab * S::nscd; idx += blockDim.x) {
                const int ab = idx / S::nscd, cd = idx % S::nscd;
                double acc = 0.0;
                for (int p = 0; p < S::nhab; p++) acc = fma(uab[ab * S::nhab + p], x[p * S::nscd + cd], acc);
                out[idx] += w * acc;
            }
        }
    }
    __syncthreads();
}

// Fock digestion of one block in LDS (factor s included), lanes over outputs.
template <int la, int lb, int lc, int ld, bool DET = false>
__device__ void digest(const PairInfo& bp, const PairInfo& kp, double s, const double* __restrict__ out,
                       const double* __restrict__ D, double* __restrict__ F, const int n, const double fj,
                       const double fk)
{
    using S = Sizes<la, lb, lc, ld>;
    constexpr int na = S::nsa, nb = S::nsb, nc = S::nsc, nd = S::nsd;
    auto I = [&](int i, int j, int k, int l) { return out[(i * nb + j) * S::nscd + k * nd + l]; };
    auto P = [&](int i) { return bp.ao_a + i * bp.str_a; };
    auto Q = [&](int j) { return bp.ao_b + j * bp.str_b; };
    auto R = [&](int k) { return kp.ao_a + k * kp.str_a; };
    auto T = [&](int l) { return kp.ao_b + l * kp.str_b; };
    const int lane = threadIdx.x;
    if (fj != 0.0) {
        for (int e = lane; e < na * nb; e += blockDim.x) {
            const int i = e / nb, j = e % nb;
            double v = 0.0;
            for (int k = 0; k < nc; k++)
                for (int l = 0; l < nd; l++) v = fma(I(i, j, k, l), D[(size_t)R(k) * n + T(l)], v);
            acc_add<DET>(F, (size_t)P(i) * n + Q(j), fj * s * v);
        }
        for (int e = lane; e < nc * nd; e += blockDim.x) {
            const int k = e / nd, l = e % nd;
            double v = 0.0;
            for (int i = 0; i < na; i++)
                for (int j = 0; j < nb; j++) v = fma(I(i, j, k, l), D[(size_t)P(i) * n + Q(j)], v);
            acc_add<DET>(F, (size_t)R(k) * n + T(l), fj * s * v);
        }
    }
    if (fk != 0.0) {
        for (int e = lane; e < na * nc; e += blockDim.x) {  // F'_ac += I D_bd
            const int i = e / nc, k = e % nc;
            double v = 0.0;
            for (int j = 0; j < nb; j++)
                for (int l = 0; l < nd; l++) v = fma(I(i, j, k, l), D[(size_t)Q(j) * n + T(l)], v);
            acc_add<DET>(F, (size_t)P(i) * n + R(k), fk * s * v);
        }
        for (int e = lane; e < na * nd; e += blockDim.x) {  // F'_ad += I D_bc
            const int i = e / nd, l = e % nd;
            double v = 0.0;
            for (int j = 0; j < nb; j++)
                for (int k = 0; k < nc; k++) v = fma(I(i, j, k, l), D[(size_t)Q(j) * n + R(k)], v);
            acc_add<DET>(F, (size_t)P(i) * n + T(l), fk * s * v);
        }
        for (int e = lane; e < nb * nc; e += blockDim.x) {  // F'_bc += I D_ad
            const int j = e / nc, k = e % nc;
            double v = 0.0;
            for (int i = 0; i < na; i++)
                for (int l = 0; l < nd; l++) v = fma(I(i, j, k, l), D[(size_t)P(i) * n + T(l)], v);
            acc_add<DET>(F, (size_t)Q(j) * n + R(k), fk * s * v);
        }
        for (int e = lane; e < nb * nd; e += blockDim.x) {  // F'_bd += I D_ac
            const int j = e / nd, l = e % nd;
            double v = 0.0;
            for (int i = 0; i < na; i++)
                for (int k = 0; k < nc; k++) v = fma(I(i, j, k, l), D[(size_t)P(i) * n + R(k)], v);
            acc_add<DET>(F, (size_t)Q(j) * n + T(l), fk * s * v);
        }
    }
}

}  // namespace md

// ------------------------------------------------------------------------
// Matrix-core McMurchie-Davidson Fock kernel.
//
// The Hermite-to-spherical matrices of every primitive pair are built once
// per screener (u_build_kernel) and stored in HBM as
//   U[prim][h][s] = k sum_{cart a, b} S_a S_b E^x E^y E^z
// (h < nherm(la + lb) Hermite index, s < (2la+1)(2lb+1) spherical index, both
// padded to multiples of 16 with zeros).  One wavefront takes one shell
// quartet at a time; per primitive quartet it builds R^0[tuv] (scaled by
// 1/sqrt(zeta + eta)) in LDS and runs two FP64 matrix-core products
//   X[p][cd]    = sum_q (-1)^{|q|} R^0[p + q] U_CD[q][cd]       (v_mfma_f64_16x16x4)
//   out[ab][cd] += sum_p U_AB[p][ab] X[p][cd]
// one 16-wide cd tile at a time; out stays in the matrix-core accumulators
// over all primitive quartets of the shell quartet.  Operand layouts of
// v_mfma_f64_16x16x4f64 (wave64): A[i][k] in lane i + 16k, B[k][j] in lane
// j + 16k, D[i][j] in register i / 4 of lane j + 16 (i % 4).

using md::hidx;
using md::nherm;

struct HermDecode {
    int e[nherm(VLX_MAX_LTOT)];  // t | u << 8 | v << 16 of Hermite index idx
};
constexpr HermDecode make_herm_decode()
{
    HermDecode d{};
    int i = 0;
    for (int n = 0; n <= VLX_MAX_LTOT; n++)
        for (int s = 0; s <= n; s++)
            for (int v = 0; v <= s; v++) d.e[i++] = (n - s) | ((s - v) << 8) | (v << 16);
    return d;
}
static __device__ constexpr HermDecode g_herm = make_herm_decode();

// The step of the R recursion that produces Hermite index idx (t, u, v), as one packed int:
// s1 | s2 << 9 | axis << 18 | c << 20.  Along the first axis with a nonzero index (x for t > 0,
// else y, else z; k the index along it), R^n[idx] = PQ_axis R^{n+1}[s1] + c R^{n+1}[s2] with
// s1 the index one lower along the axis, s2 two lower and c = k - 1 (s2 = 0 and c = 0 for k = 1;
// idx 0 takes the seed).  Indices stay below 2^9 (nherm(VLX_MAX_LTOT) = 455).
struct HermRecur {
    int e[nherm(VLX_MAX_LTOT)];
};
constexpr HermRecur make_herm_recur()
{
    HermRecur r{};
    auto lin = [](int t, int u, int v) {
        const int n = t + u + v, s = u + v;
        return n * (n + 1) * (n + 2) / 6 + s * (s + 1) / 2 + v;
    };
    int i = 0;
    for (int n = 0; n <= VLX_MAX_LTOT; n++)
        for (int s = 0; s <= n; s++)
            for (int v = 0; v <= s; v++) {
                const int t = n - s, u = s - v;
                int s1 = 0, s2 = 0, axis = 0, c = 0;
                if (t > 0) {
                    s1 = lin(t - 1, u, v);
                    if (t > 1) s2 = lin(t - 2, u, v), c = t - 1;
                } else if (u > 0) {
                    axis = 1;
                    s1 = lin(t, u - 1, v);
                    if (u > 1) s2 = lin(t, u - 2, v), c = u - 1;
                } else if (v > 0) {
                    axis = 2;
                    s1 = lin(t, u, v - 1);
                    if (v > 1) s2 = lin(t, u, v - 2), c = v - 1;
                }
                r.e[i++] = s1 | (s2 << 9) | (axis << 18) | (c << 20);
            }
    return r;
}
static __device__ constexpr HermRecur g_hrec = make_herm_recur();
static_assert(nherm(VLX_MAX_LTOT) <= 512, "R recursion table packs indices in 9 bits");

__device__ __forceinline__ int hidx_packed(int tuv)
{
    return hidx(tuv & 0xff, (tuv >> 8) & 0xff, tuv >> 16);
}

typedef double vlx_d4 __attribute__((ext_vector_type(4)));

template <int la, int lb, int lc, int ld>
struct MmdSizes {
    static constexpr int L = la + lb + lc + ld;
    static constexpr int nhab = nherm(la + lb), nhcd = nherm(lc + ld), nr = nherm(L);
    static constexpr int nsab = (2 * la + 1) * (2 * lb + 1), nscd = (2 * lc + 1) * (2 * ld + 1);
    static constexpr int HA = vlx_u_rows(la, lb), SA = vlx_u_cols(la, lb);  // bra U: HA x SA
    static constexpr int HC = vlx_u_rows(lc, ld), SC = vlx_u_cols(lc, ld);  // ket U: HC x SC
    static constexpr int MT = SA / 16, NT = SC / 16, MX = HA / 16;
    static constexpr int KX = (nhcd + 3) / 4, KO = (nhab + 3) / 4;
    static constexpr int XC = MX < 2 ? MX : 2;  // X row tiles per chunk
    // operands preloaded into registers: the X product's U_CD column when its K loop is short,
    // and the whole U_AB operand of the out product when one chunk covers all its K steps
    static constexpr bool PRE_X = KX <= 8;
    static constexpr bool PRE_O = MX <= XC && KO * MT <= 16;
};

// Work item {j0, i0, i1}: bra i0 with kets [j0, i1), one shell quartet at a time.  With
// a.ndens > 1 each block is digested into every density (density d at D0 + d * dstride, its F' at
// F0 + d * dstride, factors dfac[2 d], dfac[2 d + 1]).  a.det: F0 holds fixed-point accumulators
// (fx_add, two words per element) and the factors come from dfac also for one density.
template <int la, int lb, int lc, int ld, bool RS>
__global__ void __launch_bounds__(VLX_WAVE) __attribute__((amdgpu_waves_per_eu(4)))
    fock_kernel_mmd(const PairInfo* __restrict__ bra, const PrimPair* __restrict__ bra_prims,
                    const double* __restrict__ bra_u, const PairInfo* __restrict__ ket,
                    const PrimPair* __restrict__ ket_prims, const double* __restrict__ ket_u,
                    const WorkItem* __restrict__ items, const double* __restrict__ boys_tab,
                    const double* __restrict__ D0, double* __restrict__ F0, const MolDesc* __restrict__ mols,
                    const int* __restrict__ glim, const double* __restrict__ dfac, const FockScalars a)
{
    using S = MmdSizes<la, lb, lc, ld>;
    constexpr int L = S::L;
    __shared__ double rbuf[2 * S::nr];
    __shared__ double seeds[L + 1];
    __shared__ double xt[S::XC * 16 * 16];
    __shared__ double outs[S::nsab * S::nscd];
    const int lane = threadIdx.x;
    const int lr = lane & 15, lk = lane >> 4;
    static_assert(16 * (S::MX + S::XC) <= nherm(VLX_MAX_LTOT) && 4 * S::KX <= nherm(VLX_MAX_LTOT),
                  "X-product rows and columns index the Hermite decode table");
    // packed Hermite (t, u, v) of this lane's X-product columns q = 4 kq + lk (0 beyond nhcd): the
    // same for every quartet, kept in registers when the column loop is unrolled
    int tqv[S::PRE_X ? S::KX : 1];
#pragma unroll
    for (int kq = 0; kq < (S::PRE_X ? S::KX : 1); kq++) {
        const int q = 4 * kq + lk;
        tqv[kq] = (S::PRE_X && q < S::nhcd) ? g_herm.e[q] : 0;
    }
    for (long long it = blockIdx.x; it < a.nitems; it += gridDim.x) {
        const WorkItem wi = items[it];
        const MolDesc md = mols[wi.mol];
        const double* __restrict__ D = D0 + md.base;
        double* __restrict__ F = F0 + md.base;
        const PairInfo bp = bra[wi.i0];
        for (int j = wi.j0; j < wi.i1; j++) {
            const PairInfo kp = ket[j];
            // density-weighted screening (wave-uniform: one quartet per wavefront)
            if (glim != nullptr &&
                bp.bin + kp.bin > quartet_bin_limit(glim, a.ngroups, bp.grp_a, bp.grp_b, kp.grp_a, kp.grp_b))
                continue;
            // one 16-wide cd tile at a time: its accumulators persist over all primitive quartets
            for (int nt = 0; nt < S::NT; nt++) {
                vlx_d4 acc[S::MT];
#pragma unroll
                for (int m = 0; m < S::MT; m++) acc[m] = vlx_d4{0.0, 0.0, 0.0, 0.0};
                for (int ib = 0; ib < bp.nprim; ib++) {
                    const PrimPair pb = bra_prims[bp.prim_start + ib];
                    const double* ua = bra_u + (size_t)(bp.prim_start + ib) * (S::HA * S::SA);
                    for (int ik = 0; ik < kp.nprim; ik++) {
                        const PrimPair pk = ket_prims[kp.prim_start + ik];
                        const double* uc = ket_u + (size_t)(kp.prim_start + ik) * (S::HC * S::SC) + 16 * nt + lr;
                        // short K loops: both products' U operands go to registers here, so that
                        // their memory latency overlaps the Boys function and the R levels
                        double ucv[S::PRE_X ? S::KX : 1], uav[S::PRE_O ? S::KO : 1][S::PRE_O ? S::MT : 1];
                        if constexpr (S::PRE_X) {
#pragma unroll
                            for (int kq = 0; kq < S::KX; kq++) ucv[kq] = uc[(4 * kq + lk) * S::SC];
                        }
                        if constexpr (S::PRE_O) {
#pragma unroll
                            for (int kp2 = 0; kp2 < S::KO; kp2++)
#pragma unroll
                                for (int ms = 0; ms < S::MT; ms++) uav[kp2][ms] = ua[(4 * kp2 + lk) * S::SA + 16 * ms + lr];
                        }
                        const double sz = pb.zeta + pk.zeta;
                        const double rho_c = pb.zeta * pk.zeta / sz;
                        double rho = rho_c, w = vrsqrt<double>(sz);
                        if constexpr (RS) {
                            const double sfac = a.omega2 / (a.omega2 + rho_c);
                            rho = rho_c * sfac;
                            w *= sqrt(sfac);
                        }
                        const double pqx = pb.px - pk.px, pqy = pb.py - pk.py, pqz = pb.pz - pk.pz;
                        double Fm[L + 1];
                        boys_eval<L, double>(rho * (pqx * pqx + pqy * pqy + pqz * pqz), boys_tab, Fm);
                        // R levels n = L .. 0, ping-pong in rbuf; R^n_000 = w (-2 rho)^n F_n, lane n
                        // keeps level n's seed in LDS (the level loop then needs no register array)
                        {
                            double f = w, mine = 0.0;
#pragma unroll
                            for (int n = 0; n <= L; n++) {
                                if (lane == n) mine = f * Fm[n];
                                f *= -2.0 * rho;
                            }
                            if (lane <= L) seeds[lane] = mine;
                            if (lane == L) rbuf[(L & 1) ? S::nr : 0] = mine;
                        }
                        double* cur = rbuf + ((L & 1) ? S::nr : 0);
#pragma unroll 1
                        for (int n = L - 1; n >= 0; n--) {
                            double* prev = cur;
                            cur = (cur == rbuf) ? rbuf + S::nr : rbuf;
                            __syncthreads();
                            const int nn = nherm(L - n);
                            for (int idx = lane; idx < nn; idx += VLX_WAVE) {
                                // one branch-free step from the packed table (idx 0: the seed)
                                const int rr = g_hrec.e[idx];
                                const int ax = (rr >> 18) & 3;
                                const double pq = ax == 0 ? pqx : (ax == 1 ? pqy : pqz);
                                const double val = fma(double(rr >> 20), prev[(rr >> 9) & 511], pq * prev[rr & 511]);
                                cur[idx] = idx == 0 ? seeds[n] : val;
                            }
                        }
                        __syncthreads();
                        // X = sum_q (-1)^|q| R^0[p + q] U_CD[q][cd] for the 16 columns cd of tile nt, in
                        // chunks of XC row tiles; each chunk feeds its K range of out += U_AB^T X at once
#pragma unroll 1
                        for (int pc = 0; pc < S::MX; pc += S::XC) {
                            vlx_d4 x[S::XC];
                            int prow[S::XC];
                            bool prow_ok[S::XC];
#pragma unroll
                            for (int i = 0; i < S::XC; i++) {
                                x[i] = vlx_d4{0.0, 0.0, 0.0, 0.0};
                                const int p = 16 * (pc + i) + lr;  // < 16 (MX + XC) <= HA + 16, inside g_herm
                                prow_ok[i] = p < S::nhab;
                                const int tp = g_herm.e[p];
                                prow[i] = prow_ok[i] ? tp : 0;
                            }
                            // branch-free steps: a masked row (p >= nhab) or column (q >= nhcd) has its
                            // Hermite index zeroed, so it reads an in-range R entry, and selects an
                            // exact zero; the LDS reads can then be issued ahead of the products
#pragma unroll(S::PRE_X ? S::KX : 4)
                            for (int kq = 0; kq < S::KX; kq++) {
                                const int q = 4 * kq + lk;
                                const bool qok = q < S::nhcd;
                                int tq;
                                if constexpr (S::PRE_X) {
                                    tq = tqv[kq];
                                } else {
                                    const int t0 = g_herm.e[q];
                                    tq = qok ? t0 : 0;
                                }
                                const bool neg = ((tq & 0xff) + ((tq >> 8) & 0xff) + (tq >> 16)) & 1;
                                double bv;
                                if constexpr (S::PRE_X) bv = ucv[kq];
                                else bv = uc[q * S::SC];
#pragma unroll
                                for (int i = 0; i < S::XC; i++) {
                                    // a last chunk past the MX row tiles (odd MX) has nothing to compute:
                                    // wave-uniform skip; its X rows are never read by the next product
                                    if (S::MX % S::XC != 0 && pc + i >= S::MX) continue;
                                    const double r0 = cur[hidx_packed(prow[i] + tq)];
                                    const double sr = neg ? -r0 : r0;
                                    const double av = (qok && prow_ok[i]) ? sr : 0.0;
                                    x[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, x[i], 0, 0, 0);
                                }
                            }
#pragma unroll
                            for (int i = 0; i < S::XC; i++)
#pragma unroll
                                for (int r = 0; r < 4; r++) xt[(16 * i + lk + 4 * r) * 16 + lr] = x[i][r];
                            __syncthreads();
                            // out[ab][cd] += sum_p U_AB[p][ab] X[p][cd] over this chunk's rows p
                            const int k0 = 4 * pc, k1 = std::min(S::KO, 4 * (pc + S::XC));
                            if constexpr (S::PRE_O) {
                                // all K steps of the product are in registers (one chunk covers them)
#pragma unroll
                                for (int kp2 = 0; kp2 < S::KO; kp2++) {
                                    const int pr = 4 * kp2 + lk;
                                    const double bv = xt[(pr - 16 * pc) * 16 + lr];
#pragma unroll
                                    for (int ms = 0; ms < S::MT; ms++)
                                        acc[ms] = __builtin_amdgcn_mfma_f64_16x16x4f64(uav[kp2][ms], bv, acc[ms], 0, 0, 0);
                                }
                            } else {
                                // unrolled so that the U_AB loads of later steps issue ahead of the products
#pragma unroll 4
                                for (int kp2 = k0; kp2 < k1; kp2++) {
                                    const int pr = 4 * kp2 + lk;
                                    const double bv = xt[(pr - 16 * pc) * 16 + lr];
#pragma unroll
                                    for (int ms = 0; ms < S::MT; ms++) {
                                        const double av = ua[pr * S::SA + 16 * ms + lr];
                                        acc[ms] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[ms], 0, 0, 0);
                                    }
                                }
                            }
                            __syncthreads();
                        }
                    }
                }
                // contracted tile -> LDS [ab][cd]
#pragma unroll
                for (int ms = 0; ms < S::MT; ms++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int ab = 16 * ms + lk + 4 * r, cd = 16 * nt + lr;
                        if (ab < S::nsab && cd < S::nscd) outs[ab * S::nscd + cd] = acc[ms][r];
                    }
            }
            __syncthreads();
            double sf = 1.0;
            if (bp.same) sf *= 0.5;
            if (kp.same) sf *= 0.5;
            if (a.same_class && bp.rank == j) sf *= 0.5;
            if (a.det) {  // F0: fixed-point accumulators (two words per element), factors in dfac
                for (int dd = 0; dd < max(a.ndens, 1); dd++)
                    md::digest<la, lb, lc, ld, true>(bp, kp, sf, outs, D + dd * a.dstride,
                                                     F0 + 2 * (md.base + dd * a.dstride), md.nao, dfac[2 * dd],
                                                     dfac[2 * dd + 1]);
            } else if (a.ndens > 1) {
                for (int dd = 0; dd < a.ndens; dd++)
                    md::digest<la, lb, lc, ld>(bp, kp, sf, outs, D + dd * a.dstride, F + dd * a.dstride, md.nao,
                                               dfac[2 * dd], dfac[2 * dd + 1]);
            } else {
                md::digest<la, lb, lc, ld>(bp, kp, sf, outs, D, F, md.nao, a.fj, a.fk);
            }
            __syncthreads();
        }
    }
}


template <int la, int lb, int lc, int ld>
__global__ void __launch_bounds__(VLX_WAVE) eri_out_kernel_md(const EriArgs a)
{
    using S = md::Sizes<la, lb, lc, ld>;
    __shared__ double tab_lds[S::tab_in_lds ? BOYS_TAB_SIZE : 1];
    __shared__ double work[S::lds];
    __shared__ double out[S::nsab * S::nscd];
    const double* tab = S::tab_in_lds ? tab_lds : a.boys_tab;
    if constexpr (S::tab_in_lds) stage_boys(tab_lds, a.boys_tab);
    for (int qi = blockIdx.x; qi < a.nquartets; qi += gridDim.x) {
        const int2 qq = a.quartets[qi];
        const PairInfo bp = a.bra[qq.x];
        const PairInfo kp = a.ket[qq.y];
        const double abv[3] = {bp.abx, bp.aby, bp.abz}, cdv[3] = {kp.abx, kp.aby, kp.abz};
        md::quartet<la, lb, lc, ld>(a.bra_prims + bp.prim_start, bp.nprim, abv, a.ket_prims + kp.prim_start,
                                    kp.nprim, cdv, tab, 0.0, work, out);
        for (int k = threadIdx.x; k < S::nsab * S::nscd; k += blockDim.x)
            a.out[(size_t)qi * S::nsab * S::nscd + k] = out[k];
        __syncthreads();
    }
}

template <int la, int lb>
__global__ void __launch_bounds__(VLX_WAVE) diag_kernel_md(const DiagArgs a)
{
    using S = md::Sizes<la, lb, la, lb>;
    __shared__ double tab_lds[S::tab_in_lds ? BOYS_TAB_SIZE : 1];
    __shared__ double work[S::lds];
    __shared__ double out[S::nsab * S::nscd];
    const double* tab = S::tab_in_lds ? tab_lds : a.boys_tab;
    if constexpr (S::tab_in_lds) stage_boys(tab_lds, a.boys_tab);
    for (int pi = blockIdx.x; pi < a.npairs; pi += gridDim.x) {
        const PairInfo pp = a.pairs[pi];
        const double abv[3] = {pp.abx, pp.aby, pp.abz};
        md::quartet<la, lb, la, lb>(a.prims + pp.prim_start, pp.nprim, abv, a.prims + pp.prim_start, pp.nprim, abv,
                                    tab, 0.0, work, out);
        double mx = 0.0;
        for (int k = threadIdx.x; k < S::nsab * S::nscd; k += blockDim.x) mx = (mx < out[k]) ? out[k] : mx;
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(mx, off, VLX_WAVE);
            mx = (mx < o) ? o : mx;
        }
        if (threadIdx.x == 0) a.maxval[pi] = mx;
        __syncthreads();
    }
}

// Fock: the matrix-core kernel (U arrays of both pair sets in FockArgs); the
// cooperative LDS kernel serves explicit quartets and the screener.
#define VLX_MMD_LAUNCH(la, lb, lc, ld, RS)                                                              \
    hipLaunchKernelGGL((fock_kernel_mmd<la, lb, lc, ld, RS>), dim3(g), dim3(VLX_WAVE), 0, s, a.bra,     \
                       a.bra_prims, a.bra_u, a.ket, a.ket_prims, a.ket_u, a.items, a.boys_tab, a.D, a.F, \
                       a.mols, a.glim, a.dfac, a.sc)
// launch_mmd_: the matrix-core kernel as an alternative variant of a class
// whose default kernel is generated (same MD work items).
#define VLX_INSTANTIATE_MMD(la, lb, lc, ld)                                                             \
    void launch_mmd_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s)                        \
    {                                                                                                   \
        const int g = grid > 0 ? grid : static_cast<int>(std::min<long long>(a.sc.nitems, 16384));       \
        if (a.sc.omega2 > 0.0) VLX_MMD_LAUNCH(la, lb, lc, ld, true);                                    \
        else VLX_MMD_LAUNCH(la, lb, lc, ld, false);                                                     \
    }
#define VLX_INSTANTIATE_MD_CLASS(la, lb, lc, ld)                                                        \
    void launch_fock_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s)                       \
    {                                                                                                   \
        const int g = grid > 0 ? grid : static_cast<int>(std::min<long long>(a.sc.nitems, 16384));       \
        if (a.sc.omega2 > 0.0) VLX_MMD_LAUNCH(la, lb, lc, ld, true);                                    \
        else VLX_MMD_LAUNCH(la, lb, lc, ld, false);                                                     \
    }                                                                                                   \
    void launch_eri_##la##lb##lc##ld(const EriArgs& a, hipStream_t s)                                   \
    {                                                                                                   \
        const int g = std::min(a.nquartets, 4096);                                                      \
        hipLaunchKernelGGL((eri_out_kernel_md<la, lb, lc, ld>), dim3(g), dim3(VLX_WAVE), 0, s, a);      \
    }

#define VLX_INSTANTIATE_MD_DIAG(la, lb)                                                                 \
    void launch_diag_##la##lb(const DiagArgs& a, hipStream_t s)                                         \
    {                                                                                                   \
        const int g = std::min(a.npairs, 4096);                                                         \
        hipLaunchKernelGGL((diag_kernel_md<la, lb>), dim3(g), dim3(VLX_WAVE), 0, s, a);                 \
    }
