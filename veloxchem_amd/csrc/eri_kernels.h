// Device-side building blocks of the MI355X (gfx950) two-electron Fock build.
//
// One translation unit per shell-quartet class includes this header, supplies
// the generated eri_quartet<la,lb,lc,ld> specialisation and instantiates the
// kernels below through VLX_INSTANTIATE_CLASS / VLX_INSTANTIATE_DIAG.
//
// Work decomposition (Fock kernel): a wavefront owns one bra shell pair and
// 64 ket shell pairs (one per lane).  Bra data are wave-uniform (scalar
// loads), ket data are per lane.  The quartet set per bra pair i is the prefix
// of the bin-sorted ket list allowed by bin_i + bin_j <= ithreshold (and
// j <= i inside one pair class), i.e. the reference's Schwarz bin rule
// (src/general/OpenMPFunc.cpp:121-259), so no quartet list is materialised.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "vlx_types.h"

constexpr int VLX_WAVE = 64;
constexpr int VLX_BLOCK = 256;
constexpr int VLX_WAVES_PER_BLOCK = VLX_BLOCK / VLX_WAVE;

template <int la, int lb, int lc, int ld>
__device__ __forceinline__ void eri_quartet(const PrimPair* __restrict__ bra_prims, const int bra_np, const double ABx,
                                            const double ABy, const double ABz, const PrimPair* __restrict__ ket_prims,
                                            const int ket_np, const double CDx, const double CDy, const double CDz,
                                            const double* __restrict__ boys_tab, double* __restrict__ out);

// Boys function F_0..F_L(T).  T < BOYS_TMAX: F_L from the order-BOYS_ORDER
// Taylor expansion around the nearest grid point (table row k holds
// F_{L+j}(T_k) (-1)^j / j!), then downward recursion
// F_m = (2T F_{m+1} + e^-T) / (2m+1).  T >= BOYS_TMAX: F_0 = sqrt(pi/T)/2
// (erf(sqrt(T)) = 1 to double precision there) and upward recursion
// F_{m+1} = ((2m+1) F_m - e^-T) / (2T), which is stable for T > L.
template <int L>
__device__ __forceinline__ void boys_eval(const double T, const double* __restrict__ tab, double* __restrict__ F)
{
    const double et = exp(-T);
    if (T < BOYS_TMAX) {
        const int k = static_cast<int>(T * BOYS_INV_DT + 0.5);
        const double d = T - BOYS_DT * k;
        const double* c = tab + k * (BOYS_ORDER + 1);
        double f = c[BOYS_ORDER];
#pragma unroll
        for (int j = BOYS_ORDER - 1; j >= 0; j--) f = fma(f, d, c[j]);
        F[L] = f;
#pragma unroll
        for (int m = L - 1; m >= 0; m--) F[m] = (2.0 * T * F[m + 1] + et) * (1.0 / (2 * m + 1));
    } else {
        const double it = 1.0 / T;
        F[0] = 0.5 * sqrt(3.14159265358979323846 * it);
#pragma unroll
        for (int m = 0; m < L; m++) F[m + 1] = ((2 * m + 1) * F[m] - et) * (0.5 * it);
    }
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, VLX_WAVE);
    return v;
}

__device__ __forceinline__ void stage_boys(double* lds, const double* __restrict__ tab)
{
    for (int t = threadIdx.x; t < BOYS_TAB_SIZE; t += blockDim.x) lds[t] = tab[t];
    __syncthreads();
}

// ---------------------------------------------------------------- Fock

template <int la, int lb, int lc, int ld>
__device__ __forceinline__ void digest_quartet(const FockArgs& a, const PairInfo& bp, const PairInfo& kp, const double s,
                                               const bool active, const double* __restrict__ I)
{
    constexpr int na = 2 * la + 1, nb = 2 * lb + 1, nc = 2 * lc + 1, nd = 2 * ld + 1;
    const int n = a.nao;
    const double* __restrict__ D = a.D;
    double* __restrict__ F = a.F;
    int p[na], q[nb], r[nc], t[nd];
#pragma unroll
    for (int i = 0; i < na; i++) p[i] = bp.ao_a + i * bp.str_a;
#pragma unroll
    for (int i = 0; i < nb; i++) q[i] = bp.ao_b + i * bp.str_b;
#pragma unroll
    for (int i = 0; i < nc; i++) r[i] = kp.ao_a + i * kp.str_a;
#pragma unroll
    for (int i = 0; i < nd; i++) t[i] = kp.ao_b + i * kp.str_b;

    if (a.fj != 0.0) {
        const double fj = a.fj * s;
        double dcd[nc][nd];
#pragma unroll
        for (int k = 0; k < nc; k++)
#pragma unroll
            for (int l = 0; l < nd; l++) dcd[k][l] = D[(size_t)r[k] * n + t[l]];
        // bra side F_pq += fj s sum_rs I D_rs: reduce over the wavefront first
#pragma unroll
        for (int i = 0; i < na; i++)
#pragma unroll
            for (int j = 0; j < nb; j++) {
                double v = 0.0;
#pragma unroll
                for (int k = 0; k < nc; k++)
#pragma unroll
                    for (int l = 0; l < nd; l++) v = fma(I[((i * nb + j) * nc + k) * nd + l], dcd[k][l], v);
                v = wave_sum(v * fj);
                if ((threadIdx.x & (VLX_WAVE - 1)) == 0) unsafeAtomicAdd(&F[(size_t)p[i] * n + q[j]], v);
            }
        // ket side F_rs += fj s sum_pq I D_pq
        double dab[na][nb];
#pragma unroll
        for (int i = 0; i < na; i++)
#pragma unroll
            for (int j = 0; j < nb; j++) dab[i][j] = D[(size_t)p[i] * n + q[j]];
        if (active) {
#pragma unroll
            for (int k = 0; k < nc; k++)
#pragma unroll
                for (int l = 0; l < nd; l++) {
                    double v = 0.0;
#pragma unroll
                    for (int i = 0; i < na; i++)
#pragma unroll
                        for (int j = 0; j < nb; j++) v = fma(I[((i * nb + j) * nc + k) * nd + l], dab[i][j], v);
                    unsafeAtomicAdd(&F[(size_t)r[k] * n + t[l]], fj * v);
                }
        }
    }
    if (a.fk != 0.0 && active) {
        const double fk = a.fk * s;
        // F_pr += fk s sum_qs I D_qs
        {
            double dqs[nb][nd];
#pragma unroll
            for (int j = 0; j < nb; j++)
#pragma unroll
                for (int l = 0; l < nd; l++) dqs[j][l] = D[(size_t)q[j] * n + t[l]];
#pragma unroll
            for (int i = 0; i < na; i++)
#pragma unroll
                for (int k = 0; k < nc; k++) {
                    double v = 0.0;
#pragma unroll
                    for (int j = 0; j < nb; j++)
#pragma unroll
                        for (int l = 0; l < nd; l++) v = fma(I[((i * nb + j) * nc + k) * nd + l], dqs[j][l], v);
                    unsafeAtomicAdd(&F[(size_t)p[i] * n + r[k]], fk * v);
                }
        }
        // F_ps += fk s sum_qr I D_qr
        {
            double dqr[nb][nc];
#pragma unroll
            for (int j = 0; j < nb; j++)
#pragma unroll
                for (int k = 0; k < nc; k++) dqr[j][k] = D[(size_t)q[j] * n + r[k]];
#pragma unroll
            for (int i = 0; i < na; i++)
#pragma unroll
                for (int l = 0; l < nd; l++) {
                    double v = 0.0;
#pragma unroll
                    for (int j = 0; j < nb; j++)
#pragma unroll
                        for (int k = 0; k < nc; k++) v = fma(I[((i * nb + j) * nc + k) * nd + l], dqr[j][k], v);
                    unsafeAtomicAdd(&F[(size_t)p[i] * n + t[l]], fk * v);
                }
        }
        // F_qr += fk s sum_ps I D_ps
        {
            double dps[na][nd];
#pragma unroll
            for (int i = 0; i < na; i++)
#pragma unroll
                for (int l = 0; l < nd; l++) dps[i][l] = D[(size_t)p[i] * n + t[l]];
#pragma unroll
            for (int j = 0; j < nb; j++)
#pragma unroll
                for (int k = 0; k < nc; k++) {
                    double v = 0.0;
#pragma unroll
                    for (int i = 0; i < na; i++)
#pragma unroll
                        for (int l = 0; l < nd; l++) v = fma(I[((i * nb + j) * nc + k) * nd + l], dps[i][l], v);
                    unsafeAtomicAdd(&F[(size_t)q[j] * n + r[k]], fk * v);
                }
        }
        // F_qs += fk s sum_pr I D_pr
        {
            double dpr[na][nc];
#pragma unroll
            for (int i = 0; i < na; i++)
#pragma unroll
                for (int k = 0; k < nc; k++) dpr[i][k] = D[(size_t)p[i] * n + r[k]];
#pragma unroll
            for (int j = 0; j < nb; j++)
#pragma unroll
                for (int l = 0; l < nd; l++) {
                    double v = 0.0;
#pragma unroll
                    for (int i = 0; i < na; i++)
#pragma unroll
                        for (int k = 0; k < nc; k++) v = fma(I[((i * nb + j) * nc + k) * nd + l], dpr[i][k], v);
                    unsafeAtomicAdd(&F[(size_t)q[j] * n + t[l]], fk * v);
                }
        }
    }
}

// Each wavefront takes a contiguous range of 64-ket tiles; a tile is
// (bra pair i, kets [64 w, 64 w + 64)).  tile_prefix[i] is the first tile of
// bra pair i and nket[i] the length of its ket prefix.
template <int la, int lb, int lc, int ld>
__global__ void __launch_bounds__(VLX_BLOCK) fock_kernel(const FockArgs a)
{
    constexpr int nsph = (2 * la + 1) * (2 * lb + 1) * (2 * lc + 1) * (2 * ld + 1);
    __shared__ double tab[BOYS_TAB_SIZE];
    stage_boys(tab, a.boys_tab);

    const int lane = threadIdx.x & (VLX_WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / VLX_WAVE);
    const long gw = (long)blockIdx.x * VLX_WAVES_PER_BLOCK + wave;
    const long nw = (long)gridDim.x * VLX_WAVES_PER_BLOCK;
    const long long t0 = a.total_tiles * gw / nw;
    const long long t1 = a.total_tiles * (gw + 1) / nw;
    if (t0 >= t1) return;

    int lo = 0, hi = a.nbra;  // tile_prefix[lo] <= t0 < tile_prefix[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.tile_prefix[mid] <= t0) lo = mid;
        else hi = mid;
    }
    int i = lo;
    for (long long tile = t0; tile < t1; tile++) {
        while (a.tile_prefix[i + 1] <= tile) i++;
        const PairInfo bp = a.bra[i];
        const int nk = a.nket[i];
        const int j = static_cast<int>(tile - a.tile_prefix[i]) * VLX_WAVE + lane;
        const bool active = j < nk;
        const PairInfo kp = a.ket[active ? j : 0];
        double I[nsph];
        eri_quartet<la, lb, lc, ld>(a.bra_prims + bp.prim_start, bp.nprim, bp.abx, bp.aby, bp.abz,
                                    a.ket_prims + kp.prim_start, active ? kp.nprim : 0, kp.abx, kp.aby, kp.abz, tab, I);
        double s = active ? 1.0 : 0.0;
        if (bp.same) s *= 0.5;
        if (kp.same) s *= 0.5;
        if (a.same_class && i == j) s *= 0.5;
        digest_quartet<la, lb, lc, ld>(a, bp, kp, s, active, I);
    }
}

// ------------------------------------------------- explicit quartet lists

template <int la, int lb, int lc, int ld>
__global__ void __launch_bounds__(VLX_BLOCK) eri_out_kernel(const EriArgs a)
{
    constexpr int nsph = (2 * la + 1) * (2 * lb + 1) * (2 * lc + 1) * (2 * ld + 1);
    __shared__ double tab[BOYS_TAB_SIZE];
    stage_boys(tab, a.boys_tab);
    const int qi = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = qi < a.nquartets;
    const int2 qq = a.quartets[active ? qi : 0];
    const PairInfo bp = a.bra[qq.x];
    const PairInfo kp = a.ket[qq.y];
    double I[nsph];
    eri_quartet<la, lb, lc, ld>(a.bra_prims + bp.prim_start, bp.nprim, bp.abx, bp.aby, bp.abz,
                                a.ket_prims + kp.prim_start, active ? kp.nprim : 0, kp.abx, kp.aby, kp.abz, tab, I);
    if (active) {
        double* dst = a.out + (size_t)qi * nsph;
#pragma unroll
        for (int k = 0; k < nsph; k++) dst[k] = I[k];
    }
}

// (ab|ab) of every pair: maximum over all spherical components, starting from
// zero, as the reference screener takes it (src/t4c_func/T4CUtils.cpp:513-519).
template <int la, int lb>
__global__ void __launch_bounds__(VLX_BLOCK) diag_kernel(const DiagArgs a)
{
    constexpr int nsph = (2 * la + 1) * (2 * lb + 1) * (2 * la + 1) * (2 * lb + 1);
    __shared__ double tab[BOYS_TAB_SIZE];
    stage_boys(tab, a.boys_tab);
    const int pi = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = pi < a.npairs;
    const PairInfo pp = a.pairs[active ? pi : 0];
    double I[nsph];
    eri_quartet<la, lb, la, lb>(a.prims + pp.prim_start, pp.nprim, pp.abx, pp.aby, pp.abz, a.prims + pp.prim_start,
                                active ? pp.nprim : 0, pp.abx, pp.aby, pp.abz, tab, I);
    double mx = 0.0;
#pragma unroll
    for (int k = 0; k < nsph; k++) mx = (mx < I[k]) ? I[k] : mx;
    if (active) a.maxval[pi] = mx;
}

#define VLX_INSTANTIATE_CLASS(la, lb, lc, ld)                                                           \
    void launch_fock_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s)                       \
    {                                                                                                   \
        hipLaunchKernelGGL((fock_kernel<la, lb, lc, ld>), dim3(grid), dim3(VLX_BLOCK), 0, s, a);        \
    }                                                                                                   \
    void launch_eri_##la##lb##lc##ld(const EriArgs& a, hipStream_t s)                                   \
    {                                                                                                   \
        const int grid = (a.nquartets + VLX_BLOCK - 1) / VLX_BLOCK;                                     \
        hipLaunchKernelGGL((eri_out_kernel<la, lb, lc, ld>), dim3(grid), dim3(VLX_BLOCK), 0, s, a);     \
    }

#define VLX_INSTANTIATE_DIAG(la, lb)                                                                    \
    void launch_diag_##la##lb(const DiagArgs& a, hipStream_t s)                                         \
    {                                                                                                   \
        const int grid = (a.npairs + VLX_BLOCK - 1) / VLX_BLOCK;                                        \
        hipLaunchKernelGGL((diag_kernel<la, lb>), dim3(grid), dim3(VLX_BLOCK), 0, s, a);                \
    }
