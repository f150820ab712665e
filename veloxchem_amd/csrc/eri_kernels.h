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
//
// Ket-stationary digestion.  Lane j owns ket pair (c, d) for a whole work
// item and walks the item's bra pairs (a, b) in wave-uniform order.  With
// G = F' + F'^T and the shell-level symmetry factor s (1/2 for a == b, for
// c == d and for the diagonal pair-pair quartet), each quartet adds
//   F'_ab += fj s (ab|cd) D_cd      (wave-reduced, one atomic per component)
//   F'_cd += fj s (ab|cd) D_ab      (register accumulator, flushed per item)
//   F'_ac += fk s (ab|cd) D_bd,  F'_ad += fk s (ab|cd) D_bc
//                                   (register accumulators while shell a is
//                                    unchanged; bras are sorted by a)
//   F'_bc += fk s (ab|cd) D_ad,  F'_bd += fk s (ab|cd) D_ac   (atomics)
// which, summed over the unique quartets, gives the reference's digestion
// (src/t4c_func/T4CLocalDistributor.cpp:130-280 and siblings).

template <int na, int nc>
__device__ __forceinline__ void flush_block(double* __restrict__ F, const int n, const int* __restrict__ row,
                                            const int* __restrict__ col, double (&acc)[na][nc], const double f)
{
#pragma unroll
    for (int i = 0; i < na; i++)
#pragma unroll
        for (int k = 0; k < nc; k++) {
            unsafeAtomicAdd(&F[(size_t)row[i] * n + col[k]], f * acc[i][k]);
            acc[i][k] = 0.0;
        }
}

// Pointers are separate __restrict__ kernel arguments so that wave-uniform
// loads (work items, bra pairs and primitives, D_ab) become scalar loads.
template <int la, int lb, int lc, int ld>
__global__ void __launch_bounds__(VLX_BLOCK)
    fock_kernel(const PairInfo* __restrict__ bra, const PrimPair* __restrict__ bra_prims,
                const PairInfo* __restrict__ ket, const PrimPair* __restrict__ ket_prims,
                const WorkItem* __restrict__ items, const double* __restrict__ boys_tab,
                const double* __restrict__ D, double* __restrict__ F, const FockScalars a)
{
    constexpr int na = 2 * la + 1, nb = 2 * lb + 1, nc = 2 * lc + 1, nd = 2 * ld + 1;
    constexpr int nsph = na * nb * nc * nd;
    __shared__ double tab[BOYS_TAB_SIZE];
    stage_boys(tab, boys_tab);

    const int n = a.nao;
    const bool do_j = a.fj != 0.0, do_k = a.fk != 0.0;
    const int lane = threadIdx.x & (VLX_WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / VLX_WAVE);
    const long long gw = (long long)blockIdx.x * VLX_WAVES_PER_BLOCK + wave;
    const long long nw = (long long)gridDim.x * VLX_WAVES_PER_BLOCK;
    const long long w0 = a.nitems * gw / nw, w1 = a.nitems * (gw + 1) / nw;

    for (long long it = w0; it < w1; it++) {
        const WorkItem wi = items[it];
        const int j = wi.j0 + lane;
        const bool lane_ok = j < a.nket;
        const PairInfo kp = ket[lane_ok ? j : wi.j0];
        int r[nc], t[nd];
#pragma unroll
        for (int k = 0; k < nc; k++) r[k] = kp.ao_a + k * kp.str_a;
#pragma unroll
        for (int l = 0; l < nd; l++) t[l] = kp.ao_b + l * kp.str_b;
        double dcd[nc][nd], jcd[nc][nd];
#pragma unroll
        for (int k = 0; k < nc; k++)
#pragma unroll
            for (int l = 0; l < nd; l++) {
                dcd[k][l] = D[(size_t)r[k] * n + t[l]];
                jcd[k][l] = 0.0;
            }
        double kac[na][nc], kad[na][nd];
#pragma unroll
        for (int i = 0; i < na; i++) {
#pragma unroll
            for (int k = 0; k < nc; k++) kac[i][k] = 0.0;
#pragma unroll
            for (int l = 0; l < nd; l++) kad[i][l] = 0.0;
        }
        int p[na];
        int cur_a = -1;
        bool pending = false;  // this lane holds kac/kad contributions for cur_a

        for (int ib = wi.i0; ib < wi.i1; ib++) {
            const PairInfo bp = bra[ib];
            if (bp.shell_a != cur_a) {
                if (do_k && pending) {
                    flush_block<na, nc>(F, n, p, r, kac, a.fk);
                    flush_block<na, nd>(F, n, p, t, kad, a.fk);
                }
                pending = false;
                cur_a = bp.shell_a;
#pragma unroll
                for (int i = 0; i < na; i++) p[i] = bp.ao_a + i * bp.str_a;
            }
            bool active = lane_ok;
            if (a.ithreshold >= 0) active = active && (bp.bin + kp.bin <= a.ithreshold);
            if (a.same_class) active = active && (bp.rank >= j);
            double s = active ? 1.0 : 0.0;
            if (bp.same) s *= 0.5;
            if (kp.same) s *= 0.5;
            if (a.same_class && bp.rank == j) s *= 0.5;
            double I[nsph];
            eri_quartet<la, lb, lc, ld>(bra_prims + bp.prim_start, bp.nprim, bp.abx, bp.aby, bp.abz,
                                        ket_prims + kp.prim_start, active ? kp.nprim : 0, kp.abx, kp.aby, kp.abz,
                                        tab, I);
#pragma unroll
            for (int x = 0; x < nsph; x++) I[x] *= s;
            int q[nb];
#pragma unroll
            for (int jb = 0; jb < nb; jb++) q[jb] = bp.ao_b + jb * bp.str_b;
            if (do_j) {
                // bra side: F'_ab, reduced over the wavefront
#pragma unroll
                for (int i = 0; i < na; i++)
#pragma unroll
                    for (int jb = 0; jb < nb; jb++) {
                        double v = 0.0;
#pragma unroll
                        for (int k = 0; k < nc; k++)
#pragma unroll
                            for (int l = 0; l < nd; l++) v = fma(I[((i * nb + jb) * nc + k) * nd + l], dcd[k][l], v);
                        v = wave_sum(v);
                        if (lane == 0 && v != 0.0) unsafeAtomicAdd(&F[(size_t)p[i] * n + q[jb]], a.fj * v);
                    }
                // ket side: F'_cd accumulated in registers
#pragma unroll
                for (int i = 0; i < na; i++)
#pragma unroll
                    for (int jb = 0; jb < nb; jb++) {
                        const double dab = D[(size_t)p[i] * n + q[jb]];
#pragma unroll
                        for (int k = 0; k < nc; k++)
#pragma unroll
                            for (int l = 0; l < nd; l++) jcd[k][l] = fma(I[((i * nb + jb) * nc + k) * nd + l], dab, jcd[k][l]);
                    }
            }
            if (do_k && active) {
                pending = true;
                // F'_ac += I D_bd and F'_ad += I D_bc (sticky while a is unchanged)
#pragma unroll
                for (int jb = 0; jb < nb; jb++) {
                    double dbd[nd], dbc[nc];
#pragma unroll
                    for (int l = 0; l < nd; l++) dbd[l] = D[(size_t)q[jb] * n + t[l]];
#pragma unroll
                    for (int k = 0; k < nc; k++) dbc[k] = D[(size_t)q[jb] * n + r[k]];
#pragma unroll
                    for (int i = 0; i < na; i++)
#pragma unroll
                        for (int k = 0; k < nc; k++)
#pragma unroll
                            for (int l = 0; l < nd; l++) {
                                const double v = I[((i * nb + jb) * nc + k) * nd + l];
                                kac[i][k] = fma(v, dbd[l], kac[i][k]);
                                kad[i][l] = fma(v, dbc[k], kad[i][l]);
                            }
                }
                // F'_bc += I D_ad and F'_bd += I D_ac (b changes every bra)
                double dad[na][nd], dac[na][nc];
#pragma unroll
                for (int i = 0; i < na; i++) {
#pragma unroll
                    for (int l = 0; l < nd; l++) dad[i][l] = D[(size_t)p[i] * n + t[l]];
#pragma unroll
                    for (int k = 0; k < nc; k++) dac[i][k] = D[(size_t)p[i] * n + r[k]];
                }
#pragma unroll
                for (int jb = 0; jb < nb; jb++) {
#pragma unroll
                    for (int k = 0; k < nc; k++) {
                        double v = 0.0;
#pragma unroll
                        for (int i = 0; i < na; i++)
#pragma unroll
                            for (int l = 0; l < nd; l++) v = fma(I[((i * nb + jb) * nc + k) * nd + l], dad[i][l], v);
                        unsafeAtomicAdd(&F[(size_t)q[jb] * n + r[k]], a.fk * v);
                    }
#pragma unroll
                    for (int l = 0; l < nd; l++) {
                        double v = 0.0;
#pragma unroll
                        for (int i = 0; i < na; i++)
#pragma unroll
                            for (int k = 0; k < nc; k++) v = fma(I[((i * nb + jb) * nc + k) * nd + l], dac[i][k], v);
                        unsafeAtomicAdd(&F[(size_t)q[jb] * n + t[l]], a.fk * v);
                    }
                }
            }
        }
        if (do_k && pending) {
            flush_block<na, nc>(F, n, p, r, kac, a.fk);
            flush_block<na, nd>(F, n, p, t, kad, a.fk);
        }
        if (do_j && lane_ok) flush_block<nc, nd>(F, n, r, t, jcd, a.fj);
    }
}

// Bra-stationary digestion, for classes whose integrals need most of the
// register file.  A wavefront takes one bra pair (wave-uniform) and 64 kets
// (one per lane) from the prefix of the ket list its bins allow; work item
// {j0, i0, i1}: kets [j0, j0 + 64) of bra i0, i1 = number of kets of bra i0.
// F'_ab is reduced over the wavefront; F'_cd, F'_ac, F'_ad, F'_bc and F'_bd
// are per-lane atomics.
template <int la, int lb, int lc, int ld>
__global__ void __launch_bounds__(VLX_BLOCK)
    fock_kernel_bs(const PairInfo* __restrict__ bra, const PrimPair* __restrict__ bra_prims,
                   const PairInfo* __restrict__ ket, const PrimPair* __restrict__ ket_prims,
                   const WorkItem* __restrict__ items, const double* __restrict__ boys_tab,
                   const double* __restrict__ D, double* __restrict__ F, const FockScalars a)
{
    constexpr int na = 2 * la + 1, nb = 2 * lb + 1, nc = 2 * lc + 1, nd = 2 * ld + 1;
    constexpr int nsph = na * nb * nc * nd;
    __shared__ double tab[BOYS_TAB_SIZE];
    stage_boys(tab, boys_tab);

    const int n = a.nao;
    const int lane = threadIdx.x & (VLX_WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / VLX_WAVE);
    const long long gw = (long long)blockIdx.x * VLX_WAVES_PER_BLOCK + wave;
    const long long nw = (long long)gridDim.x * VLX_WAVES_PER_BLOCK;
    const long long w0 = a.nitems * gw / nw, w1 = a.nitems * (gw + 1) / nw;

    for (long long it = w0; it < w1; it++) {
        const WorkItem wi = items[it];
        const PairInfo bp = bra[wi.i0];
        const int j = wi.j0 + lane;
        const bool active = j < wi.i1;
        const PairInfo kp = ket[active ? j : wi.j0];
        double s = active ? 1.0 : 0.0;
        if (bp.same) s *= 0.5;
        if (kp.same) s *= 0.5;
        if (a.same_class && bp.rank == j) s *= 0.5;
        double I[nsph];
        eri_quartet<la, lb, lc, ld>(bra_prims + bp.prim_start, bp.nprim, bp.abx, bp.aby, bp.abz,
                                    ket_prims + kp.prim_start, active ? kp.nprim : 0, kp.abx, kp.aby, kp.abz, tab, I);
#pragma unroll
        for (int x = 0; x < nsph; x++) I[x] *= s;
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
#pragma unroll
            for (int i = 0; i < na; i++)
#pragma unroll
                for (int jb = 0; jb < nb; jb++) {
                    double v = 0.0;
#pragma unroll
                    for (int k = 0; k < nc; k++)
#pragma unroll
                        for (int l = 0; l < nd; l++)
                            v = fma(I[((i * nb + jb) * nc + k) * nd + l], D[(size_t)r[k] * n + t[l]], v);
                    v = wave_sum(v);
                    if (lane == 0 && v != 0.0) unsafeAtomicAdd(&F[(size_t)p[i] * n + q[jb]], a.fj * v);
                }
            if (active) {
#pragma unroll
                for (int k = 0; k < nc; k++)
#pragma unroll
                    for (int l = 0; l < nd; l++) {
                        double v = 0.0;
#pragma unroll
                        for (int i = 0; i < na; i++)
#pragma unroll
                            for (int jb = 0; jb < nb; jb++)
                                v = fma(I[((i * nb + jb) * nc + k) * nd + l], D[(size_t)p[i] * n + q[jb]], v);
                        unsafeAtomicAdd(&F[(size_t)r[k] * n + t[l]], a.fj * v);
                    }
            }
        }
        if (a.fk != 0.0 && active) {
#pragma unroll
            for (int i = 0; i < na; i++)
#pragma unroll
                for (int k = 0; k < nc; k++) {
                    double v = 0.0;
#pragma unroll
                    for (int jb = 0; jb < nb; jb++)
#pragma unroll
                        for (int l = 0; l < nd; l++)
                            v = fma(I[((i * nb + jb) * nc + k) * nd + l], D[(size_t)q[jb] * n + t[l]], v);
                    unsafeAtomicAdd(&F[(size_t)p[i] * n + r[k]], a.fk * v);
                }
#pragma unroll
            for (int i = 0; i < na; i++)
#pragma unroll
                for (int l = 0; l < nd; l++) {
                    double v = 0.0;
#pragma unroll
                    for (int jb = 0; jb < nb; jb++)
#pragma unroll
                        for (int k = 0; k < nc; k++)
                            v = fma(I[((i * nb + jb) * nc + k) * nd + l], D[(size_t)q[jb] * n + r[k]], v);
                    unsafeAtomicAdd(&F[(size_t)p[i] * n + t[l]], a.fk * v);
                }
#pragma unroll
            for (int jb = 0; jb < nb; jb++)
#pragma unroll
                for (int k = 0; k < nc; k++) {
                    double v = 0.0;
#pragma unroll
                    for (int i = 0; i < na; i++)
#pragma unroll
                        for (int l = 0; l < nd; l++)
                            v = fma(I[((i * nb + jb) * nc + k) * nd + l], D[(size_t)p[i] * n + t[l]], v);
                    unsafeAtomicAdd(&F[(size_t)q[jb] * n + r[k]], a.fk * v);
                }
#pragma unroll
            for (int jb = 0; jb < nb; jb++)
#pragma unroll
                for (int l = 0; l < nd; l++) {
                    double v = 0.0;
#pragma unroll
                    for (int i = 0; i < na; i++)
#pragma unroll
                        for (int k = 0; k < nc; k++)
                            v = fma(I[((i * nb + jb) * nc + k) * nd + l], D[(size_t)p[i] * n + r[k]], v);
                    unsafeAtomicAdd(&F[(size_t)q[jb] * n + t[l]], a.fk * v);
                }
        }
    }
}

// LDS-tiled digestion.  A work-group takes one TileItem: bras of atom groups
// (ga, gb) and up to 64 kets of groups (gc, gd).  Every F' contribution of
// the item lands in six small dense tiles indexed by group-local AOs,
//   J_AB, J_CD, K_AC, K_AD, K_BC, K_BD,
// accumulated with LDS atomics and flushed to HBM once per item.  The four
// wavefronts share the kets and split the bras.
__device__ __forceinline__ void flush_tile(double* __restrict__ F, const int n, const double* __restrict__ tile,
                                           const int nr, const int nc, const int* __restrict__ rows,
                                           const int* __restrict__ cols, const double f)
{
    for (int e = threadIdx.x; e < nr * nc; e += blockDim.x) {
        const double v = tile[e];
        if (v != 0.0) unsafeAtomicAdd(&F[(size_t)rows[e / nc] * n + cols[e % nc]], f * v);
    }
}

template <int la, int lb, int lc, int ld>
__global__ void __launch_bounds__(VLX_BLOCK)
    fock_kernel_tiled(const PairInfo* __restrict__ pairs_bra, const PrimPair* __restrict__ bra_prims,
                      const PairInfo* __restrict__ pairs_ket, const PrimPair* __restrict__ ket_prims,
                      const TileItem* __restrict__ items, const int2* __restrict__ kentries,
                      const double* __restrict__ boys_tab, const double* __restrict__ D, double* __restrict__ F,
                      const int* __restrict__ grp_off, const int* __restrict__ grp_n,
                      const int* __restrict__ grp_ao, const FockScalars a)
{
    constexpr int na = 2 * la + 1, nb = 2 * lb + 1, nc = 2 * lc + 1, nd = 2 * ld + 1;
    constexpr int nsph = na * nb * nc * nd;
    __shared__ double tab[BOYS_TAB_SIZE];
    extern __shared__ double tiles[];
    stage_boys(tab, boys_tab);

    const int n = a.nao;
    const bool do_j = a.fj != 0.0, do_k = a.fk != 0.0;
    const int lane = threadIdx.x & (VLX_WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / VLX_WAVE);

    for (long long it = blockIdx.x; it < a.nitems; it += gridDim.x) {
        const TileItem ti = items[it];
        const int nA = grp_n[ti.ga * 4 + la], nB = grp_n[ti.gb * 4 + lb];
        const int nC = grp_n[ti.gc * 4 + lc], nD = grp_n[ti.gd * 4 + ld];
        double* t_jab = tiles;
        double* t_jcd = t_jab + nA * nB;
        double* t_kac = t_jcd + nC * nD;
        double* t_kad = t_kac + nA * nC;
        double* t_kbc = t_kad + nA * nD;
        double* t_kbd = t_kbc + nB * nC;
        const int ntile = nA * nB + nC * nD + nA * nC + nA * nD + nB * nC + nB * nD;
        for (int e = threadIdx.x; e < ntile; e += blockDim.x) tiles[e] = 0.0;
        __syncthreads();

        // lane = one primitive pair of a ket pair: the contraction over ket
        // primitives happens in the LDS tiles (every step after the VRR is linear)
        const int e = ti.j0 + lane;
        const bool lane_ok = e < ti.j1;
        const int2 ke = kentries[lane_ok ? e : ti.j0];
        const PairInfo kp = pairs_ket[ke.x];
        int r[nc], t[nd], rl[nc], tl[nd];
#pragma unroll
        for (int k = 0; k < nc; k++) {
            r[k] = kp.ao_a + k * kp.str_a;
            rl[k] = kp.loc_a + k;
        }
#pragma unroll
        for (int l = 0; l < nd; l++) {
            t[l] = kp.ao_b + l * kp.str_b;
            tl[l] = kp.loc_b + l;
        }
        double dcd[nc][nd];
#pragma unroll
        for (int k = 0; k < nc; k++)
#pragma unroll
            for (int l = 0; l < nd; l++) dcd[k][l] = D[(size_t)r[k] * n + t[l]];

        for (int ib = ti.i0 + wave; ib < ti.i1; ib += VLX_WAVES_PER_BLOCK) {
            const PairInfo bp = pairs_bra[ib];
            bool active = lane_ok;
            if (a.ithreshold >= 0) active = active && (bp.bin + kp.bin <= a.ithreshold);
            if (a.same_class) active = active && (bp.rank >= kp.rank);
            double s = active ? 1.0 : 0.0;
            if (bp.same) s *= 0.5;
            if (kp.same) s *= 0.5;
            if (a.same_class && bp.rank == kp.rank) s *= 0.5;
            double I[nsph];
            eri_quartet<la, lb, lc, ld>(bra_prims + bp.prim_start, bp.nprim, bp.abx, bp.aby, bp.abz,
                                        ket_prims + kp.prim_start + ke.y, active ? 1 : 0, kp.abx, kp.aby, kp.abz,
                                        tab, I);
#pragma unroll
            for (int x = 0; x < nsph; x++) I[x] *= s;
            int p[na], q[nb];
#pragma unroll
            for (int i = 0; i < na; i++) p[i] = bp.ao_a + i * bp.str_a;
#pragma unroll
            for (int jb = 0; jb < nb; jb++) q[jb] = bp.ao_b + jb * bp.str_b;
            if (do_j) {
#pragma unroll
                for (int i = 0; i < na; i++)
#pragma unroll
                    for (int jb = 0; jb < nb; jb++) {
                        double v = 0.0;
#pragma unroll
                        for (int k = 0; k < nc; k++)
#pragma unroll
                            for (int l = 0; l < nd; l++) v = fma(I[((i * nb + jb) * nc + k) * nd + l], dcd[k][l], v);
                        v = wave_sum(v);
                        if (lane == 0 && v != 0.0) unsafeAtomicAdd(&t_jab[(bp.loc_a + i) * nB + bp.loc_b + jb], v);
                    }
                if (active) {
                    double vcd[nc][nd];
#pragma unroll
                    for (int k = 0; k < nc; k++)
#pragma unroll
                        for (int l = 0; l < nd; l++) vcd[k][l] = 0.0;
#pragma unroll
                    for (int i = 0; i < na; i++)
#pragma unroll
                        for (int jb = 0; jb < nb; jb++) {
                            const double dab = D[(size_t)p[i] * n + q[jb]];
#pragma unroll
                            for (int k = 0; k < nc; k++)
#pragma unroll
                                for (int l = 0; l < nd; l++)
                                    vcd[k][l] = fma(I[((i * nb + jb) * nc + k) * nd + l], dab, vcd[k][l]);
                        }
#pragma unroll
                    for (int k = 0; k < nc; k++)
#pragma unroll
                        for (int l = 0; l < nd; l++) unsafeAtomicAdd(&t_jcd[rl[k] * nD + tl[l]], vcd[k][l]);
                }
            }
            if (do_k && active) {
#pragma unroll
                for (int i = 0; i < na; i++) {
#pragma unroll
                    for (int k = 0; k < nc; k++) {  // K_ac += I D_bd
                        double v = 0.0;
#pragma unroll
                        for (int jb = 0; jb < nb; jb++)
#pragma unroll
                            for (int l = 0; l < nd; l++)
                                v = fma(I[((i * nb + jb) * nc + k) * nd + l], D[(size_t)q[jb] * n + t[l]], v);
                        unsafeAtomicAdd(&t_kac[(bp.loc_a + i) * nC + rl[k]], v);
                    }
#pragma unroll
                    for (int l = 0; l < nd; l++) {  // K_ad += I D_bc
                        double v = 0.0;
#pragma unroll
                        for (int jb = 0; jb < nb; jb++)
#pragma unroll
                            for (int k = 0; k < nc; k++)
                                v = fma(I[((i * nb + jb) * nc + k) * nd + l], D[(size_t)q[jb] * n + r[k]], v);
                        unsafeAtomicAdd(&t_kad[(bp.loc_a + i) * nD + tl[l]], v);
                    }
                }
#pragma unroll
                for (int jb = 0; jb < nb; jb++) {
#pragma unroll
                    for (int k = 0; k < nc; k++) {  // K_bc += I D_ad
                        double v = 0.0;
#pragma unroll
                        for (int i = 0; i < na; i++)
#pragma unroll
                            for (int l = 0; l < nd; l++)
                                v = fma(I[((i * nb + jb) * nc + k) * nd + l], D[(size_t)p[i] * n + t[l]], v);
                        unsafeAtomicAdd(&t_kbc[(bp.loc_b + jb) * nC + rl[k]], v);
                    }
#pragma unroll
                    for (int l = 0; l < nd; l++) {  // K_bd += I D_ac
                        double v = 0.0;
#pragma unroll
                        for (int i = 0; i < na; i++)
#pragma unroll
                            for (int k = 0; k < nc; k++)
                                v = fma(I[((i * nb + jb) * nc + k) * nd + l], D[(size_t)p[i] * n + r[k]], v);
                        unsafeAtomicAdd(&t_kbd[(bp.loc_b + jb) * nD + tl[l]], v);
                    }
                }
            }
        }
        __syncthreads();
        const int* ra = grp_ao + grp_off[ti.ga * 4 + la];
        const int* rb = grp_ao + grp_off[ti.gb * 4 + lb];
        const int* rc = grp_ao + grp_off[ti.gc * 4 + lc];
        const int* rd = grp_ao + grp_off[ti.gd * 4 + ld];
        if (do_j) {
            flush_tile(F, n, t_jab, nA, nB, ra, rb, a.fj);
            flush_tile(F, n, t_jcd, nC, nD, rc, rd, a.fj);
        }
        if (do_k) {
            flush_tile(F, n, t_kac, nA, nC, ra, rc, a.fk);
            flush_tile(F, n, t_kad, nA, nD, ra, rd, a.fk);
            flush_tile(F, n, t_kbc, nB, nC, rb, rc, a.fk);
            flush_tile(F, n, t_kbd, nB, nD, rb, rd, a.fk);
        }
        __syncthreads();
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

// kernel: fock_kernel (ket-stationary) or fock_kernel_bs (bra-stationary)
#define VLX_INSTANTIATE_CLASS(la, lb, lc, ld, kernel)                                                   \
    void launch_fock_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s)                       \
    {                                                                                                   \
        hipLaunchKernelGGL((kernel<la, lb, lc, ld>), dim3(grid), dim3(VLX_BLOCK), 0, s, a.bra,          \
                           a.bra_prims, a.ket, a.ket_prims, a.items, a.boys_tab, a.D, a.F, a.sc);       \
    }                                                                                                   \
    void launch_eri_##la##lb##lc##ld(const EriArgs& a, hipStream_t s)                                   \
    {                                                                                                   \
        const int grid = (a.nquartets + VLX_BLOCK - 1) / VLX_BLOCK;                                     \
        hipLaunchKernelGGL((eri_out_kernel<la, lb, lc, ld>), dim3(grid), dim3(VLX_BLOCK), 0, s, a);     \
    }

// LDS-tiled launcher (dynamic LDS = a.tile_lds_bytes); grid-stride over items
#define VLX_INSTANTIATE_TILED(la, lb, lc, ld)                                                           \
    void launch_tiled_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s)                      \
    {                                                                                                   \
        hipLaunchKernelGGL((fock_kernel_tiled<la, lb, lc, ld>), dim3(grid), dim3(VLX_BLOCK),            \
                           a.tile_lds_bytes, s, a.bra, a.bra_prims, a.ket, a.ket_prims, a.titems,        \
                           a.kentries, a.boys_tab, a.D, a.F, a.grp_off, a.grp_n, a.grp_ao, a.sc);       \
    }

#define VLX_INSTANTIATE_DIAG(la, lb)                                                                    \
    void launch_diag_##la##lb(const DiagArgs& a, hipStream_t s)                                         \
    {                                                                                                   \
        const int grid = (a.npairs + VLX_BLOCK - 1) / VLX_BLOCK;                                        \
        hipLaunchKernelGGL((diag_kernel<la, lb>), dim3(grid), dim3(VLX_BLOCK), 0, s, a);                \
    }
