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

// Arithmetic helpers of the generated recursions, per real type.  The FP64
// forms are plain IEEE operations; the FP32 forms use the hardware
// reciprocal and exponential (1 ulp), which the mixed-precision path accepts
// for quartets whose Schwarz bound is small.
template <typename R> __device__ __forceinline__ R vdiv(R a, R b);
template <> __device__ __forceinline__ double vdiv<double>(double a, double b) { return a / b; }
template <> __device__ __forceinline__ float vdiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <typename R> __device__ __forceinline__ R vsqrt(R x);
template <> __device__ __forceinline__ double vsqrt<double>(double x) { return sqrt(x); }
template <> __device__ __forceinline__ float vsqrt<float>(float x) { return __builtin_sqrtf(x); }
template <typename R> __device__ __forceinline__ R vexp(R x);
template <> __device__ __forceinline__ double vexp<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float vexp<float>(float x) { return __expf(x); }
// 1/x and 1/sqrt(x) for the positive, normal arguments of the set-up
// (zeta, eta, zeta + eta, T >= BOYS_TMAX): hardware estimate refined by two
// Newton steps in FP64 (each step squares the relative error), the estimate
// alone in FP32.
template <typename R> __device__ __forceinline__ R vrcp(R x);
template <> __device__ __forceinline__ double vrcp<double>(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
template <> __device__ __forceinline__ float vrcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <typename R> __device__ __forceinline__ R vrsqrt(R x);
template <> __device__ __forceinline__ double vrsqrt<double>(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    double e = fma(-h * y, y, 0.5);
    y = fma(y, e, y);
    e = fma(-h * y, y, 0.5);
    return fma(y, e, y);
}
template <> __device__ __forceinline__ float vrsqrt<float>(float x) { return __builtin_amdgcn_rsqf(x); }

// Per-class primitive-quartet recursion, generated (one specialisation per
// class): EriClass<la,lb,lc,ld>::quartet<R> writes the spherical block
// (ab|cd) in real type R.
template <int la, int lb, int lc, int ld>
struct EriClass;

// Boys function F_0..F_L(T).  T < BOYS_TMAX: F_L and F_{L-1} from their
// order-BOYS_ORDER Taylor expansions around the nearest grid point (the two
// blocks of table row k hold F_{L+j}(T_k) (-1)^j / j! and F_{L-1+j}(T_k)
// (-1)^j / j!), then downward recursion F_m = (2T F_{m+1} + e^-T) / (2m+1).
// T >= BOYS_TMAX: F_0 = sqrt(pi/T)/2 (erf(sqrt(T)) = 1 to double precision
// there) and upward recursion F_{m+1} = ((2m+1) F_m - e^-T) / (2T), stable
// for T > L.
// FP32 quartets keep Boys arguments near zero in FP64: below BOYS_T_FP64 the FP32 kernels take F_L
// from its series at T = 0, F_L(T) = sum_k (-T)^k / (k! (2L + 2k + 1)), and e^-T from its own series
// (18 terms each: 1^18 / 18! < 2e-16), then the downward recursion in FP64.
constexpr double BOYS_T_FP64 = 1.0;
constexpr int BOYS_NEAR0_TERMS = 18;

template <int L>
__device__ __forceinline__ void boys_near_zero_fp64(const double t, float* __restrict__ F)
{
    double f = 0.0, e = 0.0, inv_fact = 1.0;
    // Horner from the highest term: coefficient k of F_L is 1 / (k! (2L + 2k + 1)) with sign (-1)^k
    double cf[BOYS_NEAR0_TERMS + 1], ce[BOYS_NEAR0_TERMS + 1];
#pragma unroll
    for (int k = 0; k <= BOYS_NEAR0_TERMS; k++) {
        if (k > 0) inv_fact /= k;
        cf[k] = inv_fact / (2 * L + 2 * k + 1);
        ce[k] = inv_fact;
    }
#pragma unroll
    for (int k = BOYS_NEAR0_TERMS; k >= 0; k--) {
        f = fma(f, -t, cf[k]);
        e = fma(e, -t, ce[k]);
    }
    double fm = f;
    F[L] = static_cast<float>(fm);
#pragma unroll
    for (int m = L - 1; m >= 0; m--) {
        fm = (2.0 * t * fm + e) * (1.0 / (2 * m + 1));
        F[m] = static_cast<float>(fm);
    }
}

template <int L, typename R>
__device__ __forceinline__ void boys_eval(const R T, const R* __restrict__ tab, R* __restrict__ F)
{
    // Taylor orders: full table order in FP64; 5 terms reach FP32 rounding
    // (0.125^6 / 6! ~ 5e-9)
    constexpr int ORD = (sizeof(R) == 8) ? BOYS_ORDER : 5;
    if constexpr (sizeof(R) == 4) {
        if (T < R(BOYS_T_FP64)) {
            boys_near_zero_fp64<L>(static_cast<double>(T), F);
            return;
        }
    }
    if (T < R(BOYS_TMAX)) {
        const int k = static_cast<int>(T * R(BOYS_INV_DT) + R(0.5));
        const R d = T - R(BOYS_DT) * R(k);
        const R* c = tab + k * BOYS_ROW;
        R f = c[ORD];
#pragma unroll
        for (int j = ORD - 1; j >= 0; j--) f = fma(f, d, c[j]);
        F[L] = f;
        if constexpr (L > 0) {
            // F_{L-1} from its own Taylor block of the row; then e^-T = (2L - 1) F_{L-1} - 2T F_L
            // (the recursion at m = L - 1) for the downward steps.  At large T the difference
            // cancels, but its error enters F_m with weight <= (2L - 1) F_{L-1} / ((2m + 1) F_m) <= 1
            const R* c1 = c + (BOYS_ORDER + 1);
            R g = c1[ORD];
#pragma unroll
            for (int j = ORD - 1; j >= 0; j--) g = fma(g, d, c1[j]);
            F[L - 1] = g;
            if constexpr (L > 1) {
                const R tt = R(2.0) * T;
                const R et = R(2 * L - 1) * g - tt * f;
#pragma unroll
                for (int m = L - 2; m >= 0; m--) F[m] = (tt * F[m + 1] + et) * R(1.0 / (2 * m + 1));
            }
        }
    } else {
        const R rs = vrsqrt<R>(T);
        const R it = rs * rs;
        F[0] = R(0.88622692545275801365) * rs;  // sqrt(pi) / 2 / sqrt(T)
        if constexpr (L > 0) {
            // F_{m+1} = ((2m+1) F_m - e^-T) / (2T).  Next to (2m+1) F_m the e^-T term
            // weighs at most e^-36 / ((2m+1) F_m(36)): 3e-10 for m < 6, 1e-6 for m = 11.
            // L <= 6: e^-T to ~1e-7 relative is enough, as 2^x with x = -T log2(e)
            // split in FP64 into n + f (0 <= f < 1) and 2^f from the hardware FP32
            // exp2; in FP32 the term is below rounding and is dropped.  Higher L
            // takes the full-precision exponential.
            R et = R(0);
            if constexpr (L > 6) {
                et = vexp<R>(-T);
            } else if constexpr (sizeof(R) == 8) {
                const double x = -T * 1.4426950408889634;
                const double nf = floor(x);
                const float f = static_cast<float>(x - nf);
                et = ldexp(static_cast<double>(__builtin_amdgcn_exp2f(f)), static_cast<int>(nf));
            }
#pragma unroll
            for (int m = 0; m < L; m++) F[m + 1] = (R(2 * m + 1) * F[m] - et) * (R(0.5) * it);
        }
    }
}

// FP64 quartet (all FP64 kernels); RS selects the erf(omega r)/r operator
template <int la, int lb, int lc, int ld, bool RS = false>
__device__ __forceinline__ void eri_quartet(const PrimPair* __restrict__ bra_prims, const int bra_np, const double ABx,
                                            const double ABy, const double ABz, const PrimPair* __restrict__ ket_prims,
                                            const int ket_np, const double CDx, const double CDy, const double CDz,
                                            const double* __restrict__ boys_tab, const double omega2,
                                            double* __restrict__ out)
{
    EriClass<la, lb, lc, ld>::template quartet<double, RS>(bra_prims, bra_np, ABx, ABy, ABz, ket_prims, ket_np, CDx, CDy,
                                                       CDz, boys_tab, omega2, out);
}

// Cross-lane moves within rows of 16 lanes (DPP): quad_perm [1,0,3,2] (0xB1),
// quad_perm [2,3,0,1] (0x4E), row_half_mirror (0x141), row_mirror (0x140).
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp(static_cast<int>(b), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(static_cast<int>(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
__device__ __forceinline__ float read_lane(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
__device__ __forceinline__ double read_lane(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), l);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// Sum over the 64 lanes of a wavefront (all lanes must be active); the
// result is wave-uniform.  Butterfly within each row of 16 lanes by DPP
// moves, then the four row sums are read out with readlane.
// Sum over each row of 16 lanes (left in every lane of the row).
template <typename R>
__device__ __forceinline__ R row_sum(R v)
{
    v += dpp_move<0xB1>(v);
    v += dpp_move<0x4E>(v);
    v += dpp_move<0x141>(v);
    v += dpp_move<0x140>(v);
    return v;
}

template <typename R>
__device__ __forceinline__ R wave_sum(R v)
{
    v += dpp_move<0xB1>(v);
    v += dpp_move<0x4E>(v);
    v += dpp_move<0x141>(v);
    v += dpp_move<0x140>(v);
    return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

template <typename R>
__device__ __forceinline__ void stage_boys(R* lds, const R* __restrict__ tab)
{
    for (int t = threadIdx.x; t < BOYS_TAB_SIZE; t += blockDim.x) lds[t] = tab[t];
    __syncthreads();
}

// Density-weighted screening (a non-reference fast mode): glim[g1 * ng + g2] is the largest bin sum
// a quartet may have when it reads the D block of atom groups (g1, g2), from that block's max|D|;
// a quartet of groups (A, B|C, D) reads D_CD, D_AB (Coulomb) and D_BD, D_BC, D_AD, D_AC (exchange),
// so it is kept up to the largest of those six limits.
__device__ __forceinline__ int quartet_bin_limit(const int* __restrict__ glim, const int ng, const int A, const int B,
                                                 const int C, const int D)
{
    const int l0 = max(glim[C * ng + D], glim[A * ng + B]);
    const int l1 = max(glim[B * ng + D], glim[B * ng + C]);
    const int l2 = max(glim[A * ng + D], glim[A * ng + C]);
    return max(l0, max(l1, l2));
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
template <int la, int lb, int lc, int ld, bool RS>
__global__ void __launch_bounds__(VLX_BLOCK)
    fock_kernel(const PairInfo* __restrict__ bra, const PrimPair* __restrict__ bra_prims,
                const PairInfo* __restrict__ ket, const PrimPair* __restrict__ ket_prims,
                const WorkItem* __restrict__ items, const double* __restrict__ boys_tab,
                const double* __restrict__ D0, double* __restrict__ F0, const MolDesc* __restrict__ mols,
                const int* __restrict__ glim, const FockScalars a)
{
    constexpr int na = 2 * la + 1, nb = 2 * lb + 1, nc = 2 * lc + 1, nd = 2 * ld + 1;
    constexpr int nsph = na * nb * nc * nd;
    __shared__ double tab[BOYS_TAB_SIZE];
    stage_boys(tab, boys_tab);

    const bool do_j = a.fj != 0.0, do_k = a.fk != 0.0;
    const int lane = threadIdx.x & (VLX_WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / VLX_WAVE);
    const long long gw = (long long)blockIdx.x * VLX_WAVES_PER_BLOCK + wave;
    const long long nw = (long long)gridDim.x * VLX_WAVES_PER_BLOCK;
    const long long w0 = a.nitems * gw / nw, w1 = a.nitems * (gw + 1) / nw;

    for (long long it = w0; it < w1; it++) {
        const WorkItem wi = items[it];
        const MolDesc md = mols[wi.mol];
        const double* __restrict__ D = D0 + md.base;
        double* __restrict__ F = F0 + md.base;
        const int n = md.nao;
        const int j = wi.j0 + lane;
        const bool lane_ok = j < wi.jend;
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
            if (glim != nullptr)
                active = active && (bp.bin + kp.bin <=
                                     quartet_bin_limit(glim, a.ngroups, bp.grp_a, bp.grp_b, kp.grp_a, kp.grp_b));
            if (a.same_class) active = active && (bp.rank >= j);
            double s = active ? 1.0 : 0.0;
            if (bp.same) s *= 0.5;
            if (kp.same) s *= 0.5;
            if (a.same_class && bp.rank == j) s *= 0.5;
            double I[nsph];
            eri_quartet<la, lb, lc, ld, RS>(bra_prims + bp.prim_start, bp.nprim, bp.abx, bp.aby, bp.abz,
                                            ket_prims + kp.prim_start, active ? kp.nprim : 0, kp.abx, kp.aby, kp.abz,
                                            tab, a.omega2, I);
#pragma unroll
            for (int x = 0; x < nsph; x++) I[x] = active ? I[x] * s : 0.0;
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

// Deterministic accumulation (FockScalars::det): an F' element is a 128-bit fixed-point integer of
// resolution 2^-96, two 64-bit words (low word first), added with integer atomics.  Integer sums do
// not depend on the order of the additions, so a build gives the same bits on every run and with
// any number of streams.  x = hi 2^-32 + lo 2^-96 with hi = floor(x 2^32) and lo = (x 2^32 - hi) 2^64:
// both scalings are by powers of two and x 2^32 - floor(x 2^32) is exact, so only the bits of x
// below 2^-96 are dropped (|x| < 2^31).  The carry out of the low word is counted by the atomic
// whose addition wraps it, once per wrap, so it too is independent of the order.
__device__ __forceinline__ void fx_add(unsigned long long* __restrict__ acc, const size_t idx, const double x)
{
    if (x == 0.0) return;
    const double y = x * 4294967296.0;
    const double h = floor(y);
    const unsigned long long lo = static_cast<unsigned long long>((y - h) * 18446744073709551616.0);
    unsigned long long hi = static_cast<unsigned long long>(static_cast<long long>(h));
    unsigned long long* p = acc + 2 * idx;
    if (lo != 0ull) {
        const unsigned long long old = atomicAdd(p, lo);
        if (old + lo < old) hi += 1ull;
    }
    if (hi != 0ull) atomicAdd(p + 1, hi);
}

// F'[idx] += v: an FP64 atomic, or the fixed-point add above into the accumulator array that F
// then points to (two words per element)
template <bool DET>
__device__ __forceinline__ void acc_add(double* __restrict__ F, const size_t idx, const double v)
{
    if constexpr (DET) fx_add(reinterpret_cast<unsigned long long*>(F), idx, v);
    else unsafeAtomicAdd(F + idx, v);
}

// Lane-private LDS accumulators of fock_kernel_ksl: slot e of the calling thread at p[e * stride]
// (stride = threads per block, so the lanes of a wave hit consecutive banks).
struct LaneAcc {
    double* jcd;  // J'_cd of the lane's ket, nc x nd (factor fj applied)
    double* kac;  // K'_ac for the current shell a, na x nc (factor fk applied)
    double* kad;  // K'_ad, na x nd
    int stride;
};

// Direct digestion of one quartet block per lane (bra pair wave-uniform, ket pair per lane; every
// lane of the wavefront calls it): F'_ab is reduced over the wavefront, F'_cd, F'_ac, F'_ad, F'_bc
// and F'_bd are per-lane atomics.  Masked lanes hold exact zeros in I.
// LACC: J'_cd, K'_ac and K'_ad go to the lane's LDS accumulators instead (no atomics: only the lane
// touches them).
template <int la, int lb, int lc, int ld, bool DET = false, bool LACC = false>
__device__ __forceinline__ void digest_direct(const double (&I)[(2 * la + 1) * (2 * lb + 1) * (2 * lc + 1) * (2 * ld + 1)],
                                              const PairInfo& bp, const PairInfo& kp, const bool active,
                                              const int lane, const double* __restrict__ D,
                                              double* __restrict__ F, const int n, const double fj,
                                              const double fk, const LaneAcc* lacc = nullptr)
{
    constexpr int na = 2 * la + 1, nb = 2 * lb + 1, nc = 2 * lc + 1, nd = 2 * ld + 1;
    int p[na], q[nb], r[nc], t[nd];
#pragma unroll
    for (int i = 0; i < na; i++) p[i] = bp.ao_a + i * bp.str_a;
#pragma unroll
    for (int i = 0; i < nb; i++) q[i] = bp.ao_b + i * bp.str_b;
#pragma unroll
    for (int i = 0; i < nc; i++) r[i] = kp.ao_a + i * kp.str_a;
#pragma unroll
    for (int i = 0; i < nd; i++) t[i] = kp.ao_b + i * kp.str_b;
    if (fj != 0.0) {
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
                if (lane == 0 && v != 0.0) acc_add<DET>(F, (size_t)p[i] * n + q[jb], fj * v);
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
                    if constexpr (LACC) lacc->jcd[(k * nd + l) * lacc->stride] += fj * v;
                    else acc_add<DET>(F, (size_t)r[k] * n + t[l], fj * v);
                }
        }
    }
    if (fk != 0.0 && active) {
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
                if constexpr (LACC) lacc->kac[(i * nc + k) * lacc->stride] += fk * v;
                else acc_add<DET>(F, (size_t)p[i] * n + r[k], fk * v);
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
                if constexpr (LACC) lacc->kad[(i * nd + l) * lacc->stride] += fk * v;
                else acc_add<DET>(F, (size_t)p[i] * n + t[l], fk * v);
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
                acc_add<DET>(F, (size_t)q[jb] * n + r[k], fk * v);
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
                acc_add<DET>(F, (size_t)q[jb] * n + t[l], fk * v);
            }
    }
}

// Bra-stationary digestion, for classes whose integrals need most of the
// register file.  A wavefront takes one bra pair (wave-uniform) and 64 kets
// (one per lane) from the prefix of the ket list its bins allow; work item
// {j0, i0, i1}: kets [j0, j0 + 64) of bra i0, i1 = number of kets of bra i0.
// F'_ab is reduced over the wavefront; F'_cd, F'_ac, F'_ad, F'_bc and F'_bd
// are per-lane atomics.
template <int la, int lb, int lc, int ld, bool RS>
__global__ void __launch_bounds__(VLX_BLOCK)
    fock_kernel_bs(const PairInfo* __restrict__ bra, const PrimPair* __restrict__ bra_prims,
                   const PairInfo* __restrict__ ket, const PrimPair* __restrict__ ket_prims,
                   const WorkItem* __restrict__ items, const double* __restrict__ boys_tab,
                   const double* __restrict__ D0, double* __restrict__ F0, const MolDesc* __restrict__ mols,
                   const int* __restrict__ glim, const FockScalars a)
{
    constexpr int na = 2 * la + 1, nb = 2 * lb + 1, nc = 2 * lc + 1, nd = 2 * ld + 1;
    constexpr int nsph = na * nb * nc * nd;
    __shared__ double tab[BOYS_TAB_SIZE];
    stage_boys(tab, boys_tab);

    const int lane = threadIdx.x & (VLX_WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / VLX_WAVE);
    const long long gw = (long long)blockIdx.x * VLX_WAVES_PER_BLOCK + wave;
    const long long nw = (long long)gridDim.x * VLX_WAVES_PER_BLOCK;
    const long long w0 = a.nitems * gw / nw, w1 = a.nitems * (gw + 1) / nw;

    for (long long it = w0; it < w1; it++) {
        const WorkItem wi = items[it];
        const MolDesc md = mols[wi.mol];
        const double* __restrict__ D = D0 + md.base;
        double* __restrict__ F = F0 + md.base;
        const int n = md.nao;
        const PairInfo bp = bra[wi.i0];
        const int j = wi.j0 + lane;
        bool active = j < wi.i1;
        const PairInfo kp = ket[active ? j : wi.j0];
        if (glim != nullptr) {
            active = active &&
                     (bp.bin + kp.bin <= quartet_bin_limit(glim, a.ngroups, bp.grp_a, bp.grp_b, kp.grp_a, kp.grp_b));
            if (__ballot(active) == 0) continue;  // wave-uniform: the item has no quartet left
        }
        double s = active ? 1.0 : 0.0;
        if (bp.same) s *= 0.5;
        if (kp.same) s *= 0.5;
        if (a.same_class && bp.rank == j) s *= 0.5;
        double I[nsph];
        eri_quartet<la, lb, lc, ld, RS>(bra_prims + bp.prim_start, bp.nprim, bp.abx, bp.aby, bp.abz,
                                        ket_prims + kp.prim_start, active ? kp.nprim : 0, kp.abx, kp.aby, kp.abz, tab,
                                        a.omega2, I);
#pragma unroll
        for (int x = 0; x < nsph; x++) I[x] = active ? I[x] * s : 0.0;
        digest_direct<la, lb, lc, ld>(I, bp, kp, active, lane, D, F, n, a.fj, a.fk);
    }
}

// Multi-density digestion: the integrals of each quartet, computed once, are digested into
// a.ndens densities (density d at D0 + d * dstride, its F' at F0 + d * dstride, factors
// dfac[2 d] (Coulomb) and dfac[2 d + 1] (exchange)), every contribution by direct digestion:
// per-density register accumulators would not fit beside the integrals.  The work items are
// those of the class's untiled schedule: ket-stationary (a.bs_layout = 0, bras [i0, i1) against kets
// [j0, jend)) or bra-stationary (a.bs_layout = 1, bra i0 against kets [j0, i1)).  a.det: F0 is the
// fixed-point accumulator array of fx_add (two words per element, same offsets doubled).
template <int la, int lb, int lc, int ld, bool RS>
__global__ void __launch_bounds__(VLX_BLOCK)
    fock_kernel_multi(const PairInfo* __restrict__ bra, const PrimPair* __restrict__ bra_prims,
                      const PairInfo* __restrict__ ket, const PrimPair* __restrict__ ket_prims,
                      const WorkItem* __restrict__ items, const double* __restrict__ boys_tab,
                      const double* __restrict__ D0, double* __restrict__ F0, const MolDesc* __restrict__ mols,
                      const double* __restrict__ dfac, const FockScalars a)
{
    constexpr int nsph = (2 * la + 1) * (2 * lb + 1) * (2 * lc + 1) * (2 * ld + 1);
    __shared__ double tab[BOYS_TAB_SIZE];
    stage_boys(tab, boys_tab);

    const int lane = threadIdx.x & (VLX_WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / VLX_WAVE);
    const long long gw = (long long)blockIdx.x * VLX_WAVES_PER_BLOCK + wave;
    const long long nw = (long long)gridDim.x * VLX_WAVES_PER_BLOCK;
    const long long w0 = a.nitems * gw / nw, w1 = a.nitems * (gw + 1) / nw;
    const int nds = a.ndens > 1 ? a.ndens : 1;
    const bool BSL = a.bs_layout != 0;

    for (long long it = w0; it < w1; it++) {
        const WorkItem wi = items[it];
        const MolDesc md = mols[wi.mol];
        const int n = md.nao;
        const int j = wi.j0 + lane;
        const bool lane_ok = j < (BSL ? wi.i1 : wi.jend);
        const PairInfo kp = ket[lane_ok ? j : wi.j0];
        const int ib1 = BSL ? wi.i0 + 1 : wi.i1;
        for (int ib = wi.i0; ib < ib1; ib++) {
            const PairInfo bp = bra[ib];
            bool active = lane_ok;
            // bra-stationary items hold only kets within the bin and rank limits
            if (!BSL && a.ithreshold >= 0) active = active && (bp.bin + kp.bin <= a.ithreshold);
            if (!BSL && a.same_class) active = active && (bp.rank >= j);
            double s = 1.0;
            if (bp.same) s *= 0.5;
            if (kp.same) s *= 0.5;
            if (a.same_class && bp.rank == j) s *= 0.5;
            double I[nsph];
            eri_quartet<la, lb, lc, ld, RS>(bra_prims + bp.prim_start, bp.nprim, bp.abx, bp.aby, bp.abz,
                                            ket_prims + kp.prim_start, active ? kp.nprim : 0, kp.abx, kp.aby, kp.abz,
                                            tab, a.omega2, I);
#pragma unroll
            for (int x = 0; x < nsph; x++) I[x] = active ? I[x] * s : 0.0;
            for (int dd = 0; dd < nds; dd++) {
                const long long off = md.base + dd * a.dstride;
                if (a.det)  // F0 is the fixed-point accumulator array, two words per element
                    digest_direct<la, lb, lc, ld, true>(I, bp, kp, active, lane, D0 + off, F0 + 2 * off, n,
                                                        dfac[2 * dd], dfac[2 * dd + 1]);
                else
                    digest_direct<la, lb, lc, ld>(I, bp, kp, active, lane, D0 + off, F0 + off, n, dfac[2 * dd],
                                                  dfac[2 * dd + 1]);
            }
        }
    }
}

// Ket-stationary digestion with lane-private LDS accumulators, for classes whose integrals leave no
// registers for the accumulators of fock_kernel.  Work items as in fock_kernel (64 kets, one per
// lane, against a range of bras sorted by shell A).  J'_cd, K'_ac and K'_ad of a lane's ket
// accumulate in LDS slots only that lane touches (no atomics); K'_ac and K'_ad go to F' when shell A
// changes, J'_cd at the end of the item.  F'_ab is reduced over the wavefront, F'_bc and F'_bd are
// per-lane atomics, as in the bra-stationary kernel, which sends all six per quartet.  Dynamic LDS:
// (na nc + na nd + nc nd) doubles per thread.
template <int la, int lb, int lc, int ld, bool RS, int NW>
__global__ void __launch_bounds__(NW * VLX_WAVE)
    fock_kernel_ksl(const PairInfo* __restrict__ bra, const PrimPair* __restrict__ bra_prims,
                    const PairInfo* __restrict__ ket, const PrimPair* __restrict__ ket_prims,
                    const WorkItem* __restrict__ items, const double* __restrict__ boys_tab,
                    const double* __restrict__ D0, double* __restrict__ F0, const MolDesc* __restrict__ mols,
                    const int* __restrict__ glim, const FockScalars a)
{
    constexpr int na = 2 * la + 1, nb = 2 * lb + 1, nc = 2 * lc + 1, nd = 2 * ld + 1;
    constexpr int nsph = na * nb * nc * nd;
    constexpr int NT = NW * VLX_WAVE;
    constexpr int NACC = na * nc + na * nd + nc * nd;
    __shared__ double tab[BOYS_TAB_SIZE];
    extern __shared__ double lslots[];
    const LaneAcc acc{lslots + (na * nc + na * nd) * NT + threadIdx.x, lslots + threadIdx.x,
                      lslots + na * nc * NT + threadIdx.x, NT};
    for (int e = 0; e < NACC; e++) lslots[e * NT + threadIdx.x] = 0.0;
    stage_boys(tab, boys_tab);

    const int lane = threadIdx.x & (VLX_WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / VLX_WAVE);
    const long long gw = (long long)blockIdx.x * NW + wave;
    const long long nw = (long long)gridDim.x * NW;
    const long long w0 = a.nitems * gw / nw, w1 = a.nitems * (gw + 1) / nw;

    for (long long it = w0; it < w1; it++) {
        const WorkItem wi = items[it];
        const MolDesc md = mols[wi.mol];
        const double* __restrict__ D = D0 + md.base;
        double* __restrict__ F = F0 + md.base;
        const int n = md.nao;
        const int j = wi.j0 + lane;
        const bool lane_ok = j < wi.jend;
        const PairInfo kp = ket[lane_ok ? j : wi.j0];
        int r[nc], t[nd], p[na];
#pragma unroll
        for (int k = 0; k < nc; k++) r[k] = kp.ao_a + k * kp.str_a;
#pragma unroll
        for (int l = 0; l < nd; l++) t[l] = kp.ao_b + l * kp.str_b;
        int cur_a = -1;
        bool pending = false;  // this lane's K slots hold contributions for shell cur_a
        for (int ib = wi.i0; ib < wi.i1; ib++) {
            const PairInfo bp = bra[ib];
            if (bp.shell_a != cur_a) {
                if (pending) {
#pragma unroll
                    for (int i = 0; i < na; i++) {
#pragma unroll
                        for (int k = 0; k < nc; k++) {
                            double& v = acc.kac[(i * nc + k) * NT];
                            if (v != 0.0) unsafeAtomicAdd(&F[(size_t)p[i] * n + r[k]], v);
                            v = 0.0;
                        }
#pragma unroll
                        for (int l = 0; l < nd; l++) {
                            double& v = acc.kad[(i * nd + l) * NT];
                            if (v != 0.0) unsafeAtomicAdd(&F[(size_t)p[i] * n + t[l]], v);
                            v = 0.0;
                        }
                    }
                }
                pending = false;
                cur_a = bp.shell_a;
#pragma unroll
                for (int i = 0; i < na; i++) p[i] = bp.ao_a + i * bp.str_a;
            }
            bool active = lane_ok;
            if (a.ithreshold >= 0) active = active && (bp.bin + kp.bin <= a.ithreshold);
            if (glim != nullptr)
                active = active && (bp.bin + kp.bin <=
                                     quartet_bin_limit(glim, a.ngroups, bp.grp_a, bp.grp_b, kp.grp_a, kp.grp_b));
            if (a.same_class) active = active && (bp.rank >= j);
            double s = 1.0;
            if (bp.same) s *= 0.5;
            if (kp.same) s *= 0.5;
            if (a.same_class && bp.rank == j) s *= 0.5;
            double I[nsph];
            eri_quartet<la, lb, lc, ld, RS>(bra_prims + bp.prim_start, bp.nprim, bp.abx, bp.aby, bp.abz,
                                            ket_prims + kp.prim_start, active ? kp.nprim : 0, kp.abx, kp.aby, kp.abz,
                                            tab, a.omega2, I);
#pragma unroll
            for (int x = 0; x < nsph; x++) I[x] = active ? I[x] * s : 0.0;
            digest_direct<la, lb, lc, ld, false, true>(I, bp, kp, active, lane, D, F, n, a.fj, a.fk, &acc);
            pending = pending || active;
        }
        if (pending) {
#pragma unroll
            for (int i = 0; i < na; i++) {
#pragma unroll
                for (int k = 0; k < nc; k++) {
                    double& v = acc.kac[(i * nc + k) * NT];
                    if (v != 0.0) unsafeAtomicAdd(&F[(size_t)p[i] * n + r[k]], v);
                    v = 0.0;
                }
#pragma unroll
                for (int l = 0; l < nd; l++) {
                    double& v = acc.kad[(i * nd + l) * NT];
                    if (v != 0.0) unsafeAtomicAdd(&F[(size_t)p[i] * n + t[l]], v);
                    v = 0.0;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < nc; k++)
#pragma unroll
            for (int l = 0; l < nd; l++) {
                double& v = acc.jcd[(k * nd + l) * NT];
                if (v != 0.0) unsafeAtomicAdd(&F[(size_t)r[k] * n + t[l]], v);
                v = 0.0;
            }
    }
}

// LDS-tiled digestion.  A work-group takes one TileItem: bras of atom groups
// (ga, gb) and up to 64 kets of groups (gc, gd).  Every F' contribution of
// the item lands in six small dense tiles indexed by group-local AOs,
//   J_AB, J_CD, K_AC, K_AD, K_BC, K_BD,
// accumulated with LDS atomics and flushed to HBM once per item.  The four
// wavefronts share the kets and split the bras.  MULTI: the item's integrals
// are digested into a.ndens densities (layout and factors as in
// fock_kernel_multi), with one set of six tiles per density in LDS.  DLDS: the
// six D blocks the item reads (D_AB, D_CD, D_AC, D_AD, D_BC, D_BD, shapes of the
// F' tiles) are staged in LDS after the F' tiles once per item, and every lane
// reads D from there instead of gathering it from HBM (single density only).
__device__ __forceinline__ void flush_tile(double* __restrict__ F, const int n, const double* __restrict__ tile,
                                           const int nr, const int nc, const int* __restrict__ rows,
                                           const int* __restrict__ cols, const double f)
{
    for (int e = threadIdx.x; e < nr * nc; e += blockDim.x) {
        const double v = tile[e];
        if (v != 0.0) unsafeAtomicAdd(&F[(size_t)rows[e / nc] * n + cols[e % nc]], f * v);
    }
}

template <int la, int lb, int lc, int ld, typename R, int NW, bool RS, bool MULTI = false, bool DLDS = false>
__global__ void __launch_bounds__(NW * VLX_WAVE)
    fock_kernel_tiled(const PairInfo* __restrict__ pairs_bra, const PrimPair* __restrict__ bra_prims,
                      const PairInfo* __restrict__ pairs_ket, const PrimPair* __restrict__ ket_prims,
                      const TileItem* __restrict__ items, const int2* __restrict__ kentries,
                      const R* __restrict__ boys_tab, const double* __restrict__ D0, double* __restrict__ F0,
                      const MolDesc* __restrict__ mols, const int* __restrict__ k32, const int* __restrict__ glim,
                      const int* __restrict__ grp_off, const int* __restrict__ grp_n,
                      const int* __restrict__ grp_ao, const double* __restrict__ dfac, const FockScalars a)
{
    constexpr int na = 2 * la + 1, nb = 2 * lb + 1, nc = 2 * lc + 1, nd = 2 * ld + 1;
    constexpr int nsph = na * nb * nc * nd;
    __shared__ R tab[BOYS_TAB_SIZE];
    extern __shared__ double tiles[];
    stage_boys<R>(tab, boys_tab);

    const int lane = threadIdx.x & (VLX_WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / VLX_WAVE);
    const int nds = (MULTI && a.ndens > 1) ? a.ndens : 1;

    for (long long it = blockIdx.x; it < a.nitems; it += gridDim.x) {
        const TileItem ti = items[it];
        const MolDesc md = mols[ti.mol];
        const double* __restrict__ D = D0 + md.base;
        double* __restrict__ F = F0 + md.base;
        const int n = md.nao;
        // bras [b0, b1) of the item: all of them, or, with the precision split on the device, those
        // whose bin plus the smallest ket bin of the item is below k32 (FP64 kernel) or not (FP32
        // kernel).  The item's bras are sorted by bin, so the split is one point, found by bisection.
        int b0 = ti.i0, b1 = ti.i1;
        if (k32 != nullptr) {
            const int lim = k32[ti.mol] - ti.bin_lo;  // FP32 from the first bra with bin >= lim
            int lo = ti.i0, hi = ti.i1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (pairs_bra[mid].bin < lim) lo = mid + 1;
                else hi = mid;
            }
            if (sizeof(R) == sizeof(double)) b1 = lo;
            else b0 = lo;
            if (b0 >= b1) continue;  // block-uniform: every thread skips the item together
        }
        // density-weighted screening: every quartet of the item reads the D blocks of its four atom
        // groups, so the item has one bin-sum limit; bras past it (sorted by bin) are cut off
        int ilim = a.ithreshold;
        if (glim != nullptr) {
            int ql = quartet_bin_limit(glim, a.ngroups, ti.ga, ti.gb, ti.gc, ti.gd);
            if (ti.gd2 >= 0) ql = max(ql, quartet_bin_limit(glim, a.ngroups, ti.ga, ti.gb, ti.gc, ti.gd2));
            ilim = min(ilim, ql);
            const int bl = ilim - ti.bin_lo;  // a bra needs bin <= bl for one quartet with the item's kets
            int lo = b0, hi = b1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (pairs_bra[mid].bin <= bl) lo = mid + 1;
                else hi = mid;
            }
            b1 = lo;
            if (b0 >= b1) continue;  // block-uniform
        }
        const int nA = grp_n[ti.ga * 4 + la], nB = grp_n[ti.gb * 4 + lb];
        const int nC = grp_n[ti.gc * 4 + lc];
        // D groups: gd, and gd2 for the entries [jm, j1) of a chunk spanning two ket segments
        const bool two = ti.gd2 >= 0;
        const int nD0 = grp_n[ti.gd * 4 + ld], nD1 = two ? grp_n[ti.gd2 * 4 + ld] : 0;
        // tiles of one density: J_AB, K_AC, K_BC, then per D group J_CD, K_AD, K_BD
        const int nshared = nA * nB + nA * nC + nB * nC;
        const int nseg0 = (nC + nA + nB) * nD0;
        const int ntile = nshared + nseg0 + (nC + nA + nB) * nD1;
        for (int e = threadIdx.x; e < nds * ntile; e += blockDim.x) tiles[e] = 0.0;
        const int* ra = grp_ao + grp_off[ti.ga * 4 + la];
        const int* rb = grp_ao + grp_off[ti.gb * 4 + lb];
        const int* rc = grp_ao + grp_off[ti.gc * 4 + lc];
        const int* rd0 = grp_ao + grp_off[ti.gd * 4 + ld];
        const int* rd1 = two ? grp_ao + grp_off[ti.gd2 * 4 + ld] : rd0;
        // D blocks in LDS (DLDS) after the F' tiles, in the same layout
        double* dtiles = tiles + ntile;
        if constexpr (DLDS) {
            auto stage = [&](double* dst, const int* rows, const int nr, const int* cols, const int ncol) {
                for (int x = threadIdx.x; x < nr * ncol; x += blockDim.x)
                    dst[x] = D[(size_t)rows[x / ncol] * n + cols[x % ncol]];
            };
            stage(dtiles, ra, nA, rb, nB);
            stage(dtiles + nA * nB, ra, nA, rc, nC);
            stage(dtiles + nA * nB + nA * nC, rb, nB, rc, nC);
            for (int sgi = 0; sgi < (two ? 2 : 1); sgi++) {
                const int nDs = sgi ? nD1 : nD0;
                const int* rds = sgi ? rd1 : rd0;
                double* base = dtiles + nshared + (sgi ? nseg0 : 0);
                stage(base, rc, nC, rds, nDs);
                stage(base + nC * nDs, ra, nA, rds, nDs);
                stage(base + nC * nDs + nA * nDs, rb, nB, rds, nDs);
            }
        }
        __syncthreads();

        // lane = a run of primitive pairs of one ket pair (first | count << 16):
        // the contraction over the pair's runs happens in the LDS tiles (every
        // step after the VRR is linear)
        const int e = ti.j0 + lane;
        const bool lane_ok = e < ti.j1;
        const int esel = lane_ok ? e : ti.j0;
        const int2 ke = kentries[esel];
        const int kq0 = ke.y & 0xffff, kqn = ke.y >> 16;
        const PairInfo kp = pairs_ket[ke.x];
        // the lane's D group: the tiles J_CD, K_AD, K_BD (and D_CD, D_AD, D_BD) of its segment
        const int sg = (two && esel >= ti.jm) ? 1 : 0;
        const int nD = sg ? nD1 : nD0;
        const int dofs = nshared + (sg ? nseg0 : 0);
        const double* d_ab = dtiles;
        const double* d_ac = d_ab + nA * nB;
        const double* d_bc = d_ac + nA * nC;
        const double* d_cd = dtiles + dofs;
        const double* d_ad = d_cd + nC * nD;
        const double* d_bd = d_ad + nA * nD;
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
        R dcd[nc][nd];
        if constexpr (!MULTI) {
#pragma unroll
            for (int k = 0; k < nc; k++)
#pragma unroll
                for (int l = 0; l < nd; l++)
                    dcd[k][l] = DLDS ? R(d_cd[rl[k] * nD + tl[l]]) : R(D[(size_t)r[k] * n + t[l]]);
        }
        // J_CD depends on the lane's ket only: small ket blocks accumulate it in
        // registers over all bras of the item (one LDS atomic per element at the end)
        constexpr bool jcd_in_regs = !MULTI && nc * nd <= 9;
        double jcd[jcd_in_regs ? nc : 1][jcd_in_regs ? nd : 1];
#pragma unroll
        for (int k = 0; k < (jcd_in_regs ? nc : 1); k++)
#pragma unroll
            for (int l = 0; l < (jcd_in_regs ? nd : 1); l++) jcd[k][l] = 0.0;

        for (int ib = b0 + wave; ib < b1; ib += NW) {
            const PairInfo bp = pairs_bra[ib];
            bool active = lane_ok;
            if (a.ithreshold >= 0) active = active && (bp.bin + kp.bin <= ilim);
            if (a.same_class) active = active && (bp.rank >= kp.rank);
            R s = active ? R(1) : R(0);
            if (bp.same) s *= R(0.5);
            if (kp.same) s *= R(0.5);
            if (a.same_class && bp.rank == kp.rank) s *= R(0.5);
            R I[nsph];
            EriClass<la, lb, lc, ld>::template quartet<R, RS>(bra_prims + bp.prim_start, bp.nprim, R(bp.abx), R(bp.aby),
                                                          R(bp.abz), ket_prims + kp.prim_start + kq0, active ? kqn : 0,
                                                          R(kp.abx), R(kp.aby), R(kp.abz), tab, R(a.omega2), I);
#pragma unroll
            // masked lanes contribute exact zeros: multiplying their integrals by s = 0 would turn
            // a non-finite value left in a masked lane's registers into NaN (seen in the L = 5
            // tiled kernels, depending on what ran before them)
            for (int x = 0; x < nsph; x++) I[x] = active ? I[x] * s : R(0);
            int p[na], q[nb];
#pragma unroll
            for (int i = 0; i < na; i++) p[i] = bp.ao_a + i * bp.str_a;
#pragma unroll
            for (int jb = 0; jb < nb; jb++) q[jb] = bp.ao_b + jb * bp.str_b;
            for (int dd = 0; dd < nds; dd++) {
                const double* __restrict__ Dd = MULTI ? D + dd * a.dstride : D;
                double* t_jab = tiles + dd * ntile;
                double* t_kac = t_jab + nA * nB;
                double* t_kbc = t_kac + nA * nC;
                double* t_jcd = tiles + dd * ntile + dofs;
                double* t_kad = t_jcd + nC * nD;
                double* t_kbd = t_kad + nA * nD;
                const bool do_j = (MULTI ? dfac[2 * dd] : a.fj) != 0.0;
                const bool do_k = (MULTI ? dfac[2 * dd + 1] : a.fk) != 0.0;
                if constexpr (MULTI) {
#pragma unroll
                    for (int k = 0; k < nc; k++)
#pragma unroll
                        for (int l = 0; l < nd; l++) dcd[k][l] = R(Dd[(size_t)r[k] * n + t[l]]);
                }
                if (do_j) {
#pragma unroll
                    for (int i = 0; i < na; i++)
#pragma unroll
                        for (int jb = 0; jb < nb; jb++) {
                            R v = R(0);
#pragma unroll
                            for (int k = 0; k < nc; k++)
#pragma unroll
                                for (int l = 0; l < nd; l++) v = fma(I[((i * nb + jb) * nc + k) * nd + l], dcd[k][l], v);
                            // row sums by DPP; the four rows meet in the tile through LDS atomics
                            // (the LDS pipe has room where the VALU does not)
                            v = row_sum(v);
                            if ((lane & 15) == 0 && v != R(0))
                                unsafeAtomicAdd(&t_jab[(bp.loc_a + i) * nB + bp.loc_b + jb], static_cast<double>(v));
                        }
                    if (jcd_in_regs || active) {
                        R vcd[nc][nd];
#pragma unroll
                        for (int k = 0; k < nc; k++)
#pragma unroll
                            for (int l = 0; l < nd; l++) vcd[k][l] = R(0);
#pragma unroll
                        for (int i = 0; i < na; i++)
#pragma unroll
                            for (int jb = 0; jb < nb; jb++) {
                                const R dab = DLDS ? R(d_ab[(bp.loc_a + i) * nB + bp.loc_b + jb])
                                               : R(Dd[(size_t)p[i] * n + q[jb]]);
#pragma unroll
                                for (int k = 0; k < nc; k++)
#pragma unroll
                                    for (int l = 0; l < nd; l++)
                                        vcd[k][l] = fma(I[((i * nb + jb) * nc + k) * nd + l], dab, vcd[k][l]);
                            }
                        if constexpr (jcd_in_regs) {
                            // inactive lanes add zeros (their integrals were replaced by exact zeros)
#pragma unroll
                            for (int k = 0; k < nc; k++)
#pragma unroll
                                for (int l = 0; l < nd; l++) jcd[k][l] += static_cast<double>(vcd[k][l]);
                        } else {
#pragma unroll
                            for (int k = 0; k < nc; k++)
#pragma unroll
                                for (int l = 0; l < nd; l++)
                                    unsafeAtomicAdd(&t_jcd[rl[k] * nD + tl[l]], static_cast<double>(vcd[k][l]));
                        }
                    }
                }
                if (do_k && active) {
#pragma unroll
                    for (int i = 0; i < na; i++) {
#pragma unroll
                        for (int k = 0; k < nc; k++) {  // K_ac += I D_bd
                            R v = R(0);
#pragma unroll
                            for (int jb = 0; jb < nb; jb++)
#pragma unroll
                                for (int l = 0; l < nd; l++)
                                    v = fma(I[((i * nb + jb) * nc + k) * nd + l],
                                            DLDS ? R(d_bd[(bp.loc_b + jb) * nD + tl[l]]) : R(Dd[(size_t)q[jb] * n + t[l]]),
                                            v);
                            unsafeAtomicAdd(&t_kac[(bp.loc_a + i) * nC + rl[k]], static_cast<double>(v));
                        }
#pragma unroll
                        for (int l = 0; l < nd; l++) {  // K_ad += I D_bc
                            R v = R(0);
#pragma unroll
                            for (int jb = 0; jb < nb; jb++)
#pragma unroll
                                for (int k = 0; k < nc; k++)
                                    v = fma(I[((i * nb + jb) * nc + k) * nd + l],
                                            DLDS ? R(d_bc[(bp.loc_b + jb) * nC + rl[k]]) : R(Dd[(size_t)q[jb] * n + r[k]]),
                                            v);
                            unsafeAtomicAdd(&t_kad[(bp.loc_a + i) * nD + tl[l]], static_cast<double>(v));
                        }
                    }
#pragma unroll
                    for (int jb = 0; jb < nb; jb++) {
#pragma unroll
                        for (int k = 0; k < nc; k++) {  // K_bc += I D_ad
                            R v = R(0);
#pragma unroll
                            for (int i = 0; i < na; i++)
#pragma unroll
                                for (int l = 0; l < nd; l++)
                                    v = fma(I[((i * nb + jb) * nc + k) * nd + l],
                                            DLDS ? R(d_ad[(bp.loc_a + i) * nD + tl[l]]) : R(Dd[(size_t)p[i] * n + t[l]]),
                                            v);
                            unsafeAtomicAdd(&t_kbc[(bp.loc_b + jb) * nC + rl[k]], static_cast<double>(v));
                        }
#pragma unroll
                        for (int l = 0; l < nd; l++) {  // K_bd += I D_ac
                            R v = R(0);
#pragma unroll
                            for (int i = 0; i < na; i++)
#pragma unroll
                                for (int k = 0; k < nc; k++)
                                    v = fma(I[((i * nb + jb) * nc + k) * nd + l],
                                            DLDS ? R(d_ac[(bp.loc_a + i) * nC + rl[k]]) : R(Dd[(size_t)p[i] * n + r[k]]),
                                            v);
                            unsafeAtomicAdd(&t_kbd[(bp.loc_b + jb) * nD + tl[l]], static_cast<double>(v));
                        }
                    }
                }
            }
        }
        if constexpr (jcd_in_regs) {
            if (a.fj != 0.0 && lane_ok) {
#pragma unroll
                for (int k = 0; k < nc; k++)
#pragma unroll
                    for (int l = 0; l < nd; l++)
                        if (jcd[k][l] != 0.0) unsafeAtomicAdd(&tiles[dofs + rl[k] * nD + tl[l]], jcd[k][l]);
            }
        }
        __syncthreads();
        for (int dd = 0; dd < nds; dd++) {
            double* __restrict__ Fd = MULTI ? F + dd * a.dstride : F;
            const double fj = MULTI ? dfac[2 * dd] : a.fj, fk = MULTI ? dfac[2 * dd + 1] : a.fk;
            const double* t_jab = tiles + dd * ntile;
            const double* t_kac = t_jab + nA * nB;
            const double* t_kbc = t_kac + nA * nC;
            if (fj != 0.0) flush_tile(Fd, n, t_jab, nA, nB, ra, rb, fj);
            if (fk != 0.0) {
                flush_tile(Fd, n, t_kac, nA, nC, ra, rc, fk);
                flush_tile(Fd, n, t_kbc, nB, nC, rb, rc, fk);
            }
            for (int sgi = 0; sgi < (two ? 2 : 1); sgi++) {
                const int nDs = sgi ? nD1 : nD0;
                const int* rds = sgi ? rd1 : rd0;
                const double* t_jcd = tiles + dd * ntile + nshared + (sgi ? nseg0 : 0);
                const double* t_kad = t_jcd + nC * nDs;
                const double* t_kbd = t_kad + nA * nDs;
                if (fj != 0.0) flush_tile(Fd, n, t_jcd, nC, nDs, rc, rds, fj);
                if (fk != 0.0) {
                    flush_tile(Fd, n, t_kad, nA, nDs, ra, rds, fk);
                    flush_tile(Fd, n, t_kbd, nB, nDs, rb, rds, fk);
                }
            }
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
                                a.ket_prims + kp.prim_start, active ? kp.nprim : 0, kp.abx, kp.aby, kp.abz, tab, 0.0, I);
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
                                active ? pp.nprim : 0, pp.abx, pp.aby, pp.abz, tab, 0.0, I);
    double mx = 0.0;
#pragma unroll
    for (int k = 0; k < nsph; k++) mx = (mx < I[k]) ? I[k] : mx;
    if (active) a.maxval[pi] = mx;
}

// kernel: fock_kernel (ket-stationary) or fock_kernel_bs (bra-stationary);
// the erf(omega r)/r instantiation runs when FockScalars::omega2 > 0
#define VLX_INSTANTIATE_CLASS(la, lb, lc, ld, kernel)                                                   \
    void launch_fock_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s)                       \
    {                                                                                                   \
        if (a.sc.omega2 > 0.0)                                                                          \
            hipLaunchKernelGGL((kernel<la, lb, lc, ld, true>), dim3(grid), dim3(VLX_BLOCK), 0, s, a.bra, \
                               a.bra_prims, a.ket, a.ket_prims, a.items, a.boys_tab, a.D, a.F, a.mols, \
                               a.glim, a.sc);                                                           \
        else                                                                                            \
            hipLaunchKernelGGL((kernel<la, lb, lc, ld, false>), dim3(grid), dim3(VLX_BLOCK), 0, s,      \
                               a.bra, a.bra_prims, a.ket, a.ket_prims, a.items, a.boys_tab, a.D, a.F,   \
                               a.mols, a.glim, a.sc);                                                   \
    }                                                                                                   \
    void launch_eri_##la##lb##lc##ld(const EriArgs& a, hipStream_t s)                                   \
    {                                                                                                   \
        const int grid = (a.nquartets + VLX_BLOCK - 1) / VLX_BLOCK;                                     \
        hipLaunchKernelGGL((eri_out_kernel<la, lb, lc, ld>), dim3(grid), dim3(VLX_BLOCK), 0, s, a);     \
    }

// LDS-tiled launcher (dynamic LDS = a.tile_lds_bytes); grid-stride over items
// Tiled kernels: NWB waves per block (the item's bras split over them; the
// LDS tiles are shared), or 4 waves when FockArgs::tiled_waves == 4.  The
// FP64 launcher picks the erf(omega r)/r instantiation when omega2 > 0; the
// FP32 launcher has the Coulomb one only (range-separated builds run no FP32
// items).
#define VLX_TILED_LAUNCH(la, lb, lc, ld, R, NW, RS, tab, MULTI, DLDS)                                   \
    hipLaunchKernelGGL((fock_kernel_tiled<la, lb, lc, ld, R, NW, RS, MULTI, DLDS>), dim3(grid),           \
                       dim3((NW) * VLX_WAVE), a.tile_lds_bytes, s, a.bra, a.bra_prims, a.ket, a.ket_prims,   \
                       a.titems, a.kentries, tab, a.D, a.F, a.mols, a.k32, a.glim, a.grp_off, a.grp_n,      \
                       a.grp_ao, a.dfac, a.sc)
#define VLX_INSTANTIATE_TILED(la, lb, lc, ld, NWB)                                                      \
    void launch_tiled_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s)                      \
    {                                                                                                   \
        const bool w4 = a.tiled_waves == 4 || (NWB) == 4;                                               \
        if (a.sc.omega2 > 0.0) {                                                                        \
            if (w4) VLX_TILED_LAUNCH(la, lb, lc, ld, double, 4, true, a.boys_tab, false, false);               \
            else VLX_TILED_LAUNCH(la, lb, lc, ld, double, NWB, true, a.boys_tab, false, false);                \
        } else {                                                                                        \
            if (w4) VLX_TILED_LAUNCH(la, lb, lc, ld, double, 4, false, a.boys_tab, false, false);              \
            else VLX_TILED_LAUNCH(la, lb, lc, ld, double, NWB, false, a.boys_tab, false, false);               \
        }                                                                                               \
    }                                                                                                   \
    void launch_tiled32_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s)                    \
    {                                                                                                   \
        if (a.tiled_waves == 4 || (NWB) == 4) VLX_TILED_LAUNCH(la, lb, lc, ld, float, 4, false, a.boys_tab_f, false, false); \
        else VLX_TILED_LAUNCH(la, lb, lc, ld, float, NWB, false, a.boys_tab_f, false, false);                  \
    }                                                                                                   \
    void launch_tiled_multi_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s)                \
    {                                                                                                   \
        if (a.sc.omega2 > 0.0) VLX_TILED_LAUNCH(la, lb, lc, ld, double, NWB, true, a.boys_tab, true, false);   \
        else VLX_TILED_LAUNCH(la, lb, lc, ld, double, NWB, false, a.boys_tab, true, false);                    \
    }                                                                                                   \
    void launch_tiled_dlds_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s)                 \
    {                                                                                                   \
        if (a.sc.omega2 > 0.0) VLX_TILED_LAUNCH(la, lb, lc, ld, double, NWB, true, a.boys_tab, false, true);   \
        else VLX_TILED_LAUNCH(la, lb, lc, ld, double, NWB, false, a.boys_tab, false, true);                    \
    }

// multi-density form of a class's untiled kernels (work-item layout in FockScalars::bs_layout)
#define VLX_INSTANTIATE_MULTI(la, lb, lc, ld)                                                           \
    void launch_fock_multi_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s)                 \
    {                                                                                                   \
        if (a.sc.omega2 > 0.0)                                                                          \
            hipLaunchKernelGGL((fock_kernel_multi<la, lb, lc, ld, true>), dim3(grid), dim3(VLX_BLOCK), 0, \
                               s, a.bra, a.bra_prims, a.ket, a.ket_prims, a.items, a.boys_tab, a.D, a.F, \
                               a.mols, a.dfac, a.sc);                                                   \
        else                                                                                            \
            hipLaunchKernelGGL((fock_kernel_multi<la, lb, lc, ld, false>), dim3(grid), dim3(VLX_BLOCK), 0, \
                               s, a.bra, a.bra_prims, a.ket, a.ket_prims, a.items, a.boys_tab, a.D, a.F, \
                               a.mols, a.dfac, a.sc);                                                   \
    }

// ket-stationary kernel with LDS accumulators (NW waves per block); `grid` counts wavefronts
#define VLX_INSTANTIATE_KSL(la, lb, lc, ld, NW)                                                         \
    void launch_ksl_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s)                        \
    {                                                                                                   \
        constexpr int lds = ((2 * la + 1) * (2 * lc + 1) + (2 * la + 1) * (2 * ld + 1) +                 \
                             (2 * lc + 1) * (2 * ld + 1)) * (NW) * VLX_WAVE * 8;                        \
        const int blocks = (grid + (NW) - 1) / (NW);                                                    \
        if (a.sc.omega2 > 0.0)                                                                          \
            hipLaunchKernelGGL((fock_kernel_ksl<la, lb, lc, ld, true, NW>), dim3(blocks), dim3((NW) * VLX_WAVE), \
                               lds, s, a.bra, a.bra_prims, a.ket, a.ket_prims, a.items, a.boys_tab, a.D, \
                               a.F, a.mols, a.glim, a.sc);                                              \
        else                                                                                            \
            hipLaunchKernelGGL((fock_kernel_ksl<la, lb, lc, ld, false, NW>), dim3(blocks),              \
                               dim3((NW) * VLX_WAVE), lds, s, a.bra, a.bra_prims, a.ket, a.ket_prims,   \
                               a.items, a.boys_tab, a.D, a.F, a.mols, a.glim, a.sc);                    \
    }

#define VLX_INSTANTIATE_DIAG(la, lb)                                                                    \
    void launch_diag_##la##lb(const DiagArgs& a, hipStream_t s)                                         \
    {                                                                                                   \
        const int grid = (a.npairs + VLX_BLOCK - 1) / VLX_BLOCK;                                        \
        hipLaunchKernelGGL((diag_kernel<la, lb>), dim3(grid), dim3(VLX_BLOCK), 0, s, a);                \
    }
