// Data layout shared by the host driver and the gfx950 kernels.
//
// Shell pairs of one pair class (la >= lb) live in HBM as an array of
// PairInfo (48 B) plus an array of PrimPair (one 64 B line per primitive
// pair, so a lane's primitive data is a single cache-line read).  Pairs of a
// class are sorted by (Schwarz bin, primitive count) so that the kets a bra
// pair meets under bin_bra + bin_ket <= ithreshold form a prefix.
#pragma once

#include <hip/hip_runtime.h>

#include "vlx_boys.h"

// Sets the calling thread's message of vlx_last_error() (every translation unit of the library
// reports through it).
void vlx_set_last_error(const char* msg);

// Wavefront width and the block shape of the untiled Fock kernels (4 waves).
constexpr int VLX_WAVE = 64;
constexpr int VLX_BLOCK = 256;
constexpr int VLX_WAVES_PER_BLOCK = VLX_BLOCK / VLX_WAVE;

// One primitive pair (a on centre A, b on centre B), A = first shell of the pair.
struct alignas(64) PrimPair {
    double zeta;          // a + b
    double px, py, pz;    // P = (a A + b B) / zeta
    double pax, pay, paz; // P - A
    double k;             // sqrt(2) pi^(5/4) N_a N_b exp(-a b |AB|^2 / zeta) / zeta
};

// One contracted shell pair (A, B) with l_A >= l_B (80 B).
struct alignas(16) PairInfo {
    int ao_a, str_a;  // AO index of component 0 of A and stride between components
    int ao_b, str_b;
    int prim_start;   // first PrimPair of this pair
    int nprim;        // number of primitive pairs
    int same;         // 1 when A and B are the same shell
    int bin;          // Schwarz bin (0..15)
    int rank;         // position in the ket ordering of its pair class
    int shell_a;      // shell index of A
    int grp_a, grp_b; // atom groups of A and B
    int loc_a, loc_b; // local AO index of component 0 within (group, l) blocks
    double abx, aby, abz;  // A - B
};

// Tiled Fock work: bras [i0, i1) of atom groups (ga, gb) against ket
// primitive entries [j0, j1) (at most 64, one per lane) of atom groups (gc, gd),
// of molecule `mol` of the launch (MolDesc).  bin_lo is the smallest Schwarz
// bin of the item's kets (the precision split on the device, FockArgs::k32).
// A chunk may span two ket segments with the same C group: entries [jm, j1)
// then belong to D group gd2 (gd2 = -1 and jm = j1 for one segment).
struct TileItem {
    int i0, i1, j0, j1;
    int ga, gb, gc, gd;
    int mol, bin_lo, gd2, jm;
};

// One unit of Fock work: kets [j0, j0 + 64) of the ket list (one per lane,
// lanes at or beyond jend idle) against bras [i0, i1) of the bra list, of
// molecule `mol` of the launch.
struct WorkItem {
    int j0, i0, i1, mol;
    int jend, pad0, pad1, pad2;
};

// Where one molecule's D and F' live: nao x nao row-major at element offset
// `base` of the launch's D and F arrays.  A single-molecule build has one
// entry {0, nao}; a batched build packs its molecules back to back.
struct MolDesc {
    long long base;
    int nao, pad;
};

// Scalar parameters of one Fock class launch.
struct FockScalars {
    long long nitems;
    int nket;
    int ithreshold;                // < 0: no Schwarz selection
    int same_class;                // bra and ket lists hold the same pairs (quartets with rank_bra >= rank_ket)
    int nao;
    int ngroups;                   // atom groups of the molecule (density-weighted screening table side)
    double fj;                     // coefficient of the Coulomb terms (0: none)
    double fk;                     // coefficient of the exchange terms (0: none)
    double omega2;                 // omega^2 of the erf(omega r)/r operator (range-separated labels), 0: 1/r
    int ndens;                     // densities digested per quartet (multi-density kernels; 0 or 1: one)
    int det;                       // 1: deterministic fixed-point accumulation (multi-density and MD kernels)
    int bs_layout;                 // multi-density kernel: 1 bra-stationary work items, 0 ket-stationary
    int pad_sc;
    long long dstride;             // elements between consecutive densities (and F' accumulators)
};

struct FockArgs {
    const PairInfo* bra;           // bra list, sorted by (bin, shell A, primitive count)
    const PrimPair* bra_prims;
    const PairInfo* ket;           // ket list, sorted by (bin, primitive count)
    const PrimPair* ket_prims;
    const WorkItem* items;
    const TileItem* titems;        // tiled variant
    const int2* kentries;          // tiled variant: (ket pair, first primitive | count << 16) per lane
    const int* grp_off;            // (ngroups x 4): offset of the (group, l) AO list in grp_ao
    const int* grp_n;              // (ngroups x 4): number of AOs in the (group, l) block
    const int* grp_ao;             // global AO index of each local AO
    int tile_lds_bytes;            // dynamic LDS of the tiled variant
    int tiled_waves;               // 4: force 4-wave blocks for the tiled kernel; else the class default
    const double* boys_tab;        // BOYS_TAB_SIZE doubles for L = la+lb+lc+ld
    const float* boys_tab_f;       // the same table in FP32 (mixed-precision tiled kernel)
    const double* bra_u;           // matrix-core MD kernel: Hermite-to-spherical U of every bra primitive pair
    const double* ket_u;           // ... and of every ket primitive pair
    const double* D;               // densities of the launch's molecules (MolDesc layout)
    double* F;                     // F' accumulators (G = F' + F'^T), same layout
    const MolDesc* mols;           // per molecule: offset and size of its D and F'
    const int* glim;               // density-weighted screening: ngroups x ngroups largest bin sums per
                                   // pair of atom groups (from max|D| of the block), or nullptr (off)
    const double* dfac;            // multi-density kernels: (fj, fk) of each density, 2 x ndens doubles
    const int* k32;                // per molecule: FP32 bin-sum threshold on the device, or nullptr
                                   // (tiled items: the FP64 kernel runs the item's bras with
                                   // bin + bin_lo < k32, the FP32 kernel the others)
    FockScalars sc;
};

// Hermite functions of total degree <= l, and the padded U array of a pair
// class (la, lb) of the matrix-core MD kernel: rows = Hermite index, columns
// = spherical index of the pair, both padded to multiples of 16.
__host__ __device__ constexpr int vlx_nherm(int l) { return (l + 1) * (l + 2) * (l + 3) / 6; }
__host__ __device__ constexpr int vlx_pad16(int x) { return (x + 15) / 16 * 16; }
__host__ __device__ constexpr int vlx_u_rows(int la, int lb) { return vlx_pad16(vlx_nherm(la + lb)); }
__host__ __device__ constexpr int vlx_u_cols(int la, int lb) { return vlx_pad16((2 * la + 1) * (2 * lb + 1)); }

// U arrays ([prim][vlx_u_rows][vlx_u_cols]) of the npairs pairs of one pair set.
void vlx_launch_u_build(const PairInfo* pairs, const PrimPair* prims, int npairs, int la, int lb, double* u,
                        hipStream_t s);

struct EriArgs {
    const PairInfo* bra;
    const PrimPair* bra_prims;
    const PairInfo* ket;
    const PrimPair* ket_prims;
    const int2* quartets;  // (bra pair, ket pair)
    int nquartets;
    const double* boys_tab;
    double* out;           // nquartets x nsph
};

struct DiagArgs {
    const PairInfo* pairs;
    const PrimPair* prims;
    int npairs;
    const double* boys_tab;
    double* maxval;
};

using FockLaunchFn = void (*)(const FockArgs&, int, hipStream_t);
using EriLaunchFn = void (*)(const EriArgs&, hipStream_t);
using DiagLaunchFn = void (*)(const DiagArgs&, hipStream_t);

enum { VLX_BRA_STATIONARY = 0, VLX_KET_STATIONARY = 1, VLX_WAVE_MD = 2, VLX_TILED = 3 };

struct ClassLaunchers {
    int x, y;             // bra / ket pair-class indices (x >= y)
    int la, lb, lc, ld;
    int vrr_flops;        // generated VRR + contraction FLOPs per primitive quartet
    int hrr_flops;        // generated HRR + cart->sph FLOPs per contracted quartet
    int variant;          // VLX_BRA_STATIONARY, VLX_KET_STATIONARY or VLX_WAVE_MD (work-item layout)
    FockLaunchFn fock;
    EriLaunchFn eri;
    FockLaunchFn tiled;   // LDS-tiled Fock kernel, or nullptr
    FockLaunchFn tiled32; // the tiled kernel with FP32 recursions (mixed precision), or nullptr
    FockLaunchFn mmd;     // matrix-core McMurchie-Davidson Fock kernel (MD work items), or nullptr
    FockLaunchFn fock_multi;   // multi-density form of `fock` (same work items)
    FockLaunchFn tiled_multi;  // multi-density form of `tiled` (FP64), or nullptr
    FockLaunchFn ksl;          // ket-stationary kernel with LDS accumulators (grid in wavefronts), or nullptr
    FockLaunchFn tiled_dlds;   // `tiled` (FP64, wide block) with the item's D blocks staged in LDS, or nullptr
};

struct DiagLauncher {
    int x;
    int la, lb;
    DiagLaunchFn diag;
};

#define VLX_DECLARE_CLASS(la, lb, lc, ld)                                        \
    void launch_fock_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s); \
    void launch_eri_##la##lb##lc##ld(const EriArgs& a, hipStream_t s);
#define VLX_DECLARE_TILED(la, lb, lc, ld)                                          \
    void launch_tiled_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s); \
    void launch_tiled32_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s); \
    void launch_tiled_multi_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s); \
    void launch_tiled_dlds_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s);
#define VLX_DECLARE_MULTI(la, lb, lc, ld) \
    void launch_fock_multi_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s);
#define VLX_DECLARE_KSL(la, lb, lc, ld) void launch_ksl_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s);
#define VLX_DECLARE_DIAG(la, lb) void launch_diag_##la##lb(const DiagArgs& a, hipStream_t s);
#define VLX_DECLARE_MMD(la, lb, lc, ld) void launch_mmd_##la##lb##lc##ld(const FockArgs& a, int grid, hipStream_t s);
