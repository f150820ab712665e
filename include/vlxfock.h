/*
 * C ABI of the MI355X-native two-electron Fock build (libvlxfock.so).
 *
 * Plain pointers and sizes only.  Every function returns 0 on success and a
 * negative code on failure; vlx_last_error() then describes the failure
 * (thread-local).  Matrices are dense nao x nao, row-major, float64, in the
 * reference AO order (L-major, then spherical component, then function;
 * src/orbdata/MolecularBasis.cpp:400-410,452-470).
 */
#ifndef VLXFOCK_H
#define VLXFOCK_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vlx_context vlx_context;
typedef struct vlx_screener vlx_screener;

enum {
    VLX_OK = 0,
    VLX_ERR_ARG = -1,      /* bad argument (shape, label, threshold, null pointer) */
    VLX_ERR_UNSUPPORTED = -2, /* angular momentum or operator not built */
    VLX_ERR_HIP = -3,      /* HIP runtime error */
    VLX_ERR_NOMEM = -4
};

/* Last error message of the calling thread (never NULL). */
const char* vlx_last_error(void);

/* Library version string. */
const char* vlx_version(void);

/* Highest angular momentum per shell the built kernels cover. */
int vlx_max_angular_momentum(void);

/* Device context: HIP device, stream, Boys table in HBM.
 * Replaces the implicit OpenMP state of CFockDriver
 * (src/t4c_func/FockDriver.hpp:36-60). */
int vlx_context_create(int device, vlx_context** out);
void vlx_context_destroy(vlx_context* ctx);

/* Block-size factor of the reference's task partitioning
 * (src/t4c_func/FockDriver.cpp:356-376).  Accepted and stored for interface
 * compatibility; the GPU schedule does not depend on it. */
int vlx_set_block_size_factor(vlx_context* ctx, int factor);

/*
 * Precision of the recursions (no reference counterpart: the reference
 * computes every quartet in FP64).  VLX_PRECISION_FP64 (default) keeps the
 * FP64 path.  VLX_PRECISION_MIXED runs the Boys function and recursions of a
 * quartet in FP32 when its Schwarz bound times max|D| of the build is below
 * `threshold` (decided per tile of kets at the screener's decimal bins:
 * bin_ab + bin_cd >= ceil(2 + log10(max|D| / threshold))), for the classes
 * with the LDS-tiled kernel (la+lb+lc+ld <= 4); J/K accumulation stays FP64.
 * Range-separated labels (*_rs) always run in FP64.
 */
enum { VLX_PRECISION_FP64 = 0, VLX_PRECISION_MIXED = 1 };
int vlx_set_precision(vlx_context* ctx, int mode, double threshold);
/*
 * Density-weighted screening (no reference counterpart: a fast mode whose
 * error is stated against the FP64 build).  threshold > 0: a quartet
 * (ab|cd) is skipped when its Schwarz bound 10^(2 - bin_ab - bin_cd) times
 * the largest |D| of the six D blocks its J and K terms read (D_cd, D_ab,
 * D_bd, D_bc, D_ad, D_ac, taken per pair of the screener's atom groups) is
 * below threshold; 0 (default) turns it off, and the build is the
 * reference's again.  Applies to single-molecule builds with ithreshold >= 0
 * (not to batched builds, nor to vlx_fock_compute_device_noalloc, which
 * refuses it).  With a difference density (incremental SCF builds) the
 * bound follows |Delta D| block by block.
 */
int vlx_set_density_screening(vlx_context* ctx, double threshold);
/*
 * Deterministic reduction mode (no reference counterpart: the reference
 * merges task-local Fock blocks under an OpenMP critical section,
 * src/t4c_func/T4CMatrixDistributor.cpp:142, whose order varies from run to
 * run; this mode is for diffing against the atomic path).  enable != 0: every
 * later build of the context (single, general and multi-density) runs the
 * direct-digestion kernels and adds each F' contribution as a 128-bit
 * fixed-point integer (resolution 2^-96) with integer atomics, whose sums do
 * not depend on their order: the same bits on every run and for any number of
 * streams.  FP64 with the Schwarz selection only (the mixed-precision and
 * density-weighted modes do not apply); slower than the default kernels.
 * Batched builds, the no-allocation build and the autotuner refuse it.
 */
int vlx_set_deterministic(vlx_context* ctx, int enable);
/* Bin-sum threshold the last build used (99 = no FP32 work). */
int vlx_context_last_fp32_bin_sum(vlx_context* ctx, int* k32);
/* Primitive slots of one build run in FP64 and in FP32 at bin-sum threshold k32. */
int vlx_fock_precision_split(vlx_screener* scr, int ithreshold, int k32, long long* prim_slots_fp64,
                             long long* prim_slots_fp32);
/* The FP32 primitive slots of the same split per angular class: *nrecords
 * records of (la, lb, lc, ld) (4 ints each) and their FP32 primitive slots;
 * at most max_records are written. */
int vlx_fock_precision_classes(vlx_screener* scr, int ithreshold, int k32, int max_records, int* nrecords,
                               int* classes, long long* prim_slots_fp32);

/*
 * T4CScreener::partition(basis, molecule, "eri")
 * (src/t4c_func/T4CScreener.cpp:63-118, src/orbdata/BlockedGtoPairBlock.cpp:22-152).
 * Builds all shell pairs, computes max over components of (ab|ab) on the GPU,
 * bins them by -floor(log10 sqrt(max)) and drops pairs with sqrt <= 1e-15.
 *
 * Shell table (one entry per contracted shell, atom order then file order):
 *   l[s], nprim[s], prim_offset[s] into exps/norms (normalised contraction
 *   coefficients), centers[3 s .. 3 s + 2] in bohr, ao_base[s] = AO index of
 *   component 0, ao_stride[s] = AO distance between components (n_L).
 */
int vlx_screener_create(vlx_context* ctx, int nshells, const int* l, const int* nprim, const int* prim_offset,
                        const double* exps, const double* norms, const double* centers, const int* ao_base,
                        const int* ao_stride, int nao, vlx_screener** out);
void vlx_screener_destroy(vlx_screener* scr);

/* Pair maxima max_comp (ab|ab) as used for binning, nshells x nshells
 * (symmetric), for comparison with CT4CScreener's values. */
int vlx_screener_pair_maxima(const vlx_screener* scr, double* out);

/* Number of kept (binned) shell pairs. */
int vlx_screener_number_of_pairs(const vlx_screener* scr, long long* out);

/*
 * CFockDriver::compute(screener, D, label, exchange_factor, omega, ithreshold)
 * (src/t4c_func/FockDriver.cpp:120-171; rank/nodes variant :234-289).
 * label: "2jk" (2J - K), "2jkx" (2J - xf K), "j" (J), "k" (K), "kx" (xf K),
 * and the range-separated "j_rs", "k_rs", "kx_rs" over erf(omega r)/r
 * (omega = 0 gives a zero matrix there; omega is ignored by the other labels).
 * ithreshold >= 0 applies the binned Schwarz selection bin_bra + bin_ket <=
 * ithreshold; ithreshold < 0 computes every shell quartet (the reference's
 * unscreened _compute_fock_omp(basis, molecule, ...) path).  D must be
 * symmetric (vlx_fock_compute_general takes any D).
 * Host buffers: density in, G out (both nao x nao).
 */
int vlx_fock_compute(vlx_context* ctx, vlx_screener* scr, const double* density, const char* label,
                     double exchange_factor, double omega, int ithreshold, double* fock_out);

/* Same as vlx_fock_compute on device buffers already resident in HBM; runs on
 * the context's stream and returns without synchronising. */
int vlx_fock_compute_device(vlx_context* ctx, vlx_screener* scr, const double* d_density, const char* label,
                            double exchange_factor, double omega, int ithreshold, double* d_fock_out);

/*
 * vlx_fock_compute_device without any allocation (no reference counterpart:
 * it is for callers that cannot allocate device memory, such as code under
 * HIP graph capture or tight serving loops).  The F' accumulator and the
 * max|D| slot of a mixed build live in a file-scope device array of the
 * library (VLX_NOALLOC_MAX_NAO^2 doubles + 256 B, 72 MiB per device, held
 * until the process ends), so nao must not exceed VLX_NOALLOC_MAX_NAO
 * (VLX_ERR_ARG otherwise).  The work schedule of (screener, ithreshold,
 * precision split) must exist: one vlx_fock_compute_device call with the
 * same screener and settings (and, in mixed mode, a density of the same
 * max|D| decade) comes first, else VLX_ERR_ARG.  Per-class timing is not
 * recorded.  Every call on one device shares the array: calls that overlap
 * on one device (two contexts, or two host threads) race and must be
 * serialised by the caller.  Same arguments, results and stream behaviour
 * as vlx_fock_compute_device.
 */
#define VLX_NOALLOC_MAX_NAO 3072
int vlx_fock_compute_device_noalloc(vlx_context* ctx, vlx_screener* scr, const double* d_density, const char* label,
                                    double exchange_factor, double omega, int ithreshold, double* d_fock_out);

/*
 * Many molecules in one call (no reference counterpart: the reference builds
 * one molecule per CFockDriver::compute call, src/t4c_func/FockDriver.cpp:120-171).
 * Molecule m has its own screener scrs[m] (same context), symmetric density
 * d_densities[m] and output d_focks[m] (device buffers, nao_m x nao_m); label,
 * exchange_factor, omega and ithreshold as in vlx_fock_compute, in the
 * context's precision mode (mixed: each molecule's FP32 split from its own
 * max|D|, decided on the device).  The molecules' pair lists are concatenated
 * per angular class, so that one kernel launch per (class, kernel variant)
 * covers the whole batch; the build (density gather, F' clear, precision
 * split, class launches, symmetrisation into d_focks) is captured into a HIP
 * graph on the first call with these screeners, label, thresholds and
 * precision, and replayed by later calls (up to 8 such plans are kept per
 * context; destroying a screener drops the plans that hold it).  A screener
 * may be listed more than once.  At most 65535 molecules.  The d_focks must
 * not overlap.  Runs on the context's stream and returns without
 * synchronising.
 */
int vlx_fock_compute_batch_device(vlx_context* ctx, int nmol, vlx_screener* const* scrs,
                                  const double* const* d_densities, const char* label, double exchange_factor,
                                  double omega, int ithreshold, double* const* d_focks);

/*
 * CFockDriver::compute for a general (non-symmetric) density, the reference's
 * mat_t.general path (src/t4c_func/FockDriver.cpp:120-171 with the general
 * distributors of src/t4c_func/T4CLocalDistributor.cpp; its tests compare K
 * with the transposed k.gen fixture, tests/test_fock_driver.py:320-360).
 * J(D) = J(S) and K(D) = K(S) + K(A) for S = (D + D^T)/2, A = (D - D^T)/2:
 * one build of S plus, for labels with exchange, an exchange-only build of A
 * whose F' enters antisymmetrised.  Same labels and arguments as
 * vlx_fock_compute; host buffers (nao x nao).
 */
int vlx_fock_compute_general(vlx_context* ctx, vlx_screener* scr, const double* density, const char* label,
                             double exchange_factor, double omega, int ithreshold, double* fock_out);
/* The same on device buffers (context stream, no synchronisation). */
int vlx_fock_compute_general_device(vlx_context* ctx, vlx_screener* scr, const double* d_density, const char* label,
                                    double exchange_factor, double omega, int ithreshold, double* d_fock_out);

/*
 * CFockDriver::compute(screener, CMatrices densities, labels, exchange_factor,
 * omega, ithreshold) (src/t4c_func/FockDriver.cpp:173-229 with
 * CT4CMatricesDistributor, src/t4c_func/T4CMatricesDistributor.cpp): one pass
 * over the integrals digests all nmat densities, each with its own label.
 * densities: nmat x nao x nao, back to back; focks_out the same.  general:
 * nmat flags (NULL: all symmetric); a general density costs a second digested
 * density (its antisymmetric part, exchange labels only), not a second pass.
 * The integrals are erf(omega r)/r only when every label is range-separated
 * (CT4CMatricesDistributor::need_omega, :55-66), else 1/r for all, *_rs labels
 * then digesting like j, k, kx.  FP64 with the Schwarz selection; the
 * context's mixed-precision and density-weighted modes do not apply.  Host
 * buffers; symmetric densities are checked like vlx_fock_compute's.
 */
int vlx_fock_compute_multi(vlx_context* ctx, vlx_screener* scr, int nmat, const double* densities, const int* general,
                           const char* const* labels, double exchange_factor, double omega, int ithreshold,
                           double* focks_out);
/* The same on device buffers (context stream, no synchronisation, no symmetry check). */
int vlx_fock_compute_multi_device(vlx_context* ctx, vlx_screener* scr, int nmat, const double* d_densities,
                                  const int* general, const char* const* labels, double exchange_factor, double omega,
                                  int ithreshold, double* d_focks_out);

/* Device memory helpers (HBM of the context's device) for callers that keep
 * D and G resident between builds. */
int vlx_device_malloc(vlx_context* ctx, unsigned long long bytes, void** out);
int vlx_device_free(vlx_context* ctx, void* ptr);
int vlx_copy_to_device(vlx_context* ctx, void* dst, const void* src, unsigned long long bytes);
int vlx_copy_to_host(vlx_context* ctx, void* dst, const void* src, unsigned long long bytes);

/* Synchronise the context's stream. */
int vlx_context_synchronize(vlx_context* ctx);

/* Side streams the class kernels of later builds fan out to (1..4; the
 * VLX_STREAMS environment variable sets the initial value, default 3).  No
 * reference counterpart: one stream serialises the class kernels, for timing
 * each in isolation. */
int vlx_context_set_streams(vlx_context* ctx, int nstreams);

/* Kernel variant every class of later builds runs (no reference counterpart;
 * diagnostics and tests): the autotuner's numbering, 0..5, where a class that
 * cannot run it falls back to its untiled kernel; -1 restores the per-class
 * choice (autotuned, else the default).  Results do not depend on it. */
int vlx_context_set_variant(vlx_context* ctx, int variant);

/* Raw HIP stream of the context (hipStream_t), for event timing by callers. */
void* vlx_context_stream(vlx_context* ctx);

/*
 * Work of one build at ithreshold: one record per angular class
 * (la, lb, lc, ld) with the number of contracted and primitive shell quartets
 * and the FP64 FLOPs of one 2jk build of that class:
 *   prim_quartets (V + 53 + 4 L) + quartets (C + 12 n_sph),
 * V = VRR + contraction FLOPs per primitive quartet and C = HRR + cart->sph
 * FLOPs per contracted quartet of the generated recursion.
 * Writes at most max_records; *nrecords receives the number available.
 */
int vlx_fock_work(vlx_screener* scr, int ithreshold, int max_records, int* nrecords, int* classes /* 4 per */,
                  long long* quartets, long long* prim_quartets, double* flops);

/* Schedule efficiency of one build at ithreshold, per class record in the
 * same order as vlx_fock_work: lane slots (work items x bras x 64 lanes, or
 * per-lane quartets of the cooperative kernel) and primitive slots (the
 * primitive-pair iterations the wavefronts execute, masked lanes included). */
int vlx_fock_schedule_stats(vlx_screener* scr, int ithreshold, int max_records, int* nrecords,
                            long long* lane_slots, long long* prim_slots);

/*
 * Autotuner (no reference counterpart): times every angular class of one
 * 2jk build at ithreshold on `d_density` under each kernel variant
 * (0 = untiled, 1 = LDS-tiled with 4-wave blocks, 2 = LDS-tiled with the
 * class's wide block, 3 = matrix-core McMurchie-Davidson, for classes with
 * la+lb+lc+ld >= 4, 4 and 5 = variants 1 and 2 with one ket primitive per lane
 * instead of runs of up to three, 6 = ket-stationary work items with lane-private
 * LDS accumulators for the bra-stationary classes, 7 = variant 2 with the item's
 * D blocks staged in LDS, 8 = variant 2 with chunks spanning two ket segments),
 * `repeats` builds each on a
 * single stream, and keeps
 * the fastest variant per class for later builds with this screener (and as
 * the default of screeners created later on this context).
 * Optional outputs: *nchosen records of classes (4 ints each) and variants
 * (arrays of at least 64 classes).  Results do not depend on the variant;
 * a variant whose warm-up G disagrees with the untiled one is not chosen.
 */
int vlx_fock_autotune(vlx_context* ctx, vlx_screener* scr, const double* d_density, int ithreshold, int repeats,
                      int* nchosen, int* classes, int* variants);
/* The per-class kernel variants a screener uses (set by vlx_fock_autotune or
 * vlx_fock_set_variants): *nrecords records of classes (4 ints), variants and
 * the measured milliseconds of each class under its variant (0 when not
 * measured); at most max_records are written.  No reference counterpart. */
int vlx_fock_get_variants(vlx_screener* scr, int max_records, int* nrecords, int* classes, int* variants, double* ms);
/* Applies saved autotuner choices to a screener (e.g. from another process of
 * the same workload): variants per class and, optionally (ms may be NULL),
 * their measured times, which balance the classes over the side streams.
 * Classes not listed use the default rule.  No reference counterpart. */
int vlx_fock_set_variants(vlx_context* ctx, vlx_screener* scr, int nrecords, const int* classes, const int* variants,
                          const double* ms);
/* Variants the last vlx_fock_autotune on this context rejected (bit v set for
 * variant v): the G of a variant's warm-up build is checked on the host against
 * the untiled variant's (non-finite elements, or a difference above 1e-10), and
 * a rejected variant is never chosen.  No reference counterpart. */
int vlx_fock_autotune_rejected(vlx_context* ctx, int* mask);

/* Kernel (template name as profilers report it) that a class (la lb|lc ld)
 * runs with at ithreshold, written to name[0..len). */
int vlx_fock_class_kernel(vlx_context* ctx, vlx_screener* scr, int ithreshold, int la, int lb, int lc, int ld,
                          char* name, int len);

/* Per-class kernel timing with HIP events on the context's stream
 * (off by default).  vlx_context_class_times synchronises the stream and
 * returns, per class (la, lb, lc, ld), the summed kernel time in ms and the
 * number of launches since the last reset. */
int vlx_context_set_timing(vlx_context* ctx, int enable);
int vlx_context_reset_times(vlx_context* ctx);
int vlx_context_class_times(vlx_context* ctx, int max_records, int* nrecords, int* classes /* 4 per */,
                            double* total_ms, long long* launches);

/*
 * Spherical integral blocks (ab|cd) for explicit shell quartets
 * (a, b, c, d) = quartets[4 q .. 4 q + 3]; block q is written to out at the
 * offset sum of the previous block sizes, index ((i nb + j) nc + k) nd + l.
 * Used by compute_eri (src/t4c_func/FockDriver.cpp:76-117) and by tests.
 */
int vlx_eri_quartets(vlx_context* ctx, vlx_screener* scr, int nquartets, const int* quartets, double* out);

/*
 * One-electron matrices of the SCF loop around the Fock build, host-side
 * (O(N^2) work, threaded): overlap S, kinetic energy T and nuclear
 * attraction V = -sum_C Z_C <a|1/|r - C||b> (electron charge included), each
 * dense nao x nao row-major in the AO order of the shell table (same table as
 * vlx_screener_create).  Replaces COverlapDriver::compute
 * (src/t2c_overlap/OverlapDriver.hpp:50), CKineticEnergyDriver::compute
 * (src/t2c_kinetic_energy/KineticEnergyDriver.hpp) and
 * compute_nuclear_potential_integrals (src/pymodule/oneeints.py:71-95).
 */
int vlx_one_electron(int nshells, const int* l, const int* nprim, const int* prim_offset, const double* exps,
                     const double* norms, const double* centers, const int* ao_base, const int* ao_stride, int nao,
                     int natoms, const double* charges, const double* atom_xyz, double* S, double* T, double* V);

#ifdef __cplusplus
}
#endif

#endif /* VLXFOCK_H */
