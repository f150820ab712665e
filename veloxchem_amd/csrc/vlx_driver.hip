// Host driver and C ABI of the MI355X two-electron Fock build.
//
// Replaces CFockDriver::compute (src/t4c_func/FockDriver.cpp:120-289),
// CT4CScreener::partition (src/t4c_func/T4CScreener.cpp:63-118) and the pair
// data model (src/orbdata/GtoPairBlock.cpp:49-305) with a GPU pipeline:
//
//   screener: all shell pairs -> PairInfo/PrimPair in HBM -> diag_kernel
//             (max over components of (ab|ab)) -> bins -> kept pairs sorted by
//             (bin, primitive count) per pair class;
//   build:    F' = 0 -> one fock_kernel launch per class pair (bra class >=
//             ket class) -> G = F' + F'^T.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <queue>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/vlxfock.h"
#include "vlx_types.h"

#include "gen/dispatch.inc"

namespace {

thread_local std::string g_last_error = "";

int fail(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t err_ = (expr);                                                                \
        if (err_ != hipSuccess) return fail(VLX_ERR_HIP, "%s: %s", #expr, hipGetErrorString(err_)); \
    } while (0)

constexpr int NBINS = 16;

// Scratch of vlx_fock_compute_device_noalloc (callers that cannot allocate): the F' accumulator
// (up to VLX_NOALLOC_MAX_NAO^2 doubles) and one 8-byte slot for max|D|, 72 MiB held by every
// device the library runs on for the life of the process.  Host code never passes `scratch`
// itself to a kernel: it takes the device address from hipGetSymbolAddress (g_scratch_ptr).
constexpr size_t SCRATCH_FACC = (size_t)VLX_NOALLOC_MAX_NAO * VLX_NOALLOC_MAX_NAO;
__device__ __align__(256) double scratch[SCRATCH_FACC + 32];
constexpr int MAX_DEVICES = 64;
double* g_scratch_ptr[MAX_DEVICES] = {};  // device address of `scratch` per device (set at context creation)
std::mutex g_scratch_mutex;

int npair_classes() { return (VLX_GEN_LMAX + 1) * (VLX_GEN_LMAX + 2) / 2; }

int pair_class_index(int la, int lb)  // la >= lb; order (0,0),(1,0),(1,1),(2,0),...
{
    return la * (la + 1) / 2 + lb;
}

template <typename T>
struct DevBuf {
    T* ptr = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), n(o.n)
    {
        o.ptr = nullptr;
        o.n = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            ptr = o.ptr;
            n = o.n;
            o.ptr = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count)
    {
        if (count <= n && ptr) return hipSuccess;
        release();
        n = count;
        return hipMalloc(&ptr, std::max<size_t>(count, 1) * sizeof(T));
    }
    hipError_t upload(const std::vector<T>& v)
    {
        hipError_t e = alloc(v.size());
        if (e != hipSuccess) return e;
        if (v.empty()) return hipSuccess;
        return hipMemcpy(ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    }
};

// Non-owning device pointer into an arena (a DevBuf<char> that holds several arrays).
template <typename T>
struct DevView {
    T* ptr = nullptr;
};

}  // namespace

void vlx_set_last_error(const char* msg) { g_last_error = msg ? msg : ""; }

struct PairSet {
    int la = 0, lb = 0;
    std::vector<PairInfo> pairs;
    std::vector<PrimPair> prims;
    std::vector<int> shell_a, shell_b;  // shells of each pair
    std::vector<int> bins;
    int cnt_le[NBINS] = {0};            // pairs with bin <= b
    std::vector<long long> nprim_prefix;
    struct Segment {
        int start, end, ga, gb;
        int cnt_le[NBINS];  // pairs of the segment with bin <= b
    };
    std::vector<Segment> segs;  // tiled order only: runs of equal (group A, group B)
    // tiled ket order: lane entries (pair, first primitive | count << 16) and the entry range
    // of each segment, for runs of up to ket_prims primitives and for single primitives
    struct Entries {
        std::vector<int2> kentries;
        std::vector<int> seg_e0, seg_e1;
        DevView<int2> d_kentries;
    };
    Entries runs, singles;
    // device copies of pairs, prims and both entry lists, in one allocation
    DevBuf<char> d_arena;
    DevView<PairInfo> d_pairs;
    DevView<PrimPair> d_prims;
    DevBuf<double> d_u;   // matrix-core MD kernel: U of every primitive pair (built on first use)
    bool u_ready = false;
    int size() const { return static_cast<int>(pairs.size()); }
};

struct ClassSchedule {
    int entry = -1;  // index into g_class_table
    long long lane_slots = 0, prim_slots = 0;  // schedule efficiency diagnostics
    int stream = 0;  // side stream the class runs on; -1: alone on the main stream before the others
    int variant = 0;      // VARIANT_* the schedule runs
    bool tiled = false;
    bool singles = false; // tiled over single-primitive ket entries
    bool mmd = false;     // matrix-core MD kernel (MD work items)
    bool ksl = false;     // ket-stationary kernel with LDS accumulators (ket-stationary work items)
    bool dlds = false;    // tiled, the item's D blocks staged in LDS (twice the tile LDS)
    int tiled_waves = 0;  // 4: 4-wave blocks; 0: the class's wide block
    int tile_lds = 0;  // bytes of dynamic LDS (tiled)
    std::vector<TileItem> titems;
    DevBuf<TileItem> d_titems;
    std::vector<TileItem> titems32;  // tiled items run with FP32 recursions (mixed precision)
    DevBuf<TileItem> d_titems32;
    long long prim_slots32 = 0;      // primitive slots of the FP32 items
    std::vector<WorkItem> items;
    long long quartets = 0, prim_quartets = 0;
    DevBuf<WorkItem> d_items;
};

struct vlx_screener;

// One launch of a batched build: the work items of one (class, kernel variant) of every
// molecule of the batch, their indices moved into the batch's concatenated pair lists.
struct BatchLaunch {
    int entry = -1;       // index into g_class_table
    int variant = 0, tiled_waves = 0, tile_lds = 0, stream = 0;
    bool tiled = false, singles = false, mmd = false, ksl = false;
    std::vector<TileItem> titems;
    std::vector<WorkItem> items;
    DevBuf<TileItem> d_titems;
    DevBuf<WorkItem> d_items;
    double cost = 0.0;    // estimated FLOPs (stream balance, launch order)
};

// A batched build of many molecules.  Per pair class and order (ket, bra, tiled ket, tiled
// bra) the molecules' pair lists are concatenated, with primitive ranges, ranks (positions in
// the ket orders), atom groups and tiled lane entries moved by each molecule's offsets.  Work
// items move the same way and name their molecule, so one launch per (class, variant) covers
// the batch; D and F' of molecule m sit at offset mols[m].base of packed arrays.  The whole
// build (density gather, F' clear, precision split, class launches on the side streams,
// symmetrisation into the callers' outputs) is captured once into a HIP graph and replayed by
// later calls with the same screeners and settings.
struct BatchPlan {
    std::vector<vlx_screener*> scrs;
    std::vector<int> epochs;                 // the screeners' variant epochs the plan was built at
    int ithreshold = 0, force_variant = -1, nstreams = 0;
    double fj = 0.0, fk = 0.0, omega2 = 0.0; // label factors (kernel arguments of the graph)
    double tau = 0.0;                        // mixed precision threshold; 0: FP64
    std::vector<PairSet> ket, bra, tket, tbra;  // concatenated pair sets per pair class
    std::vector<int> grp_off, grp_n, grp_ao;
    DevBuf<int> d_grp_off, d_grp_n, d_grp_ao;
    std::vector<MolDesc> mols;
    DevBuf<MolDesc> d_mols;
    size_t total = 0;                        // elements of the packed D and F'
    int max_nao = 0;
    DevBuf<double> d_D, d_F;
    std::vector<const double*> h_src;        // the callers' densities and outputs of the last call
    std::vector<double*> h_dst;
    DevBuf<const double*> d_src;
    DevBuf<double*> d_dst;
    DevBuf<unsigned long long> d_dmax;       // mixed: max|D| per molecule (bit patterns)
    DevBuf<int> d_k32;                       // mixed: FP32 bin-sum threshold per molecule
    std::vector<std::unique_ptr<BatchLaunch>> launches;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    BatchPlan() = default;
    BatchPlan(const BatchPlan&) = delete;
    BatchPlan& operator=(const BatchPlan&) = delete;
    ~BatchPlan()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

constexpr int BRA_CHUNK = 128;  // default upper bound of bras per work item

constexpr int NSIDE = 4;  // side streams that run class kernels concurrently

constexpr int NO_FP32 = 99;  // bin-sum threshold that sends no quartet to FP32

// Kernel variant of a class: untiled (generated ket-/bra-stationary kernel),
// tiled with 4-wave blocks, tiled with the class's wide block, matrix-core MD,
// and the two tiled forms over single-primitive ket entries.  Classes without
// generated recursions run the matrix-core MD kernel only.
enum {
    VARIANT_UNTILED = 0,
    VARIANT_TILED4 = 1,
    VARIANT_TILED_WIDE = 2,
    VARIANT_MMD = 3,
    VARIANT_TILED4_P1 = 4,
    VARIANT_TILED_WIDE_P1 = 5,
    VARIANT_KSL = 6,  // ket-stationary items with lane-private LDS accumulators (bra-stationary classes)
    VARIANT_TILED_DLDS = 7,  // tiled wide block with the item's D blocks staged in LDS
    VARIANT_TILED_2SEG = 8,  // tiled wide block, chunks spanning two ket segments of one C group
    NVARIANTS = 9
};
bool variant_tiled(int v)
{
    return v == VARIANT_TILED4 || v == VARIANT_TILED_WIDE || v == VARIANT_TILED4_P1 || v == VARIANT_TILED_WIDE_P1 ||
           v == VARIANT_TILED_DLDS || v == VARIANT_TILED_2SEG;
}

struct vlx_context {
    int device = 0;
    // side streams in use (VLX_STREAMS, 1..NSIDE): three measured faster than four on the 64-water
    // def2-TZVP build (0.852 against 0.815 builds/s) and the 32-water def2-SVP build (45.7 against 43.8)
    int nstreams = 3;
    int max_chunk = BRA_CHUNK;      // bras per work item upper bound (VLX_BRA_CHUNK)
    int min_items = 16384;          // work items a class launch aims for (VLX_MIN_ITEMS)
    int use_tiled = 1;              // LDS-tiled kernels where built (VLX_TILED)
    int tiled_lmax = 5;             // largest la+lb+lc+ld that may use the tiled kernel (VLX_TILED_LMAX)
    int force_variant = -1;         // autotuning: every class runs this variant (VARIANT_*)
    std::map<int, int> tuned_variant;  // class entry -> VARIANT_* of the last autotune (default for new screeners)
    int autotune_rejected = 0;      // bit v: variant v's G failed the last autotune's check (never chosen)
    // atoms per group of the tiled kernels (VLX_GROUP_ATOMS); 0: the largest of 12, 8, 6, 4, 3,
    // 2, 1 whose (ps|ss) tiles fit max_tile_lds (12 for def2-SVP waters, 8 for def2-TZVP)
    int group_atoms = 0;
    int tiled_waves = 0;            // 4: 4-wave blocks for every tiled class; 0: per-class default (VLX_TILED_WAVES)
    int max_tile_lds = 96 * 1024;   // larger tiles fall back to the untiled kernel
    // primitives of one ket pair per lane of the tiled kernels (VLX_KET_PRIMS, 1..16): a ket
    // pair of n primitives takes ceil(n / ket_prims) lanes with near-equal shares
    int ket_prims = 3;
    // stream balance by the autotuner's measured class times when every class has one (VLX_BALANCE=1,
    // default; 40.4 against 40.1 builds/s on the 32-water build) or by estimated FLOPs (0)
    int balance_measured = 1;
    bool launch_by_cost = true;  // VLX_LAUNCH_ORDER=0: class-table order on every stream
    // largest grid of a tiled launch, blocks striding over the items (VLX_TILED_GRID); 2048 measured
    // best on the 32-water build (42.2 builds/s against 41.7 at 4096 and 40.2 at 8192)
    int tiled_grid = 2048;
    int untiled_wave_cap = 8192;  // largest wavefront count of an untiled launch (VLX_UNTILED_WAVES)
    // largest grid (one wavefront per block) of a matrix-core MD launch (VLX_MMD_GRID): on the 64-water
    // def2-TZVP build 16384 / 32768 / 65536 measured 0.775 / 0.786 / 0.793 builds/s
    int mmd_grid = 65536;
    // kets per matrix-core MD work item (VLX_MMD_KETS): on the 64-water def2-TZVP build the MD classes
    // took 333.2 / 327.1 / 332.1 / 342.9 ms with 1 / 2 / 4 / 8 kets
    int mmd_kets = 2;
    // untiled items reordered so each wave's contiguous range has equal estimated cost
    // (VLX_BALANCE_ITEMS=0: the plain build order)
    int balance_items = 1;
    // classes whose estimated share of the build reaches solo_frac run alone (VLX_SOLO_FRAC; 0: none)
    double solo_frac = 0.0;
    // order of equal-(bin, nprim) kets in the untiled lists (VLX_KET_ORDER): 2, along diagonals
    // of equal C - D offset, gives the lanes of a wave distinct and mostly consecutive F' targets
    // for their exchange atomics (shell order, 0, puts lanes that share C on one address)
    int ket_order = 2;
    hipStream_t stream = nullptr;
    hipStream_t side[NSIDE] = {};
    hipEvent_t ev_start = nullptr;
    hipEvent_t ev_done[NSIDE] = {};
    DevBuf<double> d_boys;
    DevBuf<float> d_boys_f;         // the Boys table in FP32 (mixed precision)
    DevBuf<double> d_density, d_fock, d_faccum;
    DevBuf<double> d_sym, d_anti;   // symmetric and antisymmetric parts of a general density
    // multi-density builds: (fj, fk) per digested density, the densities when general ones are
    // split, and the F' accumulators of all of them
    DevBuf<double> d_dfac, d_mdens, d_mfacc;
    std::vector<double> h_dfac;
    // deterministic mode: every build through the direct-digestion kernels with fixed-point
    // accumulation into d_fx (two 64-bit words per F' element), the same bits on every run
    bool deterministic = false;
    DevBuf<unsigned long long> d_fx;
    DevBuf<unsigned long long> d_dmax;  // max |D| of the current build (bit pattern)
    double mixed_tau = 0.0;         // mixed precision: FP32 below this Schwarz x max|D| bound; 0 = FP64 only
    int last_k32 = NO_FP32;         // bin-sum threshold of the last build
    // density-weighted screening (a non-reference fast mode): quartets whose Schwarz bound times the
    // largest |D| of the blocks they read is below dscreen_tau are skipped; 0 = off
    double dscreen_tau = 0.0;
    DevBuf<unsigned long long> d_gmax;  // per pair of atom groups: max|D| (bit patterns)
    DevBuf<int> d_glim;                 // per pair of atom groups: largest bin sum kept
    int block_size_factor = 8;
    // optional per-class timing
    bool timing = false;
    struct Pending {
        int entry;
        hipEvent_t start, stop;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<int, std::pair<double, long long>> class_ms;  // entry -> (ms, launches)
    std::vector<std::unique_ptr<BatchPlan>> plans;  // batched builds, most recently used last
};

struct vlx_screener {
    vlx_context* ctx = nullptr;
    int nshells = 0, nao = 0;
    std::vector<int> l, nprim, prim_offset, ao_base, ao_stride;
    std::vector<double> exps, norms, centers;
    std::vector<PairSet> all;       // every pair, per pair class (ket order = construction order)
    std::vector<PairSet> all_bra;   // the same pairs sorted by shell A (bra order)
    std::vector<PairSet> kept;      // binned pairs sorted by (bin, nprim): ket order
    std::vector<PairSet> kept_bra;  // binned pairs sorted by (bin, shell A, nprim): bra order
    std::vector<PairSet> kept_tiled;      // binned pairs sorted by (group A, group B, nprim, bin): ket order
    std::vector<PairSet> kept_tiled_bra;  // the same sorted by (group A, group B, bin, nprim): bra order
    std::vector<PairSet> all_tiled, all_tiled_bra;  // every pair, same two orders
    // atom groups of the tiled kernels
    std::vector<int> shell_group, shell_loc;  // per shell: group, local AO index of component 0
    int ngroups = 0;
    int group_atoms = 0;  // atoms per group in use
    std::vector<int> grp_off, grp_n, grp_ao;  // (ngroups x 4) offsets/sizes, flat global AO lists
    int grp_max[4] = {0, 0, 0, 0};           // largest (group, l) block per l
    DevBuf<int> d_grp_off, d_grp_n, d_grp_ao;
    std::vector<double> pair_max;   // nshells x nshells
    struct PairRef {
        int cls = -1, idx = -1, swapped = 0;
    };
    std::vector<PairRef> pair_ref;  // nshells x nshells -> position in all[]
    std::map<int, std::vector<std::unique_ptr<ClassSchedule>>> schedules;  // by (ithreshold, FP32 bin-sum threshold)
    std::map<int, int> variant;  // class entry -> VARIANT_* chosen by vlx_fock_autotune
    std::map<int, double> tuned_ms;  // class entry -> its measured time (ms) under that variant
    DevBuf<double> d_faccum;  // F' of this screener's molecule in batched builds
    DevBuf<MolDesc> d_mol;    // {0, nao}: the molecule table of a single-molecule build
    DevBuf<int> d_ao_group;   // atom group of every AO (density-weighted screening)
    int epoch = 0;            // incremented when the autotuner changes the variants (batch plans)
};

namespace {

void build_prims(const vlx_screener& s, int sa, int sb, std::vector<PrimPair>& prims)
{
    const double* A = &s.centers[3 * sa];
    const double* B = &s.centers[3 * sb];
    const double ab2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]);
    const double kfac = std::sqrt(2.0) * std::pow(M_PI, 1.25);
    for (int i = 0; i < s.nprim[sa]; i++)
        for (int j = 0; j < s.nprim[sb]; j++) {
            const double a = s.exps[s.prim_offset[sa] + i], b = s.exps[s.prim_offset[sb] + j];
            const double na = s.norms[s.prim_offset[sa] + i], nb = s.norms[s.prim_offset[sb] + j];
            const double z = a + b;
            PrimPair p;
            p.zeta = z;
            p.px = (a * A[0] + b * B[0]) / z;
            p.py = (a * A[1] + b * B[1]) / z;
            p.pz = (a * A[2] + b * B[2]) / z;
            p.pax = p.px - A[0];
            p.pay = p.py - A[1];
            p.paz = p.pz - A[2];
            p.k = kfac * na * nb * std::exp(-a * b / z * ab2) / z;
            prims.push_back(p);
        }
}

void add_pair(const vlx_screener& s, PairSet& ps, int sa, int sb)
{
    PairInfo pi{};
    pi.ao_a = s.ao_base[sa];
    pi.str_a = s.ao_stride[sa];
    pi.ao_b = s.ao_base[sb];
    pi.str_b = s.ao_stride[sb];
    pi.prim_start = static_cast<int>(ps.prims.size());
    pi.nprim = s.nprim[sa] * s.nprim[sb];
    pi.same = (sa == sb) ? 1 : 0;
    pi.bin = 0;
    pi.rank = static_cast<int>(ps.pairs.size());
    pi.shell_a = sa;
    pi.grp_a = s.shell_group.empty() ? 0 : s.shell_group[sa];
    pi.grp_b = s.shell_group.empty() ? 0 : s.shell_group[sb];
    pi.loc_a = s.shell_loc.empty() ? 0 : s.shell_loc[sa];
    pi.loc_b = s.shell_loc.empty() ? 0 : s.shell_loc[sb];
    pi.abx = s.centers[3 * sa] - s.centers[3 * sb];
    pi.aby = s.centers[3 * sa + 1] - s.centers[3 * sb + 1];
    pi.abz = s.centers[3 * sa + 2] - s.centers[3 * sb + 2];
    ps.pairs.push_back(pi);
    ps.shell_a.push_back(sa);
    ps.shell_b.push_back(sb);
    build_prims(s, sa, sb, ps.prims);
}

// Wavefronts of an untiled class launch (fock_kernel, fock_kernel_bs): at most `cap`, in blocks
// of VLX_WAVES_PER_BLOCK; wave g takes the contiguous items [n g / nw, n (g + 1) / nw).
long long untiled_waves(long long nitems, int cap)
{
    const long long waves = std::min<long long>(nitems, cap);
    return (waves + VLX_WAVES_PER_BLOCK - 1) / VLX_WAVES_PER_BLOCK * VLX_WAVES_PER_BLOCK;
}

// Reorders work items so that the kernel's equal-count contiguous split over nw waves is
// balanced by estimated cost: longest item first onto the least-loaded wave that still has a
// free slot (greedy LPT with slot counts fixed by the split).  The set of items is unchanged.
void balance_items(std::vector<WorkItem>& items, const std::vector<double>& cost, long long nw)
{
    const long long n = static_cast<long long>(items.size());
    if (n <= nw) return;
    std::vector<int> order(n);
    for (long long i = 0; i < n; i++) order[i] = static_cast<int>(i);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    using Slot = std::pair<double, long long>;  // (load, wave)
    std::priority_queue<Slot, std::vector<Slot>, std::greater<Slot>> heap;
    std::vector<long long> fill(nw, 0), cap(nw);
    for (long long g = 0; g < nw; g++) {
        cap[g] = n * (g + 1) / nw - n * g / nw;
        if (cap[g] > 0) heap.push({0.0, g});
    }
    std::vector<WorkItem> out(n);
    for (int idx : order) {
        const auto [load, g] = heap.top();
        heap.pop();
        out[n * g / nw + fill[g]++] = items[idx];
        if (fill[g] < cap[g]) heap.push({load + cost[idx], g});
    }
    items.swap(out);
}

int pair_bin(double maxval)
{
    // src/orbdata/BlockedGtoPairBlock.cpp:32-42
    const double fact = std::sqrt(maxval);
    if (fact > 1.0e-15) return (fact < 1.0) ? static_cast<int>(-std::floor(std::log10(fact))) : 0;
    return -1;
}

void finalize_set(PairSet& ps)
{
    for (int b = 0; b < NBINS; b++) ps.cnt_le[b] = 0;
    for (int k = 0; k < ps.size(); k++)
        for (int b = std::max(ps.pairs[k].bin, 0); b < NBINS; b++) ps.cnt_le[b]++;
    ps.nprim_prefix.assign(ps.size() + 1, 0);
    for (int k = 0; k < ps.size(); k++) ps.nprim_prefix[k + 1] = ps.nprim_prefix[k] + ps.pairs[k].nprim;
}

// Copy of `src` reordered by `order`; PairInfo.rank keeps the position in src.
void reorder_set(const vlx_screener& s, const PairSet& src, const std::vector<int>& order, PairSet& dst)
{
    dst.pairs.clear();
    dst.prims.clear();
    dst.shell_a.clear();
    dst.shell_b.clear();
    dst.bins.clear();
    (void)s;
    // the pair records and their primitive pairs are copied (they do not depend on the order),
    // with the primitive range moved to the new position
    dst.pairs.reserve(order.size());
    dst.shell_a.reserve(order.size());
    dst.shell_b.reserve(order.size());
    dst.bins.reserve(order.size());
    size_t np = 0;
    for (int k : order) np += src.pairs[k].nprim;
    dst.prims.reserve(np);
    for (int k : order) {
        PairInfo pi = src.pairs[k];
        const auto p0 = src.prims.begin() + pi.prim_start;
        pi.prim_start = static_cast<int>(dst.prims.size());
        dst.prims.insert(dst.prims.end(), p0, p0 + pi.nprim);
        dst.pairs.push_back(pi);
        dst.shell_a.push_back(src.shell_a[k]);
        dst.shell_b.push_back(src.shell_b[k]);
        dst.bins.push_back(src.pairs[k].bin);
    }
    finalize_set(dst);
}

// Tiled orders of `src`, both by (group A, group B, bin, nprim): the bras a
// ket chunk's bins allow are a prefix of each segment.  Kets are expanded to
// lane entries of at most run_cap consecutive primitives of one pair (the
// contraction over the rest of the pair's primitives happens in the LDS
// tiles).  Segments are the runs of equal (group A, group B).
void build_tiled(const vlx_screener& s, const PairSet& src, PairSet& ket, PairSet& bra)
{
    std::vector<int> order(src.size());
    std::iota(order.begin(), order.end(), 0);
    auto by = [&](bool nprim_first) {
        return [&, nprim_first](int i, int j) {
            const PairInfo &x = src.pairs[i], &y = src.pairs[j];
            if (x.grp_a != y.grp_a) return x.grp_a < y.grp_a;
            if (x.grp_b != y.grp_b) return x.grp_b < y.grp_b;
            if (nprim_first && x.nprim != y.nprim) return x.nprim < y.nprim;
            if (x.bin != y.bin) return x.bin < y.bin;
            return x.nprim < y.nprim;
        };
    };
    std::stable_sort(order.begin(), order.end(), by(false));
    reorder_set(s, src, order, ket);
    std::vector<int> rank_of(src.size());
    for (int k = 0; k < ket.size(); k++) {
        ket.pairs[k].rank = k;
        rank_of[order[k]] = k;
    }
    std::vector<int> border(src.size());
    std::iota(border.begin(), border.end(), 0);
    std::stable_sort(border.begin(), border.end(), by(false));
    reorder_set(s, src, border, bra);
    for (int k = 0; k < bra.size(); k++) bra.pairs[k].rank = rank_of[border[k]];
    for (PairSet* ps : {&ket, &bra}) {
        ps->segs.clear();
        for (int k = 0; k < ps->size(); k++) {
            const PairInfo& pk = ps->pairs[k];
            if (ps->segs.empty() || ps->segs.back().ga != pk.grp_a || ps->segs.back().gb != pk.grp_b) {
                PairSet::Segment sg{};
                sg.start = k;
                sg.ga = pk.grp_a;
                sg.gb = pk.grp_b;
                ps->segs.push_back(sg);
            }
            PairSet::Segment& sg = ps->segs.back();
            sg.end = k + 1;
            for (int b = std::max(pk.bin, 0); b < NBINS; b++) sg.cnt_le[b]++;
        }
    }
    // lane entries for a run cap; inside a segment the entries go by (bin, run length
    // descending), so that a 64-entry chunk holds runs of equal length where it can (its
    // lanes loop to the chunk's longest run)
    auto make_entries = [&](int pmax, PairSet::Entries& out) {
        out.kentries.clear();
        out.seg_e0.clear();
        out.seg_e1.clear();
        std::vector<long long> eoff(ket.size() + 1, 0);
        for (int k = 0; k < ket.size(); k++) {
            const int np = ket.pairs[k].nprim;
            const int m = (np + pmax - 1) / pmax;
            for (int t = 0, q = 0; t < m; t++) {
                const int cnt = np / m + (t < np % m ? 1 : 0);
                out.kentries.push_back(make_int2(k, q | (cnt << 16)));
                q += cnt;
            }
            eoff[k + 1] = eoff[k] + m;
        }
        for (const auto& sg : ket.segs) {
            const int e0 = static_cast<int>(eoff[sg.start]), e1 = static_cast<int>(eoff[sg.end]);
            out.seg_e0.push_back(e0);
            out.seg_e1.push_back(e1);
            std::stable_sort(out.kentries.begin() + e0, out.kentries.begin() + e1,
                             [&](const int2& x, const int2& y) {
                                 const int bx = ket.pairs[x.x].bin, by = ket.pairs[y.x].bin;
                                 if (bx != by) return bx < by;
                                 return (x.y >> 16) > (y.y >> 16);
                             });
        }
    };
    // runs of up to ket_prims primitives (they share a pair's contracted work: HRR,
    // digestion), and single primitives (chunks of pairs with mixed primitive counts, such
    // as the 1-, 3- and 5-primitive d-s pairs of def2 bases, then fill every lane)
    make_entries(s.ctx->ket_prims, ket.runs);
    make_entries(1, ket.singles);
}

// U arrays of a pair set for the matrix-core MD Fock kernel (once per set).
int ensure_u(vlx_context* ctx, PairSet& ps)
{
    if (ps.u_ready || ps.size() == 0) return VLX_OK;
    const size_t per = (size_t)vlx_u_rows(ps.la, ps.lb) * vlx_u_cols(ps.la, ps.lb);
    HIP_TRY(ps.d_u.alloc(per * ps.prims.size()));
    vlx_launch_u_build(ps.d_pairs.ptr, ps.d_prims.ptr, ps.size(), ps.la, ps.lb, ps.d_u.ptr, ctx->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ps.u_ready = true;
    return VLX_OK;
}

int upload_set(PairSet& ps)
{
    // one allocation per set (a screener holds dozens of sets; hipMalloc per array dominated
    // its set-up), each array at a 256-byte boundary
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t b_pairs = up(ps.pairs.size() * sizeof(PairInfo)), b_prims = up(ps.prims.size() * sizeof(PrimPair));
    const size_t b_runs = up(ps.runs.kentries.size() * sizeof(int2));
    const size_t b_singles = up(ps.singles.kentries.size() * sizeof(int2));
    HIP_TRY(ps.d_arena.alloc(b_pairs + b_prims + b_runs + b_singles));
    char* base = ps.d_arena.ptr;
    ps.d_pairs.ptr = reinterpret_cast<PairInfo*>(base);
    ps.d_prims.ptr = reinterpret_cast<PrimPair*>(base + b_pairs);
    ps.runs.d_kentries.ptr = b_runs ? reinterpret_cast<int2*>(base + b_pairs + b_prims) : nullptr;
    ps.singles.d_kentries.ptr = b_singles ? reinterpret_cast<int2*>(base + b_pairs + b_prims + b_runs) : nullptr;
    auto copy = [](void* dst, const void* src, size_t bytes) {
        return bytes ? hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    HIP_TRY(copy(ps.d_pairs.ptr, ps.pairs.data(), ps.pairs.size() * sizeof(PairInfo)));
    HIP_TRY(copy(ps.d_prims.ptr, ps.prims.data(), ps.prims.size() * sizeof(PrimPair)));
    HIP_TRY(copy(ps.runs.d_kentries.ptr, ps.runs.kentries.data(), ps.runs.kentries.size() * sizeof(int2)));
    HIP_TRY(copy(ps.singles.d_kentries.ptr, ps.singles.kentries.data(), ps.singles.kentries.size() * sizeof(int2)));
    return VLX_OK;
}

const DiagLauncher* find_diag(int x)
{
    for (const auto& d : g_diag_table)
        if (d.x == x) return &d;
    return nullptr;
}

// Digestion factors of a label and whether it uses erf(omega r)/r.
// src/t4c_func/T4CLocalDistributor.cpp: 2jk (4, -1), 2jkx (4, -xf),
// j (2, 0), k (0, 1), kx (0, xf); then G = F' + F'^T.  The range-separated
// labels j_rs, k_rs, kx_rs digest like j, k, kx on erf-attenuated integrals
// (src/t4c_func/T4CMatrixDistributor.cpp:51-54); omega is ignored otherwise.  Labels are
// case-insensitive, as the distributors lower-case them (T4CMatrixDistributor.cpp:18).
int label_factors(const char* label, double xf, double* fj, double* fk, bool* range_separated)
{
    if (!label) return fail(VLX_ERR_ARG, "label is NULL");
    std::string s(label);
    for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    *range_separated = false;
    if (s == "2jk") { *fj = 4.0; *fk = -1.0; }
    else if (s == "2jkx") { *fj = 4.0; *fk = -xf; }
    else if (s == "j" || s == "j_rs") { *fj = 2.0; *fk = 0.0; }
    else if (s == "k" || s == "k_rs") { *fj = 0.0; *fk = 1.0; }
    else if (s == "kx" || s == "kx_rs") { *fj = 0.0; *fk = xf; }
    else
        return fail(VLX_ERR_ARG, "unknown Fock label '%s'", label);
    *range_separated = s.size() > 3 && s.compare(s.size() - 3, 3, "_rs") == 0;
    return VLX_OK;
}

__global__ void zero_kernel(double* __restrict__ p, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

// max |x| as the bit pattern of a non-negative double (ordered like the value)
__global__ void absmax_kernel(const double* __restrict__ x, size_t n, unsigned long long* __restrict__ out)
{
    double m = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs(x[i]));
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, static_cast<unsigned long long>(__double_as_longlong(m)));
}

__global__ void clear_u64_kernel(unsigned long long* __restrict__ p) { *p = 0ull; }

__global__ void clear_u64_n_kernel(unsigned long long* __restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// max |D_pq| per pair of atom groups (both orders, bit patterns of non-negative doubles)
__global__ void group_dmax_kernel(const double* __restrict__ d, int n, const int* __restrict__ ao_group, int ng,
                                  unsigned long long* __restrict__ gmax)
{
    const size_t nn = (size_t)n * n;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nn; i += (size_t)gridDim.x * blockDim.x) {
        const double v = fabs(d[i]);
        if (v == 0.0) continue;
        const int gp = ao_group[i / n], gq = ao_group[i % n];
        const auto bits = static_cast<unsigned long long>(__double_as_longlong(v));
        atomicMax(gmax + gp * ng + gq, bits);
        atomicMax(gmax + gq * ng + gp, bits);
    }
}

// The largest bin sum kept for a quartet reading a D block of max|D| = M: Schwarz bins bound the
// quartet by 10^(2 - bin sum), so it stays while 10^(2 - s) M >= tau, s <= floor(2 + log10(M / tau));
// never above ithreshold; -1000 (nothing kept) for a zero block.
__global__ void group_lim_kernel(const unsigned long long* __restrict__ gmax, int n, double tau, int ithreshold,
                                 int* __restrict__ glim)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = __longlong_as_double(static_cast<long long>(gmax[i]));
    int lim = -1000;
    if (m > 0.0) lim = static_cast<int>(fmin(static_cast<double>(ithreshold), floor(2.0 + log10(m / tau))));
    glim[i] = lim;
}

// G = F' + F'^T
__global__ void symmetrize_kernel(const double* __restrict__ f, double* __restrict__ g, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (q < n) g[(size_t)p * n + q] = f[(size_t)p * n + q] + f[(size_t)q * n + p];
}

// G += c (F' - F'^T)
__global__ void antisym_add_kernel(const double* __restrict__ f, double* __restrict__ g, int n, double c)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (q < n) g[(size_t)p * n + q] += c * (f[(size_t)p * n + q] - f[(size_t)q * n + p]);
}

// F' from the fixed-point accumulators: hi 2^-32 + lo 2^-96 (hi signed)
__global__ void fx_to_double_kernel(const unsigned long long* __restrict__ fx, double* __restrict__ f, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const long long hi = static_cast<long long>(fx[2 * i + 1]);
        f[i] = static_cast<double>(hi) * 0x1p-32 + static_cast<double>(fx[2 * i]) * 0x1p-96;
    }
}

// S = (D + D^T) / 2, A = (D - D^T) / 2
__global__ void split_kernel(const double* __restrict__ d, double* __restrict__ s, double* __restrict__ a, int n)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (q < n) {
        const double x = d[(size_t)p * n + q], y = d[(size_t)q * n + p];
        s[(size_t)p * n + q] = 0.5 * (x + y);
        a[(size_t)p * n + q] = 0.5 * (x - y);
    }
}

// Work items of every class at ithreshold.  k32 < NO_FP32 also splits the
// tiled items: the bras of an item whose bin plus the smallest ket bin of
// its chunk reaches k32 form an FP32 item (every quartet in it has
// Schwarz bound < 10^(2 - k32)); the rest stay FP64.
// create = false: only a schedule built earlier is returned (no host or device allocation);
// without one the call fails.
int get_schedules(vlx_screener* s, int ithr, int k32, std::vector<std::unique_ptr<ClassSchedule>>** out,
                  bool create = true)
{
    const int key = ((ithr + 1) * 128 + k32) * (NVARIANTS + 1) + (s->ctx->force_variant + 1);
    auto it = s->schedules.find(key);
    if (it != s->schedules.end()) {
        *out = &it->second;
        return VLX_OK;
    }
    if (!create)
        return fail(VLX_ERR_ARG, "no work schedule of this screener for ithreshold %d and FP32 bin sum %d: one "
                                 "allocating build of the screener at these settings must come first", ithr, k32);
    std::vector<std::unique_ptr<ClassSchedule>> list;
    const int nentries = static_cast<int>(sizeof(g_class_table) / sizeof(g_class_table[0]));
    const bool unscreened = ithr < 0;
    for (int e = 0; e < nentries; e++) {
        const auto& ce = g_class_table[e];
        const PairSet& bra = unscreened ? s->all_bra[ce.x] : s->kept_bra[ce.x];
        const PairSet& ket = unscreened ? s->all[ce.y] : s->kept[ce.y];
        if (bra.size() == 0 || ket.size() == 0) continue;
        const bool same = ce.x == ce.y;
        auto sch = std::make_unique<ClassSchedule>();
        sch->entry = e;
        // exact quartet counts, from the ket side (ket order is sorted by bin)
        for (int j = 0; j < ket.size(); j++) {
            long long lo = 0, hi;
            if (unscreened) {
                hi = bra.size();
            } else {
                const int lim = ithr - ket.pairs[j].bin;
                hi = (lim < 0) ? 0 : bra.cnt_le[std::min(lim, NBINS - 1)];
            }
            if (same) lo = j;  // ranks of the same-class bras run over the ket order
            if (hi <= lo) continue;
            const PairSet& rk = same ? ket : bra;  // prefix sums in an order where [lo, hi) is contiguous
            sch->quartets += hi - lo;
            sch->prim_quartets += (long long)ket.pairs[j].nprim * (rk.nprim_prefix[hi] - rk.nprim_prefix[lo]);
        }
        const int* gm = s->grp_max;
        const int tile_lds = 8 * (gm[ce.la] * gm[ce.lb] + gm[ce.lc] * gm[ce.ld] + gm[ce.la] * gm[ce.lc] +
                                   gm[ce.la] * gm[ce.ld] + gm[ce.lb] * gm[ce.lc] + gm[ce.lb] * gm[ce.ld]);
        const int L = ce.la + ce.lb + ce.lc + ce.ld;
        const int nsph = (2 * ce.la + 1) * (2 * ce.lb + 1) * (2 * ce.lc + 1) * (2 * ce.ld + 1);
        const bool can_tile = s->ctx->use_tiled && ce.tiled && L <= s->ctx->tiled_lmax && tile_lds <= s->ctx->max_tile_lds;
        // default: tiled for L <= 4 (and for L = 5 blocks of <= 75 integrals when tiled_lmax
        // allows it); the autotuner's choice, or the one being timed, overrides it
        int var = (can_tile && (L <= 4 || nsph <= 75)) ? (s->ctx->tiled_waves == 4 ? VARIANT_TILED4 : VARIANT_TILED_WIDE)
                                                      : VARIANT_UNTILED;
        if (s->ctx->force_variant >= 0) {
            var = s->ctx->force_variant;
        } else {
            const auto vit = s->variant.find(e);
            const auto cit = s->ctx->tuned_variant.find(e);
            if (vit != s->variant.end()) var = vit->second;
            else if (cit != s->ctx->tuned_variant.end()) var = cit->second;
        }
        if (variant_tiled(var) && !can_tile) var = VARIANT_UNTILED;
        if (var == VARIANT_MMD && !ce.mmd) var = VARIANT_UNTILED;
        if (var == VARIANT_KSL && !ce.ksl) var = VARIANT_UNTILED;
        // D tiles double the LDS of the F' tiles: beside the FP64 Boys table, or the plain wide tiles
        if (var == VARIANT_TILED_DLDS &&
            (!ce.tiled_dlds || 2 * tile_lds > 160 * 1024 - static_cast<int>(BOYS_TAB_SIZE * sizeof(double)) - 1024))
            var = VARIANT_TILED_WIDE;
        // two D groups per chunk: the J_CD, K_AD and K_BD tiles twice
        const int tile_lds2 = 8 * (gm[ce.la] * gm[ce.lb] + gm[ce.la] * gm[ce.lc] + gm[ce.lb] * gm[ce.lc] +
                                   2 * (gm[ce.lc] + gm[ce.la] + gm[ce.lb]) * gm[ce.ld]);
        if (var == VARIANT_TILED_2SEG && tile_lds2 > s->ctx->max_tile_lds) var = VARIANT_TILED_WIDE;
        if (ce.variant == VLX_WAVE_MD) var = VARIANT_MMD;  // no generated kernel
        sch->variant = var;
        sch->mmd = var == VARIANT_MMD;
        sch->ksl = var == VARIANT_KSL;
        sch->dlds = var == VARIANT_TILED_DLDS;
        if (variant_tiled(var)) {
            // tiled items: (bra segment prefix, chunk of <= 64 kets of a ket segment)
            const PairSet& tb = unscreened ? s->all_tiled_bra[ce.x] : s->kept_tiled_bra[ce.x];
            const PairSet& tk = unscreened ? s->all_tiled[ce.y] : s->kept_tiled[ce.y];
            sch->tiled = true;
            sch->singles = var == VARIANT_TILED4_P1 || var == VARIANT_TILED_WIDE_P1;
            const PairSet::Entries& en = sch->singles ? tk.singles : tk.runs;
            sch->tiled_waves = (var == VARIANT_TILED4 || var == VARIANT_TILED4_P1) ? 4 : 0;
            sch->tile_lds = var == VARIANT_TILED_2SEG ? tile_lds2 : tile_lds;
            // chunks of <= 64 lane entries of one ket segment; with VARIANT_TILED_2SEG the last, partly
            // filled chunk of a segment takes the first entries of the next segment when both have the
            // same C group (their entry ranges are adjacent): {j0, j1, first segment, gd2, jm}
            std::vector<std::array<int, 5>> chunks;
            int open = -1;  // a one-segment chunk with room left
            for (size_t sk = 0; sk < tk.segs.size(); sk++) {
                int j = en.seg_e0[sk];
                const int e1 = en.seg_e1[sk];
                if (var == VARIANT_TILED_2SEG && open >= 0 && tk.segs[chunks[open][2]].ga == tk.segs[sk].ga &&
                    chunks[open][1] == j) {
                    auto& c = chunks[open];
                    const int take = std::min(64 - (c[1] - c[0]), e1 - j);
                    c[3] = tk.segs[sk].gb;
                    c[4] = j;
                    c[1] = j + take;
                    j += take;
                }
                open = -1;
                for (; j < e1; j += 64) {
                    const int j1 = std::min(e1, j + 64);
                    chunks.push_back({j, j1, static_cast<int>(sk), -1, j1});
                    if (j1 - j < 64) open = static_cast<int>(chunks.size()) - 1;
                }
            }
            for (const auto& ch : chunks) {
                const size_t sk = static_cast<size_t>(ch[2]);
                const auto& ks = tk.segs[sk];
                const int j0 = ch[0], j1 = ch[1], gd2 = ch[3], jm = ch[4];
                {
                    int bin_lo = NBINS, kmin = en.kentries[j0].x;
                    for (int j = j0; j < j1; j++) {
                        bin_lo = std::min(bin_lo, std::max(tk.pairs[en.kentries[j].x].bin, 0));
                        kmin = std::min(kmin, en.kentries[j].x);
                    }
                    const int lim = unscreened ? NBINS - 1 : ithr - bin_lo;
                    if (lim < 0) continue;
                    // same class: bra segments at or after the (first) ket segment (segment order is
                    // identical in both lists); the diagonal segment is masked by rank
                    for (size_t sb = same ? sk : 0; sb < tb.segs.size(); sb++) {
                        const auto& bs = tb.segs[sb];
                        int i0 = bs.start;
                        const int i1 = bs.start + (unscreened ? bs.end - bs.start : bs.cnt_le[std::min(lim, NBINS - 1)]);
                        // diagonal segment: the tiled bra and ket orders are the same permutation
                        // (position = rank), so bras before the chunk's lowest ket pair have a lower
                        // rank than every ket of the chunk and would run fully masked
                        if (same && sb == sk) i0 = std::max(i0, kmin);
                        if (i1 <= i0) continue;
                        // bras [im, i1) have bin >= k32 - bin_lo: FP32 for every lane of the chunk
                        int im = i1;
                        if (k32 < NO_FP32 && !unscreened) {
                            const int b = k32 - bin_lo - 1;
                            im = (b < 0) ? i0 : std::min(i1, bs.start + bs.cnt_le[std::min(b, NBINS - 1)]);
                            im = std::max(im, i0);
                        }
                        if (im > i0) sch->titems.push_back({i0, im, j0, j1, bs.ga, bs.gb, ks.ga, ks.gb, 0, bin_lo, gd2, jm});
                        if (i1 > im)
                            sch->titems32.push_back({im, i1, j0, j1, bs.ga, bs.gb, ks.ga, ks.gb, 0, bin_lo, gd2, jm});
                    }
                }
            }
        } else if (!sch->mmd && (ce.variant == VLX_KET_STATIONARY || sch->ksl)) {
            // work items: 64 consecutive kets x a chunk of bras; the chunk shrinks
            // for classes with little work so that the launch still has enough items
            std::vector<std::pair<int, int>> ranges;
            long long span = 0;
            for (int j0 = 0; j0 < ket.size(); j0 += 64) {
                const int bin_lo = unscreened ? 0 : ket.pairs[j0].bin;
                int hi;
                if (unscreened) {
                    hi = bra.size();
                } else {
                    const int lim = ithr - bin_lo;
                    hi = (lim < 0) ? 0 : bra.cnt_le[std::min(lim, NBINS - 1)];
                }
                int lo = 0;
                if (same && !unscreened) lo = (bin_lo > 0) ? bra.cnt_le[bin_lo - 1] : 0;
                ranges.push_back({lo, hi});
                span += std::max(0, hi - lo);
            }
            const int chunk = static_cast<int>(
                std::max<long long>(1, std::min<long long>(s->ctx->max_chunk, span / s->ctx->min_items)));
            for (size_t t = 0; t < ranges.size(); t++)
                for (int i0 = ranges[t].first; i0 < ranges[t].second; i0 += chunk)
                    sch->items.push_back({static_cast<int>(64 * t), i0, std::min(ranges[t].second, i0 + chunk), 0,
                                          ket.size(), 0, 0, 0});
        } else {
            // work items: one bra x a run of the ket prefix its bins allow
            // (64 kets, one per lane; or a few kets for the matrix-core MD kernel)
            const int kstep = sch->mmd ? s->ctx->mmd_kets : 64;
            for (int i = 0; i < bra.size(); i++) {
                int nk;
                if (unscreened) {
                    nk = ket.size();
                } else {
                    const int lim = ithr - bra.pairs[i].bin;
                    nk = (lim < 0) ? 0 : ket.cnt_le[std::min(lim, NBINS - 1)];
                }
                if (same) nk = std::min(nk, bra.pairs[i].rank + 1);
                for (int j0 = 0; j0 < nk; j0 += kstep) {
                    const int j1 = sch->mmd ? std::min(nk, j0 + kstep) : nk;
                    sch->items.push_back({j0, i, j1, 0, j1, 0, 0, 0});
                }
            }
        }
        // schedule efficiency: slots the wavefronts execute, masked lanes included
        if (sch->tiled) {
            const PairSet& tb = unscreened ? s->all_tiled_bra[ce.x] : s->kept_tiled_bra[ce.x];
            const PairSet& tk = unscreened ? s->all_tiled[ce.y] : s->kept_tiled[ce.y];
            // a chunk's lanes loop over the largest primitive count among its entries
            const PairSet::Entries& en = sch->singles ? tk.singles : tk.runs;
            auto kmax = [&](const TileItem& ti) {
                int m = 1;
                for (int j = ti.j0; j < ti.j1; j++) m = std::max(m, en.kentries[j].y >> 16);
                return static_cast<long long>(m);
            };
            for (const auto& ti : sch->titems) {
                sch->lane_slots += 64LL * (ti.i1 - ti.i0);
                sch->prim_slots += (tb.nprim_prefix[ti.i1] - tb.nprim_prefix[ti.i0]) * 64 * kmax(ti);
            }
            for (const auto& ti : sch->titems32) {
                sch->lane_slots += 64LL * (ti.i1 - ti.i0);
                sch->prim_slots32 += (tb.nprim_prefix[ti.i1] - tb.nprim_prefix[ti.i0]) * 64 * kmax(ti);
            }
            sch->prim_slots += sch->prim_slots32;
        } else if (!sch->mmd) {
            // estimated cost of an item in FLOPs of one lane: primitive iterations times the
            // per-primitive work, plus the contracted work per quartet; the items are then
            // spread so that every wave's contiguous range costs about the same
            const int nsph = (2 * ce.la + 1) * (2 * ce.lb + 1) * (2 * ce.lc + 1) * (2 * ce.ld + 1);
            const double pcost = ce.vrr_flops + 60.0, qcost = ce.hrr_flops + 12.0 * nsph;
            std::vector<double> cost;
            cost.reserve(sch->items.size());
            const bool ks = ce.variant == VLX_KET_STATIONARY || sch->ksl;
            for (const auto& wi : sch->items) {
                const int j1 = ks ? std::min(wi.j0 + 64, ket.size()) : std::min(wi.j0 + 64, wi.i1);
                int kmax = 0;
                for (int j = wi.j0; j < j1; j++) kmax = std::max(kmax, ket.pairs[j].nprim);
                const long long bprims = ks ? bra.nprim_prefix[wi.i1] - bra.nprim_prefix[wi.i0] : bra.pairs[wi.i0].nprim;
                const long long nbra = ks ? wi.i1 - wi.i0 : 1;
                sch->lane_slots += 64LL * nbra;
                sch->prim_slots += (long long)kmax * bprims * 64;
                cost.push_back(kmax * bprims * pcost + nbra * qcost);
            }
            if (s->ctx->balance_items)
                balance_items(sch->items, cost,
                              untiled_waves(static_cast<long long>(sch->items.size()), s->ctx->untiled_wave_cap));
        } else {
            for (const auto& wi : sch->items) {  // matrix-core MD: one ket per wavefront pass
                int kmax = 0;
                for (int j = wi.j0; j < wi.i1; j++) kmax = std::max(kmax, ket.pairs[j].nprim);
                const int lanes = wi.i1 - wi.j0;
                sch->lane_slots += lanes;
                sch->prim_slots += (long long)kmax * bra.pairs[wi.i0].nprim * lanes;
            }
        }
        if (sch->items.empty() && sch->titems.empty() && sch->titems32.empty()) continue;
        if (sch->mmd) {
            int rc = ensure_u(s->ctx, unscreened ? s->all_bra[ce.x] : s->kept_bra[ce.x]);
            if (!rc) rc = ensure_u(s->ctx, unscreened ? s->all[ce.y] : s->kept[ce.y]);
            if (rc) return rc;
        }
        if (sch->tiled) {
            if (!sch->titems.empty()) HIP_TRY(sch->d_titems.upload(sch->titems));
            if (!sch->titems32.empty()) HIP_TRY(sch->d_titems32.upload(sch->titems32));
        } else {
            HIP_TRY(sch->d_items.upload(sch->items));
        }
        list.push_back(std::move(sch));
    }
    // spread classes over the side streams, most expensive first: by the autotuner's measured
    // times when every class has one (the kernel families differ ~5x in FLOP rate), else by
    // estimated FLOPs
    bool measured = s->ctx->balance_measured && !s->tuned_ms.empty();
    for (auto& sp : list) measured = measured && s->tuned_ms.count(sp->entry);
    std::vector<std::pair<double, ClassSchedule*>> cost;
    for (auto& sp : list) {
        const auto& ce = g_class_table[sp->entry];
        const double c = measured ? s->tuned_ms[sp->entry]
                                  : (double)sp->prim_quartets * (ce.vrr_flops + 60) + (double)sp->quartets * ce.hrr_flops;
        cost.push_back({c, sp.get()});
    }
    std::sort(cost.begin(), cost.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
    // the classes above solo_frac of the build fill the chip by themselves: they run one after
    // another alone, and the rest share the side streams (overlap fills the tails of small launches)
    double total = 0.0;
    for (auto& c : cost) total += c.first;
    double load[NSIDE] = {0.0};
    const int nst = s->ctx->nstreams;
    for (auto& c : cost) {
        if (nst > 1 && s->ctx->solo_frac > 0.0 && c.first >= s->ctx->solo_frac * total) {
            c.second->stream = -1;
            continue;
        }
        const int k = static_cast<int>(std::min_element(load, load + nst) - load);
        c.second->stream = k;
        load[k] += c.first;
    }
    // launch order: most expensive first on every stream, so that the streams end on short launches
    if (s->ctx->launch_by_cost) {
        std::vector<std::unique_ptr<ClassSchedule>> sorted;
        for (auto& c : cost)
            for (auto& sp : list)
                if (sp.get() == c.second) sorted.push_back(std::move(sp));
        list.swap(sorted);
    }
    auto res = s->schedules.emplace(key, std::move(list));
    *out = &res.first->second;
    return VLX_OK;
}

int check_symmetric_host(const double* d, int n)
{
    for (int p = 0; p < n; p++)
        for (int q = p + 1; q < n; q++) {
            const double a = d[(size_t)p * n + q], b = d[(size_t)q * n + p];
            if (std::fabs(a - b) > 1.0e-12 * std::max(1.0, std::max(std::fabs(a), std::fabs(b))))
                return fail(VLX_ERR_ARG, "density is not symmetric at (%d, %d): use vlx_fock_compute_general", p, q);
        }
    return VLX_OK;
}

}  // namespace

extern "C" {

const char* vlx_last_error(void) { return g_last_error.c_str(); }

const char* vlx_version(void) { return "veloxchem_amd 0.1 (gfx950)"; }

int vlx_max_angular_momentum(void) { return VLX_GEN_LMAX; }

int vlx_context_create(int device, vlx_context** out)
{
    if (!out) return fail(VLX_ERR_ARG, "out is NULL");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(VLX_ERR_ARG, "device %d out of range (%d devices)", device, ndev);
    auto ctx = std::make_unique<vlx_context>();
    ctx->device = device;
    if (const char* e = std::getenv("VLX_STREAMS")) ctx->nstreams = std::max(1, std::min(NSIDE, std::atoi(e)));
    if (const char* e = std::getenv("VLX_BRA_CHUNK")) ctx->max_chunk = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("VLX_MIN_ITEMS")) ctx->min_items = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("VLX_TILED")) ctx->use_tiled = std::atoi(e) != 0;
    if (const char* e = std::getenv("VLX_TILED_LMAX")) ctx->tiled_lmax = std::atoi(e);
    if (const char* e = std::getenv("VLX_GROUP_ATOMS")) ctx->group_atoms = std::max(0, std::atoi(e));
    if (const char* e = std::getenv("VLX_TILED_WAVES")) ctx->tiled_waves = std::atoi(e);
    if (const char* e = std::getenv("VLX_KET_PRIMS")) ctx->ket_prims = std::max(1, std::min(16, std::atoi(e)));
    if (const char* e = std::getenv("VLX_BALANCE")) ctx->balance_measured = std::atoi(e) != 0;
    if (const char* e = std::getenv("VLX_BALANCE_ITEMS")) ctx->balance_items = std::atoi(e) != 0;
    if (const char* e = std::getenv("VLX_LAUNCH_ORDER")) ctx->launch_by_cost = std::atoi(e) != 0;
    if (const char* e = std::getenv("VLX_TILED_GRID")) ctx->tiled_grid = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("VLX_MMD_GRID")) ctx->mmd_grid = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("VLX_MMD_KETS")) ctx->mmd_kets = std::max(1, std::min(64, std::atoi(e)));
    // dynamic LDS of a tiled block: the Boys table (static) and the tiles share the 160 KB of a CU
    if (const char* e = std::getenv("VLX_MAX_TILE_LDS"))
        ctx->max_tile_lds = std::max(16 * 1024, std::min(160 * 1024 - (int)(BOYS_TAB_SIZE * sizeof(double)), std::atoi(e)));
    if (const char* e = std::getenv("VLX_UNTILED_WAVES"))
        ctx->untiled_wave_cap = std::max(VLX_WAVES_PER_BLOCK, std::atoi(e));
    if (const char* e = std::getenv("VLX_SOLO_FRAC")) ctx->solo_frac = std::max(0.0, std::atof(e));
    if (const char* e = std::getenv("VLX_KET_ORDER")) ctx->ket_order = std::max(0, std::min(2, std::atoi(e)));
    if (const char* e = std::getenv("VLX_VARIANT")) ctx->force_variant = std::max(-1, std::min(NVARIANTS - 1, std::atoi(e)));
    HIP_TRY(hipSetDevice(device));
    if (device >= MAX_DEVICES) return fail(VLX_ERR_ARG, "device %d beyond the %d the library supports", device, MAX_DEVICES);
    {
        // each device holds its own copy of `scratch`: its address, once per device, with the device current
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        if (!g_scratch_ptr[device]) {
            void* ptr = nullptr;
            HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
            g_scratch_ptr[device] = static_cast<double*>(ptr);
        }
    }
    HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    for (int k = 0; k < NSIDE; k++) {
        HIP_TRY(hipStreamCreateWithFlags(&ctx->side[k], hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&ctx->ev_done[k], hipEventDisableTiming));
    }
    HIP_TRY(hipEventCreateWithFlags(&ctx->ev_start, hipEventDisableTiming));
    std::vector<double> tab((size_t)(VLX_MAX_LTOT + 1) * BOYS_TAB_SIZE);
    vlx_build_boys_table(tab.data());
    HIP_TRY(ctx->d_boys.upload(tab));
    std::vector<float> tabf(tab.begin(), tab.end());
    HIP_TRY(ctx->d_boys_f.upload(tabf));
    *out = ctx.release();
    return VLX_OK;
}

void vlx_context_destroy(vlx_context* ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& p : ctx->pending) {
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    for (auto ev : ctx->event_pool) (void)hipEventDestroy(ev);
    for (int k = 0; k < NSIDE; k++) {
        if (ctx->side[k]) (void)hipStreamSynchronize(ctx->side[k]);
        if (ctx->side[k]) (void)hipStreamDestroy(ctx->side[k]);
        if (ctx->ev_done[k]) (void)hipEventDestroy(ctx->ev_done[k]);
    }
    if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int vlx_set_block_size_factor(vlx_context* ctx, int factor)
{
    if (!ctx) return fail(VLX_ERR_ARG, "context is NULL");
    if (factor < 1 || factor > 128) return fail(VLX_ERR_ARG, "block size factor %d outside 1..128", factor);
    ctx->block_size_factor = factor;
    return VLX_OK;
}

int vlx_set_precision(vlx_context* ctx, int mode, double threshold)
{
    if (!ctx) return fail(VLX_ERR_ARG, "context is NULL");
    if (mode == VLX_PRECISION_FP64) {
        ctx->mixed_tau = 0.0;
    } else if (mode == VLX_PRECISION_MIXED) {
        if (!(threshold > 0.0) || !std::isfinite(threshold))
            return fail(VLX_ERR_ARG, "mixed-precision threshold must be positive and finite");
        ctx->mixed_tau = threshold;
    } else {
        return fail(VLX_ERR_ARG, "unknown precision mode %d", mode);
    }
    return VLX_OK;
}

int vlx_set_density_screening(vlx_context* ctx, double threshold)
{
    if (!ctx) return fail(VLX_ERR_ARG, "context is NULL");
    if (!(threshold >= 0.0) || !std::isfinite(threshold))
        return fail(VLX_ERR_ARG, "density screening threshold must be finite and >= 0 (0 = off)");
    ctx->dscreen_tau = threshold;
    return VLX_OK;
}

int vlx_set_deterministic(vlx_context* ctx, int enable)
{
    if (!ctx) return fail(VLX_ERR_ARG, "context is NULL");
    ctx->deterministic = enable != 0;
    return VLX_OK;
}

int vlx_context_last_fp32_bin_sum(vlx_context* ctx, int* k32)
{
    if (!ctx || !k32) return fail(VLX_ERR_ARG, "NULL argument");
    *k32 = ctx->last_k32;
    return VLX_OK;
}

int vlx_fock_precision_split(vlx_screener* scr, int ithreshold, int k32, long long* prim_slots_fp64,
                             long long* prim_slots_fp32)
{
    if (!scr || !prim_slots_fp64 || !prim_slots_fp32) return fail(VLX_ERR_ARG, "NULL argument");
    if (k32 < 0) return fail(VLX_ERR_ARG, "bin-sum threshold must be >= 0");
    std::vector<std::unique_ptr<ClassSchedule>>* list = nullptr;
    int rc = get_schedules(scr, ithreshold, std::min(k32, NO_FP32), &list);
    if (rc) return rc;
    *prim_slots_fp64 = *prim_slots_fp32 = 0;
    for (const auto& sp : *list) {
        *prim_slots_fp64 += sp->prim_slots - sp->prim_slots32;
        *prim_slots_fp32 += sp->prim_slots32;
    }
    return VLX_OK;
}

int vlx_fock_precision_classes(vlx_screener* scr, int ithreshold, int k32, int max_records, int* nrecords,
                               int* classes, long long* prim_slots_fp32)
{
    if (!scr || !nrecords) return fail(VLX_ERR_ARG, "NULL argument");
    if (k32 < 0) return fail(VLX_ERR_ARG, "bin-sum threshold must be >= 0");
    std::vector<std::unique_ptr<ClassSchedule>>* list = nullptr;
    int rc = get_schedules(scr, ithreshold, std::min(k32, NO_FP32), &list);
    if (rc) return rc;
    *nrecords = static_cast<int>(list->size());
    for (int r = 0; r < std::min(max_records, *nrecords); r++) {
        const auto& ce = g_class_table[(*list)[r]->entry];
        if (classes) {
            classes[4 * r] = ce.la;
            classes[4 * r + 1] = ce.lb;
            classes[4 * r + 2] = ce.lc;
            classes[4 * r + 3] = ce.ld;
        }
        if (prim_slots_fp32) prim_slots_fp32[r] = (*list)[r]->prim_slots32;
    }
    return VLX_OK;
}

int vlx_context_synchronize(vlx_context* ctx)
{
    if (!ctx) return fail(VLX_ERR_ARG, "context is NULL");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return VLX_OK;
}

void* vlx_context_stream(vlx_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int vlx_device_malloc(vlx_context* ctx, unsigned long long bytes, void** out)
{
    if (!ctx || !out) return fail(VLX_ERR_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipMalloc(out, std::max<unsigned long long>(bytes, 1)));
    return VLX_OK;
}

int vlx_device_free(vlx_context* ctx, void* ptr)
{
    if (!ctx) return fail(VLX_ERR_ARG, "NULL argument");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ptr) HIP_TRY(hipFree(ptr));
    return VLX_OK;
}

int vlx_copy_to_device(vlx_context* ctx, void* dst, const void* src, unsigned long long bytes)
{
    if (!ctx || !dst || !src) return fail(VLX_ERR_ARG, "NULL argument");
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return VLX_OK;
}

int vlx_copy_to_host(vlx_context* ctx, void* dst, const void* src, unsigned long long bytes)
{
    if (!ctx || !dst || !src) return fail(VLX_ERR_ARG, "NULL argument");
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return VLX_OK;
}

int vlx_screener_create(vlx_context* ctx, int nshells, const int* l, const int* nprim, const int* prim_offset,
                        const double* exps, const double* norms, const double* centers, const int* ao_base,
                        const int* ao_stride, int nao, vlx_screener** out)
{
    if (!ctx || !out || !l || !nprim || !prim_offset || !exps || !norms || !centers || !ao_base || !ao_stride)
        return fail(VLX_ERR_ARG, "NULL argument to vlx_screener_create");
    if (nshells <= 0 || nao <= 0) return fail(VLX_ERR_ARG, "empty basis");
    HIP_TRY(hipSetDevice(ctx->device));
    auto s = std::make_unique<vlx_screener>();
    s->ctx = ctx;
    s->nshells = nshells;
    s->nao = nao;
    int nprims_total = 0;
    for (int i = 0; i < nshells; i++) {
        if (l[i] < 0 || l[i] > VLX_GEN_LMAX)
            return fail(VLX_ERR_UNSUPPORTED, "shell %d has L=%d; kernels are built up to L=%d", i, l[i], VLX_GEN_LMAX);
        if (nprim[i] <= 0) return fail(VLX_ERR_ARG, "shell %d has no primitives", i);
        const int stride = ao_stride[i];
        if (ao_base[i] < 0 || stride <= 0 || ao_base[i] + 2 * l[i] * stride >= nao)
            return fail(VLX_ERR_ARG, "shell %d AO indices outside 0..%d", i, nao - 1);
        nprims_total = std::max(nprims_total, prim_offset[i] + nprim[i]);
    }
    s->l.assign(l, l + nshells);
    s->nprim.assign(nprim, nprim + nshells);
    s->prim_offset.assign(prim_offset, prim_offset + nshells);
    s->ao_base.assign(ao_base, ao_base + nshells);
    s->ao_stride.assign(ao_stride, ao_stride + nshells);
    s->exps.assign(exps, exps + nprims_total);
    s->norms.assign(norms, norms + nprims_total);
    s->centers.assign(centers, centers + 3 * nshells);

    // atoms: runs of shells with identical centres; groups: consecutive atoms
    s->shell_group.assign(nshells, 0);
    s->shell_loc.assign(nshells, 0);
    int group_atoms = ctx->group_atoms;
    if (group_atoms <= 0) {
        // functions per (atom, l), atoms = runs of shells on one centre
        std::vector<std::array<int, 4>> per_atom;
        for (int sh = 0; sh < nshells; sh++) {
            const bool new_atom = sh == 0 || centers[3 * sh] != centers[3 * sh - 3] ||
                                  centers[3 * sh + 1] != centers[3 * sh - 2] ||
                                  centers[3 * sh + 2] != centers[3 * sh - 1];
            if (new_atom) per_atom.push_back({0, 0, 0, 0});
            per_atom.back()[l[sh]] += 2 * l[sh] + 1;
        }
        for (int g : {12, 8, 6, 4, 3, 2, 1}) {
            group_atoms = g;
            int gm0 = 0, gm1 = 0;
            for (size_t a0 = 0; a0 < per_atom.size(); a0 += g) {
                int n0 = 0, n1 = 0;
                for (size_t a = a0; a < std::min(per_atom.size(), a0 + g); a++) {
                    n0 += per_atom[a][0];
                    n1 += per_atom[a][1];
                }
                gm0 = std::max(gm0, n0);
                gm1 = std::max(gm1, n1);
            }
            // tiles of (ps|ss): AB, AC, AD = p x s; CD, BC, BD = s x s
            if (8LL * (3LL * gm1 * gm0 + 3LL * gm0 * gm0) <= ctx->max_tile_lds) break;
        }
    }
    s->group_atoms = group_atoms;
    {
        int atom = 0, group = 0, atoms_in_group = 0;
        std::vector<int> count(4, 0);
        std::vector<std::vector<int>> ao_lists;  // per (group, l)
        ao_lists.assign(4, {});
        std::vector<std::vector<int>> all_lists;
        for (int sh = 0; sh < nshells; sh++) {
            const bool new_atom = sh > 0 && (centers[3 * sh] != centers[3 * sh - 3] ||
                                             centers[3 * sh + 1] != centers[3 * sh - 2] ||
                                             centers[3 * sh + 2] != centers[3 * sh - 1]);
            if (new_atom) {
                atom++;
                atoms_in_group++;
                if (atoms_in_group >= group_atoms) {
                    for (int l = 0; l < 4; l++) all_lists.push_back(ao_lists[l]);
                    ao_lists.assign(4, {});
                    group++;
                    atoms_in_group = 0;
                }
            }
            s->shell_group[sh] = group;
            s->shell_loc[sh] = static_cast<int>(ao_lists[l[sh]].size());
            for (int m = 0; m < 2 * l[sh] + 1; m++) ao_lists[l[sh]].push_back(ao_base[sh] + m * ao_stride[sh]);
        }
        for (int ll = 0; ll < 4; ll++) all_lists.push_back(ao_lists[ll]);
        s->ngroups = group + 1;
        s->grp_off.assign(s->ngroups * 4, 0);
        s->grp_n.assign(s->ngroups * 4, 0);
        s->grp_ao.clear();
        for (int g = 0; g < s->ngroups; g++)
            for (int ll = 0; ll < 4; ll++) {
                const auto& v = all_lists[g * 4 + ll];
                s->grp_off[g * 4 + ll] = static_cast<int>(s->grp_ao.size());
                s->grp_n[g * 4 + ll] = static_cast<int>(v.size());
                s->grp_max[ll] = std::max(s->grp_max[ll], static_cast<int>(v.size()));
                s->grp_ao.insert(s->grp_ao.end(), v.begin(), v.end());
            }
        (void)atom;
        (void)count;
        HIP_TRY(s->d_grp_off.upload(s->grp_off));
        HIP_TRY(s->d_grp_n.upload(s->grp_n));
        HIP_TRY(s->d_grp_ao.upload(s->grp_ao));
    }
    HIP_TRY(s->d_mol.upload(std::vector<MolDesc>{MolDesc{0, s->nao, 0}}));
    {
        std::vector<int> ao_group(nao, 0);
        for (int sh = 0; sh < nshells; sh++)
            for (int m = 0; m < 2 * l[sh] + 1; m++) ao_group[ao_base[sh] + m * ao_stride[sh]] = s->shell_group[sh];
        HIP_TRY(s->d_ao_group.upload(ao_group));
    }
    const int npc = npair_classes();
    s->all.resize(npc);
    s->kept_tiled.resize(npc);
    s->all_tiled.resize(npc);
    s->kept_tiled_bra.resize(npc);
    s->all_tiled_bra.resize(npc);
    s->all_bra.resize(npc);
    s->kept.resize(npc);
    s->kept_bra.resize(npc);
    for (int la = 0; la <= VLX_GEN_LMAX; la++)
        for (int lb = 0; lb <= la; lb++) {
            s->all[pair_class_index(la, lb)].la = la;
            s->all[pair_class_index(la, lb)].lb = lb;
            for (auto* v : {&s->all_bra, &s->kept, &s->kept_bra, &s->kept_tiled, &s->all_tiled, &s->kept_tiled_bra,
                            &s->all_tiled_bra}) {
                (*v)[pair_class_index(la, lb)].la = la;
                (*v)[pair_class_index(la, lb)].lb = lb;
            }
        }
    s->pair_ref.assign((size_t)nshells * nshells, {});
    // every unordered shell pair once, oriented l_A >= l_B (A >= B by index on ties)
    for (int sa = 0; sa < nshells; sa++)
        for (int sb = 0; sb < nshells; sb++) {
            const bool keep = (l[sa] > l[sb]) || (l[sa] == l[sb] && sa >= sb);
            if (!keep) continue;
            const int x = pair_class_index(l[sa], l[sb]);
            PairSet& ps = s->all[x];
            const int idx = ps.size();
            add_pair(*s, ps, sa, sb);
            s->pair_ref[(size_t)sa * nshells + sb] = {x, idx, 0};
            s->pair_ref[(size_t)sb * nshells + sa] = {x, idx, 1};
        }
    // diagonal integrals on the GPU
    s->pair_max.assign((size_t)nshells * nshells, 0.0);
    for (int x = 0; x < npc; x++) {
        PairSet& ps = s->all[x];
        if (ps.size() == 0) continue;
        int rc = upload_set(ps);
        if (rc) return rc;
        const DiagLauncher* dl = find_diag(x);
        if (!dl) return fail(VLX_ERR_UNSUPPORTED, "no diagonal kernel for pair class %d", x);
        DevBuf<double> d_max;
        HIP_TRY(d_max.alloc(ps.size()));
        DiagArgs da{ps.d_pairs.ptr, ps.d_prims.ptr, ps.size(),
                     ctx->d_boys.ptr + (size_t)(2 * (ps.la + ps.lb)) * BOYS_TAB_SIZE, d_max.ptr};
        dl->diag(da, ctx->stream);
        HIP_TRY(hipGetLastError());
        std::vector<double> mx(ps.size());
        HIP_TRY(hipMemcpyAsync(mx.data(), d_max.ptr, mx.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        ps.bins.resize(ps.size());
        for (int i = 0; i < ps.size(); i++) {
            const int sa = ps.shell_a[i], sb = ps.shell_b[i];
            s->pair_max[(size_t)sa * nshells + sb] = mx[i];
            s->pair_max[(size_t)sb * nshells + sa] = mx[i];
            ps.bins[i] = pair_bin(mx[i]);
            ps.pairs[i].bin = ps.bins[i];
        }
        // kept set (ket order): bin >= 0, sorted by (bin, nprim); inside equal (bin, nprim)
        // the pairs go in shell order (ket_order 0), scattered (1) or along diagonals of equal
        // (C - D) AO offset (2)
        std::vector<int> order;
        for (int i = 0; i < ps.size(); i++)
            if (ps.bins[i] >= 0) order.push_back(i);
        const int ko = ctx->ket_order;
        auto mix = [](int i) { return static_cast<unsigned>(i) * 2654435761u; };
        std::stable_sort(order.begin(), order.end(), [&](int i, int j) {
            const PairInfo &x = ps.pairs[i], &y = ps.pairs[j];
            if (ps.bins[i] != ps.bins[j]) return ps.bins[i] < ps.bins[j];
            if (x.nprim != y.nprim) return x.nprim < y.nprim;
            if (ko == 1) return mix(i) < mix(j);
            if (ko == 2) {
                if (x.ao_a - x.ao_b != y.ao_a - y.ao_b) return x.ao_a - x.ao_b < y.ao_a - y.ao_b;
                return x.ao_a < y.ao_a;
            }
            return false;
        });
        PairSet& ks = s->kept[x];
        reorder_set(*s, ps, order, ks);
        for (int k = 0; k < ks.size(); k++) ks.pairs[k].rank = k;
        // bra order: (bin, shell A, nprim) over the kept set, ranks from the ket order
        std::vector<int> border(ks.size());
        std::iota(border.begin(), border.end(), 0);
        std::stable_sort(border.begin(), border.end(), [&](int i, int j) {
            if (ks.pairs[i].bin != ks.pairs[j].bin) return ks.pairs[i].bin < ks.pairs[j].bin;
            if (ks.shell_a[i] != ks.shell_a[j]) return ks.shell_a[i] < ks.shell_a[j];
            return ks.pairs[i].nprim < ks.pairs[j].nprim;
        });
        reorder_set(*s, ks, border, s->kept_bra[x]);
        // unscreened lists: every pair, bins ignored
        for (int k = 0; k < ps.size(); k++) ps.pairs[k].rank = k;
        finalize_set(ps);
        std::vector<int> aorder(ps.size());
        std::iota(aorder.begin(), aorder.end(), 0);
        std::stable_sort(aorder.begin(), aorder.end(), [&](int i, int j) { return ps.shell_a[i] < ps.shell_a[j]; });
        reorder_set(*s, ps, aorder, s->all_bra[x]);
        build_tiled(*s, ks, s->kept_tiled[x], s->kept_tiled_bra[x]);
        build_tiled(*s, ps, s->all_tiled[x], s->all_tiled_bra[x]);
        for (PairSet* up : {&ps, &ks, &s->kept_bra[x], &s->all_bra[x], &s->kept_tiled[x], &s->all_tiled[x],
                            &s->kept_tiled_bra[x], &s->all_tiled_bra[x]}) {
            rc = upload_set(*up);
            if (rc) return rc;
        }
    }
    *out = s.release();
    return VLX_OK;
}

void vlx_screener_destroy(vlx_screener* scr)
{
    if (!scr) return;
    (void)hipSetDevice(scr->ctx->device);
    (void)hipStreamSynchronize(scr->ctx->stream);
    // batch plans that hold this screener's pairs go with it
    auto& plans = scr->ctx->plans;
    plans.erase(std::remove_if(plans.begin(), plans.end(),
                               [scr](const std::unique_ptr<BatchPlan>& p) {
                                   return std::find(p->scrs.begin(), p->scrs.end(), scr) != p->scrs.end();
                               }),
                plans.end());
    delete scr;
}

int vlx_screener_pair_maxima(const vlx_screener* scr, double* out)
{
    if (!scr || !out) return fail(VLX_ERR_ARG, "NULL argument");
    std::copy(scr->pair_max.begin(), scr->pair_max.end(), out);
    return VLX_OK;
}

int vlx_screener_number_of_pairs(const vlx_screener* scr, long long* out)
{
    if (!scr || !out) return fail(VLX_ERR_ARG, "NULL argument");
    long long n = 0;
    for (const auto& ks : scr->kept) n += ks.size();
    *out = n;
    return VLX_OK;
}

int vlx_fock_schedule_stats(vlx_screener* scr, int ithreshold, int max_records, int* nrecords,
                            long long* lane_slots, long long* prim_slots)
{
    if (!scr || !nrecords) return fail(VLX_ERR_ARG, "NULL argument");
    std::vector<std::unique_ptr<ClassSchedule>>* list = nullptr;
    int rc = get_schedules(scr, ithreshold, NO_FP32, &list);
    if (rc) return rc;
    *nrecords = static_cast<int>(list->size());
    for (int r = 0; r < std::min(max_records, *nrecords); r++) {
        if (lane_slots) lane_slots[r] = (*list)[r]->lane_slots;
        if (prim_slots) prim_slots[r] = (*list)[r]->prim_slots;
    }
    return VLX_OK;
}

int vlx_context_set_streams(vlx_context* ctx, int nstreams)
{
    if (!ctx) return fail(VLX_ERR_ARG, "NULL context");
    if (nstreams < 1 || nstreams > NSIDE) return fail(VLX_ERR_ARG, "nstreams %d outside 1..%d", nstreams, NSIDE);
    ctx->nstreams = nstreams;
    return VLX_OK;
}

int vlx_context_set_variant(vlx_context* ctx, int variant)
{
    if (!ctx) return fail(VLX_ERR_ARG, "NULL context");
    if (variant < -1 || variant >= NVARIANTS) return fail(VLX_ERR_ARG, "variant %d outside -1..%d", variant, NVARIANTS - 1);
    ctx->force_variant = variant;
    return VLX_OK;
}

int vlx_context_set_timing(vlx_context* ctx, int enable)
{
    if (!ctx) return fail(VLX_ERR_ARG, "context is NULL");
    ctx->timing = enable != 0;
    return VLX_OK;
}

static int drain_timing(vlx_context* ctx)
{
    if (ctx->pending.empty()) return VLX_OK;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < NSIDE; k++) HIP_TRY(hipStreamSynchronize(ctx->side[k]));
    for (auto& p : ctx->pending) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, p.start, p.stop));
        auto& acc = ctx->class_ms[p.entry];
        acc.first += ms;
        acc.second += 1;
        ctx->event_pool.push_back(p.start);
        ctx->event_pool.push_back(p.stop);
    }
    ctx->pending.clear();
    return VLX_OK;
}

int vlx_context_reset_times(vlx_context* ctx)
{
    if (!ctx) return fail(VLX_ERR_ARG, "context is NULL");
    int rc = drain_timing(ctx);
    ctx->class_ms.clear();
    return rc;
}

int vlx_context_class_times(vlx_context* ctx, int max_records, int* nrecords, int* classes, double* total_ms,
                            long long* launches)
{
    if (!ctx || !nrecords) return fail(VLX_ERR_ARG, "NULL argument");
    int rc = drain_timing(ctx);
    if (rc) return rc;
    *nrecords = static_cast<int>(ctx->class_ms.size());
    int r = 0;
    for (const auto& kv : ctx->class_ms) {
        if (r >= max_records) break;
        const auto& ce = g_class_table[kv.first];
        if (classes) {
            classes[4 * r] = ce.la;
            classes[4 * r + 1] = ce.lb;
            classes[4 * r + 2] = ce.lc;
            classes[4 * r + 3] = ce.ld;
        }
        if (total_ms) total_ms[r] = kv.second.first;
        if (launches) launches[r] = kv.second.second;
        r++;
    }
    return VLX_OK;
}

int vlx_fock_autotune(vlx_context* ctx, vlx_screener* scr, const double* d_density, int ithreshold, int repeats,
                      int* nchosen, int* classes, int* variants)
{
    if (!ctx || !scr || !d_density) return fail(VLX_ERR_ARG, "NULL argument to vlx_fock_autotune");
    if (scr->ctx != ctx) return fail(VLX_ERR_ARG, "screener belongs to another context");
    if (ctx->deterministic) return fail(VLX_ERR_ARG, "no autotuning in the deterministic mode (one kernel family)");
    if (repeats < 1) repeats = 1;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t nn = (size_t)scr->nao * scr->nao;
    HIP_TRY(ctx->d_fock.alloc(nn));
    // time every class under each variant on one stream (no overlap between classes)
    const int saved_streams = ctx->nstreams;
    const bool saved_timing = ctx->timing;
    const double saved_tau = ctx->mixed_tau;
    const int saved_force = ctx->force_variant;
    ctx->nstreams = 1;
    ctx->mixed_tau = 0.0;
    std::map<int, double> best_ms;
    std::map<int, int> best_var;
    int rc = VLX_OK;
    // the warm-up G of every variant is checked against the untiled variant's (variant 0, run
    // first): a variant whose G has a non-finite element or differs by more than 1e-10 is never
    // chosen (the choice is made on timing alone otherwise)
    std::vector<double> g_ref(nn), g_var(nn);
    ctx->autotune_rejected = 0;
    for (int v = 0; v < NVARIANTS && rc == VLX_OK; v++) {
        ctx->force_variant = v;
        scr->schedules.clear();
        ctx->timing = false;
        rc = vlx_fock_compute_device(ctx, scr, d_density, "2jk", 1.0, 0.0, ithreshold, ctx->d_fock.ptr);  // warm-up
        if (rc) break;
        rc = drain_timing(ctx);
        if (rc) break;
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        std::vector<double>& g = v == 0 ? g_ref : g_var;
        HIP_TRY(hipMemcpy(g.data(), ctx->d_fock.ptr, nn * sizeof(double), hipMemcpyDeviceToHost));
        bool ok = true;
        for (size_t e = 0; e < nn && ok; e++)
            ok = std::isfinite(g[e]) && (v == 0 || std::fabs(g[e] - g_ref[e]) <= 1.0e-10);
        if (!ok) {
            ctx->autotune_rejected |= 1 << v;
            if (v == 0) rc = fail(VLX_ERR_HIP, "autotune: the untiled variant's G has non-finite elements");
            continue;
        }
        ctx->class_ms.clear();
        ctx->timing = true;
        for (int r = 0; r < repeats && rc == VLX_OK; r++)
            rc = vlx_fock_compute_device(ctx, scr, d_density, "2jk", 1.0, 0.0, ithreshold, ctx->d_fock.ptr);
        if (rc) break;
        rc = drain_timing(ctx);
        if (rc) break;
        // classes that cannot run this variant ran another one: they count only for the one they ran
        std::map<int, int> ran;
        std::vector<std::unique_ptr<ClassSchedule>>* list = nullptr;
        rc = get_schedules(scr, ithreshold, NO_FP32, &list);
        if (rc) break;
        for (const auto& sp : *list) ran[sp->entry] = sp->variant;
        for (const auto& kv : ctx->class_ms) {
            if (ran[kv.first] != v) continue;
            const double ms = kv.second.first / repeats;
            auto it = best_ms.find(kv.first);
            if (it == best_ms.end() || ms < it->second) {
                best_ms[kv.first] = ms;
                best_var[kv.first] = v;
            }
        }
    }
    ctx->force_variant = saved_force;
    ctx->nstreams = saved_streams;
    ctx->timing = saved_timing;
    ctx->mixed_tau = saved_tau;
    ctx->class_ms.clear();
    scr->schedules.clear();
    if (rc) return rc;
    scr->variant = best_var;
    scr->tuned_ms = best_ms;
    scr->epoch++;
    for (const auto& kv : best_var) ctx->tuned_variant[kv.first] = kv.second;
    if (nchosen) *nchosen = static_cast<int>(best_var.size());
    int r = 0;
    for (const auto& kv : best_var) {
        const auto& ce = g_class_table[kv.first];
        if (classes) {
            classes[4 * r] = ce.la;
            classes[4 * r + 1] = ce.lb;
            classes[4 * r + 2] = ce.lc;
            classes[4 * r + 3] = ce.ld;
        }
        if (variants) variants[r] = kv.second;
        r++;
    }
    return VLX_OK;
}

int vlx_fock_get_variants(vlx_screener* scr, int max_records, int* nrecords, int* classes, int* variants, double* ms)
{
    if (!scr || !nrecords) return fail(VLX_ERR_ARG, "NULL argument");
    int r = 0;
    for (const auto& kv : scr->variant) {
        if (r < max_records) {
            const auto& ce = g_class_table[kv.first];
            if (classes) {
                classes[4 * r] = ce.la;
                classes[4 * r + 1] = ce.lb;
                classes[4 * r + 2] = ce.lc;
                classes[4 * r + 3] = ce.ld;
            }
            if (variants) variants[r] = kv.second;
            if (ms) {
                const auto it = scr->tuned_ms.find(kv.first);
                ms[r] = it != scr->tuned_ms.end() ? it->second : 0.0;
            }
        }
        r++;
    }
    *nrecords = r;
    return VLX_OK;
}

int vlx_fock_set_variants(vlx_context* ctx, vlx_screener* scr, int nrecords, const int* classes, const int* variants,
                          const double* ms)
{
    if (!ctx || !scr || nrecords < 0 || (nrecords > 0 && (!classes || !variants)))
        return fail(VLX_ERR_ARG, "NULL argument to vlx_fock_set_variants");
    if (scr->ctx != ctx) return fail(VLX_ERR_ARG, "screener belongs to another context");
    std::map<int, int> var;
    std::map<int, double> tms;
    const int nentries = static_cast<int>(sizeof(g_class_table) / sizeof(g_class_table[0]));
    for (int r = 0; r < nrecords; r++) {
        int e = -1;
        for (int k = 0; k < nentries && e < 0; k++) {
            const auto& ce = g_class_table[k];
            if (ce.la == classes[4 * r] && ce.lb == classes[4 * r + 1] && ce.lc == classes[4 * r + 2] &&
                ce.ld == classes[4 * r + 3])
                e = k;
        }
        if (e < 0)
            return fail(VLX_ERR_ARG, "no class (%d %d|%d %d) in the kernel table", classes[4 * r], classes[4 * r + 1],
                        classes[4 * r + 2], classes[4 * r + 3]);
        if (variants[r] < 0 || variants[r] >= NVARIANTS) return fail(VLX_ERR_ARG, "variant %d out of range", variants[r]);
        var[e] = variants[r];
        if (ms) tms[e] = ms[r];
    }
    scr->variant = var;
    scr->tuned_ms = tms;
    scr->epoch++;
    scr->schedules.clear();
    return VLX_OK;
}

int vlx_fock_autotune_rejected(vlx_context* ctx, int* mask)
{
    if (!ctx || !mask) return fail(VLX_ERR_ARG, "NULL argument");
    *mask = ctx->autotune_rejected;
    return VLX_OK;
}

int vlx_fock_class_kernel(vlx_context* ctx, vlx_screener* scr, int ithreshold, int la, int lb, int lc, int ld,
                          char* name, int len)
{
    if (!ctx || !scr || !name || len < 1) return fail(VLX_ERR_ARG, "NULL argument");
    std::vector<std::unique_ptr<ClassSchedule>>* list = nullptr;
    int rc = get_schedules(scr, ithreshold, NO_FP32, &list);
    if (rc) return rc;
    for (const auto& sp : *list) {
        const auto& ce = g_class_table[sp->entry];
        if (ce.la != la || ce.lb != lb || ce.lc != lc || ce.ld != ld) continue;
        if (sp->tiled) {
            const int L = la + lb + lc + ld;
            const int wide = (L <= 2) ? 16 : (L == 3 ? 8 : 4);
            // template arguments as the profiler prints them: <..., R, NW, RS, MULTI, DLDS>
            std::snprintf(name, len, "fock_kernel_tiled<%d,%d,%d,%d,double,%d,false,false,%s>", la, lb, lc, ld,
                          sp->tiled_waves == 4 ? 4 : wide, sp->dlds ? "true" : "false");
        } else {
            const char* k = sp->mmd                             ? "fock_kernel_mmd"
                            : sp->ksl                           ? "fock_kernel_ksl"
                            : (ce.variant == VLX_KET_STATIONARY) ? "fock_kernel"
                                                                : "fock_kernel_bs";
            if (sp->ksl) std::snprintf(name, len, "%s<%d,%d,%d,%d,false,4>", k, la, lb, lc, ld);
            else std::snprintf(name, len, "%s<%d,%d,%d,%d,false>", k, la, lb, lc, ld);
        }
        return VLX_OK;
    }
    return fail(VLX_ERR_ARG, "class (%d%d|%d%d) has no work at this threshold", la, lb, lc, ld);
}

int vlx_fock_work(vlx_screener* scr, int ithreshold, int max_records, int* nrecords, int* classes,
                  long long* quartets, long long* prim_quartets, double* flops)
{
    if (!scr || !nrecords) return fail(VLX_ERR_ARG, "NULL argument");
    if (ithreshold > 30) return fail(VLX_ERR_ARG, "ithreshold %d outside the accepted range (< 0 or 0..30)", ithreshold);
    std::vector<std::unique_ptr<ClassSchedule>>* list = nullptr;
    int rc = get_schedules(scr, ithreshold, NO_FP32, &list);
    if (rc) return rc;
    *nrecords = static_cast<int>(list->size());
    for (int r = 0; r < std::min(max_records, *nrecords); r++) {
        const auto& sch = *(*list)[r];
        const auto& ce = g_class_table[sch.entry];
        if (classes) {
            classes[4 * r] = ce.la;
            classes[4 * r + 1] = ce.lb;
            classes[4 * r + 2] = ce.lc;
            classes[4 * r + 3] = ce.ld;
        }
        if (quartets) quartets[r] = sch.quartets;
        if (prim_quartets) prim_quartets[r] = sch.prim_quartets;
        if (flops) {
            const int L = ce.la + ce.lb + ce.lc + ce.ld;
            const int nsph = (2 * ce.la + 1) * (2 * ce.lb + 1) * (2 * ce.lc + 1) * (2 * ce.ld + 1);
            flops[r] = (double)sch.prim_quartets * (ce.vrr_flops + 53 + 4 * L) +
                       (double)sch.quartets * (ce.hrr_flops + 12.0 * nsph);
        }
    }
    return VLX_OK;
}

namespace {

// One build on a device density: F' by the class kernels, then
// G = F' + F'^T into d_out (anti_coef == 0), or d_out += anti_coef (F' - F'^T)
// (the exchange of the antisymmetric part of a general density).
// Enqueues the class kernels of one build that accumulate F' (nao x nao) from D: on `stream`, or,
// when it is null, on the side stream each class schedule was given.  solo: 1 only the classes
// that run alone (stream -1), 0 only the others, -1 all.  No synchronisation.
// ndens > 1: the multi-density kernels digest each quartet into ndens densities, D and F' of
// density c at d_density + c nao^2 and F + c nao^2, factors (fj, fk) of density c at dfac[2 c]
// (device, and on the host in ctx->h_dfac); fj and fk are then unused.  det: the same kernels with
// fixed-point accumulation, F being the accumulator array (two words per element); the schedules
// must be untiled.
int enqueue_classes(vlx_context* ctx, vlx_screener* scr, const std::vector<std::unique_ptr<ClassSchedule>>& list,
                    const double* d_density, double* F, double fj, double fk, double omega2, int ithreshold,
                    hipStream_t stream, int solo = -1, bool allow_timing = true, const int* glim = nullptr,
                    int ndens = 1, const double* dfac = nullptr, bool det = false)
{
    const bool multi = ndens > 1 || det;
    const bool timing = ctx->timing && allow_timing;
    const bool unscreened = ithreshold < 0;
    const long long nn = (long long)scr->nao * scr->nao;
    for (const auto& schp : list) {
        if (!multi && fj == 0.0 && fk == 0.0) break;  // nothing to digest: G = 0
        const ClassSchedule& sch = *schp;
        if (solo >= 0 && (sch.stream < 0) != (solo == 1)) continue;
        const auto& ce = g_class_table[sch.entry];
        const PairSet& bra = sch.tiled ? (unscreened ? scr->all_tiled_bra[ce.x] : scr->kept_tiled_bra[ce.x])
                                       : (unscreened ? scr->all_bra[ce.x] : scr->kept_bra[ce.x]);
        const PairSet& ket = sch.tiled ? (unscreened ? scr->all_tiled[ce.y] : scr->kept_tiled[ce.y])
                                       : (unscreened ? scr->all[ce.y] : scr->kept[ce.y]);
        FockArgs a{};
        a.bra = bra.d_pairs.ptr;
        a.bra_prims = bra.d_prims.ptr;
        a.ket = ket.d_pairs.ptr;
        a.ket_prims = ket.d_prims.ptr;
        a.bra_u = bra.d_u.ptr;
        a.ket_u = ket.d_u.ptr;
        a.items = sch.d_items.ptr;
        a.boys_tab = ctx->d_boys.ptr + (size_t)(ce.la + ce.lb + ce.lc + ce.ld) * BOYS_TAB_SIZE;
        a.boys_tab_f = ctx->d_boys_f.ptr + (size_t)(ce.la + ce.lb + ce.lc + ce.ld) * BOYS_TAB_SIZE;
        a.D = d_density;
        a.F = F;
        a.mols = scr->d_mol.ptr;
        a.k32 = nullptr;
        a.glim = glim;
        a.sc.ngroups = scr->ngroups;
        a.titems = sch.d_titems.ptr;
        a.kentries = sch.singles ? ket.singles.d_kentries.ptr : ket.runs.d_kentries.ptr;
        a.grp_off = scr->d_grp_off.ptr;
        a.grp_n = scr->d_grp_n.ptr;
        a.grp_ao = scr->d_grp_ao.ptr;
        a.tile_lds_bytes = sch.tile_lds;
        a.tiled_waves = sch.tiled_waves;
        a.sc.nitems = static_cast<long long>(sch.tiled ? sch.titems.size() : sch.items.size());
        a.sc.nket = ket.size();
        a.sc.ithreshold = ithreshold;
        a.sc.same_class = (ce.x == ce.y) ? 1 : 0;
        a.sc.nao = scr->nao;
        a.sc.fj = fj;
        a.sc.fk = fk;
        a.sc.omega2 = omega2;
        a.sc.ndens = ndens;
        a.sc.dstride = nn;
        a.sc.det = det ? 1 : 0;
        a.sc.bs_layout = (ce.variant == VLX_BRA_STATIONARY && !sch.ksl) ? 1 : 0;
        a.dfac = dfac;
        const int grid = sch.tiled ? static_cast<int>(std::min<long long>(a.sc.nitems, ctx->tiled_grid))
                         : sch.mmd ? static_cast<int>(std::min<long long>(a.sc.nitems, ctx->mmd_grid))
                                   : static_cast<int>(untiled_waves(a.sc.nitems, ctx->untiled_wave_cap) / VLX_WAVES_PER_BLOCK);
        hipStream_t st = stream ? stream : (sch.stream < 0 ? ctx->stream : ctx->side[sch.stream % ctx->nstreams]);
        vlx_context::Pending pend{sch.entry, nullptr, nullptr};
        if (timing) {
            for (hipEvent_t* ev : {&pend.start, &pend.stop}) {
                if (ctx->event_pool.empty()) {
                    HIP_TRY(hipEventCreate(ev));
                } else {
                    *ev = ctx->event_pool.back();
                    ctx->event_pool.pop_back();
                }
            }
            HIP_TRY(hipEventRecord(pend.start, st));
        }
        if (multi) {
            if (det && sch.tiled) return fail(VLX_ERR_ARG, "deterministic builds run the untiled kernels only");
            // Kernels whose single-density form keeps accumulators the multi-density form cannot
            // hold run once per density instead (the integrals are recomputed): ks-lds (one density
            // of LDS slots), the ket-stationary register accumulators (the multi-density direct
            // digestion measured 4.8x slower than four launches) and the tiled classes above L = 2
            // or over single primitives (k tile sets cut the occupancy: 1.1-1.9x slower).  MD,
            // bra-stationary and L <= 2 tiled classes share the integrals (0.5-0.95x of k launches,
            // def2-SVP and def2-TZVP water clusters).
            const int L = ce.la + ce.lb + ce.lc + ce.ld;
            const bool per_density =
                !det && (sch.ksl || (!sch.tiled && !sch.mmd && ce.variant == VLX_KET_STATIONARY) ||
                         (sch.tiled && (sch.singles || L > 2)));
            if (per_density) {
                const int waves = static_cast<int>(untiled_waves(a.sc.nitems, ctx->untiled_wave_cap));
                for (int c = 0; c < ndens; c++) {
                    FockArgs a1 = a;
                    a1.sc.ndens = 1;
                    a1.sc.fj = ctx->h_dfac[2 * c];
                    a1.sc.fk = ctx->h_dfac[2 * c + 1];
                    a1.D = d_density + c * nn;
                    a1.F = F + c * nn;
                    if (a1.sc.fj == 0.0 && a1.sc.fk == 0.0) continue;
                    if (sch.ksl) ce.ksl(a1, waves, st);
                    else if (sch.tiled) {
                        if (!sch.titems.empty()) ce.tiled(a1, grid, st);
                    } else ce.fock(a1, grid, st);
                }
            } else if (sch.tiled) {
                // one set of six tiles per density in LDS: as many densities per launch as fit
                // beside the FP64 Boys table; the integrals are recomputed per group of densities
                const int lds_max = 160 * 1024 - static_cast<int>(BOYS_TAB_SIZE * sizeof(double)) - 1024;
                const int chunk = std::max(1, std::min(ndens, lds_max / std::max(1, sch.tile_lds)));
                for (int c0 = 0; c0 < ndens && !sch.titems.empty(); c0 += chunk) {
                    FockArgs am = a;
                    am.sc.ndens = std::min(chunk, ndens - c0);
                    am.D = d_density + c0 * nn;
                    am.F = F + c0 * nn;
                    am.dfac = dfac + 2 * c0;
                    am.tile_lds_bytes = am.sc.ndens * sch.tile_lds;
                    ce.tiled_multi(am, grid, st);
                }
            } else if (sch.mmd) {
                ce.mmd(a, grid, st);
            } else {
                ce.fock_multi(a, grid, st);
            }
        } else if (sch.tiled) {
            if (!sch.titems.empty()) {
                if (sch.dlds) {
                    FockArgs ad = a;
                    ad.tile_lds_bytes = 2 * sch.tile_lds;
                    ce.tiled_dlds(ad, grid, st);
                } else {
                    ce.tiled(a, grid, st);
                }
            }
            if (!sch.titems32.empty()) {
                FockArgs a32 = a;
                a32.titems = sch.d_titems32.ptr;
                a32.sc.nitems = static_cast<long long>(sch.titems32.size());
                const int grid32 = static_cast<int>(std::min<long long>(a32.sc.nitems, ctx->tiled_grid));
                ce.tiled32(a32, grid32, st);
            }
        } else if (sch.mmd) {
            ce.mmd(a, grid, st);
        } else if (sch.ksl) {
            ce.ksl(a, static_cast<int>(untiled_waves(a.sc.nitems, ctx->untiled_wave_cap)), st);
        } else {
            ce.fock(a, grid, st);
        }
        HIP_TRY(hipGetLastError());
        if (timing) {
            HIP_TRY(hipEventRecord(pend.stop, st));
            ctx->pending.push_back(pend);
        }
    }
    return VLX_OK;
}

// Makes the context stream wait for everything queued so far on the side streams.
int join_side_streams(vlx_context* ctx)
{
    for (int k = 0; k < NSIDE; k++) {
        HIP_TRY(hipEventRecord(ctx->ev_done[k], ctx->side[k]));
        HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->ev_done[k], 0));
    }
    return VLX_OK;
}

// One pass over the integrals for m densities packed at dens (nao x nao each) with factors h_fac
// (fj, fk of density c at h_fac[2 c]): F' of density c into facc + c nao^2.  FP64 with the Schwarz
// selection.  m = 1 runs the single-density kernels; in the context's deterministic mode every class
// runs the direct-digestion kernels (schedules of the untiled variant) with fixed-point accumulation,
// converted to FP64 at the end.
int multi_pass(vlx_context* ctx, vlx_screener* scr, const double* dens, int m, const double* h_fac, double omega2,
               int ithreshold, double* facc)
{
    const size_t nn = (size_t)scr->nao * scr->nao;
    const bool det = ctx->deterministic;
    ctx->h_dfac.assign(h_fac, h_fac + 2 * (size_t)m);
    HIP_TRY(ctx->d_dfac.alloc(2 * (size_t)m));
    HIP_TRY(hipMemcpyAsync(ctx->d_dfac.ptr, ctx->h_dfac.data(), 2 * (size_t)m * sizeof(double), hipMemcpyHostToDevice,
                           ctx->stream));
    double* acc = facc;
    if (det) {
        HIP_TRY(ctx->d_fx.alloc(2 * (size_t)m * nn));
        acc = reinterpret_cast<double*>(ctx->d_fx.ptr);
    }
    const size_t nacc = (det ? 2 : 1) * (size_t)m * nn;
    zero_kernel<<<std::min<size_t>((nacc + 255) / 256, 4096), 256, 0, ctx->stream>>>(acc, nacc);
    HIP_TRY(hipGetLastError());
    std::vector<std::unique_ptr<ClassSchedule>>* list = nullptr;
    const int saved_variant = ctx->force_variant;
    if (det) ctx->force_variant = VARIANT_UNTILED;
    int rc = get_schedules(scr, ithreshold, NO_FP32, &list, true);
    ctx->force_variant = saved_variant;
    if (rc) return rc;
    const bool single = m == 1 && !det;
    const double fj = single ? h_fac[0] : 0.0, fk = single ? h_fac[1] : 0.0;
    const int nd = single ? 1 : m;
    const double* dfac = single ? nullptr : ctx->d_dfac.ptr;
    rc = enqueue_classes(ctx, scr, *list, dens, acc, fj, fk, omega2, ithreshold, nullptr, 1, true, nullptr, nd, dfac,
                         det);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(ctx->ev_start, ctx->stream));
    for (int k = 0; k < NSIDE; k++) HIP_TRY(hipStreamWaitEvent(ctx->side[k], ctx->ev_start, 0));
    rc = enqueue_classes(ctx, scr, *list, dens, acc, fj, fk, omega2, ithreshold, nullptr, 0, true, nullptr, nd, dfac,
                         det);
    const int rj = join_side_streams(ctx);
    if (rc) return rc;
    if (rj) return rj;
    if (det) {
        fx_to_double_kernel<<<std::min<size_t>((m * nn + 255) / 256, 4096), 256, 0, ctx->stream>>>(ctx->d_fx.ptr, facc,
                                                                                                  m * nn);
        HIP_TRY(hipGetLastError());
    }
    return VLX_OK;
}

// scratch = nullptr: F' and the max|D| slot live in context buffers (allocated on first use) and
// missing schedules are built.  Otherwise F' is `scratch` (nao x nao doubles) and max|D| the
// slot `scratch_dmax`, the schedules must exist already and nothing is allocated.
int fock_build(vlx_context* ctx, vlx_screener* scr, const double* d_density, double fj, double fk, double omega2,
               int ithreshold, double* d_fock_out, double anti_coef, double* scratch = nullptr,
               unsigned long long* scratch_dmax = nullptr)
{
    int rc = VLX_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    const bool no_alloc = scratch != nullptr;
    const int n = scr->nao;
    const size_t nn = (size_t)n * n;
    if (!no_alloc) HIP_TRY(ctx->d_faccum.alloc(nn));
    double* facc = no_alloc ? scratch : ctx->d_faccum.ptr;
    dim3 sg((n + 255) / 256, n);
    if (ctx->deterministic) {
        // FP64, Schwarz selection only (the mixed-precision and density-weighted modes do not apply)
        if (no_alloc) return fail(VLX_ERR_ARG, "the no-allocation build has no deterministic mode");
        const double fac[2] = {fj, fk};
        rc = multi_pass(ctx, scr, d_density, 1, fac, omega2, ithreshold, facc);
        if (rc) return rc;
        ctx->last_k32 = NO_FP32;
        if (anti_coef == 0.0) symmetrize_kernel<<<sg, 256, 0, ctx->stream>>>(facc, d_fock_out, n);
        else antisym_add_kernel<<<sg, 256, 0, ctx->stream>>>(facc, d_fock_out, n, anti_coef);
        HIP_TRY(hipGetLastError());
        return VLX_OK;
    }
    zero_kernel<<<std::min<size_t>((nn + 255) / 256, 4096), 256, 0, ctx->stream>>>(facc, nn);
    HIP_TRY(hipGetLastError());
    // mixed precision: FP32 where Schwarz bound x max|D| < tau, i.e. for bin sums
    // >= k32 = ceil(2 + log10(max|D| / tau)) (bins: sqrt(max) < 10^(1 - bin));
    // range-separated builds stay in FP64 (the FP32 kernels are Coulomb-only)
    int k32 = NO_FP32;
    if (ctx->mixed_tau > 0.0 && ithreshold >= 0 && omega2 == 0.0) {
        if (!no_alloc) HIP_TRY(ctx->d_dmax.alloc(1));
        unsigned long long* dmax_slot = no_alloc ? scratch_dmax : ctx->d_dmax.ptr;
        clear_u64_kernel<<<1, 1, 0, ctx->stream>>>(dmax_slot);
        absmax_kernel<<<std::min<size_t>((nn + 255) / 256, 1024), 256, 0, ctx->stream>>>(d_density, nn, dmax_slot);
        HIP_TRY(hipGetLastError());
        unsigned long long bits = 0;
        HIP_TRY(hipMemcpyAsync(&bits, dmax_slot, sizeof(bits), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        double dmax;
        std::memcpy(&dmax, &bits, sizeof(dmax));
        if (dmax > 0.0) {
            const double kk = std::ceil(2.0 + std::log10(dmax / ctx->mixed_tau));
            k32 = static_cast<int>(std::max(0.0, std::min<double>(NO_FP32, kk)));
        }
    }
    ctx->last_k32 = k32;
    // density-weighted screening: bin-sum limits per pair of atom groups from this D
    const int* glim = nullptr;
    if (ctx->dscreen_tau > 0.0 && ithreshold >= 0) {
        if (no_alloc)
            return fail(VLX_ERR_ARG, "density-weighted screening is not available in the no-allocation build");
        const int ng = scr->ngroups;
        const size_t ng2 = (size_t)ng * ng;
        HIP_TRY(ctx->d_gmax.alloc(ng2));
        HIP_TRY(ctx->d_glim.alloc(ng2));
        clear_u64_n_kernel<<<static_cast<int>((ng2 + 255) / 256), 256, 0, ctx->stream>>>(ctx->d_gmax.ptr, (int)ng2);
        group_dmax_kernel<<<std::min<size_t>((nn + 255) / 256, 2048), 256, 0, ctx->stream>>>(
            d_density, n, scr->d_ao_group.ptr, ng, ctx->d_gmax.ptr);
        group_lim_kernel<<<static_cast<int>((ng2 + 255) / 256), 256, 0, ctx->stream>>>(
            ctx->d_gmax.ptr, (int)ng2, ctx->dscreen_tau, ithreshold, ctx->d_glim.ptr);
        HIP_TRY(hipGetLastError());
        glim = ctx->d_glim.ptr;
    }
    std::vector<std::unique_ptr<ClassSchedule>>* list = nullptr;
    rc = get_schedules(scr, ithreshold, k32, &list, !no_alloc);
    if (rc) return rc;
    rc = enqueue_classes(ctx, scr, *list, d_density, facc, fj, fk, omega2, ithreshold, nullptr, 1, !no_alloc, glim);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(ctx->ev_start, ctx->stream));
    for (int k = 0; k < NSIDE; k++) HIP_TRY(hipStreamWaitEvent(ctx->side[k], ctx->ev_start, 0));
    // the side streams join the context stream also when a launch fails part-way, so that
    // work ordered on the context stream (copies, frees) never overtakes a queued kernel
    rc = enqueue_classes(ctx, scr, *list, d_density, facc, fj, fk, omega2, ithreshold, nullptr, 0, !no_alloc, glim);
    const int rj = join_side_streams(ctx);
    if (rc) return rc;
    if (rj) return rj;
    if (anti_coef == 0.0) symmetrize_kernel<<<sg, 256, 0, ctx->stream>>>(facc, d_fock_out, n);
    else antisym_add_kernel<<<sg, 256, 0, ctx->stream>>>(facc, d_fock_out, n, anti_coef);
    HIP_TRY(hipGetLastError());
    return VLX_OK;
}

int parse_build(vlx_context* ctx, vlx_screener* scr, const void* d_density, const void* d_out, const char* label,
                double exchange_factor, double omega, int ithreshold, double* fj, double* fk, double* omega2)
{
    if (!ctx || !scr || !d_density || !d_out) return fail(VLX_ERR_ARG, "NULL argument to the Fock build");
    if (scr->ctx != ctx) return fail(VLX_ERR_ARG, "screener belongs to another context");
    if (ithreshold > 30) return fail(VLX_ERR_ARG, "ithreshold %d outside the accepted range (< 0 or 0..30)", ithreshold);
    bool range_separated = false;
    int rc = label_factors(label, exchange_factor, fj, fk, &range_separated);
    if (rc) return rc;
    if (range_separated && !std::isfinite(omega)) return fail(VLX_ERR_ARG, "omega must be finite");
    // erf(0 r)/r = 0: a range-separated build with omega = 0 is zero, as in the reference
    *omega2 = range_separated ? omega * omega : 0.0;
    if (range_separated && *omega2 == 0.0) *fj = *fk = 0.0;
    return VLX_OK;
}

// ------------------------------------------------------------- batched builds

// Packed densities: molecule blockIdx.y's D into its MolDesc slot.
__global__ void batch_gather_kernel(const double* const* __restrict__ src, const MolDesc* __restrict__ mols,
                                    double* __restrict__ dst)
{
    const MolDesc md = mols[blockIdx.y];
    const double* __restrict__ d = src[blockIdx.y];
    const size_t nn = (size_t)md.nao * md.nao;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nn; i += (size_t)gridDim.x * blockDim.x)
        dst[md.base + i] = d[i];
}

// G_m = F'_m + F'_m^T into the caller's output of molecule blockIdx.y.
__global__ void batch_symmetrize_kernel(const double* __restrict__ f, const MolDesc* __restrict__ mols,
                                        double* const* __restrict__ dst)
{
    const MolDesc md = mols[blockIdx.y];
    double* __restrict__ g = dst[blockIdx.y];
    const int n = md.nao;
    const double* __restrict__ fm = f + md.base;
    const size_t nn = (size_t)n * n;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nn; i += (size_t)gridDim.x * blockDim.x) {
        const size_t p = i / n, q = i % n;
        g[i] = fm[i] + fm[q * n + p];
    }
}

// max |D_m| per molecule (bit patterns of non-negative doubles order like their values)
__global__ void batch_absmax_kernel(const double* __restrict__ d, const MolDesc* __restrict__ mols,
                                    unsigned long long* __restrict__ out)
{
    const MolDesc md = mols[blockIdx.y];
    const size_t nn = (size_t)md.nao * md.nao;
    double m = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nn; i += (size_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs(d[md.base + i]));
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out + blockIdx.y, static_cast<unsigned long long>(__double_as_longlong(m)));
}

// k32_m = ceil(2 + log10(max|D_m| / tau)) clamped to [0, NO_FP32] (NO_FP32 for a zero density),
// the bin-sum threshold of the single-molecule build
__global__ void batch_k32_kernel(const unsigned long long* __restrict__ dmax, int nmol, double tau, int* __restrict__ k32)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    const double v = __longlong_as_double(static_cast<long long>(dmax[m]));
    int k = NO_FP32;
    if (v > 0.0) k = static_cast<int>(fmax(0.0, fmin(static_cast<double>(NO_FP32), ceil(2.0 + log10(v / tau)))));
    k32[m] = k;
}

// Concatenated pair lists, moved items and packed layout of a batch (see BatchPlan).
int build_batch_plan(vlx_context* ctx, BatchPlan& P)
{
    const int nmol = static_cast<int>(P.scrs.size());
    const int npc = npair_classes();
    const bool unscreened = P.ithreshold < 0;
    std::vector<std::vector<std::unique_ptr<ClassSchedule>>*> lists(nmol, nullptr);
    for (int m = 0; m < nmol; m++) {
        const int rc = get_schedules(P.scrs[m], P.ithreshold, NO_FP32, &lists[m]);
        if (rc) return rc;
    }
    // the four orders of every pair class, and each molecule's offsets in them
    enum { KET = 0, BRA = 1, TKET = 2, TBRA = 3, NKIND = 4 };
    auto sets_of = [&](vlx_screener* sc, int kind) -> std::vector<PairSet>& {
        switch (kind) {
        case KET: return unscreened ? sc->all : sc->kept;
        case BRA: return unscreened ? sc->all_bra : sc->kept_bra;
        case TKET: return unscreened ? sc->all_tiled : sc->kept_tiled;
        default: return unscreened ? sc->all_tiled_bra : sc->kept_tiled_bra;
        }
    };
    std::vector<PairSet>* dst_sets[NKIND] = {&P.ket, &P.bra, &P.tket, &P.tbra};
    for (auto* v : dst_sets) {
        v->clear();
        v->resize(npc);
    }
    // poff[kind][m * npc + c]: offset of molecule m's pairs of class c in the concatenated set
    std::vector<int> poff[NKIND], eoff_runs(nmol * npc), eoff_singles(nmol * npc), goff(nmol);
    for (int k = 0; k < NKIND; k++) poff[k].assign(nmol * npc, 0);
    for (int c = 0; c < npc; c++)
        for (int k = 0; k < NKIND; k++) {
            int off = 0;
            for (int m = 0; m < nmol; m++) {
                const std::vector<PairSet>& v = sets_of(P.scrs[m], k);
                poff[k][m * npc + c] = off;
                off += c < static_cast<int>(v.size()) ? v[c].size() : 0;
            }
        }
    int ngroups = 0;
    for (int m = 0; m < nmol; m++) {
        goff[m] = ngroups;
        ngroups += P.scrs[m]->ngroups;
    }
    for (int c = 0; c < npc; c++)
        for (int k = 0; k < NKIND; k++) {
            PairSet& dst = (*dst_sets[k])[c];
            for (int m = 0; m < nmol; m++) {
                const std::vector<PairSet>& v = sets_of(P.scrs[m], k);
                if (c >= static_cast<int>(v.size())) continue;
                const PairSet& src = v[c];
                dst.la = src.la;
                dst.lb = src.lb;
                const int pair0 = static_cast<int>(dst.pairs.size()), prim0 = static_cast<int>(dst.prims.size());
                // ranks are positions in the ket order of the same kind (untiled or tiled)
                const int rank0 = poff[(k == KET || k == BRA) ? KET : TKET][m * npc + c];
                for (const PairInfo& pi0 : src.pairs) {
                    PairInfo pi = pi0;
                    pi.prim_start += prim0;
                    pi.rank += rank0;
                    pi.grp_a += goff[m];
                    pi.grp_b += goff[m];
                    dst.pairs.push_back(pi);
                }
                dst.prims.insert(dst.prims.end(), src.prims.begin(), src.prims.end());
                if (k == TKET) {
                    eoff_runs[m * npc + c] = static_cast<int>(dst.runs.kentries.size());
                    eoff_singles[m * npc + c] = static_cast<int>(dst.singles.kentries.size());
                    for (int2 e : src.runs.kentries) dst.runs.kentries.push_back(make_int2(e.x + pair0, e.y));
                    for (int2 e : src.singles.kentries) dst.singles.kentries.push_back(make_int2(e.x + pair0, e.y));
                }
            }
            finalize_set(dst);
            if (dst.size() > 0) {
                const int rc = upload_set(dst);
                if (rc) return rc;
            }
        }
    // atom groups: (group, l) blocks of every molecule, AO lists local to the molecule
    for (int m = 0; m < nmol; m++) {
        const vlx_screener* sc = P.scrs[m];
        const int ao0 = static_cast<int>(P.grp_ao.size());
        for (size_t i = 0; i < sc->grp_off.size(); i++) {
            P.grp_off.push_back(sc->grp_off[i] + ao0);
            P.grp_n.push_back(sc->grp_n[i]);
        }
        P.grp_ao.insert(P.grp_ao.end(), sc->grp_ao.begin(), sc->grp_ao.end());
    }
    HIP_TRY(P.d_grp_off.upload(P.grp_off));
    HIP_TRY(P.d_grp_n.upload(P.grp_n));
    HIP_TRY(P.d_grp_ao.upload(P.grp_ao));
    // packed D and F'
    for (int m = 0; m < nmol; m++) {
        const int n = P.scrs[m]->nao;
        P.mols.push_back(MolDesc{static_cast<long long>(P.total), n, 0});
        P.total += (size_t)n * n;
        P.max_nao = std::max(P.max_nao, n);
    }
    HIP_TRY(P.d_mols.upload(P.mols));
    HIP_TRY(P.d_D.alloc(P.total));
    HIP_TRY(P.d_F.alloc(P.total));
    P.h_src.assign(nmol, nullptr);
    P.h_dst.assign(nmol, nullptr);
    HIP_TRY(P.d_src.alloc(nmol));
    HIP_TRY(P.d_dst.alloc(nmol));
    if (P.tau > 0.0) {
        HIP_TRY(P.d_dmax.alloc(nmol));
        HIP_TRY(P.d_k32.alloc(nmol));
    }
    // launches: every molecule's class schedules, grouped by (class, variant, lane entries, block)
    std::map<std::array<int, 4>, BatchLaunch*> by_key;
    for (int m = 0; m < nmol; m++) {
        for (const auto& sp : *lists[m]) {
            const ClassSchedule& sch = *sp;
            const auto& ce = g_class_table[sch.entry];
            const std::array<int, 4> key{sch.entry, sch.variant, sch.singles ? 1 : 0, sch.tiled_waves};
            BatchLaunch*& L = by_key[key];
            if (!L) {
                P.launches.push_back(std::make_unique<BatchLaunch>());
                L = P.launches.back().get();
                L->entry = sch.entry;
                L->variant = sch.variant;
                L->tiled = sch.tiled;
                L->singles = sch.singles;
                L->mmd = sch.mmd;
                L->ksl = sch.ksl;
                L->tiled_waves = sch.tiled_waves;
            }
            L->tile_lds = std::max(L->tile_lds, sch.tile_lds);
            L->cost += (double)sch.prim_quartets * (ce.vrr_flops + 60) + (double)sch.quartets * ce.hrr_flops;
            const int x = m * npc + ce.x, y = m * npc + ce.y;
            if (sch.tiled) {
                const int eo = sch.singles ? eoff_singles[y] : eoff_runs[y];
                for (const auto* lst : {&sch.titems, &sch.titems32})
                    for (TileItem ti : *lst) {
                        ti.i0 += poff[TBRA][x];
                        ti.i1 += poff[TBRA][x];
                        ti.j0 += eo;
                        ti.j1 += eo;
                        ti.ga += goff[m];
                        ti.gb += goff[m];
                        ti.gc += goff[m];
                        ti.gd += goff[m];
                        if (ti.gd2 >= 0) ti.gd2 += goff[m];
                        ti.jm += eo;
                        ti.mol = m;
                        L->titems.push_back(ti);
                    }
            } else {
                const bool ks = !sch.mmd && (ce.variant == VLX_KET_STATIONARY || sch.ksl);  // i1 is a bra index
                for (WorkItem wi : sch.items) {
                    wi.j0 += poff[KET][y];
                    wi.i0 += poff[BRA][x];
                    wi.i1 += ks ? poff[BRA][x] : poff[KET][y];
                    wi.jend += poff[KET][y];
                    wi.mol = m;
                    L->items.push_back(wi);
                }
            }
        }
    }
    for (auto& Lp : P.launches) {
        BatchLaunch& L = *Lp;
        const auto& ce = g_class_table[L.entry];
        if (L.tiled) {
            HIP_TRY(L.d_titems.upload(L.titems));
            continue;
        }
        const PairSet& bra = P.bra[ce.x];
        const PairSet& ket = P.ket[ce.y];
        if (L.mmd) {
            int rc = ensure_u(ctx, P.bra[ce.x]);
            if (!rc) rc = ensure_u(ctx, P.ket[ce.y]);
            if (rc) return rc;
        } else if (ctx->balance_items) {
            // the wavefronts' contiguous ranges of equal estimated cost, over the whole batch
            const int nsph = (2 * ce.la + 1) * (2 * ce.lb + 1) * (2 * ce.lc + 1) * (2 * ce.ld + 1);
            const double pcost = ce.vrr_flops + 60.0, qcost = ce.hrr_flops + 12.0 * nsph;
            const bool ks = ce.variant == VLX_KET_STATIONARY || L.ksl;
            std::vector<double> cost;
            cost.reserve(L.items.size());
            for (const auto& wi : L.items) {
                const int j1 = std::min(wi.j0 + 64, wi.jend);
                int kmax = 0;
                for (int j = wi.j0; j < j1; j++) kmax = std::max(kmax, ket.pairs[j].nprim);
                const long long bprims = ks ? bra.nprim_prefix[wi.i1] - bra.nprim_prefix[wi.i0] : bra.pairs[wi.i0].nprim;
                const long long nbra = ks ? wi.i1 - wi.i0 : 1;
                cost.push_back(kmax * bprims * pcost + nbra * qcost);
            }
            balance_items(L.items, cost, untiled_waves(static_cast<long long>(L.items.size()), ctx->untiled_wave_cap));
        }
        HIP_TRY(L.d_items.upload(L.items));
    }
    // side streams by estimated cost, most expensive launches first
    std::sort(P.launches.begin(), P.launches.end(),
              [](const std::unique_ptr<BatchLaunch>& a, const std::unique_ptr<BatchLaunch>& b) { return a->cost > b->cost; });
    std::vector<double> load(std::max(1, ctx->nstreams), 0.0);
    for (auto& Lp : P.launches) {
        const int k = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
        Lp->stream = k;
        load[k] += Lp->cost;
    }
    return VLX_OK;
}

// Enqueues one batched build (see BatchPlan) on the context stream and its side streams.
int record_batch(vlx_context* ctx, BatchPlan& P)
{
    const int nmol = static_cast<int>(P.scrs.size());
    const size_t nn_max = (size_t)P.max_nao * P.max_nao;
    const dim3 gm(static_cast<unsigned>(std::min<size_t>(std::max<size_t>((nn_max + 255) / 256, 1), 256)), nmol);
    hipStream_t st = ctx->stream;
    batch_gather_kernel<<<gm, 256, 0, st>>>(P.d_src.ptr, P.d_mols.ptr, P.d_D.ptr);
    zero_kernel<<<std::min<size_t>((P.total + 255) / 256, 8192), 256, 0, st>>>(P.d_F.ptr, P.total);
    HIP_TRY(hipGetLastError());
    const bool mixed = P.tau > 0.0;
    if (mixed) {
        clear_u64_n_kernel<<<(nmol + 255) / 256, 256, 0, st>>>(P.d_dmax.ptr, nmol);
        batch_absmax_kernel<<<gm, 256, 0, st>>>(P.d_D.ptr, P.d_mols.ptr, P.d_dmax.ptr);
        batch_k32_kernel<<<(nmol + 255) / 256, 256, 0, st>>>(P.d_dmax.ptr, nmol, P.tau, P.d_k32.ptr);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(ctx->ev_start, st));
    for (int k = 0; k < NSIDE; k++) HIP_TRY(hipStreamWaitEvent(ctx->side[k], ctx->ev_start, 0));
    int rc = VLX_OK;
    for (auto& Lp : P.launches) {
        const BatchLaunch& L = *Lp;
        const auto& ce = g_class_table[L.entry];
        const PairSet& bra = L.tiled ? P.tbra[ce.x] : P.bra[ce.x];
        const PairSet& ket = L.tiled ? P.tket[ce.y] : P.ket[ce.y];
        FockArgs a{};
        a.bra = bra.d_pairs.ptr;
        a.bra_prims = bra.d_prims.ptr;
        a.ket = ket.d_pairs.ptr;
        a.ket_prims = ket.d_prims.ptr;
        a.bra_u = bra.d_u.ptr;
        a.ket_u = ket.d_u.ptr;
        a.items = L.d_items.ptr;
        a.titems = L.d_titems.ptr;
        a.kentries = L.singles ? ket.singles.d_kentries.ptr : ket.runs.d_kentries.ptr;
        a.grp_off = P.d_grp_off.ptr;
        a.grp_n = P.d_grp_n.ptr;
        a.grp_ao = P.d_grp_ao.ptr;
        a.tile_lds_bytes = L.tile_lds;
        a.tiled_waves = L.tiled_waves;
        a.boys_tab = ctx->d_boys.ptr + (size_t)(ce.la + ce.lb + ce.lc + ce.ld) * BOYS_TAB_SIZE;
        a.boys_tab_f = ctx->d_boys_f.ptr + (size_t)(ce.la + ce.lb + ce.lc + ce.ld) * BOYS_TAB_SIZE;
        a.D = P.d_D.ptr;
        a.F = P.d_F.ptr;
        a.mols = P.d_mols.ptr;
        a.k32 = nullptr;
        a.glim = nullptr;  // density-weighted screening: single-molecule builds only
        a.sc.ngroups = 0;
        a.sc.nitems = static_cast<long long>(L.tiled ? L.titems.size() : L.items.size());
        a.sc.nket = ket.size();
        a.sc.ithreshold = P.ithreshold;
        a.sc.same_class = (ce.x == ce.y) ? 1 : 0;
        a.sc.nao = 0;
        a.sc.fj = P.fj;
        a.sc.fk = P.fk;
        a.sc.omega2 = P.omega2;
        if (a.sc.nitems == 0) continue;
        hipStream_t s = ctx->side[L.stream % NSIDE];
        if (L.tiled) {
            const int grid = static_cast<int>(std::min<long long>(a.sc.nitems, ctx->tiled_grid));
            if (mixed && ce.tiled32) {
                // both precisions over every item: each kernel runs its side of the device split
                a.k32 = P.d_k32.ptr;
                ce.tiled(a, grid, s);
                ce.tiled32(a, grid, s);
            } else {
                ce.tiled(a, grid, s);
            }
        } else if (L.mmd) {
            ce.mmd(a, static_cast<int>(std::min<long long>(a.sc.nitems, ctx->mmd_grid)), s);
        } else if (L.ksl) {
            ce.ksl(a, static_cast<int>(untiled_waves(a.sc.nitems, ctx->untiled_wave_cap)), s);
        } else {
            ce.fock(a, static_cast<int>(untiled_waves(a.sc.nitems, ctx->untiled_wave_cap) / VLX_WAVES_PER_BLOCK), s);
        }
        if (hipGetLastError() != hipSuccess) {
            rc = fail(VLX_ERR_HIP, "batched class launch failed");
            break;
        }
    }
    const int rj = join_side_streams(ctx);
    if (rc) return rc;
    if (rj) return rj;
    batch_symmetrize_kernel<<<gm, 256, 0, st>>>(P.d_F.ptr, P.d_mols.ptr, P.d_dst.ptr);
    HIP_TRY(hipGetLastError());
    return VLX_OK;
}

// Captures the plan's build into a HIP graph (no allocation, copy or host wait inside).
int capture_batch(vlx_context* ctx, BatchPlan& P)
{
    HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = record_batch(ctx, P);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(VLX_ERR_HIP, "graph capture of the batched build: %s", hipGetErrorString(e));
    P.graph = g;
    HIP_TRY(hipGraphInstantiate(&P.exec, g, nullptr, nullptr, 0));
    return VLX_OK;
}

}  // namespace

int vlx_fock_compute_batch_device(vlx_context* ctx, int nmol, vlx_screener* const* scrs, const double* const* d_densities,
                                  const char* label, double exchange_factor, double omega, int ithreshold,
                                  double* const* d_focks)
{
    if (!ctx || nmol < 0 || (nmol > 0 && (!scrs || !d_densities || !d_focks)))
        return fail(VLX_ERR_ARG, "NULL argument to vlx_fock_compute_batch_device");
    if (nmol == 0) return VLX_OK;
    if (nmol > 65535) return fail(VLX_ERR_ARG, "%d molecules in one batch (at most 65535)", nmol);
    if (ctx->deterministic) return fail(VLX_ERR_ARG, "batched builds have no deterministic mode");
    double fj = 0.0, fk = 0.0, omega2 = 0.0;
    for (int m = 0; m < nmol; m++) {
        const int rc = parse_build(ctx, scrs[m], d_densities[m], d_focks[m], label, exchange_factor, omega, ithreshold,
                                   &fj, &fk, &omega2);
        if (rc) return rc;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    // mixed precision as in single builds (FP32 recursions below tau, Coulomb builds with Schwarz
    // selection only), the split decided on the device per molecule
    const double tau = (ctx->mixed_tau > 0.0 && ithreshold >= 0 && omega2 == 0.0) ? ctx->mixed_tau : 0.0;
    BatchPlan* plan = nullptr;
    for (auto it = ctx->plans.begin(); it != ctx->plans.end(); ++it) {
        BatchPlan& P = **it;
        if (static_cast<int>(P.scrs.size()) != nmol || P.ithreshold != ithreshold || P.tau != tau || P.fj != fj ||
            P.fk != fk || P.omega2 != omega2 || P.force_variant != ctx->force_variant || P.nstreams != ctx->nstreams)
            continue;
        bool same = true;
        for (int m = 0; m < nmol && same; m++) same = P.scrs[m] == scrs[m] && P.epochs[m] == scrs[m]->epoch;
        if (!same) continue;
        // most recently used last
        std::unique_ptr<BatchPlan> keep = std::move(*it);
        ctx->plans.erase(it);
        ctx->plans.push_back(std::move(keep));
        plan = ctx->plans.back().get();
        break;
    }
    if (!plan) {
        auto P = std::make_unique<BatchPlan>();
        P->scrs.assign(scrs, scrs + nmol);
        for (int m = 0; m < nmol; m++) P->epochs.push_back(scrs[m]->epoch);
        P->ithreshold = ithreshold;
        P->force_variant = ctx->force_variant;
        P->nstreams = ctx->nstreams;
        P->fj = fj;
        P->fk = fk;
        P->omega2 = omega2;
        P->tau = tau;
        int rc = build_batch_plan(ctx, *P);
        if (!rc) rc = capture_batch(ctx, *P);
        if (rc) return rc;
        constexpr size_t MAX_PLANS = 8;
        if (ctx->plans.size() >= MAX_PLANS) {
            HIP_TRY(hipStreamSynchronize(ctx->stream));  // the oldest plan's graph may still run
            ctx->plans.erase(ctx->plans.begin());
        }
        ctx->plans.push_back(std::move(P));
        plan = ctx->plans.back().get();
    }
    // the callers' buffers: pointer tables in HBM, read by the graph's gather and scatter kernels
    bool changed = false;
    for (int m = 0; m < nmol; m++) changed = changed || plan->h_src[m] != d_densities[m] || plan->h_dst[m] != d_focks[m];
    if (changed) {
        plan->h_src.assign(d_densities, d_densities + nmol);
        plan->h_dst.assign(d_focks, d_focks + nmol);
        HIP_TRY(hipMemcpyAsync(plan->d_src.ptr, plan->h_src.data(), nmol * sizeof(const double*), hipMemcpyHostToDevice,
                               ctx->stream));
        HIP_TRY(hipMemcpyAsync(plan->d_dst.ptr, plan->h_dst.data(), nmol * sizeof(double*), hipMemcpyHostToDevice,
                               ctx->stream));
    }
    HIP_TRY(hipGraphLaunch(plan->exec, ctx->stream));
    return VLX_OK;
}

int vlx_fock_compute_device(vlx_context* ctx, vlx_screener* scr, const double* d_density, const char* label,
                            double exchange_factor, double omega, int ithreshold, double* d_fock_out)
{
    double fj = 0.0, fk = 0.0, omega2 = 0.0;
    int rc = parse_build(ctx, scr, d_density, d_fock_out, label, exchange_factor, omega, ithreshold, &fj, &fk, &omega2);
    if (rc) return rc;
    return fock_build(ctx, scr, d_density, fj, fk, omega2, ithreshold, d_fock_out, 0.0);
}

int vlx_fock_compute_device_noalloc(vlx_context* ctx, vlx_screener* scr, const double* d_density, const char* label,
                                    double exchange_factor, double omega, int ithreshold, double* d_fock_out)
{
    double fj = 0.0, fk = 0.0, omega2 = 0.0;
    int rc = parse_build(ctx, scr, d_density, d_fock_out, label, exchange_factor, omega, ithreshold, &fj, &fk, &omega2);
    if (rc) return rc;
    if (scr->nao > VLX_NOALLOC_MAX_NAO)
        return fail(VLX_ERR_ARG, "vlx_fock_compute_device_noalloc: %d AOs exceed its scratch limit of %d", scr->nao,
                    VLX_NOALLOC_MAX_NAO);
    double* base = ctx->device < MAX_DEVICES ? g_scratch_ptr[ctx->device] : nullptr;
    if (!base) return fail(VLX_ERR_HIP, "no scratch address for device %d", ctx->device);
    return fock_build(ctx, scr, d_density, fj, fk, omega2, ithreshold, d_fock_out, 0.0, base,
                      reinterpret_cast<unsigned long long*>(base + SCRATCH_FACC));
}

int vlx_fock_compute_general_device(vlx_context* ctx, vlx_screener* scr, const double* d_density, const char* label,
                                    double exchange_factor, double omega, int ithreshold, double* d_fock_out)
{
    double fj = 0.0, fk = 0.0, omega2 = 0.0;
    int rc = parse_build(ctx, scr, d_density, d_fock_out, label, exchange_factor, omega, ithreshold, &fj, &fk, &omega2);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    const int n = scr->nao;
    const size_t nn = (size_t)n * n;
    HIP_TRY(ctx->d_sym.alloc(nn));
    HIP_TRY(ctx->d_anti.alloc(nn));
    dim3 sg((n + 255) / 256, n);
    split_kernel<<<sg, 256, 0, ctx->stream>>>(d_density, ctx->d_sym.ptr, ctx->d_anti.ptr, n);
    HIP_TRY(hipGetLastError());
    // J(D) = J(S); K(D) = K(S) + K(A) with K(A) = F'_A - F'_A^T
    rc = fock_build(ctx, scr, ctx->d_sym.ptr, fj, fk, omega2, ithreshold, d_fock_out, 0.0);
    if (rc || fk == 0.0) return rc;
    return fock_build(ctx, scr, ctx->d_anti.ptr, 0.0, 1.0, omega2, ithreshold, d_fock_out, fk);
}

int vlx_fock_compute(vlx_context* ctx, vlx_screener* scr, const double* density, const char* label,
                     double exchange_factor, double omega, int ithreshold, double* fock_out)
{
    if (!ctx || !scr || !density || !fock_out) return fail(VLX_ERR_ARG, "NULL argument to vlx_fock_compute");
    const int n = scr->nao;
    int rc = check_symmetric_host(density, n);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t nn = (size_t)n * n;
    HIP_TRY(ctx->d_density.alloc(nn));
    HIP_TRY(ctx->d_fock.alloc(nn));
    HIP_TRY(hipMemcpyAsync(ctx->d_density.ptr, density, nn * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    rc = vlx_fock_compute_device(ctx, scr, ctx->d_density.ptr, label, exchange_factor, omega, ithreshold,
                                 ctx->d_fock.ptr);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(fock_out, ctx->d_fock.ptr, nn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return VLX_OK;
}

int vlx_fock_compute_general(vlx_context* ctx, vlx_screener* scr, const double* density, const char* label,
                             double exchange_factor, double omega, int ithreshold, double* fock_out)
{
    if (!ctx || !scr || !density || !fock_out) return fail(VLX_ERR_ARG, "NULL argument to vlx_fock_compute_general");
    const int n = scr->nao;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t nn = (size_t)n * n;
    HIP_TRY(ctx->d_density.alloc(nn));
    HIP_TRY(ctx->d_fock.alloc(nn));
    HIP_TRY(hipMemcpyAsync(ctx->d_density.ptr, density, nn * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    int rc = vlx_fock_compute_general_device(ctx, scr, ctx->d_density.ptr, label, exchange_factor, omega, ithreshold,
                                             ctx->d_fock.ptr);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(fock_out, ctx->d_fock.ptr, nn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return VLX_OK;
}

// Multi-density build (src/t4c_func/FockDriver.cpp:173-229, CT4CMatricesDistributor): one pass over
// the integrals digests every density with its own label; the integrals are erf(omega r)/r only
// when every label is range-separated (CT4CMatricesDistributor::need_omega), else 1/r for all, the
// *_rs labels then digesting like their Coulomb forms.  A general density contributes two
// digested densities: its symmetric part with the label's (fj, fk) and its antisymmetric part with
// (0, fk), G = F'_S + F'_S^T + F'_A - F'_A^T.  FP64 with the Schwarz selection (the context's
// mixed-precision and density-weighted modes do not apply).
int vlx_fock_compute_multi_device(vlx_context* ctx, vlx_screener* scr, int nmat, const double* d_densities,
                                  const int* general, const char* const* labels, double exchange_factor, double omega,
                                  int ithreshold, double* d_focks_out)
{
    if (!ctx || !scr || !d_densities || !labels || !d_focks_out || nmat < 1)
        return fail(VLX_ERR_ARG, "bad argument to vlx_fock_compute_multi_device");
    if (scr->ctx != ctx) return fail(VLX_ERR_ARG, "screener belongs to another context");
    if (ithreshold > 30) return fail(VLX_ERR_ARG, "ithreshold %d outside the accepted range (< 0 or 0..30)", ithreshold);
    std::vector<double> fj(nmat), fk(nmat);
    bool all_rs = true;
    for (int i = 0; i < nmat; i++) {
        bool rs = false;
        int rc = label_factors(labels[i], exchange_factor, &fj[i], &fk[i], &rs);
        if (rc) return rc;
        all_rs = all_rs && rs;
    }
    if (all_rs && !std::isfinite(omega)) return fail(VLX_ERR_ARG, "omega must be finite");
    const double omega2 = all_rs ? omega * omega : 0.0;
    HIP_TRY(hipSetDevice(ctx->device));
    const int n = scr->nao;
    const size_t nn = (size_t)n * n;
    dim3 sg((n + 255) / 256, n);
    // erf(0 r)/r = 0: every output is zero
    if (all_rs && omega2 == 0.0) {
        zero_kernel<<<std::min<size_t>((nmat * nn + 255) / 256, 4096), 256, 0, ctx->stream>>>(d_focks_out, nmat * nn);
        HIP_TRY(hipGetLastError());
        return VLX_OK;
    }
    // digested densities: c_sym[i] (and c_anti[i] for a general density with exchange)
    std::vector<int> c_sym(nmat), c_anti(nmat, -1);
    bool any_general = false;
    int m = 0;
    ctx->h_dfac.clear();
    for (int i = 0; i < nmat; i++) {
        c_sym[i] = m++;
        ctx->h_dfac.push_back(fj[i]);
        ctx->h_dfac.push_back(fk[i]);
        if (general && general[i] && fk[i] != 0.0) {
            c_anti[i] = m++;
            ctx->h_dfac.push_back(0.0);
            ctx->h_dfac.push_back(fk[i]);
        }
        any_general = any_general || (general && general[i]);
    }
    const double* dens = d_densities;
    if (any_general) {
        HIP_TRY(ctx->d_mdens.alloc((size_t)m * nn));
        HIP_TRY(ctx->d_sym.alloc(nn));
        for (int i = 0; i < nmat; i++) {
            const double* di = d_densities + (size_t)i * nn;
            double* ds = ctx->d_mdens.ptr + (size_t)c_sym[i] * nn;
            if (general && general[i]) {
                double* da = c_anti[i] >= 0 ? ctx->d_mdens.ptr + (size_t)c_anti[i] * nn : ctx->d_sym.ptr;
                split_kernel<<<sg, 256, 0, ctx->stream>>>(di, ds, da, n);
                HIP_TRY(hipGetLastError());
            } else {
                HIP_TRY(hipMemcpyAsync(ds, di, nn * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
            }
        }
        dens = ctx->d_mdens.ptr;
    }
    HIP_TRY(ctx->d_mfacc.alloc((size_t)m * nn));
    double* facc = ctx->d_mfacc.ptr;
    const std::vector<double> fac = ctx->h_dfac;
    int rc = multi_pass(ctx, scr, dens, m, fac.data(), omega2, ithreshold, facc);
    if (rc) return rc;
    for (int i = 0; i < nmat; i++) {
        double* g = d_focks_out + (size_t)i * nn;
        symmetrize_kernel<<<sg, 256, 0, ctx->stream>>>(facc + (size_t)c_sym[i] * nn, g, n);
        if (c_anti[i] >= 0) antisym_add_kernel<<<sg, 256, 0, ctx->stream>>>(facc + (size_t)c_anti[i] * nn, g, n, 1.0);
        HIP_TRY(hipGetLastError());
    }
    return VLX_OK;
}

int vlx_fock_compute_multi(vlx_context* ctx, vlx_screener* scr, int nmat, const double* densities, const int* general,
                           const char* const* labels, double exchange_factor, double omega, int ithreshold,
                           double* focks_out)
{
    if (!ctx || !scr || !densities || !labels || !focks_out || nmat < 1)
        return fail(VLX_ERR_ARG, "bad argument to vlx_fock_compute_multi");
    const int n = scr->nao;
    const size_t nn = (size_t)n * n;
    for (int i = 0; i < nmat; i++) {
        if (general && general[i]) continue;
        int rc = check_symmetric_host(densities + i * nn, n);
        if (rc) return rc;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(ctx->d_density.alloc(nmat * nn));
    HIP_TRY(ctx->d_fock.alloc(nmat * nn));
    HIP_TRY(hipMemcpyAsync(ctx->d_density.ptr, densities, nmat * nn * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    int rc = vlx_fock_compute_multi_device(ctx, scr, nmat, ctx->d_density.ptr, general, labels, exchange_factor, omega,
                                           ithreshold, ctx->d_fock.ptr);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(focks_out, ctx->d_fock.ptr, nmat * nn * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return VLX_OK;
}

int vlx_eri_quartets(vlx_context* ctx, vlx_screener* scr, int nquartets, const int* quartets, double* out)
{
    if (!ctx || !scr || !quartets || !out || nquartets < 0) return fail(VLX_ERR_ARG, "bad argument to vlx_eri_quartets");
    HIP_TRY(hipSetDevice(ctx->device));
    const int ns = scr->nshells;
    struct Req {
        int q;
        int bra_x, bra_i, ket_x, ket_i;
        int perm[4];  // requested position -> canonical position (0 A, 1 B, 2 C, 3 D)
        size_t offset;
    };
    std::vector<Req> reqs(nquartets);
    size_t off = 0;
    for (int q = 0; q < nquartets; q++) {
        const int* sh = &quartets[4 * q];
        for (int k = 0; k < 4; k++)
            if (sh[k] < 0 || sh[k] >= ns) return fail(VLX_ERR_ARG, "quartet %d: shell index out of range", q);
        auto r1 = scr->pair_ref[(size_t)sh[0] * ns + sh[1]];
        auto r2 = scr->pair_ref[(size_t)sh[2] * ns + sh[3]];
        Req r{};
        r.q = q;
        r.offset = off;
        int pos[4];  // canonical position of requested shells a, b, c, d
        pos[0] = r1.swapped ? 1 : 0;
        pos[1] = r1.swapped ? 0 : 1;
        pos[2] = r2.swapped ? 3 : 2;
        pos[3] = r2.swapped ? 2 : 3;
        if (r1.cls >= r2.cls) {
            r.bra_x = r1.cls; r.bra_i = r1.idx; r.ket_x = r2.cls; r.ket_i = r2.idx;
        } else {
            r.bra_x = r2.cls; r.bra_i = r2.idx; r.ket_x = r1.cls; r.ket_i = r1.idx;
            for (int k = 0; k < 4; k++) pos[k] = (pos[k] + 2) % 4;
        }
        for (int k = 0; k < 4; k++) r.perm[k] = pos[k];
        reqs[q] = r;
        size_t sz = 1;
        for (int k = 0; k < 4; k++) sz *= 2 * scr->l[sh[k]] + 1;
        off += sz;
    }
    const int nentries = static_cast<int>(sizeof(g_class_table) / sizeof(g_class_table[0]));
    for (int e = 0; e < nentries; e++) {
        const auto& ce = g_class_table[e];
        std::vector<int> sel;
        std::vector<int2> list;
        for (const auto& r : reqs)
            if (r.bra_x == ce.x && r.ket_x == ce.y) {
                sel.push_back(r.q);
                list.push_back(make_int2(r.bra_i, r.ket_i));
            }
        if (list.empty()) continue;
        const int nsph = (2 * ce.la + 1) * (2 * ce.lb + 1) * (2 * ce.lc + 1) * (2 * ce.ld + 1);
        DevBuf<int2> d_list;
        DevBuf<double> d_out;
        HIP_TRY(d_list.upload(list));
        HIP_TRY(d_out.alloc(list.size() * nsph));
        const PairSet& bra = scr->all[ce.x];
        const PairSet& ket = scr->all[ce.y];
        EriArgs a{bra.d_pairs.ptr, bra.d_prims.ptr, ket.d_pairs.ptr, ket.d_prims.ptr, d_list.ptr,
                  static_cast<int>(list.size()),
                  ctx->d_boys.ptr + (size_t)(ce.la + ce.lb + ce.lc + ce.ld) * BOYS_TAB_SIZE, d_out.ptr};
        ce.eri(a, ctx->stream);
        HIP_TRY(hipGetLastError());
        std::vector<double> h(list.size() * nsph);
        HIP_TRY(hipMemcpyAsync(h.data(), d_out.ptr, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        const int cdim[4] = {2 * ce.la + 1, 2 * ce.lb + 1, 2 * ce.lc + 1, 2 * ce.ld + 1};
        for (size_t k = 0; k < sel.size(); k++) {
            const Req& r = reqs[sel[k]];
            int rdim[4];
            for (int t = 0; t < 4; t++) rdim[t] = cdim[r.perm[t]];
            const double* src = &h[k * nsph];
            double* dst = out + r.offset;
            int idx[4];
            for (idx[0] = 0; idx[0] < rdim[0]; idx[0]++)
                for (idx[1] = 0; idx[1] < rdim[1]; idx[1]++)
                    for (idx[2] = 0; idx[2] < rdim[2]; idx[2]++)
                        for (idx[3] = 0; idx[3] < rdim[3]; idx[3]++) {
                            int c[4];
                            for (int t = 0; t < 4; t++) c[r.perm[t]] = idx[t];
                            dst[((idx[0] * rdim[1] + idx[1]) * rdim[2] + idx[2]) * rdim[3] + idx[3]] =
                                src[((c[0] * cdim[1] + c[1]) * cdim[2] + c[2]) * cdim[3] + c[3]];
                        }
        }
    }
    return VLX_OK;
}

}  // extern "C"
