"""Benchmark of the two-electron J+K Fock build on one MI355X.

Headline workload (one "step" = one build): seeded 64-water cluster in a
12.8 Angstrom box (seed 64), def2-TZVP (2752 AOs, f shells on O), FP64
G = 2J - K ('2jk', exchange factor 1, omega 0, ithreshold 12) on the SAD
guess density, after a per-class kernel-variant autotuner pass.  D and G stay
resident in HBM; the timed region is K builds between stream
synchronisations.  value = FP64 builds/s.

Printed: one JSON line with the roofline of the dominant class kernel (FP64
vector-ALU bound; per-launch algorithmic FLOPs of the reference scheme over
the HIP-event kernel time measured inside the timed region), the mixed
precision sweep on the same density, and a CPU baseline (the reference's own
CFockDriver::compute, built from its sources into oracle/_ref, on the host
cores for one full build of the same workload).  Extra keys: the 32-water
def2-SVP build and its RHF SCF, and the batched many-molecule workload.

Algorithmic FLOPs per class (reference scheme): per screened contracted
quartet nprim_bra nprim_ket (V + 53 + 4 L) + C + 12 n_sph, with V and C of
the reference's recursions (tests/golden/ref_scheme_flops.json).
"""

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FP64_VALU_PEAK_TFLOPS = 78.6
# HBM bytes per launch per kernel and workload, from FETCH_SIZE/WRITE_SIZE passes scaled by a copy
# calibration (profiles/*_traffic.json)
TRAFFIC_FILES = {'water64_tzvp': os.path.join(ROOT, 'profiles', 'fock_build_water64_tzvp_traffic.json'),
                 'water32_svp': os.path.join(ROOT, 'profiles', 'fock_build_water32_svp_traffic.json')}
REF_FLOPS_FILE = os.path.join(ROOT, 'tests', 'golden', 'ref_scheme_flops.json')


def log(msg):
    """Progress line on stderr (the JSON line alone goes to stdout)."""
    print(f'bench: {msg}', file=sys.stderr, flush=True)


def kernel_traffic(workload, name):
    try:
        with open(TRAFFIC_FILES[workload]) as fh:
            return json.load(fh)['bytes_per_launch'].get(name)
    except (OSError, ValueError, KeyError):
        return None


def ref_scheme_key(cls):
    """Class (la lb|lc ld) of this build (la >= lb, lc >= ld) in the reference's orientation: pairs
    (l_first <= l_second), the lexicographically lower pair on the bra side, as its triangular
    loops over angular-momentum-ordered pair blocks run them."""
    la, lb, lc, ld = cls
    p1, p2 = (lb, la), (ld, lc)
    lo, hi = min(p1, p2), max(p1, p2)
    return '%d%d%d%d' % (lo + hi)


def ref_scheme_flops(work):
    """{class: FLOPs of one 2jk build} from the per-class quartet counts of Screener.work() and the
    reference scheme's V (per primitive quartet) and C (per contracted quartet):
    prim_quartets (V + 53 + 4 L) + quartets (C + 12 n_sph)."""
    with open(REF_FLOPS_FILE) as fh:
        table = json.load(fh)['classes']
    out = {}
    for cls, nq, npq, _ in work:
        vc = table[ref_scheme_key(cls)]
        L = sum(cls)
        nsph = np.prod([2 * l + 1 for l in cls])
        out[cls] = npq * (vc['V'] + 53 + 4 * L) + nq * (vc['C'] + 12 * nsph)
    return out


def roofline(scr, ithr, times, ms_per_step, workload):
    """Dominant class kernel (largest summed HIP-event time in the timed region) with its
    reference-scheme FLOPs per launch over its average launch time, and the whole build's rate."""
    from veloxchem_amd import _native
    flops = ref_scheme_flops(scr.work(ithr))
    dom = max(times, key=lambda c: times[c][0])
    dom_ms = times[dom][0] / times[dom][1]
    achieved = flops[dom] / (dom_ms * 1e-3) / 1e12
    kernel = _native.class_kernel(scr, ithr, dom)
    build = sum(flops.values()) / (ms_per_step * 1e-3) / 1e12
    return {'bound': 'valu', 'kernel': kernel, 'class': '%d%d%d%d' % dom, 'achieved': achieved,
            'peak': FP64_VALU_PEAK_TFLOPS, 'unit': 'TFLOP/s', 'frac': achieved / FP64_VALU_PEAK_TFLOPS,
            'traffic': kernel_traffic(workload, kernel), 'kernel_ms': dom_ms,
            'flops_per_launch': flops[dom], 'kernel_share_of_build': times[dom][0] / sum(v[0] for v in times.values()),
            'build_achieved': build, 'build_frac': build / FP64_VALU_PEAK_TFLOPS,
            'build_flops': sum(flops.values())}


def build_workload(nmol, box, seed, basis):
    from veloxchem_amd import MolecularBasis, T4CScreener
    from veloxchem_amd.sadguess import sad_density
    from veloxchem_amd.workloads import water_cluster
    mol = water_cluster(nmol, box, seed)
    bas = MolecularBasis.read(mol, basis)
    t0 = time.perf_counter()
    scr = T4CScreener()
    scr.partition(bas, mol, 'eri')
    t_screen = time.perf_counter() - t0
    dm = sad_density(mol, MolecularBasis.read(mol, 'ao-start-guess'), bas)
    return mol, bas, scr, dm, t_screen


def run_scf(mol, bas, conv_thresh, precision='fp64', tau=0.0, incremental=False, density_screening=0.0):
    from veloxchem_amd.scfdriver import ScfRestrictedDriver
    drv = ScfRestrictedDriver()
    drv.conv_thresh = conv_thresh
    drv.precision = precision
    drv.mixed_threshold = tau
    drv.incremental_fock = incremental
    drv.density_screening = density_screening
    res = drv.compute(mol, bas)
    tm = res['timings']
    out = {'scf_time_s': round(tm['total'], 3), 'iterations': res['num_iter'], 'converged': res['is_converged'],
           'energy': res['scf_energy'], 'conv_thresh': conv_thresh, 'fock_time_s': round(tm['fock'], 3),
           'one_electron_s': round(tm['one_electron'], 3), 'linear_algebra_s': round(tm['linear_algebra'], 3),
           'screener_s': round(tm['screener'], 3), 'guess_s': round(tm['guess'], 3), 'precision': precision,
           'incremental_fock': incremental}
    if incremental:
        out['incremental_builds'] = sum(1 for h in drv.history if h['incremental'])
    if precision == 'mixed':
        out['mixed_threshold'] = tau
    if density_screening > 0.0:
        out['density_screening'] = density_screening
    # per-SCF average rate of its Fock builds (every build of the run, on the SCF's own densities)
    out['fock_builds'] = len(drv.history)
    out['fock_builds_per_s'] = len(drv.history) / max(tm['fock'], 1e-12)
    return out


def fast_vs_fp64_scf(mol, bas, conv_thresh, tau_mixed, tau_dens):
    """The plain FP64 SCF (full builds) against the fast one (mixed precision, incremental
    difference-density builds, density-weighted screening; confirmed by an unscreened full build):
    per-SCF-average Fock builds/s, Fock time, iterations and the energy difference."""
    ref = run_scf(mol, bas, conv_thresh)
    fast = run_scf(mol, bas, conv_thresh, 'mixed', tau_mixed, incremental=True, density_screening=tau_dens)
    screened = run_scf(mol, bas, conv_thresh, incremental=True, density_screening=tau_dens)
    for r in (fast, screened):
        r['dE_vs_fp64'] = r['energy'] - ref['energy']
        r['builds_per_s_vs_fp64'] = r['fock_builds_per_s'] / ref['fock_builds_per_s']
        r['fock_time_vs_fp64'] = ref['fock_time_s'] / max(r['fock_time_s'], 1e-12)
    return {'fp64': ref, 'fast': fast, 'fp64_incremental_density_screened': screened}


def measure_builds(scr, dm, ithr, steps, warmup, taus, label=b'2jk', precision='fp64', tau=0.0):
    """Timed builds on device-resident D (HIP events per class kernel), then the mixed sweep on the
    same density: rate and error of each threshold against the FP64 build."""
    from veloxchem_amd import _native
    lib = _native.lib()
    ctx = _native.context(0)
    n = dm.shape[0]
    d_dm = _native.DeviceArray((n, n))
    d_g = _native.DeviceArray((n, n))
    d_dm.upload(dm)

    def one_build():
        _native.check(lib.vlx_fock_compute_device(ctx, scr.handle, d_dm.ptr, label, 1.0, 0.0, ithr, d_g.ptr))

    def timed(k, w=warmup):
        for _ in range(max(1, w)):
            one_build()
        _native.check(lib.vlx_context_synchronize(ctx))
        t0 = time.perf_counter()
        for _ in range(k):
            one_build()
        _native.check(lib.vlx_context_synchronize(ctx))
        return (time.perf_counter() - t0) * 1e3 / k

    try:
        _native.set_precision(precision, tau)
        _native.check(lib.vlx_context_set_timing(ctx, 1))
        _native.check(lib.vlx_context_reset_times(ctx))
        ms = timed(steps)
        _native.check(lib.vlx_context_set_timing(ctx, 0))
        times = _native.class_times(0)
        g = d_g.download()
        _native.set_precision('fp64')
        if precision == 'fp64':
            g64, ms64 = g, ms
        else:  # the FP64 reference of the sweep
            ms64 = timed(steps)
            g64 = d_g.download()
        sweep = []
        sweep_steps = min(steps, 3)  # the sweep's rates: three builds per threshold after one warm-up
        for t in taus:
            _native.set_precision('mixed', t)
            ms_m = timed(sweep_steps, 1)
            gm = d_g.download()
            k32 = _native.last_fp32_bin_sum()
            p64, p32 = scr.precision_split(ithr, k32)
            entry = {'threshold': t, 'builds_per_s': 1e3 / ms_m, 'ms_per_step': ms_m, 'fp32_bin_sum': k32,
                     'fp32_prim_fraction': p32 / max(1, p64 + p32), 'max_abs_dG': float(np.max(np.abs(gm - g64))),
                     'dE2_vs_fp64': float(np.sum(dm * (gm - g64)))}
            entry['speedup_vs_fp64'] = ms64 / ms_m
            entry['steps'] = sweep_steps
            sweep.append(entry)
    finally:
        _native.set_precision('fp64')
        d_dm.free()
        d_g.free()
    return ms, times, g, sweep


def multi_density(scr, dm, ithr, counts=(2, 4, 8), steps=3, seed=7):
    """One pass over the integrals for k densities (vlx_fock_compute_multi_device) against k single
    builds of the same densities, all FP64 on device-resident D: times, the cost of the pass in
    units of one build, and max|dG| of every density against its own single build.  The densities
    are the workload's D plus seeded symmetric perturbations of 1 % of its largest element."""
    from veloxchem_amd import _native
    lib, ctx = _native.lib(), _native.context(0)
    n = dm.shape[0]
    kmax = max(counts)
    rng = np.random.default_rng(seed)
    dens = np.empty((kmax, n, n))
    for i in range(kmax):
        p = rng.standard_normal((n, n))
        dens[i] = dm + 0.01 * np.max(np.abs(dm)) * 0.5 * (p + p.T) / np.sqrt(n)
    d_d = _native.DeviceArray((kmax, n, n))
    d_g = _native.DeviceArray((kmax, n, n))
    d_one = _native.DeviceArray((n, n))
    d_d.upload(dens)
    nn8 = n * n * 8

    def single(i, out):
        _native.check(lib.vlx_fock_compute_device(ctx, scr.handle, ctypes.c_void_p(d_d.ptr.value + i * nn8), b'2jk',
                                                  1.0, 0.0, ithr, out))

    def timed(fn):
        fn()
        _native.check(lib.vlx_context_synchronize(ctx))
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        _native.check(lib.vlx_context_synchronize(ctx))
        return (time.perf_counter() - t0) * 1e3 / steps

    rows, class_rows = [], {}
    try:
        ms_one = timed(lambda: single(0, d_one.ptr))
        singles = []
        for i in range(kmax):
            single(i, d_one.ptr)
            singles.append(d_one.download())
        for k in counts:
            labels = (ctypes.c_char_p * k)(*([b'2jk'] * k))

            def multi():
                _native.check(lib.vlx_fock_compute_multi_device(ctx, scr.handle, k, d_d.ptr, None, labels, 1.0, 0.0,
                                                                ithr, d_g.ptr))
            ms_k = timed(multi)
            ms_sep = timed(lambda: [single(i, d_one.ptr) for i in range(k)])
            if k == max(counts):
                # per class, one stream: the pass's kernel time against k single-density launches
                _native.check(lib.vlx_context_set_streams(ctx, 1))
                try:
                    per = {}
                    for name, fn in (('one', lambda: single(0, d_one.ptr)), ('multi', multi)):
                        fn()
                        _native.check(lib.vlx_context_synchronize(ctx))
                        _native.check(lib.vlx_context_set_timing(ctx, 1))
                        _native.check(lib.vlx_context_reset_times(ctx))
                        fn()
                        _native.check(lib.vlx_context_synchronize(ctx))
                        _native.check(lib.vlx_context_set_timing(ctx, 0))
                        per[name] = {'%d%d%d%d' % c: v[0] for c, v in _native.class_times(0).items()}
                finally:
                    _native.check(lib.vlx_context_set_streams(ctx, _native.default_streams()))
                class_rows = {c: {'ms_one_density': round(per['one'][c], 3), 'ms_pass': round(per['multi'].get(c, 0.0), 3),
                                  'pass_vs_separate': round(per['multi'].get(c, 0.0) / max(k * per['one'][c], 1e-9), 3)}
                              for c in per['one']}
            g = d_g.download()
            err = max(float(np.max(np.abs(g[i] - singles[i]))) for i in range(k))
            rows.append({'densities': k, 'ms_one_pass': ms_k, 'ms_separate_builds': ms_sep,
                         'cost_in_single_builds': ms_k / ms_one, 'speedup_vs_separate': ms_sep / ms_k,
                         'max_abs_dG_vs_single_builds': err})
            log(f'multi-density: {k} densities in {ms_k:.1f} ms ({ms_k / ms_one:.2f} builds), separately '
                f'{ms_sep:.1f} ms, max|dG| {err:.1e}')
    finally:
        d_d.free()
        d_g.free()
        d_one.free()
    return {'ms_single_build': ms_one, 'label': '2jk', 'steps': steps, 'rows': rows,
            f'classes_one_stream_{max(counts)}_densities': class_rows}


def cpu_baseline(mol, bas, dm, ithr, target_s, basis_name, g_gpu, repeats=1):
    """The reference's own C++ Fock path (oracle/_ref, CFockDriver::compute) on the host cores,
    median of `repeats` full builds, with its elementwise difference from the GPU build of the same
    D; where that harness is not built, the oracle over every k-th bra pair, extrapolated."""
    from oracle import refharness
    cores = min(os.cpu_count() or 1, 16)
    if refharness.available():
        g_ref, info = refharness.fock(mol, basis_name, dm, '2jk', 1.0, 0.0, ithr, repeats=repeats, nthreads=cores)
        return {'value': 1.0 / info['median_s'], 'unit': 'builds/s', 'cores': cores, 'kind': 'reference',
                'sample': f'{repeats} full 2jk build(s) of the same workload by the reference '
                          f'CFockDriver::compute (OpenMP, {cores} threads), median {info["median_s"]:.2f} s; '
                          f'its screener ({info["screener_s"]:.2f} s) not included',
                'max_abs_dG_gpu_vs_reference': float(np.max(np.abs(g_ref - g_gpu)))}
    from oracle import oracle as orc
    ob = orc.basis_from(bas, mol)
    _, bins = orc.pair_maxima(ob)  # screening set-up, not timed
    # calibrate with a coarse sample, then size the sample to ~target_s
    stride = 1024
    t0 = time.perf_counter()
    nq, total = orc.fock_sampled(ob, dm, '2jk', 1.0, ithr, cores, stride, bins)
    dt = time.perf_counter() - t0
    if dt < target_s / 4 and nq > 0:
        stride = max(1, int(stride * dt / target_s))
        t0 = time.perf_counter()
        nq, total = orc.fock_sampled(ob, dm, '2jk', 1.0, ithr, cores, stride, bins)
        dt = time.perf_counter() - t0
    builds_per_s = (nq / total) / dt if nq else 0.0
    return {'value': builds_per_s, 'unit': 'builds/s', 'cores': cores, 'kind': 'port',
            'sample': f'oracle 2jk over every {stride}-th bra shell pair ({nq} of {total} shell quartets, '
                      f'{dt:.1f} s), extrapolated by quartet count'}


def save_variants(scr, path):
    """The screener's per-class kernel variants and measured class times, as JSON."""
    import ctypes
    from veloxchem_amd import _native
    lib = _native.lib()
    n = ctypes.c_int()
    _native.check(lib.vlx_fock_get_variants(scr.handle, 0, ctypes.byref(n), None, None, None))
    cls = np.zeros(4 * max(n.value, 1), dtype=np.int32)
    var = np.zeros(max(n.value, 1), dtype=np.int32)
    ms = np.zeros(max(n.value, 1))
    _native.check(lib.vlx_fock_get_variants(scr.handle, n.value, ctypes.byref(n), cls.ctypes.data_as(_native._c_int_p),
                                            var.ctypes.data_as(_native._c_int_p), ms.ctypes.data_as(_native._c_double_p)))
    with open(path, 'w') as fh:
        json.dump({'classes': cls[:4 * n.value].tolist(), 'variants': var[:n.value].tolist(),
                   'ms': ms[:n.value].tolist()}, fh)


def load_variants(scr, path):
    """Applies variants saved by save_variants (a profiling run then times the same kernels
    without an autotuner pass of its own)."""
    from veloxchem_amd import _native
    with open(path) as fh:
        d = json.load(fh)
    cls, pc = _native.as_int(d['classes'])
    var, pv = _native.as_int(d['variants'])
    ms, pm = _native.as_double(d['ms'])
    _native.check(_native.lib().vlx_fock_set_variants(_native.context(0), scr.handle, len(var), pc, pv, pm))
    return {tuple(int(x) for x in cls[4 * r:4 * r + 4]): int(var[r]) for r in range(len(var))}


def workload_builds(nmol, box, seed, basis, ithr, steps, warmup, taus, autotune_repeats, workload_key,
                    save_tune=None, load_tune=None, single_stream=True):
    """FP64 builds of one seeded water cluster after its own autotuner pass: rate, per-class kernel
    times and the roofline of the dominant class kernel (timed region, three streams), the same
    kernel alone on one stream, and the mixed sweep with its error against FP64."""
    from veloxchem_amd import FockDriver, _native
    mol, bas, scr, dm, t_screen = build_workload(nmol, box, seed, basis)
    n = bas.get_dimensions_of_basis()
    log(f'{workload_key}: {n} AOs, screener {t_screen:.2f} s')
    tuned, t_tune = None, 0.0
    if load_tune:
        tuned = load_variants(scr, load_tune)
        log(f'{workload_key}: kernel variants of {len(tuned)} classes from {load_tune}')
    elif autotune_repeats > 0:
        t0 = time.perf_counter()
        tuned = FockDriver()._autotune(scr, dm, ithr, autotune_repeats)
        t_tune = time.perf_counter() - t0
        log(f'{workload_key}: autotuned in {t_tune:.1f} s')
        if save_tune:
            save_variants(scr, save_tune)
    ms, times, g, sweep = measure_builds(scr, dm, ithr, steps, warmup, taus)
    log(f'{workload_key}: {1e3 / ms:.3f} builds/s over {steps} builds; mixed sweep of {len(taus)} thresholds done')
    work = scr.work(ithr)
    class_flops = ref_scheme_flops(work)
    rf = roofline(scr, ithr, times, ms, workload_key)
    # the classes timed alone: one stream serialises the class kernels (in the timed region four
    # streams share the chip, so each launch's time includes its neighbours)
    lib, ctx = _native.lib(), _native.context(0)
    rf1 = None
    if single_stream:
        _native.check(lib.vlx_context_set_streams(ctx, 1))
        try:
            ms1, times1, _, _ = measure_builds(scr, dm, ithr, 1, 1, [])
        finally:
            _native.check(lib.vlx_context_set_streams(ctx, _native.default_streams()))
        rf1 = roofline(scr, ithr, times1, ms1, workload_key)
    out = {'workload': f'{nmol}-water cluster (seed {seed}, {box} A box), {basis}, {n} AOs, 2jk FP64 build on the '
                       f'SAD density, ithreshold {ithr}',
           'nao': n, 'builds_per_s': 1e3 / ms, 'ms_per_step': ms, 'steps': steps, 'warmup': warmup,
           'shell_quartets': int(sum(w[1] for w in work)), 'primitive_quartets': int(sum(w[2] for w in work)),
           'screener_setup_s': round(t_screen, 3), 'autotune_s': round(t_tune, 2),
           'roofline': rf,
           'class_ms': {'%d%d%d%d' % c: round(v[0] / v[1], 4) for c, v in sorted(times.items())},
           # reference-scheme FLOPs of each class per build, and its TFLOP/s in the timed region
           'class_flops': {'%d%d%d%d' % c: float(f) for c, f in sorted(class_flops.items())},
           'class_tflops': {'%d%d%d%d' % c: round(class_flops[c] / (v[0] / v[1] * 1e-3) / 1e12, 3)
                            for c, v in sorted(times.items()) if c in class_flops},
           'checksum': float(np.sum(g * dm))}
    if rf1 is not None:
        out['roofline_single_stream'] = {k: rf1[k] for k in ('kernel', 'class', 'achieved', 'frac', 'kernel_ms')} | {
            'build_ms_one_stream': ms1}
        out['class_ms_one_stream'] = {'%d%d%d%d' % c: round(v[0] / v[1], 4) for c, v in sorted(times1.items())}
        out['class_tflops_one_stream'] = {'%d%d%d%d' % c: round(class_flops[c] / (v[0] / v[1] * 1e-3) / 1e12, 3)
                                          for c, v in sorted(times1.items()) if c in class_flops}
    if tuned:
        out['autotune'] = {'%d%d%d%d' % c: v for c, v in sorted(tuned.items())}
    if sweep:
        out['mixed_sweep'] = sweep
    return out, (mol, bas, scr, dm, g)


def batched_clusters(args):
    """Batched uneven builds: 256 water clusters of 1-12 molecules (sizes drawn with seed 4,
    cluster i with seed 4000 + i, box scaled to the 32-water density), def2-SVP, SAD densities,
    resident in HBM.  Aggregate builds/s of one batched call (one launch per angular class over
    every molecule, replayed as a captured HIP graph) against the same molecules built one after
    another, in FP64 and in mixed precision."""
    import ctypes
    from veloxchem_amd import MolecularBasis, T4CScreener, _native
    from veloxchem_amd.sadguess import sad_density
    from veloxchem_amd.workloads import water_cluster
    lib = _native.lib()
    ctx = _native.context(0)
    sizes = np.random.Generator(np.random.PCG64(4)).integers(1, 13, args.batch_clusters)
    scrs, d_dm, d_g, naos = [], [], [], []
    t0 = time.perf_counter()
    for i, nw in enumerate(sizes):
        mol = water_cluster(int(nw), 10.0 * (nw / 32.0) ** (1.0 / 3.0), 4000 + i)
        bas = MolecularBasis.read(mol, 'def2-svp')
        scr = T4CScreener()
        scr.partition(bas, mol, 'eri')
        dm = sad_density(mol, MolecularBasis.read(mol, 'ao-start-guess'), bas)
        n = bas.get_dimensions_of_basis()
        a, b = _native.DeviceArray((n, n)), _native.DeviceArray((n, n))
        a.upload(dm)
        scrs.append(scr)
        d_dm.append(a)
        d_g.append(b)
        naos.append(n)
    t_setup = time.perf_counter() - t0
    nmol = len(scrs)
    arr = ctypes.c_void_p * nmol
    h_scr = arr(*[s.handle.value for s in scrs])
    h_dm = arr(*[a.ptr.value for a in d_dm])
    h_g = arr(*[b.ptr.value for b in d_g])

    def one_by_one():
        for s_, a, b in zip(scrs, d_dm, d_g):
            _native.check(lib.vlx_fock_compute_device(ctx, s_.handle, a.ptr, b'2jk', 1.0, 0.0, args.ithreshold, b.ptr))

    def batched():
        _native.check(lib.vlx_fock_compute_batch_device(ctx, nmol, h_scr, h_dm, b'2jk', 1.0, 0.0, args.ithreshold, h_g))

    def timed(fn, reps):
        fn()
        _native.check(lib.vlx_context_synchronize(ctx))
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        _native.check(lib.vlx_context_synchronize(ctx))
        return (time.perf_counter() - t) / reps

    t_seq = timed(one_by_one, args.batch_reps)
    g_seq = [b.download() for b in d_g]
    t_bat = timed(batched, args.batch_reps)
    g_bat = [b.download() for b in d_g]
    diff = max(float(np.max(np.abs(x - y))) for x, y in zip(g_seq, g_bat))
    # mixed precision at the loosest swept threshold: one-by-one (host reads max|D| per build) and
    # batched (split on the device), both against the FP64 one-by-one G
    tau = max([float(x) for x in args.mixed.split(',') if x.strip()] or [1e-6])
    try:
        _native.set_precision('mixed', tau)
        t_seq32 = timed(one_by_one, args.batch_reps)
        g_seq32 = [b.download() for b in d_g]
        t_bat32 = timed(batched, args.batch_reps)
        g_bat32 = [b.download() for b in d_g]
    finally:
        _native.set_precision('fp64')
    diff32_seq = max(float(np.max(np.abs(x - y))) for x, y in zip(g_seq, g_seq32))
    diff32_bat = max(float(np.max(np.abs(x - y))) for x, y in zip(g_seq, g_bat32))
    quartets = sum(sum(w[1] for w in s_.work(args.ithreshold)) for s_ in scrs)
    for a, b in zip(d_dm, d_g):
        a.free()
        b.free()
    return {'workload': f'{nmol} water clusters of 1-12 molecules (sizes seed 4, cluster seeds 4000+i, '
                        f'box 10 A x (n/32)^(1/3)), def2-SVP, {min(naos)}-{max(naos)} AOs, '
                        f'{sum(naos)} AOs in all, 2jk on SAD densities, ithreshold {args.ithreshold}',
            'batched_builds_per_s': nmol / t_bat, 'batched_ms_per_batch': t_bat * 1e3,
            'one_by_one_builds_per_s': nmol / t_seq, 'one_by_one_ms_per_batch': t_seq * 1e3,
            'speedup': t_seq / t_bat, 'max_abs_dG_batched_vs_one_by_one': diff,
            'mixed': {'threshold': tau, 'batched_builds_per_s': nmol / t_bat32,
                      'one_by_one_builds_per_s': nmol / t_seq32, 'speedup': t_seq32 / t_bat32,
                      'batched_vs_fp64_one_by_one': t_seq / t_bat32,
                      'max_abs_dG_batched_vs_fp64': diff32_bat, 'max_abs_dG_one_by_one_vs_fp64': diff32_seq},
            'shell_quartets': int(quartets), 'reps': args.batch_reps, 'setup_s': round(t_setup, 2)}


def learned_guess(args):
    """The learned SCF guess: a network trained on the GPU (BF16 matrix-core kernels) on converged
    RHF densities of seeded water clusters of 4-16 molecules (sizes seed 5, cluster i seed 10000 + i,
    box scaled to the 32-water density), from this package's own SCF; then the SCF of held-out
    clusters of 16-32 molecules (cluster i seed 20000 + i) from SAD and from the learned guess, with
    iterations and the converged energies."""
    from veloxchem_amd import MolecularBasis
    from veloxchem_amd.learnedguess import LearnedGuess, make_dataset
    from veloxchem_amd.scfdriver import ScfRestrictedDriver
    from veloxchem_amd.workloads import water_cluster

    def cluster(nw, seed):
        return water_cluster(int(nw), 10.0 * (nw / 32.0) ** (1.0 / 3.0), seed)

    t0 = time.perf_counter()
    sizes = np.random.Generator(np.random.PCG64(5)).integers(4, 17, args.guess_train)
    X, T, S = make_dataset([cluster(nw, 10000 + i) for i, nw in enumerate(sizes)])
    t_data = time.perf_counter() - t0
    log(f'learned guess: {len(X)} pair samples from {len(sizes)} converged SCFs in {t_data:.1f} s')
    guess = LearnedGuess(max_batch=4096, seed=1)
    t0 = time.perf_counter()
    loss = guess.train(X, T, S, epochs=args.guess_epochs, batch=2048, lr=2e-3)
    t_train = time.perf_counter() - t0
    log(f'learned guess: trained {args.guess_epochs} epochs in {t_train:.1f} s, loss {loss:.3e}')
    esizes = np.random.Generator(np.random.PCG64(6)).integers(16, 33, args.guess_eval)
    rows = []
    for i, nw in enumerate(esizes):
        mol = cluster(nw, 20000 + i)
        bas = MolecularBasis.read(mol, 'def2-svp')
        res = {}
        for name, g in (('sad', None), ('learned', guess)):
            drv = ScfRestrictedDriver()
            drv.acc_type = 'DIIS'
            drv.conv_thresh = 1.0e-8
            drv.guess = g
            r = drv.compute(mol, bas)
            res[name] = (r['num_iter'], r['scf_energy'], r['is_converged'], r['timings']['total'])
        rows.append({'waters': int(nw), 'seed': 20000 + i, 'sad_iterations': res['sad'][0],
                     'learned_iterations': res['learned'][0], 'dE': res['learned'][1] - res['sad'][1],
                     'converged': bool(res['sad'][2] and res['learned'][2]),
                     'sad_scf_s': round(res['sad'][3], 3), 'learned_scf_s': round(res['learned'][3], 3)})
        log(f'learned guess: {nw} waters, SAD {res["sad"][0]} vs learned {res["learned"][0]} iterations')
    it_sad = sum(r['sad_iterations'] for r in rows)
    it_learned = sum(r['learned_iterations'] for r in rows)
    return {'workload': f'train: {len(sizes)} water clusters of 4-16 molecules (sizes seed 5, seeds 10000+i), '
                        f'def2-SVP converged RHF densities of this package\'s SCF; evaluate: {len(esizes)} held-out '
                        f'clusters of 16-32 molecules (sizes seed 6, seeds 20000+i), SCF to 1e-8',
            'pair_samples': int(len(X)), 'data_s': round(t_data, 2), 'train_s': round(t_train, 2),
            'epochs': args.guess_epochs, 'final_loss': loss,
            'iteration_cut': 1.0 - it_learned / max(1, it_sad), 'sad_iterations': it_sad,
            'learned_iterations': it_learned, 'max_abs_dE': max(abs(r['dE']) for r in rows),
            'clusters': rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--ithreshold', type=int, default=12)
    ap.add_argument('--cpu-seconds', type=float, default=15.0)
    ap.add_argument('--cpu-repeats', type=int, default=1,
                    help='reference CPU builds of the headline workload (median); about 90 s each on 16 cores')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-autotune', action='store_true', help='skip the per-class kernel-variant autotuner')
    ap.add_argument('--mixed', default='1e-6,1e-8,1e-10',
                    help='Schwarz x max|D| thresholds of the mixed-precision sweep (comma list, empty: none)')
    ap.add_argument('--scf-conv', type=float, default=1.0e-8)
    ap.add_argument('--only-headline', action='store_true',
                    help='the 64-water def2-TZVP builds only (profiling runs): no sweep, SCF, extras or CPU baseline')
    ap.add_argument('--svp-only', action='store_true',
                    help='the 32-water def2-SVP builds only (profiling runs)')
    ap.add_argument('--no-svp', action='store_true', help='skip the 32-water def2-SVP builds and SCF')
    ap.add_argument('--no-scf', action='store_true')
    ap.add_argument('--svp-steps', type=int, default=10)
    ap.add_argument('--save-tune', default=None, help='write the headline workload\'s autotuned variants here')
    ap.add_argument('--load-tune', default=None,
                    help='take the headline workload\'s variants from this file instead of an autotuner pass '
                         '(profiling runs: the profile then holds the timed builds only)')
    ap.add_argument('--no-single-stream', action='store_true',
                    help='skip the single-stream measurement of the dominant class (profiling runs)')
    ap.add_argument('--no-batch', action='store_true',
                    help='skip the batched workload of 256 small water clusters')
    ap.add_argument('--batch-clusters', type=int, default=256)
    ap.add_argument('--batch-reps', type=int, default=3)
    ap.add_argument('--no-guess', action='store_true', help='skip the learned-guess training and evaluation')
    ap.add_argument('--no-multi', action='store_true', help='skip the multi-density measurements')
    ap.add_argument('--density-tau', type=float, default=1e-10,
                    help='density-weighted screening threshold of the fast SCF mode')
    ap.add_argument('--no-tzvp-scf', action='store_true',
                    help='skip the FP64 and fast SCFs of the 64-water def2-TZVP cluster')
    ap.add_argument('--guess-train', type=int, default=160, help='training clusters of the learned guess')
    ap.add_argument('--guess-eval', type=int, default=6, help='held-out evaluation clusters of the learned guess')
    ap.add_argument('--guess-epochs', type=int, default=150)
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py: this build runs on one GPU; the multi-GPU sharding is design only '
              '(one process per GPU over RCCL is not built). Use --gpus 1.', file=sys.stderr)
        sys.exit(2)

    from veloxchem_amd import _native
    _native.lib()
    taus = [float(x) for x in args.mixed.split(',') if x.strip()]
    if args.only_headline or args.svp_only:
        taus = []
    atr = 0 if args.no_autotune else 1
    if args.svp_only:
        svp, _ = workload_builds(32, 10.0, 32, 'def2-svp', args.ithreshold, args.steps, args.warmup, [],
                                 0 if args.no_autotune else 2, 'water32_svp', args.save_tune, args.load_tune,
                                 not args.no_single_stream)
        print(json.dumps({'svp32': svp}, default=_plain))
        return
    head, (mol, bas, scr, dm, g) = workload_builds(64, 12.8, 64, 'def2-tzvp', args.ithreshold, args.steps,
                                                   args.warmup, taus, atr, 'water64_tzvp', args.save_tune,
                                                   args.load_tune, not args.no_single_stream)
    result = {
        'metric': 'J+K Fock builds/s + SCF time, mixed vs FP64 on GPU, at |\u0394E|<=1e-6 Eh',
        'mode': 'fp64 2jk build of the 64-water def2-TZVP cluster (value); mixed sweep on the same density',
        'value': head['builds_per_s'],
        'unit': 'builds/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': head['ms_per_step'],
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic (seeded water cluster, SAD guess density)',
        'config': {
            'workload': head['workload'],
            'nao': head['nao'],
            'shell_quartets': head['shell_quartets'],
            'primitive_quartets': head['primitive_quartets'],
            'screener_setup_s': head['screener_setup_s'],
            'autotune_s': head['autotune_s'],
        },
        'roofline': head['roofline'],
    }
    for k in ('roofline_single_stream', 'class_ms', 'class_tflops', 'class_ms_one_stream', 'class_tflops_one_stream',
              'class_flops', 'checksum', 'autotune', 'mixed_sweep'):
        if k in head:
            result[k] = head[k]
    if not args.only_headline:
        if not args.no_svp:
            svp, (smol, sbas, scr32, dm32, _) = workload_builds(32, 10.0, 32, 'def2-svp', args.ithreshold,
                                                                args.svp_steps, 2, taus,
                                                                0 if args.no_autotune else 2, 'water32_svp')
            result['svp32'] = svp
            if not args.no_scf:
                # the dense-algebra libraries load their kernels for each matrix size on first use: one
                # untimed SCF of the same molecule first (every run below builds its own screener)
                log('SCF runs')
                run_scf(smol, sbas, 1.0e-4)
                result['scf'] = run_scf(smol, sbas, args.scf_conv)
                inc = run_scf(smol, sbas, args.scf_conv, incremental=True)
                inc['dE_vs_fp64'] = inc['energy'] - result['scf']['energy']
                result['scf_incremental'] = inc
                if taus:
                    tau = max(taus)
                    for key, incr in (('scf_mixed', False), ('scf_mixed_incremental', True)):
                        mixed = run_scf(smol, sbas, args.scf_conv, 'mixed', tau, incremental=incr)
                        mixed['dE_vs_fp64'] = mixed['energy'] - result['scf']['energy']
                        result[key] = mixed
                    result['scf_fast_vs_fp64'] = fast_vs_fp64_scf(smol, sbas, args.scf_conv, tau, args.density_tau)
            if not args.no_multi:
                log('multi-density pass, 32-water def2-SVP')
                result['svp32']['multi_density'] = multi_density(scr32, dm32, args.ithreshold)
        if not args.no_multi:
            log('multi-density pass, 64-water def2-TZVP')
            result['multi_density'] = multi_density(scr, dm, args.ithreshold, counts=(2, 4), steps=2)
        if not args.no_tzvp_scf and taus:
            log('64-water def2-TZVP SCFs: FP64 and fast')
            result['scf_tzvp_fast_vs_fp64'] = fast_vs_fp64_scf(mol, bas, args.scf_conv, max(taus), args.density_tau)
        if not args.no_batch:
            log('batched clusters')
            result['batched_clusters'] = batched_clusters(args)
        if not args.no_guess:
            log('learned guess')
            result['learned_guess'] = learned_guess(args)
        if not args.no_cpu_baseline:
            log('reference CPU build of the headline workload')
            result['cpu_baseline'] = cpu_baseline(mol, bas, dm, args.ithreshold, args.cpu_seconds, 'def2-tzvp', g,
                                                  repeats=args.cpu_repeats)
    print(json.dumps(result, default=_plain))


def _plain(o):
    """numpy scalars and arrays as plain JSON values."""
    if isinstance(o, np.generic):
        return o.item()
    if isinstance(o, np.ndarray):
        return o.tolist()
    raise TypeError(f'{type(o).__name__} is not JSON serialisable')


if __name__ == '__main__':
    main()
