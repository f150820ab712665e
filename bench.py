"""Benchmark of the two-electron J+K Fock build on one MI355X.

Workload (one "step" = one build): seeded 32-water cluster in a 10 Angstrom
box (seed 32), def2-SVP (768 AOs), FP64 G = 2J - K ('2jk', exchange factor
1, omega 0, ithreshold 12) on the SAD guess density.  D and G stay resident
in HBM; the timed region is K builds between stream synchronisations.  After
the timed builds, the full RHF SCF of the same cluster (L2-DIIS, gradient
norm 1e-8) is run once and its wall time reported beside the build rate.

Printed: one JSON line with value = builds/s, the roofline of the dominant
class kernel (FP64 vector-ALU bound; per-launch algorithmic FLOPs over the
HIP-event kernel time measured inside the timed region), and a CPU baseline
(the project's C oracle on the host cores over a bounded strided sample of
the same build, extrapolated to builds/s).
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FP64_VALU_PEAK_TFLOPS = 78.6
# HBM bytes per launch per kernel, from FETCH_SIZE/WRITE_SIZE passes scaled by a copy calibration
TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'fock_build_water32_svp_traffic.json')


def kernel_traffic(name):
    try:
        with open(TRAFFIC_FILE) as fh:
            return json.load(fh)['bytes_per_launch'].get(name)
    except (OSError, ValueError, KeyError):
        return None


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


def run_scf(mol, bas, conv_thresh, precision='fp64', tau=0.0, incremental=False):
    from veloxchem_amd.scfdriver import ScfRestrictedDriver
    drv = ScfRestrictedDriver()
    drv.conv_thresh = conv_thresh
    drv.precision = precision
    drv.mixed_threshold = tau
    drv.incremental_fock = incremental
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
    return out


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

    def timed(k):
        for _ in range(max(1, warmup)):
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
        for t in taus:
            _native.set_precision('mixed', t)
            ms_m = timed(steps)
            gm = d_g.download()
            k32 = _native.last_fp32_bin_sum()
            p64, p32 = scr.precision_split(ithr, k32)
            entry = {'threshold': t, 'builds_per_s': 1e3 / ms_m, 'ms_per_step': ms_m, 'fp32_bin_sum': k32,
                     'fp32_prim_fraction': p32 / max(1, p64 + p32), 'max_abs_dG': float(np.max(np.abs(gm - g64))),
                     'dE2_vs_fp64': float(np.sum(dm * (gm - g64)))}
            entry['speedup_vs_fp64'] = ms64 / ms_m
            sweep.append(entry)
    finally:
        _native.set_precision('fp64')
        d_dm.free()
        d_g.free()
    return ms, times, g, sweep


def cpu_baseline(mol, bas, dm, ithr, target_s, basis_name, g_gpu):
    """The reference's own C++ Fock path (oracle/_ref, CFockDriver::compute) on the host cores for
    one full build, with its elementwise difference from the GPU build of the same D; where that
    harness is not built, the oracle over every k-th bra pair, extrapolated."""
    from oracle import refharness
    cores = min(os.cpu_count() or 1, 16)
    if refharness.available():
        g_ref, info = refharness.fock(mol, basis_name, dm, '2jk', 1.0, 0.0, ithr, repeats=1, nthreads=cores)
        return {'value': 1.0 / info['median_s'], 'unit': 'builds/s', 'cores': cores, 'kind': 'reference',
                'sample': f'one full 2jk build of the same workload by the reference CFockDriver::compute '
                          f'(OpenMP, {cores} threads): {info["median_s"]:.2f} s; its screener '
                          f'{info["screener_s"]:.2f} s not included',
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


def tzvp_builds(args, taus):
    """Mixed-precision workload: seeded 64-water cluster (12.8 A box, seed 64), def2-TZVP
    (2752 AOs, f shells): FP64 build rate and the mixed sweep with its error against FP64."""
    mol, bas, scr, dm, t_screen = build_workload(64, 12.8, 64, 'def2-tzvp')
    t_tune = 0.0
    if not args.no_tzvp_autotune:
        from veloxchem_amd import FockDriver
        t0 = time.perf_counter()
        FockDriver()._autotune(scr, dm, args.ithreshold, 1)
        t_tune = time.perf_counter() - t0
    ms, times, _, sweep = measure_builds(scr, dm, args.ithreshold, args.tzvp_steps, 1, taus)
    work = scr.work(args.ithreshold)
    flops = sum(w[3] for w in work)
    return {'workload': '64-water cluster (seed 64, 12.8 A box), def2-TZVP, '
                        f'{bas.get_dimensions_of_basis()} AOs, 2jk on the SAD density, ithreshold {args.ithreshold}',
            'fp64_builds_per_s': 1e3 / ms, 'fp64_ms_per_step': ms, 'steps': args.tzvp_steps,
            'build_achieved_tflops': flops / (ms * 1e-3) / 1e12, 'screener_setup_s': round(t_screen, 3),
            'autotune_s': round(t_tune, 1),
            'shell_quartets': int(sum(w[1] for w in work)), 'mixed_sweep': sweep}


def batched_clusters(args):
    """Batched uneven builds: 256 water clusters of 1-12 molecules (sizes drawn with seed 4,
    cluster i with seed 4000 + i, box scaled to the 32-water density), def2-SVP, SAD densities,
    resident in HBM.  Aggregate builds/s of one batched call (each molecule whole on one stream,
    molecules concurrent) against the same molecules built one after another."""
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
            'shell_quartets': int(quartets), 'reps': args.batch_reps, 'setup_s': round(t_setup, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--nmol', type=int, default=32)
    ap.add_argument('--box', type=float, default=10.0)
    ap.add_argument('--seed', type=int, default=32)
    ap.add_argument('--basis', default='def2-svp')
    ap.add_argument('--ithreshold', type=int, default=12)
    ap.add_argument('--cpu-seconds', type=float, default=15.0)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-scf', action='store_true')
    ap.add_argument('--no-autotune', action='store_true', help='skip the per-class kernel-variant autotuner')
    ap.add_argument('--precision', default='fp64', choices=['fp64', 'mixed'],
                    help='precision of the main timed builds (value)')
    ap.add_argument('--tau', type=float, default=1e-6, help='mixed-precision threshold of --precision mixed')
    ap.add_argument('--mixed', default='1e-6,1e-8,1e-10',
                    help='Schwarz x max|D| thresholds of the mixed-precision sweep (comma list, empty: none)')
    ap.add_argument('--scf-conv', type=float, default=1.0e-8)
    ap.add_argument('--label', default='2jk', choices=['2jk', 'j', 'k'])
    ap.add_argument('--no-tzvp', action='store_true',
                    help='skip the 64-water def2-TZVP FP64 build and mixed-precision sweep')
    ap.add_argument('--tzvp-steps', type=int, default=2)
    ap.add_argument('--no-tzvp-autotune', action='store_true',
                    help='run the def2-TZVP builds with the default kernel variants (no autotuner pass)')
    ap.add_argument('--no-batch', action='store_true',
                    help='skip the batched workload of 256 small water clusters')
    ap.add_argument('--batch-clusters', type=int, default=256)
    ap.add_argument('--batch-reps', type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py: this build runs on one GPU; the multi-GPU sharding is design only '
              '(one process per GPU over RCCL is not built). Use --gpus 1.', file=sys.stderr)
        sys.exit(2)

    from veloxchem_amd import _native
    _native.lib()
    mol, bas, scr, dm, t_screen = build_workload(args.nmol, args.box, args.seed, args.basis)
    n = bas.get_dimensions_of_basis()

    tuned, t_tune = None, 0.0
    if not args.no_autotune:
        from veloxchem_amd import FockDriver
        t0 = time.perf_counter()
        tuned = FockDriver()._autotune(scr, dm, args.ithreshold, repeats=2)
        t_tune = time.perf_counter() - t0

    taus = [float(x) for x in args.mixed.split(',') if x.strip()]
    ms_per_step, times, g, sweep = measure_builds(scr, dm, args.ithreshold, args.steps, args.warmup, taus,
                                                  args.label.encode(), args.precision,
                                                  args.tau if args.precision == 'mixed' else 0.0)
    work = scr.work(args.ithreshold)
    flops_build = sum(w[3] for w in work)
    by_class = {w[0]: w for w in work}
    dom = max(times, key=lambda c: times[c][0])
    dom_ms = times[dom][0] / times[dom][1]
    dom_flops = by_class[dom][3]
    achieved = dom_flops / (dom_ms * 1e-3) / 1e12
    dom_kernel = _native.class_kernel(scr, args.ithreshold, dom)
    build_tflops = flops_build / (ms_per_step * 1e-3) / 1e12

    # the same kernel timed alone: one stream serialises the class kernels (in the timed region
    # above four streams share the chip, so each kernel's launch time includes its neighbours)
    lib = _native.lib()
    ctx = _native.context(0)
    _native.check(lib.vlx_context_set_streams(ctx, 1))
    try:
        ms1, times1, _, _ = measure_builds(scr, dm, args.ithreshold, 2, 1, [], args.label.encode())
    finally:
        _native.check(lib.vlx_context_set_streams(ctx, max(1, min(4, int(os.environ.get('VLX_STREAMS', 4))))))
    dom1 = max(times1, key=lambda c: times1[c][0])
    dom1_ms = times1[dom1][0] / times1[dom1][1]
    achieved1 = by_class[dom1][3] / (dom1_ms * 1e-3) / 1e12

    result = {
        'metric': 'J+K Fock builds/s + SCF time, mixed vs FP64 on GPU, at |\u0394E|<=1e-6 Eh',
        'mode': 'fp64 2jk build (value); fp64 RHF SCF (scf)',
        'value': 1e3 / ms_per_step,
        'unit': 'builds/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': ms_per_step,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic (seeded water cluster, SAD guess density)',
        'config': {
            'workload': f'{args.nmol}-water cluster (seed {args.seed}, {args.box} A box), {args.basis}, '
                        f'{n} AOs, 2jk FP64 build on the SAD density, ithreshold {args.ithreshold}',
            'nao': n,
            'shell_quartets': int(sum(w[1] for w in work)),
            'primitive_quartets': int(sum(w[2] for w in work)),
            'screener_setup_s': round(t_screen, 3),
            'autotune_s': round(t_tune, 3),
        },
        'roofline': {
            'bound': 'valu',
            'kernel': dom_kernel,
            'achieved': achieved,
            'peak': FP64_VALU_PEAK_TFLOPS,
            'unit': 'TFLOP/s',
            'frac': achieved / FP64_VALU_PEAK_TFLOPS,
            'traffic': kernel_traffic(dom_kernel),
            'kernel_ms': dom_ms,
            'kernel_share_of_build': times[dom][0] / sum(v[0] for v in times.values()),
            'build_achieved': build_tflops,
            'build_frac': build_tflops / FP64_VALU_PEAK_TFLOPS,
        },
        'roofline_single_stream': {
            'kernel': _native.class_kernel(scr, args.ithreshold, dom1),
            'achieved': achieved1,
            'peak': FP64_VALU_PEAK_TFLOPS,
            'unit': 'TFLOP/s',
            'frac': achieved1 / FP64_VALU_PEAK_TFLOPS,
            'kernel_ms': dom1_ms,
            'build_ms_one_stream': ms1,
        },
        'class_ms': {'%d%d%d%d' % c: round(v[0] / v[1], 4) for c, v in sorted(times.items())},
        'checksum': float(np.sum(g * dm)),
    }
    if tuned:
        result['autotune'] = {'%d%d%d%d' % c: v for c, v in sorted(tuned.items())}
    if sweep:
        result['mixed_sweep'] = sweep
    if not args.no_scf:
        # the dense-algebra libraries load their kernels for each matrix size on first use: one
        # untimed SCF of the same molecule first (every run below builds its own screener)
        run_scf(mol, bas, 1.0e-4)
        result['scf'] = run_scf(mol, bas, args.scf_conv)
        inc = run_scf(mol, bas, args.scf_conv, incremental=True)
        inc['dE_vs_fp64'] = inc['energy'] - result['scf']['energy']
        result['scf_incremental'] = inc
        if taus:
            tau = max(taus)
            for key, incr in (('scf_mixed', False), ('scf_mixed_incremental', True)):
                mixed = run_scf(mol, bas, args.scf_conv, 'mixed', tau, incremental=incr)
                mixed['dE_vs_fp64'] = mixed['energy'] - result['scf']['energy']
                result[key] = mixed
    if not args.no_tzvp:
        result['tzvp64'] = tzvp_builds(args, taus)
    if not args.no_batch:
        result['batched_clusters'] = batched_clusters(args)
    if not args.no_cpu_baseline:
        result['cpu_baseline'] = cpu_baseline(mol, bas, dm, args.ithreshold, args.cpu_seconds, args.basis, g)
    print(json.dumps(result))


if __name__ == '__main__':
    main()
