"""Generator of the per-class electron-repulsion device functions (gfx950).

For every shell-quartet class (la lb | lc ld) with la >= lb, lc >= ld and
pair class (la, lb) >= (lc, ld) it writes one HIP source file
``gen/eri_<la><lb><lc><ld>.hip`` that specialises

    eri_quartet<la, lb, lc, ld>(bra pair, ket pair, Boys table, out[])

as straight-line code the compiler keeps in registers:

1. per primitive quartet: Boys F_0..F_L from the Taylor table, then the
   Obara-Saika vertical recursion on centre A and electron transfer to
   centre C (Head-Gordon-Pople form): [e0|00]^(m) -> [e0|f0]^(m), every
   intermediate emitted once (memoised DAG, unused branches never emitted);
2. contraction: [e0|f0]^(0) summed over bra and ket primitive pairs;
3. horizontal recursions on the contracted values, ket first then bra,
   each followed by the Cartesian -> real-solid-harmonic transform with the
   factors and component order of src/general/SphericalMomentum.hpp:18-134
   (p = y,z,x; d = xy,yz,z2,xz,x2-y2; f as listed there).

The semantics of the produced integrals are those of the reference kernels
erirec::comp_electron_repulsion_<abcd>
(src/t4c_electron_repulsion/ElectronRepulsionRec*.hpp) -- contracted,
normalised spherical (ab|cd) -- but the code is generated from the
recurrences themselves, not from those files.

It also writes ``gen/dispatch.inc`` with the table of launchers.

Usage: python gen_eri.py OUTDIR LMAX
"""

import importlib.util
import math
import sys
from pathlib import Path


def cart_list(lang):
    return [(i, j, lang - i - j) for i in range(lang, -1, -1) for j in range(lang - i, -1, -1)]


def ncart(lang):
    return (lang + 1) * (lang + 2) // 2


def nsph(lang):
    return 2 * lang + 1


# Cartesian -> spherical factors: sph[l][m] = [(cart_index, coefficient)] (veloxchem_amd/sphericals.py)
# (loaded by file path: the package itself needs the built library)
_spec = importlib.util.spec_from_file_location('_vlx_sphericals', Path(__file__).resolve().parents[1] / 'sphericals.py')
_sph = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(_sph)
SPH = _sph.SPH


AX = 'xyz'


def lit(x):
    r = repr(float(x))
    return r if ('e' in r or '.' in r) else r + '.0'


def add(t, i, d):
    t = list(t)
    t[i] += d
    return tuple(t)


def count_flops(expr):
    """Binary floating-point operations in one generated expression."""
    return expr.count('*') + expr.count(' + ') + expr.count(' - ')


class Code:
    def __init__(self, prefix):
        self.lines = []
        self.n = 0
        self.prefix = prefix

    def new(self, expr):
        name = f'{self.prefix}{self.n}'
        self.n += 1
        self.lines.append(f'const R {name} = {expr};')
        return name


def gen_vrr(code, la, lb, lc, ld, L):
    """Emit the primitive VRR; returns {(e, f): var} for the contracted targets."""
    memo = {}

    def node(a, c, m):
        key = (a, c, m)
        if key in memo:
            return memo[key]
        if sum(c) == 0:
            if sum(a) == 0:
                memo[key] = f'b{m}'
                return memo[key]
            # raise on A: pick the axis with the smallest non-zero exponent
            cands = [i for i in range(3) if a[i] > 0]
            i = min(cands, key=lambda k: (a[k], k))
            am = add(a, i, -1)
            terms = [f'PA{AX[i]} * {node(am, c, m)}', f'WP{AX[i]} * {node(am, c, m + 1)}']
            if am[i] > 0:
                amm = add(am, i, -1)
                terms.append(f'R({lit(am[i])}) * oo2z * ({node(amm, c, m)} - rz * {node(amm, c, m + 1)})')
            memo[key] = code.new(' + '.join(terms))
            return memo[key]
        cands = [i for i in range(3) if c[i] > 0]
        i = min(cands, key=lambda k: (a[k] > 0, c[k], k))
        cm = add(c, i, -1)
        terms = [f'QC{AX[i]} * {node(a, cm, m)}', f'WQ{AX[i]} * {node(a, cm, m + 1)}']
        if cm[i] > 0:
            cmm = add(cm, i, -1)
            terms.append(f'R({lit(cm[i])}) * oo2e * ({node(a, cmm, m)} - re * {node(a, cmm, m + 1)})')
        if a[i] > 0:
            terms.append(f'R({lit(a[i])}) * oo2ze * {node(add(a, i, -1), cm, m + 1)}')
        memo[key] = code.new(' + '.join(terms))
        return memo[key]

    targets = {}
    for le in range(la, la + lb + 1):
        for e in cart_list(le):
            for lf in range(lc, lc + ld + 1):
                for f in cart_list(lf):
                    targets[(e, f)] = node(e, f, 0)
    return targets


def gen_class(la, lb, lc, ld):
    L = la + lb + lc + ld
    tag = f'{la}{lb}{lc}{ld}'
    e_list = [e for le in range(la, la + lb + 1) for e in cart_list(le)]
    f_list = [f for lf in range(lc, lc + ld + 1) for f in cart_list(lf)]
    acc_index = {}
    for ie, e in enumerate(e_list):
        for jf, f in enumerate(f_list):
            acc_index[(e, f)] = ie * len(f_list) + jf
    nacc = len(acc_index)

    vrr = Code('v')
    targets = gen_vrr(vrr, la, lb, lc, ld, L)

    out = []
    w = out.append
    w(f'// class ({la}{lb}|{lc}{ld}), L = {L}: {nacc} contracted [e0|f0] values, '
      f'{len(vrr.lines)} VRR intermediates per primitive quartet')
    w('template <>')
    w(f'struct EriClass<{la}, {lb}, {lc}, {ld}> {{')
    # RS: erf(omega r)/r (range-separated) attenuation, a compile-time switch so
    # that plain Coulomb builds carry no per-primitive attenuation arithmetic
    w('template <typename R, bool RS>')
    w('static __device__ __forceinline__ void quartet(')
    w('    const PrimPair* __restrict__ bra_prims, const int bra_np, const R ABx, const R ABy, const R ABz,')
    w('    const PrimPair* __restrict__ ket_prims, const int ket_np, const R CDx, const R CDy, const R CDz,')
    w('    const R* __restrict__ boys_tab, const R omega2, R* __restrict__ out)')
    w('{')
    w(f'    R acc[{nacc}];')
    w('#pragma unroll')
    w(f'    for (int i = 0; i < {nacc}; i++) acc[i] = R(0);')
    w('    for (int ik = 0; ik < ket_np; ik++) {')
    w('        const PrimPair kp = ket_prims[ik];')
    w('        const R eta = R(kp.zeta);')
    w('        for (int ib = 0; ib < bra_np; ib++) {')
    w('            const PrimPair bp = bra_prims[ib];')
    w('            const R zeta = R(bp.zeta);')
    w('            const R oo2z = R(0.5) * vrcp<R>(zeta);')
    w('            const R PAx = R(bp.pax), PAy = R(bp.pay), PAz = R(bp.paz);')
    w('            const R rs = vrsqrt<R>(zeta + eta);  // (zeta + eta)^(-1/2)')
    w('            const R inv_s = rs * rs;')
    w('            const R rho = zeta * eta * inv_s;')
    w('            const R PQx = R(bp.px - kp.px), PQy = R(bp.py - kp.py), PQz = R(bp.pz - kp.pz);')
    w('            const R T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);')
    w('            const R pref = R(bp.k * kp.k) * rs;')
    w(f'            R F[{L + 1}];')
    # erf(omega r)/r: F_m(T) -> s^(m+1/2) F_m(s T), s = omega^2 / (omega^2 + rho)
    w('            if constexpr (RS) {')
    w('                const R sfac = omega2 * vrcp<R>(omega2 + rho);')
    w(f'                boys_eval<{L}, R>(T * sfac, boys_tab, F);')
    w('                R fm = vsqrt<R>(sfac);')
    w('#pragma unroll')
    w(f'                for (int m = 0; m <= {L}; m++) {{')
    w('                    F[m] *= fm;')
    w('                    fm *= sfac;')
    w('                }')
    w('            } else {')
    w(f'                boys_eval<{L}, R>(T, boys_tab, F);')
    w('                (void)omega2;')
    w('            }')
    for m in range(L + 1):
        w(f'            const R b{m} = pref * F[{m}];')
    if L > 0:
        w('            const R re = zeta * inv_s, rz = eta * inv_s;')
        # W - P = -(eta / (zeta + eta)) (P - Q),  W - Q = (zeta / (zeta + eta)) (P - Q)
        w('            const R WPx = -rz * PQx, WPy = -rz * PQy, WPz = -rz * PQz;')
        w('            (void)WPx; (void)WPy; (void)WPz; (void)rz; (void)re; (void)PAx; (void)PAy; (void)PAz; (void)oo2z;')
        w('            const R WQx = re * PQx, WQy = re * PQy, WQz = re * PQz;')
        w('            (void)WQx; (void)WQy; (void)WQz;')
        w('            const R QCx = R(kp.pax), QCy = R(kp.pay), QCz = R(kp.paz);')
        w('            (void)QCx; (void)QCy; (void)QCz;')
        w('            const R oo2e = R(0.5) * vrcp<R>(eta), oo2ze = R(0.5) * inv_s;')
        w('            (void)oo2e; (void)oo2ze;')
    for ln in vrr.lines:
        w('            ' + ln)
    for (e, f), var in targets.items():
        w(f'            acc[{acc_index[(e, f)]}] += {var};')
    w('        }')
    w('    }')

    # ---- ket HRR + ket cart->sph, for every bra e component
    hrr = Code('h')
    ket_sph = {}  # (e, kc, kd) -> var
    for e in e_list:
        memo = {}

        def knode(c, d):
            key = (c, d)
            if key in memo:
                return memo[key]
            if sum(d) == 0:
                memo[key] = f'acc[{acc_index[(e, c)]}]'
                return memo[key]
            i = [k for k in range(3) if d[k] > 0][0]
            memo[key] = hrr.new(f'{knode(add(c, i, 1), add(d, i, -1))} + CD{AX[i]} * {knode(c, add(d, i, -1))}')
            return memo[key]

        cc, dd = cart_list(lc), cart_list(ld)
        for kc in range(nsph(lc)):
            for kd in range(nsph(ld)):
                terms = []
                for ic, fc in SPH[lc][kc]:
                    for idd, fd in SPH[ld][kd]:
                        coef = fc * fd
                        v = knode(cc[ic], dd[idd])
                        terms.append(v if coef == 1.0 else f'R({lit(coef)}) * {v}')
                ket_sph[(e, kc, kd)] = hrr.new(' + '.join(terms)) if len(terms) > 1 or '*' in terms[0] else terms[0]

    # ---- bra HRR + bra cart->sph, for every ket spherical pair
    for kc in range(nsph(lc)):
        for kd in range(nsph(ld)):
            memo = {}

            def bnode(a, b):
                key = (a, b)
                if key in memo:
                    return memo[key]
                if sum(b) == 0:
                    memo[key] = ket_sph[(a, kc, kd)]
                    return memo[key]
                i = [k for k in range(3) if b[k] > 0][0]
                memo[key] = hrr.new(f'{bnode(add(a, i, 1), add(b, i, -1))} + AB{AX[i]} * {bnode(a, add(b, i, -1))}')
                return memo[key]

            ca, cb = cart_list(la), cart_list(lb)
            for ka in range(nsph(la)):
                for kb in range(nsph(lb)):
                    terms = []
                    for ia, fa in SPH[la][ka]:
                        for ib, fb in SPH[lb][kb]:
                            coef = fa * fb
                            v = bnode(ca[ia], cb[ib])
                            terms.append(v if coef == 1.0 else f'R({lit(coef)}) * {v}')
                    idx = ((ka * nsph(lb) + kb) * nsph(lc) + kc) * nsph(ld) + kd
                    hrr.lines.append(f'out[{idx}] = {" + ".join(terms)};')
    w('    (void)ABx; (void)ABy; (void)ABz; (void)CDx; (void)CDy; (void)CDz;')
    for ln in hrr.lines:
        w('    ' + ln)
    w('}')
    w('};')
    vrr_flops = sum(count_flops(ln.split(' = ', 1)[1]) for ln in vrr.lines) + nacc
    hrr_flops = sum(count_flops(ln.split(' = ', 1)[1]) for ln in hrr.lines)
    return tag, '\n'.join(out), vrr_flops, hrr_flops


# classes whose straight-line recursion exceeds this many generated lines use
# the wavefront-cooperative McMurchie-Davidson kernel (md_kernels.h) instead
MD_LINES = 5500


def variant(la, lb, lc, ld, nlines):
    """Kernel of a class: 'md' (wavefront-cooperative, LDS) for the largest
    classes; 'ks' (ket-stationary, sticky exchange accumulators, few atomics)
    where the integrals are cheap and atomics dominate; 'bs'
    (bra-stationary) where the integrals need most of the register file."""
    if nlines > MD_LINES:
        return 'md'
    if lb == 0 and la + lb + lc + ld <= 4:
        return 'ks'
    return 'bs'


def tiled_waves(la, lb, lc, ld):
    """Waves per block of the tiled kernel: more waves share one set of LDS
    tiles where the integrals need few registers (the VGPR budget per lane is
    512 / waves-per-SIMD)."""
    L = la + lb + lc + ld
    return 16 if L <= 2 else (8 if L == 3 else 4)


def ksl_waves(la, lb, lc, ld, var):
    """Waves per block of the ket-stationary kernel with LDS accumulators, an alternative variant of
    the bra-stationary classes, or 0 (none): (na nc + na nd + nc nd) doubles per thread must fit in
    LDS beside the Boys table with 4 waves (one wave per SIMD, as the register use allows)."""
    na, nc, nd = 2 * la + 1, 2 * lc + 1, 2 * ld + 1
    nacc = na * nc + na * nd + nc * nd
    return 4 if var == 'bs' and nacc <= KSL_MAX_SLOTS else 0


# lane-private accumulator slots of fock_kernel_ksl with 4-wave blocks: 64 x 256 threads x 8 B =
# 128 KiB, plus the 25.5 KB Boys table, within the 160 KB of LDS
KSL_MAX_SLOTS = 64


# classes with la+lb+lc+ld >= MMD_LMIN also get the matrix-core MD Fock kernel as
# an alternative variant (the autotuner picks it per class)
MMD_LMIN = 4


def mmd_eligible(la, lb, lc, ld, var):
    return var == 'md' or la + lb + lc + ld >= MMD_LMIN


def tiled_eligible(la, lb, lc, ld):
    """Classes that also get the LDS-tiled Fock kernel (L <= 5; VLX_TILED_LMAX
    lowers the limit at run time)."""
    return la + lb + lc + ld <= 5


def pair_classes(lmax):
    return [(a, b) for a in range(lmax + 1) for b in range(a + 1)]


def class_pairs(lmax):
    pcs = pair_classes(lmax)
    return [(pcs[x], pcs[y]) for x in range(len(pcs)) for y in range(x + 1)]


def main(outdir, lmax):
    outdir = Path(outdir)
    outdir.mkdir(parents=True, exist_ok=True)
    pcs = pair_classes(lmax)
    written = []
    flops = {}
    variants = {}
    for (la, lb), (lc, ld) in class_pairs(lmax):
        tag, body, nv, nh = gen_class(la, lb, lc, ld)
        flops[(la, lb, lc, ld)] = (nv, nh)
        var = variant(la, lb, lc, ld, body.count('\n'))
        variants[(la, lb, lc, ld)] = var
        if var == 'md':
            src = ['// generated by gen_eri.py -- do not edit', '#include "../md_kernels.h"', '',
                   f'VLX_INSTANTIATE_MD_CLASS({la}, {lb}, {lc}, {ld})']
            if (la, lb) == (lc, ld):
                src.append(f'VLX_INSTANTIATE_MD_DIAG({la}, {lb})')
        else:
            kern = 'fock_kernel' if var == 'ks' else 'fock_kernel_bs'
            hdr = '../md_kernels.h' if mmd_eligible(la, lb, lc, ld, var) else '../eri_kernels.h'
            src = ['// generated by gen_eri.py -- do not edit', f'#include "{hdr}"', '', body, '',
                   f'VLX_INSTANTIATE_CLASS({la}, {lb}, {lc}, {ld}, {kern})',
                   f'VLX_INSTANTIATE_MULTI({la}, {lb}, {lc}, {ld})']
            if ksl_waves(la, lb, lc, ld, var):
                src.append(f'VLX_INSTANTIATE_KSL({la}, {lb}, {lc}, {ld}, {ksl_waves(la, lb, lc, ld, var)})')
            if mmd_eligible(la, lb, lc, ld, var):
                src.append(f'VLX_INSTANTIATE_MMD({la}, {lb}, {lc}, {ld})')
            if (la, lb) == (lc, ld):
                src.append(f'VLX_INSTANTIATE_DIAG({la}, {lb})')
            if tiled_eligible(la, lb, lc, ld):
                src.append(f'VLX_INSTANTIATE_TILED({la}, {lb}, {lc}, {ld}, {tiled_waves(la, lb, lc, ld)})')
        path = outdir / f'eri_{tag}.hip'
        text = '\n'.join(src) + '\n'
        if not path.exists() or path.read_text() != text:
            path.write_text(text)
        written.append((la, lb, lc, ld))
    # dispatch table
    d = ['// generated by gen_eri.py -- do not edit', f'#define VLX_GEN_LMAX {lmax}']
    for (la, lb, lc, ld) in written:
        d.append(f'VLX_DECLARE_CLASS({la}, {lb}, {lc}, {ld})')
        if variants[(la, lb, lc, ld)] != 'md':
            d.append(f'VLX_DECLARE_MULTI({la}, {lb}, {lc}, {ld})')
        if ksl_waves(la, lb, lc, ld, variants[(la, lb, lc, ld)]):
            d.append(f'VLX_DECLARE_KSL({la}, {lb}, {lc}, {ld})')
        if tiled_eligible(la, lb, lc, ld) and variants[(la, lb, lc, ld)] != 'md':
            d.append(f'VLX_DECLARE_TILED({la}, {lb}, {lc}, {ld})')
        if mmd_eligible(la, lb, lc, ld, variants[(la, lb, lc, ld)]) and variants[(la, lb, lc, ld)] != 'md':
            d.append(f'VLX_DECLARE_MMD({la}, {lb}, {lc}, {ld})')
        if (la, lb) == (lc, ld):
            d.append(f'VLX_DECLARE_DIAG({la}, {lb})')
    d.append('static const ClassLaunchers g_class_table[] = {')
    for (la, lb, lc, ld) in written:
        x = pcs.index((la, lb))
        y = pcs.index((lc, ld))
        nv, nh = flops[(la, lb, lc, ld)]
        ks = {'bs': 'VLX_BRA_STATIONARY', 'ks': 'VLX_KET_STATIONARY', 'md': 'VLX_WAVE_MD'}[variants[(la, lb, lc, ld)]]
        has_tiled = tiled_eligible(la, lb, lc, ld) and variants[(la, lb, lc, ld)] != 'md'
        tl = f'&launch_tiled_{la}{lb}{lc}{ld}' if has_tiled else 'nullptr'
        tl32 = f'&launch_tiled32_{la}{lb}{lc}{ld}' if has_tiled else 'nullptr'
        var = variants[(la, lb, lc, ld)]
        mm = (f'&launch_fock_{la}{lb}{lc}{ld}' if var == 'md' else
              f'&launch_mmd_{la}{lb}{lc}{ld}' if mmd_eligible(la, lb, lc, ld, var) else 'nullptr')
        # multi-density forms: the matrix-core MD kernel loops over the densities itself
        fm = f'&launch_fock_{la}{lb}{lc}{ld}' if var == 'md' else f'&launch_fock_multi_{la}{lb}{lc}{ld}'
        tlm = f'&launch_tiled_multi_{la}{lb}{lc}{ld}' if has_tiled else 'nullptr'
        ksl = f'&launch_ksl_{la}{lb}{lc}{ld}' if ksl_waves(la, lb, lc, ld, var) else 'nullptr'
        tld = f'&launch_tiled_dlds_{la}{lb}{lc}{ld}' if has_tiled else 'nullptr'
        d.append(f'    {{{x}, {y}, {la}, {lb}, {lc}, {ld}, {nv}, {nh}, {ks}, &launch_fock_{la}{lb}{lc}{ld}, '
                 f'&launch_eri_{la}{lb}{lc}{ld}, {tl}, {tl32}, {mm}, {fm}, {tlm}, {ksl}, {tld}}},')
    d.append('};')
    d.append('static const DiagLauncher g_diag_table[] = {')
    for (la, lb) in pcs:
        d.append(f'    {{{pcs.index((la, lb))}, {la}, {lb}, &launch_diag_{la}{lb}}},')
    d.append('};')
    text = '\n'.join(d) + '\n'
    path = outdir / 'dispatch.inc'
    if not path.exists() or path.read_text() != text:
        path.write_text(text)
    print(f'generated {len(written)} classes up to L={lmax} in {outdir}')


if __name__ == '__main__':
    main(sys.argv[1], int(sys.argv[2]))
