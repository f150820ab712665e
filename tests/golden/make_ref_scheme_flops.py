"""Writes ref_scheme_flops.json: the reference scheme's FLOPs per angular class.

Reads the reference's generated recursion sources as text (nothing is
imported or run) and counts, for each of the 100 classes (la lb|lc ld) with
l <= 3:

  V = FLOPs per primitive quartet of the vertical recursion: the binary
      '*', '/', '+', '-' of every statement in the SIMD loop bodies of each
      comp_prim_electron_repulsion_* function, summed over the calls that the
      class driver makes inside its primitive loop
      (src/t4c_electron_repulsion/ElectronRepulsionRec<ABCD>.hpp, calls into
      ElectronRepulsionPrimRec*.cpp);
  C = FLOPs per contracted quartet of the horizontal recursions and the
      Cartesian-to-spherical transforms: the SIMD loop bodies of the
      comp_ket_hrr_* / comp_bra_hrr_* functions
      (ElectronRepulsionContrRec*.cpp) times their component loops, plus two
      FLOPs (multiply, add) per transformation term of ket_transform and
      bra_transform (src/t4c_func/T4CUtils.hpp:405-518) with the term counts
      of spher_mom::transformation_factors
      (src/general/SphericalMomentum.hpp).

The class's algorithmic FLOPs of a build are then, per screened contracted
quartet, nprim_bra nprim_ket (V + 53 + 4 L) + C + w n_sph with w = 12 for
2jk.  Usage: python make_ref_scheme_flops.py [/root/reference]
"""
import json
import re
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / 'ref_scheme_flops.json'
LETTERS = 'SPDFGHIKLM'


def ncart(l):
    return (l + 1) * (l + 2) // 2


def nsph(l):
    return 2 * l + 1


def statement_flops(stmt):
    """Binary arithmetic operators of one C++ statement (unary minus excluded)."""
    rhs = stmt.split('=', 1)[1] if '=' in stmt else stmt
    if re.search(r'[^=!<>+\-*/]=', stmt) is None:
        return 0
    # strip array subscripts and numeric literals' exponents
    rhs = re.sub(r'\[[^\]]*\]', '', rhs)
    rhs = re.sub(r'(\d)[eE][+-](\d)', r'\1e\2', rhs)
    count = 0
    prev = '('  # start of expression: a leading '-' is unary
    for ch in rhs:
        if ch.isspace():
            continue
        if ch in '*/':
            count += 1
        elif ch in '+-':
            if prev not in '(*/+-=,':
                count += 1
        prev = ch
    return count


def loop_bodies(text, loop_re):
    """Bodies of the loops that match loop_re (brace matched)."""
    bodies = []
    for m in re.finditer(loop_re, text):
        i = text.index('{', m.end())
        depth, j = 0, i
        while True:
            if text[j] == '{':
                depth += 1
            elif text[j] == '}':
                depth -= 1
                if depth == 0:
                    break
            j += 1
        bodies.append(text[i + 1:j])
    return bodies


def body_flops(body):
    return sum(statement_flops(s) for s in body.split(';') if s.strip())


def function_flops(src_dir, kind, name):
    """SIMD-loop FLOPs of one recursion function, per element."""
    fname = {'prim': 'ElectronRepulsionPrimRec', 'contr': 'ElectronRepulsionContrRec'}[kind] + name.upper() + '.cpp'
    text = (src_dir / fname).read_text()
    loop = r'for \(size_t i = 0; i < nelems; i\+\+\)' if kind == 'prim' else r'for \(size_t k = 0; k < nelems; k\+\+\)'
    return sum(body_flops(b) for b in loop_bodies(text, loop))


def transform_terms(ref_root):
    """Number of Cartesian terms of each spherical component, per l (SphericalMomentum.hpp)."""
    text = (ref_root / 'src' / 'general' / 'SphericalMomentum.hpp').read_text()
    terms = {}
    # one 'if constexpr (N == l)' block per angular momentum; inside it, one
    # 'if (component == k) return {{index, factor}, ...};' per spherical component
    blocks = re.split(r'if constexpr \(N == (\d+)\)', text)
    for k in range(1, len(blocks), 2):
        l = int(blocks[k])
        comps = {}
        for br in re.finditer(r'if \(component == (\d+)\)\s*return\s*\{(.*?)\};', blocks[k + 1], re.S):
            comps[int(br.group(1))] = len(re.findall(r'\{\s*\d+\s*,', br.group(2)))
        if comps:
            terms[l] = [comps[i] for i in range(nsph(l))]
    return terms


def class_counts(src_dir, cls, fcache, tterms):
    text = (src_dir / f'ElectronRepulsionRec{cls}.hpp').read_text()
    # primitive loop: calls between 'for (int k = 0; k < bra_npgtos; k++)' and its end
    prim_body = loop_bodies(text, r'for \(int k = 0; k < bra_npgtos; k\+\+\)')[0]
    v = 0
    for m in re.finditer(r'comp_prim_electron_repulsion_(\w+)\(', prim_body):
        key = ('prim', m.group(1))
        if key not in fcache:
            fcache[key] = function_flops(src_dir, 'prim', m.group(1))
        v += fcache[key]
    rest = text[text.index(prim_body) + len(prim_body):]
    c = 0
    for m in re.finditer(r'comp_(ket|bra)_hrr_electron_repulsion_(\w+)\(([^;]*)\);', rest):
        side, name, args = m.group(1), m.group(2), [a.strip() for a in m.group(3).split(',')]
        key = ('contr', name)
        if key not in fcache:
            fcache[key] = function_flops(src_dir, 'contr', name)
        x, y = int(args[-2]), int(args[-1])
        outer = ncart(x) * ncart(y) if side == 'ket' else nsph(x) * nsph(y)
        c += fcache[key] * outer
    for m in re.finditer(r'(ket|bra)_transform<(\d+), (\d+)>\(([^;]*)\);', rest):
        side, n1, n2 = m.group(1), int(m.group(2)), int(m.group(3))
        args = [a.strip() for a in m.group(4).split(',')]
        x, y = int(args[-2]), int(args[-1])
        nterm = sum(tterms[n1]) * sum(tterms[n2])
        outer = ncart(x) * ncart(y) if side == 'ket' else nsph(x) * nsph(y)
        c += 2 * nterm * outer
    return v, c


def main(ref_root):
    ref_root = Path(ref_root)
    src_dir = ref_root / 'src' / 't4c_electron_repulsion'
    tterms = transform_terms(ref_root)
    fcache = {}
    table = {}
    for la in range(4):
        for lb in range(4):
            for lc in range(4):
                for ld in range(4):
                    cls = LETTERS[la] + LETTERS[lb] + LETTERS[lc] + LETTERS[ld]
                    if not (src_dir / f'ElectronRepulsionRec{cls}.hpp').exists():
                        continue
                    v, c = class_counts(src_dir, cls, fcache, tterms)
                    table[f'{la}{lb}{lc}{ld}'] = {'V': v, 'C': c}
    doc = {
        'description': 'Reference-scheme FLOPs per angular class (la lb|lc ld): V per primitive quartet '
                       '(vertical recursion), C per contracted quartet (horizontal recursions and '
                       'Cartesian-to-spherical transforms), counted from the reference sources '
                       'src/t4c_electron_repulsion/ElectronRepulsion{Rec,PrimRec,ContrRec}*.{hpp,cpp} '
                       'and src/t4c_func/T4CUtils.hpp:405-518 by make_ref_scheme_flops.py',
        'transform_terms': {str(k): v for k, v in sorted(tterms.items())},
        'classes': table,
    }
    OUT.write_text(json.dumps(doc, indent=1, sort_keys=True) + '\n')
    print(f'{len(table)} classes -> {OUT}')


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
