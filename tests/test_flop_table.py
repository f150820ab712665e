"""The reference-scheme FLOP table (tests/golden/ref_scheme_flops.json), the
numerator of the roofline figures.

Pinned by: the Cartesian term counts of the real solid harmonics, fitted
numerically from scipy's spherical harmonics (independent of the reference's
table); closed forms for the s classes; the horizontal-recursion and
transform counts of an independent per-line count of the same sources; and,
where the reference sources are present, regeneration from them.
"""
import json
import sys
from pathlib import Path

import numpy as np
import pytest
import scipy.special as sp

GOLDEN = Path(__file__).resolve().parent / 'golden'
TABLE = json.loads((GOLDEN / 'ref_scheme_flops.json').read_text())
sys.path.insert(0, str(GOLDEN))
import make_ref_scheme_flops as gen  # noqa: E402


def _solid_harmonic_terms(l):
    """Non-zero Cartesian monomials x^a y^b z^c (a + b + c = l) of r^l Y_lm, m = -l..l, by a
    least-squares fit on random points."""
    rng = np.random.default_rng(1)
    pts = rng.normal(size=(200, 3))
    x, y, z = pts.T
    r = np.linalg.norm(pts, axis=1)
    polar, azim = np.arccos(z / r), np.arctan2(y, x)
    mons = [(a, b, l - a - b) for a in range(l, -1, -1) for b in range(l - a, -1, -1)]
    A = np.stack([x ** a * y ** b * z ** c for a, b, c in mons], axis=1)
    counts = []
    for m in range(-l, l + 1):
        Y = sp.sph_harm_y(l, abs(m), polar, azim)
        f = (np.imag(Y) if m < 0 else np.real(Y)) * r ** l
        coef, *_ = np.linalg.lstsq(A, f, rcond=None)
        counts.append(int(np.sum(np.abs(coef) > 1e-8 * np.max(np.abs(coef)))))
    return counts


@pytest.mark.parametrize('l', [0, 1, 2, 3])
def test_transform_terms_match_solid_harmonics(l):
    assert TABLE['transform_terms'][str(l)] == _solid_harmonic_terms(l)


def test_statement_counter():
    assert gen.statement_flops('g[i] = a[i] * b[i] + c[i] * d[i]') == 3
    assert gen.statement_flops('g[k] = -a[k] * b[k] + c[k]') == 2
    assert gen.statement_flops('const double f = 0.5 / (a_exp + b_exp)') == 2
    assert gen.statement_flops('x[i] = 1.0e-5 * y[i]') == 1
    assert gen.statement_flops('auto p = buf.data(idx + 3)') == 1
    assert gen.statement_flops('if (a < b) return') == 0


def test_s_classes_closed_form():
    c = TABLE['classes']
    # (ss|ss): one product per primitive quartet; two transforms of one term (2 FLOPs each)
    assert c['0000'] == {'V': 1, 'C': 4}
    # (ss|sp): two (ss|ss) orders and three components of 2 products + 1 sum; ket transform over 3
    # p terms and a bra transform of 3 spherical components
    assert c['0001'] == {'V': 11, 'C': 12}


def test_counts_agree_with_independent_line_count():
    """A per-line count of the same sources (it reads one line per statement and so misses the
    continuation lines of long statements) gave these; horizontal recursions and transforms have
    no continuation lines, the vertical recursions of s and p pairs neither."""
    c = TABLE['classes']
    for key, C in {'0101': 36, '1111': 648, '0202': 176, '2222': 16576, '3333': 158424, '1133': 13356}.items():
        assert c[key]['C'] == C, key
    assert c['0101']['V'] == 66
    for key, v_line in {'1111': 716, '0202': 563, '2222': 12240, '3333': 83806, '1133': 6080, '1212': 3361,
                        '1122': 2686, '2233': 27846, '2323': 32578, '0303': 2563}.items():
        assert v_line <= c[key]['V'] <= 1.4 * v_line, key


def test_table_covers_every_class():
    keys = set(TABLE['classes'])
    assert len(keys) == 100
    assert all(len(k) == 4 and max(int(x) for x in k) <= 3 for k in keys)
    # V grows with the total angular momentum inside each family
    c = TABLE['classes']
    assert c['0000']['V'] < c['0101']['V'] < c['1111']['V'] < c['2222']['V'] < c['3333']['V']


@pytest.mark.skipif(not Path('/root/reference/src/t4c_electron_repulsion').is_dir(),
                    reason='reference sources not present')
def test_regenerates_from_reference(tmp_path, monkeypatch):
    monkeypatch.setattr(gen, 'OUT', tmp_path / 'flops.json')
    gen.main('/root/reference')
    assert json.loads((tmp_path / 'flops.json').read_text()) == TABLE


def test_bench_orientation_maps_every_build_class():
    """Every class the build runs (la >= lb, lc >= ld, bra pair class >= ket pair class) maps onto a
    class of the table."""
    sys.path.insert(0, str(GOLDEN.parents[1]))
    import bench
    for la in range(4):
        for lb in range(la + 1):
            for lc in range(4):
                for ld in range(lc + 1):
                    if la * (la + 1) // 2 + lb >= lc * (lc + 1) // 2 + ld:
                        assert bench.ref_scheme_key((la, lb, lc, ld)) in TABLE['classes']
    assert bench.ref_scheme_key((2, 0, 1, 1)) == '0211'
    assert bench.ref_scheme_key((1, 0, 0, 0)) == '0001'
