"""Real solid harmonics in Cartesian monomials, in the reference's component order.

SPH[l][m] = [(cartesian index, coefficient)] for spherical component m = 0..2l
(order -l..l), with Cartesian monomials x^a y^b z^c indexed in the order
a = l..0, b = l-a..0 (xx, xy, xz, yy, yz, zz for l = 2).  The coefficients
are those of spher_mom::transformation_factors (src/general/SphericalMomentum.hpp:18-176):
they define the AO layout of every matrix, so they are restated, not chosen.
Used by the kernel generator (csrc/gen_eri.py) and by the learned guess
(rotations of density blocks).
"""
import math

_F3 = 2.0 * math.sqrt(3.0)
_G5 = math.sqrt(2.5)
_G15 = 2.0 * math.sqrt(15.0)
_G3 = math.sqrt(1.5)

SPH = {
    0: [[(0, 1.0)]],
    1: [[(1, 1.0)], [(2, 1.0)], [(0, 1.0)]],
    2: [[(1, _F3)], [(4, _F3)], [(0, -1.0), (3, -1.0), (5, 2.0)], [(2, _F3)],
        [(0, 0.5 * _F3), (3, -0.5 * _F3)]],
    3: [[(1, 3.0 * _G5), (6, -_G5)], [(4, _G15)], [(8, 4.0 * _G3), (1, -_G3), (6, -_G3)],
        [(9, 2.0), (2, -3.0), (7, -3.0)], [(5, 4.0 * _G3), (0, -_G3), (3, -_G3)],
        [(2, 0.5 * _G15), (7, -0.5 * _G15)], [(0, _G5), (3, -3.0 * _G5)]],
}


def cartesian_exponents(lang):
    """(a, b, c) of the Cartesian monomials of angular momentum lang, in index order."""
    return [(a, b, lang - a - b) for a in range(lang, -1, -1) for b in range(lang - a, -1, -1)]
