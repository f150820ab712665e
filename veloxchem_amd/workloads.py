"""Synthetic inputs of the Fock-build workloads.

* water: RHF water, O (0,0,0), H (0,+-0.7572,0.5865) Angstrom;
* water clusters: numpy Generator(PCG64(seed)); oxygen positions uniform in
  a cube (10 Angstrom for 32 molecules, 12.8 for 64), rejected when closer
  than 2.6 Angstrom to an accepted oxygen; orientation a uniform random unit
  quaternion; rigid water with O-H 0.9572 Angstrom and H-O-H 104.52 degrees;
* densities: deterministic symmetric pseudo-densities D = C C^T / n_occ-ish
  with C a seeded random n_ao x n_occ matrix (parity does not depend on D
  being an SCF density; the SCF-path densities come from the SCF loop).
"""

import math

import numpy as np

from .molecule import Molecule

WATER_XYZ = """O 0.000000 0.000000 0.000000
H 0.000000 0.757200 0.586500
H 0.000000 -0.757200 0.586500"""

OH = 0.9572
HOH = math.radians(104.52)


def water():
    return Molecule.read_str(WATER_XYZ, 'angstrom')


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def water_cluster(nmol, box, seed, min_oo=2.6):
    """Seeded random water cluster (Angstrom box edge ``box``)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    oxy = []
    tries = 0
    while len(oxy) < nmol:
        tries += 1
        if tries > 200000:
            raise RuntimeError('water_cluster: could not place molecules')
        p = rng.uniform(0.0, box, size=3)
        if all(np.linalg.norm(p - o) >= min_oo for o in oxy):
            oxy.append(p)
    h1 = np.array([0.0, OH * math.sin(HOH / 2), OH * math.cos(HOH / 2)])
    h2 = np.array([0.0, -OH * math.sin(HOH / 2), OH * math.cos(HOH / 2)])
    labels, coords = [], []
    for o in oxy:
        rot = _random_rotation(rng)
        labels += ['O', 'H', 'H']
        coords += [o, o + rot @ h1, o + rot @ h2]
    return Molecule(labels, np.array(coords), 'angstrom')


def pseudo_density(nao, nocc, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.normal(size=(nao, nocc)) / math.sqrt(nao)
    return np.ascontiguousarray(c @ c.T)
