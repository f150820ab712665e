"""Re-encode the basis-set data this package ships as one JSON file.

The published Gaussian basis sets (STO-3G, def2-SVP, def2-SVPD, def2-TZVP, def2-QZVP and
the minimal AO-START-GUESS set used by the SAD guess) are read from the plain
text files of a VeloxChem checkout (format: ``@BASIS_SET name``, then per element
``@ATOMBASIS X`` blocks of shells ``L nprim 1`` followed by ``exponent coeff``
lines, closed by ``@END``; parsed the same way as
src/pymodule/molecularbasis.py:112-169) and written, for elements H..Ne, to
``basis_sets.json`` beside this script.  The JSON keeps the raw exponents and
contraction coefficients; normalisation happens at load time.

Usage: python make_basis_sets.py /path/to/veloxchem/basis
"""

import json
import sys
from pathlib import Path

ELEMENTS = ['H', 'HE', 'LI', 'BE', 'B', 'C', 'N', 'O', 'F', 'NE']
BASIS_FILES = ['STO-3G', 'DEF2-SVP', 'DEF2-SVPD', 'DEF2-TZVP', 'DEF2-QZVP', 'AO-START-GUESS']
ANGULAR = {'S': 0, 'P': 1, 'D': 2, 'F': 3, 'G': 4, 'H': 5, 'I': 6}


def parse_basis_file(path):
    lines = [ln.strip() for ln in Path(path).read_text().splitlines()]
    lines = [ln for ln in lines if ln and not ln.startswith('!')]
    out = {}
    elem = None
    shells = []
    i = 0
    while i < len(lines):
        ln = lines[i]
        if ln.startswith('@ATOMBASIS'):
            elem = ln.split()[1].upper()
            shells = []
        elif ln.startswith('@END'):
            if elem is not None:
                out[elem] = shells
            elem = None
        elif elem is not None:
            tok = ln.split()
            if tok[0] == 'ECP':
                raise ValueError('ECP shells are not supported')
            lang = ANGULAR[tok[0].upper()]
            nprim = int(tok[1])
            if len(tok) == 3 and int(tok[2]) != 1:
                raise ValueError('general contractions are not supported')
            exps, coefs = [], []
            for k in range(nprim):
                e, c = lines[i + 1 + k].split()
                exps.append(float(e))
                coefs.append(float(c))
            shells.append({'l': lang, 'exps': exps, 'coefs': coefs})
            i += nprim
        i += 1
    return out


def main(basis_dir):
    data = {}
    for name in BASIS_FILES:
        parsed = parse_basis_file(Path(basis_dir) / name)
        data[name] = {e: parsed[e] for e in ELEMENTS if e in parsed}
    dest = Path(__file__).with_name('basis_sets.json')
    dest.write_text(json.dumps(data, indent=0, sort_keys=True))
    print('wrote', dest)


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference/basis')
