"""Reference CPU harness wrapper: test infrastructure only.

Runs oracle/_ref/vlx_ref_fock, the reference's own C++ Fock path built by
oracle/ref.mk (CT4CScreener::partition + CFockDriver::compute,
src/t4c_func/FockDriver.cpp:120-171), on a molecule, a packaged basis set and
a density.  The basis goes to the harness as the published exponents and raw
contraction coefficients, so the reference normalises them itself
(src/orbdata/BasisFunction.cpp:129-215).  Only tests/ and bench.py's CPU
baseline use it.
"""
import json
import os
import struct
import subprocess
import tempfile
from pathlib import Path

import numpy as np

REF_BIN = Path(__file__).resolve().parent / '_ref' / 'vlx_ref_fock'
# the reference-side drop-in (veloxchem_amd/refbind/GpuFockDriver.cpp) beside CFockDriver::compute
BIND_BIN = Path(__file__).resolve().parent / '_ref' / 'vlx_ref_gpu_fock'


def available():
    return REF_BIN.exists()


def binding_available():
    return BIND_BIN.exists()


def _input(molecule, basis_name, density, label, exchange_factor, omega, ithreshold, repeats):
    from veloxchem_amd.molecularbasis import _ALIASES, _basis_library
    lib = _basis_library()[_ALIASES[basis_name.upper()]]
    ids = list(molecule.get_element_ids())
    xyz = np.asarray(molecule.get_coordinates_in_bohr(), dtype=np.float64)
    buf = bytearray(struct.pack('<i', len(ids)))
    for z, r in zip(ids, xyz):
        buf += struct.pack('<i3d', int(z), float(r[0]), float(r[1]), float(r[2]))
    for lab in molecule.get_labels():
        shells = lib[lab.upper()]
        buf += struct.pack('<i', len(shells))
        for sh in shells:
            n = len(sh['exps'])
            buf += struct.pack('<ii', int(sh['l']), n)
            buf += struct.pack(f'<{n}d', *sh['exps']) + struct.pack(f'<{n}d', *sh['coefs'])
    d = np.ascontiguousarray(density, dtype='<f8')
    nao = d.shape[0]
    buf += struct.pack('<i', nao) + d.tobytes()
    lb = label.encode()
    buf += struct.pack('<i', len(lb)) + lb + struct.pack('<ddii', exchange_factor, omega, ithreshold, repeats)
    return bytes(buf), nao


def fock(molecule, basis_name, density, label='2jk', exchange_factor=1.0, omega=0.0, ithreshold=12, repeats=1,
         nthreads=None):
    """Reference G for a symmetric density; returns (G, timing dict)."""
    buf, nao = _input(molecule, basis_name, density, label, exchange_factor, omega, ithreshold, repeats)
    env = dict(os.environ)
    if nthreads:
        env['OMP_NUM_THREADS'] = str(nthreads)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = Path(td) / 'in.bin', Path(td) / 'out.bin'
        fin.write_bytes(buf)
        res = subprocess.run([str(REF_BIN), str(fin), str(fout)], capture_output=True, text=True, env=env)
        if res.returncode != 0:
            raise RuntimeError(f'vlx_ref_fock failed: {res.stderr.strip()}')
        g = np.fromfile(fout, dtype='<f8').reshape(nao, nao)
    return g, json.loads(res.stdout.strip().splitlines()[-1])


def binding_fock(molecule, basis_name, density, label='2jk', exchange_factor=1.0, omega=0.0, ithreshold=12,
                 general=False, label2='k'):
    """CFockDriver::compute and the drop-in CGpuFockDriver::compute on the same reference objects
    (CT4CScreener, CMatrix of type symmetric or general); returns (G_reference, G_gpu, summary dict with
    the CMatrices two-density comparison, labels {label, label2})."""
    buf, nao = _input(molecule, basis_name, density, label, exchange_factor, omega, ithreshold, 1)
    env = dict(os.environ)
    env['VLX_HARNESS_GENERAL'] = '1' if general else '0'
    env['VLX_HARNESS_LABEL2'] = label2
    with tempfile.TemporaryDirectory() as td:
        fin, fout = Path(td) / 'in.bin', Path(td) / 'out.bin'
        fin.write_bytes(buf)
        res = subprocess.run([str(BIND_BIN), str(fin), str(fout)], capture_output=True, text=True, env=env)
        if res.returncode != 0:
            raise RuntimeError(f'vlx_ref_gpu_fock failed: {res.stderr.strip()}')
        both = np.fromfile(fout, dtype='<f8').reshape(2, nao, nao)
    return both[0], both[1], json.loads(res.stdout.strip().splitlines()[-1])
