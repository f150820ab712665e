"""Closed-shell Hartree-Fock SCF around the GPU Fock build.

Mirrors ScfRestrictedDriver of the reference (src/pymodule/scfdriver.py,
scfrestdriver.py, c2diis.py):

- one-electron S, T, V (host code of libvlxfock.so);
- Lowdin/canonical orthogonaliser X = U s^-1/2, dropping overlap eigenvalues
  at or below ovl_thresh (scfdriver.py:1283-1300);
- SAD guess from the AO-START-GUESS minimal basis (sadguessdriver.py:841-971);
- 'L2_DIIS' (default): up to 5 iterations to gradient 1e-3 in the valence
  basis, then the valence MOs are inserted into the full basis as the
  starting density (scfdriver.py:788-815, 923-990); 'DIIS': SAD in the full
  basis (:777-786);
- per iteration (scfdriver.py:1347-1540): G = FockDriver.compute(screener, D,
  '2jk', 1.0, 0.0, thresh) on the GPU with the gradient-dependent threshold
  of _get_dyn_threshold (:2410-2435); E = sum D*G + 2 sum D*(T + V) + E_nuc
  (:2115-2171); F = G + T + V; gradient 2 ||X^T (FDS - SDF) X||_F
  (scfrestdriver.py:70-107); convergence when the gradient is below
  conv_thresh from the second iteration on (:2437-2470); C2-DIIS (Sellers)
  over at most max_err_vecs Fock matrices (c2diis.py:45-243,
  scfrestdriver.py:133-200); MOs from eigh(X^T F X), density C_occ C_occ^T.

The dense algebra (N x N products, eigh) runs in numpy on the host; the
two-electron part is the only GPU work.
"""

import math
import time
from collections import deque

import numpy as np

from .fockdriver import FockDriver
from .matrix import make_matrix, mat_t
from .molecularbasis import MolecularBasis
from .oneeints import compute_one_electron_integrals
from .sadguess import sad_density
from .screener import T4CScreener


def c2diis_weights(errors):
    """C2-DIIS weights (H. Sellers, Int. J. Quantum Chem. 45, 31 (1993)).

    Eigenvectors of the B matrix B_ij = <e_i, e_j>, each rescaled to unit
    coefficient sum (skipped if the sum is below 1e-6), and the one whose
    combined error vector has the smallest norm is taken (c2diis.py:156-243).
    """
    n = len(errors)
    B = np.empty((n, n))
    for i in range(n):
        for j in range(i, n):
            B[i, j] = B[j, i] = np.vdot(errors[i], errors[j])
    _, vecs = np.linalg.eigh(B)
    sums = vecs.sum(axis=0)
    cands = [vecs[:, k] / sums[k] for k in range(n) if abs(sums[k]) > 1.0e-6]
    best, fmin = cands[0], 1.0e8
    for w in cands:
        e = sum(f * v for f, v in zip(w, errors))
        r = np.vdot(e, e)
        if fmin > r:
            fmin, best = r, w
    return best


class ScfRestrictedDriver:
    """Spin-restricted closed-shell Hartree-Fock with the GPU Fock build."""

    def __init__(self, comm=None, ostream=None):
        self.comm = comm
        self.ostream = ostream
        self.acc_type = 'L2_DIIS'
        self.max_err_vecs = 10
        self.max_iter = 50
        self.conv_thresh = 1.0e-6
        self.ovl_thresh = 1.0e-6
        self.diis_thresh = 1000.0
        self.eri_thresh = 1.0e-12
        self.fock_driver = None
        self.screener_factory = None
        self.history = []
        self.timings = {}
        self._scf_energy = 0.0
        self._is_converged = False
        self._num_iter = 0

    # --- reference-style accessors ---
    def get_scf_energy(self):
        return self._scf_energy

    @property
    def scf_energy(self):
        return self._scf_energy

    @property
    def is_converged(self):
        return self._is_converged

    @property
    def num_iter(self):
        return self._num_iter

    def _get_dyn_threshold(self, e_grad):
        # scfdriver.py:2410-2435, same floating-point steps
        if e_grad < 1.0e-6:
            return self.eri_thresh
        nteri = math.pow(10, math.floor(math.log10(e_grad)))
        nteri = 1.0e-10 * nteri
        if nteri > 1.0e-10:
            return 1.0e-10
        if nteri < self.eri_thresh:
            return self.eri_thresh
        return nteri

    def compute(self, molecule, ao_basis, min_basis=None):
        """Runs the SCF; returns a dict with 'scf_energy', 'C_alpha', 'E_alpha', 'D_alpha', 'F_alpha', 'S'."""
        if molecule.get_multiplicity() != 1 or molecule.number_of_electrons() % 2:
            raise ValueError('ScfRestrictedDriver: closed-shell molecules only')
        if min_basis is None:
            min_basis = MolecularBasis.read(molecule, 'AO-START-GUESS')
        if self.fock_driver is None:
            self.fock_driver = FockDriver(self.comm)
        t0 = time.perf_counter()
        self.history = []
        self.timings = {'fock': 0.0, 'one_electron': 0.0, 'linear_algebra': 0.0, 'screener': 0.0}
        acc = self.acc_type.upper()
        if acc == 'DIIS':
            den = sad_density(molecule, min_basis, ao_basis)
            res = self._comp_diis(molecule, ao_basis, den, self.conv_thresh, self.max_iter)
        elif acc == 'L2_DIIS':
            val_basis = ao_basis.reduce_to_valence_basis()
            den = sad_density(molecule, min_basis, val_basis)
            first = self._comp_diis(molecule, val_basis, den, 1.0e-3, 5)
            den = self._project_density(molecule, ao_basis, val_basis, first['C_alpha'])
            res = self._comp_diis(molecule, ao_basis, den, self.conv_thresh, self.max_iter)
        else:
            raise ValueError(f'ScfRestrictedDriver: unknown acc_type {self.acc_type!r}')
        self.timings['total'] = time.perf_counter() - t0
        res['timings'] = dict(self.timings)
        res['num_iter'] = self._num_iter
        res['is_converged'] = self._is_converged
        return res

    def _project_density(self, molecule, ao_basis, val_basis, c_val):
        """Valence MOs inserted into the full AO basis (scfdriver.py:923-990), aufbau density."""
        nao = ao_basis.get_dimensions_of_basis()
        nmo = c_val.shape[1]
        c = np.zeros((nao, nmo))
        r1 = r2 = 0
        for lang in range(val_basis.max_angular_momentum() + 1):
            for _ in range(2 * lang + 1):
                for a in range(molecule.number_of_atoms()):
                    n1 = val_basis.number_of_atom_basis_functions(a, lang)
                    n2 = ao_basis.number_of_atom_basis_functions(a, lang)
                    if n1 > 0:
                        c[r2:r2 + n2, :] = c_val[r1:r1 + n1, :]
                    r1 += n1
                    r2 += n2
        nocc = molecule.number_of_alpha_electrons()
        return c[:, :nocc] @ c[:, :nocc].T

    def _fock_2e(self, screener, den, thresh_int, basis):
        dm = make_matrix(basis, mat_t.symmetric)
        dm.set_values(den)
        t = time.perf_counter()
        g = self.fock_driver.compute(screener, dm, '2jk', 1.0, 0.0, thresh_int).to_numpy()
        self.timings['fock'] += time.perf_counter() - t
        return g

    def _comp_diis(self, molecule, basis, den, conv_thresh, max_iter):
        t = time.perf_counter()
        S, T, V = compute_one_electron_integrals(molecule, basis)
        self.timings['one_electron'] += time.perf_counter() - t
        hcore = T + V
        e_nuc = molecule.nuclear_repulsion_energy()
        t = time.perf_counter()
        w, U = np.linalg.eigh(S)
        keep = w > self.ovl_thresh
        X = U[:, keep] * (1.0 / np.sqrt(w[keep]))
        self.timings['linear_algebra'] += time.perf_counter() - t
        t = time.perf_counter()
        if self.screener_factory is None:
            screener = T4CScreener()
            screener.partition(basis, molecule, 'eri')
        else:
            screener = self.screener_factory(basis, molecule)
        self.timings['screener'] += time.perf_counter() - t
        nocc = molecule.number_of_alpha_electrons()
        focks, dens = deque(), deque()
        e_grad = None
        e_old = 0.0
        self._is_converged = False
        C = eps = F = None
        for it in range(max_iter + 1):
            self._num_iter = it
            thresh = self.eri_thresh if e_grad is None else self._get_dyn_threshold(e_grad)
            thresh_int = int(-math.log10(thresh))
            G = self._fock_2e(screener, den, thresh_int, basis)
            t = time.perf_counter()
            e_el = np.sum(den * G) + 2.0 * np.sum(den * hcore)
            F = G + hcore
            fds = F @ den @ S
            err = X.T @ (fds - fds.T) @ X
            e_grad = 2.0 * np.linalg.norm(err)
            e_scf = e_el + e_nuc
            self.history.append({'energy': e_scf, 'gradient_norm': e_grad, 'max_gradient': np.max(np.abs(err)),
                                 'diff_energy': e_scf - e_old, 'eri_threshold': thresh_int})
            e_old = e_scf
            self._scf_energy = e_scf
            if it > 0 and e_grad < conv_thresh:
                self._is_converged = True
                self.timings['linear_algebra'] += time.perf_counter() - t
                break
            if e_grad < self.diis_thresh:
                if len(focks) == self.max_err_vecs:
                    focks.popleft()
                    dens.popleft()
                focks.append(F.copy())
                dens.append(den.copy())
            if len(focks) > 1:
                errs = []
                for f, d in zip(focks, dens):
                    x = f @ d @ S
                    errs.append(X.T @ (x - x.T) @ X)
                wts = c2diis_weights(errs)
                f_eff = sum(wi * fi for wi, fi in zip(wts, focks))
            elif len(focks) == 1:
                f_eff = focks[0].copy()
            else:
                f_eff = F
            eps, vecs = np.linalg.eigh(X.T @ f_eff @ X)
            C = X @ vecs
            # drop MOs with coefficients above 1/sqrt(ovl_thresh) (scfdriver.py:2724-2747)
            big = np.max(np.abs(C), axis=0) > 1.0 / math.sqrt(self.ovl_thresh)
            if np.any(big):
                C, eps = C[:, ~big], eps[~big]
            den = C[:, :nocc] @ C[:, :nocc].T
            self.timings['linear_algebra'] += time.perf_counter() - t
        if C is None:
            eps, vecs = np.linalg.eigh(X.T @ F @ X)
            C = X @ vecs
        return {'scf_energy': self._scf_energy, 'C_alpha': C, 'E_alpha': eps, 'D_alpha': den, 'F_alpha': F,
                'S': S}
