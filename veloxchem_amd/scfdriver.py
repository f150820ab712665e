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

The density, Fock and orthogonaliser stay on the GPU as float64 torch
tensors (torch is used as device memory and for the dense N x N products and
eigh); the two-electron part goes through vlx_fock_compute_device on the
same buffers.  DIIS error vectors are formed once per stored Fock matrix
(the reference recomputes them each iteration from the same F, D, S, X).
"""

import ctypes
import math
import time
from collections import deque
from concurrent.futures import Future, ThreadPoolExecutor

import numpy as np
import torch

from . import _native
from .molecularbasis import MolecularBasis
from .oneeints import compute_one_electron_integrals
from .sadguess import sad_density
from .screener import T4CScreener


def c2diis_weights(errors):
    """C2-DIIS weights (H. Sellers, Int. J. Quantum Chem. 45, 31 (1993)).

    Eigenvectors of the B matrix B_ij = <e_i, e_j>, each rescaled to unit
    coefficient sum (skipped if the sum is below 1e-6), and the one whose
    combined error vector has the smallest norm (w^T B w) is taken
    (c2diis.py:156-243).  errors: sequence of arrays or tensors, or an
    (n x n) B matrix given as ``bmat=``.
    """
    if isinstance(errors, dict):
        B = np.asarray(errors['bmat'], dtype=np.float64)
    else:
        n = len(errors)
        B = np.empty((n, n))
        for i in range(n):
            for j in range(i, n):
                B[i, j] = B[j, i] = float(np.vdot(np.asarray(errors[i]), np.asarray(errors[j])))
    n = B.shape[0]
    _, vecs = np.linalg.eigh(B)
    sums = vecs.sum(axis=0)
    cands = [vecs[:, k] / sums[k] for k in range(n) if abs(sums[k]) > 1.0e-6]
    best, fmin = cands[0], 1.0e8
    for w in cands:
        r = float(w @ B @ w)
        if fmin > r:
            fmin, best = r, w
    return best


def canonical_purification(fo, nocc, tol=1e-13, max_iter=200):
    """Occupied-space projector of a symmetric matrix (orthonormal basis) without diagonalisation.

    Trace-conserving canonical purification (A. H. R. Palser, D. E. Manolopoulos, Phys. Rev. B 58,
    12704 (1998)): start from a linear map of fo with trace nocc and spectrum in [0, 1] (Gershgorin
    bounds), then iterate the cubic that fixes the trace, two matrix products per step, until
    tr(D - D^2) ~ 0.  Returns None when it does not converge (no gap), so the caller falls back to
    eigh.
    """
    m = fo.shape[0]
    if nocc <= 0 or nocc >= m:
        return None
    diag = torch.diagonal(fo)
    rad = fo.abs().sum(dim=1) - diag.abs()
    bounds = torch.stack([(diag - rad).min(), (diag + rad).max(), torch.trace(fo) / m]).tolist()
    emin, emax, mu = bounds
    lam = min(nocc / (emax - mu), (m - nocc) / (mu - emin))
    eye = torch.eye(m, dtype=fo.dtype, device=fo.device)
    d = (lam / m) * (mu * eye - fo) + (nocc / m) * eye
    for _ in range(max_iter):
        d2 = d @ d
        d3 = d2 @ d
        t = torch.stack([torch.trace(d), torch.trace(d2), torch.trace(d3)]).tolist()
        denom = t[0] - t[1]
        if abs(denom) <= tol * m:
            return d
        c = (t[1] - t[2]) / denom
        if c >= 0.5:
            d = ((1.0 + c) * d2 - d3) / c
        else:
            d = ((1.0 - 2.0 * c) * d + (1.0 + c) * d2 - d3) / (1.0 - c)
    return None


class _DeviceFock:
    """G = 2J - K of a device density through vlx_fock_compute_device (same buffers, no host copies)."""

    def __init__(self, device):
        self.device = device
        self.lib = _native.lib()
        self.ctx = _native.context(device)

    def __call__(self, screener, D, thresh_int):
        G = torch.empty_like(D)
        torch.cuda.current_stream(D.device).synchronize()
        _native.check(self.lib.vlx_fock_compute_device(self.ctx, screener.handle, ctypes.c_void_p(D.data_ptr()),
                                                        b'2jk', 1.0, 0.0, int(thresh_int),
                                                        ctypes.c_void_p(G.data_ptr())))
        _native.check(self.lib.vlx_context_synchronize(self.ctx))
        return G


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
        # density from the effective Fock matrix: 'purification' (matrix products only, eigh once
        # per stage for the orbitals) or 'eigh' every iteration (as the reference)
        self.density_method = 'purification'
        # fock_2e(screener, D, thresh_int) -> G on D's device; default: the GPU build
        self.fock_2e = None
        # 'fp64', or 'mixed': FP32 recursions where Schwarz bound x max|D| < the smaller of
        # mixed_threshold and mixed_gradient_factor x the last gradient norm (the FP32 errors of a
        # build at threshold tau left a gradient floor of about 1.4 tau on a 2752-AO cluster)
        self.precision = 'fp64'
        self.mixed_threshold = 1.0e-8
        self.mixed_gradient_factor = 0.05
        # incremental (difference-density) Fock builds: G_n = G_{n-1} + G(D_n - D_{n-1}) with the
        # screening threshold tightened by max|D_n - D_{n-1}|; a full build every full_fock_every
        # iterations, when the ERI threshold changes, and to confirm convergence
        self.incremental_fock = False
        self.full_fock_every = 8
        # per-quartet error bound of an incremental build, relative to conv_thresh (see _comp_diis)
        self.incremental_drop_factor = 1.0e-6
        # density-weighted screening of the builds (a fast mode, no reference counterpart): quartets
        # whose Schwarz bound x the largest |D| (|dD| in incremental builds) of the blocks they read
        # is below min(density_screening, 1e-6 x the last gradient norm) are skipped (in incremental
        # builds also below incremental_drop_factor x conv_thresh, as their errors add up); 0 = off.
        # The skipped terms leave an error in G, so the threshold follows the gradient down and the
        # SCF keeps converging; the converged density is confirmed by one unscreened full build,
        # which gives the reported energy.
        self.density_screening = 0.0
        # deterministic reduction mode of the GPU builds (fixed-point integer accumulation): the same
        # bits on every run; FP64 with Schwarz selection only, so precision and density_screening
        # do not apply while it is on
        self.deterministic = False
        self.density_screening_gradient_factor = 1.0e-6
        self.screener_factory = None
        # starting density: None for the reference's guesses (SAD, L2-DIIS valence stage), or an
        # object with supports(molecule, basis) and density(molecule, basis, overlap) (the learned
        # guess, learnedguess.LearnedGuess); the full-basis DIIS then starts from its density, and
        # molecules it does not support fall back to SAD
        self.guess = None
        self.device = None
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
        """Runs the SCF; returns a dict with 'scf_energy', 'C_alpha', 'E_alpha', 'D_alpha', 'F_alpha', 'S' (numpy)."""
        if molecule.get_multiplicity() != 1 or molecule.number_of_electrons() % 2:
            raise ValueError('ScfRestrictedDriver: closed-shell molecules only')
        if min_basis is None:
            min_basis = MolecularBasis.read(molecule, 'AO-START-GUESS')
        if self.fock_2e is None or isinstance(self.fock_2e, _DeviceFock):
            dev = _native.current_device()
            self.fock_2e = _DeviceFock(dev)
            self.device = torch.device('cuda', dev)
            _native.set_precision(self.precision, self.mixed_threshold if self.precision == 'mixed' else 0.0, dev)
            _native.set_density_screening(self.density_screening, dev)
            _native.set_deterministic(self.deterministic, dev)
        elif self.device is None:
            self.device = torch.device('cpu')
        t0 = time.perf_counter()
        self.history = []
        self.timings = {'fock': 0.0, 'one_electron': 0.0, 'linear_algebra': 0.0, 'screener': 0.0, 'guess': 0.0}
        try:
            res = self._run(molecule, ao_basis, min_basis)
        finally:
            if isinstance(self.fock_2e, _DeviceFock):
                _native.set_precision('fp64', 0.0, self.fock_2e.device)
                _native.set_density_screening(0.0, self.fock_2e.device)
                _native.set_deterministic(False, self.fock_2e.device)
        self.timings['total'] = time.perf_counter() - t0
        res['timings'] = dict(self.timings)
        res['num_iter'] = self._num_iter
        res['is_converged'] = self._is_converged
        return res

    def _run(self, molecule, ao_basis, min_basis):
        acc = self.acc_type.upper()
        if self.guess is not None and self.guess.supports(molecule, ao_basis):
            t = time.perf_counter()
            one = Future()
            one.set_result(compute_one_electron_integrals(molecule, ao_basis))
            self.timings['one_electron'] += time.perf_counter() - t
            t = time.perf_counter()
            den = self.guess.density(molecule, ao_basis, one.result()[0])
            self.timings['guess'] += time.perf_counter() - t
            return self._comp_diis(molecule, ao_basis, den, self.conv_thresh, self.max_iter, one)
        if acc == 'DIIS':
            t = time.perf_counter()
            den = sad_density(molecule, min_basis, ao_basis)
            self.timings['guess'] += time.perf_counter() - t
            return self._comp_diis(molecule, ao_basis, den, self.conv_thresh, self.max_iter)
        if acc == 'L2_DIIS':
            val_basis = ao_basis.reduce_to_valence_basis()
            t = time.perf_counter()
            den = sad_density(molecule, min_basis, val_basis)
            self.timings['guess'] += time.perf_counter() - t
            # the one-electron integrals of both stages (host threads, the GIL released) are
            # computed while the screeners are built and the valence stage runs on the GPU
            with ThreadPoolExecutor(max_workers=2) as pool:
                one_val = pool.submit(compute_one_electron_integrals, molecule, val_basis)
                one_full = pool.submit(compute_one_electron_integrals, molecule, ao_basis)
                first = self._comp_diis(molecule, val_basis, den, 1.0e-3, 5, one_val)
                den = self._project_density(molecule, ao_basis, val_basis, first['C_alpha'])
                return self._comp_diis(molecule, ao_basis, den, self.conv_thresh, self.max_iter, one_full)
        raise ValueError(f'ScfRestrictedDriver: unknown acc_type {self.acc_type!r}')

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

    def _comp_diis(self, molecule, basis, den_np, conv_thresh, max_iter, one_e=None):
        dev = self.device
        t = time.perf_counter()
        if self.screener_factory is None:
            screener = T4CScreener()
            screener.partition(basis, molecule, 'eri')
        else:
            screener = self.screener_factory(basis, molecule)
        self.timings['screener'] += time.perf_counter() - t
        t = time.perf_counter()
        # one_e: a future of the integrals started earlier (only the wait is timed)
        S_np, T_np, V_np = one_e.result() if one_e is not None else compute_one_electron_integrals(molecule, basis)
        self.timings['one_electron'] += time.perf_counter() - t
        t = time.perf_counter()
        S = torch.from_numpy(S_np).to(dev)
        hcore = torch.from_numpy(T_np + V_np).to(dev)
        den = torch.from_numpy(np.ascontiguousarray(den_np)).to(dev)
        e_nuc = molecule.nuclear_repulsion_energy()
        w, U = torch.linalg.eigh(S)
        keep = w > self.ovl_thresh
        X = (U[:, keep] * (1.0 / torch.sqrt(w[keep]))).contiguous()
        self.timings['linear_algebra'] += time.perf_counter() - t
        nocc = molecule.number_of_alpha_electrons()
        focks, errs = deque(), deque()
        bmat = np.zeros((0, 0))
        e_grad = None
        e_old = 0.0
        self._is_converged = False
        C = eps = F = last_f_eff = None
        fmax = 1.0 / math.sqrt(self.ovl_thresh)
        g_prev = d_prev = None
        full_thresh, since_full, confirm = None, 0, False
        # set when the fast modes' error floor is found above conv_thresh (below): the rest of the
        # SCF runs full, unscreened FP64 builds (else it would alternate between drifting
        # incremental builds and failed confirmations)
        full_only = False
        best_grad, best_it = math.inf, 0
        for it in range(max_iter + 1):
            self._num_iter = it
            thresh = self.eri_thresh if e_grad is None else self._get_dyn_threshold(e_grad)
            thresh_int = int(-math.log10(thresh))
            t = time.perf_counter()
            incremental = (self.incremental_fock and g_prev is not None and not confirm and not full_only and
                           thresh_int == full_thresh and since_full < self.full_fock_every)
            if self.precision == 'mixed' and isinstance(self.fock_2e, _DeviceFock):
                # FP32 in full builds only: a difference-density build splits by max|dD|, so nearly all
                # of it would run in FP32, and those errors would add up over the incremental builds
                # (a gradient floor of 1.4e-6 on the 64-water def2-TZVP cluster)
                tau_m = 0.0 if (full_only or incremental) else self.mixed_threshold
                if e_grad is not None:
                    tau_m = min(tau_m, self.mixed_gradient_factor * e_grad)
                _native.set_precision('mixed' if tau_m > 0.0 else 'fp64', tau_m, self.fock_2e.device)
            if self.density_screening > 0.0 and isinstance(self.fock_2e, _DeviceFock) and not confirm:
                tau_d = 0.0 if full_only else self.density_screening
                if e_grad is not None:
                    tau_d = min(tau_d, self.density_screening_gradient_factor * e_grad)
                if incremental:
                    # skipped difference-density terms add up over the incremental builds like the
                    # threshold shift's (below): the same per-quartet budget
                    tau_d = min(tau_d, self.incremental_drop_factor * conv_thresh)
                _native.set_density_screening(tau_d, self.fock_2e.device)
            build_thresh = thresh_int
            if incremental:
                # a quartet of bin sum s is bounded by 10^(1 - s) max|dD| once s > build_thresh.  The
                # errors of skipped quartets add up over the incremental builds (unlike a full build's,
                # which are the same every iteration), and their sum in the gradient norm measured
                # about 1e5 times the per-quartet bound (64-water def2-TZVP), so the bound is kept at
                # incremental_drop_factor x conv_thresh: s <= 1 + log10(max|dD| / (factor conv)), at
                # most the full build's thresh_int
                dd = (den - d_prev).contiguous()
                dmax = float(dd.abs().max())
                if dmax > 0.0:
                    need = math.ceil(1.0 + math.log10(dmax / (self.incremental_drop_factor * conv_thresh)))
                    build_thresh = min(thresh_int, max(need, 0))
                else:
                    build_thresh = -1
                G = g_prev + self.fock_2e(screener, dd, build_thresh) if build_thresh >= 0 else g_prev.clone()
                since_full += 1
            elif confirm and self.density_screening > 0.0 and isinstance(self.fock_2e, _DeviceFock):
                # the confirming build of a density-screened SCF runs unscreened
                _native.set_density_screening(0.0, self.fock_2e.device)
                try:
                    G = self.fock_2e(screener, den, thresh_int)
                finally:
                    _native.set_density_screening(self.density_screening, self.fock_2e.device)
                full_thresh, since_full = thresh_int, 0
            else:
                G = self.fock_2e(screener, den, thresh_int)
                full_thresh, since_full = thresh_int, 0
            g_prev, d_prev = G, den
            self.timings['fock'] += time.perf_counter() - t
            t = time.perf_counter()
            e_el = float(torch.sum(den * G) + 2.0 * torch.sum(den * hcore))
            F = G + hcore
            fds = F @ den @ S
            err = X.T @ (fds - fds.T) @ X
            e_grad = 2.0 * float(torch.linalg.norm(err))
            e_scf = e_el + e_nuc
            self.history.append({'energy': e_scf, 'gradient_norm': e_grad, 'max_gradient': float(err.abs().max()),
                                 'diff_energy': e_scf - e_old, 'eri_threshold': thresh_int,
                                 'incremental': incremental, 'build_threshold': build_thresh})
            e_old = e_scf
            self._scf_energy = e_scf
            if it > 0 and e_grad < conv_thresh:
                self.timings['linear_algebra'] += time.perf_counter() - t
                if incremental or (self.density_screening > 0.0 and not confirm and not full_only):
                    confirm = True  # the same density once more with a full (unscreened) build
                    continue
                self._is_converged = True
                break
            # a confirming full build that found the SCF not converged after all, or a gradient that
            # has not halved in four iterations while incremental builds, density screening or FP32
            # recursions ran, means that their error floor is above conv_thresh: the rest of the SCF
            # runs full, unscreened FP64 builds, and the DIIS history, whose Fock matrices carry
            # those errors, starts afresh
            fast = self.incremental_fock or self.density_screening > 0.0 or self.precision == 'mixed'
            if e_grad < 0.5 * best_grad:
                best_grad, best_it = e_grad, it
            elif it - best_it >= 4 and fast and not full_only:
                confirm = True
            if confirm and not full_only:
                full_only = True
                focks.clear()
                errs.clear()
                bmat = np.zeros((0, 0))
            confirm = False
            if e_grad < self.diis_thresh:
                if len(focks) == self.max_err_vecs:
                    focks.popleft()
                    errs.popleft()
                    bmat = bmat[1:, 1:]
                # the error vector of a stored (F, D) pair is this iteration's err; the B matrix
                # grows by one row of dot products (elementwise products and sums: a k x N^2 GEMM
                # with k <= 10 is a pathological BLAS shape)
                dots = torch.stack([torch.sum(e * err) for e in errs] + [torch.sum(err * err)]).cpu().numpy()
                k = len(errs)
                grown = np.zeros((k + 1, k + 1))
                grown[:k, :k] = bmat
                grown[k, :] = dots
                grown[:, k] = dots
                bmat = grown
                focks.append(F)
                errs.append(err)
            if len(focks) > 1:
                wts = c2diis_weights({'bmat': bmat})
                f_eff = torch.zeros_like(F)
                for wi, fi in zip(wts, focks):
                    f_eff.add_(fi, alpha=float(wi))
            elif len(focks) == 1:
                f_eff = focks[0]
            else:
                f_eff = F
            fo = X.T @ f_eff @ X
            dp = canonical_purification(fo, nocc) if self.density_method == 'purification' else None
            if dp is not None:
                den = (X @ dp @ X.T).contiguous()
                C = None
            else:
                eps, vecs = torch.linalg.eigh(fo)
                C = X @ vecs
                # drop MOs with coefficients above 1/sqrt(ovl_thresh) (scfdriver.py:2724-2747)
                big = C.abs().amax(dim=0) > fmax
                if bool(big.any()):
                    C, eps = C[:, ~big], eps[~big]
                den = (C[:, :nocc] @ C[:, :nocc].T).contiguous()
            last_f_eff = f_eff
            self.timings['linear_algebra'] += time.perf_counter() - t
        if C is None:
            # orbitals of the last effective Fock matrix (the one that produced the density)
            t = time.perf_counter()
            eps, vecs = torch.linalg.eigh(X.T @ (F if last_f_eff is None else last_f_eff) @ X)
            C = X @ vecs
            self.timings['linear_algebra'] += time.perf_counter() - t
        return {'scf_energy': self._scf_energy, 'C_alpha': C.cpu().numpy(), 'E_alpha': eps.cpu().numpy(),
                'D_alpha': den.cpu().numpy(), 'F_alpha': F.cpu().numpy(), 'S': S_np}
