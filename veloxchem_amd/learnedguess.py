"""Learned SCF starting density, predicted atom-pair block by atom-pair block.

No reference counterpart: the reference starts every SCF from the SAD guess
(src/pymodule/sadguessdriver.py:841-971; scfdriver.py:777-815), which stays
the fallback here for molecules or bases the network was not trained on.

Design
------
* One sample per atom pair (A, B) within RCUT (on-site pairs A = A
  included), ordered so that A has the larger nuclear charge (the lower
  index on ties).  Its target is the density block D[AOs of A, AOs of B] of a
  converged RHF density, expressed in a local frame of the pair: e_z along
  A -> B (on site: towards the nearest neighbour), e_x towards the nearest
  other atom (on site: the second nearest), e_y = e_z x e_x.  In that frame
  the block no longer depends on the molecule's orientation, so
  rotation-invariant inputs can predict it; the block is rotated back with
  the real-solid-harmonic rotation matrices of each shell.
* Inputs (70): element-pair one-hot over H, C, N, O, F (25) and an on-site
  flag; 32 Gaussian radial features of |AB| on [0, RCUT]; 8 neighbour
  counts within 3.5 Angstrom (H, C, N, O/F around A and around B); the
  local-frame geometry (distance and cosine of the frame-defining atoms).
* Output: the local block flattened as 14 x 14 (def2-SVP: 14 functions on
  C..F, 5 on H), rows and columns beyond the atoms' functions masked.
* Network: 70 -> 256 -> 256 -> 196, SiLU, BF16 on the matrix cores with
  FP32 accumulation, Adam (densitynet.DensityNet, the library's kernels).
* Training data: converged densities of seeded water clusters from this
  package's own GPU SCF.
* Use: ScfRestrictedDriver(guess=LearnedGuess) starts the full-basis DIIS
  from the predicted density, scaled so that tr(D S) = N_alpha; the SCF
  itself, its convergence test and its energy are unchanged.
"""
import math
import time

import numpy as np

from .molecularbasis import MolecularBasis
from .sphericals import SPH, cartesian_exponents

ELEMENTS = (1, 6, 7, 8, 9)          # H, C, N, O, F
BASIS = 'DEF2-SVP'
MAX_BF = 14                          # functions per atom (C..F in def2-SVP)
RCUT = 6.0                           # Angstrom; pairs farther apart are predicted zero
NRADIAL = 32
NEIGH_R = 3.5
N_IN = len(ELEMENTS) ** 2 + 1 + NRADIAL + 8 + 4
N_OUT = MAX_BF * MAX_BF
BOHR = 0.529177210903                # Angstrom per bohr

_rng_pts = np.random.Generator(np.random.PCG64(12345)).normal(size=(24, 3))


def _sph_values(lang, pts):
    """Real solid harmonics (reference order and scale) of angular momentum lang at points (..., 3)."""
    mons = np.stack([pts[..., 0] ** a * pts[..., 1] ** b * pts[..., 2] ** c
                     for a, b, c in cartesian_exponents(lang)], axis=-1)
    coef = np.zeros((len(cartesian_exponents(lang)), 2 * lang + 1))
    for m, terms in enumerate(SPH[lang]):
        for ic, f in terms:
            coef[ic, m] = f
    return mons @ coef


_PINV = {lang: np.linalg.pinv(_sph_values(lang, _rng_pts)) for lang in range(4)}


def shell_rotations(Q, lang):
    """U[..., m, m'] with f_m(Q r) = sum_m' U[m, m'] f_m'(r) for rotations Q (..., 3, 3)."""
    rotated = np.einsum('...ij,pj->...pi', Q, _rng_pts)
    vals = _sph_values(lang, rotated)                         # (..., npts, nm)
    return np.swapaxes(np.einsum('mp,...pn->...mn', _PINV[lang], vals), -1, -2)


def atom_ao_layout(basis, atom):
    """(l, m, shell) of each AO of an atom, in the order of MolecularBasis.atom_ao_indices."""
    out = []
    for lang in range(basis.max_angular_momentum() + 1):
        nf = basis.number_of_atom_basis_functions(atom, lang)
        for m in range(2 * lang + 1):
            for f in range(nf):
                out.append((lang, m, f))
    return out


def atom_rotation(layout, U_by_l):
    """AO-block rotations of one atom for a batch of frames: U_l[..., m, m'] between the AOs of one
    shell, zero elsewhere (U_by_l[l] has shape (..., 2l+1, 2l+1))."""
    n = len(layout)
    batch = U_by_l[0].shape[:-2]
    R = np.zeros(batch + (n, n))
    for i, (li, mi, fi) in enumerate(layout):
        js = [j for j, (lj, _, fj) in enumerate(layout) if lj == li and fj == fi]
        R[..., i, js] = U_by_l[li][..., mi, [layout[j][1] for j in js]]
    return R


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _perp(ez):
    """A unit vector perpendicular to ez (for frames with no second direction)."""
    t = np.where(np.abs(ez[..., :1]) < 0.9, np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0]))
    return _unit(t - np.sum(t * ez, axis=-1, keepdims=True) * ez)


def _frame(ez, toward):
    """Columns (e_x, e_y, e_z): e_x is `toward` made orthogonal to e_z (any perpendicular when
    `toward` is None or parallel to e_z)."""
    if toward is None:
        ex = _perp(ez)
    else:
        w = toward - np.dot(toward, ez) * ez
        ex = w / np.linalg.norm(w) if np.linalg.norm(w) > 1e-6 else _perp(ez)
    ey = np.cross(ez, ex)
    return np.stack([ex, ey, ez], axis=1)


class PairData:
    """Samples of one molecule: pair list, local frames, features, and the AO maps."""

    def __init__(self, molecule, basis):
        self.z = np.asarray(molecule.get_element_ids())
        if not set(self.z.tolist()) <= set(ELEMENTS):
            raise ValueError('learned guess: element outside H, C, N, O, F')
        xyz = np.asarray(molecule.get_coordinates_in_bohr()) * BOHR   # Angstrom
        self.natoms = len(self.z)
        self.ao = basis.atom_ao_indices()
        self.layout = [atom_ao_layout(basis, a) for a in range(self.natoms)]
        if max(len(x) for x in self.ao) > MAX_BF:
            raise ValueError('learned guess: more than 14 functions on an atom')
        dist = np.linalg.norm(xyz[:, None, :] - xyz[None, :, :], axis=-1)
        order = np.argsort(dist + np.eye(self.natoms) * 1e9, axis=1, kind='stable')   # neighbours by distance
        counts = np.zeros((self.natoms, 4))
        for k, zz in enumerate((1, 6, 7)):
            counts[:, k] = np.sum((dist < NEIGH_R) & (self.z[None, :] == zz), axis=1)
        counts[:, 3] = np.sum((dist < NEIGH_R) & (self.z[None, :] >= 8), axis=1) - (self.z >= 8)
        counts[:, 0] -= (self.z == 1)
        counts[:, 1] -= (self.z == 6)
        counts[:, 2] -= (self.z == 7)
        pairs, frames, feats = [], [], []
        centers = np.linspace(0.0, RCUT, NRADIAL)
        width = centers[1] - centers[0]
        eidx = {z: k for k, z in enumerate(ELEMENTS)}
        for a in range(self.natoms):
            for b in range(a, self.natoms):
                if dist[a, b] > RCUT:
                    continue
                p, q = (a, b) if (self.z[a] > self.z[b] or (self.z[a] == self.z[b])) else (b, a)
                f = np.zeros(N_IN)
                f[eidx[self.z[p]] * len(ELEMENTS) + eidx[self.z[q]]] = 1.0
                o = len(ELEMENTS) ** 2
                f[o] = 1.0 if p == q else 0.0
                f[o + 1:o + 1 + NRADIAL] = np.exp(-0.5 * ((dist[p, q] - centers) / width) ** 2)
                o += 1 + NRADIAL
                f[o:o + 4] = counts[p] / 4.0
                f[o + 4:o + 8] = counts[q] / 4.0
                o += 8
                if p == q:
                    nb = [k for k in order[p] if k != p][:2]
                    if nb:
                        ez = _unit(xyz[nb[0]] - xyz[p])
                        tw = xyz[nb[1]] - xyz[p] if len(nb) > 1 else None
                        Q = _frame(ez, tw)
                        f[o] = dist[p, nb[0]] / 3.0
                        if len(nb) > 1:
                            f[o + 1] = dist[p, nb[1]] / 3.0
                            f[o + 2] = float(np.dot(ez, _unit(tw)))
                    else:
                        Q = np.eye(3)
                else:
                    ez = _unit(xyz[q] - xyz[p])
                    kp = [k for k in order[p] if k not in (p, q)][:1]
                    kq = [k for k in order[q] if k not in (p, q)][:1]
                    Q = _frame(ez, xyz[kp[0]] - xyz[p] if kp else None)
                    if kp:
                        f[o] = dist[p, kp[0]] / 3.0
                        f[o + 1] = float(np.dot(ez, _unit(xyz[kp[0]] - xyz[p])))
                    if kq:
                        f[o + 2] = dist[q, kq[0]] / 3.0
                        f[o + 3] = float(np.dot(-ez, _unit(xyz[kq[0]] - xyz[q])))
                pairs.append((p, q))
                frames.append(Q)
                feats.append(f)
        self.pairs = np.array(pairs, dtype=np.int64)
        self.frames = np.array(frames)
        self.features = np.array(feats, dtype=np.float32)
        self.shape = np.array([[len(self.ao[p]), len(self.ao[q])] for p, q in pairs], dtype=np.int32)
        U = {lang: shell_rotations(self.frames, lang) for lang in range(basis.max_angular_momentum() + 1)}
        # per pair: the AO-block rotations of both atoms in the pair's frame
        self._rot = [[None, None] for _ in pairs]
        for side in (0, 1):
            atoms = self.pairs[:, side]
            for key in {tuple(self.layout[a]) for a in atoms.tolist()}:
                sel = np.array([k for k in range(len(pairs)) if tuple(self.layout[atoms[k]]) == key])
                R = atom_rotation(list(key), {lang: U[lang][sel] for lang in U})
                for r, k in enumerate(sel):
                    self._rot[k][side] = R[r]

    def targets(self, density):
        """Local-frame blocks R_A^T D[A, B] R_B of a density, flattened as 14 x 14."""
        t = np.zeros((len(self.pairs), N_OUT), dtype=np.float32)
        for k, (p, q) in enumerate(self.pairs):
            Rp, Rq = self._rot[k]
            blk = Rp.T @ density[np.ix_(self.ao[p], self.ao[q])] @ Rq
            tt = np.zeros((MAX_BF, MAX_BF))
            tt[:blk.shape[0], :blk.shape[1]] = blk
            t[k] = tt.ravel()
        return t

    def assemble(self, blocks, nao):
        """The symmetric density whose pair blocks are the local blocks rotated back."""
        D = np.zeros((nao, nao))
        for k, (p, q) in enumerate(self.pairs):
            Rp, Rq = self._rot[k]
            ni, nj = self.shape[k]
            loc = np.asarray(blocks[k], dtype=np.float64).reshape(MAX_BF, MAX_BF)[:ni, :nj]
            blk = Rp @ loc @ Rq.T
            if p == q:
                blk = 0.5 * (blk + blk.T)
                D[np.ix_(self.ao[p], self.ao[p])] = blk
            else:
                D[np.ix_(self.ao[p], self.ao[q])] = blk
                D[np.ix_(self.ao[q], self.ao[p])] = blk.T
        return D


def converged_density(molecule, basis, conv_thresh=1.0e-6):
    """Converged RHF alpha density of this package's GPU SCF (training data)."""
    from .scfdriver import ScfRestrictedDriver
    drv = ScfRestrictedDriver()
    drv.conv_thresh = conv_thresh
    res = drv.compute(molecule, basis)
    if not res['is_converged']:
        return None
    return res['D_alpha']


def sample_keys(features):
    """Baseline class of each sample: (element pair, on-site flag, nearest radial centre), one int."""
    ne = len(ELEMENTS) ** 2
    pair = np.argmax(features[:, :ne], axis=1)
    onsite = (features[:, ne] > 0.5).astype(np.int64)
    rbin = np.argmax(features[:, ne + 1:ne + 1 + NRADIAL], axis=1)
    return (pair * 2 + onsite) * NRADIAL + rbin


class LearnedGuess:
    """The trained network and its use as an SCF starting density.

    The network learns a scaled residual: the target block minus the mean block of its baseline
    class (element pair, on-site flag, distance to the nearest 0.19 Angstrom), divided by the
    residuals' root mean square over the training set (one scale for all elements: a scale per
    element would blow up the elements that are zero by symmetry up to rounding).  BF16 carries
    about three significant digits, so its rounding then falls on the residual, not on the whole
    block."""

    def __init__(self, n_hidden=256, max_batch=8192, seed=0, device=0):
        from .densitynet import DensityNet
        self.net = DensityNet(N_IN, n_hidden, N_OUT, MAX_BF, max_batch, seed=seed, device=device)
        self.history = []
        self.base_keys = np.zeros(0, dtype=np.int64)
        self.base_mean = np.zeros((0, N_OUT), dtype=np.float32)
        self.scale = np.ones(N_OUT, dtype=np.float32)

    def fit_baseline(self, features, targets):
        """Mean block per baseline class and the residual scale per output element."""
        keys = sample_keys(features)
        self.base_keys, inv = np.unique(keys, return_inverse=True)
        mean = np.zeros((len(self.base_keys), N_OUT))
        np.add.at(mean, inv, targets.astype(np.float64))
        mean /= np.bincount(inv)[:, None]
        resid = targets - mean[inv]
        self.base_mean = mean.astype(np.float32)
        rms = float(np.sqrt(np.mean(resid.astype(np.float64) ** 2)))
        self.scale = np.full(N_OUT, max(rms, 1e-7), dtype=np.float32)

    def baseline(self, features):
        """Mean blocks of the samples' classes; a class unseen in training takes the nearest radial
        centre of its (element pair, on-site) group, or zeros."""
        keys = sample_keys(features)
        out = np.zeros((len(keys), N_OUT), dtype=np.float32)
        if len(self.base_keys) == 0:
            return out
        idx = np.clip(np.searchsorted(self.base_keys, keys), 0, len(self.base_keys) - 1)
        hit = self.base_keys[idx] == keys
        out[hit] = self.base_mean[idx[hit]]
        groups = self.base_keys // NRADIAL
        for k in np.nonzero(~hit)[0]:
            same = np.nonzero(groups == keys[k] // NRADIAL)[0]
            if len(same):
                near = same[np.argmin(np.abs(self.base_keys[same] % NRADIAL - keys[k] % NRADIAL))]
                out[k] = self.base_mean[near]
        return out

    def supports(self, molecule, basis):
        return (basis.get_label().upper() == BASIS and
                set(np.asarray(molecule.get_element_ids()).tolist()) <= set(ELEMENTS))

    def train(self, features, targets, shapes, epochs=40, batch=4096, lr=2e-3, seed=0, log=None):
        """Baseline fit, then Adam over shuffled mini-batches of the normalised residuals with a
        cosine learning-rate decay; returns the last epoch's mean loss (on the normalised scale)."""
        self.fit_baseline(features, targets)
        targets = ((targets - self.baseline(features)) / self.scale).astype(np.float32)
        rng = np.random.Generator(np.random.PCG64(seed))
        n = len(features)
        steps = epochs * max(1, n // batch)
        step = 0
        mean = 0.0
        t0 = time.perf_counter()
        for ep in range(epochs):
            perm = rng.permutation(n)
            tot, nb = 0.0, 0
            for s in range(0, n - batch + 1 if n >= batch else 1, batch):
                idx = perm[s:s + batch]
                loss, _ = self.net.loss_grad(features[idx], targets[idx], shapes[idx])
                cur = lr * 0.5 * (1.0 + math.cos(math.pi * step / max(1, steps)))
                self.net.adam_step(cur)
                step += 1
                tot += loss
                nb += 1
            mean = tot / max(1, nb)
            self.history.append(mean)
            if log:
                log(f'epoch {ep + 1}/{epochs}: loss {mean:.3e} ({time.perf_counter() - t0:.1f} s)')
        return mean

    def predict_blocks(self, pdata):
        y = self.net.forward(pdata.features).cpu().numpy()
        return y * self.scale + self.baseline(pdata.features)

    def density(self, molecule, basis, overlap=None, purify=3):
        """Predicted alpha density; with the overlap given, scaled so that tr(D S) = N_alpha and made
        idempotent in the overlap metric by up to `purify` McWeeny steps D <- 3 DSD - 2 DSDSD (each
        kept only while it lowers |DSD - D|; the predicted blocks are close to a converged density, so
        the eigenvalues of DS start near 0 and 1, where the step converges quadratically)."""
        pd = PairData(molecule, basis)
        D = pd.assemble(self.predict_blocks(pd), basis.get_dimensions_of_basis())
        if overlap is not None:
            nocc = molecule.number_of_alpha_electrons()
            ne = float(np.sum(D * overlap))
            if ne > 0:
                D *= nocc / ne
            err = np.linalg.norm(D @ overlap @ D - D)
            for _ in range(purify):
                DS = D @ overlap
                DSD = DS @ D
                Dn = 3.0 * DSD - 2.0 * DS @ DSD
                en = np.linalg.norm(Dn @ overlap @ Dn - Dn)
                if not en < err:
                    break
                D, err = 0.5 * (Dn + Dn.T), en
            ne = float(np.sum(D * overlap))
            if ne > 0:
                D *= nocc / ne
        return D


def make_dataset(molecules, basis_name='def2-svp', log=None):
    """(features, targets, shapes) of every atom pair of the molecules, from their converged SCF
    densities (molecules whose SCF does not converge are skipped)."""
    feats, targs, shapes = [], [], []
    t0 = time.perf_counter()
    for i, mol in enumerate(molecules):
        bas = MolecularBasis.read(mol, basis_name)
        D = converged_density(mol, bas)
        if D is None:
            continue
        pd = PairData(mol, bas)
        feats.append(pd.features)
        targs.append(pd.targets(D))
        shapes.append(pd.shape)
        if log and (i + 1) % 50 == 0:
            log(f'{i + 1}/{len(molecules)} training densities ({time.perf_counter() - t0:.1f} s)')
    return np.concatenate(feats), np.concatenate(targs), np.concatenate(shapes)
