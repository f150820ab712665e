"""AO matrices split into angular-momentum blocks.

Mirrors VeloxChem's ``CMatrix`` / ``CSubMatrix`` / ``make_matrix`` surface
(src/math/MatrixFunc.cpp:36-78, src/math/Matrix.cpp:348-373,
src/python/ExportMath.cpp:243-310): a matrix over a molecular basis is a set
of sub-matrices, one per pair of angular momenta (la <= lb for
``mat_t.symmetric``, all pairs for ``mat_t.general``); ``set_values`` takes a
dense N x N array and ``to_numpy`` / ``full_matrix`` return the dense matrix.
Values are held dense here (row-major float64) since the native library works
on dense N x N device arrays; the block view is reconstructed on demand.

Element comparison follows ``mathfunc::equal`` with rtol = atol = 1e-12
(src/math/SubMatrix.cpp:105-115, src/math/MathFunc.hpp:72-77).
"""

import enum

import numpy as np


class mat_t(enum.Enum):
    symmetric = 0
    antisymmetric = 1
    general = 2


def _equal(a, b, rtol=1.0e-12, atol=1.0e-12):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.shape != b.shape:
        return False
    tol = np.maximum(atol, rtol * np.maximum(np.abs(a), np.abs(b)))
    return bool(np.all(np.abs(a - b) < tol))


class SubMatrix:
    """Dense block with dimensions ``[row_offset, col_offset, nrows, ncols]``."""

    def __init__(self, dims, values=None):
        self._dims = tuple(int(x) for x in dims)
        if values is None:
            values = np.zeros((self._dims[2], self._dims[3]))
        self._values = np.array(values, dtype=np.float64).reshape(self._dims[2], self._dims[3])

    def set_values(self, values):
        self._values = np.array(values, dtype=np.float64).reshape(self._dims[2], self._dims[3])

    def to_numpy(self):
        return self._values.copy()

    def get_dimensions(self):
        return list(self._dims)

    def __eq__(self, other):
        return self._dims == other._dims and _equal(self._values, other._values)

    def __ne__(self, other):
        return not self == other


class Matrix:
    """AO matrix with angular block structure of a molecular basis."""

    def __init__(self, block_dims, mtype=mat_t.symmetric, values=None):
        self._block_dims = list(block_dims)
        self._type = mtype
        n = self._block_dims[-1]
        self._values = np.zeros((n, n)) if values is None else np.array(values, dtype=np.float64)

    def get_type(self):
        return self._type

    def number_of_rows(self):
        return self._block_dims[-1]

    number_of_columns = number_of_rows

    def set_values(self, values):
        values = np.asarray(values, dtype=np.float64)
        n = self._block_dims[-1]
        if values.shape != (n, n):
            raise ValueError(f'Matrix.set_values: expected shape {(n, n)}, got {values.shape}')
        if self._type == mat_t.symmetric:
            # the reference keeps only la <= lb blocks and mirrors them on export
            full = np.triu(np.ones((n, n), dtype=bool))
            bd = self._block_dims
            keep = np.zeros((n, n), dtype=bool)
            for i in range(len(bd) - 1):
                for j in range(i, len(bd) - 1):
                    keep[bd[i]:bd[i + 1], bd[j]:bd[j + 1]] = True
            del full
            vals = np.where(keep, values, values.T)
            self._values = np.ascontiguousarray(vals)
        else:
            self._values = np.ascontiguousarray(values.copy())

    def zero(self):
        self._values[:] = 0.0

    def to_numpy(self):
        return self._values.copy()

    def full_matrix(self):
        n = self._block_dims[-1]
        return SubMatrix([0, 0, n, n], self._values)

    def submatrix(self, angpair):
        i, j = angpair
        bd = self._block_dims
        return SubMatrix([bd[i], bd[j], bd[i + 1] - bd[i], bd[j + 1] - bd[j]],
                         self._values[bd[i]:bd[i + 1], bd[j]:bd[j + 1]])

    sub_matrix = submatrix

    def angular_pairs(self):
        nb = len(self._block_dims) - 1
        if self._type == mat_t.general:
            return [(i, j) for i in range(nb) for j in range(nb)]
        return [(i, j) for i in range(nb) for j in range(i, nb)]

    def __eq__(self, other):
        return self._type == other._type and _equal(self._values, other._values)


def make_matrix(basis, mtype=mat_t.symmetric):
    """Zero matrix with the angular block layout of ``basis``."""
    return Matrix(basis.angular_block_dims(), mtype)


class Matrices:
    """Named collection of matrices (multi-density Fock calls)."""

    def __init__(self, mats=None):
        self._mats = {}
        if mats:
            for k, m in enumerate(mats):
                self._mats[str(k)] = m

    def add(self, matrix, key):
        self._mats[str(key)] = matrix

    def matrix(self, key):
        return self._mats[str(key)]

    def keys(self):
        return list(self._mats.keys())

    def number_of_matrices(self):
        return len(self._mats)
