"""DOF numbering, VIJ (COO) pattern and COO->CSR scatter (oracle side).

* DOF index of node position p, component c: 3 p + c (single displacement
  field, node-major interleaved; dofmanager.py:259-300).
* Element DOF list: concatenation over the element's nodes (dofmanager.py:455-471).
* VIJ blocks back to back in element order; inside a block of n DOFs,
  I[o+f] = idcs[f mod n], J[o+f] = idcs[f div n] (dofmanager.py:543-555).
* CSR pattern = SciPy's canonical csr_matrix((0,(I,J))) with int32 indptr /
  indices (csrgenerator.pyx:65-77); x[k] = CSR slot of COO pair k
  (csrgenerator.pyx:81-98).
* updateCSR: data = 0; data[x[k]] += V[k] in k order (csrgenerator.pyx:100-115).
"""

import numpy as np
import scipy.sparse as sp


def vij_pattern(dof_map):
    dof_map = np.asarray(dof_map, dtype=np.int64)
    n = dof_map.shape[1]
    f = np.arange(n * n)
    I = dof_map[:, f % n].ravel()  # noqa: E741
    J = dof_map[:, f // n].ravel()
    return I, J


def csr_pattern(I, J, nDof):
    A = sp.csr_matrix((np.zeros(I.size, dtype=np.double), (I, J)), shape=(nDof, nDof))
    if not A.has_canonical_format:
        A.sum_duplicates()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32)


def x_map(I, J, indptr, indices):
    nDof = indptr.size - 1
    rows = np.repeat(np.arange(nDof, dtype=np.int64), np.diff(indptr))
    keys = rows * nDof + indices.astype(np.int64)
    x = np.searchsorted(keys, I * nDof + J)
    assert np.array_equal(keys[x], I * nDof + J)
    return x.astype(np.int32)


def vij_values_reference(Ke):
    """V blocks as DisplacementElement writes them: Ke row-major into the
    column-major VIJ block (element.py:318), i.e. the global matrix receives Ke^T."""
    return np.ascontiguousarray(Ke).reshape(-1)


def vij_values_colmajor(Ke):
    """V blocks in the column-major (native element ABI) convention: global receives Ke."""
    return np.ascontiguousarray(np.transpose(Ke, (0, 2, 1))).reshape(-1)


def update_csr(V, x, nnz):
    """data[x[k]] += V[k] sequentially in k order."""
    return np.bincount(x, weights=V, minlength=nnz)


def csr_matrix(data, indptr, indices):
    n = indptr.size - 1
    return sp.csr_matrix((data.copy(), indices.copy(), indptr.copy()), shape=(n, n))
