"""Matrix conversions (sucemfem/Utilities/Converters.py:27-51)."""
import numpy as np


def dolfin_ublassparse_to_scipy_csr(A, dtype=None, imagify=False):
    """Assembled device matrix -> ``scipy.sparse.csr_matrix`` (full pattern, explicit zeros kept)."""
    A_sp = A.to_scipy()
    if imagify:
        A_sp = A_sp * 1j
    if dtype is not None:
        A_sp = A_sp.astype(dtype)
    return A_sp


def as_dolfin_vector(a):
    """The reference wraps arrays in dolfin Vectors; numpy arrays are used directly here."""
    assert len(a.shape) == 1
    return np.require(a, requirements=['C'])
