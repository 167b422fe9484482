"""Save / load scipy matrices as MATLAB .mat files (sucemfem/Utilities/MatrixIO.py:23-82)."""
import os

import numpy as np


def check_path(path, create=True):
    """True when ``path`` exists afterwards (created when ``create``)."""
    if not os.path.exists(path):
        if not create:
            return True
        os.makedirs(path)
    return True


def save_scipy_matrix_as_mat(path, name, matrix):
    import scipy.io
    if not check_path(path):
        return False
    # <name>.mat, the file name the reference's SciPy wrote
    scipy.io.savemat(os.path.join(path, name) + '.mat', {name: matrix}, oned_as='column')
    return True


def load_scipy_matrix_from_mat(path, name):
    """The stored matrix (CSR for sparse data), or None when the file is absent."""
    import scipy.io
    filename = os.path.join(path, name)
    if not os.path.exists(filename + '.mat'):
        return None
    data = scipy.io.loadmat(filename + '.mat')[name]
    return data if isinstance(data, np.ndarray) else data.tocsr()
