"""The allocation-free point-source entry point (sucem_rhs_point_sources_noalloc) against the
allocating one it mirrors (sucem_rhs_point_sources), called through the C-ABI: equal
results on two calls in a row, the second with fewer sources and right-hand sides (the
staging array keeps the first call's contents), and its limits."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rhs(entry, prob, pos, val, idx, nrhs, scale):
    from sucemfem import _hip
    b = prob.ctx.tensor(nrhs * prob.ndofs)
    _hip.call(entry, prob.ctx.handle, prob.space.handle.h, len(pos), _hip.ptr(pos),
              _hip.ptr(val), _hip.ptr(idx), nrhs, float(scale.real), float(scale.imag),
              _hip.ptr(b))
    return b.cpu().numpy()


def test_noalloc_point_sources_match_allocating_entry():
    import torch
    torch.cuda.set_device(0)
    from sucemfem import _hip
    from sucemfem.Utilities.Workloads import DeviceProblem
    print("no-alloc point sources", flush=True)
    sys.stdout.flush()
    prob = DeviceProblem(4, 1.0, 2, _hip.MAT_S | _hip.MAT_T)
    rng = np.random.default_rng(11)
    scale = complex(0.0, -3.25)
    for nsrc, nrhs in ((300, 5), (17, 2)):
        pos = np.ascontiguousarray(rng.uniform(0.05, 0.95, size=(nsrc, 3)))
        val = np.ascontiguousarray(rng.standard_normal((nsrc, 3))
                                   + 1j * rng.standard_normal((nsrc, 3)))
        idx = np.ascontiguousarray(rng.integers(0, nrhs, size=nsrc), np.int64)
        ref = _rhs("sucem_rhs_point_sources", prob, pos, val, idx, nrhs, scale)
        got = _rhs("sucem_rhs_point_sources_noalloc", prob, pos, val, idx, nrhs, scale)
        assert np.abs(ref).max() > 0
        assert np.array_equal(got, ref), (nsrc, np.abs(got - ref).max())
        # without indices every source goes to the first right-hand side
        ref0 = _rhs("sucem_rhs_point_sources", prob, pos, val, None, 1, scale)
        got0 = _rhs("sucem_rhs_point_sources_noalloc", prob, pos, val, None, 1, scale)
        assert np.array_equal(got0, ref0)
    # limits: more than 4096 sources, a point outside the mesh
    big = np.zeros((4097, 3)) + 0.5
    bval = np.zeros((4097, 3), np.complex128)
    with pytest.raises(ValueError, match="4096"):
        _rhs("sucem_rhs_point_sources_noalloc", prob, big, bval, None, 1, scale)
    out = np.array([[0.5, 0.5, 0.5], [1.5, 0.5, 0.5]])
    oval = np.ones((2, 3), np.complex128)
    with pytest.raises(AssertionError, match="outside"):
        _rhs("sucem_rhs_point_sources_noalloc", prob, out, oval, None, 1, scale)
    # after the error the staging is reset by the next call
    pos = np.array([[0.3, 0.4, 0.6]])
    val = np.array([[1.0, 2.0, -1.0]], np.complex128)
    assert np.array_equal(_rhs("sucem_rhs_point_sources_noalloc", prob, pos, val, None, 1, scale),
                          _rhs("sucem_rhs_point_sources", prob, pos, val, None, 1, scale))
