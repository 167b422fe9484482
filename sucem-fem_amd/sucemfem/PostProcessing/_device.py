"""DOF vectors handed to the device post-processing entry points."""
import numpy as np


def device_dofs(space, dofs):
    """``dofs`` (numpy array or device tensor) as a contiguous complex128 device tensor of
    length ``space.dim()``; host arrays are copied once."""
    import torch
    n = space.dim()
    if isinstance(dofs, torch.Tensor) and dofs.is_cuda:
        t = dofs.to(torch.complex128).contiguous()
    else:
        t = space.ctx.to_device(np.asarray(dofs), "complex128")
    if t.numel() != n:
        raise ValueError("expected %d DOF values, got %d" % (n, t.numel()))
    return t
