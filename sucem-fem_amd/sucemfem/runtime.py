"""Device runtime: one library context per (device, stream) and device-resident handles.

PyTorch-ROCm provides device memory for vectors and the stream; every kernel
runs inside libsucem_hip.
"""
import ctypes
import threading

import numpy as np

from . import _hip

_lock = threading.Lock()
_contexts = {}


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("sucemfem needs a HIP device (torch.cuda.is_available() is False)")
    return torch


class Context(object):
    """A libsucem_hip context bound to a device and a stream."""

    def __init__(self, device=None, stream=None):
        torch = _torch()
        self.device = torch.cuda.current_device() if device is None else int(device)
        with torch.cuda.device(self.device):
            self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        self.handle = _hip.out_handle()
        _hip.call("sucem_ctx_create", self.device, ctypes.c_void_p(self.stream.cuda_stream),
                  ctypes.byref(self.handle))

    def synchronize(self):
        _hip.call("sucem_ctx_synchronize", self.handle)

    def tensor(self, n, dtype="complex128"):
        torch = _torch()
        return torch.zeros(n, dtype=getattr(torch, dtype), device="cuda:%d" % self.device)

    def to_device(self, a, dtype="complex128"):
        torch = _torch()
        t = torch.as_tensor(np.ascontiguousarray(a), device="cuda:%d" % self.device)
        return t.to(getattr(torch, dtype)).contiguous()

    def __del__(self):
        try:
            if self.handle:
                _hip.load().sucem_ctx_destroy(self.handle)
        except Exception:
            pass


def get_context(device=None):
    """The default context of the current (or given) device on torch's current stream."""
    torch = _torch()
    dev = torch.cuda.current_device() if device is None else int(device)
    key = (dev, torch.cuda.current_stream(dev).cuda_stream)
    with _lock:
        ctx = _contexts.get(key)
        if ctx is None:
            ctx = Context(dev)
            _contexts[key] = ctx
    return ctx


class Handle(object):
    """Owner of one library object; destroyed with the given C function."""

    def __init__(self, ctx, destroy):
        self.ctx = ctx
        self.h = _hip.out_handle()
        self._destroy = destroy

    def __del__(self):
        try:
            if self.h:
                getattr(_hip.load(), self._destroy)(self.h)
                self.h = ctypes.c_void_p()
        except Exception:
            pass
