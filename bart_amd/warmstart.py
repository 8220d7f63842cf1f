"""Learned CG warm start: the network of the C ABI (bart_amd_warmnet_*) and
the host-side steps around it -- training pairs from this solver, training,
and CG from the network's guess with the zero start as fallback.

No counterpart in BART: pics takes an initial guess with -W
(src/pics.c:580-594, iter/itop.c:54-66); this module proposes it.  The
network runs on the device through the library (3x3 convolutions on the
matrix cores); torch only holds device memory.
"""
import ctypes

import numpy as np

from . import BartError, _check, _fptr, lib

DIMS = 16


def _lib():
    L = lib()
    if not getattr(L, "_warmnet_ready", False):
        vp = ctypes.c_void_p
        ci = ctypes.c_int
        L.bart_amd_warmnet_create.restype = vp
        L.bart_amd_warmnet_create.argtypes = [ci, ctypes.c_ulong]
        L.bart_amd_warmnet_free.restype = None
        L.bart_amd_warmnet_free.argtypes = [vp]
        L.bart_amd_warmnet_num_params.restype = ctypes.c_long
        L.bart_amd_warmnet_num_params.argtypes = [vp]
        L.bart_amd_warmnet_params.restype = ci
        L.bart_amd_warmnet_params.argtypes = [vp, vp, ci]
        L.bart_amd_warmnet_forward.restype = ci
        L.bart_amd_warmnet_forward.argtypes = [vp, vp, ci, ci, ci, vp]
        L.bart_amd_warmnet_train_step.restype = ctypes.c_double
        L.bart_amd_warmnet_train_step.argtypes = [vp, vp, vp, ci, ci, ci, ctypes.c_float]
        L._warmnet_ready = True
    return L


class WarmNet:
    """2 -> channels (x5) -> 2 channel residual CNN (ReLU, 3x3 convolutions)."""

    def __init__(self, channels=48, seed=0):
        self.h = _lib().bart_amd_warmnet_create(channels, seed)
        if not self.h:
            _check()
            raise BartError("warm-start network creation failed")
        self.channels = channels

    def free(self):
        if getattr(self, "h", None):
            _lib().bart_amd_warmnet_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def nparams(self):
        return int(_lib().bart_amd_warmnet_num_params(self.h))

    def _params(self, mode, values=None):
        buf = np.empty(self.nparams, np.float32) if values is None else np.ascontiguousarray(values, np.float32)
        _lib().bart_amd_warmnet_params(self.h, buf.ctypes.data_as(ctypes.c_void_p), mode)
        _check()
        return buf

    def get_params(self):
        return self._params(0)

    def set_params(self, p):
        assert np.size(p) == self.nparams
        self._params(1, p)

    def get_grads(self):
        """Gradients of the last training step."""
        return self._params(2)

    def forward(self, x):
        """x [B, 2, H, W] float32 CUDA tensor -> x + net(x)."""
        import torch
        B, _, H, W = x.shape
        y = torch.empty_like(x)
        _lib().bart_amd_warmnet_forward(self.h, _fptr(x), B, H, W, _fptr(y))
        _check()
        return y

    def train_step(self, x, target, lr):
        B, _, H, W = x.shape
        loss = _lib().bart_amd_warmnet_train_step(self.h, _fptr(x), _fptr(target), B, H, W, lr)
        _check()
        return float(loss)


def to_channels(img, n):
    """Complex [n*n] device image -> [1, 2, n, n] float32 (real, imaginary)."""
    import torch
    return torch.view_as_real(img.reshape(n, n)).permute(2, 0, 1).contiguous().unsqueeze(0)


def from_channels(t, n):
    """[1, 2, n, n] float32 -> complex [n*n] device image."""
    import torch
    return torch.view_as_complex(t[0].permute(1, 2, 0).contiguous()).reshape(n * n).contiguous()
