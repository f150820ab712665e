"""Device-side density network of the learned SCF guess (vlx_mlp_* of libvlxfock.so).

The network n_in -> n_hidden -> n_hidden -> n_out (SiLU) runs its forward
pass, gradients and Adam steps in the library's BF16 matrix-core kernels; this
class holds its handle, seeds its parameters and moves batches through torch
device tensors (device memory only: every product, activation, loss and
update is the library's).  No reference counterpart (the reference starts
from the SAD guess, src/pymodule/sadguessdriver.py:841-971).
"""
import ctypes

import numpy as np
import torch

from . import _native

_fp = ctypes.POINTER(ctypes.c_float)


def _bind(L):
    if getattr(L, '_mlp_bound', False):
        return L
    L.vlx_mlp_create.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_void_p)]
    L.vlx_mlp_destroy.argtypes = [ctypes.c_void_p]
    L.vlx_mlp_num_params.argtypes = [ctypes.c_void_p, _native._c_ll_p]
    L.vlx_mlp_set_params.argtypes = [ctypes.c_void_p, _fp]
    L.vlx_mlp_get_params.argtypes = [ctypes.c_void_p, _fp]
    L.vlx_mlp_forward.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L.vlx_mlp_loss_grad.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                    ctypes.c_void_p, _fp, _fp]
    L.vlx_mlp_adam_step.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]
    L._mlp_bound = True
    return L


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class DensityNet:
    """The network on device `device` (its library context's stream)."""

    def __init__(self, n_in, n_hidden, n_out, block_cols, max_batch, seed=0, device=0):
        self.lib = _bind(_native.lib())
        self.device = device
        self.ctx = _native.context(device)
        self.n_in, self.n_hidden, self.n_out, self.block_cols = n_in, n_hidden, n_out, block_cols
        self.max_batch = max_batch
        h = ctypes.c_void_p()
        _native.check(self.lib.vlx_mlp_create(self.ctx, n_in, n_hidden, n_out, block_cols, max_batch,
                                              ctypes.byref(h)))
        self.handle = h
        n = ctypes.c_longlong()
        _native.check(self.lib.vlx_mlp_num_params(h, ctypes.byref(n)))
        self.num_params = n.value
        self.set_params(self.initial_params(seed))

    def __del__(self):
        h = getattr(self, 'handle', None)
        if h is not None and h.value:
            self.lib.vlx_mlp_destroy(h)
            self.handle = None

    def initial_params(self, seed):
        """He-normal weights (fan-in scaling, suited to SiLU) and zero biases; the output layer is
        scaled down so that a fresh network predicts near zero."""
        rng = np.random.Generator(np.random.PCG64(seed))
        parts = []
        dims = [(self.n_in, self.n_hidden), (self.n_hidden, self.n_hidden), (self.n_hidden, self.n_out)]
        for k, (fi, fo) in enumerate(dims):
            scale = np.sqrt(2.0 / fi) * (0.1 if k == 2 else 1.0)
            parts += [rng.normal(size=fi * fo) * scale, np.zeros(fo)]
        return np.concatenate(parts).astype(np.float32)

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != self.num_params:
            raise ValueError(f'expected {self.num_params} parameters, got {p.size}')
        _native.check(self.lib.vlx_mlp_set_params(self.handle, p.ctypes.data_as(_fp)))

    def get_params(self):
        p = np.empty(self.num_params, dtype=np.float32)
        _native.check(self.lib.vlx_mlp_get_params(self.handle, p.ctypes.data_as(_fp)))
        return p

    def _dev(self, a, dtype=torch.float32):
        """A device copy made on torch's stream, complete before the library's stream reads it."""
        t = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a))
        t = t.to(device=f'cuda:{self.device}', dtype=dtype).contiguous()
        torch.cuda.current_stream(t.device).synchronize()
        return t

    def forward(self, x):
        """Outputs (m x n_out, torch on the device) of x (m x n_in), in batches of max_batch."""
        xd = self._dev(x)
        m = xd.shape[0]
        y = torch.empty((m, self.n_out), dtype=torch.float32, device=xd.device)
        torch.cuda.current_stream(xd.device).synchronize()
        for s in range(0, m, self.max_batch):
            e = min(m, s + self.max_batch)
            _native.check(self.lib.vlx_mlp_forward(self.handle, _ptr(xd[s:e]), e - s, _ptr(y[s:e])))
        _native.check(_native.lib().vlx_context_synchronize(self.ctx))
        return y

    def loss_grad(self, x, target, shape, want_grad=False):
        """(loss, gradient (host, packed) or None) of one batch; the gradient stays on the device for
        adam_step."""
        xd, td, sd = self._dev(x), self._dev(target), self._dev(shape, torch.int32)
        m = xd.shape[0]
        y = torch.empty((m, self.n_out), dtype=torch.float32, device=xd.device)
        torch.cuda.current_stream(xd.device).synchronize()
        loss = ctypes.c_float()
        g = np.empty(self.num_params, dtype=np.float32) if want_grad else None
        _native.check(self.lib.vlx_mlp_loss_grad(self.handle, _ptr(xd), _ptr(td), _ptr(sd), m, _ptr(y),
                                                 ctypes.byref(loss), g.ctypes.data_as(_fp) if want_grad else None))
        return loss.value, g

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        _native.check(self.lib.vlx_mlp_adam_step(self.handle, lr, beta1, beta2, eps))
