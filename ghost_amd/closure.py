"""The learned subgrid closure's network (no counterpart in GHOST): a
10 -> 64 -> 64 -> 6 GELU perceptron on the gfx950 matrix cores (BF16 operands,
FP32 accumulation) with its forward, backward and Adam kernels in libghx.

Inputs are the nine resolved velocity gradients d u_i / d x_j (i major) and
the filter width; outputs are tau_xx, tau_yy, tau_zz, tau_xy, tau_xz, tau_yz.
Device buffers are torch tensors; every computation runs in the library.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib

NI, NH, NO = 10, 64, 6


def _f32(t, shape=None):
    assert t.is_cuda and t.dtype == torch.float32 and t.is_contiguous(), "float32 contiguous CUDA tensor expected"
    if shape is not None:
        assert tuple(t.shape) == tuple(shape), f"shape {tuple(t.shape)} != {shape}"
    return C.c_void_p(t.data_ptr())


def _ptrs(fields, n, count):
    assert len(fields) == count
    for f in fields:
        assert f.is_cuda and f.dtype == torch.float64 and f.is_contiguous() and f.numel() == n
    arr = (C.c_void_p * count)(*[f.data_ptr() for f in fields])
    return C.cast(arr, _lib._PP), arr


class MLP:
    def __init__(self, seed=0):
        L = _lib.load()
        h = C.c_void_p()
        _lib.check(L.ghx_mlp_create(C.byref(h), int(seed)))
        self._h = h
        self.nparams = L.ghx_mlp_nparams()

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _lib.load().ghx_mlp_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def _pptr(self):
        p = C.c_void_p()
        _lib.check(_lib.load().ghx_mlp_params(self._h, C.byref(p)))
        return p.value

    def params(self):
        """Copy of the FP32 master parameters (W1, b1, W2, b2, W3, b3)."""
        out = torch.empty(self.nparams, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        _lib.check(_lib.load().ghx_memcpy_d2d(C.c_void_p(out.data_ptr()), C.c_void_p(self._pptr()), 4 * self.nparams))
        return out

    def set_params(self, p):
        p = torch.as_tensor(p, dtype=torch.float32).cuda().contiguous()
        assert p.numel() == self.nparams
        torch.cuda.synchronize()
        _lib.check(_lib.load().ghx_memcpy_d2d(C.c_void_p(self._pptr()), C.c_void_p(p.data_ptr()), 4 * self.nparams))

    def norm(self):
        buf = (C.c_float * 32)()
        _lib.check(_lib.load().ghx_mlp_get_norm(self._h, buf))
        return np.array(buf[:], dtype=np.float32)

    def set_norm(self, norm):
        norm = np.ascontiguousarray(norm, dtype=np.float32)
        assert norm.size == 32
        _lib.check(_lib.load().ghx_mlp_set_norm(self._h, norm.ctypes.data_as(C.POINTER(C.c_float))))

    def fit_norm(self, grad9, delta, tau6):
        n = grad9[0].numel()
        g, ka = _ptrs(grad9, n, 9)
        t, kb = _ptrs(tau6, n, 6)
        _lib.check(_lib.load().ghx_mlp_fit_norm(self._h, n, g, float(delta), t))

    def pack(self, grad9, delta, tau6=None):
        n = grad9[0].numel()
        g, ka = _ptrs(grad9, n, 9)
        X = torch.empty((n, NI), dtype=torch.float32, device="cuda")
        T = None
        tp = None
        if tau6 is not None:
            tp, kb = _ptrs(tau6, n, 6)
            T = torch.empty((n, NO), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        _lib.check(_lib.load().ghx_mlp_pack(self._h, n, g, float(delta), tp, _f32(X),
                                            _f32(T) if T is not None else None))
        torch.cuda.synchronize()
        return X, T

    def forward(self, X):
        n = X.shape[0]
        Y = torch.empty((n, NO), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        _lib.check(_lib.load().ghx_mlp_forward(self._h, n, _f32(X, (n, NI)), _f32(Y)))
        torch.cuda.synchronize()
        return Y

    def grad(self, X, T):
        n = X.shape[0]
        g = torch.empty(self.nparams, dtype=torch.float32, device="cuda")
        loss = C.c_double()
        torch.cuda.synchronize()
        _lib.check(_lib.load().ghx_mlp_grad(self._h, n, _f32(X, (n, NI)), _f32(T, (n, NO)), _f32(g), C.byref(loss)))
        return loss.value, g

    def train(self, X, T, batch=65536, steps=100, lr=1e-3, seed=0):
        n = X.shape[0]
        losses = np.zeros(max(steps, 1), dtype=np.float32)
        torch.cuda.synchronize()
        _lib.check(_lib.load().ghx_mlp_train(self._h, n, _f32(X, (n, NI)), _f32(T, (n, NO)), int(batch), int(steps),
                                             float(lr), int(seed), losses.ctypes.data_as(C.POINTER(C.c_float))))
        return losses[:steps]

    def stress(self, grad9, delta, tau6):
        n = grad9[0].numel()
        g, ka = _ptrs(grad9, n, 9)
        t, kb = _ptrs(tau6, n, 6)
        torch.cuda.synchronize()
        _lib.check(_lib.load().ghx_mlp_stress(self._h, n, g, float(delta), t))
