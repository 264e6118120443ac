"""Learned path-guiding distribution (lmgpu_guide_*): an MLP from (position,
normal, outgoing direction) to an 8-lobe vMF mixture, trained online on the
GPU with hand-written MFMA forward / backward / Adam kernels."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .accel import _stream_ptr

NPARAMS = 17664
MIX = 40
RECORD = 16
N_LOBES = 8
# packed layout offsets (transposed weights, output-neuron major)
OFF_W1, OFF_B1 = 0, 64 * 80
OFF_W2, OFF_B2 = OFF_B1 + 64, OFF_B1 + 64 + 64 * 64
OFF_W3, OFF_B3 = OFF_B2 + 64, OFF_B2 + 64 + 64 * 64
OFF_W4, OFF_B4 = OFF_B3 + 64, OFF_B3 + 64 + 64 * 64


def unpack_params(p: np.ndarray):
    """Packed FP32 params -> [(W [nout,nin], b [nout])] per layer (real sizes)."""
    p = np.asarray(p, dtype=np.float32)
    W1 = p[OFF_W1:OFF_B1].reshape(64, 80)[:, :78]
    W2 = p[OFF_W2:OFF_B2].reshape(64, 64)
    W3 = p[OFF_W3:OFF_B3].reshape(64, 64)
    W4 = p[OFF_W4:OFF_B4].reshape(64, 64)[:40]
    return [(W1, p[OFF_B1:OFF_B1 + 64]), (W2, p[OFF_B2:OFF_B2 + 64]), (W3, p[OFF_B3:OFF_B3 + 64]),
            (W4, p[OFF_B4:OFF_B4 + 40])]


class Guide:
    def __init__(self, box_lo, box_hi, seed: int = 1):
        lo = (ctypes.c_float * 3)(*[float(x) for x in box_lo])
        hi = (ctypes.c_float * 3)(*[float(x) for x in box_hi])
        h = ctypes.c_void_p()
        _lib.call('lmgpu_guide_create', ctypes.byref(h), ctypes.cast(lo, ctypes.c_void_p),
                  ctypes.cast(hi, ctypes.c_void_p), int(seed))
        self._h = h
        self.box_lo = np.asarray(box_lo, dtype=np.float32)
        self.box_hi = np.asarray(box_hi, dtype=np.float32)

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.lmgpu_guide_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def get_params(self) -> np.ndarray:
        p = np.zeros(NPARAMS, dtype=np.float32)
        _lib.call('lmgpu_guide_params', self._h, p.ctypes.data_as(ctypes.c_void_p), None)
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (NPARAMS,)
        _lib.call('lmgpu_guide_params', self._h, None, p.ctypes.data_as(ctypes.c_void_p))

    def infer(self, query, mix=None, stream=None):
        """query: [n,9] float32 CUDA tensor (pos, nrm, wo) -> mix [n,40]."""
        import torch
        qp = _lib.device_tensor(query, 'query', torch.float32, cols=9)
        n = query.shape[0]
        if mix is None:
            mix = torch.empty((n, MIX), dtype=torch.float32, device=query.device)
        mp = _lib.device_tensor(mix, 'mix', torch.float32, cols=MIX, min_rows=n, device=query.device)
        _lib.call('lmgpu_guide_infer', self._h, ctypes.c_void_p(qp), ctypes.c_void_p(mp),
                  n, ctypes.c_void_p(_stream_ptr(stream)))
        return mix

    def stats(self):
        """(Adam steps taken, loss of the last step)."""
        steps = ctypes.c_int64(0)
        loss = ctypes.c_float(0.0)
        _lib.call('lmgpu_guide_stats', self._h, ctypes.byref(steps), ctypes.byref(loss))
        return steps.value, loss.value

    def train(self, records, lr: float = 1e-3, want_grad: bool = False, stream=None):
        """records: [n,16] float32 CUDA tensor. Returns (loss, grad or None)."""
        import torch
        rp = _lib.device_tensor(records, 'records', torch.float32, cols=RECORD)
        loss = ctypes.c_float(0.0)
        g = np.zeros(NPARAMS, dtype=np.float32) if want_grad else None
        _lib.call('lmgpu_guide_train', self._h, ctypes.c_void_p(rp), records.shape[0],
                  float(lr), ctypes.byref(loss), g.ctypes.data_as(ctypes.c_void_p) if g is not None else None,
                  ctypes.c_void_p(_stream_ptr(stream)))
        return loss.value, g


def sample(mix, u, stream=None):
    import torch
    mp = _lib.device_tensor(mix, 'mix', torch.float32, cols=MIX)
    n = mix.shape[0]
    up = _lib.device_tensor(u, 'u', torch.float32, cols=3, min_rows=n, device=mix.device)
    dirs = torch.empty((n, 3), dtype=torch.float32, device=mix.device)
    pdf = torch.empty((n,), dtype=torch.float32, device=mix.device)
    _lib.call('lmgpu_guide_sample', ctypes.c_void_p(mp), ctypes.c_void_p(up),
              ctypes.c_void_p(dirs.data_ptr()), ctypes.c_void_p(pdf.data_ptr()), n,
              ctypes.c_void_p(_stream_ptr(stream)))
    return dirs, pdf


def pdf(mix, dirs, stream=None):
    import torch
    mp = _lib.device_tensor(mix, 'mix', torch.float32, cols=MIX)
    n = mix.shape[0]
    dp = _lib.device_tensor(dirs, 'dirs', torch.float32, cols=3, min_rows=n, device=mix.device)
    out = torch.empty((n,), dtype=torch.float32, device=mix.device)
    _lib.call('lmgpu_guide_pdf', ctypes.c_void_p(mp), ctypes.c_void_p(dp),
              ctypes.c_void_p(out.data_ptr()), n, ctypes.c_void_p(_stream_ptr(stream)))
    return out
