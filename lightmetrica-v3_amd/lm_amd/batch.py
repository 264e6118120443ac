"""Batched multi-job rendering (lmgpu_batch_*): many renderer::pt jobs of
uneven scene size, resolution and spp in one wavefront loop over a pool of
path slots fed from a device work queue, with the jobs' BVHs in one device
arena; BVH width, path capacity and block size autotuned when left 0."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .accel import _stream_ptr
from .renderer import scene_desc


class Batch:
    def __init__(self):
        h = ctypes.c_void_p()
        _lib.call('lmgpu_batch_create', ctypes.byref(h))
        self._h = h
        self.sizes = []

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.lmgpu_batch_destroy(h)
            self._h = None

    def add_job(self, flat, spp: int, max_verts: int = 16, sampling_mode: int = 2, rr_depth: int = 5,
                seed: int = 0) -> int:
        keep, d = scene_desc(flat)
        j = _lib.JobDesc(ctypes.cast(ctypes.pointer(d), ctypes.c_void_p), int(spp), int(max_verts),
                         int(sampling_mode), int(rr_depth), int(seed))
        jid = ctypes.c_int32(-1)
        _lib.call('lmgpu_batch_add_job', self._h, ctypes.byref(j), ctypes.byref(jid))
        del keep   # the library copied the job's arrays
        self.sizes.append((int(flat.h), int(flat.w)))
        return jid.value

    def run(self, width: int = 0, path_capacity: int = 0, block: int = 0, leaf_size: int = 0, sumsq: bool = False,
            interleaved: bool = False, stream=None) -> dict:
        """Renders every job; interleaved=True orders the work queue sample
        by sample across jobs instead of job by job."""
        o = _lib.BatchOpts(int(width), int(path_capacity), int(block), int(leaf_size), 1 if sumsq else 0,
                           _lib.QUEUE_INTERLEAVED if interleaved else _lib.QUEUE_JOB_MAJOR)
        self._sumsq = bool(sumsq)
        st = _lib.BatchStats()
        _lib.call('lmgpu_batch_run', self._h, ctypes.byref(o), ctypes.byref(st), ctypes.c_void_p(_stream_ptr(stream)))
        return {f: getattr(st, f) for f, _ in _lib.BatchStats._fields_ if not f.startswith('reserved')}

    def latency_ms(self, job: int) -> float:
        """Latency of a job in the last run (start of the run to its last sample)."""
        lat = ctypes.c_double(0.0)
        _lib.call('lmgpu_batch_job_film', self._h, int(job), None, None, ctypes.byref(lat))
        return lat.value

    def film(self, job: int):
        """(sum [h,w,3], sumsq [h,w,3] or None, latency_ms) of a job in the
        last run (sumsq when the run was made with sumsq=True)."""
        h, w = self.sizes[job]
        s = np.zeros((h, w, 3), dtype=np.float64)
        q = np.zeros((h, w, 3), dtype=np.float64) if getattr(self, '_sumsq', False) else None
        lat = ctypes.c_double(0.0)
        _lib.call('lmgpu_batch_job_film', self._h, int(job), s.ctypes.data_as(ctypes.c_void_p),
                  q.ctypes.data_as(ctypes.c_void_p) if q is not None else None, ctypes.byref(lat))
        return s, q, lat.value
