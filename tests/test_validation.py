"""Host-side checks of the device buffers handed to the C ABI: the kernels
index rays, hits, occlusion flags, films and guide buffers by the row count
they are given, so a wrong dtype, a short or non-contiguous tensor or host
memory must raise ValueError before any library call (a device write out of
bounds would otherwise follow). These run without a GPU: every rejected
tensor lives in host memory, and the checks of dtype, shape and contiguity
come before the device check."""
import ctypes

import pytest
import torch

from lm_amd import _lib
from lm_amd.accel import Accel
from lm_amd.renderer import GpuScene


class _Calls:
    """Replaces _lib.call so a test fails if any entry point is reached."""

    def __init__(self, monkeypatch):
        self.names = []
        monkeypatch.setattr(_lib, 'call', self._call)

    def _call(self, name, *args):
        self.names.append(name)
        raise AssertionError(f'{name} reached with an invalid buffer')


def _accel():
    a = Accel.__new__(Accel)   # no library handle needed: nothing may be called
    a._h = ctypes.c_void_p()
    a.width, a.leaf_size, a.refill, a.builder, a.radius, a.collapse, a.sah_tri_cost = 4, 2, 32, 1, 16, 0, 0.0
    return a


@pytest.mark.parametrize('rays,hits,what', [
    (torch.zeros((10, 8), dtype=torch.float64), torch.zeros((10, 4)), 'rays must be torch.float32'),
    (torch.zeros((10, 7)), torch.zeros((10, 4)), r'rays must be \[>=0, 8\]'),
    (torch.zeros((8, 10)).t(), torch.zeros((10, 4)), 'rays must be contiguous'),
    (torch.zeros((10, 8)), torch.zeros((9, 4)), r'hits must be \[>=10, 4\]'),
    (torch.zeros((10, 8)), torch.zeros((10, 4), dtype=torch.int32), 'hits must be torch.float32'),
    (torch.zeros((10, 8)), torch.zeros((10, 4)), 'must be a CUDA'),
    ([[0.0] * 8], torch.zeros((1, 4)), 'must be a torch tensor'),
])
def test_intersect_rejects_bad_buffers(monkeypatch, rays, hits, what):
    calls = _Calls(monkeypatch)
    a = _accel()
    with pytest.raises(ValueError, match=what):
        a.intersect_device(rays, hits)
    with pytest.raises(ValueError):
        a.intersect_counted(rays, hits)
    assert calls.names == []


@pytest.mark.parametrize('occl,what', [
    (torch.zeros(9, dtype=torch.uint8), 'at least 10'),
    (torch.zeros(10, dtype=torch.int32), 'occl must be torch.uint8'),
    (torch.zeros((10, 1), dtype=torch.uint8), '1-D'),
])
def test_occluded_rejects_bad_buffers(monkeypatch, occl, what):
    calls = _Calls(monkeypatch)
    a = _accel()
    with pytest.raises(ValueError, match=what):
        a.occluded_device(torch.zeros((10, 8)), occl)
    assert calls.names == []


def test_build_rejects_bad_vertex_tensor(monkeypatch):
    calls = _Calls(monkeypatch)
    a = _accel()
    for v, what in ((torch.zeros((5, 8)), r'\[>=0, 9\]'), (torch.zeros(45), r'\[>=0, 9\]'),
                    (torch.zeros((5, 9), dtype=torch.float64), 'float32'), (torch.zeros((5, 9)), 'CUDA')):
        with pytest.raises(ValueError, match=what):
            a.build(v)
    assert calls.names == []


def _scene(h=6, w=8):
    s = GpuScene.__new__(GpuScene)
    s._h = ctypes.c_void_p()
    s.h, s.w = h, w
    return s


@pytest.mark.parametrize('film,what', [
    (torch.zeros((6, 8, 3), dtype=torch.float32), 'float64'),
    (torch.zeros((8, 6, 3), dtype=torch.float64), r'shape \(6, 8, 3\)'),
    (torch.zeros((6, 8, 4), dtype=torch.float64)[..., :3], 'contiguous'),
    (torch.zeros((6, 8, 3), dtype=torch.float64), 'CUDA'),
])
def test_films_rejected_before_render(monkeypatch, film, what):
    calls = _Calls(monkeypatch)
    s = _scene()
    with pytest.raises(ValueError, match=what):
        s.render_raw(spp=1, film_sum=film)
    with pytest.raises(ValueError, match=what):
        s.render_lt(num_samples=1, film=film)
    with pytest.raises(ValueError, match=what):
        s.render_bdpt(num_samples=1, film=film)
    assert calls.names == []


def test_host_array_films_rejected(monkeypatch):
    """LT and BDPT accumulate only into device films: a host array is an
    error, not silently replaced by a fresh film."""
    import numpy as np
    calls = _Calls(monkeypatch)
    s = _scene()
    film = np.zeros((6, 8, 3))
    with pytest.raises(ValueError, match='torch CUDA tensor'):
        s.render_lt(num_samples=1, film=film)
    with pytest.raises(ValueError, match='torch CUDA tensor'):
        s.render_bdpt(num_samples=1, film=film)
    assert calls.names == []
