"""Loader of liblmgpu.so, the HIP (gfx950) library behind the lmgpu C ABI.

The library is built in-tree (``make -C lightmetrica-v3_amd``). There is no
fallback: if it is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, 'liblmgpu.so')

# Status codes (include/lmgpu.h)
OK = 0
ERR_INVALID_ARGUMENT = -1
ERR_UNSUPPORTED = -2
ERR_FAILED_TO_RENDER = -3
ERR_DEVICE = -4
ERR_OUT_OF_MEMORY = -5
ERR_STACK_OVERFLOW = -6

_ERR_NAMES = {
    ERR_INVALID_ARGUMENT: 'InvalidArgument',
    ERR_UNSUPPORTED: 'Unsupported',
    ERR_FAILED_TO_RENDER: 'FailedToRender',
    ERR_DEVICE: 'Device',
    ERR_OUT_OF_MEMORY: 'OutOfMemory',
    ERR_STACK_OVERFLOW: 'StackOverflow',
}

# Every symbol include/lmgpu.h declares (checked by the CPU tests).
EXPORTED = [
    'lmgpu_last_error', 'lmgpu_version', 'lmgpu_device_count',
    'lmgpu_accel_create', 'lmgpu_accel_destroy', 'lmgpu_accel_build', 'lmgpu_accel_info_get',
    'lmgpu_accel_intersect', 'lmgpu_accel_intersect_host', 'lmgpu_accel_occluded', 'lmgpu_accel_status',
    'lmgpu_accel_intersect_counted', 'lmgpu_accel_intersect_host_noalloc', 'lmgpu_accel_serialize',
    'lmgpu_accel_deserialize',
    'lmgpu_scene_create', 'lmgpu_scene_destroy', 'lmgpu_scene_accel', 'lmgpu_pt_render',
    'lmgpu_guide_create', 'lmgpu_guide_destroy', 'lmgpu_guide_params', 'lmgpu_guide_infer',
    'lmgpu_guide_sample', 'lmgpu_guide_pdf', 'lmgpu_guide_train', 'lmgpu_guide_stats', 'lmgpu_pt_render_guided',
    'lmgpu_raycast_render', 'lmgpu_lt_render', 'lmgpu_bdpt_render',
    'lmgpu_batch_create', 'lmgpu_batch_destroy', 'lmgpu_batch_add_job', 'lmgpu_batch_run', 'lmgpu_batch_job_film',
]


class LmError(RuntimeError):
    """Error raised for a failing lmgpu call (mirrors LM_THROW_EXCEPTION)."""

    def __init__(self, code: int, msg: str, func: str):
        self.code = code
        super().__init__(f'{func}: {_ERR_NAMES.get(code, code)}: {msg}')


class BuildOpts(ctypes.Structure):
    _fields_ = [('width', ctypes.c_int32), ('leaf_size', ctypes.c_int32), ('refill', ctypes.c_int32),
                ('builder', ctypes.c_int32), ('ploc_radius', ctypes.c_int32), ('collapse', ctypes.c_int32),
                ('sah_tri_cost', ctypes.c_float), ('reserved', ctypes.c_int32 * 1)]


BUILDER_PLOC = 0
BUILDER_LBVH = 1
COLLAPSE_SAH = 0
COLLAPSE_GREEDY = 1


class AccelInfo(ctypes.Structure):
    _fields_ = [
        ('num_tris', ctypes.c_int64), ('num_nodes', ctypes.c_int64), ('num_leaves', ctypes.c_int64),
        ('width', ctypes.c_int32), ('leaf_size', ctypes.c_int32), ('depth', ctypes.c_int32),
        ('stack_capacity', ctypes.c_int32), ('node_bytes', ctypes.c_int64), ('tri_bytes', ctypes.c_int64),
        ('sah_cost', ctypes.c_double), ('build_ms', ctypes.c_double),
    ]


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ('num_tris', ctypes.c_int64),
        ('verts', ctypes.c_void_p),
        ('normals', ctypes.c_void_p),
        ('tri_prim', ctypes.c_void_p),
        ('num_prims', ctypes.c_int32),
        ('mat_type', ctypes.c_void_p),
        ('mat_param', ctypes.c_void_p),
        ('prim_Ke', ctypes.c_void_p),
        ('num_area_lights', ctypes.c_int32),
        ('light_prims', ctypes.c_void_p),
        ('has_env', ctypes.c_int32),
        ('env_Le', ctypes.c_float * 3),
        ('cam_pos', ctypes.c_float * 3),
        ('cam_center', ctypes.c_float * 3),
        ('cam_up', ctypes.c_float * 3),
        ('vfov', ctypes.c_float),
        ('aspect', ctypes.c_float),
        ('w', ctypes.c_int32),
        ('h', ctypes.c_int32),
        ('mat_ks', ctypes.c_void_p),
        ('mat_tex', ctypes.c_void_p),
        ('uvs', ctypes.c_void_p),
        ('num_textures', ctypes.c_int32),
        ('tex_info', ctypes.c_void_p),
        ('tex_data', ctypes.c_void_p),
        ('tex_data_len', ctypes.c_int64),
    ]


class PTParams(ctypes.Structure):
    _fields_ = [
        ('spp', ctypes.c_int32), ('max_verts', ctypes.c_int32), ('sampling_mode', ctypes.c_int32),
        ('rr_depth', ctypes.c_int32), ('seed', ctypes.c_uint64), ('sample_offset', ctypes.c_int32),
        ('guiding', ctypes.c_int32), ('render_time', ctypes.c_double), ('primary_mode', ctypes.c_int32),
        ('reserved0', ctypes.c_int32), ('num_samples', ctypes.c_int64), ('reserved', ctypes.c_int32 * 4),
    ]


class PTStats(ctypes.Structure):
    _fields_ = [
        ('samples', ctypes.c_int64), ('spp', ctypes.c_int64), ('extension_rays', ctypes.c_int64),
        ('shadow_rays', ctypes.c_int64), ('elapsed_s', ctypes.c_double), ('guide_records', ctypes.c_int64),
        ('reserved', ctypes.c_int64 * 3),
    ]


class RaycastParams(ctypes.Structure):
    _fields_ = [('bg_color', ctypes.c_float * 3), ('color', ctypes.c_float * 3), ('use_color', ctypes.c_int32),
                ('use_constant_color', ctypes.c_int32), ('visualize_normal', ctypes.c_int32),
                ('reserved', ctypes.c_int32 * 5)]


class LTParams(ctypes.Structure):
    _fields_ = [('num_samples', ctypes.c_int64), ('max_verts', ctypes.c_int32), ('pass_offset', ctypes.c_int32),
                ('seed', ctypes.c_uint64), ('render_time', ctypes.c_double), ('min_verts', ctypes.c_int32),
                ('reserved', ctypes.c_int32 * 7)]


class JobDesc(ctypes.Structure):
    _fields_ = [('scene', ctypes.c_void_p), ('spp', ctypes.c_int32), ('max_verts', ctypes.c_int32),
                ('sampling_mode', ctypes.c_int32), ('rr_depth', ctypes.c_int32), ('seed', ctypes.c_uint64),
                ('reserved', ctypes.c_int32 * 4)]


class BatchOpts(ctypes.Structure):
    _fields_ = [('width', ctypes.c_int32), ('path_capacity', ctypes.c_int32), ('block', ctypes.c_int32),
                ('leaf_size', ctypes.c_int32), ('film_sumsq', ctypes.c_int32), ('queue_order', ctypes.c_int32),
                ('reserved', ctypes.c_int32 * 2)]


QUEUE_JOB_MAJOR = 0
QUEUE_INTERLEAVED = 1


class BatchStats(ctypes.Structure):
    _fields_ = [('jobs', ctypes.c_int64), ('samples', ctypes.c_int64), ('extension_rays', ctypes.c_int64),
                ('shadow_rays', ctypes.c_int64), ('iterations', ctypes.c_int64), ('elapsed_s', ctypes.c_double),
                ('build_s', ctypes.c_double), ('tune_s', ctypes.c_double), ('width', ctypes.c_int32),
                ('path_capacity', ctypes.c_int32), ('block', ctypes.c_int32), ('reserved0', ctypes.c_int32),
                ('reserved', ctypes.c_int64 * 4)]


class GuideOpts(ctypes.Structure):
    _fields_ = [('alpha', ctypes.c_float), ('learning_rate', ctypes.c_float), ('batch', ctypes.c_int32),
                ('steps_per_pass', ctypes.c_int32), ('record_depth', ctypes.c_int32), ('weight_clamp', ctypes.c_float),
                ('reserved', ctypes.c_int32 * 6)]


_lib = None


def build(force: bool = False) -> str:
    """Compile liblmgpu.so in-tree with hipcc (gfx950)."""
    args = ['make', '-s', '-C', PKG_DIR, f'-j{min(8, os.cpu_count() or 1)}']
    if force:
        args.insert(2, '-B')
    subprocess.check_call(args)
    return LIB_PATH


def lib():
    """The loaded library; raises if it is not built (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f'{LIB_PATH} is not built; run lm_amd._lib.build() '
                               '(make -C lightmetrica-v3_amd) first')
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        i64 = ctypes.c_int64
        L.lmgpu_last_error.restype = ctypes.c_char_p
        L.lmgpu_version.restype = ctypes.c_char_p
        L.lmgpu_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
        L.lmgpu_accel_create.argtypes = [ctypes.POINTER(vp)]
        L.lmgpu_accel_destroy.argtypes = [vp]
        L.lmgpu_accel_build.argtypes = [vp, vp, i64, ctypes.c_int, ctypes.POINTER(BuildOpts), vp]
        L.lmgpu_accel_info_get.argtypes = [vp, ctypes.POINTER(AccelInfo)]
        L.lmgpu_accel_intersect.argtypes = [vp, vp, vp, i64, vp]
        L.lmgpu_accel_intersect_host.argtypes = [vp, vp, vp, i64]
        L.lmgpu_accel_intersect_host_noalloc.argtypes = [vp, vp, vp, i64]
        L.lmgpu_accel_occluded.argtypes = [vp, vp, vp, i64, vp]
        L.lmgpu_accel_status.argtypes = [vp, vp]
        L.lmgpu_accel_intersect_counted.argtypes = [vp, vp, vp, i64, ctypes.POINTER(i64), ctypes.POINTER(i64), vp]
        L.lmgpu_accel_serialize.argtypes = [vp, vp, ctypes.POINTER(ctypes.c_size_t)]
        L.lmgpu_accel_deserialize.argtypes = [vp, vp, ctypes.c_size_t, ctypes.POINTER(BuildOpts)]
        L.lmgpu_scene_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(SceneDesc), ctypes.POINTER(BuildOpts)]
        L.lmgpu_scene_destroy.argtypes = [vp]
        L.lmgpu_scene_accel.argtypes = [vp, ctypes.POINTER(vp)]
        L.lmgpu_pt_render.argtypes = [vp, ctypes.POINTER(PTParams), vp, vp, ctypes.c_int, ctypes.POINTER(PTStats), vp]
        f32p = ctypes.POINTER(ctypes.c_float)
        L.lmgpu_guide_create.argtypes = [ctypes.POINTER(vp), vp, vp, ctypes.c_uint64]
        L.lmgpu_guide_destroy.argtypes = [vp]
        L.lmgpu_guide_params.argtypes = [vp, vp, vp]
        L.lmgpu_guide_infer.argtypes = [vp, vp, vp, i64, vp]
        L.lmgpu_guide_sample.argtypes = [vp, vp, vp, vp, i64, vp]
        L.lmgpu_guide_pdf.argtypes = [vp, vp, vp, i64, vp]
        L.lmgpu_guide_stats.argtypes = [vp, ctypes.POINTER(i64), f32p]
        L.lmgpu_raycast_render.argtypes = [vp, ctypes.POINTER(RaycastParams), vp, ctypes.c_int, vp]
        L.lmgpu_lt_render.argtypes = [vp, ctypes.POINTER(LTParams), vp, ctypes.c_int, ctypes.POINTER(PTStats), vp]
        L.lmgpu_bdpt_render.argtypes = [vp, ctypes.POINTER(LTParams), vp, ctypes.c_int, ctypes.POINTER(PTStats), vp]
        L.lmgpu_batch_create.argtypes = [ctypes.POINTER(vp)]
        L.lmgpu_batch_destroy.argtypes = [vp]
        L.lmgpu_batch_add_job.argtypes = [vp, ctypes.POINTER(JobDesc), ctypes.POINTER(ctypes.c_int32)]
        L.lmgpu_batch_run.argtypes = [vp, ctypes.POINTER(BatchOpts), ctypes.POINTER(BatchStats), vp]
        L.lmgpu_batch_job_film.argtypes = [vp, ctypes.c_int32, vp, vp, ctypes.POINTER(ctypes.c_double)]
        L.lmgpu_pt_render_guided.argtypes = [vp, ctypes.POINTER(PTParams), vp, ctypes.POINTER(GuideOpts), vp, vp,
                                             ctypes.c_int, ctypes.POINTER(PTStats), vp]
        L.lmgpu_guide_train.argtypes = [vp, vp, i64, ctypes.c_float, f32p, vp, vp]
        for name in EXPORTED:
            if name not in ('lmgpu_last_error', 'lmgpu_version'):
                getattr(L, name).restype = ctypes.c_int
        _lib = L
    return _lib


def check(rc: int, func: str):
    if rc != OK:
        msg = lib().lmgpu_last_error()
        raise LmError(rc, msg.decode() if msg else '', func)


def call(func: str, *args):
    check(getattr(lib(), func)(*args), func)


def device_tensor(t, name: str, dtype, cols=None, min_rows: int = 0, shape=None, device=None,
                  cuda: bool = True) -> int:
    """Checks a torch tensor handed to a device entry point and returns its
    data pointer. The kernels behind the C ABI index the buffer by the row
    count they are given, so a wrong dtype, a short or non-contiguous buffer,
    or host memory would be read or written out of bounds on the device:
    every such case raises ValueError here, before any library call.
    `cols`: required size of the last dimension of a 2-D tensor (None for a
    1-D tensor); `min_rows`: rows the call touches; `shape`: an exact shape
    instead; `device`: the torch.device the tensor must live on; `cuda`:
    False defers the device check to on_device()."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f'{name} must be a torch tensor, got {type(t).__name__}')
    if t.dtype != dtype:
        raise ValueError(f'{name} must be {dtype}, got {t.dtype}')
    if shape is not None:
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f'{name} must have shape {tuple(shape)}, got {tuple(t.shape)}')
    elif cols is None:
        if t.dim() != 1 or t.shape[0] < min_rows:
            raise ValueError(f'{name} must be 1-D with at least {min_rows} entries, got {tuple(t.shape)}')
    elif t.dim() != 2 or t.shape[1] != cols or t.shape[0] < min_rows:
        raise ValueError(f'{name} must be [>={min_rows}, {cols}], got {tuple(t.shape)}')
    if not t.is_contiguous():
        raise ValueError(f'{name} must be contiguous')
    if cuda:
        on_device((name, t))
    if device is not None and t.device != device:
        raise ValueError(f'{name} is on {t.device}, expected {device}')
    return t.data_ptr()


def on_device(*named):
    """(name, tensor) pairs: all CUDA (HIP device) tensors on one device."""
    dev = None
    for name, t in named:
        if not t.is_cuda:
            raise ValueError(f'{name} must be a CUDA (HIP device) tensor')
        if dev is not None and t.device != dev:
            raise ValueError(f'{name} is on {t.device}, expected {dev}')
        dev = t.device
