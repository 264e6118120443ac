"""renderer::raycast on the GPU accel, and the reference's accel-consistency
harness (functest/func_accel_consistency.py:54-116): the image of every GPU
accel variant against the image of accel::sahbvh (the FP64 oracle's SAH
BVH), pixel-wise RMSE ~ 0 outside the edge/tie set; plus the C++ host
components' own test program."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, cached

from lm_amd import scenes
from lm_amd.renderer import Film, GpuScene, RendererRaycast
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def reference_raycast(flat, bg, mode='shade', color=None):
    """Ray cast of the flattened scene with the oracle's accel::sahbvh, shaded
    as renderer::raycast (renderer_raycast.cpp:72-95). Returns the image and
    the per-pixel edge/tie flags."""
    o, d = scenes.primary_rays(flat.camera, flat.w, flat.h)
    n = o.shape[0]
    rays = {'o': o, 'd': d, 'tmin': np.full(n, 1e-3, np.float32), 'tmax': np.full(n, 1e10, np.float32)}
    oacc = cached(('raycast_oacc', id(flat)), lambda: orc.Accel(flat.verts))
    ref = oacc.intersect(rays)
    flags = oacc.classify(rays, ref)
    img = np.tile(np.asarray(bg, dtype=np.float64), (n, 1))
    if flat.has_env:
        img[:] = 0.0 if mode == 'normal' else np.asarray(flat.env_Le, dtype=np.float64)
    hit = ref['prim'] >= 0
    tri = ref['prim'][hit]
    u, v = ref['u'][hit][:, None], ref['v'][hit][:, None]
    nv = np.asarray(flat.normals, dtype=np.float64).reshape(-1, 3, 3)[tri]
    nrm = nv[:, 0] * (1 - u - v) + nv[:, 1] * u + nv[:, 2] * v
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    if mode == 'normal':
        C = np.abs(nrm)
    else:
        prim = np.asarray(flat.tri_prim)[tri]
        mt = np.asarray(flat.mat_type)[prim]
        R = np.asarray(flat.mat_param, dtype=np.float64)[prim, :3] * np.isin(mt, [0, 1])[:, None]
        if color is not None:
            R = np.tile(np.asarray(color, dtype=np.float64), (tri.shape[0], 1))
        dd = d[hit].astype(np.float64)
        C = R if mode == 'constant' else R * (0.2 + 0.8 * np.abs((nrm * -dd).sum(1)))[:, None]
    img[hit] = C
    return img.reshape(flat.h, flat.w, 3), flags.reshape(flat.h, flat.w)


def c2_small():
    return cached('raycast_c2', lambda: scenes.config2(width=320, height=180).flatten())


@pytest.mark.parametrize('width,builder', [(2, 1), (4, 1), (8, 1), (4, 0)])
def test_accel_consistency_c2(device, width, builder):
    """Config 2 (1.32M triangles, all materials, env light) at 320x180."""
    flat = c2_small()
    ref, flags = cached(('raycast_ref_c2',), lambda: reference_raycast(flat, (0.1, 0.2, 0.3)))
    sc = GpuScene(flat, width=width, builder=builder)
    film = Film(flat.w, flat.h)
    RendererRaycast(sc, film, bg_color=(0.1, 0.2, 0.3)).render()
    diff = np.sqrt(((film.data - ref) ** 2).sum(-1) / 3)
    clean = flags == 0
    rmse = float(np.sqrt(np.mean((film.data - ref) ** 2)))
    print(f'W{width} builder {builder}: rmse {rmse:.3e}, flagged {1 - clean.mean():.4f}, '
          f'max clean diff {diff[clean].max():.2e}, clean pixels off {(diff[clean] > 1e-4).sum()}')
    assert (diff[clean] > 1e-4).sum() == 0
    assert rmse < 1e-2


@pytest.mark.parametrize('mode', ['shade', 'normal', 'constant'])
def test_raycast_modes_c0(device, mode):
    flat = scenes.config0().flatten()
    flat.w, flat.h = 160, 120
    flat.camera.aspect = 160 / 120
    color = (0.4, 0.5, 0.6) if mode == 'constant' else None
    ref, flags = reference_raycast(flat, (0.05, 0.05, 0.05), mode, color)
    film = Film(flat.w, flat.h)
    RendererRaycast(GpuScene(flat), film, bg_color=(0.05, 0.05, 0.05), visualize_normal=mode == 'normal',
                    use_constant_color=mode == 'constant', color=color).render()
    diff = np.abs(film.data - ref).max(-1)
    clean = flags == 0
    print(f'{mode}: flagged {1 - clean.mean():.4f}, max clean diff {diff[clean].max():.2e}')
    assert (diff[clean] > 1e-4).sum() == 0


def test_host_components_program(device):
    """The C++ host components (accel::mi355x, renderer::pt_mi355x,
    renderer::raycast_mi355x over the C ABI) pass their own tests, run as a
    child process."""
    exe = os.path.join(ROOT, 'lightmetrica-v3_amd', 'bin', 'test_host')
    assert os.path.exists(exe), 'bin/test_host not built (make -C lightmetrica-v3_amd)'
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout[-6000:])
    print(r.stderr[-2000:])
    assert r.returncode == 0
