"""GPU renderer::bdpt_mi355x (lmgpu_bdpt_render): same-stream parity with the
FP64 oracle of renderer::bdpt (whose MIS weights are the explicit
power-heuristic sums over all strategies, while the kernels accumulate them
along the subpaths), the furnace closed form, agreement with the GPU path
tracer, and the renderer API."""
import numpy as np
import pytest

from lm_amd import _lib, scenes
from lm_amd.renderer import Film, GpuScene, RendererBDPT
from oracle import oracle as orc

from lt_common import (batch_mean_z, env_cube_scene, furnace_scene, glass_lamp_scene, mixture_room_scene,
                       smooth_sphere_scene)

pytestmark = pytest.mark.gpu


def bdpt_batches(sc, npix, K, per, M, seed, min_verts=2):
    out = []
    for k in range(K):
        s, st = sc.render_bdpt(num_samples=per * npix, max_verts=M, seed=seed, pass_offset=per * k,
                               min_verts=min_verts)
        out.append(s * (npix / st['samples']))
    return np.asarray(out)


def room_env():
    sc = scenes.config0()
    sc.env_Le = (0.6, 0.7, 0.8)   # the open front lets the environment in
    return sc


def mixture_env():
    sc = mixture_room_scene()
    sc.env_Le = (0.5, 0.5, 0.5)
    return sc


SAME_STREAM = {'config0': scenes.config0, 'mixture_room': mixture_room_scene, 'room_env': room_env,
               'mixture_env': mixture_env}


@pytest.mark.parametrize('name', list(SAME_STREAM))
def test_bdpt_same_stream_matches_oracle(device, name):
    """Same Philox streams as the oracle; room_env and mixture_env add
    light::envconst (light paths from the bounding-sphere disk, eye paths
    ending on the environment)."""
    flat = SAME_STREAM[name]().flatten()
    flat.w = flat.h = 48
    flat.camera.aspect = 1.0
    npix = flat.w * flat.h
    n, M, seed = 8 * npix, 7, 2024
    s, st = GpuScene(flat).render_bdpt(num_samples=n, max_verts=M, seed=seed)
    ref = orc.bdpt_render(flat, n, M, seed=seed)
    b = ref['sum']
    close = np.abs(s - b) <= 1e-3 * (1.0 + np.abs(b))
    frac = close.all(-1).mean()
    rel = abs(s.sum() - b.sum()) / b.sum()
    print(f'BDPT {name} same-stream: pixels within 1e-3 {frac:.4f}, total rel diff {rel:.2e}; gpu {st}; '
          f'oracle {ref["stats"]}')
    assert frac >= 0.97 and rel < 2e-3


def test_bdpt_same_stream_pass_size(device):
    """Passes of 2^20 samples (max_verts 12: 2^21 paths would store more
    than 2 GiB of light vertices) on a 1200x900 image: sample i of a render
    is path i mod 2^20 of pass i div 2^20 on the GPU and in the oracle alike,
    so a render of 2^20 + 4096 samples stays on the same streams."""
    flat = scenes.config0().flatten()
    flat.w, flat.h = 1200, 900
    flat.camera.aspect = 1200 / 900
    n, M, seed = (1 << 20) + 4096, 12, 77
    s, st = GpuScene(flat).render_bdpt(num_samples=n, max_verts=M, seed=seed)
    assert st['passes'] == 2
    ref = orc.bdpt_render(flat, n, M, seed=seed)
    b = ref['sum']
    close = np.abs(s - b) <= 1e-3 * (1.0 + np.abs(b))
    frac = close.all(-1).mean()
    rel = abs(s.sum() - b.sum()) / b.sum()
    print(f'BDPT 1200x900 M=12 same-stream: pixels within 1e-3 {frac:.4f}, total rel diff {rel:.2e}; gpu {st}')
    assert frac >= 0.97 and rel < 2e-3


def test_bdpt_furnace_closed_form(device):
    rho, Le, M = 0.5, 1.0, 8
    flat = furnace_scene(rho, Le).flatten()
    sc = GpuScene(flat)
    npix = flat.w * flat.h
    means = bdpt_batches(sc, npix, 16, 4, M, 3)[..., 0].mean((1, 2))
    exp = Le * sum(rho ** k for k in range(M - 1))
    z = batch_mean_z(means, exp)
    print(f'BDPT furnace {means.mean():.5f} expected {exp:.5f} z {z:.2f}')
    assert abs(z) < 5


SCENES = {'config0': scenes.config0, 'glass_lamp': glass_lamp_scene, 'smooth_sphere': smooth_sphere_scene,
          'mixture_room': mixture_room_scene, 'room_env': room_env, 'env_cube': env_cube_scene}


@pytest.mark.parametrize('name', list(SCENES))
def test_bdpt_equals_pt(device, name):
    flat = SCENES[name]().flatten()
    flat.w = flat.h = 64
    flat.camera.aspect = 1.0
    npix = flat.w * flat.h
    M, spp = 8, 8192
    sc = GpuScene(flat)
    s, q, _ = sc.render_raw(spp=spp, max_verts=M, seed=1)
    img = s / spp
    b = bdpt_batches(sc, npix, 32, 32, M, 5)
    K = b.shape[0]

    def bs(a, blk):
        return a.reshape(a.shape[:-3] + (64 // blk, blk, 64 // blk, blk, 3)).sum((-4, -2))

    res = {}
    for blk in (8, 64):
        gb = bs(b, blk)
        se = np.sqrt(gb.var(0, ddof=1) / K + bs(np.maximum(q / spp - img ** 2, 0) / spp, blk))
        ok = se > 0
        res[blk] = (gb.mean(0) - bs(img, blk))[ok] / se[ok]
    z, zt = res[8], res[64]
    print(f'BDPT {name}: mean {b.mean():.5f} vs PT {img.mean():.5f} (z {zt}); max|z| {np.abs(z).max():.2f}, '
          f'mean z^2 {np.mean(z ** 2):.2f}')
    assert np.all(np.abs(zt) < 4.5)
    assert np.abs(z).max() < 5.5 and np.mean(z ** 2) < 2.0


def test_renderer_bdpt_api(device):
    import torch
    flat = scenes.config0().flatten()
    flat.w = flat.h = 32
    sc = GpuScene(flat)
    film = Film(32, 32)
    out = RendererBDPT(sc, film, max_verts=6, seed=7, scheduler='sample', num_samples=3000).render()
    assert out['processed'] == 3000 and out['stats']['passes'] == 3
    s, _ = sc.render_bdpt(num_samples=3000, max_verts=6, seed=7)
    assert np.allclose(film.data, s * (1024 / 3000), rtol=1e-9, atol=1e-12)
    # progressive: two calls with consecutive pass offsets equal one call
    f = torch.zeros((32, 32, 3), dtype=torch.float64, device=device)
    sc.render_bdpt(num_samples=1024, max_verts=6, seed=7, film=f)
    sc.render_bdpt(num_samples=1024, max_verts=6, seed=7, pass_offset=1, film=f)
    both, _ = sc.render_bdpt(num_samples=2048, max_verts=6, seed=7)
    assert np.allclose(f.cpu().numpy(), both, rtol=1e-9, atol=1e-12)
    t = RendererBDPT(sc, film, max_verts=6, seed=7, scheduler='time', render_time=0.2).render()
    assert t['processed'] % 1024 == 0 and t['processed'] >= 1024
    with pytest.raises(_lib.LmError):
        sc.render_bdpt(num_samples=10, max_verts=1)
