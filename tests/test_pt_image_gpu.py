"""GPU renderer::pt with primary_ray_sampling_mode = "image"
(renderer_pt.cpp:96-107, 130-188, 209-212, 307-312; scheduler::spi::sample,
scheduler.cpp): same-stream parity with the FP64 oracle (same Philox streams:
walk j of a render draws (j mod w*h, j div w*h), raster uniform over the
image, the camera vertex's light sample splatted at the light's raster
position), agreement with the pixel sampling mode, the renderer API and the
errors."""
import numpy as np
import pytest

from lm_amd import _lib, scenes
from lm_amd.renderer import Film, GpuScene, RendererPT
from oracle import oracle as orc

from lt_common import mixture_room_scene

pytestmark = pytest.mark.gpu


def room_env():
    sc = scenes.config0()
    sc.env_Le = (0.6, 0.7, 0.8)   # the open front lets the environment in
    return sc


SCENES = {'config0': scenes.config0, 'mixture_room': mixture_room_scene, 'room_env': room_env}


@pytest.mark.parametrize('name', list(SCENES))
@pytest.mark.parametrize('mode', [2, 1, 0])
def test_image_mode_same_stream_matches_oracle(device, name, mode):
    flat = SCENES[name]().flatten()
    flat.w, flat.h = 48, 40
    flat.camera.aspect = 48 / 40
    npix = flat.w * flat.h
    ns, M, seed = 8 * npix + 77, 7, 31   # 8 full passes and one of 77 samples
    s, q, st = GpuScene(flat).render_raw(max_verts=M, sampling_mode=mode, seed=seed, primary_mode='image',
                                         num_samples=ns)
    assert q is None and st['samples'] == ns and st['spp'] == 9
    ref = orc.pt_render(flat, 0, M, sampling_mode=mode, seed=seed, primary_mode='image', num_samples=ns)
    b = ref['sum']
    close = np.abs(s - b) <= 1e-3 * (1.0 + np.abs(b))
    frac = close.all(-1).mean()
    rel = abs(s.sum() - b.sum()) / b.sum()
    print(f'image mode {name} mode {mode} same-stream: pixels within 1e-3 {frac:.4f}, total rel diff {rel:.2e}; '
          f'gpu {st}; oracle {ref["stats"]}')
    assert frac >= 0.97 and rel < 2e-3


@pytest.mark.parametrize('name', ['config0', 'room_env'])
def test_image_mode_equals_pixel_mode(device, name):
    """E[image mode] = E[pixel mode] per pixel: block z-tests of K image-mode
    batches against a long pixel-mode render with its per-pixel variance."""
    flat = SCENES[name]().flatten()
    flat.w = flat.h = 64
    flat.camera.aspect = 1.0
    npix = flat.w * flat.h
    M, spp = 8, 4096
    sc = GpuScene(flat)
    s, q, _ = sc.render_raw(spp=spp, max_verts=M, seed=1)
    img = s / spp
    K, per = 32, 64
    b = []
    for k in range(K):
        f, _, st = sc.render_raw(max_verts=M, seed=9, sample_offset=per * k, primary_mode='image',
                                 num_samples=per * npix)
        b.append(f * (npix / st['samples']))
    b = np.asarray(b)

    def bs(a, blk):
        return a.reshape(a.shape[:-3] + (64 // blk, blk, 64 // blk, blk, 3)).sum((-4, -2))

    res = {}
    for blk in (8, 64):
        gb = bs(b, blk)
        se = np.sqrt(gb.var(0, ddof=1) / K + bs(np.maximum(q / spp - img ** 2, 0) / spp, blk))
        ok = se > 0
        res[blk] = (gb.mean(0) - bs(img, blk))[ok] / se[ok]
    z, zt = res[8], res[64]
    print(f'image vs pixel mode {name}: mean {b.mean():.5f} vs {img.mean():.5f} (z {zt}); '
          f'max|z| {np.abs(z).max():.2f}, mean z^2 {np.mean(z ** 2):.2f}')
    assert np.all(np.abs(zt) < 4.5)
    assert np.abs(z).max() < 5.5 and np.mean(z ** 2) < 2.0


def test_image_mode_renderer_api_and_errors(device):
    import torch
    flat = scenes.config0().flatten()
    flat.w = flat.h = 32
    sc = GpuScene(flat)
    film = Film(32, 32)
    out = RendererPT(sc, film, max_verts=6, seed=7, primary_ray_sampling_mode='image', num_samples=3000).render()
    assert out['processed'] == 3000 and out['stats']['spp'] == 3 and film.sumsq is None
    s, _, _ = sc.render_raw(max_verts=6, seed=7, primary_mode='image', num_samples=3000)
    assert np.allclose(film.data, s * (1024 / 3000), rtol=1e-9, atol=1e-12)
    # device film, progressive: two calls with consecutive sample offsets equal one call
    f = torch.zeros((32, 32, 3), dtype=torch.float64, device=device)
    sc.render_raw(max_verts=6, seed=7, film_sum=f, primary_mode='image', num_samples=1024)
    sc.render_raw(max_verts=6, seed=7, film_sum=f, sample_offset=1, primary_mode='image', num_samples=1024)
    both, _, _ = sc.render_raw(max_verts=6, seed=7, primary_mode='image', num_samples=2048)
    assert np.allclose(f.cpu().numpy(), both, rtol=1e-9, atol=1e-12)
    t = RendererPT(sc, film, max_verts=6, seed=7, scheduler='time', render_time=0.2,
                   primary_ray_sampling_mode='image').render()
    assert t['processed'] % 1024 == 0 and t['processed'] >= 1024
    with pytest.raises(ValueError):
        RendererPT(sc, film, max_verts=6, primary_ray_sampling_mode='image', spp=4)   # spi::sample needs num_samples
    with pytest.raises(ValueError):
        RendererPT(sc, film, max_verts=6, spp=4, primary_ray_sampling_mode='film')
    with pytest.raises(_lib.LmError):
        sc.render_raw(max_verts=6, primary_mode='image', num_samples=0)
    g = torch.zeros((32, 32, 3), dtype=torch.float64, device=device)
    with pytest.raises(ValueError):
        sc.render_raw(max_verts=6, film_sum=g, film_sumsq=g.clone(), primary_mode='image', num_samples=10)
