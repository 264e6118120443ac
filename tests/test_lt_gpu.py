"""GPU renderer::lt_mi355x (lmgpu_lt_render) against closed forms, the FP64
oracle of renderer::lt on the same Philox streams, and the path-space
identity LT(M) = PT(M) - PT(2) with the GPU path tracer."""
import math

import numpy as np
import pytest

from lm_amd import _lib, scenes
from lm_amd.renderer import Film, GpuScene, RendererLT
from oracle import oracle as orc

from lt_common import (batch_mean_z, block_z, env_cube_scene, furnace_expected, furnace_scene, glass_lamp_scene,
                       smooth_sphere_scene)

pytestmark = pytest.mark.gpu


def lt_batches(sc, npix, batches, passes, M, seed):
    out = []
    for k in range(batches):
        s, st = sc.render_lt(num_samples=passes * npix, max_verts=M, seed=seed, pass_offset=passes * k)
        assert st['samples'] == passes * npix and st['passes'] == passes
        out.append(s * (npix / st['samples']))
    return np.asarray(out)


def test_lt_furnace_known_answer(device):
    rho, Le, M = 0.5, 1.0, 8
    flat = furnace_scene(rho, Le).flatten()
    sc = GpuScene(flat)
    npix = flat.w * flat.h
    imgs = lt_batches(sc, npix, 16, 16, M, seed=3)
    means = imgs[..., 0].mean((1, 2))
    exp = furnace_expected(rho, Le, M)
    z = batch_mean_z(means, exp)
    print(f'LT furnace {means.mean():.5f} expected {exp:.5f} z {z:.2f}')
    assert abs(z) < 5


def test_lt_env_cube_closed_form(device):
    rho = np.array([0.6, 0.4, 0.2])
    flat = env_cube_scene(tuple(rho)).flatten()
    sc = GpuScene(flat)
    npix = flat.w * flat.h
    imgs = lt_batches(sc, npix, 8, 32, 8, seed=5)
    c = slice(flat.h // 2 - 6, flat.h // 2 + 6)
    vals = imgs[:, c, c].reshape(8, -1, 3).mean(1)
    z = [batch_mean_z(vals[:, i], rho[i]) for i in range(3)]
    print('LT env cube', vals.mean(0), z)
    assert np.all(np.abs(z) < 5)
    assert np.all(imgs[:, 0, 0] == 0) and np.all(imgs[:, -1, -1] == 0)


def test_lt_same_stream_matches_oracle(device):
    """Same Philox streams: the FP32 GPU light paths and the FP64 oracle's
    take the same decisions almost everywhere, so the splat sums agree."""
    flat = scenes.config0().flatten()
    flat.w = flat.h = 64
    npix = flat.w * flat.h
    n, M, seed = 16 * npix, 8, 4321
    sc = GpuScene(flat)
    s, st = sc.render_lt(num_samples=n, max_verts=M, seed=seed)
    ref = orc.lt_render(flat, n, M, seed=seed)
    a, b = s, ref['sum']
    close = np.abs(a - b) <= 1e-3 * (1.0 + np.abs(b))
    frac = close.all(-1).mean()
    rel = abs(a.sum() - b.sum()) / b.sum()
    print(f'LT same-stream: pixels within 1e-3: {frac:.4f}; total rel diff {rel:.2e}; gpu {st}; oracle {ref["stats"]}')
    assert frac >= 0.97
    assert rel < 2e-3


def test_lt_welch_vs_oracle(device):
    """Independent seeds: block means of GPU and oracle LT batches agree."""
    flat = scenes.config0().flatten()
    flat.w = flat.h = 32
    npix = flat.w * flat.h
    K, P, M = 24, 32, 8
    sc = GpuScene(flat)
    g = lt_batches(sc, npix, K, P, M, seed=11)
    acc = orc.Accel(flat.verts, 0)
    o = np.asarray([orc.lt_render(flat, P * npix, M, seed=12, pass_offset=P * k, accel=acc)['image']
                    for k in range(K)])
    bs = 4
    gb = g.reshape(K, 8, bs, 8, bs, 3).sum((2, 4))
    ob = o.reshape(K, 8, bs, 8, bs, 3).sum((2, 4))
    se = np.sqrt(gb.var(0, ddof=1) / K + ob.var(0, ddof=1) / K)
    ok = se > 0
    z = (gb.mean(0) - ob.mean(0))[ok] / se[ok]
    print(f'LT GPU vs oracle: {ok.sum()} blocks, max|z| {np.abs(z).max():.2f}, mean z^2 {np.mean(z ** 2):.2f}')
    assert ok.sum() >= 150
    assert np.abs(z).max() < 5.5
    assert np.mean(z ** 2) < 2.0


@pytest.mark.parametrize('name', ['config0', 'glass_lamp', 'smooth_sphere'])
def test_lt_equals_pt_minus_direct(device, name):
    make = {'config0': scenes.config0, 'glass_lamp': glass_lamp_scene, 'smooth_sphere': smooth_sphere_scene}[name]
    flat = make().flatten()
    flat.w = flat.h = 64
    flat.camera.aspect = 1.0
    npix = flat.w * flat.h
    M, spp = 8, 4096
    sc = GpuScene(flat)
    s, q, _ = sc.render_raw(spp=spp, max_verts=M, seed=1)
    s2, _, _ = sc.render_raw(spp=spp, max_verts=2, seed=1)
    lts = lt_batches(sc, npix, 32, 64, M, seed=101)
    z, nb = block_z(lts, s / spp, q, spp, s2 / spp, block=8)
    zt, _ = block_z(lts, s / spp, q, spp, s2 / spp, block=flat.w)
    print(f'LT GPU {name}: mean {lts.mean():.5f} vs PT-PT2 {(s - s2).mean() / spp:.5f} (z {zt}); {nb} blocks, '
          f'max|z| {np.abs(z).max():.2f}, mean z^2 {np.mean(z ** 2):.2f}')
    assert nb >= 100
    assert np.all(np.abs(zt) < 4.0)
    assert np.abs(z).max() < 5.5
    assert np.mean(z ** 2) < 2.0


def test_renderer_lt_api(device):
    """RendererLT mirrors renderer::lt: film = splat sum * w*h / processed;
    scheduler::spi::time renders whole passes; device films accumulate."""
    import torch
    flat = scenes.config0().flatten()
    flat.w = flat.h = 32
    sc = GpuScene(flat)
    film = Film(32, 32)
    r = RendererLT(sc, film, max_verts=6, seed=7, scheduler='sample', num_samples=5000)
    out = r.render()
    assert out['processed'] == 5000 and out['stats']['passes'] == 5
    s, _ = sc.render_lt(num_samples=5000, max_verts=6, seed=7)
    assert np.allclose(film.data, s * (1024 / 5000), rtol=1e-12, atol=0)
    # the same two passes, one call or two progressive ones, on a device film
    f = torch.zeros((32, 32, 3), dtype=torch.float64, device=device)
    sc.render_lt(num_samples=1024, max_verts=6, seed=7, film=f)
    sc.render_lt(num_samples=1024, max_verts=6, seed=7, pass_offset=1, film=f)
    both, _ = sc.render_lt(num_samples=2048, max_verts=6, seed=7)
    assert np.allclose(f.cpu().numpy(), both, rtol=1e-9, atol=1e-12)
    t = RendererLT(sc, film, max_verts=6, seed=7, scheduler='time', render_time=0.2).render()
    assert t['processed'] % 1024 == 0 and t['processed'] >= 1024
    # the time scheduler renders chunks of whole passes on consecutive pass indices:
    # its film equals a num_samples render of as many light paths
    ts, tst = sc.render_lt(max_verts=6, seed=8, render_time=0.3)
    assert tst['samples'] == tst['passes'] * 1024 and tst['elapsed'] >= 0.3
    fs_, _ = sc.render_lt(num_samples=tst['samples'], max_verts=6, seed=8)
    assert np.allclose(ts, fs_, rtol=1e-9, atol=1e-12)
    with pytest.raises(_lib.LmError):
        sc.render_lt(num_samples=10, max_verts=1)
    with pytest.raises(_lib.LmError):
        sc.render_lt(num_samples=0, max_verts=6)
    with pytest.raises(_lib.LmError, match='beyond'):
        sc.render_lt(num_samples=1024 * 2**32 + 1, max_verts=6)
    with pytest.raises(ValueError):
        RendererLT(sc, film, max_verts=6, scheduler='sample')
