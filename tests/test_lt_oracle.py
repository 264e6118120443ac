"""The FP64 oracle of renderer::lt (renderer_lt.cpp:56-170) pinned on the CPU:
closed forms (furnace, convex object under a constant environment) and the
path-space identity LT(M) = PT(M) - PT(2) against the separately pinned PT
oracle, on scenes that exercise area lights, glass refraction under
importance transport and interpolated shading normals."""
import numpy as np
import pytest

from lm_amd import scenes
from oracle import oracle as orc

from lt_common import (batch_mean_z, block_z, env_cube_scene, furnace_expected, furnace_scene, glass_lamp_scene,
                       masked_quad_expected, masked_quad_scene, mixture_room_scene, smooth_sphere_scene)


def test_lt_furnace_closed_form():
    rho, Le, M = 0.5, 1.0, 8
    flat = furnace_scene(rho, Le).flatten()
    acc = orc.Accel(flat.verts, 0)
    npix = flat.w * flat.h
    means = [orc.lt_render(flat, 4 * npix, M, seed=3, pass_offset=4 * k, accel=acc)['image'][..., 0].mean()
             for k in range(16)]
    exp = furnace_expected(rho, Le, M)
    z = batch_mean_z(means, exp)
    print(f'LT oracle furnace {np.mean(means):.5f} expected {exp:.5f} z {z:.2f}')
    assert abs(z) < 5


def test_lt_env_cube_closed_form():
    """light::envconst rays (a disk on the scene's bounding sphere) onto a
    convex diffuse cube: L = rho * Le on the cube; the directly seen
    environment (two vertices) is not sampled, so the corners stay 0."""
    rho = np.array([0.6, 0.4, 0.2])
    flat = env_cube_scene(tuple(rho)).flatten()
    acc = orc.Accel(flat.verts, 0)
    npix = flat.w * flat.h
    c = slice(flat.h // 2 - 6, flat.h // 2 + 6)
    vals = [orc.lt_render(flat, 16 * npix, 8, seed=5, pass_offset=16 * k, accel=acc)['image'][c, c]
            .reshape(-1, 3).mean(0) for k in range(8)]
    vals = np.asarray(vals)
    z = [batch_mean_z(vals[:, i], rho[i]) for i in range(3)]
    print('LT oracle env cube', vals.mean(0), z)
    assert np.all(np.abs(z) < 5)
    r = orc.lt_render(flat, npix, 8, seed=5, accel=acc)
    assert np.all(r['image'][0, 0] == 0) and np.all(r['image'][-1, -1] == 0)


def _identity(flat, spp, batches, per_batch, seed=1, M=8):
    acc = orc.Accel(flat.verts, 0)
    npix = flat.w * flat.h
    pt = orc.pt_render(flat, spp, M, seed=seed, accel=acc)
    pt2 = orc.pt_render(flat, spp, 2, seed=seed, accel=acc, sumsq=False)
    lts = [orc.lt_render(flat, per_batch * npix, M, seed=seed + 100, pass_offset=per_batch * k,
                         accel=acc)['image'] for k in range(batches)]
    z, nb = block_z(lts, pt['image'], pt['sumsq'], spp, pt2['image'])
    zt, _ = block_z(lts, pt['image'], pt['sumsq'], spp, pt2['image'], block=flat.w)
    return z, nb, zt, np.mean(lts, 0).mean(), (pt['image'] - pt2['image']).mean()


SCENES = {'config0': scenes.config0, 'glass_lamp': glass_lamp_scene, 'smooth_sphere': smooth_sphere_scene,
          'mixture_room': mixture_room_scene}


@pytest.mark.parametrize('name', list(SCENES))
def test_lt_equals_pt_minus_direct(name):
    """mixture_room: material::mixture_wavefrontobj walls and cube and a
    texture::bitmap floor, in both estimators."""
    make = SCENES[name]
    flat = make().flatten()
    flat.w = flat.h = 24
    flat.camera.aspect = 1.0
    z, nb, zt, lt, ref = _identity(flat, 4096, 32, 128)
    print(f'LT oracle {name}: mean {lt:.5f} vs PT-PT2 {ref:.5f} (z {zt}); {nb} blocks, '
          f'max|z| {np.abs(z).max():.2f}, mean z^2 {np.mean(z ** 2):.2f}')
    assert nb >= 60
    assert np.all(np.abs(zt) < 4.0)
    assert np.abs(z).max() < 5.0
    # 36 blocks x 3 correlated channels; z of a 32-batch variance estimate
    assert np.mean(z ** 2) < 2.0


def test_pt_masked_texture_quad_closed_form():
    """PT oracle on a textured mixture quad with a binary alpha mask under a
    constant environment: texel Kd where alpha = 1, the environment (exactly)
    where alpha = 0, on every pixel whose footprint lies in one texel."""
    s, tex = masked_quad_scene()
    flat = s.flatten()
    spp = 64
    r = orc.pt_render(flat, spp, 8, seed=3)
    exp, same = masked_quad_expected(flat, tex)
    img = r['image']
    se = np.sqrt(np.maximum(r['sumsq'] / spp - img ** 2, 0) / spp)
    hole = same & (exp == 1.0).all(-1)
    assert hole.sum() > 100 and (same & ~hole).sum() > 100
    assert np.all(img[hole] == 1.0)
    z = (img - exp)[same & ~hole] / se[same & ~hole]
    print(f'masked quad: {same.sum()} pixels, max|z| {np.abs(z).max():.2f}, mean z^2 {np.mean(z ** 2):.2f}')
    assert np.abs(z).max() < 5 and np.mean(z ** 2) < 1.5
