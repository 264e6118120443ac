"""renderer::pt with primary_ray_sampling_mode = "image" in the FP64 oracle
(renderer_pt.cpp:96-107, 130-136, 147-152, 209-212, 307-312): raster
positions uniform over the whole image, contributions splatted at the
primary ray's raster position, the camera vertex sampled by NEE (Camera
J(wo) times the light's weight, splatted at the light's raster position)
with MIS against hitting the light from the camera. Pinned by the closed
form of an environment-only view (every pixel = Le, so the image mean is Le
whatever the splat pattern) and by agreement with the pinned pixel-mode
oracle pixel block by pixel block (E[image mode] = E[pixel mode] per pixel,
since both integrate the same measurement over the pixel's footprint)."""
import numpy as np
import pytest

from lm_amd import scenes
from oracle import oracle as orc

from lt_common import env_cube_scene, glass_lamp_scene, mixture_room_scene


def env_only_scene(Le=(0.7, 0.5, 0.3), w=16, h=12):
    """A diffuse cube behind the camera under a constant environment: every
    camera ray escapes, so every pixel's value is Le."""
    c = scenes.cube(0.5, 0.0, (0, 0, 6.0))
    cam = scenes.Camera((0, 0, 4.0), (0, 0, 0), (0, 1, 0), 30.0, w / h)
    return scenes.Scene([scenes.Primitive(c, scenes.Material(scenes.MAT_DIFFUSE, (0.5, 0.5, 0.5)))], cam, Le, w, h)


@pytest.mark.parametrize('mode', [0, 1, 2])
def test_image_mode_env_only_closed_form(mode):
    """Naive: only direct hits (weight 1); NEE: only the camera vertex's
    light sample (J Le / p_light inside the frustum); MIS: both, balance
    weights. Each has image mean Le; with NEE and MIS the light sample also
    lands outside the frustum (dropped), so a wrong J or p_light biases it."""
    Le = np.array([0.7, 0.5, 0.3])
    flat = env_only_scene(tuple(Le)).flatten()
    acc = orc.Accel(flat.verts, 0)
    means = np.asarray([orc.pt_render(flat, 32, 4, sampling_mode=mode, seed=3, accel=acc, sample_offset=32 * k,
                                      primary_mode='image')['image'].mean((0, 1)) for k in range(16)])
    if mode == 0:
        # naive: every sample adds exactly Le (as the scene stores it, float32) to one pixel
        assert np.allclose(means, Le.astype(np.float32), rtol=1e-12)
        return
    se = means.std(0, ddof=1) / np.sqrt(len(means))
    z = (means.mean(0) - Le) / se
    print(f'env only, mode {mode}: mean {means.mean(0)} expected {Le} z {z}')
    assert np.all(np.abs(z) < 4)


def test_image_mode_pixel_sums_are_sample_counts():
    """Naive mode under an environment alone: a pixel's sum is Le times the
    number of samples whose uniform raster position fell in it, and the
    counts over the image add up to w*h*spp."""
    Le = (1.0, 1.0, 1.0)
    flat = env_only_scene(Le, 8, 8).flatten()
    r = orc.pt_render(flat, 64, 3, sampling_mode=0, seed=9, primary_mode='image')
    counts = r['sum'][..., 0]
    assert np.allclose(counts, np.round(counts)) and counts.sum() == 64 * 64
    # roughly uniform: chi-square of 64 cells with mean 64
    chi2 = ((counts - 64) ** 2 / 64).sum()
    assert chi2 < 120
    assert 'sumsq' not in r


def block_z(batches, ref_img, ref_sumsq, spp, block):
    K = batches.shape[0]
    h, w = ref_img.shape[:2]

    def bs(a):
        return a.reshape(a.shape[:-3] + (h // block, block, w // block, block, 3)).sum((-4, -2))

    gb = bs(batches)
    se = np.sqrt(gb.var(0, ddof=1) / K + bs(np.maximum(ref_sumsq / spp - ref_img ** 2, 0) / spp))
    ok = se > 0
    return (gb.mean(0) - bs(ref_img))[ok] / se[ok]


SCENES = {'config0': scenes.config0, 'mixture_room': mixture_room_scene, 'env_cube': env_cube_scene,
          'glass_lamp': glass_lamp_scene}


@pytest.mark.parametrize('name', list(SCENES))
@pytest.mark.parametrize('mode', [1, 2])
def test_image_mode_equals_pixel_mode(name, mode):
    flat = SCENES[name]().flatten()
    flat.w = flat.h = 24
    flat.camera.aspect = 1.0
    acc = orc.Accel(flat.verts, 0)
    spp, M = (4096 if name == 'glass_lamp' else 1024), 6
    ref = orc.pt_render(flat, spp, M, sampling_mode=mode, seed=1, accel=acc)
    K, per = 24, 64
    b = np.asarray([orc.pt_render(flat, per, M, sampling_mode=mode, seed=5, accel=acc, sample_offset=per * k,
                                  primary_mode='image')['image'] for k in range(K)])
    z = block_z(b, ref['image'], ref['sumsq'], spp, 4)
    zt = block_z(b, ref['image'], ref['sumsq'], spp, 24)
    print(f'image vs pixel mode, {name}, mode {mode}: mean {b.mean():.5f} vs {ref["image"].mean():.5f} (z {zt}); '
          f'max|z| {np.abs(z).max():.2f}, mean z^2 {np.mean(z ** 2):.2f}')
    assert np.all(np.abs(zt) < 4)
    assert np.abs(z).max() < 5 and np.mean(z ** 2) < 2.0


def test_image_mode_num_samples_passes():
    """scheduler::spi::sample with num_samples not a multiple of w*h: walk j
    draws the stream (j mod w*h, offset + j div w*h), so a render of
    3 w*h + 50 samples is the sum of a 3-pass render and a 50-sample render
    starting at pass 3; the film value is sum * w*h / num_samples."""
    flat = scenes.config0().flatten()
    flat.w, flat.h = 12, 10
    acc = orc.Accel(flat.verts, 0)
    npix = flat.w * flat.h
    a = orc.pt_render(flat, 0, 5, seed=4, accel=acc, primary_mode='image', num_samples=3 * npix + 50)
    b = orc.pt_render(flat, 0, 5, seed=4, accel=acc, primary_mode='image', num_samples=3 * npix)
    c = orc.pt_render(flat, 0, 5, seed=4, accel=acc, primary_mode='image', num_samples=50, sample_offset=3)
    assert np.allclose(a['sum'], b['sum'] + c['sum'], rtol=1e-12, atol=1e-14)
    assert a['stats']['samples'] == 3 * npix + 50
    assert np.allclose(a['image'], a['sum'] * npix / (3 * npix + 50))


def test_image_mode_rejects_sumsq_and_bad_mode():
    flat = env_only_scene().flatten()
    with pytest.raises(ValueError):
        orc.pt_render(flat, 1, 3, primary_mode='film')
    # sums of squares are dropped in image mode
    assert 'sumsq' not in orc.pt_render(flat, 1, 3, primary_mode='image', sumsq=True)
