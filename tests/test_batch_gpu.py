"""Batched multi-job rendering (lmgpu_batch_*): every job of a batch equals
the single-scene render (lmgpu_pt_render) of the same scene, seed and BVH
width up to the FP64 order of accumulation, whatever the path capacity,
block size or autotuned choice; plus latency bookkeeping and errors."""
import numpy as np
import pytest

from conftest import cached

from lm_amd import scenes
from lm_amd.batch import Batch
from lm_amd.renderer import GpuScene

pytestmark = pytest.mark.gpu


def small_jobs():
    def make():
        jobs = []
        f0 = scenes.config0().flatten()
        f0.w, f0.h = 64, 48
        f0.camera.aspect = 64 / 48
        jobs.append((f0, dict(spp=6, max_verts=8, sampling_mode=2, seed=3)))
        f1 = scenes.config0(seed=77).flatten()
        f1.w, f1.h = 40, 40
        jobs.append((f1, dict(spp=5, max_verts=6, sampling_mode=1, seed=9)))
        j = {'job': 0, 'tris': 30000, 'res': 56, 'spp': 4, 'seed': 1234}
        f2 = scenes.config3_scene(j).flatten()
        jobs.append((f2, dict(spp=4, max_verts=16, sampling_mode=2, seed=5)))
        f3 = scenes.config0(seed=5).flatten()
        f3.w, f3.h = 33, 17
        f3.camera.aspect = 33 / 17
        jobs.append((f3, dict(spp=3, max_verts=8, sampling_mode=0, seed=11)))
        return jobs
    return cached('batch_small_jobs', make)


def single_renders(jobs, width):
    def make():
        out = []
        for flat, p in jobs:
            sc = GpuScene(flat, width=width)
            s, q, _ = sc.render_raw(spp=p['spp'], max_verts=p['max_verts'], sampling_mode=p['sampling_mode'],
                                    seed=p['seed'])
            out.append((s, q))
        return out
    return cached(('batch_single', width), make)


@pytest.mark.parametrize('width,cap,block', [(4, 4096, 256), (8, 1 << 20, 128), (2, 1024, 256)])
def test_batch_jobs_equal_single_renders(device, width, cap, block):
    jobs = small_jobs()
    b = Batch()
    for flat, p in jobs:
        b.add_job(flat, **p)
    st = b.run(width=width, path_capacity=cap, block=block, sumsq=True)
    print(st)
    assert st['width'] == width and st['path_capacity'] == cap and st['block'] == block
    ref = single_renders(jobs, width)
    total = 0
    for j, (flat, p) in enumerate(jobs):
        s, q, lat = b.film(j)
        rs, rq = ref[j]
        err = np.abs(s - rs).max() / max(1.0, np.abs(rs).max())
        print(f'job {j}: {flat.w}x{flat.h} spp {p["spp"]} max rel diff {err:.2e} latency {lat:.3f} ms')
        assert np.allclose(s, rs, rtol=1e-9, atol=1e-9)
        assert np.allclose(q, rq, rtol=1e-9, atol=1e-9)
        assert 0.0 < lat <= st['elapsed_s'] * 1e3 * 1.01 + 1.0
        total += flat.w * flat.h * p['spp']
    assert st['samples'] == total and st['jobs'] == len(jobs)


def test_batch_interleaved_queue(device):
    """The interleaved work queue (sample 0 of every job, then sample 1, ...)
    renders the same films as the job-major queue and the single-scene
    renderer, and moves every job's completion toward the end of the run:
    the earliest job finishes later than under the job-major order."""
    jobs = small_jobs()
    b = Batch()
    for flat, p in jobs:
        b.add_job(flat, **p)
    st = b.run(width=4, path_capacity=4096, block=256)
    lat_major = [b.latency_ms(j) for j in range(len(jobs))]
    st_i = b.run(width=4, path_capacity=4096, block=256, interleaved=True)
    lat_inter = [b.latency_ms(j) for j in range(len(jobs))]
    print('job-major', st, lat_major)
    print('interleaved', st_i, lat_inter)
    assert st_i['samples'] == st['samples']
    assert st['iterations'] > 0 and st_i['iterations'] > 0
    ref = single_renders(jobs, 4)
    for j in range(len(jobs)):
        s, _, _ = b.film(j)
        assert np.allclose(s, ref[j][0], rtol=1e-9, atol=1e-9)
    assert min(lat_inter) > min(lat_major)


def test_batch_autotune(device):
    """Fields left 0 are chosen by the autotuner; the films do not depend on
    the choice beyond the BVH width, and a second run reuses it."""
    jobs = small_jobs()
    b = Batch()
    for flat, p in jobs:
        b.add_job(flat, **p)
    st = b.run()
    print('autotuned', st)
    assert st['width'] in (2, 4, 8) and st['block'] in (128, 256) and st['path_capacity'] in (1 << 19, 1 << 20, 1 << 21)
    assert st['tune_s'] > 0
    ref = single_renders(jobs, st['width'])
    for j in range(len(jobs)):
        s, _, _ = b.film(j)
        assert np.allclose(s, ref[j][0], rtol=1e-9, atol=1e-9)
    st2 = b.run()
    assert (st2['width'], st2['path_capacity'], st2['block']) == (st['width'], st['path_capacity'], st['block'])
    assert st2['tune_s'] == 0


def test_batch_errors(device):
    from lm_amd._lib import LmError
    b = Batch()
    with pytest.raises(LmError):
        b.run(width=4, path_capacity=4096, block=256)   # no jobs
    flat, p = small_jobs()[0]
    b.add_job(flat, **p)
    with pytest.raises(LmError):
        b.film(0)                                          # not run yet
    with pytest.raises(LmError):
        b.run(width=3)
    with pytest.raises(LmError):
        b.run(width=4, path_capacity=100)
    with pytest.raises(LmError):
        b.add_job(flat, spp=0)
    from lm_amd import _lib
    o = _lib.BatchOpts(4, 4096, 256, 0, 0, 7)
    st = _lib.BatchStats()
    with pytest.raises(LmError, match='queue_order'):
        _lib.call('lmgpu_batch_run', b._h, o, st, None)


def test_c3_jobs_welch_vs_oracle(device):
    """Config-3 jobs through the batched renderer against the FP64 oracle of
    renderer::pt (src/renderer/renderer_pt.cpp:80-315), independent seeds:
    the largest job of the 64 (about 2M triangles) and two smaller ones, at
    64x64 and 64 spp with 16 vertices, rendered in one batch; every job's
    per-pixel Welch fail fraction (alpha 0.01) stays within the binomial
    bound and its mean radiance within 4 standard errors of the oracle's."""
    from lm_amd import stats
    from oracle import oracle as orc
    jobs = scenes.config3_jobs()
    order = sorted(range(len(jobs)), key=lambda j: jobs[j]['tris'])
    pick = [order[-1], order[len(order) // 2], order[0]]
    spp, M = 64, 16
    b = Batch()
    flats = []
    for j in pick:
        jj = dict(jobs[j], res=64)
        flat = scenes.config3_scene(jj).flatten()
        flats.append(flat)
        b.add_job(flat, spp=spp, max_verts=M, seed=jobs[j]['seed'])
    st = b.run(width=4, path_capacity=1 << 20, block=256, sumsq=True)
    print('batch', st)
    for k, (j, flat) in enumerate(zip(pick, flats)):
        s, q, _ = b.film(k)
        ref = orc.pt_render(flat, spp=spp, max_verts=M, sampling_mode=2, seed=jobs[j]['seed'] ^ 0x5A5A)
        frac, _, _ = stats.welch(s, q, spp, ref['sum'], ref['sumsq'], spp)
        bound = stats.binomial_bound(0.01, flat.w * flat.h)
        n = flat.w * flat.h * spp
        mg, mo = s.sum() / n, ref['sum'].sum() / n
        se = np.sqrt((q.sum() / n - (s.sum(axis=(0, 1)) / n) ** 2).sum() / n +
                     (ref['sumsq'].sum() / n - (ref['sum'].sum(axis=(0, 1)) / n) ** 2).sum() / n)
        print(f'job {j}: {jobs[j]["tris"]} tris (flattened {flat.verts.shape[0]}), Welch fail {frac:.4f} '
              f'(bound {bound:.4f}), mean {mg:.5f} vs oracle {mo:.5f}')
        assert frac <= bound
        assert abs(mg - mo) <= 4 * se + 1e-12
