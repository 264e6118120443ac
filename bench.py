"""Benchmark of the MI355X path-tracing hot path (one GPU).

Headline: closest-hit Mrays/s on the config-1 mesh (displaced icosphere, 8
subdivision levels, + 10K lognormal clutter = 1,320,720 triangles) for the
2,000,000 incoherent rays (origins uniform on the surface, directions
uniform over the normal's hemisphere). A step is one closest-hit pass over
the whole ray batch, rays and BVH resident in HBM. Also reported: coherent
1920x1080 primaries, the GPU BVH build time, and wavefront path-tracing
samples/s at 1920x1080 on config 2 (max 16 vertices, NEE + MIS), and the
batched rendering of config 3 (64 jobs of 10K-2M triangles, 256-1024^2
pixels, 16-256 spp, one wavefront loop over a device work queue with
autotuned BVH width / path capacity / block size): jobs/s and p50 / p99 job
latency; and a reduced equal-time comparison of guided and plain PT on the
slit-lit room (config 4 at 960x540, 10 s each against a 2048-spp plain
reference; the 1920x1080 / 30 s / 4096-spp run is
lightmetrica-v3_amd/lm_amd/equal_time.py).

Prints one JSON line. `roofline` is for the dominant kernel (closest-hit
traversal, k_trace): algorithmic bytes per launch = rays x B_ray, with
B_ray = 32 + 16 + N_node * 32 + N_tri * 36 from the reference SAH BVH's
mean node visits / triangle tests on the same rays
(tests/golden/c1_traversal_stats.json), divided by the mean kernel time
measured with HIP events on the launch stream. `cpu_baseline` times the FP64
oracle restatement of accel::sahbvh on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'lightmetrica-v3_amd'))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profiles():
    """HBM bytes per launch of the traversal kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy
    kernel calibration); None if absent."""
    p = os.path.join(ROOT, 'profiles', 'trace_c1_incoherent_pmc.json')
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            d = json.load(f)
        return float(d['hbm_bytes_per_launch'])
    except Exception:
        return None


def cpu_baseline(flat, rays, n_rays):
    """The oracle's FP64 restatement of accel::sahbvh, renderer::pt and
    renderer::lt and renderer::bdpt on all host cores: incoherent (the
    headline unit) and coherent closest-hit Mrays/s on the config-1 rays, the
    SAH build time, renderer::pt samples/s, renderer::lt light paths/s and
    renderer::bdpt samples/s on config 0 (256x256, 8 vertices)."""
    from oracle import oracle as orc
    import numpy as np
    from lm_amd import scenes
    cores = os.cpu_count() or 1
    t0 = time.time()
    acc = orc.Accel(flat.verts, cores)
    build_s = time.time() - t0
    out = {}
    for label in ('incoherent', 'coherent'):
        sub = {k: np.ascontiguousarray(v[:n_rays]) for k, v in rays[label].items()}
        acc.intersect(sub, nthreads=cores)  # warm
        t0 = time.time()
        acc.intersect(sub, nthreads=cores)
        out[label] = n_rays / (time.time() - t0) / 1e6
    f0 = scenes.config0().flatten()
    t0 = time.time()
    r = orc.pt_render(f0, spp=16, max_verts=8, seed=1, nthreads=cores, sumsq=False)
    pt_s = r['stats']['samples'] / (time.time() - t0)
    t0 = time.time()
    r = orc.lt_render(f0, 16 * f0.w * f0.h, 8, seed=1, nthreads=cores)
    lt_s = r['stats']['samples'] / (time.time() - t0)
    t0 = time.time()
    r = orc.bdpt_render(f0, 4 * f0.w * f0.h, 8, seed=1, nthreads=cores)
    bdpt_s = r['stats']['samples'] / (time.time() - t0)
    return {'value': out['incoherent'], 'unit': 'Mrays/s', 'cores': cores, 'kind': 'port',
            'sample': f'FP64 accel::sahbvh restatement (oracle), OpenMP over {n_rays} of the 2,000,000 '
                      f'incoherent config-1 rays; SAH build of 1,320,720 tris {build_s:.1f} s excluded',
            'build_s': round(build_s, 2), 'coherent_mrays_s': round(out['coherent'], 2),
            'pt_c0_samples_s': round(pt_s, 1),
            'pt_c0_sample': 'renderer::pt restatement (oracle), config 0 256x256 16 spp, max 8 vertices, MIS',
            'lt_c0_light_paths_s': round(lt_s, 1),
            'bdpt_c0_samples_s': round(bdpt_s, 1),
            'bdpt_c0_sample': 'renderer::bdpt restatement (oracle, explicit MIS), config 0, 4 x 256x256 samples, '
                              'max 8 vertices',
            'lt_c0_sample': 'renderer::lt restatement (oracle), config 0, 16 x 256x256 light paths, max 8 vertices'}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--width', type=int, default=4)
    ap.add_argument('--leaf', type=int, default=2)
    ap.add_argument('--pt-spp', type=int, default=4)
    ap.add_argument('--no-pt', action='store_true')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-c3', action='store_true')
    ap.add_argument('--no-c4', action='store_true')
    ap.add_argument('--cpu-rays', type=int, default=2_000_000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({'error': 'this build runs on one GPU only; --gpus must be 1 '
                          '(rays and samples would shard by index range over a replicated BVH, '
                          'with one sum of the FP64 films at the end)'}))
        sys.exit(2)

    import numpy as np
    import torch
    from lm_amd import scenes
    from lm_amd.accel import Accel, rays_from_dict

    dev = torch.device('cuda:0')
    log('generating config-1 scene and rays')
    scene, rays = scenes.config1()
    flat = scene.flatten()
    acc = Accel(args.width, args.leaf)
    acc.build(flat)
    acc.build(flat)
    info = acc.info()
    log('accel', info)

    def time_rays(label, steps, warmup):
        r = torch.from_numpy(rays_from_dict(rays[label])).to(dev)
        h = torch.zeros((r.shape[0], 4), dtype=torch.float32, device=dev)
        for _ in range(warmup):
            acc.intersect_device(r, h)
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        t_start = torch.cuda.Event(enable_timing=True)
        t_end = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t_start.record()
        for e0, e1 in ev:
            e0.record()
            acc.intersect_device(r, h)
            e1.record()
        t_end.record()
        torch.cuda.synchronize()
        acc.status()
        total_ms = t_start.elapsed_time(t_end)
        kern_ms = sum(a.elapsed_time(b) for a, b in ev) / steps
        return r.shape[0], total_ms, kern_ms

    n_inc, total_ms, kern_ms = time_rays('incoherent', args.steps, args.warmup)
    value = n_inc * args.steps / (total_ms / 1e3) / 1e6
    n_coh, tot_c, kern_c = time_rays('coherent', args.steps, args.warmup)
    coherent = n_coh * args.steps / (tot_c / 1e3) / 1e6
    log(f'incoherent {value:.1f} Mrays/s (kernel {kern_ms:.3f} ms), coherent {coherent:.1f} Mrays/s')

    with open(os.path.join(ROOT, 'tests', 'golden', 'c1_traversal_stats.json')) as f:
        ref_stats = json.load(f)
    b_ray = ref_stats['incoherent']['bytes_per_ray']
    achieved = n_inc * b_ray / (kern_ms / 1e3) / 1e9
    traffic = traffic_from_profiles()
    roofline = {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                'frac': round(achieved / HBM_PEAK_GBS, 4), 'traffic': traffic,
                'kernel': f'k_trace<{args.width},closest> over {n_inc} rays',
                'bytes_per_ray': round(b_ray, 1), 'kernel_ms': round(kern_ms, 4)}

    extra = {'coherent_mrays_s': round(coherent, 1), 'bvh_build_ms': round(info['build_ms'], 2),
             'bvh': {k: info[k] for k in ('width', 'leaf_size', 'num_nodes', 'depth', 'sah_cost')}}

    if not args.no_pt:
        from lm_amd.renderer import GpuScene
        log('generating config-2 scene (1920x1080)')
        s2 = scenes.config2().flatten()
        sc = GpuScene(s2, width=args.width, leaf_size=args.leaf)
        fs = torch.zeros((s2.h, s2.w, 3), dtype=torch.float64, device=dev)
        fq = torch.zeros_like(fs)
        sc.render_raw(spp=1, max_verts=16, seed=1, film_sum=fs, film_sumsq=fq)
        torch.cuda.synchronize()
        t0 = time.time()
        _, _, st = sc.render_raw(spp=args.pt_spp, max_verts=16, seed=1, sample_offset=1, film_sum=fs, film_sumsq=fq)
        torch.cuda.synchronize()
        dt = time.time() - t0
        extra['pt_1080p'] = {'samples_s': round(st['samples'] / dt, 1), 'spp': args.pt_spp,
                             'secondary_mrays_s': round((st['extension_rays'] + st['shadow_rays']) / dt / 1e6, 1),
                             'extension_rays': st['extension_rays'], 'shadow_rays': st['shadow_rays'],
                             'ms_per_spp': round(dt / args.pt_spp * 1e3, 2)}
        log('pt', extra['pt_1080p'])
        # renderer::lt on the same scene: passes of 1920x1080 light paths
        sc.render_lt(num_samples=s2.w * s2.h, max_verts=16, seed=2, film=fs)
        torch.cuda.synchronize()
        t0 = time.time()
        _, st = sc.render_lt(num_samples=args.pt_spp * s2.w * s2.h, max_verts=16, seed=2, pass_offset=1, film=fs)
        torch.cuda.synchronize()
        dt = time.time() - t0
        extra['lt_1080p'] = {'light_paths_s': round(st['samples'] / dt, 1), 'passes': st['passes'],
                             'mrays_s': round((st['extension_rays'] + st['shadow_rays']) / dt / 1e6, 1),
                             'ms_per_pass': round(dt / st['passes'] * 1e3, 2)}
        log('lt', extra['lt_1080p'])
        # renderer::bdpt on config 0 (area light only) at 1920x1080, max 8 vertices
        f0 = scenes.config0().flatten()
        f0.w, f0.h = 1920, 1080
        f0.camera.aspect = 1920 / 1080
        sb = GpuScene(f0, width=args.width, leaf_size=args.leaf)
        fb = torch.zeros((f0.h, f0.w, 3), dtype=torch.float64, device=dev)
        sb.render_bdpt(num_samples=1 << 18, max_verts=8, seed=3, film=fb)
        torch.cuda.synchronize()
        t0 = time.time()
        _, st = sb.render_bdpt(num_samples=8 << 18, max_verts=8, seed=3, pass_offset=1, film=fb)
        torch.cuda.synchronize()
        dt = time.time() - t0
        extra['bdpt_c0_1080p'] = {'samples_s': round(st['samples'] / dt, 1), 'samples': st['samples'],
                                  'mrays_s': round((st['extension_rays'] + st['shadow_rays']) / dt / 1e6, 1),
                                  'max_verts': 8}
        log('bdpt', extra['bdpt_c0_1080p'])
        del sb, fb

    if not args.no_c3:
        from lm_amd.batch import Batch
        log('generating config-3 jobs (64 scenes)')
        jobs = scenes.config3_jobs()
        b = Batch()
        t0 = time.time()
        for j in jobs:
            b.add_job(scenes.config3_scene(j).flatten(), spp=j['spp'], max_verts=16, seed=j['seed'])
        gen_s = time.time() - t0
        st = b.run()
        lat = np.array([b.latency_ms(i) for i in range(len(jobs))])
        extra['c3_batch'] = {
            'jobs': len(jobs), 'jobs_s': round(len(jobs) / st['elapsed_s'], 3),
            'p50_latency_ms': round(float(np.percentile(lat, 50)), 1),
            'p99_latency_ms': round(float(np.percentile(lat, 99)), 1),
            'samples_s': round(st['samples'] / st['elapsed_s'], 1), 'render_s': round(st['elapsed_s'], 3),
            'build_s': round(st['build_s'], 3), 'tune_s': round(st['tune_s'], 3),
            'triangles': int(sum(j['tris'] for j in jobs)), 'samples': int(st['samples']),
            'autotuned': {'bvh_width': st['width'], 'path_capacity': st['path_capacity'], 'block': st['block']},
            'scene_generation_s': round(gen_s, 1)}
        log('c3', extra['c3_batch'])
        del b

    if not args.no_c4:
        from lm_amd import equal_time
        log('config-4 equal-time comparison (960x540, 10 s each)')
        et = equal_time.run(width=960, height=540, seconds=10.0, ref_spp=2048)

        def row(d):
            return {'spp': d['spp'], 'rr_depth': d['rr_depth'], 'rmse': d['rmse'],
                    'rmse_corrected': d['rmse_corrected']}
        extra['c4_equal_time'] = {
            'resolution': '960x540', 'seconds_each': 10.0, 'reference_spp': 2048,
            'plain': row(et['plain']), 'plain_control': row(et['control']),
            'guided_defaults': dict(row(et['guided']), adam_steps=et['guided']['adam_steps']),
            'ratio_plain_over_guided': round(et['rmse_ratio_plain_over_guided'], 3),
            'ratio_control_over_guided': round(et['rmse_ratio_control_over_guided'], 3),
            'full_size_result': 'profiles/guided_c4_equal_time.md'}
        log('c4', extra['c4_equal_time'])

    cpu = None
    if not args.no_cpu:
        log('cpu baseline (oracle)')
        cpu = cpu_baseline(flat, rays, min(args.cpu_rays, 2_000_000))

    out = {
        'metric': 'Mrays/s (incoherent) & samples/s at 1080p; equal-time MSE guided vs plain PT',
        'value': round(value, 2), 'unit': 'Mrays/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': round(total_ms / args.steps, 4), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic (seeded procedural scene and rays)',
        'config': {'workload': 'closest-hit ray cast, config-1 mesh 1,320,720 tris, 2,000,000 incoherent rays',
                   'bvh_width': args.width, 'leaf_size': args.leaf},
        'roofline': roofline,
        'cpu_baseline': cpu,
        'extra': extra,
    }
    print(json.dumps(out))


if __name__ == '__main__':
    main()
