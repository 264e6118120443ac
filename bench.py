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
# Guide settings exactly as config 4 states them: an even one-sample MIS
# split, Adam at lr 1e-3 on 2^18 records per step (one step per pass), and
# the reference renderer's Russian roulette (from vertex 5).
C4_CONFIG_EXACT = {'alpha': 0.5, 'learning_rate': 1e-3, 'batch': 1 << 18, 'steps_per_pass': 1, 'rr_depth': 5}
# Bytes a path vertex moves outside traversal (k_shade and the queues): the
# ray and hit it reads (32 + 16), its shading triangle (96) and material
# (48), the path state read and written (2 x 64), the extension and shadow
# queue appends (36 + 48).
B_VERTEX = 32 + 16 + 96 + 48 + 2 * 64 + 36 + 48


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


def host_cores():
    """Threads the oracle runs on: OMP_NUM_THREADS when set (the CPU share
    of the machine), else the CPUs this process may run on."""
    try:
        n = int(os.environ.get('OMP_NUM_THREADS', '0'))
    except ValueError:
        n = 0
    return n if n > 0 else len(os.sched_getaffinity(0))


def cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def cpu_baseline(flat, rays, n_rays, c2_flat=None):
    """The oracle's FP64 restatement of accel::sahbvh, renderer::pt,
    renderer::lt and renderer::bdpt on the host cores: incoherent (the
    headline unit) and coherent closest-hit Mrays/s on the config-1 rays with
    all cores and with one thread, the SAH build time, renderer::pt
    samples/s on config 2 (one 1920x1080 1-spp pass; config 2 is 64 spp, so
    its full render takes 64 such passes) and on config 0 (256x256, 8
    vertices, all cores and one thread), renderer::lt light paths/s and
    renderer::bdpt samples/s on config 0."""
    from oracle import oracle as orc
    import numpy as np
    from lm_amd import scenes
    cores = host_cores()
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
    n1 = min(n_rays, 200_000)
    sub = {k: np.ascontiguousarray(v[:n1]) for k, v in rays['incoherent'].items()}
    t0 = time.time()
    acc.intersect(sub, nthreads=1)
    inc_1t = n1 / (time.time() - t0) / 1e6
    f0 = scenes.config0().flatten()
    t0 = time.time()
    r = orc.pt_render(f0, spp=16, max_verts=8, seed=1, nthreads=cores, sumsq=False)
    pt_s = r['stats']['samples'] / (time.time() - t0)
    t0 = time.time()
    r = orc.pt_render(f0, spp=2, max_verts=8, seed=1, nthreads=1, sumsq=False)
    pt_1t = r['stats']['samples'] / (time.time() - t0)
    t0 = time.time()
    r = orc.lt_render(f0, 16 * f0.w * f0.h, 8, seed=1, nthreads=cores)
    lt_s = r['stats']['samples'] / (time.time() - t0)
    t0 = time.time()
    r = orc.bdpt_render(f0, 4 * f0.w * f0.h, 8, seed=1, nthreads=cores)
    bdpt_s = r['stats']['samples'] / (time.time() - t0)
    res = {'value': out['incoherent'], 'unit': 'Mrays/s', 'cores': cores, 'kind': 'port',
           'cpu_model': cpu_model(),
           'sample': f'FP64 accel::sahbvh restatement (oracle), OpenMP over {n_rays} of the 2,000,000 '
                     f'incoherent config-1 rays on {cores} threads; SAH build of 1,320,720 tris '
                     f'{build_s:.1f} s excluded',
           'build_s': round(build_s, 2), 'coherent_mrays_s': round(out['coherent'], 2),
           'incoherent_mrays_s_1_thread': round(inc_1t, 3),
           'incoherent_1_thread_sample': f'first {n1} incoherent config-1 rays, one thread',
           'pt_c0_samples_s': round(pt_s, 1),
           'pt_c0_samples_s_1_thread': round(pt_1t, 1),
           'pt_c0_sample': 'renderer::pt restatement (oracle), config 0 256x256 16 spp (1 thread: 2 spp), '
                           'max 8 vertices, MIS',
           'lt_c0_light_paths_s': round(lt_s, 1),
           'bdpt_c0_samples_s': round(bdpt_s, 1),
           'bdpt_c0_sample': 'renderer::bdpt restatement (oracle, explicit MIS), config 0, 4 x 256x256 samples, '
                             'max 8 vertices',
           'lt_c0_sample': 'renderer::lt restatement (oracle), config 0, 16 x 256x256 light paths, max 8 vertices'}
    if c2_flat is not None:
        acc2 = orc.Accel(c2_flat.verts, cores)
        t0 = time.time()
        r = orc.pt_render(c2_flat, spp=1, max_verts=16, seed=1, nthreads=cores, accel=acc2, sumsq=False)
        dt = time.time() - t0
        res['pt_c2_samples_s'] = round(r['stats']['samples'] / dt, 1)
        res['pt_c2_64spp_render_s'] = round(64 * dt, 1)
        res['pt_c2_sample'] = ('renderer::pt restatement (oracle), config 2 one 1920x1080 1-spp pass, max 16 '
                               'vertices, NEE + MIS; the 64-spp render time is 64 x that pass (extrapolated)')
    return res


def pt_roofline(samples_s, spp_stats):
    """Algorithmic HBM bytes of a config-2 path-tracing sample: every
    extension and shadow segment priced like a ray-cast ray (32 B ray + 16 B
    result + 32 B per node visit + 36 B per triangle test) with the node
    visits and triangle tests of the reference's SAH BVH on the same sample
    distribution (tests/golden/c2_pt_stats.json, from the oracle), plus
    B_VERTEX per path vertex and 48 B of FP64 film (sum and sum of squares)."""
    ext = (spp_stats['extension_rays_per_sample'] * 48 + spp_stats['extension_node_visits_per_sample'] * 32 +
           spp_stats['extension_tri_tests_per_sample'] * 36)
    shadow = (spp_stats['shadow_rays_per_sample'] * 48 + spp_stats['shadow_node_visits_per_sample'] * 32 +
              spp_stats['shadow_tri_tests_per_sample'] * 36)
    vert = spp_stats['path_vertices_per_sample'] * B_VERTEX
    b = ext + shadow + vert + 48
    achieved = b * samples_s / 1e9
    return {'bound': 'hbm', 'bytes_per_sample': round(b, 1), 'traversal_bytes_per_sample': round(ext + shadow, 1),
            'shading_bytes_per_sample': round(vert + 48, 1), 'achieved': round(achieved, 1),
            'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--width', type=int, default=4)
    ap.add_argument('--leaf', type=int, default=2)
    ap.add_argument('--pt-width', type=int, default=8,
                    help='BVH width of the path tracer (its bounce kernels run 3.5%% faster at 8 than at 4)')
    ap.add_argument('--pt-spp', type=int, default=64)
    ap.add_argument('--no-pt', action='store_true')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-c3', action='store_true')
    ap.add_argument('--no-c4', action='store_true')
    ap.add_argument('--cpu-rays', type=int, default=2_000_000)
    ap.add_argument('--c4-width', type=int, default=960)
    ap.add_argument('--c4-height', type=int, default=540)
    ap.add_argument('--c4-seconds', type=float, default=10.0)
    ap.add_argument('--c4-ref-spp', type=int, default=2048,
                    help='config 4 as configured: --c4-width 1920 --c4-height 1080 --c4-seconds 30 '
                         '--c4-ref-spp 4096')
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
    # the same bytes priced on this build's own tree: its node visits and
    # triangle tests on these rays, at its node size (quantised nodes for
    # widths 4 and 8) and 48-B triangles
    r_inc = torch.from_numpy(rays_from_dict(rays['incoherent'])).to(dev)
    h_inc = torch.zeros((r_inc.shape[0], 4), dtype=torch.float32, device=dev)
    own_nv, own_nt = acc.intersect_counted(r_inc, h_inc)
    node_b = 16 * args.width if args.width in (4, 8) else 32 * args.width
    own_b_ray = 32 + 16 + own_nv / n_inc * node_b + own_nt / n_inc * 48
    own_achieved = n_inc * own_b_ray / (kern_ms / 1e3) / 1e9
    del r_inc, h_inc
    roofline = {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                'frac': round(achieved / HBM_PEAK_GBS, 4), 'traffic': traffic,
                'kernel': f'k_trace<{args.width},closest> over {n_inc} rays',
                'bytes_model': 'reference algorithm: node visits and triangle tests of the FP64 binary SAH BVH '
                               '(accel::sahbvh) on the same rays, 32 B per node and 36 B per triangle',
                'bytes_per_ray': round(b_ray, 1), 'kernel_ms': round(kern_ms, 4),
                'own_tree': {'node_visits_per_ray': round(own_nv / n_inc, 2),
                             'tri_tests_per_ray': round(own_nt / n_inc, 2), 'node_bytes': node_b,
                             'bytes_per_ray': round(own_b_ray, 1), 'achieved': round(own_achieved, 1),
                             'frac': round(own_achieved / HBM_PEAK_GBS, 4)}}

    extra = {'coherent_mrays_s': round(coherent, 1), 'bvh_build_ms': round(info['build_ms'], 2),
             'bvh': {k: info[k] for k in ('width', 'leaf_size', 'num_nodes', 'depth', 'sah_cost')}}

    if not args.no_pt:
        from lm_amd.renderer import GpuScene
        log('generating config-2 scene (1920x1080)')
        s2 = scenes.config2().flatten()
        sc = GpuScene(s2, width=args.pt_width, leaf_size=args.leaf)
        fs = torch.zeros((s2.h, s2.w, 3), dtype=torch.float64, device=dev)
        fq = torch.zeros_like(fs)
        # warm-up of the same (regenerating, fixed-spp) path: path pool and graph
        sc.render_raw(spp=2, max_verts=16, seed=1, film_sum=fs, film_sumsq=fq)
        torch.cuda.synchronize()
        t0 = time.time()
        _, _, st = sc.render_raw(spp=args.pt_spp, max_verts=16, seed=1, sample_offset=2, film_sum=fs, film_sumsq=fq)
        torch.cuda.synchronize()
        dt = time.time() - t0
        with open(os.path.join(ROOT, 'tests', 'golden', 'c2_pt_stats.json')) as f:
            c2_stats = json.load(f)
        extra['pt_1080p'] = {'samples_s': round(st['samples'] / dt, 1), 'spp': args.pt_spp, 'bvh_width': args.pt_width,
                             'render_s': round(dt, 3),
                             'secondary_mrays_s': round((st['extension_rays'] + st['shadow_rays']) / dt / 1e6, 1),
                             'extension_rays': st['extension_rays'], 'shadow_rays': st['shadow_rays'],
                             'ms_per_spp': round(dt / args.pt_spp * 1e3, 2),
                             'roofline': pt_roofline(st['samples'] / dt, c2_stats)}
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
        st_i = b.run(interleaved=True)
        lat_i = np.array([b.latency_ms(i) for i in range(len(jobs))])
        extra['c3_batch'] = {
            'jobs': len(jobs), 'jobs_s': round(len(jobs) / st['elapsed_s'], 3),
            'p50_latency_ms': round(float(np.percentile(lat, 50)), 1),
            'p99_latency_ms': round(float(np.percentile(lat, 99)), 1),
            'queue_order': 'job-major (each job sample by sample, jobs in submission order)',
            'iterations': st['iterations'],
            'interleaved': {'jobs_s': round(len(jobs) / st_i['elapsed_s'], 3),
                            'p50_latency_ms': round(float(np.percentile(lat_i, 50)), 1),
                            'p99_latency_ms': round(float(np.percentile(lat_i, 99)), 1),
                            'render_s': round(st_i['elapsed_s'], 3),
                            'queue_order': 'sample 0 of every job, then sample 1, ...'},
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
        et = equal_time.run(width=args.c4_width, height=args.c4_height, seconds=args.c4_seconds,
                            ref_spp=args.c4_ref_spp, variants=[C4_CONFIG_EXACT])

        def row(d):
            return {'spp': d['spp'], 'rr_depth': d['rr_depth'], 'rmse': d['rmse'],
                    'rmse_corrected': d['rmse_corrected']}
        ce = et['variants'][0]
        extra['c4_equal_time'] = {
            'resolution': f'{args.c4_width}x{args.c4_height}', 'seconds_each': args.c4_seconds,
            'reference_spp': args.c4_ref_spp,
            'plain': row(et['plain']), 'plain_control': row(et['control']),
            'guided_defaults': dict(row(et['guided']), adam_steps=et['guided']['adam_steps']),
            'ratio_plain_over_guided': round(et['rmse_ratio_plain_over_guided'], 3),
            'ratio_control_over_guided': round(et['rmse_ratio_control_over_guided'], 3),
            'guided_config_exact': dict(row(ce), adam_steps=ce['adam_steps'], options=C4_CONFIG_EXACT),
            'ratio_plain_over_guided_config_exact': round(ce['rmse_ratio_plain_over_guided'], 3),
            'ratio_plain_over_guided_raw': round(et['plain']['rmse'] / et['guided']['rmse'], 3),
            'ratio_plain_over_guided_config_exact_raw': round(et['plain']['rmse'] / ce['rmse'], 3),
            'full_size_result': 'profiles/guided_c4_equal_time.md'}
        log('c4', extra['c4_equal_time'])

    cpu = None
    if not args.no_cpu:
        log('cpu baseline (oracle)')
        cpu = cpu_baseline(flat, rays, min(args.cpu_rays, 2_000_000),
                           None if args.no_pt else s2)

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
