"""Profile driver: a fixed number of closest-hit launches over the config-1
incoherent rays, preceded by a plain device copy of known size that
calibrates the HBM byte counters; optionally one 1080p config-2
path-tracing pass (--pt), 4 + 4 plain (roulette from vertex 5 / none) and 4 guided (no
roulette) 1080p renders of config 4 (--c4), a config-3 batch (--c3 N), light-tracing / BDPT passes (--lt-bdpt) and
the guide MLP's training and inference kernels (--guide).

Run under rocprofv3, e.g.
  rocprofv3 --kernel-trace --stats -d OUT -o kt -- python3 lightmetrica-v3_amd/lm_amd/profile_trace.py --pt
  rocprofv3 --pmc FETCH_SIZE -d OUT -o fetch -- python3 lightmetrica-v3_amd/lm_amd/profile_trace.py
and summarise with lightmetrica-v3_amd/lm_amd/prof_summary.py OUT/kt_results.db.
"""
import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

COPY_BYTES = 512 * 1024 * 1024


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--launches', type=int, default=5)
    ap.add_argument('--width', type=int, default=4)
    ap.add_argument('--leaf', type=int, default=2)
    ap.add_argument('--rays', default='incoherent')
    ap.add_argument('--pt', action='store_true',
                    help='also run one 1080p config-2 PT pass (one-sample pass graph) and a 4-spp render of it '
                         '(regenerating loop)')
    ap.add_argument('--pt-width', type=int, default=8, help='BVH width of the path-tracing scene')
    ap.add_argument('--c4', action='store_true', help='also render 1080p config 4 at 32 spp: plain, plain without roulette, guided')
    ap.add_argument('--no-trace', action='store_true', help='skip the config-1 closest-hit launches')
    ap.add_argument('--c3', type=int, default=0, help='also render the first N config-3 jobs as one batch')
    ap.add_argument('--guide', action='store_true',
                    help='also run 20 guide training steps of 2^16 records and 20 inferences of 2M queries')
    ap.add_argument('--lt-bdpt', action='store_true',
                    help='also run 2 LT passes of config 2 and 2 BDPT passes of config 0 at 1080p')
    args = ap.parse_args()
    import torch
    from lm_amd import scenes
    from lm_amd.accel import Accel, rays_from_dict

    dev = torch.device('cuda:0')
    # calibration: one copy of COPY_BYTES (reads and writes COPY_BYTES)
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print(f'copy calibration: {COPY_BYTES} bytes read + written', flush=True)
    del src, dst
    if not args.no_trace:
        scene, rays = scenes.config1()
        flat = scene.flatten()
        acc = Accel(args.width, args.leaf)
        acc.build(flat)
        r = torch.from_numpy(rays_from_dict(rays[args.rays])).to(dev)
        h = torch.zeros((r.shape[0], 4), dtype=torch.float32, device=dev)
        for i in range(args.launches):
            acc.intersect_device(r, h)
        torch.cuda.synchronize()
        print(f'{args.launches} x k_trace over {r.shape[0]} {args.rays} rays', flush=True)
    if args.pt:
        from lm_amd.renderer import GpuScene
        s2 = scenes.config2().flatten()
        sc = GpuScene(s2, width=args.pt_width, leaf_size=args.leaf)
        sc.render_raw(spp=1, max_verts=16, seed=1)
        torch.cuda.synchronize()
        print('1 x 1080p config-2 PT pass', flush=True)
        _, _, st = sc.render_raw(spp=4, max_verts=16, seed=1, sample_offset=1)
        torch.cuda.synchronize()
        print('4-spp 1080p config-2 PT render (regenerating loop)', st, flush=True)
    if args.lt_bdpt:
        from lm_amd.renderer import GpuScene
        s2 = scenes.config2().flatten()
        sc = GpuScene(s2, width=args.width, leaf_size=args.leaf)
        _, st = sc.render_lt(num_samples=2 * s2.w * s2.h, max_verts=16, seed=1)
        print('2 x 1080p config-2 LT passes', st, flush=True)
        f0 = scenes.config0().flatten()
        f0.w, f0.h = 1920, 1080
        f0.camera.aspect = 1920 / 1080
        sb = GpuScene(f0, width=args.width, leaf_size=args.leaf)
        _, st = sb.render_bdpt(num_samples=2 << 18, max_verts=8, seed=1)
        print('2^19 samples of config-0 BDPT at 1080p', st, flush=True)
    if args.c3:
        from lm_amd.batch import Batch
        jobs = scenes.config3_jobs()[:args.c3]
        b = Batch()
        for j in jobs:
            b.add_job(scenes.config3_scene(j).flatten(), spp=j['spp'], max_verts=16, seed=j['seed'])
        st = b.run(width=args.width, path_capacity=1 << 21, block=128, leaf_size=args.leaf)
        print(f'batch of {len(jobs)} config-3 jobs', st, flush=True)
    if args.c4:
        from lm_amd.guide import Guide
        from lm_amd.renderer import GpuScene
        s4 = scenes.config4().flatten()
        sc = GpuScene(s4, width=args.width, leaf_size=args.leaf)
        fs = torch.zeros((s4.h, s4.w, 3), dtype=torch.float64, device=dev)
        _, _, st = sc.render_raw(spp=32, max_verts=16, seed=1, film_sum=fs)
        torch.cuda.synchronize()
        print('32-spp 1080p config-4 plain PT render (roulette from vertex 5)', st, flush=True)
        _, _, st = sc.render_raw(spp=32, max_verts=16, rr_depth=16, seed=1, film_sum=fs)
        torch.cuda.synchronize()
        print('32-spp 1080p config-4 plain PT render (no roulette)', st, flush=True)
        lo, hi = sc.bounds()
        g = Guide(lo - 1e-3, hi + 1e-3, seed=2)
        _, _, st = sc.render_raw(spp=32, max_verts=16, rr_depth=16, seed=2, film_sum=fs, guide=g)
        torch.cuda.synchronize()
        print('32-spp 1080p config-4 guided PT render (default guide options, no roulette)', st, flush=True)
    if args.guide:
        guide_workload(dev)


def guide_workload(dev):
    """20 Adam steps of the guide MLP on 2^16 random records and 20
    inferences of 2M queries (the training kernels and the query forward)."""
    import numpy as np
    import torch
    from lm_amd.guide import Guide
    rng = np.random.default_rng(0)
    n = 1 << 16

    def unit(k):
        v = rng.normal(size=(k, 3))
        return v / np.linalg.norm(v, axis=1, keepdims=True)

    rec = np.zeros((n, 16), dtype=np.float32)
    rec[:, 0:3] = rng.uniform(-1, 1, (n, 3))
    rec[:, 3:6], rec[:, 6:9], rec[:, 9:12] = unit(n), unit(n), unit(n)
    rec[:, 12] = 1.0 / (4 * np.pi)
    rec[:, 13:16] = rng.uniform(0.1, 1.0, (n, 3))
    g = Guide(np.array([-1.1] * 3), np.array([1.1] * 3), seed=1)
    r = torch.from_numpy(rec).to(dev)
    q = torch.cat([torch.from_numpy(np.ascontiguousarray(rec[:, :9])).to(dev)] * 32)
    for _ in range(20):
        g.train(r, lr=1e-3)
        g.infer(q)
    torch.cuda.synchronize()
    print('20 guide training steps of 2^16 records and 20 inferences of 2M queries', flush=True)


if __name__ == '__main__':
    main()
