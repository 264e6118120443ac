"""Equal-time comparison of guided and plain wavefront PT on the slit-lit room
(config 4: light through a 0.3-wide wall slit, GGX floor alpha 0.1, 200K
clutter triangles, max 16 vertices).

Renders a converged plain-PT reference (FP64 film sum and sum of squares,
`--ref-spp` samples per pixel, in chunks that print progress), then, for
`--time` seconds each with the time scheduler:
- plain PT with the reference renderer's Russian roulette (from vertex 5,
  renderer_pt.cpp:281-287), the method as the reference defines it;
- plain PT with the same roulette setting as the guided run (the control
  that isolates what the guide itself gains);
- guided PT (a fresh guide trained online during the render).
Russian roulette on throughput alone terminates the low-throughput paths
that carry most of the light of a room lit through a slit, so the guided
renderer runs without it by default (`--guided-rr-depth`, default
max_verts = no roulette); every setting is unbiased.

rMSE = mean((I - R)^2 / (R^2 + 1e-2)) against the reference R is reported
raw and with the reference's own noise removed (stats.rel_mse_corrected);
the ratios use the corrected values. One JSON line.

    python3 lightmetrica-v3_amd/lm_amd/equal_time.py --width 1920 --height 1080 --time 30 --ref-spp 4096
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

PLAIN_RR_DEPTH = 5   # Renderer_PT: roulette once num_verts > 5


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run(width=1920, height=1080, seconds=30.0, ref_spp=4096, max_verts=16, chunk=128, guide_opts=None,
        save_dir=None, variants=None, guided_rr_depth=None):
    """Returns the result dict of the first guide variant (guide_opts) and, under
    'variants', one entry per element of `variants` (option dicts, each
    rendered with a fresh guide for `seconds`)."""
    import numpy as np
    import torch
    from lm_amd import scenes, stats
    from lm_amd.guide import Guide
    from lm_amd.renderer import GpuScene

    rr_g = max_verts if guided_rr_depth is None else int(guided_rr_depth)
    dev = torch.device('cuda:0')
    t0 = time.time()
    flat = scenes.config4(width=width, height=height).flatten()
    sc = GpuScene(flat)
    log(f'config 4 {width}x{height}: {flat.num_triangles} triangles, scene ready in {time.time() - t0:.1f} s')
    # reference: plain PT without roulette, FP64 device film (sum and sum of squares), in chunks
    fs = torch.zeros((height, width, 3), dtype=torch.float64, device=dev)
    fq = torch.zeros_like(fs)
    done = 0
    t0 = time.time()
    while done < ref_spp:
        n = min(chunk, ref_spp - done)
        sc.render_raw(spp=n, max_verts=max_verts, rr_depth=max_verts, seed=987654321, sample_offset=done,
                      film_sum=fs, film_sumsq=fq)
        torch.cuda.synchronize()
        done += n
        log(f'reference {done}/{ref_spp} spp, {time.time() - t0:.1f} s')
    ref_s = time.time() - t0
    ref_sum, ref_sq = fs.cpu().numpy(), fq.cpu().numpy()
    ref = ref_sum / ref_spp

    def errors(img):
        return stats.rel_mse(img, ref), stats.rel_mse_corrected(img, ref_sum, ref_sq, ref_spp)

    def entry(st, e, **kw):
        d = {'spp': int(st['spp']), 'seconds': round(st['elapsed'], 2), 'rmse': e[0], 'rmse_corrected': e[1],
             'ms_per_pass': round(st['elapsed'] / st['spp'] * 1e3, 2)}
        d.update(kw)
        return d

    def plain_run(rr, seed):
        ps, _, pst = sc.render_raw(render_time=seconds, max_verts=max_verts, rr_depth=rr, seed=seed)
        img = ps / pst['spp']
        e = errors(img)
        log(f'plain rr_depth {rr}: {pst["spp"]} spp in {pst["elapsed"]:.2f} s, rMSE {e[0]:.4e} '
            f'(corrected {e[1]:.4e})')
        return img, entry(pst, e, rr_depth=rr)

    plain, plain_d = plain_run(PLAIN_RR_DEPTH, 5)
    if rr_g != PLAIN_RR_DEPTH:
        _, ctrl_d = plain_run(rr_g, 8)
    else:
        ctrl_d = plain_d
    lo, hi = sc.bounds()

    def guided_run(opts):
        # a variant may set its own roulette start ('rr_depth'); the rest are lmgpu_guide_opts fields
        opts = dict(opts) if opts else None
        rr = opts.pop('rr_depth', rr_g) if opts else rr_g
        g = Guide(lo - 1e-3, hi + 1e-3, seed=2)
        gs, _, gst = sc.render_raw(render_time=seconds, max_verts=max_verts, rr_depth=rr, seed=6, guide=g,
                                   guide_opts=opts or None)
        img = gs / gst['spp']
        e = errors(img)
        steps, _ = g.stats()
        d = entry(gst, e, rr_depth=rr, adam_steps=int(steps), training_records=int(gst['guide_records']),
                  options=opts or 'defaults',
                  rmse_ratio_plain_over_guided=plain_d['rmse_corrected'] / e[1],
                  rmse_ratio_control_over_guided=ctrl_d['rmse_corrected'] / e[1])
        log(f'guided {opts or "defaults"}: {gst["spp"]} spp in {gst["elapsed"]:.2f} s, rMSE {e[0]:.4e} '
            f'(corrected {e[1]:.4e}), ratio vs plain {d["rmse_ratio_plain_over_guided"]:.3f}, vs control '
            f'{d["rmse_ratio_control_over_guided"]:.3f}, {steps} Adam steps')
        return img, d

    guided, guided_d = guided_run(guide_opts)
    extra = [guided_run(v)[1] for v in variants or []]
    if save_dir:
        os.makedirs(save_dir, exist_ok=True)
        for name, img in (('reference', ref), ('plain', plain), ('guided', guided)):
            np.save(os.path.join(save_dir, f'c4_{width}x{height}_{name}.npy'), img.astype(np.float32))
    return {
        'scene': f'config 4 (slit-lit room), {width}x{height}, {flat.num_triangles} triangles, max {max_verts} vertices',
        'seconds': seconds,
        'reference': {'spp': ref_spp, 'seconds': round(ref_s, 1), 'rr_depth': max_verts},
        'plain': plain_d,
        'control': ctrl_d,
        'guided': guided_d,
        'rmse_ratio_plain_over_guided': guided_d['rmse_ratio_plain_over_guided'],
        'rmse_ratio_control_over_guided': guided_d['rmse_ratio_control_over_guided'],
        'variants': extra,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--width', type=int, default=1920)
    ap.add_argument('--height', type=int, default=1080)
    ap.add_argument('--time', type=float, default=30.0)
    ap.add_argument('--ref-spp', type=int, default=4096)
    ap.add_argument('--guide-opts', default='', help='JSON dict of lmgpu_guide_opts fields')
    ap.add_argument('--guided-rr-depth', type=int, default=-1,
                    help='Russian roulette start of the guided run (default: max_verts, i.e. none)')
    ap.add_argument('--save-dir', default='')
    ap.add_argument('--variants', default='', help='JSON list of further guide option dicts')
    a = ap.parse_args()
    opts = json.loads(a.guide_opts) if a.guide_opts else None
    out = run(a.width, a.height, a.time, a.ref_spp, guide_opts=opts, save_dir=a.save_dir or None,
              variants=json.loads(a.variants) if a.variants else None,
              guided_rr_depth=None if a.guided_rr_depth < 0 else a.guided_rr_depth)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
