"""Writes c1_traversal_stats.json: mean node visits and triangle tests per
ray of the reference's binary SAH BVH + unordered traversal
(accel::sahbvh, restated by the oracle) on the config-1 ray sets. These are
the N_node / N_tri of the algorithmic bytes per ray used by bench.py:
B_ray = 32 + 16 + N_node * 32 + N_tri * 36 bytes.

Run: python tests/golden/make_c1_traversal_stats.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'lightmetrica-v3_amd'))

from lm_amd import scenes  # noqa: E402
from oracle import oracle as orc  # noqa: E402


def main():
    scene, rays = scenes.config1()
    flat = scene.flatten()
    t0 = time.time()
    acc = orc.Accel(flat.verts)
    build_s = time.time() - t0
    out = {'scene': 'config1 (seed 0x5EED): displaced icosphere L8 + 10K clutter',
           'num_tris': flat.num_triangles, 'sah_nodes': acc.num_nodes, 'oracle_build_s': build_s}
    for label in ('coherent', 'incoherent'):
        r = rays[label]
        res = acc.intersect(r, stats=True)
        n = r['tmin'].shape[0]
        out[label] = {'rays': n, 'node_visits_per_ray': res['node_visits'] / n,
                      'tri_tests_per_ray': res['tri_tests'] / n,
                      'hit_fraction': float((res['prim'] >= 0).mean())}
        nb = out[label]['node_visits_per_ray']
        nt = out[label]['tri_tests_per_ray']
        out[label]['bytes_per_ray'] = 32 + 16 + nb * 32 + nt * 36
        out[label]['flop_per_ray'] = nb * 20 + nt * 40
    with open(os.path.join(HERE, 'c1_traversal_stats.json'), 'w') as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == '__main__':
    main()
