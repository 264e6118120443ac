"""Quantised BVH nodes (csrc/bvh.h): the host build of quantise_node and
qplane, compiled with hipcc, decodes 4- and 8-wide boxes that contain the
exact child boxes on 200,000 random nodes of each width (1e-7..1e3 extents,
offsets up to 5e4, flat boxes, partly empty nodes), stays within one grid
step of them, and encodes stack entries and empty slots as the traversal
reads them. The traversal's slab test of a quantised node (child_keys: one
fma per plane from the exact grid, the near and far plane picked by the
direction's sign, offsets widened by a rounding bound) never misses a child
whose exact box the ray meets: 2.3 million rays aimed at faces, edges and
corners, starting on the box's planes or parallel to them, decided in exact
arithmetic (with the widening removed, the first node already fails). GPU
hit parity with the quantised traversal is covered by the accel tests."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'


@pytest.mark.skipif(not os.path.exists(HIPCC), reason='hipcc not available')
def test_quantised_node_contains_children(tmp_path):
    exe = tmp_path / 'qnode_check'
    src = os.path.join(ROOT, 'tests', 'native', 'qnode_check.cpp')
    subprocess.run([HIPCC, '-O2', '-std=c++17', '-x', 'hip', '--offload-arch=gfx950', src, '-o', str(exe)],
                   check=True, capture_output=True, timeout=600)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.startswith('ok'), r.stdout + r.stderr
