"""Watertight ray/triangle test (csrc/bvh.h, tri_intersect): the host build,
compiled by hipcc with FMA instructions and HIP's default floating-point
contraction (as the device code), hits a closed icosphere with every ray
fired through a shared edge or vertex toward its centre (9 rotations,
scales 0.05-250, offsets up to 1500; 109,458 rays). The reference's
Moller-Trumbore test (src/accel/accel_sahbvh.cpp:58-75) has closed edges, so
these rays hit there too; the GPU counterpart is
tests/test_accel_gpu.py::test_shared_edges_and_vertices_always_hit."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'


@pytest.mark.skipif(not os.path.exists(HIPCC), reason='hipcc not available')
def test_shared_edges_and_vertices_hit_on_host(tmp_path):
    exe = tmp_path / 'watertight_check'
    src = os.path.join(ROOT, 'tests', 'native', 'watertight_check.cpp')
    subprocess.run([HIPCC, '-O2', '-std=c++17', '-x', 'hip', '--offload-arch=gfx950', '-Xarch_host', '-mfma',
                    src, '-o', str(exe)], check=True, capture_output=True, timeout=600)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.startswith('ok'), r.stdout + r.stderr
