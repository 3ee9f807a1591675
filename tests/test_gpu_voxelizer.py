"""The voxelizer's inside tests on the GPU (lbm_voxelize, reached through
lbmh_voxelize_off with threads < 0) against the host voxelizer and the
reference's committed voxelization: the same grid, bit for bit."""
import os
import time

import numpy as np
import pytest

from lb_amd import host
from lb_amd.geometry import obstacle_scene, write_off
from test_voxelizer import GOLD_OFF, OUTSIDE, golden_grid

pytestmark = pytest.mark.gpu


def test_golden_grid_on_gpu(gpu):
    dims, gold = golden_grid()
    g, mn, mx = host.voxelize_off(GOLD_OFF, 0.25, 1.0, OUTSIDE, device=True)
    assert list(g.shape[::-1]) == dims
    assert np.array_equal(g, gold)


@pytest.mark.parametrize("res", [0.25, 0.5])
def test_obstacle_scene_gpu_equals_host(gpu, tmp_path, res):
    v, f, _ = obstacle_scene()
    path = os.path.join(tmp_path, "obstacle.off")
    write_off(path, v, f)
    out = (-3.1416, -2.7183, -1.4142)
    t0 = time.time()
    gh, mn, mx = host.voxelize_off(path, res, 0.0, out)
    t1 = time.time()
    gd, mn2, mx2 = host.voxelize_off(path, res, 0.0, out, device=True)
    t2 = time.time()
    print(f"[voxelizer res={res}] grid {gh.shape[::-1]} triangles {len(f)}: host {t1 - t0:.3f} s, "
          f"gpu {t2 - t1:.3f} s (incl. upload and hierarchy copy); inside fraction {gh.mean():.4f}", flush=True)
    assert (mn, mx) == (mn2, mx2)
    assert 0.5 < gh.mean() < 1.0
    assert np.array_equal(gh, gd)


def test_obstacle_workload_full_size_gpu_equals_host(gpu, tmp_path):
    """The obstacle workload's 512 x 256 x 256 grid (resolution 1)."""
    v, f, _ = obstacle_scene()
    path = os.path.join(tmp_path, "obstacle.off")
    write_off(path, v, f)
    out = (-3.1416, -2.7183, -1.4142)
    t0 = time.time()
    gh, _, _ = host.voxelize_off(path, 1.0, 0.0, out)
    t1 = time.time()
    gd, _, _ = host.voxelize_off(path, 1.0, 0.0, out, device=True)
    t2 = time.time()
    print(f"[voxelizer 512x256x256] host {t1 - t0:.3f} s, gpu {t2 - t1:.3f} s", flush=True)
    assert gh.shape == (256, 256, 512)
    assert np.array_equal(gh, gd)
