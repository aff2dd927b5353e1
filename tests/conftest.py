import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def kf():
    import kinect_fusion_slam_amd as k
    return k


@pytest.fixture(scope="session")
def oracle():
    from oracle import kf_oracle
    return kf_oracle


def say(msg):
    """Progress line that names the case before a GPU step (flushed)."""
    print(msg, flush=True)


@pytest.fixture(scope="session")
def room_stream(kf):
    """640x480 room, 360x360x300 @10mm: first 10 frames of the synthetic orbit."""
    vol, prm, K, H, W = kf.preset(kf.PRESET_ROOM)
    scene = kf.preset_scene(kf.PRESET_ROOM)
    frames, gt = kf.stream(scene, K, H, W, 0, 10)
    return vol, prm, K, H, W, frames, gt


def small_setup(kf, H=48, W=64, dims=(40, 40, 36), start=(-20, -20, 4), voxel=20.0):
    """A small volume/camera for fast exhaustive parity tests."""
    vol, prm, K, _, _ = kf.preset(kf.PRESET_ROOM)
    vol.dims[:] = list(dims)
    vol.start[:] = list(start)
    vol.voxel_mm = voxel
    prm.truncation_mm = 60.0
    prm.sdf_min_mm = -60.0
    prm.sdf_max_mm = 600.0
    K = np.array([60.0, 0, 31.5, 0, 60.0, 23.5, 0, 0, 1], np.float32)
    return vol, prm, K, H, W
