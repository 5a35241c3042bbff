"""Shared test setup: import paths, the ``gpu`` marker and scene fixtures."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parents[1]
PKG_DIR = REPO / "stochastic-gaussian-splatting_amd"
for p in (str(REPO), str(PKG_DIR)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built libsgs.so")


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.build()
    return oracle


def small_scene(n=300, seed=0, **kw):
    from StochasticGSCudaBackend.synthetic import SceneSpec, make_scene
    spec = SceneSpec(n=n, seed=seed, **kw)
    return make_scene(spec)


def small_camera(width=64, height=48, distance=3.0, angle=0.0, **flags):
    from StochasticGSCudaBackend.synthetic import camera, orbit_eye
    return camera(orbit_eye(angle, distance), width, height, **flags)


def oracle_scene(arrays):
    from oracle.oracle import OracleScene
    return OracleScene(arrays["means"], arrays["scales"], arrays["rotations"], arrays["opacities"], arrays["sh"])


def ulp_distance(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """|a - b| in float32 ulps of b."""
    a = np.asarray(a, np.float32).astype(np.float64)
    b64 = np.asarray(b, np.float64)
    spacing = np.spacing(np.abs(b64).astype(np.float32)).astype(np.float64)
    return np.abs(a - b64) / spacing
