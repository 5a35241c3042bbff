"""ctypes front end of the scalar CPU oracle (liboracle_sgs.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` as the checker; the renderer never calls it.
Arrays are numpy; the C side is documented in ``sgs_oracle.h``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
LIB_PATH = _DIR / "liboracle_sgs.so"

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


class Scene(C.Structure):
    _fields_ = [
        ("means", C.c_void_p), ("scales", C.c_void_p), ("rotations", C.c_void_p),
        ("opacities", C.c_void_p), ("sh", C.c_void_p),
        ("n_gaussians", C.c_uint32), ("n_sh_bases", C.c_uint32),
    ]


class Camera(C.Structure):
    _fields_ = [
        ("w2c", C.c_float * 16), ("cam_position", C.c_float * 3),
        ("width", C.c_uint32), ("height", C.c_uint32),
        ("focal_x", C.c_float), ("focal_y", C.c_float), ("center_x", C.c_float), ("center_y", C.c_float),
        ("near_plane", C.c_float), ("far_plane", C.c_float),
        ("proper_antialiasing", C.c_int32), ("supersampling_factor", C.c_int32),
        ("use_gps", C.c_int32), ("spatial_reuse", C.c_int32), ("max_gps_points", C.c_uint32),
    ]


class Proj(C.Structure):
    _fields_ = [
        ("active", C.c_int32), ("in_depth_range", C.c_int32),
        ("depth", C.c_float), ("opacity", C.c_float), ("determinant", C.c_float), ("power_threshold", C.c_float),
        ("mean2d", C.c_float * 2), ("conic", C.c_float * 3), ("cov2d", C.c_float * 3), ("bounds", C.c_float * 4),
    ]


class Stats(C.Structure):
    _fields_ = [("n_in_depth_range", C.c_uint64), ("n_projected", C.c_uint64),
                ("n_tested", C.c_uint64), ("n_accepted", C.c_uint64)]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


def build() -> Path:
    """Compiles liboracle_sgs.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", str(_DIR), "liboracle_sgs.so"], check=True)
    return LIB_PATH


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        f, u32, i32, vp = C.c_float, C.c_uint32, C.c_int32, C.c_void_p
        sig = {
            "orc_exp2f": (f, [f]), "orc_expf": (f, [f]), "orc_logf": (f, [f]),
            "orc_sincos_2pi": (None, [f, C.POINTER(f), C.POINTER(f)]),
            "orc_hash32": (u32, [u32]), "orc_make_seed": (None, [u32, u32, _u32p]),
            "orc_pcg2d_uint": (None, [_u32p, _u32p]), "orc_pcg2d": (None, [_u32p, _f32p]),
            "orc_pcg_hash": (u32, [u32]), "orc_rand": (f, [u32, u32, u32, u32]),
            "orc_blue_noise": (f, [vp, u32, u32, u32, u32]),
            "orc_gps_point_seed": (None, [u32, u32, u32, _u32p]),
            "orc_dilog": (f, [f]), "orc_inv_dilog": (f, [f]),
            "orc_box_muller": (None, [f, f, _f32p]),
            "orc_corrected_box_muller": (None, [f, f, f, f, _f32p]),
            "orc_poisson": (u32, [_u32p, f]),
            "orc_project": (None, [C.POINTER(Scene), u32, _f32p, f, f, f, f, f, f, f, f, i32, f, C.POINTER(Proj)]),
            "orc_eval_sh": (None, [_f32p, _f32p, u32, _f32p]),
            "orc_catmull_rom_weights": (None, [f, _f32p]),
            "orc_rgb_to_ycocg": (None, [_f32p, _f32p]),
            "orc_accumulate": (None, [C.POINTER(Scene), C.POINTER(Camera), u32, vp, u32, _u32p, _u64p,
                                      C.POINTER(Stats), C.c_int]),
            "orc_accumulate_st_bruteforce": (None, [C.POINTER(Scene), C.POINTER(Camera), u32, vp, u32, _u32p, _u64p]),
            "orc_gps_counts": (None, [C.POINTER(Scene), C.POINTER(Camera), u32, _u32p, _u32p, _f32p]),
            "orc_sample": (None, [C.POINTER(Scene), C.POINTER(Camera), _u64p, _f32p, vp, vp, vp]),
            "orc_resolve": (C.c_int, [C.POINTER(Camera), _f32p, vp, vp, vp, vp, vp, vp, i32,
                                      _f32p, _f32p, _f32p, _f32p]),
            "orc_render": (C.c_int, [C.POINTER(Scene), C.POINTER(Camera), u32, vp, u32, vp, vp,
                                     vp, vp, vp, vp, vp, i32, _f32p, _f32p, _f32p, _f32p, vp,
                                     C.POINTER(Stats), C.c_int]),
            "orc_blend_fp64": (None, [C.POINTER(Scene), C.POINTER(Camera), _f64p, _f64p, _f64p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray | None) -> int | None:
    return None if a is None else a.ctypes.data


@dataclass
class OracleScene:
    """numpy-owned Gaussian arrays plus the ctypes view the C side reads."""
    means: np.ndarray
    scales: np.ndarray
    rotations: np.ndarray
    opacities: np.ndarray
    sh: np.ndarray

    def __post_init__(self):
        self.means = np.ascontiguousarray(self.means, np.float32)
        self.scales = np.ascontiguousarray(self.scales, np.float32)
        self.rotations = np.ascontiguousarray(self.rotations, np.float32)
        self.opacities = np.ascontiguousarray(self.opacities, np.float32).reshape(-1)
        self.sh = np.ascontiguousarray(self.sh, np.float32)
        n = self.means.shape[0]
        assert self.means.shape == (n, 3) and self.scales.shape == (n, 3)
        assert self.rotations.shape == (n, 4) and self.opacities.shape == (n,)
        assert self.sh.ndim == 3 and self.sh.shape[0] == n and self.sh.shape[2] == 3
        self.c = Scene(_ptr(self.means), _ptr(self.scales), _ptr(self.rotations), _ptr(self.opacities),
                       _ptr(self.sh), n, self.sh.shape[1])

    @property
    def n(self) -> int:
        return self.means.shape[0]


def make_camera(cam: dict) -> Camera:
    """``cam`` holds the fields of Settings (image pixels) plus optional max_gps_points."""
    c = Camera()
    c.w2c[:] = [float(v) for v in np.asarray(cam["w2c"], np.float32).reshape(16)]
    c.cam_position[:] = [float(v) for v in np.asarray(cam["cam_position"], np.float32).reshape(3)]
    c.width, c.height = int(cam["width"]), int(cam["height"])
    c.focal_x, c.focal_y = float(cam["focal_x"]), float(cam["focal_y"])
    c.center_x, c.center_y = float(cam["center_x"]), float(cam["center_y"])
    c.near_plane, c.far_plane = float(cam["near_plane"]), float(cam["far_plane"])
    c.proper_antialiasing = int(bool(cam.get("proper_antialiasing", False)))
    c.supersampling_factor = int(cam.get("supersampling_factor", 1))
    c.use_gps = int(bool(cam.get("use_gps", False)))
    c.spatial_reuse = int(bool(cam.get("spatial_reuse", False)))
    c.max_gps_points = int(cam.get("max_gps_points", 0))
    return c


def fb_dims(cam: dict) -> tuple[int, int]:
    ss = max(1, int(cam.get("supersampling_factor", 1)))
    return int(cam["width"]) * ss, int(cam["height"]) * ss


def accumulate(scene: OracleScene, cam: dict, frame_idx: int, crop=None, blue_noise_slice=None,
               blue_noise_rotation: int = 0, n_threads: int = 1) -> tuple[np.ndarray, dict]:
    """Packed framebuffer of the crop (x0, y0, x1, y1) as (y1-y0, x1-x0) uint64 and the work counters."""
    W, H = fb_dims(cam)
    x0, y0, x1, y1 = crop if crop is not None else (0, 0, W, H)
    crop_a = np.array([x0, y0, x1, y1], np.uint32)
    fb = np.full((y1 - y0) * (x1 - x0), np.iinfo(np.uint64).max, np.uint64)
    st = Stats()
    bn = None if blue_noise_slice is None else np.ascontiguousarray(blue_noise_slice, np.uint8)
    lib().orc_accumulate(C.byref(scene.c), C.byref(make_camera(cam)), frame_idx, _ptr(bn),
                         blue_noise_rotation, crop_a, fb, C.byref(st), n_threads)
    return fb.reshape(y1 - y0, x1 - x0), st.as_dict()


def accumulate_bruteforce(scene: OracleScene, cam: dict, frame_idx: int, crop=None, blue_noise_slice=None,
                          blue_noise_rotation: int = 0) -> np.ndarray:
    W, H = fb_dims(cam)
    x0, y0, x1, y1 = crop if crop is not None else (0, 0, W, H)
    fb = np.empty((y1 - y0) * (x1 - x0), np.uint64)
    bn = None if blue_noise_slice is None else np.ascontiguousarray(blue_noise_slice, np.uint8)
    lib().orc_accumulate_st_bruteforce(C.byref(scene.c), C.byref(make_camera(cam)), frame_idx, _ptr(bn),
                                       blue_noise_rotation, np.array([x0, y0, x1, y1], np.uint32), fb)
    return fb.reshape(y1 - y0, x1 - x0)


def gps_counts(scene: OracleScene, cam: dict, frame_idx: int):
    n = scene.n
    capped = np.zeros(n, np.uint32)
    true = np.zeros(n, np.uint32)
    imp = np.zeros(n, np.float32)
    lib().orc_gps_counts(C.byref(scene.c), C.byref(make_camera(cam)), frame_idx, capped, true, imp)
    return capped, true, imp


def project(scene: OracleScene, idx: int, cam: dict) -> Proj:
    W, H = fb_dims(cam)
    ss = float(max(1, int(cam.get("supersampling_factor", 1))))
    p = Proj()
    lib().orc_project(C.byref(scene.c), idx, np.ascontiguousarray(np.asarray(cam["w2c"], np.float32).reshape(16)),
                      float(W), float(H), cam["focal_x"] * ss, cam["focal_y"] * ss, cam["center_x"] * ss,
                      cam["center_y"] * ss, cam["near_plane"], cam["far_plane"],
                      int(bool(cam.get("proper_antialiasing", False))), ss, C.byref(p))
    return p


def empty_k_buffer(cam: dict) -> tuple[np.ndarray, np.ndarray]:
    W, H = fb_dims(cam)
    k = 4 if cam.get("spatial_reuse", False) else 0
    return np.full((k, W * H, 2), -1, np.int32), np.zeros((k, W * H, 4), np.float32)


def render(scene: OracleScene, cam: dict, frame_idx: int, history=None, temporal_mode: int = 0,
           view_to_previous_view=None, k_buffer=None, blue_noise=None, n_threads: int = 1):
    """One stochastic sample: returns (image, depth, alpha, weight_mass, framebuffer, stats).

    ``k_buffer`` (int32 (K,P,2), float32 (K,P,4)) is updated in place like the renderer's."""
    W, H = int(cam["width"]), int(cam["height"])
    FW, FH = fb_dims(cam)
    img = np.zeros((H, W, 3), np.float32)
    dep = np.zeros((H, W, 1), np.float32)
    alp = np.zeros((H, W, 1), np.float32)
    wm = np.zeros((H, W, 1), np.float32)
    fb = np.empty(FW * FH, np.uint64)
    if history is None:
        history = (None, None, None, None)
    hist = [None if h is None else np.ascontiguousarray(h, np.float32) for h in history]
    v2p = np.ascontiguousarray(np.eye(4, dtype=np.float32) if view_to_previous_view is None
                               else np.asarray(view_to_previous_view, np.float32))
    if k_buffer is None:
        k_buffer = empty_k_buffer(cam)
    kp, ko = k_buffer
    bn = None if blue_noise is None or np.size(blue_noise) == 0 else np.ascontiguousarray(blue_noise, np.uint8)
    st = Stats()
    rc = lib().orc_render(C.byref(scene.c), C.byref(make_camera(cam)), frame_idx, _ptr(bn),
                          0 if bn is None else bn.size, _ptr(kp) if kp.size else None,
                          _ptr(ko) if ko.size else None, *[_ptr(h) for h in hist], _ptr(v2p),
                          int(temporal_mode), img, dep, alp, wm, _ptr(fb), C.byref(st), n_threads)
    if rc != 0:
        raise ValueError(f"invalid temporal mode {temporal_mode}")
    return img, dep, alp, wm, fb.reshape(FH, FW), st.as_dict()


def shade_resolve(scene: OracleScene, cam: dict, framebuffer: np.ndarray, history=None, temporal_mode: int = 0,
                  view_to_previous_view=None):
    """Shading and resolve of a given packed framebuffer (sample_cu + resolve_cu without
    spatial reuse): (image, depth, alpha, weight_mass)."""
    W, H = int(cam["width"]), int(cam["height"])
    FW, FH = fb_dims(cam)
    fb = np.ascontiguousarray(np.asarray(framebuffer, np.uint64).reshape(-1))
    assert fb.size == FW * FH
    c = make_camera(cam)
    c.spatial_reuse = 0
    samples = np.zeros(FW * FH * 4, np.float32)
    lib().orc_sample(C.byref(scene.c), C.byref(c), fb, samples, None, None, None)
    outs = [np.zeros((H, W, 3), np.float32)] + [np.zeros((H, W, 1), np.float32) for _ in range(3)]
    hist = [None] * 4 if history is None else [np.ascontiguousarray(h, np.float32) for h in history]
    v2p = np.ascontiguousarray(np.eye(4, dtype=np.float32) if view_to_previous_view is None
                               else np.asarray(view_to_previous_view, np.float32))
    rc = lib().orc_resolve(C.byref(c), samples, None, *[_ptr(h) for h in hist], _ptr(v2p), int(temporal_mode), *outs)
    if rc != 0:
        raise ValueError(f"invalid temporal mode {temporal_mode}")
    return tuple(outs)


def blend_fp64(scene: OracleScene, cam: dict):
    W, H = int(cam["width"]), int(cam["height"])
    img = np.zeros((H, W, 3), np.float64)
    dep = np.zeros((H, W), np.float64)
    alp = np.zeros((H, W), np.float64)
    lib().orc_blend_fp64(C.byref(scene.c), C.byref(make_camera(cam)), img, dep, alp)
    return img, dep, alp


# scalar helpers --------------------------------------------------------------------------------
def exp2f(x): return lib().orc_exp2f(x)
def expf(x): return lib().orc_expf(x)
def logf(x): return lib().orc_logf(x)


def sincos_2pi(u):
    s, c = C.c_float(), C.c_float()
    lib().orc_sincos_2pi(u, C.byref(s), C.byref(c))
    return s.value, c.value


def make_seed(idx, frame):
    out = np.zeros(2, np.uint32)
    lib().orc_make_seed(idx, frame, out)
    return int(out[0]), int(out[1])


def pcg2d_uint(a, b):
    out = np.zeros(2, np.uint32)
    lib().orc_pcg2d_uint(np.array([a, b], np.uint32), out)
    return int(out[0]), int(out[1])


def poisson(seed, lam):
    st = np.array(seed, np.uint32)
    return lib().orc_poisson(st, lam)


def eval_sh(coeffs, direction, n_bases):
    out = np.zeros(3, np.float32)
    lib().orc_eval_sh(np.ascontiguousarray(coeffs, np.float32), np.ascontiguousarray(direction, np.float32),
                      n_bases, out)
    return out


def n_threads_available() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1
