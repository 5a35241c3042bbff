"""ctypes binding of libsgs.so (C ABI in include/sgs.h).

Plays the role of the reference's pybind11 module ``_C`` (render.cu:162-165): ``render``
with the same 29 positional arguments and ``k_buffer_size``. Torch only provides device
memory and the current stream; all compute is in the HIP kernels of libsgs.so. Importing
this module raises ImportError when the library has not been built, so nothing can run on
a silent fallback.
"""
from __future__ import annotations

import ctypes as C
import itertools
import os
import threading
from pathlib import Path
from typing import NamedTuple

import torch  # loads the HIP runtime that libsgs.so then binds to

LIB_PATH = Path(os.environ.get("SGS_LIBRARY", Path(__file__).resolve().parent / "libsgs.so"))
if not LIB_PATH.is_file():
    raise ImportError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                      "or `python stochastic-gaussian-splatting_amd/csrc/build.py`")
_lib = C.CDLL(str(LIB_PATH))

SGS_SUCCESS = 0
SGS_ERROR_INVALID_ARGUMENT = -1
SGS_ERROR_WORKSPACE_TOO_SMALL = -2
SGS_ERROR_HIP = -3


class sgs_gaussians(C.Structure):
    _fields_ = [("means", C.c_void_p), ("scales", C.c_void_p), ("rotations", C.c_void_p),
                ("opacities", C.c_void_p), ("sh_coefficients", C.c_void_p),
                ("n_gaussians", C.c_uint32), ("total_sh_bases", C.c_uint32), ("cov3d", C.c_void_p)]


class sgs_settings(C.Structure):
    _fields_ = [("w2c", C.c_void_p), ("cam_position", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32),
                ("focal_x", C.c_float), ("focal_y", C.c_float), ("center_x", C.c_float), ("center_y", C.c_float),
                ("near_plane", C.c_float), ("far_plane", C.c_float), ("proper_antialiasing", C.c_int32),
                ("supersampling_factor", C.c_int32), ("use_gps", C.c_int32), ("hybrid_scheduling", C.c_int32),
                ("spatial_reuse", C.c_int32)]


class sgs_history(C.Structure):
    _fields_ = [("image", C.c_void_p), ("depth", C.c_void_p), ("alpha", C.c_void_p), ("weight_mass", C.c_void_p)]


class sgs_tuning(C.Structure):
    _fields_ = [("sequential_threshold", C.c_uint32), ("binning_pixel_threshold", C.c_uint32),
                ("binning_point_threshold", C.c_uint32), ("max_binned_gaussians", C.c_uint32),
                ("tile_size", C.c_uint32), ("max_tile_instances", C.c_uint32), ("max_gps_points", C.c_uint32),
                ("schedule", C.c_uint32), ("bin_size", C.c_uint32), ("max_workspace_mib", C.c_uint32)]


class sgs_view(C.Structure):
    _fields_ = [("w2c", C.c_float * 16), ("cam_position", C.c_float * 3), ("width", C.c_uint32),
                ("height", C.c_uint32), ("focal_x", C.c_float), ("focal_y", C.c_float), ("center_x", C.c_float),
                ("center_y", C.c_float), ("near_plane", C.c_float), ("far_plane", C.c_float),
                ("proper_antialiasing", C.c_int32), ("supersampling_factor", C.c_int32), ("spp", C.c_uint32),
                ("frame_idx", C.c_uint32), ("temporal_mode", C.c_int32), ("view_to_previous_view", C.c_float * 16),
                ("history", sgs_history), ("out", sgs_history), ("bin_size", C.c_uint32),
                ("use_gps", C.c_int32), ("hybrid_scheduling", C.c_int32), ("spatial_reuse", C.c_int32),
                ("max_gps_points", C.c_uint32), ("blue_noise", C.c_void_p), ("blue_noise_numel", C.c_uint64),
                ("k_buffer_primitives", C.c_void_p), ("k_buffer_occlusion", C.c_void_p)]


class sgs_batch_tuning(C.Structure):
    _fields_ = [("sequential_threshold", C.c_uint32), ("big_threshold", C.c_uint32),
                ("entry_fragments", C.c_uint32), ("queue_capacity", C.c_uint32), ("view_major", C.c_uint32),
                ("blocks_per_cu", C.c_uint32), ("schedule", C.c_uint32), ("max_tile_instances", C.c_uint32),
                ("tile_samples", C.c_uint32)]


_P = C.POINTER
_lib.sgs_k_buffer_size.restype = C.c_int
_lib.sgs_k_buffer_size.argtypes = []
_lib.sgs_last_error.restype = C.c_char_p
_lib.sgs_last_error.argtypes = []
_lib.sgs_render_workspace_size.restype = C.c_size_t
_lib.sgs_render_workspace_size.argtypes = [_P(sgs_settings), C.c_uint32, _P(sgs_tuning)]
_lib.sgs_render.restype = C.c_int
_lib.sgs_render.argtypes = [_P(sgs_gaussians), _P(sgs_settings), C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                            _P(sgs_history), C.c_void_p, C.c_int32, _P(sgs_history), C.c_uint32, _P(sgs_tuning),
                            C.c_void_p, C.c_size_t, C.c_void_p]
_lib.sgs_render_noalloc.restype = C.c_int
_lib.sgs_render_noalloc.argtypes = [_P(sgs_gaussians), _P(sgs_settings), C.c_void_p, C.c_uint64, C.c_void_p,
                                    C.c_void_p, _P(sgs_history), C.c_void_p, C.c_int32, _P(sgs_history), C.c_uint32,
                                    _P(sgs_tuning), C.c_void_p]
NOALLOC_SCRATCH_BYTES = 128 * 1024 * 1024  # SGS_NOALLOC_SCRATCH_BYTES of include/sgs.h
_lib.sgs_precompute_cov3d.restype = C.c_int
_lib.sgs_precompute_cov3d.argtypes = [_P(sgs_gaussians), C.c_void_p, C.c_void_p]
_lib.sgs_accumulate_workspace_size.restype = C.c_size_t
_lib.sgs_accumulate_workspace_size.argtypes = [_P(sgs_settings), C.c_uint32, _P(sgs_tuning)]
_lib.sgs_accumulate.restype = C.c_int
_lib.sgs_accumulate.argtypes = [_P(sgs_gaussians), _P(sgs_settings), C.c_void_p, C.c_uint64, C.c_void_p,
                                C.c_uint32, _P(sgs_tuning), C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]

_lib.sgs_render_batch_workspace_size.restype = C.c_size_t
_lib.sgs_render_batch_workspace_size.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, _P(sgs_batch_tuning)]
_lib.sgs_batch_view_table_bytes.restype = C.c_size_t
_lib.sgs_batch_view_table_bytes.argtypes = [C.c_uint32]
_lib.sgs_batch_encode_views.restype = C.c_int
_lib.sgs_batch_encode_views.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
_lib.sgs_render_batch.restype = C.c_int
_lib.sgs_render_batch.argtypes = [_P(sgs_gaussians), C.c_void_p, C.c_uint32, C.c_void_p, _P(sgs_batch_tuning),
                                  C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]

_lib.sgs_batch_tile_stats.restype = C.c_int
_lib.sgs_batch_tile_stats.argtypes = [C.c_void_p, C.c_void_p, _P(C.c_uint64), _P(C.c_uint64), _P(C.c_uint32)]

SGS_PLY_MAX_FIELDS = 59


class sgs_ply_layout(C.Structure):
    _fields_ = [("stride", C.c_uint32), ("big_endian", C.c_uint32), ("total_sh_bases", C.c_uint32),
                ("offset", C.c_uint32 * SGS_PLY_MAX_FIELDS), ("type", C.c_uint32 * SGS_PLY_MAX_FIELDS)]


class sgs_scene_arrays(C.Structure):
    _fields_ = [("means", C.c_void_p), ("scales", C.c_void_p), ("rotations", C.c_void_p),
                ("opacities", C.c_void_p), ("sh_coefficients", C.c_void_p), ("n_gaussians", C.c_uint64),
                ("total_sh_bases", C.c_uint32)]


_lib.sgs_ply_decode.restype = C.c_int
_lib.sgs_ply_decode.argtypes = [C.c_void_p, C.c_uint64, _P(sgs_ply_layout), C.c_uint64, _P(sgs_scene_arrays),
                                C.c_void_p]
_lib.sgs_morton_order_workspace_size.restype = C.c_size_t
_lib.sgs_morton_order_workspace_size.argtypes = [C.c_uint64]
_lib.sgs_morton_order.restype = C.c_int
_lib.sgs_morton_order.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                  C.c_void_p]
_lib.sgs_gather_gaussians.restype = C.c_int
_lib.sgs_gather_gaussians.argtypes = [_P(sgs_gaussians), C.c_void_p, _P(sgs_scene_arrays), C.c_void_p]

k_buffer_size: int = _lib.sgs_k_buffer_size()


class Tuning(NamedTuple):
    """Scheduling knobs (sgs_tuning); 0 keeps the library default. Only max_gps_points changes
    results. schedule: 0 or 1 = the tile schedule for stochastic transparency without hybrid
    scheduling (blue noise and debug counters always scatter), 2 = the reference's scatter.
    bin_size: the tile schedule's bin edge in pixels (16, 32 or 64; 0 = 16).
    max_workspace_mib: above this tile-schedule workspace a view takes the scatter schedule
    (0: render() applies default_workspace_cap_mib() of the device)."""
    sequential_threshold: int = 0
    binning_pixel_threshold: int = 0
    binning_point_threshold: int = 0
    max_binned_gaussians: int = 0
    tile_size: int = 0
    max_tile_instances: int = 0
    max_gps_points: int = 0
    schedule: int = 0
    bin_size: int = 0
    max_workspace_mib: int = 0

    def c(self) -> sgs_tuning:
        return sgs_tuning(*[int(v) for v in self])


_caps: dict = {}


def default_workspace_cap_mib(device: torch.device) -> int:
    """A quarter of the device's memory: the most a single-view render() takes for the tile
    schedule's per-call workspace before it falls back to the scatter schedule."""
    idx = device.index if device.index is not None else torch.cuda.current_device()
    if idx not in _caps:
        _caps[idx] = max(1, torch.cuda.get_device_properties(idx).total_memory // 4 >> 20)
    return _caps[idx]


def _capped(tuning: "Tuning | None", device: torch.device) -> "Tuning":
    t = Tuning() if tuning is None else tuning
    return t if t.max_workspace_mib else t._replace(max_workspace_mib=default_workspace_cap_mib(device))


def _check(rc: int, what: str) -> None:
    if rc != SGS_SUCCESS:
        msg = _lib.sgs_last_error().decode(errors="replace")
        if rc in (SGS_ERROR_INVALID_ARGUMENT, SGS_ERROR_WORKSPACE_TOO_SMALL):
            raise ValueError(f"{what}: {msg}")
        raise RuntimeError(f"{what}: {msg}")


# frame counter of the reference's process-global static (render.cu:56-57), made explicit
_frame_counter = itertools.count()
_frame_lock = threading.Lock()


def next_frame_idx() -> int:
    with _frame_lock:
        return next(_frame_counter) & 0xFFFFFFFF


def _f32(t: torch.Tensor, name: str, device: torch.device) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32, got {t.dtype}")
    if t.device != device:
        raise ValueError(f"{name} is on {t.device}, expected {device}")
    return t.contiguous()


def gaussians_struct(means, scales, rotations, opacities, sh_coefficients, keep: list,
                     cov3d_cache: torch.Tensor | None = None) -> sgs_gaussians:
    device = means.device
    if device.type != "cuda":
        raise ValueError("renderer not implemented in CPU mode: tensors must be on a ROCm device")
    n = means.shape[0]
    if means.dim() != 2 or means.shape[1] != 3:
        raise ValueError(f"means must be (N,3), got {tuple(means.shape)}")
    arrays = [_f32(means, "means", device), _f32(scales, "scales", device), _f32(rotations, "rotations", device),
              _f32(opacities, "opacities", device), _f32(sh_coefficients, "sh_coefficients", device)]
    m, s, r, o, sh = arrays
    if s.shape != (n, 3) or r.shape != (n, 4) or o.numel() != n:
        raise ValueError("scales (N,3), rotations (N,4) and opacities (N,) must match means (N,3)")
    if sh.dim() != 3 or sh.shape[0] != n or sh.shape[2] != 3:
        raise ValueError(f"sh_coefficients must be (N,B,3), got {tuple(sh.shape)}")
    keep.extend(arrays)
    cache = None
    if cov3d_cache is not None:
        if cov3d_cache.dtype != torch.float32 or cov3d_cache.shape != (n, 8) or cov3d_cache.device != device \
                or not cov3d_cache.is_contiguous() or cov3d_cache.data_ptr() % 16:
            raise ValueError("cov3d_cache must be the contiguous (N,8) float32 tensor precompute_cov3d returns")
        keep.append(cov3d_cache)
        cache = cov3d_cache.data_ptr()
    return sgs_gaussians(m.data_ptr(), s.data_ptr(), r.data_ptr(), o.data_ptr(), sh.data_ptr(), n, sh.shape[1],
                         cache)


def precompute_cov3d(means, scales, rotations, opacities) -> torch.Tensor:
    """(N,8) float32 cache of the view-independent projection terms (sgs_precompute_cov3d):
    3D covariance, sigmoid opacity and log(255 opacity) per Gaussian. Pass it as
    ``cov3d_cache=`` to render() / accumulate() / the batched renderers; recompute it when
    scales, rotations or opacities change. Renders are bit-identical with and without it."""
    keep: list = []
    n = means.shape[0]
    sh = torch.empty((n, 1, 3), dtype=torch.float32, device=means.device)
    g = gaussians_struct(means, scales, rotations, opacities, sh, keep)
    out = torch.empty((n, 8), dtype=torch.float32, device=means.device)
    with torch.cuda.device(means.device):
        _check(_lib.sgs_precompute_cov3d(C.byref(g), out.data_ptr(), torch.cuda.current_stream(means.device).cuda_stream),
               "sgs_precompute_cov3d")
    return out


def settings_struct(w2c, cam_position, width, height, focal_x, focal_y, center_x, center_y, near_plane, far_plane,
                    proper_antialiasing, supersampling_factor, use_gps, hybrid_scheduling, spatial_reuse,
                    device: torch.device, keep: list) -> sgs_settings:
    w = torch.as_tensor(w2c).to(device=device, dtype=torch.float32).reshape(4, 4).contiguous()
    cp = torch.as_tensor(cam_position).to(device=device, dtype=torch.float32).reshape(3).contiguous()
    keep.extend([w, cp])
    return sgs_settings(w.data_ptr(), cp.data_ptr(), int(width), int(height), float(focal_x), float(focal_y),
                        float(center_x), float(center_y), float(near_plane), float(far_plane),
                        int(bool(proper_antialiasing)), int(supersampling_factor), int(bool(use_gps)),
                        int(bool(hybrid_scheduling)), int(bool(spatial_reuse)))


def _ptr_or_null(t: torch.Tensor | None) -> int | None:
    return None if t is None or t.numel() == 0 else t.data_ptr()


def _workspace(nbytes: int, device: torch.device) -> torch.Tensor:
    ws = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)
    if ws.data_ptr() % 256:
        ws = torch.empty(int(nbytes) + 256, dtype=torch.uint8, device=device)
        off = (-ws.data_ptr()) % 256
        ws = ws[off:off + int(nbytes)]
    return ws


def render(means, scales, rotations, opacities, sh_coefficients, w2c, cam_position, width, height, focal_x, focal_y,
           center_x, center_y, near_plane, far_plane, proper_antialiasing, supersampling_factor, use_gps,
           hybrid_scheduling, spatial_reuse, blue_noise, k_buffer_primitives, k_buffer_occlusion, history_image,
           history_depth, history_alpha, history_weight_mass, view_to_previous_view, temporal_mode, *,
           frame_idx: int | None = None, tuning: Tuning | None = None, return_framebuffer: bool = False,
           cov3d_cache: torch.Tensor | None = None):
    """Same positional arguments and result as _C.render (render.cu:16-45): fresh
    (image (H,W,3), depth, alpha, weight_mass (H,W,1)) tensors. ``frame_idx`` defaults to the
    process-global counter; ``return_framebuffer`` appends the packed framebuffer (P,) int64;
    ``cov3d_cache`` is precompute_cov3d's output for these Gaussians. The tile schedule's
    workspace is capped by ``tuning.max_workspace_mib`` (default: default_workspace_cap_mib),
    above which the view takes the scatter schedule."""
    device = means.device
    keep: list = []
    g = gaussians_struct(means, scales, rotations, opacities, sh_coefficients, keep, cov3d_cache)
    s = settings_struct(w2c, cam_position, width, height, focal_x, focal_y, center_x, center_y, near_plane, far_plane,
                        proper_antialiasing, supersampling_factor, use_gps, hybrid_scheduling, spatial_reuse, device, keep)
    mode = int(temporal_mode)
    H, W = int(height), int(width)
    with torch.cuda.device(device):
        opts = dict(dtype=torch.float32, device=device)
        image = torch.empty((H, W, 3), **opts)
        depth = torch.empty((H, W, 1), **opts)
        alpha = torch.empty((H, W, 1), **opts)
        weight_mass = torch.empty((H, W, 1), **opts)
        hist_t = [history_image, history_depth, history_alpha, history_weight_mass]
        hist = None
        if mode != 0:
            hist_t = [_f32(h, n, device) for h, n in zip(hist_t, ("history.image", "history.depth", "history.alpha",
                                                                   "history.weight_mass"))]
            if hist_t[0].shape != (H, W, 3) or any(h.numel() != H * W for h in hist_t[1:]):
                raise ValueError("history tensors do not match the image size")
            keep.extend(hist_t)
            hist = sgs_history(*[h.data_ptr() for h in hist_t])
        v2p = None
        if mode == 2:
            v2p_t = torch.as_tensor(view_to_previous_view).to(device=device, dtype=torch.float32).reshape(4, 4)
            v2p_t = v2p_t.contiguous()
            keep.append(v2p_t)
            v2p = v2p_t.data_ptr()
        kp = ko = None
        if spatial_reuse:
            P = W * H * max(1, int(supersampling_factor)) ** 2
            if k_buffer_primitives.shape != (k_buffer_size, P, 2) or k_buffer_occlusion.shape != (k_buffer_size, P, 4):
                raise ValueError("k-buffer does not match the settings (use create_k_buffer)")
            if not (k_buffer_primitives.is_contiguous() and k_buffer_occlusion.is_contiguous()):
                raise ValueError("k-buffer tensors must be contiguous (they are updated in place)")
            kp, ko = k_buffer_primitives.data_ptr(), k_buffer_occlusion.data_ptr()
        bn = None
        bn_numel = 0
        if blue_noise is not None and blue_noise.numel() > 0:
            bn_t = blue_noise.to(device=device, dtype=torch.uint8).contiguous()
            keep.append(bn_t)
            bn, bn_numel = bn_t.data_ptr(), bn_t.numel()
        tc = _capped(tuning, device).c()
        tp = C.byref(tc)
        nbytes = _lib.sgs_render_workspace_size(C.byref(s), g.n_gaussians, tp)
        ws = _workspace(nbytes, device)
        fidx = next_frame_idx() if frame_idx is None else int(frame_idx) & 0xFFFFFFFF
        out = sgs_history(image.data_ptr(), depth.data_ptr(), alpha.data_ptr(), weight_mass.data_ptr())
        stream = torch.cuda.current_stream(device).cuda_stream
        rc = _lib.sgs_render(C.byref(g), C.byref(s), bn, bn_numel, kp, ko, C.byref(hist) if hist else None, v2p, mode,
                             C.byref(out), fidx, tp, ws.data_ptr(), ws.numel(), stream)
        _check(rc, "sgs_render")
        if return_framebuffer:
            P = W * H * max(1, int(supersampling_factor)) ** 2
            fb = ws[:8 * P].view(torch.int64).clone()
            return image, depth, alpha, weight_mass, fb
        return image, depth, alpha, weight_mass


def accumulate(means, scales, rotations, opacities, sh_coefficients, w2c, cam_position, width, height, focal_x,
               focal_y, center_x, center_y, near_plane, far_plane, proper_antialiasing, supersampling_factor, use_gps,
               hybrid_scheduling, spatial_reuse=False, blue_noise=None, *, frame_idx: int, tuning: Tuning | None = None,
               framebuffer: torch.Tensor | None = None, counters: torch.Tensor | None = None,
               workspace: torch.Tensor | None = None, cov3d_cache: torch.Tensor | None = None) -> torch.Tensor:
    """Accumulation stage only (sgs_accumulate): returns the packed framebuffer as
    (H*ss, W*ss) int64 (uint64 bits). ``counters`` (4,) int64 receives the work counts."""
    device = means.device
    keep: list = []
    g = gaussians_struct(means, scales, rotations, opacities, sh_coefficients, keep, cov3d_cache)
    s = settings_struct(w2c, cam_position, width, height, focal_x, focal_y, center_x, center_y, near_plane, far_plane,
                        proper_antialiasing, supersampling_factor, use_gps, hybrid_scheduling, spatial_reuse, device, keep)
    ss = max(1, int(supersampling_factor))
    FH, FW = int(height) * ss, int(width) * ss
    with torch.cuda.device(device):
        if framebuffer is None:
            framebuffer = torch.empty((FH, FW), dtype=torch.int64, device=device)
        elif framebuffer.numel() != FH * FW or framebuffer.dtype != torch.int64 or not framebuffer.is_contiguous():
            raise ValueError("framebuffer must be a contiguous int64 tensor of (H*ss)*(W*ss) elements")
        tc = _capped(tuning, device).c()
        tp = C.byref(tc)
        nbytes = _lib.sgs_accumulate_workspace_size(C.byref(s), g.n_gaussians, tp)
        if nbytes and (workspace is None or workspace.numel() < nbytes or workspace.data_ptr() % 256):
            workspace = _workspace(nbytes, device)
        bn = None
        bn_numel = 0
        if blue_noise is not None and blue_noise.numel() > 0:
            bn_t = blue_noise.to(device=device, dtype=torch.uint8).contiguous()
            keep.append(bn_t)
            bn, bn_numel = bn_t.data_ptr(), bn_t.numel()
        cptr = None
        if counters is not None:
            if counters.dtype != torch.int64 or counters.numel() < 4 or not counters.is_contiguous():
                raise ValueError("counters must be a contiguous int64 tensor of 4 elements")
            cptr = counters.data_ptr()
        stream = torch.cuda.current_stream(device).cuda_stream
        rc = _lib.sgs_accumulate(C.byref(g), C.byref(s), bn, bn_numel, framebuffer.data_ptr(), int(frame_idx) & 0xFFFFFFFF,
                                 tp, cptr, workspace.data_ptr() if nbytes else None, nbytes, stream)
        _check(rc, "sgs_accumulate")
    return framebuffer.view(FH, FW)


def scene_arrays(arrays: dict) -> sgs_scene_arrays:
    """sgs_scene_arrays of a dict with means, scales, rotations, opacities, sh (device float32)."""
    return sgs_scene_arrays(arrays["means"].data_ptr(), arrays["scales"].data_ptr(), arrays["rotations"].data_ptr(),
                            arrays["opacities"].data_ptr(), arrays["sh"].data_ptr(), int(arrays["means"].shape[0]),
                            int(arrays["sh"].shape[1]))


def ply_decode(records: torch.Tensor, layout: sgs_ply_layout, first: int, out: dict) -> None:
    """Decodes the vertex records held in ``records`` (device uint8, whole records) into rows
    ``first``.. of ``out`` on the current stream."""
    n = records.numel() // max(int(layout.stride), 1)
    o = scene_arrays(out)
    stream = torch.cuda.current_stream(records.device).cuda_stream
    _check(_lib.sgs_ply_decode(records.data_ptr() if n else None, n, C.byref(layout), int(first), C.byref(o), stream),
           "sgs_ply_decode")


def morton_order(means: torch.Tensor, return_codes: bool = False):
    """Stable Morton order of ``means`` (device (N,3) float32) as int64 indices (and the
    sorted 63-bit codes)."""
    n = int(means.shape[0])
    device = means.device
    order = torch.empty(n, dtype=torch.int32, device=device)
    codes = torch.empty(n, dtype=torch.int64, device=device)
    if n:
        nbytes = _lib.sgs_morton_order_workspace_size(n)
        if nbytes == 0:
            raise ValueError(f"morton_order: {n} means exceed the supported 2^31 - 1")
        ws = _workspace(nbytes, device)
        stream = torch.cuda.current_stream(device).cuda_stream
        _check(_lib.sgs_morton_order(means.data_ptr(), n, order.data_ptr(), codes.data_ptr(), ws.data_ptr(), nbytes,
                                     stream), "sgs_morton_order")
    order = order.to(torch.int64)
    return (order, codes) if return_codes else order


def gather_gaussians(src: dict, order: torch.Tensor) -> dict:
    """New arrays with row i = row order[i] of ``src`` (order: device int32, values < N)."""
    n = int(order.shape[0])
    out = {k: torch.empty((n, *v.shape[1:]), dtype=torch.float32, device=v.device) for k, v in src.items()}
    if n:
        keep: list = []
        g = gaussians_struct(src["means"], src["scales"], src["rotations"], src["opacities"], src["sh"], keep)
        o = scene_arrays(out)
        stream = torch.cuda.current_stream(order.device).cuda_stream
        _check(_lib.sgs_gather_gaussians(C.byref(g), order.data_ptr(), C.byref(o), stream), "sgs_gather_gaussians")
    return out


def library():
    """The loaded ctypes.CDLL (for symbol checks)."""
    return _lib
