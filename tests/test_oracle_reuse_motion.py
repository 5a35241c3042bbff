"""Closed-form pins of the oracle's spatial reuse (resolve.cu:81-143, 225-230) and of
REPROJECT under real camera motion (resolve.cu:260-340).

Every expected value here is derived independently of the oracle's code:
* spatial reuse: a neighbour's weight towards a pixel is the alpha of the neighbour's
  winner at that pixel over its alpha at its own pixel (FP64 EWA projection of
  test_oracle_project), times the occlusion ratios of the k-buffer entries in front of
  the sample; the k-buffer inserts new winners at the front and drops the oldest;
* REPROJECT: camera translations by whole pixels at a constant depth land the history
  fetch on one known texel; sub-pixel and rotated motion over linear-ramp histories are
  reproduced exactly by Catmull-Rom, so the fetched value is the ramp at the reprojected
  position (FP64 pinhole reprojection); the history rejection's extents follow from the
  closed-form moments of a checkerboard.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from conftest import oracle_scene
from test_oracle_project import ewa_fp64

MIN_ALPHA = 1.0 / 255.0
SH_C0 = 0.28209479177387814


# ---------------------------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------------------------
def _quat_about(axis, angle):
    axis = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    return np.concatenate([[math.cos(angle / 2)], math.sin(angle / 2) * axis])


def _two_gaussian_scene(extra_culled=False):
    """Two anisotropic, rotated Gaussians (SH degree 0) in front of a camera looking down +z,
    plus optionally a third one whose opacity is below 1/255 (never projects)."""
    means = [[-0.35, -0.1, 3.0], [0.3, 0.15, 3.6]]
    scales = [[math.log(0.36), math.log(0.14), math.log(0.2)], [math.log(0.2), math.log(0.42), math.log(0.16)]]
    rots = [_quat_about([0.2, 0.3, 1.0], 0.7), _quat_about([-0.4, 0.1, 1.0], -1.1)]
    logits = [2.2, 1.4]
    dcs = [[0.9, -0.4, 0.3], [-0.2, 0.6, 1.1]]
    if extra_culled:
        means.append([0.0, 0.0, 2.5])
        scales.append([math.log(0.05)] * 3)
        rots.append([1.0, 0.0, 0.0, 0.0])
        logits.append(-10.0)
        dcs.append([0.1, 0.1, 0.1])
    f32 = np.float32
    return {
        "means": np.asarray(means, f32), "scales": np.asarray(scales, f32),
        "rotations": np.asarray(rots, f32), "opacities": np.asarray(logits, f32),
        "sh": np.asarray(dcs, f32).reshape(-1, 1, 3),
    }


def _front_camera(width=40, height=30, ss=1, spatial_reuse=True):
    f = 36.0
    return {
        "w2c": np.eye(4, dtype=np.float32), "cam_position": np.zeros(3, np.float32),
        "width": width, "height": height, "focal_x": f, "focal_y": f * 1.1,
        "center_x": width / 2.0 + 0.3, "center_y": height / 2.0 - 0.2,
        "near_plane": 0.2, "far_plane": 100.0, "proper_antialiasing": False,
        "supersampling_factor": ss, "use_gps": False, "spatial_reuse": spatial_reuse,
    }


def _alpha64(proj, px, py):
    """FP64 alpha of a projected Gaussian at integer framebuffer pixel (px, py), zero below
    1/255 (the fragment-acceptance convention: pixel (x, y) is tested against mean2d - 0.5)."""
    dx = px - (proj["mean2d"][0] - 0.5)
    dy = py - (proj["mean2d"][1] - 0.5)
    a, b, c = proj["conic"]
    a = proj["opacity"] * np.exp(-0.5 * (a * dx * dx + 2.0 * b * dx * dy + c * dy * dy))
    return np.where(a < MIN_ALPHA, 0.0, a)


def _near_threshold(a):
    return np.abs(a - MIN_ALPHA) < 0.02 * MIN_ALPHA


def _sample(oracle_mod, sc, cam, fb, kb):
    W, H = oracle_mod.fb_dims(cam)
    P = W * H
    samples = np.zeros((P, 4), np.float32)
    rw = np.zeros((P, 4), np.float32)
    kp, ko = kb
    oracle_mod.lib().orc_sample(C.byref(sc.c), C.byref(oracle_mod.make_camera(cam)),
                                np.ascontiguousarray(fb.reshape(-1)), samples, rw.ctypes.data,
                                kp.ctypes.data, ko.ctypes.data)
    return samples, rw


def _word(depth, g):
    return (int(np.float32(depth).view(np.uint32)) << 32) | int(g)


def _projections(arrays, cam, n=2):
    ss = max(1, int(cam.get("supersampling_factor", 1)))
    out = []
    for g in range(n):
        out.append(ewa_fp64(arrays["means"][g].astype(np.float64), arrays["scales"][g].astype(np.float64),
                            arrays["rotations"][g].astype(np.float64), float(arrays["opacities"][g]), cam, ss))
    return out


# ---------------------------------------------------------------------------------------------
# spatial reuse: neighbour weights, k-buffer occlusion and eviction
# ---------------------------------------------------------------------------------------------
def test_reuse_weights_are_alpha_ratios(oracle_mod):
    """First sample on an empty k-buffer: the weight a winner stores towards its left, right,
    upper and lower neighbour is alpha(neighbour) / alpha(self) of that winner (FP64)."""
    arrays = _two_gaussian_scene()
    sc = oracle_scene(arrays)
    cam = _front_camera()
    W, H = 40, 30
    proj = _projections(arrays, cam)
    fb, _ = oracle_mod.accumulate(sc, cam, 3)
    fb = fb.reshape(-1)
    kb = oracle_mod.empty_k_buffer(cam)
    samples, rw = _sample(oracle_mod, sc, cam, fb, kb)
    covered = fb != np.iinfo(np.uint64).max
    assert covered.sum() > 40
    ys, xs = np.divmod(np.arange(W * H), W)
    checked = 0
    for pix in np.flatnonzero(covered):
        g = int(fb[pix] & np.uint64(0xFFFFFFFF))
        x, y = xs[pix], ys[pix]
        own = _alpha64(proj[g], x, y)
        nb = [_alpha64(proj[g], x - 1, y), _alpha64(proj[g], x + 1, y),
              _alpha64(proj[g], x, y - 1), _alpha64(proj[g], x, y + 1)]
        if _near_threshold(own) or any(_near_threshold(a) for a in nb):
            continue
        expect = np.array(nb) / max(own, MIN_ALPHA)
        assert np.allclose(rw[pix], expect, rtol=2e-4, atol=1e-6), (x, y, rw[pix], expect)
        # the k-buffer's first entry: this winner, its depth bits, occlusion ratios
        assert int(kb[0][0, pix, 0]) == g and int(np.uint32(kb[0][0, pix, 1])) == int(fb[pix] >> np.uint64(32))
        occ = (1.0 - np.minimum(nb, 0.99)) / (1.0 - min(own, 0.99))
        assert np.allclose(kb[1][0, pix], occ, rtol=2e-4, atol=1e-6)
        checked += 1
    assert checked > 30
    # background pixels: weight 1 towards every neighbour, nothing inserted
    assert np.all(rw[~covered] == 1.0)
    assert np.all(kb[0][:, ~covered] == -1) and np.all(kb[0][1:] == -1)


def test_reuse_occlusion_product_and_eviction(oracle_mod):
    """A prefilled k-buffer: entries whose depth lies in front of the sample multiply its
    weights by their occlusion ratios (NaN-depth empty entries never do); a winner not yet in
    the buffer is inserted at the front and the oldest entry is dropped; a winner already in
    it, a background pixel and a winner that no longer projects insert nothing."""
    arrays = _two_gaussian_scene(extra_culled=True)
    sc = oracle_scene(arrays)
    cam = _front_camera()
    W, H = 40, 30
    P = W * H
    proj = _projections(arrays, cam)
    ys, xs = np.divmod(np.arange(P), W)
    # framebuffer built by hand: Gaussian 0 where it is visible, Gaussian 2 (culled) in a
    # block of the top rows, background elsewhere; the sample depth is 3.0
    a0 = _alpha64(proj[0], xs, ys)
    fb = np.full(P, np.iinfo(np.uint64).max, np.uint64)
    g0 = a0 > 0.05
    fb[g0] = _word(3.0, 0)
    culled = (ys < 3) & (xs < 6)
    fb[culled] = _word(3.0, 2)
    # k-buffer: entry depths 1.0 (in front), 5.0 (behind), empty (-1, NaN depth), 2.0 (in front)
    rng = np.random.default_rng(5)
    kp = np.empty((4, P, 2), np.int32)
    depths = [1.0, 5.0, None, 2.0]
    for k, d in enumerate(depths):
        if d is None:
            kp[k] = -1
        else:
            kp[k, :, 0] = 100 + k
            kp[k, :, 1] = np.float32(d).view(np.int32)
    # some pixels of Gaussian 0 already hold it in entry 1 (depth 5.0: behind, no occlusion)
    dup = g0 & (xs % 3 == 0)
    kp[1, dup, 0] = 0
    ko = rng.uniform(0.2, 1.5, (4, P, 4)).astype(np.float32)
    kp_before, ko_before = kp.copy(), ko.copy()
    samples, rw = _sample(oracle_mod, sc, cam, fb, (kp, ko))

    in_front = ko_before[0].astype(np.float64) * ko_before[3]          # depths 1.0 and 2.0 < 3.0
    all_set = ko_before[0].astype(np.float64) * ko_before[1] * ko_before[3]  # < inf: every non-empty entry
    checked = 0
    for pix in np.flatnonzero(g0):
        x, y = xs[pix], ys[pix]
        own = _alpha64(proj[0], x, y)
        nb = np.array([_alpha64(proj[0], x - 1, y), _alpha64(proj[0], x + 1, y),
                       _alpha64(proj[0], x, y - 1), _alpha64(proj[0], x, y + 1)])
        if _near_threshold(own) or np.any(_near_threshold(nb)):
            continue
        expect = nb / max(own, MIN_ALPHA) * in_front[pix]
        assert np.allclose(rw[pix], expect, rtol=2e-4, atol=1e-6)
        checked += 1
        if dup[pix]:
            assert np.array_equal(kp[:, pix], kp_before[:, pix]) and np.array_equal(ko[:, pix], ko_before[:, pix])
        else:
            assert int(kp[0, pix, 0]) == 0 and kp[0, pix, 1] == np.float32(3.0).view(np.int32)
            occ = (1.0 - np.minimum(nb, 0.99)) / (1.0 - min(own, 0.99))
            assert np.allclose(ko[0, pix], occ, rtol=2e-4, atol=1e-6)
            assert np.array_equal(kp[1:, pix], kp_before[:3, pix]) and np.array_equal(ko[1:, pix], ko_before[:3, pix])
    assert checked > 20
    bg = fb == np.iinfo(np.uint64).max
    assert np.allclose(rw[bg], all_set[bg], rtol=1e-6)
    assert np.array_equal(kp[:, bg], kp_before[:, bg]) and np.array_equal(ko[:, bg], ko_before[:, bg])
    # the culled winner: weights 0, nothing inserted, but its colour is still shaded
    assert np.all(rw[culled] == 0.0)
    assert np.array_equal(kp[:, culled], kp_before[:, culled])
    c = np.maximum(0.5 + SH_C0 * arrays["sh"][2, 0].astype(np.float64), 0.0)
    assert np.allclose(samples[culled][:, :3], c, atol=1e-6) and np.all(samples[culled][:, 3] == 3.0)


@pytest.mark.parametrize("ss", [1, 2])
def test_reuse_resolve_closed_form(oracle_mod, ss):
    """A full render with spatial reuse on an empty k-buffer: every framebuffer pixel's
    estimate is the self-normalised sum of its own sample and its four neighbours' samples,
    each neighbour weighted by its winner's alpha at this pixel over its alpha at the
    neighbour's own pixel (background neighbours weigh 1 and contribute nothing); the image
    pixel averages its ss x ss subpixels."""
    arrays = _two_gaussian_scene()
    sc = oracle_scene(arrays)
    cam = _front_camera(ss=ss)
    W, H = 40, 30
    FW, FH = W * ss, H * ss
    proj = _projections(arrays, cam)
    kb = oracle_mod.empty_k_buffer(cam)
    img, dep, alp, wm, fb, _ = oracle_mod.render(sc, cam, 11, k_buffer=kb)
    fb = fb.reshape(-1)
    has = fb != np.iinfo(np.uint64).max
    g = np.where(has, (fb & np.uint64(0xFFFFFFFF)).astype(np.int64), -1)
    depth = np.where(has, (fb >> np.uint64(32)).astype(np.uint32).view(np.float32), 0.0).astype(np.float64)
    colour = np.zeros((FW * FH, 3))
    for k in range(2):
        colour[g == k] = np.maximum(0.5 + SH_C0 * arrays["sh"][k, 0].astype(np.float64), 0.0)
    fys, fxs = np.divmod(np.arange(FW * FH), FW)
    num_c = np.zeros((FW * FH, 3))
    num_d = np.zeros(FW * FH)
    num_a = np.zeros(FW * FH)
    den = np.zeros(FW * FH)
    skip = np.zeros(FW * FH, bool)
    for p in range(FW * FH):
        x, y = fxs[p], fys[p]
        nc, nd, na, nw = colour[p].copy(), depth[p], float(has[p]), 1.0
        for (qx, qy) in ((x - 1, y), (x + 1, y), (x, y - 1), (x, y + 1)):
            if not (0 <= qx < FW and 0 <= qy < FH):
                continue
            q = qy * FW + qx
            if has[q]:
                pr = proj[g[q]]
                towards, own = _alpha64(pr, x, y), _alpha64(pr, qx, qy)
                if _near_threshold(towards) or _near_threshold(own):
                    skip[p] = True
                w = towards / max(own, MIN_ALPHA)
            else:
                w = 1.0
            nc += w * colour[q]
            nd += w * depth[q]
            na += w * float(has[q])
            nw += w
        num_c[p], num_d[p], num_a[p], den[p] = nc / nw, nd / nw, na / nw, nw
    blk = lambda a: a.reshape(H, ss, W, ss, *a.shape[1:]).mean(axis=(1, 3))
    ok = ~blk(skip.astype(np.float64)).astype(bool)
    assert ok.sum() > 0.9 * W * H
    assert np.allclose(img[ok], blk(num_c)[ok], atol=2e-5)
    assert np.allclose(dep[..., 0][ok], blk(num_d)[ok], atol=2e-4)
    assert np.allclose(alp[..., 0][ok], blk(num_a)[ok], atol=2e-5)
    assert np.allclose(wm[..., 0][ok], blk(den)[ok], rtol=2e-5)
    assert has.mean() > 0.05 and (num_a[has] < 1.0).any()   # reuse actually mixed in background


# ---------------------------------------------------------------------------------------------
# REPROJECT under motion
# ---------------------------------------------------------------------------------------------
def _resolve(oracle_mod, cam, samples, hist, v2p, mode=2):
    W, H = int(cam["width"]), int(cam["height"])
    outs = [np.zeros((H, W, 3), np.float32)] + [np.zeros((H, W, 1), np.float32) for _ in range(3)]
    hist = [np.ascontiguousarray(h, np.float32) for h in hist]
    m = np.ascontiguousarray(np.asarray(v2p, np.float32).reshape(16))
    rc = oracle_mod.lib().orc_resolve(C.byref(oracle_mod.make_camera(cam)),
                                      np.ascontiguousarray(samples, np.float32).reshape(-1), None,
                                      *[h.ctypes.data for h in hist], m.ctypes.data, mode, *outs)
    assert rc == 0
    return outs


def _cam_plain(W=32, H=24, f=64.0, ss=1, far=100.0, fy=None):
    return {"w2c": np.eye(4, dtype=np.float32), "cam_position": np.zeros(3, np.float32), "width": W, "height": H,
            "focal_x": f, "focal_y": f if fy is None else fy, "center_x": W / 2.0, "center_y": H / 2.0,
            "near_plane": 0.2, "far_plane": far, "supersampling_factor": ss, "spatial_reuse": False}


def _ycocg(c):
    c = np.asarray(c, np.float64)
    return np.stack([0.25 * c[..., 0] + 0.5 * c[..., 1] + 0.25 * c[..., 2],
                     0.5 * c[..., 0] - 0.5 * c[..., 2],
                     -0.25 * c[..., 0] + 0.5 * c[..., 1] - 0.25 * c[..., 2]], -1)


def _reproject64(cam, m, z, x, y):
    """FP64 pinhole reprojection of pixel centre (x, y) at view depth z through m."""
    fx, fy, cx, cy = cam["focal_x"], cam["focal_y"], cam["center_x"], cam["center_y"]
    v = np.array([(x + 0.5 - cx) / fx * z, (y + 0.5 - cy) / fy * z, z, 1.0])
    p = np.asarray(m, np.float64).reshape(4, 4) @ v
    return fx * p[0] / p[2] + cx, fy * p[1] / p[2] + cy, p[2]


@pytest.mark.parametrize("kx,ky", [(3, 0), (-2, 1), (0, -3), (5, 4)])
def test_reproject_integer_translation(oracle_mod, kx, ky):
    """Constant depth z0 and a camera translation of (kx, ky) pixels at that depth: the
    history is read at texel (x + kx, y + ky) exactly; the current colour is constant, so the
    rejection extents are the floors (0.01 luma, 0.005 chroma) and the discount is 0.7, or
    0.7 / distance when the history lies outside them."""
    W, H, f, z0 = 32, 24, 64.0, 2.0
    cam = _cam_plain(W, H, f)
    c0 = np.array([0.3, 0.5, 0.2])
    samples = np.zeros((H, W, 4), np.float32)
    samples[..., :3] = c0
    samples[..., 3] = z0
    rng = np.random.default_rng(100 + kx * 7 + ky)
    h_img = rng.uniform(0.0, 1.0, (H, W, 3)).astype(np.float32)
    # a third of the history pixels lie within the extents of the current colour
    near = rng.random((H, W)) < 0.33
    h_img[near] = (c0 + rng.uniform(-0.002, 0.002, (near.sum(), 3))).astype(np.float32)
    h_dep = rng.uniform(1.0, 3.0, (H, W, 1)).astype(np.float32)
    h_alp = rng.uniform(0.0, 1.0, (H, W, 1)).astype(np.float32)
    h_wm = np.full((H, W, 1), 5.0, np.float32)
    m = np.eye(4)
    m[0, 3], m[1, 3] = kx * z0 / f, ky * z0 / f          # exact in binary
    img, dep, alp, wm = _resolve(oracle_mod, cam, samples, (h_img, h_dep, h_alp, h_wm), m)
    n_hist = 0
    for y in range(H):
        for x in range(W):
            px, py = x + 0.5 + kx, y + 0.5 + ky
            if not (2.0 <= px <= W - 2.0 and 2.0 <= py <= H - 2.0):
                assert np.allclose(img[y, x], c0, atol=1e-7) and wm[y, x, 0] == 1.0 and dep[y, x, 0] == z0
                continue
            hx, hy = x + kx, y + ky
            h = h_img[hy, hx].astype(np.float64)
            disc = np.abs(_ycocg(h) - _ycocg(c0)) / np.array([0.01, 0.005, 0.005])
            dist = disc.max()
            d = 0.7 / dist if dist > 1.0 else 0.7
            assert np.allclose(img[y, x], c0 + d * (h - c0), atol=3e-6), (x, y)
            assert abs(dep[y, x, 0] - (z0 + d * (h_dep[hy, hx, 0] - z0))) < 3e-6
            assert abs(alp[y, x, 0] - (1.0 + d * (h_alp[hy, hx, 0] - 1.0))) < 3e-6
            assert abs(wm[y, x, 0] - 1.0 / (1.0 - d)) < 2e-5 * wm[y, x, 0]
            n_hist += 1
    assert n_hist > 0.4 * W * H


@pytest.mark.parametrize("k", [0, 1, 2, 4, 7])
def test_reproject_velocity_sets_gamma(oracle_mod, k):
    """A luma checkerboard of step s has, in every interior 3x3 window, 5 pixels of the
    centre's parity and 4 of the other: mean luma L0 + s n/9, standard deviation
    s sqrt(20)/9. A uniform history colour is fetched exactly; a move of k pixels sets
    gamma = lerp(4, 1, min(k / 4, 1)) and the discount 0.7 / (|D| / (gamma sd))."""
    W, H, f, z0 = 32, 24, 64.0, 2.0
    cam = _cam_plain(W, H, f)
    s = 0.1
    base = np.array([0.2, 0.3, 0.25])
    ys, xs = np.mgrid[0:H, 0:W]
    odd = ((xs + ys) % 2).astype(np.float64)
    samples = np.zeros((H, W, 4), np.float32)
    samples[..., :3] = base + s * odd[..., None]                 # luma-only step: ycocg(1,1,1) = (1,0,0)
    samples[..., 3] = z0
    L0 = _ycocg(base)[0]
    D = 0.3
    hist_c = base + (0.5 * s + D)                                # luma L0 + s/2 + D, same chroma
    h_img = np.broadcast_to(hist_c, (H, W, 3)).astype(np.float32).copy()
    h_dep = np.full((H, W, 1), 1.5, np.float32)
    h_alp = np.full((H, W, 1), 0.25, np.float32)
    h_wm = np.ones((H, W, 1), np.float32)
    m = np.eye(4)
    m[0, 3] = k * z0 / f
    img, dep, alp, wm = _resolve(oracle_mod, cam, samples, (h_img, h_dep, h_alp, h_wm), m)
    gamma = 4.0 + (1.0 - 4.0) * min(k / 4.0, 1.0)
    sd = s * math.sqrt(20.0) / 9.0
    checked = 0
    for y in range(1, H - 1):
        for x in range(1, W - 1):
            if not (2.0 <= x + 0.5 + k <= W - 2.0 and 2.0 <= y + 0.5 <= H - 2.0):
                continue
            n_odd = 5 if odd[y, x] else 4
            mean_l = L0 + s * n_odd / 9.0
            dist = abs(L0 + 0.5 * s + D - mean_l) / max(gamma * sd, 0.01)
            d = 0.7 / dist if dist > 1.0 else 0.7
            cur = samples[y, x, :3].astype(np.float64)
            assert np.allclose(img[y, x], cur + d * (hist_c - cur), atol=5e-6), (k, x, y)
            assert abs(wm[y, x, 0] - 1.0 / (1.0 - d)) < 5e-5
            checked += 1
    assert checked > 100


def _ramp(H, W, a, bx, by):
    ys, xs = np.mgrid[0:H, 0:W]
    return (a + bx * (xs + 0.5) + by * (ys + 0.5)).astype(np.float32)


@pytest.mark.parametrize("case", ["subpixel", "rotation", "rotation_ss2", "far_plane"])
def test_reproject_linear_history_at_moved_position(oracle_mod, case):
    """Catmull-Rom reproduces linear functions, so a history depth / alpha that is a ramp
    a + bx u + by v of the pixel centre (u, v) is fetched as the ramp at the reprojected
    position; with a uniform colour equal to the history's the discount is 0.7. The motion
    matrix is not symmetric, so reading it transposed moves the fetch elsewhere. The
    reprojection depth is the mean of the valid subpixel depths, or the far plane when
    there is none."""
    ss = 2 if case == "rotation_ss2" else 1
    W, H = 40, 30
    cam = _cam_plain(W, H, 48.0, ss=ss, far=100.0, fy=52.0)
    FW, FH = W * ss, H * ss
    rng = np.random.default_rng(17)
    c0 = np.array([0.4, 0.35, 0.6])
    samples = np.zeros((FH, FW, 4), np.float32)
    samples[..., :3] = c0
    if case == "far_plane":
        samples[..., 3] = 0.0                                   # no valid sample anywhere
    else:
        samples[..., 3] = rng.uniform(1.5, 2.5, (FH, FW)).astype(np.float32)
        if ss == 2:
            samples[..., 3][rng.random((FH, FW)) < 0.3] = 0.0  # some background subpixels
    if case == "subpixel":
        m = np.eye(4)
        m[0, 3], m[1, 3] = 0.013, -0.021
    else:
        ay, ax = (0.12, -0.07) if case != "far_plane" else (0.05, 0.03)
        Ry = np.array([[math.cos(ay), 0, math.sin(ay)], [0, 1, 0], [-math.sin(ay), 0, math.cos(ay)]])
        Rx = np.array([[1, 0, 0], [0, math.cos(ax), -math.sin(ax)], [0, math.sin(ax), math.cos(ax)]])
        m = np.eye(4)
        m[:3, :3] = Ry @ Rx
        m[:3, 3] = [0.04, -0.03, 0.05] if case != "far_plane" else [0.0, 0.0, 0.0]
    m32 = m.astype(np.float32)
    if case != "subpixel":
        assert not np.allclose(m32[:3, :3], m32[:3, :3].T)
    h_img = np.broadcast_to(c0, (H, W, 3)).astype(np.float32).copy()
    h_dep = _ramp(H, W, 1.0, 0.031, -0.017)[..., None]
    h_alp = _ramp(H, W, 0.2, 0.011, 0.013)[..., None]
    h_wm = np.ones((H, W, 1), np.float32)
    img, dep, alp, wm = _resolve(oracle_mod, cam, samples, (h_img, h_dep, h_alp, h_wm), m32)
    sub = samples[..., 3].reshape(H, ss, W, ss).transpose(0, 2, 1, 3).reshape(H, W, ss * ss).astype(np.float64)
    valid = sub > 0
    nv = valid.sum(-1)
    zmean = np.where(nv > 0, np.where(valid, sub, 0.0).sum(-1) / np.maximum(nv, 1), 100.0)
    cur_d = sub.mean(-1)
    cur_a = valid.mean(-1)
    used = 0
    for y in range(H):
        for x in range(W):
            ppx, ppy, pz = _reproject64(cam, m32, zmean[y, x], x, y)
            margin = min(abs(ppx - 2.0), abs(ppx - (W - 2.0)), abs(ppy - 2.0), abs(ppy - (H - 2.0)))
            if margin < 1e-3 or pz <= 0:
                continue
            if not (2.0 <= ppx <= W - 2.0 and 2.0 <= ppy <= H - 2.0):
                assert abs(dep[y, x, 0] - cur_d[y, x]) < 1e-6 and wm[y, x, 0] == 1.0
                continue
            hd = 1.0 + 0.031 * ppx - 0.017 * ppy
            ha = 0.2 + 0.011 * ppx + 0.013 * ppy
            assert abs(dep[y, x, 0] - (cur_d[y, x] + 0.7 * (hd - cur_d[y, x]))) < 2e-5, (case, x, y)
            assert abs(alp[y, x, 0] - (cur_a[y, x] + 0.7 * (ha - cur_a[y, x]))) < 2e-5, (case, x, y)
            assert np.allclose(img[y, x], c0, atol=2e-6)
            assert abs(wm[y, x, 0] - 1.0 / 0.3) < 1e-4
            used += 1
    assert used > 0.5 * W * H
