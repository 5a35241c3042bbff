"""The counting pass's conservative tile range contains the exact one.

The batched tile schedule bins each (view, Gaussian) pair through gaussian_tiles_bound
(csrc/sgs_batch.hip): hardware reciprocal and square root (within 1 ulp), the covariance
diagonal with fused multiply-adds, widened by error bounds. The rasteriser then projects exactly
and clips every instance to its exact AABB (gs_utils.h:154-193 via pixel_aabb), so the results
stay bit-identical only if every pixel of the exact AABB lies inside the bound's box. These
tests restate both computations in float32 numpy:

* the exact side in the reference's operation order, checked bit for bit against the oracle's
  projection (oracle/sgs_oracle.c orc_project) on a sampled scene;
* the bound with the reciprocal and square root pushed one ulp in the worst direction at
  random, over populations chosen to stress it (far off-axis, needles seen end-on, cameras far
  from the origin, scaled world-to-view rows);

and require containment for every Gaussian the exact side keeps. Dropping the covariance error
term makes the containment check fail, so it is not vacuous.
"""
from __future__ import annotations

import numpy as np
import pytest

f32 = np.float32
MODES = ("plain", "wide", "needle", "far", "scaled")


def _fma(a, b, c):
    # float32 fused multiply-add through float64 (exact product and sum of float32 inputs,
    # one rounding for the sums that matter here)
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(f32)


def _rot(q):
    q = q / np.linalg.norm(q, axis=1, keepdims=True)
    r, i, j, k = q.T
    return np.stack([1 - 2 * (j * j + k * k), 2 * (i * j - r * k), 2 * (i * k + r * j),
                     2 * (i * j + r * k), 1 - 2 * (i * i + k * k), 2 * (j * k - r * i),
                     2 * (i * k - r * j), 2 * (j * k + r * i), 1 - 2 * (i * i + j * j)], 1).reshape(-1, 3, 3)


def fb_camera(W, H, fx, fy, cx, cy, aa, ss):
    """The framebuffer camera of make_fb_camera (supersampled pixels, clip bounds)."""
    Wf, Hf = f32(W * ss), f32(H * ss)
    fxs, fys, cxs, cys = f32(fx * ss), f32(fy * ss), f32(cx * ss), f32(cy * ss)
    return dict(W=Wf, H=Hf, fx=fxs, fy=fys, cx=cxs, cy=cys, aa=aa, ss=f32(ss),
                cl=f32((f32(-0.15) * Wf - cxs) / fxs), cr=f32((f32(1.15) * Wf - cxs) / fxs),
                ct=f32((f32(-0.15) * Hf - cys) / fys), cb=f32((f32(1.15) * Hf - cys) / fys),
                near=f32(0.2), far=f32(10000.0))


def _dot3(r, mx, my, mz):
    return (((r[0] * mx).astype(f32) + (r[1] * my).astype(f32)).astype(f32) + (r[2] * mz).astype(f32)).astype(f32)


def exact_aabb(m, C, op, pt_view, w2c, cam):
    """project_cov3d + pixel_aabb in the reference's operation order: (kept, projected,
    float bounds, pixel AABB); projected is the projection's own verdict (before the empty-AABB
    test of the scatter)."""
    r1, r2, r3 = w2c[0], w2c[1], w2c[2]
    mx, my, mz = m[:, 0], m[:, 1], m[:, 2]
    depth = (_dot3(r3, mx, my, mz) + r3[3]).astype(f32)
    X = (_dot3(r1, mx, my, mz) + r1[3]).astype(f32)
    Y = (_dot3(r2, mx, my, mz) + r2[3]).astype(f32)
    ok = (depth >= cam["near"]) & (depth <= cam["far"])
    c11, c12, c13, c22, c23, c33 = C[:, 0, 0], C[:, 0, 1], C[:, 0, 2], C[:, 1, 1], C[:, 1, 2], C[:, 2, 2]
    with np.errstate(all="ignore"):
        x = (X / depth).astype(f32)
        y = (Y / depth).astype(f32)
        xc = np.maximum(cam["cl"], np.minimum(x, cam["cr"]))
        yc = np.maximum(cam["ct"], np.minimum(y, cam["cb"]))
        j11 = (cam["fx"] / depth).astype(f32)
        j22 = (cam["fy"] / depth).astype(f32)
        j13 = (-j11 * xc).astype(f32)
        j23 = (-j22 * yc).astype(f32)
        a = [((j11 * r1[i]).astype(f32) + (j13 * r3[i]).astype(f32)).astype(f32) for i in range(3)]
        b = [((j22 * r2[i]).astype(f32) + (j23 * r3[i]).astype(f32)).astype(f32) for i in range(3)]

        def d3(u, w1, w2, w3):
            return (((u[0] * w1).astype(f32) + (u[1] * w2).astype(f32)).astype(f32) + (u[2] * w3).astype(f32)).astype(f32)
        p = [d3(a, c11, c12, c13), d3(a, c12, c22, c23), d3(a, c13, c23, c33)]
        q = [d3(b, c11, c12, c13), d3(b, c12, c22, c23), d3(b, c13, c23, c33)]
        cxx = d3(p, a[0], a[1], a[2])
        cxy = d3(p, b[0], b[1], b[2])
        cyy = d3(q, b[0], b[1], b[2])
        det_raw = ((cxx * cyy).astype(f32) - (cxy * cxy).astype(f32)).astype(f32)
        kernel = f32(f32(f32(0.1 if cam["aa"] else 0.3) * cam["ss"]) * cam["ss"])
        cxx = (cxx + kernel).astype(f32)
        cyy = (cyy + kernel).astype(f32)
        det = ((cxx * cyy).astype(f32) - (cxy * cxy).astype(f32)).astype(f32)
        ok &= ~(det < f32(1e-6))
        if cam["aa"]:
            opa = (op * np.sqrt(np.maximum((det_raw / det).astype(f32), f32(0))).astype(f32)).astype(f32)
            ok &= ~(opa < f32(1.0 / 255.0))
            pt = pt_view(opa)
        else:
            pt = pt_view(op)
        mean_x = ((x * cam["fx"]).astype(f32) + cam["cx"]).astype(f32)
        mean_y = ((y * cam["fy"]).astype(f32) + cam["cy"]).astype(f32)
        cutoff = (f32(2) * pt).astype(f32)
        ex = np.sqrt((cxx * cutoff).astype(f32)).astype(f32)
        ey = np.sqrt((cyy * cutoff).astype(f32)).astype(f32)
        bx0, bx1 = (mean_x - ex).astype(f32), (mean_x + ex).astype(f32)
        by0, by1 = (mean_y - ey).astype(f32), (mean_y + ey).astype(f32)
        ok &= (bx0 < cam["W"]) & (bx1 >= 0) & (by0 < cam["H"]) & (by1 >= 0)
        x0, x1 = np.floor(np.maximum(bx0, 0)), np.ceil(np.minimum(bx1, cam["W"]))
        y0, y1 = np.floor(np.maximum(by0, 0)), np.ceil(np.minimum(by1, cam["H"]))
    projected = ok.copy()
    ok &= (x1 > x0) & (y1 > y0)
    return ok, projected, (bx0, bx1, by0, by1), (x0, x1, y0, y1)


def _nudge(v, rng):
    d = rng.integers(-1, 2, v.shape)
    return np.where(d > 0, np.nextafter(v, f32(np.inf)), np.where(d < 0, np.nextafter(v, f32(0)), v)).astype(f32)


def bound_aabb(m, C, pt_cached, w2c, cam, rng, drop=None):
    """gaussian_tiles_bound: (kept, exact_path, x0, x1, y0, y1), reciprocal and square root
    within one ulp of the correctly rounded values (pushed at random)."""
    r1, r2, r3 = w2c[0], w2c[1], w2c[2]
    mx, my, mz = m[:, 0], m[:, 1], m[:, 2]
    n = len(mx)
    depth = (_dot3(r3, mx, my, mz) + r3[3]).astype(f32)
    keep = (depth >= cam["near"]) & (depth <= cam["far"])
    X = (_dot3(r1, mx, my, mz) + r1[3]).astype(f32)
    Y = (_dot3(r2, mx, my, mz) + r2[3]).astype(f32)
    c11, c12, c13, c22, c23, c33 = C[:, 0, 0], C[:, 0, 1], C[:, 0, 2], C[:, 1, 1], C[:, 1, 2], C[:, 2, 2]
    full = lambda v: np.full(n, v, f32)
    with np.errstate(all="ignore"):
        rz = _nudge((f32(1) / depth).astype(f32), rng)
        x, y = (X * rz).astype(f32), (Y * rz).astype(f32)
        dx, dy = (np.abs(x) * f32(2.0 ** -21)).astype(f32), (np.abs(y) * f32(2.0 ** -21)).astype(f32)
        xc = np.maximum(cam["cl"], np.minimum(x, cam["cr"]))
        yc = np.maximum(cam["ct"], np.minimum(y, cam["cb"]))
        j11, j22 = (cam["fx"] * rz).astype(f32), (cam["fy"] * rz).astype(f32)
        j13, j23 = (-j11 * xc).astype(f32), (-j22 * yc).astype(f32)
        a = [_fma(j13, full(r3[i]), (j11 * r1[i]).astype(f32)) for i in range(3)]
        b = [_fma(j23, full(r3[i]), (j22 * r2[i]).astype(f32)) for i in range(3)]

        def fd3(u, w1, w2, w3):
            return _fma(u[2], w3, _fma(u[1], w2, (u[0] * w1).astype(f32)))
        p = [fd3(a, c11, c12, c13), fd3(a, c12, c22, c23), fd3(a, c13, c23, c33)]
        q = [fd3(b, c11, c12, c13), fd3(b, c12, c22, c23), fd3(b, c13, c23, c33)]
        cxx = _fma(p[2], a[2], _fma(p[1], a[1], (p[0] * a[0]).astype(f32)))
        cyy = _fma(q[2], b[2], _fma(q[1], b[1], (q[0] * b[0]).astype(f32)))
        up = f32(1 + 2.0 ** -20)
        n1 = f32(f32(f32(abs(r1[0]) + abs(r1[1])) + abs(r1[2])) * up)
        n2 = f32(f32(f32(abs(r2[0]) + abs(r2[1])) + abs(r2[2])) * up)
        n3 = f32(f32(f32(abs(r3[0]) + abs(r3[1])) + abs(r3[2])) * up)
        sx = _fma((np.abs(xc) + dx).astype(f32), full(n3), full(n1))
        sy = _fma((np.abs(yc) + dy).astype(f32), full(n3), full(n2))
        tr = ((((c11 + c22).astype(f32) + c33).astype(f32)) * f32(1 + 2.0 ** -16)).astype(f32)
        exx = (((tr * (j11 * j11).astype(f32)).astype(f32) * sx).astype(f32) *
               _fma(sx, full(f32(2.0 ** -18)), (f32(2) * dx * n3).astype(f32))).astype(f32)
        eyy = (((tr * (j22 * j22).astype(f32)).astype(f32) * sy).astype(f32) *
               _fma(sy, full(f32(2.0 ** -18)), (f32(2) * dy * n3).astype(f32))).astype(f32)
        if drop == "cov":
            exx, eyy = exx * 0, eyy * 0
        kernel = f32(f32(f32(0.1 if cam["aa"] else 0.3) * cam["ss"]) * cam["ss"])
        cut = _fma(np.maximum((f32(2) * pt_cached).astype(f32), f32(0)), full(f32(1 + 2.0 ** -16)), full(f32(2.0 ** -20)))
        ex = (_nudge(np.sqrt((((cxx + exx).astype(f32) + kernel).astype(f32) * cut).astype(f32)).astype(f32), rng)
              * f32(1 + 2.0 ** -16)).astype(f32)
        ey = (_nudge(np.sqrt((((cyy + eyy).astype(f32) + kernel).astype(f32) * cut).astype(f32)).astype(f32), rng)
              * f32(1 + 2.0 ** -16)).astype(f32)
        mean_x = ((x * cam["fx"]).astype(f32) + cam["cx"]).astype(f32)
        mean_y = ((y * cam["fy"]).astype(f32) + cam["cy"]).astype(f32)
        wx = (ex + _fma((np.abs((x * cam["fx"]).astype(f32)) + np.abs(mean_x) + ex).astype(f32),
                        full(f32(2.0 ** -20)), full(f32(2.0 ** -20)))).astype(f32)
        wy = (ey + _fma((np.abs((y * cam["fy"]).astype(f32)) + np.abs(mean_y) + ey).astype(f32),
                        full(f32(2.0 ** -20)), full(f32(2.0 ** -20)))).astype(f32)
        if drop == "edge":
            wx, wy = ex, ey
        exact_path = ~((wx <= 3e38) & (wy <= 3e38) & (np.abs(mean_x) <= 3e38) & (np.abs(mean_y) <= 3e38))
        bx0, bx1 = (mean_x - wx).astype(f32), (mean_x + wx).astype(f32)
        by0, by1 = (mean_y - wy).astype(f32), (mean_y + wy).astype(f32)
        keep &= ~((bx0 >= cam["W"]) | (bx1 < 0) | (by0 >= cam["H"]) | (by1 < 0))
        x0, x1 = np.floor(np.maximum(bx0, 0)), np.ceil(np.minimum(bx1, cam["W"]))
        y0, y1 = np.floor(np.maximum(by0, 0)), np.ceil(np.minimum(by1, cam["H"]))
    keep &= (x1 > x0) & (y1 > y0)
    return keep, exact_path, (x0, x1, y0, y1)


def population(mode, n, rng):
    """Gaussians in front of a random camera: (means, covariances, opacities, w2c, camera)."""
    W, H = int(rng.choice([256, 800, 1920, 2048])), int(rng.choice([256, 600, 1080, 1152]))
    fx = float(rng.uniform(100, 3000))
    fy = fx * float(rng.uniform(0.8, 1.2))
    cam = fb_camera(W, H, fx, fy, W / 2 + rng.uniform(-50, 50), H / 2 + rng.uniform(-50, 50), bool(rng.integers(2)),
                    int(rng.choice([1, 2, 3])))
    R = _rot(rng.normal(size=(1, 4)))[0]
    if mode == "scaled":
        R = R * rng.uniform(0.5, 2.0)
    t = rng.normal(size=3) * (1000.0 if mode == "far" else 3.0)
    w2c = np.zeros((3, 4), f32)
    w2c[:, :3] = R
    w2c[:, 3] = t
    z = rng.uniform(0.25, 5000 if mode == "far" else 50, n)
    ang = rng.uniform(-1.5, 1.5, (n, 2)) if mode == "wide" else rng.uniform(-0.8, 0.8, (n, 2))
    v = np.stack([z * np.tan(ang[:, 0]), z * np.tan(ang[:, 1]), z], 1)
    m = ((v - t) @ np.linalg.inv(R).T).astype(f32)
    s = rng.uniform(-9, 3, (n, 3)) if mode == "needle" else rng.uniform(-6, 0, (n, 3))
    if mode == "needle":
        s[:, 0] = rng.uniform(-1, 3, n)
    Rg = _rot(rng.normal(size=(n, 4)))
    C = np.einsum("nij,nj,nkj->nik", Rg, np.exp(2 * s), Rg).astype(f32)
    op = (1.0 / (1.0 + np.exp(-rng.normal(1, 2, n)))).astype(f32)
    keep = op >= f32(1.0 / 255.0)
    return m[keep], C[keep], op[keep], w2c, cam


def _logf(v):
    return np.log(np.asarray(v, np.float64) * 255.0).astype(f32)


def _contained(seed, mode, n=20000, drop=None):
    rng = np.random.default_rng(seed)
    m, C, op, w2c, cam = population(mode, n, rng)
    pt = _logf(op)
    ok, _, _, (X0, X1, Y0, Y1) = exact_aabb(m, C, op, _logf, w2c, cam)
    keep, exact_path, (A0, A1, B0, B1) = bound_aabb(m, C, pt, w2c, cam, rng, drop)
    sel = ok & ~exact_path
    inside = keep & (A0 <= X0) & (A1 >= X1) & (B0 <= Y0) & (B1 >= Y1)
    return int(sel.sum()), int((sel & ~inside).sum())


@pytest.mark.parametrize("mode", MODES)
def test_bound_contains_exact_aabb(mode):
    total = 0
    for seed in range(4):
        n_sel, n_bad = _contained(1000 * MODES.index(mode) + seed, mode)
        assert n_bad == 0, f"{mode} seed {seed}: {n_bad} of {n_sel} exact AABBs escape the bound"
        total += n_sel
    assert total > 2000


def test_containment_check_sees_a_dropped_term():
    """Without the covariance error term some exact AABB escapes (needles seen end-on, whose
    covariance diagonal cancels most): the check is not vacuous."""
    bad = sum(_contained(s, "needle", n=200000, drop="cov")[1] for s in range(8))
    assert bad > 0


def test_exact_restatement_matches_oracle_projection(oracle_mod):
    """The numpy exact side gives the oracle's AABB bounds bit for bit (covariance and power
    threshold computed with the oracle's own exp and log)."""
    from conftest import oracle_scene, small_camera, small_scene
    arrays = small_scene(400, seed=7)
    sc = oracle_scene(arrays)
    ex = np.vectorize(lambda v: oracle_mod.expf(float(v)), otypes=[f32])
    lg = np.vectorize(lambda v: oracle_mod.logf(float(v)), otypes=[f32])
    s, q = arrays["scales"].astype(f32), arrays["rotations"].astype(f32)
    op = (f32(1) / (f32(1) + ex(-arrays["opacities"].astype(f32).reshape(-1)))).astype(f32)
    v = ex((f32(2) * s).astype(f32))
    qr, qi, qj, qk = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    ii, jj, kk = (qi * qi).astype(f32), (qj * qj).astype(f32), (qk * qk).astype(f32)
    ij, ik, jk = (qi * qj).astype(f32), (qi * qk).astype(f32), (qj * qk).astype(f32)
    ri, rj, rk = (qr * qi).astype(f32), (qr * qj).astype(f32), (qr * qk).astype(f32)
    nsq = ((((qr * qr).astype(f32) + ii).astype(f32) + jj).astype(f32) + kk).astype(f32)
    nr = (f32(1) / nsq).astype(f32)
    two = f32(2)
    R = np.empty((len(op), 3, 3), f32)
    R[:, 0, 0] = f32(1) - ((two * (jj + kk).astype(f32)).astype(f32) * nr).astype(f32)
    R[:, 0, 1] = ((two * (ij - rk).astype(f32)).astype(f32) * nr).astype(f32)
    R[:, 0, 2] = ((two * (ik + rj).astype(f32)).astype(f32) * nr).astype(f32)
    R[:, 1, 0] = ((two * (ij + rk).astype(f32)).astype(f32) * nr).astype(f32)
    R[:, 1, 1] = f32(1) - ((two * (ii + kk).astype(f32)).astype(f32) * nr).astype(f32)
    R[:, 1, 2] = ((two * (jk - ri).astype(f32)).astype(f32) * nr).astype(f32)
    R[:, 2, 0] = ((two * (ik - rj).astype(f32)).astype(f32) * nr).astype(f32)
    R[:, 2, 1] = ((two * (jk + ri).astype(f32)).astype(f32) * nr).astype(f32)
    R[:, 2, 2] = f32(1) - ((two * (ii + jj).astype(f32)).astype(f32) * nr).astype(f32)
    S = (R * v[:, None, :]).astype(f32)
    C = np.empty_like(R)
    for i in range(3):
        for k in range(3):
            C[:, i, k] = ((((S[:, i, 0] * R[:, k, 0]).astype(f32) + (S[:, i, 1] * R[:, k, 1]).astype(f32)).astype(f32)
                           + (S[:, i, 2] * R[:, k, 2]).astype(f32)).astype(f32))
    for aa in (False, True):
        c = small_camera(96, 72, distance=2.5, angle=0.4, proper_antialiasing=aa)
        w2c = np.asarray(c["w2c"], f32).reshape(4, 4)[:3]
        cam = fb_camera(c["width"], c["height"], c["focal_x"], c["focal_y"], c["center_x"], c["center_y"], aa, 1)
        _, projected, (bx0, bx1, by0, by1), _ = exact_aabb(arrays["means"].astype(f32), C, op,
                                                            lambda o: lg((o * f32(255)).astype(f32)), w2c, cam)
        checked = 0
        for g in range(len(op)):
            p = oracle_mod.project(sc, g, c)
            assert bool(p.active) == bool(projected[g]), (aa, g)
            if p.active:
                assert (p.bounds[0], p.bounds[1], p.bounds[2], p.bounds[3]) == (bx0[g], bx1[g], by0[g], by1[g]), (aa, g)
                checked += 1
        assert checked > 100
