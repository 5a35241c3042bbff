"""CPU checks of scene loading: the oracle's PLY reader and Morton order against files and
answers fixed independently of it, the product's header parser, and the host-side argument
checks of the scene-loading C ABI (no kernel launches)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import scene_oracle

from plyfiles import gaussian_rows, random_arrays, write_ply


@pytest.mark.parametrize("fmt", ["binary_little_endian", "binary_big_endian", "ascii"])
@pytest.mark.parametrize("bases", [1, 4, 9, 16])
def test_oracle_reads_exporter_layout(tmp_path, fmt, bases):
    arrays = random_arrays(37, bases, seed=bases)
    path = tmp_path / "s.ply"
    write_ply(path, gaussian_rows(arrays), fmt)
    got = scene_oracle.read_ply(path)
    for k, v in arrays.items():
        assert got[k].dtype == np.float32 and got[k].shape == v.shape, k
        np.testing.assert_array_equal(got[k], v, err_msg=k)


def test_oracle_converts_other_scalar_types(tmp_path):
    arrays = random_arrays(20, 4, seed=3)
    arrays["scales"] = arrays["scales"].astype(np.float64).astype(np.float32)

    def extra(i, r):
        # doubles holding values no float32 represents, uchar colours in the middle of the record
        r = [(n, "double", float(v) + 1e-12) if n.startswith("scale_") else (n, t, v) for n, t, v in r]
        return r[:6] + [("red", "uchar", i % 256), ("green", "uchar", 7), ("blue", "uchar", 255)] + r[6:]

    path = tmp_path / "t.ply"
    write_ply(path, gaussian_rows(arrays, extra), "binary_little_endian", trailing_element=True)
    got = scene_oracle.read_ply(path)
    np.testing.assert_array_equal(got["scales"], (arrays["scales"].astype(np.float64) + 1e-12).astype(np.float32))
    np.testing.assert_array_equal(got["sh"], arrays["sh"])


# ---------------------------------------------------------------------------------------
# Morton order
# ---------------------------------------------------------------------------------------
def _bits_of(axis_bit_positions):
    return sum(1 << p for p in axis_bit_positions)


def test_morton_known_answers():
    m = np.array([[0, 0, 0], [1, 1, 1], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0.5, 0, 0]], np.float32)
    codes = scene_oracle.morton_codes(m)
    assert codes[0] == 0
    assert codes[1] == (1 << 63) - 1
    assert codes[2] == _bits_of(range(0, 63, 3))
    assert codes[3] == _bits_of(range(1, 63, 3))
    assert codes[4] == _bits_of(range(2, 63, 3))
    assert codes[5] == 1 << 60      # x cell 2^20: only the top x bit


def _octant_order(points, idx, lo, size, depth):
    """Recursive octant traversal (x fastest, then y, then z), no bit operations."""
    if depth == 0 or len(idx) <= 1:
        return list(idx)
    half = size / 2
    out = []
    for oz in (0, 1):
        for oy in (0, 1):
            for ox in (0, 1):
                c = lo + half * np.array([ox, oy, oz])
                sel = [i for i in idx if all(c[a] <= points[i][a] < c[a] + half for a in range(3))]
                out += _octant_order(points, sel, c, half, depth - 1)
    return out


def test_morton_order_is_octant_traversal():
    g = np.arange(8, dtype=np.float64)
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pts = (pts + 0.5) / 8.0
    rng = np.random.default_rng(5)
    pts = pts[rng.permutation(len(pts))]
    pts = np.concatenate([np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]]), pts]).astype(np.float32)
    order = scene_oracle.morton_order(pts)
    inner = [int(i) for i in order if i >= 2]
    expect = _octant_order(pts.astype(np.float64), list(range(2, len(pts))), np.zeros(3), 1.0, 3)
    assert inner == expect


def test_morton_cells_match_float64_quantisation():
    rng = np.random.default_rng(1)
    m = rng.normal(size=(5000, 3)).astype(np.float32)
    codes = scene_oracle.morton_codes(m)
    q = np.zeros((len(m), 3), np.int64)
    for b in range(21):
        for a in range(3):
            q[:, a] |= ((codes >> np.uint64(3 * b + a)) & np.uint64(1)).astype(np.int64) << b
    lo = m.min(0).astype(np.float64)
    extent = float((m.max(0).astype(np.float64) - lo).max())
    ref = np.minimum(np.floor((m - lo) / extent * 2 ** 21), 2 ** 21 - 1)
    diff = np.abs(q - ref)
    # float32 rounding of (x - lo) / extent (relative error ~6e-8, i.e. ~0.1 of a 2^-21 cell)
    # moves some points across one cell boundary, never further
    assert diff.max() <= 1
    assert (diff == 0).mean() > 0.8


def test_morton_order_stable_and_non_finite():
    m = np.array([[0.5, 0.5, 0.5], [0, 0, 0], [0.5, 0.5, 0.5], [1, 1, 1], [np.nan, 1, 1], [np.inf, 0, 0],
                  [0.5, 0.5, 0.5]], np.float32)
    codes = scene_oracle.morton_codes(m)
    order = scene_oracle.morton_order(m)
    assert sorted(order.tolist()) == list(range(len(m)))
    assert np.all(np.diff(codes[order].astype(np.float64)) >= 0)
    dup = [i for i in order.tolist() if i in (0, 2, 6)]
    assert dup == [0, 2, 6]                               # equal codes keep index order
    assert codes[4] == codes[3] & np.uint64(_bits_of(range(1, 63, 3)) | _bits_of(range(2, 63, 3)))  # NaN x -> 0
    assert codes[5] == np.uint64(_bits_of(range(0, 63, 3)))   # +inf x -> last cell
    assert np.all(scene_oracle.morton_codes(np.zeros((4, 3), np.float32)) == 0)


# ---------------------------------------------------------------------------------------
# product host logic: header parsing and layout
# ---------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def ply_mod():
    from StochasticGSCudaBackend import ply
    return ply


def test_header_offsets(tmp_path, ply_mod):
    arrays = random_arrays(3, 16)

    def extra(i, r):
        return r[:6] + [("red", "uchar", 1)] + r[6:]

    path = tmp_path / "h.ply"
    write_ply(path, gaussian_rows(arrays, extra), "binary_big_endian", trailing_element=True)
    with open(path, "rb") as f:
        hdr = ply_mod.read_header(f)
    assert hdr.count == 3 and hdr.element == "vertex" and hdr.format == "binary_big_endian"
    assert hdr.stride == 4 * 62 + 1 and hdr.total_sh_bases == 16
    lay = ply_mod.layout_for(hdr)
    assert lay.big_endian == 1 and lay.stride == hdr.stride and lay.total_sh_bases == 16
    assert [lay.offset[i] for i in range(3)] == [0, 4, 8]
    assert lay.offset[11] == 6 * 4 + 1               # f_dc_0 after x..nz and red
    assert lay.offset[14] == 9 * 4 + 1               # f_rest_0
    assert lay.offset[10] == (9 + 45) * 4 + 1        # opacity
    assert all(lay.type[i] == 7 for i in range(59))


def test_header_errors(tmp_path, ply_mod):
    arrays = random_arrays(2, 4)
    rows = gaussian_rows(arrays)
    missing = tmp_path / "m.ply"
    write_ply(missing, [[c for c in r if c[0] != "rot_3"] for r in rows])
    with open(missing, "rb") as f:
        with pytest.raises(KeyError, match="rot_3"):
            ply_mod.layout_for(ply_mod.read_header(f))
    odd = tmp_path / "o.ply"   # 6 rest coefficients -> 3 bases, not a full SH degree
    write_ply(odd, [[c for c in r if not (c[0].startswith("f_rest_") and int(c[0][7:]) >= 6)] for r in rows])
    with open(odd, "rb") as f:
        with pytest.raises(ValueError, match="SH bases"):
            ply_mod.layout_for(ply_mod.read_header(f))
    lst = tmp_path / "l.ply"
    lst.write_bytes(b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty list uchar int idx\n"
                    b"end_header\n")
    with open(lst, "rb") as f:
        with pytest.raises(ValueError, match="list property"):
            ply_mod.read_header(f)
    bad = tmp_path / "b.ply"
    bad.write_bytes(b"ply\nformat binary_middle_endian 1.0\nend_header\n")
    with open(bad, "rb") as f:
        with pytest.raises(ValueError, match="format"):
            ply_mod.read_header(f)


# ---------------------------------------------------------------------------------------
# C ABI argument checks (return before any launch)
# ---------------------------------------------------------------------------------------
def test_scene_abi_rejects_bad_arguments():
    from StochasticGSCudaBackend import _native as nat
    lib = nat.library()
    out = nat.sgs_scene_arrays(0x1000, 0x2000, 0x3000, 0x4000, 0x5000, 10, 4)
    lay = nat.sgs_ply_layout()
    lay.stride, lay.total_sh_bases = 4 * 23, 4
    for i in range(23):
        lay.offset[i], lay.type[i] = 4 * i, 7
    assert lib.sgs_ply_decode(0x9000, 11, C.byref(lay), 0, C.byref(out), None) == -1      # 11 rows > 10
    assert lib.sgs_ply_decode(0x9000, 5, C.byref(lay), 6, C.byref(out), None) == -1       # 6 + 5 > 10
    assert b"exceed" in lib.sgs_last_error()
    assert lib.sgs_ply_decode(None, 5, C.byref(lay), 0, C.byref(out), None) == -1          # records NULL
    lay.type[3] = 9
    assert lib.sgs_ply_decode(0x9000, 5, C.byref(lay), 0, C.byref(out), None) == -1
    lay.type[3], lay.offset[22] = 7, 4 * 23 - 2
    assert lib.sgs_ply_decode(0x9000, 5, C.byref(lay), 0, C.byref(out), None) == -1       # outside the record
    lay.offset[22], lay.total_sh_bases = 4 * 22, 9
    assert lib.sgs_ply_decode(0x9000, 5, C.byref(lay), 0, C.byref(out), None) == -1       # bases disagree
    assert lib.sgs_ply_decode(0x9000, 0, C.byref(lay), 0, None, None) == -1
    assert lib.sgs_morton_order(0x1000, 1 << 31, 0x2000, None, 0x3000, 1 << 40, None) == -1
    assert lib.sgs_morton_order_workspace_size(1 << 31) == 0
    src = nat.sgs_gaussians(0x1000, 0x2000, 0x3000, 0x4000, 0x5000, 10, 9)
    assert lib.sgs_gather_gaussians(C.byref(src), 0x6000, C.byref(out), None) == -1       # bases disagree
    src.total_sh_bases, src.n_gaussians = 4, 5
    assert lib.sgs_gather_gaussians(C.byref(src), 0x6000, C.byref(out), None) == -1       # dst larger than src
