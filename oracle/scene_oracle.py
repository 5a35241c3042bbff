"""CPU restatement of the reference's scene loading: PLY -> arrays -> Morton order.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device loader
(sgs_ply_decode, sgs_morton_order, sgs_gather_gaussians); the product never calls it.

* ``read_ply`` follows Model.py:24-75: the first element's properties read as columns
  (plyfile's job, restated with numpy structured dtypes), each converted to float32, and the
  SH coefficients assembled as (N, B, 3) from f_dc_* and the channel-major f_rest_*
  (Model.py:70-74: reshape(-1, 3, B-1) then swap the last two axes).
* ``morton_codes`` / ``morton_order`` restate the ordering of Model.py:32-33. The NeRFICG
  ``morton_encode`` it calls is not in the reference tree (parity with it is unpinned); the
  code here is the library's documented definition: bounding box of the finite coordinates,
  its largest side as the cube edge, 2^21 cells per axis by float32 arithmetic, x in bit 3k,
  y in 3k+1, z in 3k+2; the order is a stable sort of the codes.
"""
from __future__ import annotations

import numpy as np

_NP = {"char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2",
       "ushort": "u2", "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
       "float": "f4", "float32": "f4", "double": "f8", "float64": "f8"}


def _header(data: bytes):
    end = data.index(b"end_header")
    end = data.index(b"\n", end) + 1
    lines = data[:end].decode("ascii").splitlines()
    fmt, elements = None, []
    for line in lines[1:]:
        w = line.split()
        if not w:
            continue
        if w[0] == "format":
            fmt = w[1]
        elif w[0] == "element":
            elements.append((w[1], int(w[2]), []))
        elif w[0] == "property":
            elements[-1][2].append((w[-1], w[1]))
    return fmt, elements, end


def read_columns(path) -> dict[str, np.ndarray]:
    """Columns of the first element of a PLY file, in their declared types."""
    data = open(path, "rb").read()
    fmt, elements, end = _header(data)
    _, count, props = elements[0]
    if fmt == "ascii":
        rows = data[end:].decode("ascii").splitlines()[:count]
        table = [r.split() for r in rows]
        return {name: np.array([row[j] for row in table], dtype=np.float64).astype(_NP[t]) if _NP[t][0] == "f"
                else np.array([int(row[j]) for row in table], dtype=_NP[t])
                for j, (name, t) in enumerate(props)}
    order = "<" if fmt == "binary_little_endian" else ">"
    dtype = np.dtype([(name, order + _NP[t]) for name, t in props])
    rec = np.frombuffer(data, dtype=dtype, count=count, offset=end)
    return {name: rec[name] for name, _ in props}


def read_ply(path) -> dict[str, np.ndarray]:
    """The five arrays of Model.py:26-75 in file order (float32)."""
    col = read_columns(path)
    f32 = lambda *names: np.stack([np.asarray(col[n]).astype(np.float32) for n in names], axis=1)
    n_rest = len([n for n in col if n.startswith("f_rest_")])
    bases = n_rest // 3 + 1
    means = f32("x", "y", "z")
    sh = np.empty((means.shape[0], bases, 3), np.float32)
    sh[:, 0, :] = f32("f_dc_0", "f_dc_1", "f_dc_2")
    if bases > 1:
        rest = f32(*[f"f_rest_{i}" for i in range(n_rest)])
        sh[:, 1:, :] = rest.reshape(-1, 3, bases - 1).transpose(0, 2, 1)
    return {"means": means,
            "scales": f32("scale_0", "scale_1", "scale_2"),
            "rotations": f32("rot_0", "rot_1", "rot_2", "rot_3"),
            "opacities": np.asarray(col["opacity"]).astype(np.float32),
            "sh": sh}


def morton_codes(means: np.ndarray) -> np.ndarray:
    """63-bit Morton codes (uint64) of float32 means (N,3)."""
    m = np.asarray(means, np.float32)
    n = m.shape[0]
    lo = np.zeros(3, np.float32)
    has = np.zeros(3, bool)
    extent = np.float32(0.0)
    for a in range(3):
        fin = m[:, a][np.isfinite(m[:, a])]
        if fin.size:
            has[a] = True
            lo[a] = fin.min()
            span = np.float32(fin.max() - lo[a])
            if span > extent:
                extent = span
    codes = np.zeros(n, np.uint64)
    if not extent > 0:
        return codes
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        for a in range(3):
            t = (m[:, a] - lo[a]) / extent
            ok = has[a] & (t > 0)
            q = np.where(ok, np.minimum(t * np.float32(2097152.0), np.float32(2097151.0)), 0).astype(np.uint64)
            for b in range(21):
                codes |= ((q >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + a)
    return codes


def morton_order(means: np.ndarray) -> np.ndarray:
    return np.argsort(morton_codes(means), kind="stable")


def load_scene(path, morton: bool = True) -> dict[str, np.ndarray]:
    arrays = read_ply(path)
    if morton:
        order = morton_order(arrays["means"])
        arrays = {k: np.ascontiguousarray(v[order]) for k, v in arrays.items()}
    return arrays
