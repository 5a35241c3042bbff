"""PLY files for the scene-loading tests, written independently of every reader: the layout
of the 3DGS exporter (x y z nx ny nz f_dc_0..2 f_rest_* opacity scale_0..2 rot_0..3, the
f_rest_* channel-major: all rest coefficients of red, then green, then blue), plus optional
extra properties, in any of the three PLY formats."""
from __future__ import annotations

import struct

import numpy as np

_PACK = {"float": "f", "double": "d", "uchar": "B", "char": "b", "short": "h", "ushort": "H", "int": "i",
         "uint": "I"}


def gaussian_rows(arrays, extra=None):
    """Per-vertex list of (name, type, value) in the exporter's order."""
    n, bases, _ = arrays["sh"].shape
    rows = []
    for i in range(n):
        r = [("x", "float", arrays["means"][i, 0]), ("y", "float", arrays["means"][i, 1]),
             ("z", "float", arrays["means"][i, 2]),
             ("nx", "float", 0.0), ("ny", "float", 0.0), ("nz", "float", 0.0)]
        r += [(f"f_dc_{c}", "float", arrays["sh"][i, 0, c]) for c in range(3)]
        k = 0
        for c in range(3):
            for j in range(1, bases):
                r.append((f"f_rest_{k}", "float", arrays["sh"][i, j, c]))
                k += 1
        r.append(("opacity", "float", arrays["opacities"][i]))
        r += [(f"scale_{a}", "float", arrays["scales"][i, a]) for a in range(3)]
        r += [(f"rot_{a}", "float", arrays["rotations"][i, a]) for a in range(4)]
        if extra:
            r = extra(i, r)
        rows.append(r)
    return rows


def write_ply(path, rows, fmt="binary_little_endian", trailing_element=False):
    names = [(name, t) for name, t, _ in rows[0]] if rows else []
    head = ["ply", f"format {fmt} 1.0", "comment written by the tests", f"element vertex {len(rows)}"]
    head += [f"property {t} {name}" for name, t in names]
    if trailing_element:
        head += ["element face 0", "property list uchar int vertex_indices"]
    head.append("end_header")
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode("ascii"))
        for r in rows:
            if fmt == "ascii":
                f.write((" ".join(repr(int(v)) if t not in ("float", "double") else repr(float(v))
                                  for _, t, v in r) + "\n").encode("ascii"))
            else:
                e = "<" if fmt == "binary_little_endian" else ">"
                f.write(b"".join(struct.pack(e + _PACK[t], int(v) if t not in ("float", "double") else float(v))
                                 for _, t, v in r))


def random_arrays(n, bases, seed=0):
    rng = np.random.default_rng(seed)
    f32 = np.float32
    return {"means": rng.normal(size=(n, 3)).astype(f32), "scales": rng.normal(-4, 1, (n, 3)).astype(f32),
            "rotations": rng.normal(size=(n, 4)).astype(f32), "opacities": rng.normal(size=n).astype(f32),
            "sh": rng.normal(size=(n, bases, 3)).astype(f32)}
