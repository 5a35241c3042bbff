"""3DGS PLY scenes -> renderer arrays on the device, in Morton order.

Replaces the array assembly of the reference's model loader (Model.py:22-79, one directory
above the backend package): plyfile reads the first element's columns, each is converted
to float32, the SH coefficients are re-laid from the file's channel-major ``f_rest_*``
to (N, B, 3), and every array is reordered by the Morton code of the means.

Here the host only parses the header and streams the raw vertex records (memory-mapped,
through two pinned staging buffers) to the device; the HIP kernels of libsgs.so decode the
records (sgs_ply_decode), compute the Morton order (sgs_morton_order) and permute the arrays
(sgs_gather_gaussians).
"""
from __future__ import annotations

import mmap
import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import torch

from . import _native

# PLY scalar type names (both spellings of the format) -> (sgs_scalar_type, bytes, numpy code)
_TYPES = {
    "char": (1, 1, "i1"), "int8": (1, 1, "i1"),
    "uchar": (2, 1, "u1"), "uint8": (2, 1, "u1"),
    "short": (3, 2, "i2"), "int16": (3, 2, "i2"),
    "ushort": (4, 2, "u2"), "uint16": (4, 2, "u2"),
    "int": (5, 4, "i4"), "int32": (5, 4, "i4"),
    "uint": (6, 4, "u4"), "uint32": (6, 4, "u4"),
    "float": (7, 4, "f4"), "float32": (7, 4, "f4"),
    "double": (8, 8, "f8"), "float64": (8, 8, "f8"),
}
_FORMATS = ("ascii", "binary_little_endian", "binary_big_endian")
# fields in the order of sgs_ply_layout (the property names of Model.py:26-73)
BASE_FIELDS = ("x", "y", "z", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3", "opacity",
               "f_dc_0", "f_dc_1", "f_dc_2")


@dataclass
class PlyProperty:
    name: str
    type: int      # sgs_scalar_type
    size: int      # bytes
    numpy: str     # numpy type code without byte order
    offset: int = 0


@dataclass
class PlyHeader:
    format: str
    element: str                # name of the first element (its records are the Gaussians)
    count: int                  # records of the first element
    properties: list[PlyProperty] = field(default_factory=list)
    stride: int = 0             # bytes per binary record
    data_offset: int = 0        # file offset of the first element's data

    def find(self, name: str) -> PlyProperty:
        for p in self.properties:
            if p.name == name:
                return p
        raise KeyError(name)

    @property
    def total_sh_bases(self) -> int:
        n_rest = sum(1 for p in self.properties if p.name.startswith("f_rest_"))
        return n_rest // 3 + 1   # Model.py:61-62


def read_header(f) -> PlyHeader:
    """Parses the header of an open binary file; the file position is left at the data."""
    if f.readline().strip() != b"ply":
        raise ValueError("not a PLY file (first line is not 'ply')")
    fmt = None
    elements: list[tuple[str, int, list]] = []
    while True:
        raw = f.readline()
        if not raw:
            raise ValueError("PLY header has no end_header line")
        words = raw.decode("ascii", errors="replace").split()
        if not words or words[0] in ("comment", "obj_info"):
            continue
        if words[0] == "end_header":
            break
        if words[0] == "format":
            if len(words) < 2 or words[1] not in _FORMATS:
                raise ValueError(f"unsupported PLY format line: {raw!r}")
            fmt = words[1]
        elif words[0] == "element":
            if len(words) != 3:
                raise ValueError(f"malformed element line: {raw!r}")
            elements.append((words[1], int(words[2]), []))
        elif words[0] == "property":
            if not elements:
                raise ValueError("property line before any element")
            if len(words) >= 2 and words[1] == "list":
                elements[-1][2].append(("list", words[-1]))
                continue
            if len(words) != 3 or words[1] not in _TYPES:
                raise ValueError(f"unsupported property line: {raw!r}")
            elements[-1][2].append((words[1], words[2]))
        else:
            raise ValueError(f"unknown PLY header line: {raw!r}")
    if fmt is None:
        raise ValueError("PLY header has no format line")
    if not elements:
        raise ValueError("PLY file has no element")
    name, count, props = elements[0]
    hdr = PlyHeader(format=fmt, element=name, count=count, data_offset=f.tell())
    offset = 0
    for kind, pname in props:
        if kind == "list":
            raise ValueError(f"element {name!r} has a list property {pname!r}: records must have a fixed size")
        code, size, npcode = _TYPES[kind]
        hdr.properties.append(PlyProperty(pname, code, size, npcode, offset))
        offset += size
    hdr.stride = offset
    return hdr


def layout_for(hdr: PlyHeader) -> _native.sgs_ply_layout:
    """sgs_ply_layout of the fields the model loader reads (Model.py:26-73); raises KeyError
    naming a missing property, ValueError for an SH layout the renderer cannot take."""
    n_rest = sum(1 for p in hdr.properties if p.name.startswith("f_rest_"))
    if n_rest % 3:
        raise ValueError(f"{n_rest} f_rest_* properties: not a multiple of 3")
    bases = hdr.total_sh_bases
    if bases not in (1, 4, 9, 16):
        raise ValueError(f"{n_rest} f_rest_* properties give {bases} SH bases; the renderer takes 1, 4, 9 or 16")
    names = list(BASE_FIELDS) + [f"f_rest_{i}" for i in range(n_rest)]
    lay = _native.sgs_ply_layout()
    lay.stride = hdr.stride
    lay.big_endian = 1 if hdr.format == "binary_big_endian" else 0
    lay.total_sh_bases = bases
    for i, name in enumerate(names):
        try:
            p = hdr.find(name)
        except KeyError:
            raise KeyError(f"PLY element {hdr.element!r} has no property {name!r}") from None
        lay.offset[i] = p.offset
        lay.type[i] = p.type
    return lay


def _ascii_records(f, hdr: PlyHeader) -> np.ndarray:
    """The first element of an ASCII file as packed little-endian binary records (uint8)."""
    dtype = np.dtype([(p.name, "<" + p.numpy) for p in hdr.properties])
    rows = []
    for _ in range(hdr.count):
        line = f.readline()
        if not line:
            raise ValueError(f"PLY data ends after {len(rows)} of {hdr.count} records")
        words = line.split()
        if len(words) != len(hdr.properties):
            raise ValueError(f"record {len(rows)} has {len(words)} values, expected {len(hdr.properties)}")
        rows.append(tuple(int(w) if p.numpy[0] in "iu" else float(w) for w, p in zip(words, hdr.properties)))
    return np.array(rows, dtype=dtype).view(np.uint8).reshape(-1)


def _empty_arrays(n: int, bases: int, device) -> dict[str, torch.Tensor]:
    f32 = torch.float32
    return {"means": torch.empty((n, 3), dtype=f32, device=device),
            "scales": torch.empty((n, 3), dtype=f32, device=device),
            "rotations": torch.empty((n, 4), dtype=f32, device=device),
            "opacities": torch.empty((n,), dtype=f32, device=device),
            "sh": torch.empty((n, bases, 3), dtype=f32, device=device)}


def load_ply(path, device=None, *, morton: bool = True, chunk_bytes: int = 256 << 20) -> dict[str, torch.Tensor]:
    """Reads a 3DGS PLY file into device arrays ``means (N,3)``, ``scales (N,3)`` (log),
    ``rotations (N,4)`` (r,i,j,k), ``opacities (N,)`` (logits), ``sh (N,B,3)``, in Morton order
    of the means unless ``morton`` is False (then in file order). The file is streamed in
    pieces of about ``chunk_bytes``."""
    device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    if device.type != "cuda":
        raise ValueError("scene arrays live on a ROCm device")
    path = Path(path)
    with open(path, "rb") as f:
        hdr = read_header(f)
        lay = layout_for(hdr)
        n = hdr.count
        out = _empty_arrays(n, lay.total_sh_bases, device)
        if n == 0:
            return out
        if hdr.format == "ascii":
            host = _ascii_records(f, hdr)
            lay.big_endian = 0
            _decode_host_records(host, lay, out, device, chunk_bytes)
        else:
            need = hdr.data_offset + n * hdr.stride
            size = os.fstat(f.fileno()).st_size
            if size < need:
                raise ValueError(f"{path}: {size} bytes, the header announces {need}")
            with mmap.mmap(f.fileno(), 0, access=mmap.ACCESS_READ) as mm:
                host = np.frombuffer(mm, dtype=np.uint8, count=n * hdr.stride, offset=hdr.data_offset)
                try:
                    _decode_host_records(host, lay, out, device, chunk_bytes)
                finally:
                    del host
    if morton:
        order = _native.morton_order(out["means"]).to(torch.int32)
        out = _native.gather_gaussians(out, order)
    return out


def _decode_host_records(host: np.ndarray, lay, out: dict, device, chunk_bytes: int) -> None:
    """Streams whole records from host memory through two pinned buffers to the device and
    decodes them there; the host copy of piece k+1 overlaps the upload and decode of k."""
    stride = int(lay.stride)
    n = host.size // stride
    per = max(1, min(n, chunk_bytes // max(stride, 1)))
    piece = per * stride
    stream = torch.cuda.current_stream(device)
    pinned = [torch.empty(piece, dtype=torch.uint8, pin_memory=True) for _ in range(2 if n > per else 1)]
    done = [None] * len(pinned)
    dev = torch.empty(piece, dtype=torch.uint8, device=device)
    for k, first in enumerate(range(0, n, per)):
        rows = min(per, n - first)
        j = k % len(pinned)
        if done[j] is not None:
            done[j].synchronize()  # the upload that last read this buffer has finished
        nbytes = rows * stride
        pinned[j][:nbytes].numpy()[:] = host[first * stride:first * stride + nbytes]
        dev[:nbytes].copy_(pinned[j][:nbytes], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(stream)
        done[j] = ev
        _native.ply_decode(dev[:nbytes], lay, first, out)
    stream.synchronize()  # the pinned buffers are released on return
