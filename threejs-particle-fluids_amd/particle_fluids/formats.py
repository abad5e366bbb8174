"""On-disk formats of the reference's assets.

- The PSDF signed-distance-field binary (src/sdf/writeBinary.ts:21-106),
  through the library's codec (``pf_sdf_decode`` / ``pf_sdf_encode``).
- Baked soft-body templates, ``SampledBody`` JSON (demo/presets/elastic.ts:
  26-34, written by scripts/prepare-elastic-assets.ts): positions, per-particle
  surface flags (surface first) and local edges, plus the render-mesh scale.
"""
import ctypes as C
import json
import os

import numpy as np

from . import _lib
from ._lib import PFError, call

f32p = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)
f64p = C.POINTER(C.c_double)


def decodeSdfBinary(buffer):
    """SDFData (dict of data, resolution, origin, voxelSize) from PSDF bytes;
    raises with the reference's messages on a malformed buffer."""
    raw = bytes(buffer)
    res = (C.c_uint32 * 3)()
    origin = (C.c_double * 3)()
    voxel = (C.c_double * 3)()
    count = C.c_size_t()
    call("pf_sdf_decode", raw, len(raw), res, origin, voxel, None, 0, C.byref(count))
    data = np.empty(count.value, np.float32)
    call("pf_sdf_decode", raw, len(raw), res, origin, voxel, data.ctypes.data_as(f32p), data.size, C.byref(count))
    return {"data": data, "resolution": (int(res[0]), int(res[1]), int(res[2])),
            "origin": tuple(float(v) for v in origin), "voxelSize": tuple(float(v) for v in voxel)}


def encodeSdfBinary(sdf):
    """PSDF bytes of an SDFData dict (the voxels are stored as f32)."""
    data = np.ascontiguousarray(np.asarray(sdf["data"], np.float32).reshape(-1))
    res = (C.c_uint32 * 3)(*[int(v) for v in sdf["resolution"]])
    origin = (C.c_double * 3)(*[float(v) for v in sdf["origin"]])
    voxel = (C.c_double * 3)(*[float(v) for v in sdf["voxelSize"]])
    size = C.c_size_t()
    call("pf_sdf_encode", data.ctypes.data_as(f32p), data.size, res, origin, voxel, None, 0, C.byref(size))
    out = C.create_string_buffer(size.value)
    call("pf_sdf_encode", data.ctypes.data_as(f32p), data.size, res, origin, voxel, out, size.value, C.byref(size))
    return out.raw[:size.value]


def load_sdf(path):
    """decodeSdfBinary of a .sdf.bin file."""
    with open(path, "rb") as f:
        return decodeSdfBinary(f.read())


def load_sampled_body(src):
    """A SampledBody template from a JSON path, JSON text or dict, checked:
    positions in xyz triples, one surface flag per particle with the surface
    particles first, edges as in-range local pairs (a < b). Returns dict of
    positions (count, 3) float32, surface (count,) uint8, surfaceCount, edges
    (E, 2) uint32 and scale."""
    if isinstance(src, dict):
        body = src
    elif isinstance(src, (str, os.PathLike)) and os.path.exists(src):
        with open(src) as f:
            body = json.load(f)
    else:
        body = json.loads(src)
    pos = np.asarray(body["positions"], np.float64)
    if pos.size % 3:
        raise PFError(1, "SampledBody: positions must hold xyz triples")
    count = pos.size // 3
    surface = np.asarray(body.get("surface", [1] * count), np.int64)
    if surface.size != count:
        raise PFError(1, f"SampledBody: {surface.size} surface flags for {count} particles")
    nsurf = int((surface != 0).sum())
    if not (np.all(surface[:nsurf] == 1) and np.all(surface[nsurf:] == 0)):
        raise PFError(1, "SampledBody: surface particles must come first")
    edges = np.asarray(body.get("edges", []), np.int64)
    if edges.size % 2:
        raise PFError(1, "SampledBody: edges must be index pairs")
    edges = edges.reshape(-1, 2)
    if edges.size and (edges.min() < 0 or edges.max() >= count or np.any(edges[:, 0] >= edges[:, 1])):
        raise PFError(1, "SampledBody: edges must be pairs a < b of particle indices")
    return {"positions": pos.reshape(-1, 3).astype(np.float32), "surface": surface.astype(np.uint8),
            "surfaceCount": nsurf, "edges": edges.astype(np.uint32), "scale": float(body.get("scale", 1.0))}


def seeded_random(seed):
    """The demo's deterministic [0, 1) generator (demo/presets/shared.ts:
    66-74), Tommy Ettinger's mulberry32, in 32-bit integer arithmetic."""
    state = [seed & 0xFFFFFFFF]

    def imul(a, b):
        return (a * b) & 0xFFFFFFFF

    def nxt():
        state[0] = (state[0] + 0x6D2B79F5) & 0xFFFFFFFF
        s = state[0]
        t = imul(s ^ (s >> 15), 1 | s)
        t = ((t + imul(t ^ (t >> 7), 61 | t)) & 0xFFFFFFFF) ^ t
        return ((t ^ (t >> 14)) & 0xFFFFFFFF) / 4294967296.0

    return nxt
