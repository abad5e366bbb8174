"""On-disk formats (CPU): the PSDF codec (src/sdf/writeBinary.ts:21-106) and
the baked soft-body templates (SampledBody JSON), pinned by the reference's
own checks: tests/analytical/sdf/sdf-validation.test.ts:28-34 for the codec's
validation, tests/analytical/softbody/study-body.test.ts:30-86 for the
templates (run here on the reference's 250-particle templates)."""
import os
import struct

import numpy as np
import pytest

import particle_fluids as pf
from particle_fluids import formats, scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "elastic")


def field(res, voxel=(1.0, 1.0, 1.0)):
    return {"data": np.zeros(int(np.prod(res)), np.float32), "resolution": res, "origin": (0.0, 0.0, 0.0),
            "voxelSize": voxel}


def test_psdf_layout_and_round_trip():
    rng = np.random.default_rng(3)
    sdf = {"data": rng.standard_normal(5 * 6 * 7).astype(np.float32), "resolution": (5, 6, 7),
           "origin": (-0.25, 0.125, 3.5), "voxelSize": (0.01, 0.02, 0.03)}
    raw = formats.encodeSdfBinary(sdf)
    assert len(raw) == 48 + 4 * 210
    assert raw[:4] == b"PSDF"
    assert struct.unpack("<5I", raw[4:24]) == (1, 5, 6, 7, 0)
    assert struct.unpack("<6f", raw[24:48]) == tuple(np.float32(v) for v in (-0.25, 0.125, 3.5, 0.01, 0.02, 0.03))
    assert np.array_equal(np.frombuffer(raw[48:], "<f4"), sdf["data"])
    back = formats.decodeSdfBinary(raw)
    assert back["resolution"] == (5, 6, 7)
    assert np.array_equal(back["data"], sdf["data"])
    assert back["origin"] == tuple(float(np.float32(v)) for v in sdf["origin"])
    assert formats.encodeSdfBinary(back) == raw


def test_psdf_validation_messages():
    raw = bytearray(formats.encodeSdfBinary(field((2, 2, 2))))
    raw[8:12] = struct.pack("<I", 1)
    with pytest.raises(pf.PFError, match="decodeSdfBinary: resolution must be ≥ 2"):
        formats.decodeSdfBinary(bytes(raw))
    with pytest.raises(pf.PFError, match="decodeSdfBinary: voxelSize must be positive"):
        formats.decodeSdfBinary(formats.encodeSdfBinary(field((2, 2, 2), (1.0, 0.0, 1.0))))
    good = formats.encodeSdfBinary(field((2, 2, 2)))
    with pytest.raises(pf.PFError, match=r"decodeSdfBinary: buffer too small for header \(40 < 48\)"):
        formats.decodeSdfBinary(good[:40])
    with pytest.raises(pf.PFError, match='decodeSdfBinary: bad magic 0x46445351 \\(expected 0x46445350 "PSDF"\\)'):
        formats.decodeSdfBinary(b"QSDF" + good[4:])
    with pytest.raises(pf.PFError, match="decodeSdfBinary: unsupported version 2"):
        formats.decodeSdfBinary(good[:4] + struct.pack("<I", 2) + good[8:])
    with pytest.raises(pf.PFError, match=r"decodeSdfBinary: buffer too small \(79 < 80 for 2·2·2 voxels\)"):
        formats.decodeSdfBinary(good[:-1])
    with pytest.raises(pf.PFError, match="encodeSdfBinary: resolution 2·2·3=12 does not match data length 8"):
        formats.encodeSdfBinary(dict(field((2, 2, 2)), resolution=(2, 2, 3)))


def load(name, budget=250):
    return formats.load_sampled_body(os.path.join(GOLDEN, f"{name}-{budget}.json"))


@pytest.mark.parametrize("name", scenes.ELASTIC_NAMES)
def test_templates_are_connected_and_non_overlapping(name):
    """study-body.test.ts:30-86 at the 250-particle level."""
    budget = 250
    body = load(name, budget)
    count = len(body["positions"])
    assert budget <= count <= int(np.ceil(budget * 1.2))
    e = body["edges"].astype(np.int64)
    assert np.all(e[:, 0] < e[:, 1]) and e.min() >= 0 and e.max() < count
    nbrs = [[] for _ in range(count)]
    for a, b in e:
        nbrs[a].append(b)
        nbrs[b].append(a)
    seen, todo = set(), [0]
    while todo:
        k = todo.pop()
        if k in seen:
            continue
        seen.add(k)
        todo.extend(j for j in nbrs[k] if j not in seen)
    assert len(seen) == count
    ns = body["surfaceCount"]
    assert 0 < ns < count
    assert np.all(body["surface"][:ns] == 1) and np.all(body["surface"][ns:] == 0)
    diameter = 0.03 * (200 / budget) ** (1 / 3)
    p = body["positions"].astype(np.float64)
    d = np.linalg.norm(p[:, None] - p[None], axis=-1) + np.eye(count) * 10
    assert abs(d.min() - diameter) < 5e-6
    if name == "donut":
        assert np.hypot(p[:, 0], p[:, 2]).min() > diameter


def test_template_loader_rejects_malformed_bodies():
    body = {"positions": [0, 0, 0, 1, 0, 0], "surface": [0, 1], "edges": [0, 1], "scale": 1}
    with pytest.raises(pf.PFError, match="surface particles must come first"):
        formats.load_sampled_body(body)
    with pytest.raises(pf.PFError, match="edges must be pairs"):
        formats.load_sampled_body(dict(body, surface=[1, 0], edges=[1, 0]))
    with pytest.raises(pf.PFError, match="xyz triples"):
        formats.load_sampled_body(dict(body, positions=[0, 0]))


def test_elastic_layout_places_twenty_forms():
    """elastic.ts:78-150: twenty forms at the level nearest particles / 20,
    each weighing the same, inside the tank between the plates."""
    templates = {n: load(n) for n in scenes.ELASTIC_NAMES}
    lay = scenes.elastic_layout(templates, 5000)
    assert lay["per_body"] == 250 and len(lay["bodies"]) == 20
    assert abs(lay["radius"] - 0.015 * (200 / 250) ** (1 / 3)) < 1e-15
    for k, b in enumerate(lay["bodies"]):
        st, c = b["range"]
        shape = templates[scenes.ELASTIC_NAMES[k % 4]]
        assert c == len(shape["positions"]) and np.all(lay["inv_mass"][st:st + c] == np.float32(c / 20))
        # a rigid placement: pairwise distances of the template survive
        p = lay["positions"][st:st + c].astype(np.float64)
        q = shape["positions"].astype(np.float64)
        assert abs(np.linalg.norm(p[0] - p[-1]) - np.linalg.norm(q[0] - q[-1])) < 1e-6
    x = lay["positions"]
    assert np.abs(x[:, 0]).max() < 1.05 and np.abs(x[:, 2]).max() < 0.87 and x[:, 1].min() > 0


def test_seeded_random_is_mulberry32():
    """demo/presets/shared.ts:66-74; first draws of seed 0x9e3779b9 in [0, 1),
    and the generator's 32-bit state steps by 0x6d2b79f5."""
    r = formats.seeded_random(0x9E3779B9)
    draws = [r() for _ in range(1000)]
    assert all(0.0 <= d < 1.0 for d in draws)
    assert abs(np.mean(draws) - 0.5) < 0.03
    a, b = formats.seeded_random(1), formats.seeded_random(1)
    assert [a() for _ in range(5)] == [b() for _ in range(5)]


HONEY = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "honey", "bunny.sdf.bin")


def test_reference_bunny_psdf_decodes_to_a_distance_field():
    """The reference's own bunny.sdf.bin: the codec reads it and writes the
    same bytes back, and what it reads is a signed distance field (|grad phi|
    = 1 away from the surface, positive on the grid's boundary, one connected
    negative interior), which a wrong header size, offset or byte order would
    break."""
    raw = open(HONEY, "rb").read()
    sdf = formats.decodeSdfBinary(raw)
    nx, ny, nz = sdf["resolution"]
    assert len(raw) == 48 + 4 * nx * ny * nz
    assert formats.encodeSdfBinary(sdf) == raw
    phi = sdf["data"].reshape(nz, ny, nx).astype(np.float64)
    vx, vy, vz = sdf["voxelSize"]
    gz, gy, gx = np.gradient(phi, vz, vy, vx)
    g = np.sqrt(gx * gx + gy * gy + gz * gz)
    band = (np.abs(phi) > 2 * max(vx, vy, vz)) & (np.abs(phi) < 8 * max(vx, vy, vz))
    assert band.sum() > 1000
    assert np.median(np.abs(g[band] - 1.0)) < 0.05
    inside = phi < 0
    assert inside.sum() > 100 and phi[0].min() > 0 and phi[-1].min() > 0
    assert phi[:, 0].min() > 0 and phi[:, -1].min() > 0 and phi[:, :, 0].min() > 0 and phi[:, :, -1].min() > 0
    from scipy import ndimage
    _, parts = ndimage.label(inside)
    assert parts == 1
