"""The elastic studies scene (demo/presets/elastic.ts:78-230) built from the
reference's baked templates without three.js: twenty locally shape-matched
forms with selfCollision contacts, squeezed by moving plates, against the
FP32 oracle. Local shape matching's orientation update uses sinf/cosf, so
the comparison is held to the 1e-5 relative tolerance of the local tests;
and a PSDF buffer gives the same SDF collider as the data it encodes."""
import os

import numpy as np
import pytest

import particle_fluids as pf
from particle_fluids import formats, scenes
from helpers import oracle_elastic_scene, rel_err, step_oracle_elastic

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "elastic")


def templates():
    return {n: formats.load_sampled_body(os.path.join(GOLDEN, f"{n}-250.json")) for n in scenes.ELASTIC_NAMES}


def test_elastic_studies_match_oracle():
    print("[elastic] 20 forms of 250, selfCollision contacts, moving plates vs oracle", flush=True)
    sc = scenes.build_elastic_scene(templates(), 5000)
    o, plates = oracle_elastic_scene(sc)
    errs, counts = [], []
    dt = 1.0 / 60.0
    for k in range(4):
        scenes.step_elastic_scene(sc, dt)
        step_oracle_elastic(o, plates, sc["time"], dt)
        s = sc["ps"].readback()
        ref = o.get()
        errs.append((rel_err(s.positions[:, :3], ref["x"]), rel_err(s.velocities[:, :3], ref["v"])))
        counts.append((sc["loop"].contacts.readbackCount(), o.contacts()[0]))
    print(f"[elastic] rel err per step (x, v) {errs}; contacts (gpu, oracle) {counts}", flush=True)
    assert counts[0][0] == counts[0][1] > 0
    for ex, ev in errs:
        assert ex <= 1e-5 and ev <= 1e-3, errs
    assert not sc["loop"].readbackOverflow().contacts
    for key in ("loop", "soft", "walls", "ps"):
        sc[key].dispose()


def test_sdf_collider_from_psdf_bytes():
    """A sphere SDF through encodeSdfBinary / decodeSdfBinary drives the same
    collider as the in-memory field, bit for bit."""
    res = (24, 24, 24)
    vox = 0.05
    g = (np.stack(np.meshgrid(*[np.arange(n) for n in res], indexing="ij"), -1) + 0.5) * vox - 0.6
    phi = (np.linalg.norm(g, axis=-1) - 0.3).astype(np.float32).transpose(2, 1, 0).reshape(-1)  # x fastest
    sdf = {"data": phi, "resolution": res, "origin": (-0.6, -0.6, -0.6), "voxelSize": (vox, vox, vox)}
    back = formats.decodeSdfBinary(formats.encodeSdfBinary(sdf))
    rng = np.random.default_rng(5)
    pos = rng.uniform(-0.25, 0.25, (500, 3)) + np.array([0, 0.6, 0])
    runs = []
    for field in (sdf, back):
        ps = pf.ParticleSystem(len(pos), 0.01)
        ps.uploadParticles(pos)
        col = pf.SDFCollider(ps, field)
        loop = pf.SimLoop(ps, substeps=2, iterations=2, colliders=[col])
        for _ in range(20):
            loop.step(1 / 60)
        runs.append(ps.readback().positions.copy())
        loop.dispose()
        col.dispose()
        ps.dispose()
    assert np.array_equal(runs[0], runs[1])
    # the particles came to rest on the sphere, not through it
    r = np.linalg.norm(runs[0][:, :3].astype(np.float64), axis=1)
    assert r.min() > 0.3
