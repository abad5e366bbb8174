"""Argument checks that need a live particle system (GPU)."""
import numpy as np
import pytest

import particle_fluids as pf

pytestmark = pytest.mark.gpu


def test_softbody_rest_positions_must_match_the_body():
    # SoftbodySystem.ts:171-173: a body of n particles needs 3n rest coordinates.
    ps = pf.ParticleSystem(10, 0.01)
    cloud = np.random.default_rng(1).uniform(0.0, 0.05, (10, 3)).astype(np.float32)
    ps.uploadParticles(cloud)
    with pytest.raises(pf.PFError, match="body 0 needs 30 rest coordinates"):
        pf.SoftbodySystem(ps, [{"range": (0, 10), "restPositions": np.zeros(27, np.float32)}])
    with pytest.raises(ValueError, match="9 velocities for 10 particles"):
        ps.uploadParticles(np.zeros((10, 3), np.float32), np.zeros((9, 3), np.float32))
    sb = pf.SoftbodySystem(ps, [{"range": (0, 10), "restPositions": cloud.reshape(-1)}])
    sb.dispose()
    ps.dispose()
