"""Shared test helpers: building the same scene on the GPU and in the oracle."""
import numpy as np

import oracle as orc
from particle_fluids import scenes


def oracle_dam_break(positions, radius, substeps, iterations, precision="f32", H=None, effects=None):
    """The dam break of particle_fluids.scenes in the CPU oracle."""
    n = len(positions)
    spacing = 2 * radius
    h = 2 * spacing
    if H is None:
        H = 1
        while H < 2 * n:
            H <<= 1
        H = min(H, 1 << 20)
    fx = dict(scenes.FLUID_DEFAULTS) if effects is None else effects
    s = orc.Scene(n, radius, precision)
    s.set_particles(positions)
    s.set_loop(substeps, iterations, (0.0, -9.81, 0.0), h, H)
    s.add_fluid(0, n, 1000.0, spacing, h, 1e-4, viscosity=fx.get("viscosity"), vorticity=fx.get("vorticity"),
                surface_tension=fx.get("surfaceTension"), adhesion=fx.get("adhesion"))
    mn, mx = scenes.CONTAINER[:3], scenes.CONTAINER[3:]
    for nrm, which in scenes.PLANES:
        s.add_plane(nrm, mn if which == "min" else mx)
    return s


def rel_err(a, b):
    """max |a - b| / max |b| (inf-norm relative error)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(a - b).max()
    s = np.abs(b).max()
    return float(d / s) if s > 0 else float(d)
