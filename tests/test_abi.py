"""The C-ABI library loads and exports every symbol include/pf.h declares (CPU only)."""
import ctypes
import os
import re

import numpy as np
import pytest

import particle_fluids as pf
from particle_fluids import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "pf.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert set(declared()) <= set(_lib.SIGNATURES)


def test_abi_version_and_errors():
    lib = _lib.load()
    assert lib.pf_abi_version() == 1
    with pytest.raises(pf.PFError, match="spacing must be positive"):
        _lib.call("pf_layout_fill_box", pf._d3((0, 0, 0)), pf._d3((1, 1, 1)), 0.0, None, 0, None)


def test_fill_box_order_and_count():
    # layout.ts:7-14: y outer, z middle, x inner, offset half a spacing.
    pts = pf.fill_box((0, 0, 0), (0.3, 0.2, 0.1), 0.1)
    assert pts.shape == (3 * 2 * 1, 3)
    assert pts[0].tolist() == pytest.approx([0.05, 0.05, 0.05])
    assert pts[1].tolist() == pytest.approx([0.15, 0.05, 0.05])
    assert pts[3].tolist() == pytest.approx([0.05, 0.15, 0.05])


def test_fit_budget_reproduces_simulation_sizes():
    # Simulation.fitBudget on the dam break: 14x17x21 and 82x102x123 lattices.
    c = (-0.5, 0.0, -0.3, 0.5, 0.8, 0.3)
    box = [(-0.5, 0.0, -0.3, -0.1, 0.5, 0.3)]
    r, n = pf.fit_budget(5000, c, box)
    assert n == 14 * 17 * 21 == 4998
    assert r == pytest.approx(0.0139117, abs=5e-8)
    r, n = pf.fit_budget(1 << 20, c, box)
    assert n == 82 * 102 * 123 == 1028772
    assert r == pytest.approx(0.00242753, abs=5e-9)


def test_upload_arrays_must_cover_every_particle():
    # The C ABI reads n entries of each array (particles.ts:121-160 takes one record per particle).
    pos = np.zeros((5, 3), np.float32)
    p, v, m, g = pf.upload_arrays(pos, np.ones((5, 3)), 2.0, 3)
    assert v.shape == (5, 3) and (m == 2.0).all() and (g == 3).all() and g.dtype == np.uint32
    with pytest.raises(ValueError, match="4 velocities for 5 particles"):
        pf.upload_arrays(pos, np.zeros((4, 3)))
    with pytest.raises(ValueError):
        pf.upload_arrays(pos, np.zeros(14))
    with pytest.raises(ValueError, match="need one value or 5 values"):
        pf.upload_arrays(pos, None, np.ones(3))
    with pytest.raises(ValueError, match="need one value or 5 values"):
        pf.upload_arrays(pos, None, None, np.zeros(6))
