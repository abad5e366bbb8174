"""The oracle's hash grid pinned against the reference tests' known answers.

CPU only. Known answers come from the reference's own test files; cited per test.
"""
import numpy as np
import pytest

import oracle as orc
from particle_fluids import scenes


def interleave(ux, uy, uz):
    """Morton code by explicit bit interleaving (independent of part1by2's magic masks)."""
    code = 0
    for bit in range(10):
        code |= ((ux >> bit) & 1) << (3 * bit)
        code |= ((uy >> bit) & 1) << (3 * bit + 1)
        code |= ((uz >> bit) & 1) << (3 * bit + 2)
    return code


def test_part1by2_spreads_every_third_bit():
    # mortonHash.ts:30-38: bit k of the low 10 bits lands at bit 3k.
    for n in list(range(0, 1024, 7)) + [1023, 0xFFFFFFFF, 0x400, 0x7FF]:
        expect = 0
        for k in range(10):
            expect |= ((n >> k) & 1) << (3 * k)
        assert orc.part1by2(n) == expect


def test_morton_hand_verified_cells():
    # morton-encoding.gpu.test.ts:46-61 cells, checked against explicit interleaving.
    cells = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1), (-1, 0, 0), (0, -1, 0), (0, 0, -1),
             (-1, -1, -1), (10, 20, 30), (-10, -20, -30), (511, 0, 0), (-512, 0, 0), (511, 511, 511)]
    for c in cells:
        u = [(v + 512) & 0x3FF for v in c]
        assert orc.morton(*c) == interleave(*u)
    assert orc.morton(0, 0, 0) == 0x38000000  # bias 512 = bit 9 on every axis
    assert orc.morton(-512, -512, -512) == 0


def test_morton_cells_bucket_through_grid():
    # morton-encoding.gpu.test.ts:46-95: particles at cell centres, H = 2^18, cell 0.1.
    cells = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1), (-1, 0, 0), (0, -1, 0), (0, 0, -1),
             (-1, -1, -1), (10, 20, 30), (-10, -20, -30), (511, 0, 0), (-512, 0, 0), (511, 511, 511)]
    pos = np.zeros((len(cells), 4), np.float32)
    for i, c in enumerate(cells):
        pos[i, :3] = [(v + 0.5) * 0.1 for v in c]
    g = orc.grid_build(pos, 0.1, 1 << 18)
    for i, c in enumerate(cells):
        assert g["cellIndex"][i] == orc.morton(*c) & ((1 << 18) - 1)
    assert not g["overflow"]


def test_overflow_flag_known_answers():
    # morton-encoding.gpu.test.ts:101-126: one cell past the Morton range sets the flag.
    pos = np.array([[(512 + 0.5) * 0.1, 0, 0, 0]], np.float32)
    assert orc.grid_build(pos, 0.1, 1024)["overflow"]
    pos = np.array([[(511 + 0.5) * 0.1, 0, 0, 0]], np.float32)
    assert not orc.grid_build(pos, 0.1, 1024)["overflow"]
    # hashgrid-overflow-flag.gpu.test.ts:60-82: a particle at 1e10 m (clamp fires).
    pos = np.array([[1e10, 0, 0, 0], [0.01, 0.01, 0.01, 0]], np.float32)
    assert orc.grid_build(pos, 0.05, 1024)["overflow"]


def lattice16():
    n, cell = 16, 0.05
    pos = np.zeros((n ** 3, 4), np.float32)
    k = 0
    for iz in range(n):
        for iy in range(n):
            for ix in range(n):
                pos[k, :3] = [ix * cell + cell / 2, iy * cell + cell / 2, iz * cell + cell / 2]
                k += 1
    return pos


def true_candidates16():
    n = 16
    total = 0
    for iz in range(n):
        for iy in range(n):
            for ix in range(n):
                for dz in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dx in (-1, 0, 1):
                            if dx == dy == dz == 0:
                                continue
                            if 0 <= ix + dx < n and 0 <= iy + dy < n and 0 <= iz + dz < n:
                                total += 1
    return total


# hashgrid-collision-rate.gpu.test.ts:31-38: measured false-positive rates.
COLLISION_CURVE = {1024: 370.0, 2048: 132.8, 4096: 14.2, 8192: 9.3, 16384: 4.5, 32768: 0.0}


def test_collision_rate_curve_matches_reference():
    pos = lattice16()
    true = true_candidates16()
    for H, rate in COLLISION_CURVE.items():
        cand, _ = orc.grid_query_counts(pos, 0.05, H, pos, 0.05)
        measured = (int(cand.sum()) - true) / true * 100.0
        assert round(measured, 1) == pytest.approx(rate, abs=0.051), (H, measured)


def test_lattice_invariants_and_neighbour_counts():
    # hashgrid-lattice.gpu.test.ts:24-125: 4^3 lattice, spacing 0.05, h 0.08.
    L, sp, h = 4, 0.05, 0.08
    pos = np.zeros((L ** 3, 4), np.float32)
    k = 0
    for kk in range(L):
        for j in range(L):
            for i in range(L):
                pos[k, :3] = [i * sp, j * sp, kk * sp]
                k += 1
    n = len(pos)
    H = 128  # nextPow2(2 * 64)
    g = orc.grid_build(pos, h, H)
    assert g["counts"][:H].sum() == n
    prefix = np.concatenate([[0], np.cumsum(g["counts"][:H])[:-1]])
    assert np.array_equal(g["cellStart"][:H], prefix)
    assert np.array_equal(g["cellEnd"][:H], prefix + g["counts"][:H])
    assert np.array_equal(np.sort(g["sortedIndices"]), np.arange(n))
    assert (g["cellIndex"] < H).all()
    _, within = orc.grid_query_counts(pos, h, H, pos, h)
    d = pos[:, None, :3].astype(np.float64) - pos[None, :, :3].astype(np.float64)
    brute = ((d ** 2).sum(-1) < h * h).sum(1) - 1
    assert np.array_equal(within, brute)


def test_random_stress_sampled_counts():
    # hashgrid-random-stress.gpu.test.ts:33-101: 100k LCG points (seed 0xc0ffee), h 0.05.
    n, h, margin = 100_000, 0.05, 0.01
    u = scenes.lcg_uniform(0xC0FFEE, 3 * n).reshape(n, 3)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = (margin + u * (1 - 2 * margin)).astype(np.float32)
    H = min(1 << 18, 1 << 20)
    sample = np.arange(0, n, n // 100)[:100]
    _, within = orc.grid_query_counts(pos, h, H, pos, h)  # query i is particle i: self excluded
    for i in sample:
        d = pos[:, :3] - pos[i, :3]
        dd = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        assert within[i] == int((dd < np.float32(h * h)).sum()) - 1


def test_stable_counting_sort_matches_numpy_stable_argsort():
    pos = scenes.sort_points(200_000, "blobs")
    g = orc.grid_build(pos, 0.01, 1 << 22)
    keys = g["cellIndex"]
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(g["sortedIndices"], order)
    assert np.array_equal(g["sortedKeys"], keys[order])
