"""pf_grid_sort_noalloc (the sort entry point that allocates nothing) against
the standalone HashGrid rebuild + readback, bit for bit, on two calls in a
row, the second smaller; it is called through the C ABI with device buffers
the test allocates through the HIP runtime libpf.so itself links."""
import ctypes as C

import numpy as np
import pytest

import particle_fluids as pf
from particle_fluids import _lib, scenes

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipMemset / hipFree of libamdhip64.so.7 (the
    runtime libpf.so is linked against), for the test's device buffers."""

    def __init__(self):
        _lib.load()
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.live = []

    def check(self, rc, what):
        assert rc == 0, f"{what} failed with HIP error {rc}"

    def alloc(self, nbytes, fill=None):
        p = C.c_void_p()
        self.check(self.h.hipMalloc(C.byref(p), max(nbytes, 4)), "hipMalloc")
        self.live.append(p)
        if fill is not None:
            self.check(self.h.hipMemset(p, fill, nbytes), "hipMemset")
        return p

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        p = self.alloc(a.nbytes)
        self.check(self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1), "hipMemcpy H2D")
        return p

    def download(self, p, n, dtype=np.uint32):
        out = np.empty(n, dtype)
        self.check(self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def free_all(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def standalone_sort(pos, cell, H):
    g = pf.HashGrid.standalone(len(pos), cell, H)
    g.setSortedPositions(False)
    g.uploadPositions(pos)
    g.rebuild()
    snap = g.readback()
    keys = g.readbackSortedKeys()
    over = g.readbackOverflow()
    g.dispose()
    return dict(cellIndex=snap.cellIndex, counts=snap.counts, cellStart=snap.cellStart, cellEnd=snap.cellEnd,
                sortedIndices=snap.sortedIndices, sortedKeys=keys, overflow=over)


def noalloc_sort(hip, ctx, pos, cell, H, origin=(0.0, 0.0, 0.0)):
    n = len(pos)
    hpad = (H + 1023) // 1024 * 1024
    x = hip.upload(np.ascontiguousarray(pos, np.float32))
    # outputs start as garbage (0xA5 bytes), so nothing may rely on their contents
    sizes = (("cellIndex", n), ("counts", hpad), ("cellStart", hpad), ("cellEnd", hpad), ("sortedIndices", n),
             ("sortedKeys", n), ("overflow", 1))
    out = {k: hip.alloc(4 * m, fill=0xA5) for k, m in sizes}
    o = (C.c_double * 3)(*origin)
    _lib.call("pf_grid_sort_noalloc", ctx.handle, x, n, float(cell), H, o, out["cellIndex"], out["counts"],
              out["cellStart"], out["cellEnd"], out["sortedIndices"], out["sortedKeys"], out["overflow"])
    ctx.synchronize()
    res = {k: hip.download(out[k], m) for k, m in sizes}
    res["overflow"] = bool(res["overflow"][0])
    hip.free_all()
    return res


def check_equal(got, want, what):
    for k in ("cellIndex", "counts", "cellStart", "cellEnd", "sortedIndices", "sortedKeys"):
        assert np.array_equal(got[k], want[k]), f"{what}: {k} differs"
    assert got["overflow"] == want["overflow"], f"{what}: overflow flag differs"


def test_noalloc_sort_equals_standalone_rebuild_twice():
    hip = Hip()
    ctx = pf.default_context()
    cases = [
        ("4M blobs, 2^22 buckets", scenes.sort_points(4_194_304, "blobs"), 0.01, 1 << 22),
        ("70001 uniform, 2^17 buckets", scenes.sort_points(70001, "uniform", seed=5), 0.05, 1 << 17),
    ]
    for what, pos, cell, H in cases:
        print(f"[noalloc] {what}", flush=True)
        want = standalone_sort(pos, cell, H)
        got = noalloc_sort(hip, ctx, pos, cell, H)
        check_equal(got, want, what)
    # an overflowing input: the flag comes back through the device word
    pos = scenes.sort_points(1000, "uniform", seed=3)
    pos[7, :3] = 1e10
    print("[noalloc] overflow flag", flush=True)
    want = standalone_sort(pos, 0.05, 1024)
    got = noalloc_sort(hip, ctx, pos, 0.05, 1024)
    assert want["overflow"] is True
    check_equal(got, want, "overflow")


def test_noalloc_sort_rejects_sizes_beyond_its_scratch():
    hip = Hip()
    ctx = pf.default_context()
    buf = hip.alloc(64, fill=0)
    with pytest.raises(pf.PFError, match="sorts 1 to"):
        _lib.call("pf_grid_sort_noalloc", ctx.handle, buf, (1 << 22) + 1, 0.01, 1024, None, *[buf] * 7)
    with pytest.raises(pf.PFError, match="exceeds this entry point"):
        _lib.call("pf_grid_sort_noalloc", ctx.handle, buf, 1, 0.01, 1 << 23, None, *[buf] * 7)
    hip.free_all()
