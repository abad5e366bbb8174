"""The passes' correctly rounded reciprocal and square root, and their
saturating float-to-int conversion, exhaustively.

pf_rcp_rn and pf_sqrt_rn start from the hardware approximations (v_rcp_f32,
v_sqrt_f32, within 1 ulp) and round them exactly with fused multiply-adds.
Parity with the FP32 oracle (C `1.0f / x`, `sqrtf`) rests on their being equal
to the IEEE operations bit for bit over the ranges the passes use. These
tests compare them on the device over every float bit pattern of those
ranges (about 2^31 values each, a few milliseconds).
"""
import ctypes as C

import pytest

import particle_fluids as pf
from particle_fluids import _lib

pytestmark = pytest.mark.gpu

RCP, SQRT, SQRT_FAST, F2I = 0, 1, 2, 3


def mismatches(function, lo, hi):
    ctx = pf.default_context()
    count = C.c_uint64(0)
    first = C.c_uint32(0)
    _lib.call("pf_check_exact_math", ctx.handle, function, lo, hi, C.byref(count), C.byref(first))
    return count.value, first.value


@pytest.mark.parametrize("sign", [0, 0x80000000], ids=["positive", "negative"])
def test_reciprocal_equals_ieee_over_its_range(sign):
    # 2^-126 <= |x| < 2^126: normal inputs with normal reciprocals
    lo, hi = 0x00800000, 0x7E800000
    assert mismatches(RCP, sign + lo, sign + hi) == (0, 0xFFFFFFFF)


def test_sqrt_equals_ieee_for_every_nonnegative_float():
    # +0, denormals, normals and +inf; then -0
    assert mismatches(SQRT, 0x00000000, 0x7F800001) == (0, 0xFFFFFFFF)
    assert mismatches(SQRT, 0x80000000, 0x80000001) == (0, 0xFFFFFFFF)


def test_branch_free_sqrt_range():
    # q = 0 and 2^-96 <= q <= +inf take the branch-free form alone
    assert mismatches(SQRT_FAST, 0x00000000, 0x00000001) == (0, 0xFFFFFFFF)
    assert mismatches(SQRT_FAST, 0x0F800000, 0x7F800001) == (0, 0xFFFFFFFF)
    # below 2^-96 it is not exact, which is why pf_sqrt_rn branches there
    n, first = mismatches(SQRT_FAST, 0x00000001, 0x0F800000)
    assert n > 0 and first != 0xFFFFFFFF


def test_saturating_conversion_equals_the_branchy_form_for_every_float():
    # v_cvt_i32_f32 against the explicit clamp and NaN test, every bit pattern
    # but the last (a NaN; the range end is exclusive)
    assert mismatches(F2I, 0x00000000, 0xFFFFFFFF) == (0, 0xFFFFFFFF)
