"""The oracle's fluid passes against closed forms and brute force (CPU only).

The checks themselves live in tests/fluid_pins.py, where each names the
reference code it follows; tests/test_oracle_planted_mistakes.py runs them
against builds of the oracle with mistakes planted.
"""
import pytest

import fluid_pins


@pytest.mark.parametrize("check", fluid_pins.CHECKS, ids=lambda c: c.__name__[len("check_"):])
def test_fluid_pass_pin(check):
    check(None)
