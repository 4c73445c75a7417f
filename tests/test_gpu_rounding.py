"""The fast correctly rounded arithmetic of the kernels (kscr_math.h:
div_rn for the divisions by a rest length, sqrt_rn / rsqrt_rn for lengths
and normalization, the reverse sweep's clamped sqrt_cr / rcp_cr) against the
compiler's IEEE division and square root, exhaustively on the GPU through
kscr_selftest_rounding: every float bit pattern for the square roots, every
numerator for the divisions by rest lengths of 0.1 mm to 10 cm."""
import pytest

import kamisimulator_amd as ka

pytestmark = pytest.mark.gpu


def test_fast_arithmetic_equals_ieee():
    divisors = (1e-4, 7.3e-4, 1e-3, 2.5e-3, 0.1)
    print(f"[rounding] all 2^32 patterns, divisors {divisors}", flush=True)
    sqrt_diff, rsqrt_diff, sweep_diff, div_diff = ka.selftest_rounding(divisors, device=0)
    assert sqrt_diff == 0
    assert rsqrt_diff == 3  # +0, -0, +inf: NaN instead of inf / 0; the scaled vector is NaN either way
    assert sweep_diff == 0
    assert div_diff == 0
