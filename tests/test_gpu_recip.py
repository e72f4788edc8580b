"""The kernels' reciprocal (hardware reciprocal + one fma Newton step on
|w| in [2^-125, 2^125], the IEEE division elsewhere) equals the IEEE quotient
1.f / w for every non-NaN float on the device: an exhaustive comparison, so
the bit-exact parity of everything that divides rests on it."""
import pytest

pytestmark = pytest.mark.gpu


def test_recip_equals_the_ieee_quotient_for_every_float(ctx):
    compared, bad, fast = ctx.check_recip()
    nan = 2 * ((1 << 23) - 1)
    assert compared == (1 << 32) - nan
    # |w| in [2^-125, 2^125]: exponents -125..124 (250 binades) plus 2^125,
    # both signs
    assert fast == 2 * (250 * (1 << 23) + 1)
    assert bad == 0
