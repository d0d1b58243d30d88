"""Each pin below must fail on an oracle built with one deliberate mistake
(oracle/xo_core.hpp, XO_SLIP): proof that the pin constrains the function it
is said to pin, not only that it passes.  The slips are planted in the CPU
oracle only."""
import pytest

import pyoracle as po
from test_oracle_pins import check_film_position, check_roulette_passthrough, check_two_lights

SLIPS = [
    (1, "film y not flipped", check_film_position),
    (2, "film jitter centred on the pixel centre", check_film_position),
    (3, "roulette survivors not reweighted", check_roulette_passthrough),
    (4, "light pdf without 1 / lights.size()", check_two_lights),
    (5, "light pdf without 1 / ntri", check_two_lights),
    (6, "light pdf without the area scaling", check_two_lights),
]


@pytest.mark.parametrize("k,what,check", SLIPS, ids=[s[1] for s in SLIPS])
def test_pin_catches_planted_slip(k, what, check):
    L = po.slip_lib(k)
    with pytest.raises(AssertionError):
        check(L)
