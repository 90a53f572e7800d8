"""The host evaluation of fused and unfused DM values (tests/fused.py), on
which the firstDM != 0 device tests rest."""
from fractions import Fraction

import numpy as np

from tests import fused

f32 = np.float32


def test_round_to_f32_matches_numpy_conversion():
    """numpy converts float64 to float32 with one correct rounding: every
    float64 is an exact rational, so both must agree (ties included)."""
    rng = np.random.default_rng(0)
    xs = list(rng.uniform(0, 1000, 3000)) + list(rng.uniform(0, 1e-3, 500))
    # exact ties between two float32 neighbours, even and odd significands
    for m in (2 ** 23 + 1, 2 ** 23 + 2, 3 * 2 ** 22 + 5):
        xs.append((m + 0.5) * 2.0 ** -20)
    for x in xs:
        assert fused.round_to_f32(Fraction(float(x))) == np.float32(x), x


def test_fused_and_unfused_dm_values():
    """0.1 + 3 * 0.7 in float32: two roundings give 2.2f (0x400ccccd), one
    rounding of the exact sum gives the float below it; the unfused values
    are numpy's float32 arithmetic."""
    a = fused.dm_values(0.1, 0.7, 4, fused=False)
    b = fused.dm_values(0.1, 0.7, 4, fused=True)
    assert a[3] == f32(f32(0.1) + f32(f32(3) * f32(0.7)))
    exact = Fraction(float(f32(0.1))) + 3 * Fraction(float(f32(0.7)))
    assert b[3] == fused.round_to_f32(exact) and a[3] != b[3]
    assert abs(Fraction(float(b[3])) - exact) <= abs(Fraction(float(a[3])) - exact)
    # DM 0: both are firstDM itself
    assert a[0] == b[0] == f32(0.1)


def test_hazard_tables_are_what_they_claim():
    """Every row of a hazard table has a DM whose fused and unfused shifts
    differ; zero rows stay zero; shifts stay near the requested bound."""
    first, step = 3.3, 0.7
    assert fused.hazard_dms(first, step, 40).size >= 3
    t = fused.hazard_table(16, first, step, 40, 120, seed=1, zero_rows=(5,))
    assert t[5] == 0
    for r in range(16):
        if r != 5:
            assert fused.differing(t[r:r + 1], first, step, 40) >= 1, r
    assert fused.shift_table(t, first, step, 40).max() <= 121
    # with firstDM = 0 the two rules agree everywhere
    assert fused.differing(t, 0.0, step, 40) == 0
