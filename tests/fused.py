"""The contraction hazard of the shift rule, for the parity tests.

The reference computes a shift as (unsigned)((firstDM + dm * DMStep) *
table[row]) in separately rounded float operations
(include/Dedispersion.hpp:81, 124, 168).  A compiler that contracts
dm * DMStep + firstDM into one fused multiply-add rounds once instead of
twice, so the DM value can differ by an ulp and a truncated shift by one
sample.  At firstDM = 0 both forms agree, so only tests with firstDM != 0 can
tell them apart.

This module evaluates both forms exactly on the host (the fused sum is
rounded from an exact rational) and builds shift tables whose rows are chosen
so that the two forms give different integer shifts at some DM: kernels fed
these tables with that firstDM and step are checked where the hazard is.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np

f32 = np.float32


def round_to_f32(x: Fraction) -> np.float32:
    """x rounded to the nearest float32 (ties to even), from the exact value."""
    if x == 0:
        return f32(0.0)
    sign = -1 if x < 0 else 1
    a = abs(x)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    scale = Fraction(2) ** (23 - e)  # a * scale in [2^23, 2^24)
    m = a * scale
    q, r = divmod(m.numerator, m.denominator)
    rem = Fraction(r, m.denominator)
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and q % 2 == 1):
        q += 1
    return f32(float(sign * Fraction(q) / scale))


def dm_values(first: float, step: float, dms: int, fused: bool) -> np.ndarray:
    """firstDM + dm * DMStep for dm < dms: two roundings (the reference) or
    one (a fused multiply-add)."""
    a, b = f32(first), f32(step)
    if fused:
        fa, fb = Fraction(float(a)), Fraction(float(b))
        return np.array([round_to_f32(fa + d * fb) for d in range(dms)], dtype=np.float32)
    d = np.arange(dms, dtype=np.float32)
    return (a + (d * b).astype(np.float32)).astype(np.float32)


def shift_table(table, first: float, step: float, dms: int, fused: bool = False) -> np.ndarray:
    """Integer shifts [rows][dms]: (unsigned)(dm value * table[row])."""
    v = dm_values(first, step, dms, fused)
    t = np.asarray(table, dtype=np.float32)
    return (v[None, :] * t[:, None]).astype(np.float32).astype(np.uint32)


def hazard_dms(first: float, step: float, dms: int) -> np.ndarray:
    """DM indices whose fused and unfused DM values differ."""
    return np.nonzero(dm_values(first, step, dms, False) != dm_values(first, step, dms, True))[0]


def hazard_value(first: float, step: float, dm: int, target: int):
    """A float32 delay t near target / dmValue for which the unfused and the
    fused shift of `dm` differ (None when no t within a few ulps does)."""
    a = dm_values(first, step, dm + 1, False)[dm]
    b = dm_values(first, step, dm + 1, True)[dm]
    if a == b:
        return None
    for base in (f32(target / float(a)), f32(target / float(b))):
        t = base
        for _ in range(8):
            if int(np.uint32(f32(a * t))) != int(np.uint32(f32(b * t))):
                return t
            t = np.nextafter(t, f32(np.inf), dtype=np.float32)
        t = base
        for _ in range(8):
            if int(np.uint32(f32(a * t))) != int(np.uint32(f32(b * t))):
                return t
            t = np.nextafter(t, f32(0.0), dtype=np.float32)
    return None


def hazard_table(rows: int, first: float, step: float, dms: int, max_shift: int, seed: int = 0,
                 zero_rows=()) -> np.ndarray:
    """A float32 delay table of `rows` entries, each (except `zero_rows`,
    which are 0) chosen so that its unfused and fused shifts differ at some
    DM below `dms`, with shifts up to about max_shift samples."""
    rng = np.random.default_rng(seed)
    dm_set = hazard_dms(first, step, dms)
    assert dm_set.size, f"no DM of {first} + dm * {step} (dm < {dms}) rounds differently when fused"
    out = np.zeros(rows, dtype=np.float32)
    zero = set(int(r) for r in zero_rows)
    for r in range(rows):
        if r in zero:
            continue
        for _ in range(1000):
            dm = int(dm_set[rng.integers(dm_set.size)])
            top = float(dm_values(first, step, dms, False)[-1])
            dmv = float(dm_values(first, step, dm + 1, False)[dm])
            # the row's largest shift (at the last DM) stays near max_shift
            hi = max(2, int(max_shift * dmv / max(top, 1e-6)))
            t = hazard_value(first, step, dm, int(rng.integers(1, hi)))
            if t is not None:
                out[r] = t
                break
        else:
            raise AssertionError(f"no hazard value for row {r}")
    return out


def differing(table, first: float, step: float, dms: int, rows=None) -> int:
    """(row, DM) pairs of `table` (restricted to `rows`) whose fused and
    unfused shifts differ."""
    t = np.asarray(table, dtype=np.float32)
    if rows is not None:
        t = t[np.asarray(list(rows), dtype=np.int64)]
    return int(np.count_nonzero(shift_table(t, first, step, dms) != shift_table(t, first, step, dms, fused=True)))
