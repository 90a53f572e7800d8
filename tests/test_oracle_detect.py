"""Pins the oracle's boxcar peak search (orc_detect_peaks), the step after
the path, which has no counterpart in the reference: closed forms for an
impulse and a boxcar pulse, a constant row, ties, and an independent numpy
formulation (cumulative sums in float64) on random integer rows."""
import numpy as np
import pytest

import oracle


def numpy_peaks(rows, S, nw):
    out = []
    for y in rows:
        y = np.trunc(y[:S]).astype(np.float64)
        mean = y.sum() / S
        var = (y * y).sum() / S - mean * mean
        sigma = np.sqrt(max(var, 0.0))
        c = np.concatenate([[0.0], np.cumsum(y)])
        best = None
        for j in range(nw):
            w = 1 << j
            if w > S:
                break
            b = c[w:] - c[:-w]
            t = int(np.argmax(b))
            snr = (b[t] - w * mean) / (sigma * np.sqrt(w)) if sigma > 0 else 0.0
            if best is None or snr > best[0]:
                best = (snr, t, w, mean, sigma)
        out.append(best)
    return out


def test_random_integer_rows_match_numpy():
    rng = np.random.default_rng(3)
    S, pitch, nw = 1000, 1024, 7
    rows = rng.integers(0, 4000, size=(20, pitch)).astype(np.float32)
    rows[5, 300:308] += 3000  # a pulse of width 8
    got = oracle.detect_peaks(rows, S, nw)
    want = numpy_peaks(rows, S, nw)
    for i, (snr, t, w, mean, sigma) in enumerate(want):
        assert got["sample"][i] == t and got["width"][i] == w, i
        np.testing.assert_allclose(got["snr"][i], snr, rtol=1e-6)
        np.testing.assert_allclose(got["mean"][i], mean, rtol=1e-7)
        np.testing.assert_allclose(got["sigma"][i], sigma, rtol=1e-6)
    assert got["width"][5] == 8 and got["sample"][5] == 300


def test_impulse_closed_form():
    S, h, t0 = 512, 1000.0, 77
    row = np.zeros((1, S), np.float32)
    row[0, t0] = h
    got = oracle.detect_peaks(row, S, 5)
    mean = h / S
    sigma = np.sqrt(h * h / S - mean * mean)
    assert got["width"][0] == 1 and got["sample"][0] == t0
    np.testing.assert_allclose(got["snr"][0], (h - mean) / sigma, rtol=1e-6)


@pytest.mark.parametrize("W", [2, 8, 32])
def test_boxcar_pulse_picks_its_width(W):
    S, t0 = 4096, 1000
    row = np.full((1, S), 10.0, np.float32)
    row[0, t0:t0 + W] += 50.0
    got = oracle.detect_peaks(row, S, 7)
    assert got["width"][0] == W and got["sample"][0] == t0


def test_constant_row_and_ties_and_fractions():
    S = 300
    rows = np.full((3, S), 7.0, np.float32)
    rows[1, 40] = rows[1, 200] = 50.0          # two equal impulses: the first wins
    rows[2] = 0.75                              # truncated to 0: a constant row
    got = oracle.detect_peaks(rows, S, 3)
    assert got["snr"][0] == 0 and got["sigma"][0] == 0 and got["width"][0] == 1 and got["sample"][0] == 0
    assert got["sample"][1] == 40 and got["width"][1] == 1
    assert got["mean"][2] == 0 and got["snr"][2] == 0


def test_widths_longer_than_the_row_are_skipped():
    rows = np.arange(6, dtype=np.float32)[None, :]
    got = oracle.detect_peaks(rows, 6, 7)  # widths 1, 2, 4 only
    assert got["width"][0] in (1, 2, 4)
