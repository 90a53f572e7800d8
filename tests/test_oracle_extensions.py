"""Pins the oracle's float-input branch, its absolute step-one delays and the
downsampled subband layouts before device results are checked against them.

Float input (the reference's templates with I = float,
include/Dedispersion.hpp:59-60): the sum is the reference's sequential
float32 sum over the live channels in order, so an independent numpy
restatement that adds float32 rows in that order must agree bit for bit,
while the same sum in reverse channel order must not (the check sees the
order).  Integer-valued float input must give the 8-bit path's result.

Absolute step one (the OpenCL kernel's delay, include/Dedispersion.hpp:
595-597): numpy brute force, an impulse response, and equality with the
relative delays when every subband's top channel has delay 0.

Downsampling in the subband path (:119, :148, :163, :169): step one writes
S1 / downsampling samples into rows of pad(S1); step two reads rows of
pad(S1 / downsampling); both against numpy brute force.
"""
import numpy as np
import pytest

import oracle
from dedispersion_amd import workloads
from tests.test_oracle import int_shifts

PAD = 128


def tiny_obs(**kw):
    return workloads.tiny(**kw).obs


def brute_single_f32(obs, orc, inp, zapped, mapping, reverse=False):
    """float32 rows added in channel order (or reversed) into a float32 sum."""
    C, D, S = obs.nr_channels, obs.nr_dms, obs.nr_samples_per_batch // obs.downsampling
    ish = int_shifts(orc.shifts[:C], obs.first_dm, obs.dm_step, D)
    x = inp.reshape(obs.nr_beams, C, orc.in_pitch)
    out = np.zeros((obs.nr_synthesized_beams, D, orc.out_pitch), dtype=np.float32)
    t = np.arange(S)
    order = range(C - 1, -1, -1) if reverse else range(C)
    for sb in range(obs.nr_synthesized_beams):
        for dm in range(D):
            acc = np.zeros(S, dtype=np.float32)
            for ch in order:
                if zapped[ch]:
                    continue
                acc = (acc + x[mapping[sb, ch], ch, t + ish[ch, dm]]).astype(np.float32)
            out[sb, dm, :S] = acc
    return out


def brute_step_one(obs, orc, inp, zapped, absolute=False, dtype=np.float64):
    """Step one over any input type; sums in float32 channel order when dtype
    is float32, exactly otherwise.  Rows of pad(S1), S1 / downsampling
    samples each."""
    C, Nsb, Dc = obs.nr_channels, obs.nr_subbands, obs.nr_dms_subbanding
    cps = C // Nsb
    n = orc.o.nr_samples_per_batch_sub // obs.downsampling
    s = orc.shifts[:C]
    out = np.zeros((obs.nr_beams, Dc, Nsb, orc.sub_out_pitch), dtype=np.float32)
    x = inp.reshape(obs.nr_beams, C, orc.sub_in_pitch).astype(dtype)
    t = np.arange(n)
    for b in range(obs.nr_beams):
        for sb in range(Nsb):
            top = s[(sb + 1) * cps - 1]
            delay = s[sb * cps:(sb + 1) * cps] if absolute else (s[sb * cps:(sb + 1) * cps] - top).astype(np.float32)
            ish = int_shifts(delay, obs.first_dm_subbanding, obs.dm_step_subbanding, Dc)
            for c in range(Dc):
                acc = np.zeros(n, dtype=dtype)
                for k in range(cps):
                    ch = sb * cps + k
                    if not zapped[ch]:
                        acc = (acc + x[b, ch, t + ish[k, c]]).astype(dtype)
                out[b, c, sb, :n] = acc
    return out


def brute_step_two(obs, orc, mid, mapping):
    """Rows of pad(S1 / downsampling), S / downsampling output samples."""
    Nsb, Dc, D = obs.nr_subbands, obs.nr_dms_subbanding, obs.nr_dms
    S = obs.nr_samples_per_batch // obs.downsampling
    ish = int_shifts(orc.shifts2[:Nsb], obs.first_dm, obs.dm_step, D)
    x = mid.reshape(obs.nr_beams, Dc, Nsb, orc.sub_two_in_pitch)
    out = np.zeros((obs.nr_synthesized_beams, Dc * D, orc.out_pitch), dtype=np.float32)
    t = np.arange(S)
    for sbm in range(obs.nr_synthesized_beams):
        for c in range(Dc):
            for f in range(D):
                acc = np.zeros(S, dtype=np.float32)
                for sb in range(Nsb):
                    acc = (acc + x[mapping[sbm, sb], c, sb, t + ish[sb, f]]).astype(np.float32)
                out[sbm, c * D + f, :S] = acc
    return out


# ------------------------------------------------------------ float input


def test_float_single_step_brute_force_and_order():
    obs = tiny_obs(channels=24, samples=60, dms=7, step=2.5, beams=2, sbeams=3, first=1.3)
    orc = oracle.Oracle(obs, PAD, threads=2, input_bits=32)
    assert orc.in_pitch == oracle.pad(orc.o.nr_samples_per_dispersed_batch, PAD // 4)
    rng = np.random.default_rng(11)
    inp = (rng.standard_normal(2 * 24 * orc.in_pitch) * 1000.0).astype(np.float32)
    zapped = np.zeros(oracle.pad(24, PAD // 4), dtype=np.uint32)
    zapped[[2, 13]] = 1
    mapping = orc.default_map(24)
    mapping[2, 5:15] = 0
    got = orc.dedispersion(inp, zapped, mapping)
    np.testing.assert_array_equal(got, brute_single_f32(obs, orc, inp, zapped, mapping))
    # the order matters for general floats, and the check sees it
    assert np.any(got != brute_single_f32(obs, orc, inp, zapped, mapping, reverse=True))


def test_float_single_step_integer_values_equal_the_byte_path():
    w = workloads.plain_128()
    f = oracle.Oracle(w.obs, PAD, threads=4, input_bits=32)
    b = oracle.Oracle(w.obs, PAD, threads=4)
    rng = np.random.default_rng(12)
    inp8 = rng.integers(0, 256, size=(128, b.in_pitch), dtype=np.uint8)
    n = f.o.nr_samples_per_dispersed_batch
    inpf = np.zeros((128, f.in_pitch), dtype=np.float32)
    inpf[:, :n] = inp8[:, :n]
    zapped = np.zeros(128, dtype=np.uint32)
    zapped[w.zapped] = 1
    np.testing.assert_array_equal(f.dedispersion(inpf.ravel(), zapped), b.dedispersion(inp8.ravel(), zapped))


def test_float_single_step_constant_input():
    """Constant 0.1f over 126 live channels: the sequential float32 sum of
    0.1f, 126 times (not 126 * 0.1f), in every sample."""
    w = workloads.plain_128()
    orc = oracle.Oracle(w.obs, PAD, threads=4, input_bits=32)
    inp = np.full(128 * orc.in_pitch, 0.1, dtype=np.float32)
    zapped = np.zeros(128, dtype=np.uint32)
    zapped[w.zapped] = 1
    want = np.float32(0.0)
    for _ in range(126):
        want = np.float32(want + np.float32(0.1))
    assert want != np.float32(126 * np.float32(0.1))
    out = orc.dedispersion(inp, zapped)
    assert np.all(out[0, :, :4096] == want)


def test_float_step_one_brute_force():
    obs = tiny_obs(channels=16, samples=48, dms=4, step=1.0, subbands=4, coarse=5, coarse_step=7.0, beams=2,
                   sbeams=2, coarse_first=2.2)
    orc = oracle.Oracle(obs, PAD, threads=2, input_bits=32)
    rng = np.random.default_rng(13)
    inp = (rng.standard_normal(2 * 16 * orc.sub_in_pitch) * 50.0).astype(np.float32)
    zapped = np.zeros(oracle.pad(16, PAD // 4), dtype=np.uint32)
    zapped[[0, 9]] = 1
    np.testing.assert_array_equal(orc.step_one(inp, zapped), brute_step_one(obs, orc, inp, zapped, dtype=np.float32))


# ------------------------------------------------------- absolute step one


def test_absolute_step_one_brute_force_and_impulse():
    obs = tiny_obs(channels=16, samples=48, dms=4, step=1.0, subbands=4, coarse=5, coarse_step=7.0, beams=2,
                   sbeams=2, coarse_first=1.5)
    orc = oracle.Oracle(obs, PAD, threads=2, step_one_absolute=True)
    rel = oracle.Oracle(obs, PAD, threads=2)
    rng = np.random.default_rng(14)
    inp = rng.integers(0, 256, size=2 * 16 * orc.sub_in_pitch, dtype=np.uint8)
    zapped = np.zeros(oracle.pad(16, PAD // 4), dtype=np.uint32)
    zapped[[3]] = 1
    got = orc.step_one(inp, zapped)
    np.testing.assert_array_equal(got, brute_step_one(obs, orc, inp, zapped, absolute=True))
    assert np.any(got != rel.step_one(inp, zapped))
    # impulse in channel 5 at t0 lands at t0 - shift(ch 5, c) with the channel's own delay
    imp = np.zeros_like(inp)
    t0 = 50
    imp[5 * orc.sub_in_pitch + t0] = 9
    out = orc.step_one(imp)
    S1 = orc.o.nr_samples_per_batch_sub
    for c in range(5):
        sh = orc.shift(1.5, 7.0, c, float(orc.shifts[5]))
        row = out[0, c, 1, :S1]
        assert 0 <= t0 - sh < S1 and row[t0 - sh] == 9 and row.sum() == 9, c
        assert out[0, c, [0, 2, 3]].sum() == 0


def test_absolute_equals_relative_when_top_channels_have_no_delay():
    """With every subband's top channel at delay 0, the relative delay of a
    channel is its own: both forms give the same step one."""
    obs = tiny_obs(channels=16, samples=48, dms=4, step=1.0, subbands=4, coarse=5, coarse_step=7.0)
    a = oracle.Oracle(obs, PAD, threads=2, step_one_absolute=True)
    r = oracle.Oracle(obs, PAD, threads=2)
    table = a.shifts.copy()
    table[[3, 7, 11, 15]] = 0.0
    rng = np.random.default_rng(15)
    inp = rng.integers(0, 256, size=16 * a.sub_in_pitch, dtype=np.uint8)
    np.testing.assert_array_equal(a.step_one(inp, shifts=table), r.step_one(inp, shifts=table))


# ------------------------------------------------ downsampling, subband path


@pytest.mark.parametrize("ds", [2, 3])
def test_downsampled_subband_layouts(ds):
    obs = tiny_obs(channels=16, samples=60, dms=4, step=1.0, subbands=4, coarse=3, coarse_step=5.0, beams=1,
                   sbeams=2)
    obs.downsampling = ds
    orc0 = oracle.Oracle(obs, 0)
    # S1 long enough for step two's reads of S / ds + shift (the derived S1
    # counts the step-two delay in downsampled samples on top of S)
    obs.nr_samples_per_batch_subbanding = ds * (60 // ds + int(orc0.shifts2[0] * 3.0) + 2)
    orc = oracle.Oracle(obs, 32, threads=2)
    assert orc.sub_out_pitch == oracle.pad(orc.o.nr_samples_per_batch_sub, 8)
    assert orc.sub_two_in_pitch == oracle.pad(orc.o.nr_samples_per_batch_sub // ds, 8)
    rng = np.random.default_rng(16)
    inp = rng.integers(0, 256, size=16 * orc.sub_in_pitch, dtype=np.uint8)
    zapped = np.zeros(16, dtype=np.uint32)
    zapped[6] = 1
    mid = orc.step_one(inp, zapped)
    np.testing.assert_array_equal(mid, brute_step_one(obs, orc, inp, zapped))
    # only S1 / ds samples of each pad(S1) row are written
    n = orc.o.nr_samples_per_batch_sub // ds
    assert np.all(mid[..., n:] == 0)
    x = (rng.standard_normal(obs.nr_dms_subbanding * 4 * orc.sub_two_in_pitch) * 100.0).astype(np.float32)
    mapping = orc.default_map(4)
    np.testing.assert_array_equal(orc.step_two(x, mapping), brute_step_two(obs, orc, x, mapping))
