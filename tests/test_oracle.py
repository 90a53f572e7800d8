"""Pins the CPU oracle (oracle/dd_oracle.c) before the device results are checked against it.

The reference ships no golden vectors (its only known answer is a constant
input, src/DedispersionTest.cpp:213, 260, 304), so each oracle function is
checked against what the mathematics fixes: closed forms, impulse responses,
invariants, and an independent numpy brute force on tiny shapes, chosen so
that a dropped term, a wrong sign or index, or a transposed layout fails.
"""
import numpy as np
import pytest

import oracle
from dedispersion_amd import workloads
from tests.conftest import host_threads

PAD = 128


def f32(x):
    return np.float32(x)


def shifts_float64(o):
    """Closed form of src/Shifts.cpp:26 in float64 (no rounding order)."""
    C = o.nr_channels
    f = o.min_freq + np.arange(C, dtype=np.float64) * o.channel_bandwidth
    fmax = o.min_freq + (C - 1) * o.channel_bandwidth
    s = 4148.808 * (1.0 / f**2 - 1.0 / fmax**2) * (o.nr_samples_per_batch // o.downsampling) / (
        o.nr_samples_per_batch * np.float64(np.float32(o.sampling_time)))
    s[-1] = 0.0
    return s


def shifts_numpy_f32(o):
    """src/Shifts.cpp:20-30 evaluated op by op in numpy float32."""
    C = o.nr_channels
    fmax = f32(o.min_freq) + f32(C - 1) * f32(o.channel_bandwidth)
    inv_high = f32(1.0) / (fmax * fmax)
    time_unit = f32(o.nr_samples_per_batch) * f32(o.sampling_time)
    samples = f32(o.nr_samples_per_batch // o.downsampling)
    out = np.zeros(C, dtype=np.float32)
    for ch in range(C - 1):
        f = f32(o.min_freq) + f32(ch) * f32(o.channel_bandwidth)
        inv = f32(1.0) / (f * f)
        d = f32(4148.808) * (inv - inv_high)
        d = d * samples
        out[ch] = d / time_unit
    return out


def int_shifts(table, first, step, dms):
    dm_value = (np.float32(first) + np.arange(dms, dtype=np.float32) * np.float32(step)).astype(np.float32)
    prod = (dm_value[None, :] * table[:, None].astype(np.float32)).astype(np.float32)
    return prod.astype(np.uint32)  # truncation toward zero


# ---------------------------------------------------------------- tables


@pytest.mark.parametrize("make", [workloads.plain_128, workloads.arts_single])
def test_shift_table_closed_form_and_invariants(make):
    w = make()
    orc = oracle.Oracle(w.obs, PAD)
    s = orc.shifts
    C = w.obs.nr_channels
    assert s.shape[0] == oracle.pad(C, PAD // 4)
    assert np.all(s[C:] == 0) and s[C - 1] == 0
    assert np.all(np.diff(s[:C]) <= 0), "delay must not grow with frequency"
    ref = shifts_float64(w.obs)
    # float32 evaluation of the closed form: a few ulp of the largest delay
    assert np.max(np.abs(s[:C] - ref)) <= 1e-5 * ref[0]
    # bit-exact against a numpy float32 evaluation in the same operation order
    np.testing.assert_array_equal(s[:C], shifts_numpy_f32(w.obs))


def test_shift_table_known_values():
    """Values quoted by the reference analysis for these observations."""
    s = oracle.Oracle(workloads.plain_128().obs, PAD).shifts
    assert abs(float(s[0]) - 3.306906) < 5e-6
    s = oracle.Oracle(workloads.arts_single().obs, PAD).shifts
    assert abs(float(s[0]) - 11.32728) < 5e-5


def test_step_two_table_matches_top_channel_delays():
    """Subband k is referenced to its top channel, so its delay equals the
    channel delay of channel (k+1)*cps-1 up to float rounding."""
    w = workloads.arts_two_step()
    orc = oracle.Oracle(w.obs, PAD)
    s, s2 = orc.shifts, orc.shifts2
    cps = w.obs.nr_channels // w.obs.nr_subbands
    tops = s[(np.arange(w.obs.nr_subbands) + 1) * cps - 1]
    assert s2[w.obs.nr_subbands - 1] == 0
    assert np.max(np.abs(s2[: w.obs.nr_subbands] - tops)) <= 1e-5 * s[0]
    assert np.all(np.diff(s2[: w.obs.nr_subbands]) <= 0)


def test_integer_shift_rule():
    """(unsigned)((first + dm*step) * table): unfused, truncating."""
    lib = oracle.load()
    rng = np.random.default_rng(0)
    for _ in range(2000):
        first = float(np.float32(rng.uniform(0, 50)))
        step = float(np.float32(rng.uniform(0, 3)))
        dm = int(rng.integers(0, 5000))
        t = float(np.float32(rng.uniform(0, 12)))
        expect = int(np.uint32((np.float32(first) + np.float32(dm) * np.float32(step)) * np.float32(t)))
        assert lib.orc_shift(first, step, dm, t) == expect
    # a fused multiply-add would round differently here: 0.1 + 3*0.7 -> 2.2f
    # (unfused) vs the exact product-sum; the product with 10 truncates to 22.
    assert lib.orc_shift(0.1, 0.7, 3, 10.0) == int(np.uint32((f32(0.1) + f32(3) * f32(0.7)) * f32(10.0)))


def test_dispersed_sizes():
    """src/DedispersionTest.cpp:129,145,146 at the measured workloads."""
    assert oracle.Oracle(workloads.plain_128().obs, PAD).o.nr_samples_per_dispersed_batch == 4306
    assert oracle.Oracle(workloads.arts_single().obs, PAD).o.nr_samples_per_dispersed_batch == 29214
    orc = oracle.Oracle(workloads.arts_two_step().obs, PAD)
    s1 = orc.o.nr_samples_per_batch_sub
    s1d = orc.o.nr_samples_per_dispersed_batch_sub
    assert abs(s1 - 24644) <= 2 and abs(s1d - 29212) <= 4
    # containment: every shifted read stays inside the dispersed row
    ish = int_shifts(orc.shifts[: orc.o.nr_channels], 0.0, 0.2, 2048)
    w = workloads.arts_single()
    o2 = oracle.Oracle(w.obs, PAD).o
    assert w.obs.nr_samples_per_batch + int(ish.max()) <= o2.nr_samples_per_dispersed_batch


# ------------------------------------------------------- dedispersion loops


def tiny_obs(**kw):
    return workloads.tiny(**kw).obs


def brute_single(obs, orc, inp, zapped, mapping):
    """Independent numpy formulation of the single-step sum."""
    C, D, S = obs.nr_channels, obs.nr_dms, obs.nr_samples_per_batch
    ish = int_shifts(orc.shifts[:C], obs.first_dm, obs.dm_step, D)
    x = inp.reshape(obs.nr_beams, C, orc.in_pitch).astype(np.float64)
    out = np.zeros((obs.nr_synthesized_beams, D, orc.out_pitch), dtype=np.float32)
    t = np.arange(S)
    for sb in range(obs.nr_synthesized_beams):
        for dm in range(D):
            acc = np.zeros(S)
            for ch in range(C):
                if zapped[ch]:
                    continue
                acc += x[mapping[sb, ch], ch, t + ish[ch, dm]]
            out[sb, dm, :S] = acc
    return out


def test_single_step_brute_force_tiny():
    obs = tiny_obs(channels=12, samples=40, dms=6, step=3.0, beams=2, sbeams=3)
    orc = oracle.Oracle(obs, PAD, threads=2)
    rng = np.random.default_rng(1)
    inp = rng.integers(0, 256, size=obs.nr_beams * obs.nr_channels * orc.in_pitch, dtype=np.uint8)
    zapped = np.zeros(oracle.pad(12, PAD // 4), dtype=np.uint32)
    zapped[[3, 7]] = 1
    mapping = orc.default_map(12)
    mapping[2, :6] = 0  # synthesized beam 2 takes half of its channels from beam 0
    got = orc.dedispersion(inp, zapped, mapping)
    np.testing.assert_array_equal(got, brute_single(obs, orc, inp, zapped, mapping))


def test_single_step_constant_input():
    """Known answer of the reference's test (src/DedispersionTest.cpp:213):
    constant c everywhere gives c * live channels."""
    w = workloads.plain_128()
    orc = oracle.Oracle(w.obs, PAD, threads=host_threads())
    inp = np.full(w.obs.nr_channels * orc.in_pitch, 10, dtype=np.uint8)
    zapped = np.zeros(oracle.pad(128, PAD // 4), dtype=np.uint32)
    zapped[w.zapped] = 1
    out = orc.dedispersion(inp, zapped)
    assert np.all(out[0, :, : w.obs.nr_samples_per_batch] == 10 * 126)
    assert np.all(out[0, :, w.obs.nr_samples_per_batch:] == 0)


def test_single_step_impulse_response():
    """A single non-zero input (ch0, t0) appears at t0 - shift(ch0, dm) in
    every DM row and nowhere else: pins the sign and row of the shift."""
    w = workloads.plain_128()
    orc = oracle.Oracle(w.obs, PAD, threads=host_threads())
    ch0, t0 = 5, 3000
    inp = np.zeros(128 * orc.in_pitch, dtype=np.uint8)
    inp[ch0 * orc.in_pitch + t0] = 7
    out = orc.dedispersion(inp)
    ish = int_shifts(orc.shifts[:128], 0.0, 0.5, 128)
    for dm in range(128):
        row = out[0, dm, :4096]
        t = t0 - int(ish[ch0, dm])
        assert row[t] == 7 and row.sum() == 7, dm


def test_single_step_dm0_and_linearity_and_zapped_invariance():
    w = workloads.plain_128()
    orc = oracle.Oracle(w.obs, PAD, threads=host_threads())
    rng = np.random.default_rng(2)
    a = rng.integers(0, 128, size=128 * orc.in_pitch, dtype=np.uint8)
    b = rng.integers(0, 128, size=128 * orc.in_pitch, dtype=np.uint8)
    zapped = np.zeros(128, dtype=np.uint32)
    zapped[[17, 90]] = 1
    oa, ob, oab = orc.dedispersion(a, zapped), orc.dedispersion(b, zapped), orc.dedispersion(a + b, zapped)
    np.testing.assert_array_equal(oa + ob, oab)
    # DM 0 has shift 0 in every channel: the plain sum over live channels
    x = a.reshape(128, orc.in_pitch)[:, :4096].astype(np.float64)
    x[[17, 90]] = 0
    np.testing.assert_array_equal(oa[0, 0, :4096], x.sum(axis=0))
    # data in zapped channels never reaches the output
    c = a.copy().reshape(128, orc.in_pitch)
    c[[17, 90]] = 255
    np.testing.assert_array_equal(orc.dedispersion(c.ravel(), zapped), oa)


def brute_step_one(obs, orc, inp, zapped):
    C, Nsb, Dc = obs.nr_channels, obs.nr_subbands, obs.nr_dms_subbanding
    cps = C // Nsb
    S1 = orc.o.nr_samples_per_batch_sub
    s = orc.shifts[:C]
    out = np.zeros((obs.nr_beams, Dc, Nsb, orc.sub_out_pitch), dtype=np.float32)
    x = inp.reshape(obs.nr_beams, C, orc.sub_in_pitch).astype(np.float64)
    t = np.arange(S1)
    for b in range(obs.nr_beams):
        for sb in range(Nsb):
            rel = (s[sb * cps:(sb + 1) * cps] - s[(sb + 1) * cps - 1]).astype(np.float32)
            ish = int_shifts(rel, obs.first_dm_subbanding, obs.dm_step_subbanding, Dc)
            for c in range(Dc):
                acc = np.zeros(S1)
                for k in range(cps):
                    ch = sb * cps + k
                    if not zapped[ch]:
                        acc += x[b, ch, t + ish[k, c]]
                out[b, c, sb, :S1] = acc
    return out


def brute_step_two(obs, orc, mid, mapping):
    """Sequential float32 sum over subbands, in order."""
    Nsb, Dc, D, S = obs.nr_subbands, obs.nr_dms_subbanding, obs.nr_dms, obs.nr_samples_per_batch
    ish = int_shifts(orc.shifts2[:Nsb], obs.first_dm, obs.dm_step, D)
    x = mid.reshape(obs.nr_beams, Dc, Nsb, orc.sub_out_pitch)
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


def test_step_one_brute_force_and_closed_forms():
    obs = tiny_obs(channels=16, samples=48, dms=4, step=1.0, subbands=4, coarse=5, coarse_step=7.0, beams=2,
                   sbeams=2)
    orc = oracle.Oracle(obs, PAD, threads=2)
    rng = np.random.default_rng(3)
    inp = rng.integers(0, 256, size=2 * 16 * orc.sub_in_pitch, dtype=np.uint8)
    zapped = np.zeros(oracle.pad(16, PAD // 4), dtype=np.uint32)
    zapped[[1, 9, 15]] = 1  # 15 is the top channel of subband 3: its shift still anchors the subband
    got = orc.step_one(inp, zapped)
    np.testing.assert_array_equal(got, brute_step_one(obs, orc, inp, zapped))
    # constant input: c * live channels of the subband
    const = orc.step_one(np.full_like(inp, 3), zapped)
    live = np.array([4 - 1, 4, 4 - 1, 4 - 1])  # zapped 1 in sb0, 9 in sb2, 15 in sb3
    S1 = orc.o.nr_samples_per_batch_sub
    assert np.all(const[:, :, :, :S1] == 3 * live[None, None, :, None])
    # coarse DM 0: plain per-subband channel sum
    x = inp.reshape(2, 16, orc.sub_in_pitch)[:, :, :S1].astype(np.float64) * (1 - zapped[:16])[None, :, None]
    np.testing.assert_array_equal(got[:, 0, :, :S1], x.reshape(2, 4, 4, S1).sum(axis=2))


def test_step_two_brute_force_general_floats():
    """Step two sums subbands in order; with non-integer floats the
    sequential float32 order is what makes the result reproducible."""
    obs = tiny_obs(channels=16, samples=40, dms=6, step=1.5, subbands=4, coarse=3, coarse_step=5.0, beams=2,
                   sbeams=3)
    orc = oracle.Oracle(obs, PAD, threads=2)
    rng = np.random.default_rng(4)
    mid = rng.standard_normal(2 * 3 * 4 * orc.sub_out_pitch).astype(np.float32) * np.float32(1000.0)
    mapping = orc.default_map(4)
    mapping[2, :4] = [1, 0, 1, 0]
    got = orc.step_two(mid, mapping)
    np.testing.assert_array_equal(got, brute_step_two(obs, orc, mid, mapping))
    # constant subbanded input: c * Nsb (src/DedispersionTest.cpp:304)
    const = orc.step_two(np.full_like(mid, 10.0), mapping)
    assert np.all(const[:, :, :40] == 40.0)


def test_step_two_impulse():
    obs = tiny_obs(channels=16, samples=40, dms=6, step=1.5, subbands=4, coarse=2, coarse_step=5.0)
    orc = oracle.Oracle(obs, PAD, threads=1)
    P = orc.sub_out_pitch
    mid = np.zeros(2 * 4 * P, dtype=np.float32)
    sb0, t0, c0 = 1, 30, 1
    mid[(c0 * 4 + sb0) * P + t0] = 2.5
    out = orc.step_two(mid)
    ish = int_shifts(orc.shifts2[:4], 0.0, 1.5, 6)
    for f in range(6):
        row = out[0, c0 * 6 + f, :40]
        t = t0 - int(ish[sb0, f])
        if 0 <= t < 40:
            assert row[t] == 2.5 and row.sum() == 2.5
        else:
            assert row.sum() == 0
    assert np.all(out[0, :6] == 0)


def test_two_step_equals_single_step_at_zero_dms():
    """With one coarse and one fine DM, both 0, two-step equals single step."""
    obs = tiny_obs(channels=16, samples=40, dms=1, step=0.0, subbands=4, coarse=1, coarse_step=0.0)
    orc = oracle.Oracle(obs, PAD, threads=1)
    rng = np.random.default_rng(5)
    inp = rng.integers(0, 256, size=16 * orc.sub_in_pitch, dtype=np.uint8)
    mid = orc.step_one(inp)
    two = orc.step_two(mid.ravel())
    x = inp.reshape(16, orc.sub_in_pitch)[:, :40].astype(np.float64).sum(axis=0)
    np.testing.assert_array_equal(two[0, 0, :40], x)


def test_row_parallel_driver_is_bit_identical():
    w = workloads.plain_128()
    one = oracle.Oracle(w.obs, PAD, threads=1)
    many = oracle.Oracle(w.obs, PAD, threads=host_threads())
    rng = np.random.default_rng(6)
    inp = rng.integers(0, 256, size=128 * one.in_pitch, dtype=np.uint8)
    rows = (0, 16)
    np.testing.assert_array_equal(one.dedispersion(inp, rows=rows)[0, :16], many.dedispersion(inp, rows=rows)[0, :16])
