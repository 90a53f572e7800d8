"""Pins the oracle's packed-input branch (inputBits 1, 2, 4;
include/Dedispersion.hpp:92-103 and :135-146).

The packed sums must equal the 8-bit sums (pinned in test_oracle.py by brute
force, closed forms and impulse responses) of the same sample values; the
rows are packed by an independent numpy packer (np.packbits for one bit),
so a wrong bit order, byte index, row pitch or beam stride fails.  The
reference's own test data (src/DedispersionTest.cpp:221-243) gives the
known answer (inputBits - 1) x live channels."""
import numpy as np
import pytest

import oracle
from dedispersion_amd import workloads
from tests.packing import pack_rows, random_values, unpacked_pitch

PAD = 128


def _single(obs, bits, padding, seed):
    orc8 = oracle.Oracle(obs, padding, threads=2)
    orcb = oracle.Oracle(obs, padding, threads=2, input_bits=bits)
    rng = np.random.default_rng(seed)
    rows = obs.nr_beams * obs.nr_channels
    n = orc8.o.nr_samples_per_dispersed_batch
    vals = random_values(rng, rows, n, bits)
    flat8 = pack_rows(vals, 8, orc8.in_pitch).ravel()
    packed = pack_rows(vals, bits, orcb.in_pitch).ravel()
    return orc8, orcb, flat8, packed


@pytest.mark.parametrize("bits", [1, 2, 4])
@pytest.mark.parametrize("padding", [PAD, 0])
def test_single_step_packed_equals_bytes(bits, padding):
    obs = workloads.tiny(channels=12, samples=45, dms=6, step=3.0, beams=3, sbeams=4).obs
    orc8, orcb, flat8, packed = _single(obs, bits, padding, seed=bits)
    zapped = np.zeros(oracle.pad(12, padding // 4), dtype=np.uint32)
    zapped[[2, 7]] = 1
    mapping = orc8.default_map(12)
    mapping[3, :5] = 2  # channels 0-4 of synthesized beam 3 from beam 2
    mapping[1, 6:] = 0
    want = orc8.dedispersion(flat8, zapped, mapping)
    got = orcb.dedispersion(packed, zapped, mapping)
    np.testing.assert_array_equal(got, want)
    # every beam contributes: the sums differ from those of beam 0's rows alone
    assert not np.array_equal(want[1], want[2])


@pytest.mark.parametrize("bits", [1, 2, 4])
def test_step_one_packed_equals_bytes(bits):
    obs = workloads.tiny(channels=16, samples=37, dms=4, step=1.0, subbands=4, coarse=5, coarse_step=7.0,
                         beams=2, sbeams=2).obs
    orc8 = oracle.Oracle(obs, PAD, threads=2)
    orcb = oracle.Oracle(obs, PAD, threads=2, input_bits=bits)
    rng = np.random.default_rng(10 + bits)
    rows = obs.nr_beams * obs.nr_channels
    vals = random_values(rng, rows, orc8.o.nr_samples_per_dispersed_batch_sub, bits)
    zapped = np.zeros(oracle.pad(16, PAD // 4), dtype=np.uint32)
    zapped[[1, 9]] = 1
    want = orc8.step_one(pack_rows(vals, 8, orc8.sub_in_pitch).ravel(), zapped)
    got = orcb.step_one(pack_rows(vals, bits, orcb.sub_in_pitch).ravel(), zapped)
    np.testing.assert_array_equal(got, want)
    assert got.shape == want.shape  # the intermediate keeps the 8-bit row pitch


@pytest.mark.parametrize("bits", [1, 2, 4])
def test_reference_test_data_known_answer(bits):
    """src/DedispersionTest.cpp:225: every sample = inputBits - 1."""
    w = workloads.plain_128()
    orc = oracle.Oracle(w.obs, PAD, threads=4, input_bits=bits)
    n = orc.o.nr_samples_per_dispersed_batch
    vals = np.full((128, n), bits - 1, np.uint8)
    zapped = np.zeros(oracle.pad(128, PAD // 4), dtype=np.uint32)
    zapped[w.zapped] = 1
    out = orc.dedispersion(pack_rows(vals, bits, orc.in_pitch).ravel(), zapped)
    S = w.obs.nr_samples_per_batch
    assert np.all(out[0, :, :S] == (bits - 1) * 126)


def test_top_bit_of_a_field_is_a_value_bit():
    """Fields are unsigned: a 2-bit field 0b11 is 3, a 4-bit 0b1111 is 15."""
    obs = workloads.tiny(channels=4, samples=16, dms=1, step=0.0).obs
    for bits in (2, 4):
        orc = oracle.Oracle(obs, PAD, threads=1, input_bits=bits)
        vals = np.full((4, orc.o.nr_samples_per_dispersed_batch), (1 << bits) - 1, np.uint8)
        out = orc.dedispersion(pack_rows(vals, bits, orc.in_pitch).ravel())
        assert np.all(out[0, 0, :16] == 4 * ((1 << bits) - 1))


def test_input_pitch_rule():
    """pad(ceil(S_disp / (8 / bits)), padding) bytes; equal to the reference's
    pad(S_disp / (8 / bits), padding) whenever 8 / bits divides S_disp."""
    lib = oracle.load()
    for s_disp in (1, 7, 8, 100, 1001, 4096, 4097):
        for bits in (1, 2, 4, 8):
            per = 8 // bits
            for padding in (0, 1, 32, 128):
                got = lib.orc_input_pitch(s_disp, bits, padding)
                assert got == unpacked_pitch(-(-s_disp // per), padding)
                if s_disp % per == 0:
                    assert got == unpacked_pitch(s_disp // per, padding)
                assert got * per >= s_disp
