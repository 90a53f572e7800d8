"""Parity of the HIP kernels with the CPU oracle, through the C ABI.

uint8 paths (single step, step one) must be bit-exact: every partial sum is an
integer below 2^24.  Step two must be bit-exact too, for any float input,
because it adds the subbands in the reference's order.
"""
import itertools

import numpy as np
import pytest

import oracle
from dedispersion_amd import DD_SINGLE_STEP, DD_STEP_ONE, DD_STEP_TWO, Conf, DDError, Plan, workloads
from tests.conftest import host_threads
from tests.gpu_helpers import compare_rows, make_input, single_step, stream_of, zapped_flags
from tests.test_oracle import int_shifts

pytestmark = pytest.mark.gpu


# ------------------------------------------------------------------ tables


@pytest.mark.parametrize("make", [workloads.plain_128, workloads.arts_single, workloads.arts_two_step])
def test_device_tables_bit_exact(torch_cuda, make):
    w = make()
    plan = Plan(w.obs, 128)
    orc = oracle.Oracle(w.obs, 128)
    s, s2 = plan.get_shifts()
    np.testing.assert_array_equal(s.view(np.uint32), orc.shifts.view(np.uint32))
    np.testing.assert_array_equal(s2.view(np.uint32), orc.shifts2.view(np.uint32))
    C, Nsb = w.obs.nr_channels, w.obs.nr_subbands
    np.testing.assert_array_equal(plan.get_integer_shifts(DD_SINGLE_STEP),
                                  int_shifts(orc.shifts[:C], w.obs.first_dm, w.obs.dm_step, w.obs.nr_dms))
    z = plan.sizes
    assert z.nr_dispersed_samples == orc.o.nr_samples_per_dispersed_batch
    assert z.in_pitch == orc.in_pitch and z.out_pitch == orc.out_pitch
    if w.obs.nr_dms_subbanding:
        cps = C // Nsb
        rel = (orc.shifts[:C] - np.repeat(orc.shifts[(np.arange(Nsb) + 1) * cps - 1], cps)).astype(np.float32)
        np.testing.assert_array_equal(
            plan.get_integer_shifts(DD_STEP_ONE),
            int_shifts(rel, w.obs.first_dm_subbanding, w.obs.dm_step_subbanding, w.obs.nr_dms_subbanding))
        np.testing.assert_array_equal(plan.get_integer_shifts(DD_STEP_TWO),
                                      int_shifts(orc.shifts2[:Nsb], w.obs.first_dm, w.obs.dm_step, w.obs.nr_dms))
        assert z.nr_samples_subbanding == orc.o.nr_samples_per_batch_sub
        assert z.nr_dispersed_samples_subbanding == orc.o.nr_samples_per_dispersed_batch_sub
        assert z.sub_in_pitch == orc.sub_in_pitch and z.sub_out_pitch == orc.sub_out_pitch


# -------------------------------------------------------------- single step


def test_single_step_plain_128(torch_cuda):
    """128 channels, 4096 samples, 128 DMs, channels 17 and 90 zapped."""
    w = workloads.plain_128()
    plan, inp, out, orc, flags = single_step(torch_cuda, w)
    want = orc.dedispersion(inp, flags)
    compare_rows(out.reshape(want.shape), want, w.obs.nr_samples_per_batch, "plain-128")


SWEEP_CONFS = [
    Conf(64, t1, i0, i1, u, loc)
    for (t1, i0, i1, u, loc) in itertools.product((1, 2, 4, 8), (4, 8, 16), (1, 2, 4, 8), (1, 3, 8), (1, 2, 0))
    if t1 * i1 <= 64
]
# channel-block patterns (local 3): 32-DM tiles
PATTERN_CONFS = [Conf(64, t1, i0, i1, u, 3) for (t1, i1) in ((8, 4), (4, 8)) for i0 in (8, 16) for u in (1, 3, 8)]
SWEEP_CONFS += PATTERN_CONFS


@pytest.fixture(scope="module")
def ragged_case(torch_cuda):
    """Odd sizes everywhere: 3 beams, 4 synthesized beams with a mixed
    mapping, 300 channels (> 257, so the 16-bit accumulators flush), 1000
    samples, 37 DMs, unpadded rows (padding 0)."""
    w = workloads.tiny(channels=300, samples=1000, dms=37, step=0.7, beams=3, sbeams=4, min_freq=1300.0,
                       bandwidth=0.5, zapped=(0, 5, 150, 299))
    plan = Plan(w.obs, 0)
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    mapping = (np.arange(4)[:, None] % 3 + np.zeros((1, 300), dtype=int)).astype(np.uint32)
    mapping[3, ::2] = 2
    plan.set_beam_mapping(single=mapping)
    z = plan.sizes
    d_in = make_input(torch_cuda, plan, w, z.in_pitch, z.nr_dispersed_samples, 11)
    orc = oracle.Oracle(w.obs, 0, threads=host_threads())
    want = orc.dedispersion(d_in.cpu().numpy(), flags, mapping)
    return w, plan, d_in, want


@pytest.mark.parametrize("conf", SWEEP_CONFS, ids=lambda c: f"t{c.threads_d1}i{c.items_d0}x{c.items_d1}u{c.unroll}l{c.local}")
def test_single_step_conf_sweep(torch_cuda, ragged_case, conf):
    w, plan, d_in, want = ragged_case
    plan.set_conf(DD_SINGLE_STEP, conf)
    z = plan.sizes
    d_out = torch_cuda.full((z.output_elements,), -1.0, dtype=torch_cuda.float32, device="cuda")
    plan.dedisperse(d_in.data_ptr(), d_out.data_ptr(), stream=stream_of(torch_cuda))
    got = d_out.cpu().numpy().reshape(want.shape)
    compare_rows(got, want, w.obs.nr_samples_per_batch, f"conf {conf}")
    # padding of the output rows is never written
    assert np.all(got[..., w.obs.nr_samples_per_batch:] == -1.0)


@pytest.mark.parametrize("conf", PATTERN_CONFS, ids=lambda c: f"t{c.threads_d1}i{c.items_d0}x{c.items_d1}u{c.unroll}")
def test_single_step_patterns_partial_block(torch_cuda, conf):
    """Channel-block patterns with 253 live channels (a last block of one
    channel), 70 DMs (a partial 32-DM tile) and 2100 samples (a partial
    sample tile), padding 128."""
    w = workloads.tiny(channels=257, samples=2100, dms=70, step=0.2, min_freq=1300.0, bandwidth=0.5,
                       zapped=(3, 4, 5, 6, 7, 100), seed=17)
    plan, inp, out, orc, flags = single_step(torch_cuda, w, padding=128, conf=conf, seed=17)
    want = orc.dedispersion(inp, flags)
    compare_rows(out.reshape(want.shape), want, w.obs.nr_samples_per_batch, f"patterns {conf}")


def test_single_step_patterns_fall_back_when_too_many(torch_cuda):
    """A coarse DM grid over narrow channels gives more than 16 patterns per
    block and tile; local 3 then runs the staged pipeline, same results."""
    w = workloads.tiny(channels=64, samples=600, dms=40, step=4.0, min_freq=1300.0, bandwidth=1.0, seed=18)
    plan, inp, out, orc, flags = single_step(torch_cuda, w, padding=128, conf=Conf(64, 8, 8, 4, 4, 3), seed=18)
    want = orc.dedispersion(inp, flags)
    compare_rows(out.reshape(want.shape), want, w.obs.nr_samples_per_batch, "fallback")


def test_single_step_patterns_plain_128(torch_cuda):
    w = workloads.plain_128()
    plan, inp, out, orc, flags = single_step(torch_cuda, w, conf=Conf(64, 8, 16, 4, 8, 3))
    want = orc.dedispersion(inp, flags)
    compare_rows(out.reshape(want.shape), want, w.obs.nr_samples_per_batch, "plain-128 patterns")


def test_single_step_beam_subrange(torch_cuda, ragged_case):
    """firstSynthesizedBeam offsets the beam mapping; the output is relative
    to the launch (include/Dedispersion.hpp:228, 267)."""
    w, plan, d_in, want = ragged_case
    plan.set_conf(DD_SINGLE_STEP, Conf())
    z = plan.sizes
    per_beam = w.obs.nr_dms * z.out_pitch
    d_out = torch_cuda.zeros(2 * per_beam, dtype=torch_cuda.float32, device="cuda")
    plan.dedisperse(d_in.data_ptr(), d_out.data_ptr(), first_sbeam=2, nr_sbeams=2, stream=stream_of(torch_cuda))
    got = d_out.cpu().numpy().reshape(2, w.obs.nr_dms, z.out_pitch)
    compare_rows(got, want[2:4], w.obs.nr_samples_per_batch, "beams 2..3")


_ARTS_SINGLE = {}


@pytest.mark.parametrize("conf", [None, Conf(64, 8, 16, 4, 4, 3)], ids=["default", "patterns"])
def test_single_step_arts(torch_cuda, conf):
    """ARTS-sized single step: 1536 channels, 2048 DMs, 24576 samples, 77
    zapped channels, Gaussian noise and 8 dispersed pulses.  The whole output
    (2048 DM rows) is compared with the oracle; the DM-0 row is checked as
    the plain channel sum and every pulse must stand out at its DM."""
    w = workloads.arts_single()
    plan, inp, out, orc, flags = single_step(torch_cuda, w, conf=conf)
    D, S = w.obs.nr_dms, w.obs.nr_samples_per_batch
    out = out.reshape(1, D, plan.sizes.out_pitch)
    # the seeded device input is the same for every configuration: one oracle run
    if "want" not in _ARTS_SINGLE or not np.array_equal(_ARTS_SINGLE["input"], inp):
        _ARTS_SINGLE.update(input=inp, want=orc.dedispersion(inp, flags))
    compare_rows(out, _ARTS_SINGLE["want"], S, "ARTS single step, every DM row")
    # DM 0: plain sum of the live channels
    x = inp.reshape(w.obs.nr_channels, plan.sizes.in_pitch)[:, :S].astype(np.float64)
    x[np.asarray(w.zapped)] = 0
    np.testing.assert_array_equal(out[0, 0, :S], x.sum(axis=0))
    # each pulse is recovered at its DM: the boxcar region stands out
    live = w.obs.nr_channels - len(w.zapped)
    for p in w.pulses:
        row = out[0, p.dm_index, :S]
        peak = row[p.t0:p.t0 + p.width].mean()
        assert peak > np.median(row) + 0.25 * live * p.amplitude, (p, peak, np.median(row))


# --------------------------------------------------------------- two step


def two_step_case(torch, w, padding=128, conf1=None, conf2=None, seed=None):
    plan = Plan(w.obs, padding)
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    if conf1 is not None:
        plan.set_conf(DD_STEP_ONE, conf1)
    if conf2 is not None:
        plan.set_conf(DD_STEP_TWO, conf2)
    z = plan.sizes
    d_in = make_input(torch, plan, w, z.sub_in_pitch, z.nr_dispersed_samples_subbanding,
                      w.seed if seed is None else seed)
    d_mid = torch.full((z.intermediate_elements,), -1.0, dtype=torch.float32, device="cuda")
    d_out = torch.full((z.sub_output_elements,), -1.0, dtype=torch.float32, device="cuda")
    s = stream_of(torch)
    plan.step_one(d_in.data_ptr(), d_mid.data_ptr(), stream=s)
    plan.step_two(d_mid.data_ptr(), d_out.data_ptr(), stream=s)
    torch.cuda.synchronize()
    orc = oracle.Oracle(w.obs, padding, threads=host_threads())
    return plan, flags, d_in.cpu().numpy(), d_mid.cpu().numpy(), d_out.cpu().numpy(), orc


STEP_ONE_CONFS = [
    Conf(64, t1, i0, i1, u, loc)
    for (t1, i0, i1, u, loc) in itertools.product((1, 4), (4, 8, 16), (1, 2, 8), (1, 5), (1, 2, 0))
]


@pytest.fixture(scope="module")
def tiny_two_step(torch_cuda):
    w = workloads.tiny(channels=96, samples=700, dms=12, step=0.9, subbands=8, coarse=9, coarse_step=4.5, beams=2,
                       sbeams=3, min_freq=1300.0, bandwidth=1.0, zapped=(0, 11, 12, 50, 95))
    plan = Plan(w.obs, 0)
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    z = plan.sizes
    d_in = make_input(torch_cuda, plan, w, z.sub_in_pitch, z.nr_dispersed_samples_subbanding, 21)
    orc = oracle.Oracle(w.obs, 0, threads=host_threads())
    want_mid = orc.step_one(d_in.cpu().numpy(), flags)
    return w, plan, d_in, orc, want_mid


@pytest.mark.parametrize("conf", STEP_ONE_CONFS, ids=lambda c: f"t{c.threads_d1}i{c.items_d0}x{c.items_d1}u{c.unroll}l{c.local}")
def test_step_one_conf_sweep(torch_cuda, tiny_two_step, conf):
    w, plan, d_in, orc, want = tiny_two_step
    plan.set_conf(DD_STEP_ONE, conf)
    z = plan.sizes
    d_mid = torch_cuda.full((z.intermediate_elements,), -1.0, dtype=torch_cuda.float32, device="cuda")
    plan.step_one(d_in.data_ptr(), d_mid.data_ptr(), stream=stream_of(torch_cuda))
    got = d_mid.cpu().numpy().reshape(want.shape)
    compare_rows(got, want, z.nr_samples_subbanding, f"step one {conf}")


STEP_TWO_CONFS = [
    Conf(64, t1, i0, i1, u, loc)
    for (t1, i0, i1, u, loc) in itertools.product((1, 2, 4), (4, 8, 16), (1, 4, 8), (1, 3), (1, 2, 0))
]


@pytest.fixture(scope="module")
def tiny_step_two_floats(torch_cuda, tiny_two_step):
    """General (non-integer) floats as the step-two input, mixed mapping."""
    w, plan, _, orc, _ = tiny_two_step
    z = plan.sizes
    rng = np.random.default_rng(7)
    mid = (rng.standard_normal(z.intermediate_elements) * 300.0).astype(np.float32)
    mapping = orc.default_map(8)
    mapping[1, :8] = [1, 0, 1, 0, 0, 1, 1, 0]
    mapping[2, :8] = 0
    plan.set_beam_mapping(step_two=mapping)
    want = orc.step_two(mid, mapping)
    d_mid = torch_cuda.from_numpy(mid).cuda()
    return w, plan, d_mid, want


@pytest.mark.parametrize("conf", STEP_TWO_CONFS, ids=lambda c: f"t{c.threads_d1}i{c.items_d0}x{c.items_d1}u{c.unroll}l{c.local}")
def test_step_two_conf_sweep_general_floats(torch_cuda, tiny_step_two_floats, conf):
    w, plan, d_mid, want = tiny_step_two_floats
    plan.set_conf(DD_STEP_TWO, conf)
    z = plan.sizes
    d_out = torch_cuda.full((z.sub_output_elements,), -1.0, dtype=torch_cuda.float32, device="cuda")
    plan.step_two(d_mid.data_ptr(), d_out.data_ptr(), stream=stream_of(torch_cuda))
    got = d_out.cpu().numpy().reshape(want.shape)
    compare_rows(got, want, w.obs.nr_samples_per_batch, f"step two {conf}")


def test_two_step_arts(torch_cuda):
    """ARTS two-step: 32 subbands of 48 channels, 64 coarse DMs of 6.4, 32
    fine DMs of 0.2.  The whole intermediate and the whole output are
    compared with the oracle (step two over the device's intermediate, and
    again over the oracle's own, which must give the same)."""
    w = workloads.arts_two_step()
    plan, flags, inp, mid, out, orc = two_step_case(torch_cuda, w)
    z = plan.sizes
    o = w.obs
    want_mid = orc.step_one(inp, flags)
    mid = mid.reshape(want_mid.shape)
    compare_rows(mid, want_mid, z.nr_samples_subbanding, "ARTS step one, every coarse DM and subband")
    out = out.reshape(1, o.nr_dms_subbanding * o.nr_dms, z.out_pitch)
    want = orc.step_two(want_mid.ravel())
    compare_rows(out, want, o.nr_samples_per_batch, "ARTS step two, every DM row")
    # intermediate padding is never written; step two output rows are complete
    assert np.all(mid[..., z.nr_samples_subbanding:] == -1.0)
    assert np.all(out[..., : o.nr_samples_per_batch] >= 0)


def test_step_two_beam_subrange(torch_cuda, tiny_step_two_floats):
    """firstSynthesizedBeam of step two (include/Dedispersion.hpp:906, 932,
    974): synthesized beams [1, 3) of 3, with the mixed mapping, written from
    the start of the launch's output, general floats."""
    w, plan, d_mid, want = tiny_step_two_floats
    plan.set_conf(DD_STEP_TWO, Conf(64, 2, 8, 4, 3, 1))
    z = plan.sizes
    per_beam = w.obs.nr_dms_subbanding * w.obs.nr_dms * z.out_pitch
    for first, n in ((1, 2), (2, 1), (0, 1)):
        d_out = torch_cuda.full((n * per_beam,), -1.0, dtype=torch_cuda.float32, device="cuda")
        plan.step_two(d_mid.data_ptr(), d_out.data_ptr(), first_sbeam=first, nr_sbeams=n, stream=stream_of(torch_cuda))
        got = d_out.cpu().numpy().reshape(n, -1, z.out_pitch)
        compare_rows(got, want[first:first + n], w.obs.nr_samples_per_batch, f"step two sBeams {first}..{first + n}")
        assert np.all(got[..., w.obs.nr_samples_per_batch:] == -1.0)
    with pytest.raises(DDError):
        plan.step_two(d_mid.data_ptr(), d_out.data_ptr(), first_sbeam=2, nr_sbeams=2)


# ------------------------------------------------------------ error paths


def test_invalid_configurations_are_rejected(torch_cuda):
    w = workloads.plain_128()
    plan = Plan(w.obs, 128)
    for bad in (Conf(threads_d0=32), Conf(threads_d1=0), Conf(threads_d1=9), Conf(items_d0=5), Conf(items_d1=3),
                Conf(unroll=0), Conf(unroll=17), Conf(split_batches=1), Conf(threads_d1=8, items_d1=16), Conf(local=3)):
        with pytest.raises(DDError):
            plan.set_conf(DD_SINGLE_STEP, bad)
    with pytest.raises(DDError):
        plan.set_conf(DD_STEP_ONE, Conf())  # no subband path in this plan
    with pytest.raises(DDError):
        plan.dedisperse(0, 0)
    # a dispersed batch shorter than samples + largest shift is refused
    o = workloads.plain_128().obs
    o.nr_samples_per_dispersed_batch = o.nr_samples_per_batch + 10
    short = Plan(o, 128)
    buf = torch_cuda.zeros(16, dtype=torch_cuda.float32, device="cuda")
    with pytest.raises(DDError):
        short.dedisperse(buf.data_ptr(), buf.data_ptr())


# --------------------------------------------------- buffer placement (regression)


def high_address_view(torch, nbytes: int):
    """A uint8 view of device memory whose address has bit 31 of its low 32
    bits set (the buffer resources of the kernels carry 48-bit addresses; a
    sign-extended low half once corrupted their upper bits)."""
    big = torch.zeros(2**31 + nbytes + 2**20, dtype=torch.uint8, device="cuda")
    base = big.data_ptr()
    low = base & 0xFFFFFFFF
    off = 0 if low >= 0x80000000 else ((0x80000000 - low + 255) & ~255)
    view = big[off:off + nbytes]
    assert (view.data_ptr() & 0x80000000) and view.data_ptr() % 256 == 0
    return big, view


def test_buffers_at_high_addresses(torch_cuda):
    torch = torch_cuda
    w = workloads.tiny(channels=96, samples=700, dms=12, step=0.9, subbands=8, coarse=9, coarse_step=4.5,
                       min_freq=1300.0, bandwidth=1.0, zapped=(3, 50))
    plan = Plan(w.obs, 128)
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    z = plan.sizes
    s = stream_of(torch)
    keep_in, d_in = high_address_view(torch, z.input_elements)
    src = make_input(torch, plan, w, z.in_pitch, z.nr_dispersed_samples, 5)
    d_in.copy_(src)
    keep_out, d_out = high_address_view(torch, 4 * z.output_elements)
    plan.dedisperse(d_in.data_ptr(), d_out.data_ptr(), stream=s)
    keep_sub, d_sub = high_address_view(torch, z.sub_input_elements)
    d_sub.copy_(make_input(torch, plan, w, z.sub_in_pitch, z.nr_dispersed_samples_subbanding, 6))
    keep_mid, d_mid = high_address_view(torch, 4 * z.intermediate_elements)
    keep_two, d_two = high_address_view(torch, 4 * z.sub_output_elements)
    plan.step_one(d_sub.data_ptr(), d_mid.data_ptr(), stream=s)
    plan.step_two(d_mid.data_ptr(), d_two.data_ptr(), stream=s)
    torch.cuda.synchronize()
    orc = oracle.Oracle(w.obs, 128, threads=host_threads())
    want = orc.dedispersion(d_in.cpu().numpy(), flags)
    compare_rows(d_out.view(torch.float32).cpu().numpy().reshape(want.shape), want, w.obs.nr_samples_per_batch,
                 "single step at high addresses")
    want_mid = orc.step_one(d_sub.cpu().numpy(), flags)
    mid = d_mid.view(torch.float32).cpu().numpy().reshape(want_mid.shape)
    compare_rows(mid, want_mid, z.nr_samples_subbanding, "step one at high addresses")
    want_two = orc.step_two(mid.ravel())
    compare_rows(d_two.view(torch.float32).cpu().numpy().reshape(want_two.shape), want_two,
                 w.obs.nr_samples_per_batch, "step two at high addresses")
