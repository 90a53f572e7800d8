"""Float input (the reference's templates and generators with I = float,
include/Dedispersion.hpp:59-60, 63-64) on the device against the oracle.

A float row sum adds the live channels in order into a float accumulator
that starts at 0: the reference's sequential float sum.  The device adds in
the same order, so results are compared bit for bit, which is stronger than
the relative error of 1e-6 allowed for float input; the inputs are general
floats (normal noise of large amplitude, so the channel order changes the
rounding).  Covered: every kernel selection of the single step (local 0, 1,
2; local 3 runs the staged pipeline for float rows), zapped channels, beam
mappings and subranges, firstDM != 0 with a hazard table, the 128-channel
single-step shape (config 1) and the ARTS two-step shape (config 3) in full,
step one followed by step two, and the refusals between float plans and the
8-bit entry points, batches and streams."""
import numpy as np
import pytest

import oracle
from dedispersion_amd import DD_SINGLE_STEP, DD_STEP_ONE, DD_STEP_TWO, Batch, Conf, DDError, Plan, workloads
from tests import fused
from tests.conftest import host_threads
from tests.gpu_helpers import compare_rows, stream_of, zapped_flags

pytestmark = pytest.mark.gpu

RTOL = 1e-6  # the tolerance allowed for float input; these kernels meet it with zero error


def float_rows(torch, n, seed, scale=1000.0):
    rng = np.random.default_rng(seed)
    host = (rng.standard_normal(n) * scale).astype(np.float32)
    return host, torch.from_numpy(host).cuda()


def check(got, want, valid, what):
    """Bit-exact; on a mismatch, report the relative error against RTOL too."""
    g, w = got[..., :valid], want[..., :valid]
    bad = np.argwhere(g != w)
    if bad.size:
        rel = float(np.max(np.abs(g - w) / np.maximum(np.abs(w), 1e-30)))
        raise AssertionError(f"{what}: {len(bad)} samples differ (largest relative error {rel:.3g}, "
                             f"allowed {RTOL}), first at {bad[:4].tolist()}")


SINGLE_CONFS = [None, Conf(64, 4, 4, 2, 4, 1), Conf(64, 8, 16, 8, 4, 1), Conf(64, 2, 8, 4, 3, 2),
                Conf(64, 1, 16, 8, 1, 0), Conf(64, 8, 8, 4, 8, 3), Conf(64, 4, 4, 1, 16, 1)]


@pytest.mark.parametrize("conf", SINGLE_CONFS, ids=lambda c: "default" if c is None else
                         f"t{c.threads_d1}i{c.items_d0}x{c.items_d1}u{c.unroll}l{c.local}")
@pytest.mark.parametrize("padding", [128, 0], ids=["pad128", "unpadded"])
def test_single_step_float(torch_cuda, padding, conf):
    torch = torch_cuda
    w = workloads.tiny(channels=120, samples=1000, dms=37, step=0.7, first=4.1, beams=2, sbeams=3, min_freq=1300.0,
                       bandwidth=0.5, zapped=(0, 9, 60, 119), seed=91)
    plan = Plan(w.obs, padding, input_bits=32)
    orc = oracle.Oracle(w.obs, padding, threads=host_threads(), input_bits=32)
    z = plan.sizes
    assert z.input_bits == 32 and z.in_pitch == orc.in_pitch
    assert z.input_elements == 2 * 120 * orc.in_pitch
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    mapping = orc.default_map(120)
    mapping[2, :50] = 0
    mapping[0, 70:] = 1
    plan.set_beam_mapping(single=mapping)
    if conf is not None:
        plan.set_conf(DD_SINGLE_STEP, conf)
    host, d_in = float_rows(torch, z.input_elements, 91)
    d_out = torch.full((z.output_elements,), -1.0, dtype=torch.float32, device="cuda")
    plan.dedisperse_f32(d_in.data_ptr(), d_out.data_ptr(), stream=stream_of(torch))
    torch.cuda.synchronize()
    want = orc.dedispersion(host, flags, mapping)
    got = d_out.cpu().numpy().reshape(want.shape)
    check(got, want, z.nr_samples, f"float single step, conf {conf}")
    assert np.all(got[..., z.nr_samples:] == -1.0)
    # a subrange of synthesized beams, written from the start of the output
    d_sub = torch.full((2 * w.obs.nr_dms * z.out_pitch,), -1.0, dtype=torch.float32, device="cuda")
    plan.dedisperse_f32(d_in.data_ptr(), d_sub.data_ptr(), first_sbeam=1, nr_sbeams=2, stream=stream_of(torch))
    check(d_sub.cpu().numpy().reshape(2, w.obs.nr_dms, z.out_pitch), want[1:3], z.nr_samples, "sBeams 1..2")
    plan.close()


def test_single_step_float_first_dm_hazard(torch_cuda):
    """A hazard table (fused and unfused shifts differ at some DM of every
    channel) with firstDM 3.3: the float kernels' shifts are unfused."""
    torch = torch_cuda
    w = workloads.tiny(channels=64, samples=700, dms=40, step=0.7, first=3.3, min_freq=1300.0, bandwidth=0.5,
                       zapped=(5,), seed=92)
    w.obs.nr_samples_per_dispersed_batch = 700 + 130
    table = np.zeros(64, dtype=np.float32)
    table[:] = fused.hazard_table(64, 3.3, 0.7, 40, 120, seed=92)
    plan = Plan(w.obs, 0, input_bits=32)
    plan.set_shifts(shifts=table)
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    z = plan.sizes
    host, d_in = float_rows(torch, z.input_elements, 92)
    for conf in (Conf(64, 4, 8, 2, 4, 1), Conf(64, 2, 4, 4, 2, 0)):
        plan.set_conf(DD_SINGLE_STEP, conf)
        d_out = torch.full((z.output_elements,), -1.0, dtype=torch.float32, device="cuda")
        plan.dedisperse_f32(d_in.data_ptr(), d_out.data_ptr(), stream=stream_of(torch))
        torch.cuda.synchronize()
        want = oracle.Oracle(w.obs, 0, threads=host_threads(), input_bits=32).dedispersion(host, flags, shifts=table)
        check(d_out.cpu().numpy().reshape(want.shape), want, z.nr_samples, f"hazard table, {conf}")
    plan.close()


def test_single_step_float_plain_128(torch_cuda):
    """Config 1's shape (128 channels from 1400 MHz, 4096 samples, 128 DMs,
    channels 17 and 90 zapped) with float input, whole output."""
    torch = torch_cuda
    w = workloads.plain_128()
    plan = Plan(w.obs, 128, input_bits=32)
    orc = oracle.Oracle(w.obs, 128, threads=host_threads(), input_bits=32)
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    z = plan.sizes
    host, d_in = float_rows(torch, z.input_elements, 93, scale=50.0)
    d_out = torch.full((z.output_elements,), -1.0, dtype=torch.float32, device="cuda")
    plan.dedisperse_f32(d_in.data_ptr(), d_out.data_ptr(), stream=stream_of(torch))
    torch.cuda.synchronize()
    want = orc.dedispersion(host, flags)
    check(d_out.cpu().numpy().reshape(want.shape), want, z.nr_samples, "plain-128, float input")
    plan.close()


def test_float_rows_of_integers_equal_the_byte_path(torch_cuda):
    """Integer-valued float input gives the 8-bit plan's output."""
    torch = torch_cuda
    w = workloads.tiny(channels=96, samples=900, dms=20, step=0.6, min_freq=1300.0, bandwidth=1.0, zapped=(4,),
                       seed=94)
    p8, pf = Plan(w.obs, 128), Plan(w.obs, 128, input_bits=32)
    flags = zapped_flags(p8, w.zapped)
    p8.set_zapped_channels(flags)
    pf.set_zapped_channels(flags)
    z8, zf = p8.sizes, pf.sizes
    rng = np.random.default_rng(94)
    b = rng.integers(0, 256, size=(96, z8.in_pitch), dtype=np.uint8)
    f = np.zeros((96, zf.in_pitch), dtype=np.float32)
    n = z8.nr_dispersed_samples
    f[:, :n] = b[:, :n]
    o8 = torch.full((z8.output_elements,), -1.0, dtype=torch.float32, device="cuda")
    of = torch.full((zf.output_elements,), -1.0, dtype=torch.float32, device="cuda")
    d8, df = torch.from_numpy(b.ravel()).cuda(), torch.from_numpy(f.ravel()).cuda()
    p8.dedisperse(d8.data_ptr(), o8.data_ptr(), stream=stream_of(torch))
    pf.dedisperse_f32(df.data_ptr(), of.data_ptr(), stream=stream_of(torch))
    torch.cuda.synchronize()
    assert torch.equal(o8.view(20, -1)[:, :900], of.view(20, -1)[:, :900])


STEP_ONE_CONFS = [None, Conf(64, 4, 4, 2, 4, 1), Conf(64, 1, 16, 8, 2, 1), Conf(64, 2, 8, 1, 5, 2),
                  Conf(64, 4, 4, 2, 3, 0)]


@pytest.mark.parametrize("conf", STEP_ONE_CONFS, ids=lambda c: "default" if c is None else
                         f"t{c.threads_d1}i{c.items_d0}x{c.items_d1}u{c.unroll}l{c.local}")
def test_step_one_float_then_step_two(torch_cuda, conf):
    torch = torch_cuda
    w = workloads.tiny(channels=96, samples=700, dms=12, step=0.9, first=1.1, subbands=8, coarse=9, coarse_step=4.5,
                       coarse_first=2.7, beams=2, sbeams=3, min_freq=1300.0, bandwidth=1.0, zapped=(0, 11, 12, 50),
                       seed=95)
    plan = Plan(w.obs, 0, input_bits=32)
    orc = oracle.Oracle(w.obs, 0, threads=host_threads(), input_bits=32)
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    if conf is not None:
        plan.set_conf(DD_STEP_ONE, conf)
    z = plan.sizes
    assert z.sub_in_pitch == orc.sub_in_pitch
    host, d_in = float_rows(torch, z.sub_input_elements, 95)
    d_mid = torch.full((z.intermediate_elements,), -1.0, dtype=torch.float32, device="cuda")
    d_out = torch.full((z.sub_output_elements,), -1.0, dtype=torch.float32, device="cuda")
    s = stream_of(torch)
    plan.step_one_f32(d_in.data_ptr(), d_mid.data_ptr(), stream=s)
    plan.step_two(d_mid.data_ptr(), d_out.data_ptr(), stream=s)
    torch.cuda.synchronize()
    want_mid = orc.step_one(host, flags)
    mid = d_mid.cpu().numpy().reshape(want_mid.shape)
    check(mid, want_mid, z.nr_samples_subbanding, f"float step one, {conf}")
    assert np.all(mid[..., z.nr_samples_subbanding:] == -1.0)
    want = orc.step_two(mid.ravel())
    check(d_out.cpu().numpy().reshape(want.shape), want, w.obs.nr_samples_per_batch, f"step two after, {conf}")
    plan.close()


def test_two_step_float_arts(torch_cuda):
    """Config 3's shape (1536 channels, 32 subbands, 64 coarse x 32 fine DMs,
    24576 samples, 5% zapped) with float input: the whole intermediate and
    the whole output."""
    torch = torch_cuda
    w = workloads.arts_two_step()
    plan = Plan(w.obs, 128, input_bits=32)
    orc = oracle.Oracle(w.obs, 128, threads=host_threads(), input_bits=32)
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    z = plan.sizes
    host, d_in = float_rows(torch, z.sub_input_elements, 96, scale=16.0)
    host += 128.0
    d_in = torch.from_numpy(host).cuda()
    d_mid = torch.full((z.intermediate_elements,), -1.0, dtype=torch.float32, device="cuda")
    d_out = torch.full((z.sub_output_elements,), -1.0, dtype=torch.float32, device="cuda")
    s = stream_of(torch)
    plan.step_one_f32(d_in.data_ptr(), d_mid.data_ptr(), stream=s)
    plan.step_two(d_mid.data_ptr(), d_out.data_ptr(), stream=s)
    torch.cuda.synchronize()
    want_mid = orc.step_one(host, flags)
    mid = d_mid.cpu().numpy().reshape(want_mid.shape)
    check(mid, want_mid, z.nr_samples_subbanding, "ARTS float step one")
    want = orc.step_two(mid.ravel())
    check(d_out.cpu().numpy().reshape(want.shape), want, w.obs.nr_samples_per_batch, "ARTS float step two")
    plan.close()


def test_float_plans_and_byte_entry_points_refuse_each_other(torch_cuda):
    w = workloads.tiny(channels=32, samples=200, dms=4, step=1.0, subbands=4, coarse=2, coarse_step=2.0)
    buf = torch_cuda.zeros(1 << 16, dtype=torch_cuda.float32, device="cuda")
    pf, p8 = Plan(w.obs, 0, input_bits=32), Plan(w.obs, 0)
    with pytest.raises(DDError, match="float input"):
        pf.dedisperse(buf.data_ptr(), buf.data_ptr())
    with pytest.raises(DDError, match="float input"):
        pf.step_one(buf.data_ptr(), buf.data_ptr())
    with pytest.raises(DDError, match="input_bits 32"):
        p8.dedisperse_f32(buf.data_ptr(), buf.data_ptr())
    with pytest.raises(DDError, match="input_bits 32"):
        p8.step_one_f32(buf.data_ptr(), buf.data_ptr())
    with Batch() as b:
        with pytest.raises(DDError, match="float"):
            b.add_single_step(pf, buf.data_ptr(), buf.data_ptr())
