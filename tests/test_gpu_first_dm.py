"""Device parity where firstDM != 0: the shift rule's unfused arithmetic.

A shift is (unsigned)((firstDM + dm * DMStep) * table[row]) in separately
rounded float operations (include/Dedispersion.hpp:81, 124, 168).  With
firstDM = 0 a fused multiply-add gives the same DM value, so only a nonzero
firstDM shows whether a kernel contracts the expression.  The tables here
are built (tests/fused.py) so that at some DM of every row the fused and the
unfused shifts differ; each test asserts that, then compares the device
output with the oracle over the whole array.  The device's integer tables
and every kernel kind are covered: the direct, staged and pattern single
step, step one (whose relative table is the channel table minus the
subband's top channel, DH:124) and step two, packed rows, and batches.
"""
import numpy as np
import pytest

import oracle
from dedispersion_amd import DD_SINGLE_STEP, DD_STEP_ONE, DD_STEP_TWO, Batch, Conf, Plan, workloads
from tests import fused
from tests.conftest import host_threads
from tests.gpu_helpers import compare_rows, make_input, stream_of, zapped_flags
from tests.test_gpu_packed import packed_input

pytestmark = pytest.mark.gpu

FIRST, STEP = 3.3, 0.7                 # single step and step two (fine DMs)
COARSE_FIRST, COARSE_STEP = 12.7, 4.3  # step one (coarse DMs)
MAX_SHIFT = 120


def padded(table, padding):
    out = np.zeros(oracle.pad(len(table), padding // 4), dtype=np.float32)
    out[: len(table)] = table
    return out


def single_case(channels=128, samples=900, dms=40, beams=2, sbeams=3, zapped=(3, 77), padding=128, seed=51):
    """A single-step workload with a hazard table; the dispersed length is
    fixed so that the table's largest shift fits."""
    w = workloads.tiny(channels=channels, samples=samples, dms=dms, step=STEP, first=FIRST, beams=beams,
                       sbeams=sbeams, min_freq=1300.0, bandwidth=0.5, zapped=zapped, seed=seed)
    w.obs.nr_samples_per_dispersed_batch = samples + MAX_SHIFT + 8
    table = padded(fused.hazard_table(channels, FIRST, STEP, dms, MAX_SHIFT, seed=seed), padding)
    live = [c for c in range(channels) if c not in set(zapped)]
    assert fused.differing(table[:channels], FIRST, STEP, dms, rows=live) >= len(live) // 2
    return w, table


def run_single(torch, w, table, padding, conf=None, bits=8):
    plan = Plan(w.obs, padding, input_bits=bits)
    plan.set_shifts(shifts=table)
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    if conf is not None:
        plan.set_conf(DD_SINGLE_STEP, conf)
    z = plan.sizes
    orc = oracle.Oracle(w.obs, padding, threads=host_threads(), input_bits=bits)
    if bits == 8:
        d_in = make_input(torch, plan, w, z.in_pitch, z.nr_dispersed_samples, w.seed)
        host = d_in.cpu().numpy()
    else:
        host, d_in = packed_input(torch, orc, bits, z.nr_dispersed_samples, z.in_pitch, seed=w.seed)
    d_out = torch.full((z.output_elements,), -1.0, dtype=torch.float32, device="cuda")
    plan.dedisperse(d_in.data_ptr(), d_out.data_ptr(), stream=stream_of(torch))
    torch.cuda.synchronize()
    want = orc.dedispersion(host, flags, shifts=table)
    got = d_out.cpu().numpy().reshape(want.shape)
    return plan, got, want


def test_device_integer_tables_are_unfused(torch_cuda):
    """The device-built integer tables of all three modes equal the unfused
    rule and differ from the fused one."""
    w = workloads.tiny(channels=96, samples=500, dms=40, step=STEP, first=FIRST, subbands=8, coarse=24,
                       coarse_step=COARSE_STEP, coarse_first=COARSE_FIRST, min_freq=1300.0, bandwidth=0.5)
    cps = 12
    tops = [(sb + 1) * cps - 1 for sb in range(8)]
    t1 = fused.hazard_table(96, COARSE_FIRST, COARSE_STEP, 24, MAX_SHIFT, seed=3, zero_rows=tops)
    t2 = fused.hazard_table(8, FIRST, STEP, 40, MAX_SHIFT, seed=4)
    w.obs.nr_samples_per_batch_subbanding = 500 + MAX_SHIFT + 8
    w.obs.nr_samples_per_dispersed_batch_subbanding = 500 + 2 * MAX_SHIFT + 16
    w.obs.nr_samples_per_dispersed_batch = 500 + MAX_SHIFT + 8
    plan = Plan(w.obs, 0)
    plan.set_shifts(shifts=t1, shifts_step_two=t2)
    # t1 is a hazard table for the coarse DM range (step one), t2 for the
    # fine one (step two); the single step reads t1 over the fine range
    for mode, table, first, step, dms, hazard in ((DD_SINGLE_STEP, t1, FIRST, STEP, 40, False),
                                                  (DD_STEP_ONE, t1, COARSE_FIRST, COARSE_STEP, 24, True),
                                                  (DD_STEP_TWO, t2, FIRST, STEP, 40, True)):
        got = plan.get_integer_shifts(mode)
        np.testing.assert_array_equal(got, fused.shift_table(table, first, step, dms))
        if hazard:
            differ = np.count_nonzero(got != fused.shift_table(table, first, step, dms, fused=True))
            assert differ >= (len(table) - len(tops) if mode == DD_STEP_ONE else len(table)) // 2


SINGLE_CONFS = [None, Conf(64, 4, 4, 2, 4, 1), Conf(64, 8, 16, 8, 4, 1), Conf(64, 2, 8, 4, 3, 2),
                Conf(64, 1, 16, 8, 1, 0), Conf(64, 8, 8, 4, 8, 3), Conf(64, 4, 16, 8, 2, 3)]


@pytest.mark.parametrize("conf", SINGLE_CONFS, ids=lambda c: "default" if c is None else
                         f"t{c.threads_d1}i{c.items_d0}x{c.items_d1}u{c.unroll}l{c.local}")
def test_single_step_first_dm(torch_cuda, conf):
    """Every single-step kernel selection (local 0, 1, 2 and the pattern
    kernel) over a hazard table, 3 beams from 2 with the default mapping."""
    w, table = single_case()
    plan, got, want = run_single(torch_cuda, w, table, 128, conf)
    compare_rows(got, want, w.obs.nr_samples_per_batch, f"first DM {FIRST}, conf {conf}")
    plan.close()


def test_single_step_first_dm_many_channels(torch_cuda):
    """More than 257 live channels (the kernels that flush 16-bit pair sums)."""
    w, table = single_case(channels=300, dms=24, zapped=(0, 150), padding=0, seed=52)
    plan, got, want = run_single(torch_cuda, w, table, 0, Conf(64, 2, 16, 8, 4, 1))
    compare_rows(got, want, w.obs.nr_samples_per_batch, "300 channels")
    plan.close()


@pytest.mark.parametrize("bits", [1, 2, 4], ids=lambda b: f"{b}bit")
def test_single_step_first_dm_packed(torch_cuda, bits):
    w, table = single_case(seed=53 + bits)
    plan, got, want = run_single(torch_cuda, w, table, 128, Conf(64, 4, 8, 4, 4, 1), bits=bits)
    compare_rows(got, want, w.obs.nr_samples_per_batch, f"{bits}-bit first DM {FIRST}")
    plan.close()


def two_step_case(seed=61, padding=0, channels=96, subbands=8, dms=24, coarse=16, samples=700, beams=2, sbeams=3,
                  zapped=(0, 11, 50)):
    """A two-step workload whose channel table (so the relative step-one
    table: every subband's top channel is 0) and step-two table are hazard
    tables for the coarse and the fine DM ranges."""
    w = workloads.tiny(channels=channels, samples=samples, dms=dms, step=STEP, first=FIRST, subbands=subbands,
                       coarse=coarse, coarse_step=COARSE_STEP, coarse_first=COARSE_FIRST, beams=beams, sbeams=sbeams,
                       min_freq=1300.0, bandwidth=0.5, zapped=zapped, seed=seed)
    cps = channels // subbands
    tops = [(sb + 1) * cps - 1 for sb in range(subbands)]
    t1 = fused.hazard_table(channels, COARSE_FIRST, COARSE_STEP, coarse, MAX_SHIFT, seed=seed, zero_rows=tops)
    t2 = fused.hazard_table(subbands, FIRST, STEP, dms, MAX_SHIFT, seed=seed + 1)
    live = [c for c in range(channels) if c not in set(zapped) and c not in set(tops)]
    assert fused.differing(t1, COARSE_FIRST, COARSE_STEP, coarse, rows=live) >= len(live) // 2
    assert fused.differing(t2, FIRST, STEP, dms) >= subbands // 2
    o = w.obs
    o.nr_samples_per_batch_subbanding = samples + MAX_SHIFT + 8
    o.nr_samples_per_dispersed_batch_subbanding = o.nr_samples_per_batch_subbanding + MAX_SHIFT + 8
    o.nr_samples_per_dispersed_batch = samples + MAX_SHIFT + 8
    return w, padded(t1, padding), padded(t2, padding)


TWO_STEP_CONFS = [(None, None), (Conf(64, 4, 4, 2, 4, 1), Conf(64, 2, 8, 4, 3, 1)),
                  (Conf(64, 1, 16, 8, 2, 2), Conf(64, 4, 4, 8, 3, 2)),
                  (Conf(64, 2, 8, 1, 5, 0), Conf(64, 1, 16, 1, 1, 0))]


@pytest.mark.parametrize("confs", TWO_STEP_CONFS, ids=["default", "staged", "local2", "local0"])
def test_two_step_first_dm(torch_cuda, confs):
    torch = torch_cuda
    w, t1, t2 = two_step_case()
    plan = Plan(w.obs, 0)
    plan.set_shifts(shifts=t1, shifts_step_two=t2)
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    if confs[0] is not None:
        plan.set_conf(DD_STEP_ONE, confs[0])
        plan.set_conf(DD_STEP_TWO, confs[1])
    z = plan.sizes
    d_in = make_input(torch, plan, w, z.sub_in_pitch, z.nr_dispersed_samples_subbanding, w.seed)
    d_mid = torch.full((z.intermediate_elements,), -1.0, dtype=torch.float32, device="cuda")
    d_out = torch.full((z.sub_output_elements,), -1.0, dtype=torch.float32, device="cuda")
    s = stream_of(torch)
    plan.step_one(d_in.data_ptr(), d_mid.data_ptr(), stream=s)
    plan.step_two(d_mid.data_ptr(), d_out.data_ptr(), stream=s)
    torch.cuda.synchronize()
    orc = oracle.Oracle(w.obs, 0, threads=host_threads())
    want_mid = orc.step_one(d_in.cpu().numpy(), flags, shifts=t1)
    mid = d_mid.cpu().numpy().reshape(want_mid.shape)
    compare_rows(mid, want_mid, z.nr_samples_subbanding, f"step one, confs {confs}")
    want = orc.step_two(mid.ravel(), shifts=t2)
    compare_rows(d_out.cpu().numpy().reshape(want.shape), want, w.obs.nr_samples_per_batch, f"step two, {confs}")
    plan.close()


@pytest.mark.parametrize("bits", [1, 2, 4], ids=lambda b: f"{b}bit")
def test_step_one_first_dm_packed(torch_cuda, bits):
    torch = torch_cuda
    w, t1, t2 = two_step_case(seed=70 + bits, padding=128)
    plan = Plan(w.obs, 128, input_bits=bits)
    plan.set_shifts(shifts=t1, shifts_step_two=t2)
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    plan.set_conf(DD_STEP_ONE, Conf(64, 4, 8, 2, 4, 1))
    z = plan.sizes
    orc = oracle.Oracle(w.obs, 128, threads=host_threads(), input_bits=bits)
    host, d_in = packed_input(torch, orc, bits, z.nr_dispersed_samples_subbanding, z.sub_in_pitch, seed=w.seed)
    d_mid = torch.full((z.intermediate_elements,), -1.0, dtype=torch.float32, device="cuda")
    plan.step_one(d_in.data_ptr(), d_mid.data_ptr(), stream=stream_of(torch))
    torch.cuda.synchronize()
    want = orc.step_one(host, flags, shifts=t1)
    compare_rows(d_mid.cpu().numpy().reshape(want.shape), want, z.nr_samples_subbanding, f"{bits}-bit step one")
    plan.close()


@pytest.mark.parametrize("chunks", [0, 2], ids=["persistent", "pipelined"])
def test_batch_first_dm(torch_cuda, chunks):
    """Batches (uint16 intermediate) of two-step jobs with hazard tables
    beside a job with the observation's own tables, and a single-step job."""
    torch = torch_cuda
    jobs = []
    for seed in (81, 82):
        w, t1, t2 = two_step_case(seed=seed, padding=0, channels=96 if seed == 81 else 64,
                                  subbands=8 if seed == 81 else 4, coarse=12 if seed == 81 else 14)
        jobs.append((w, t1, t2))
    plain = workloads.tiny(channels=96, samples=600, dms=11, step=0.45, first=2.9, subbands=8, coarse=5,
                           coarse_step=3.7, coarse_first=10.1, min_freq=1300.0, bandwidth=1.0, zapped=(7,), seed=83)
    jobs.append((plain, None, None))
    s = stream_of(torch)
    plans, bufs, wants = [], [], []
    for w, t1, t2 in jobs:
        plan = Plan(w.obs, 0)
        if t1 is not None:
            plan.set_shifts(shifts=t1, shifts_step_two=t2)
        flags = zapped_flags(plan, w.zapped)
        plan.set_zapped_channels(flags)
        z = plan.sizes
        d_in = make_input(torch, plan, w, z.sub_in_pitch, z.nr_dispersed_samples_subbanding, w.seed)
        orc = oracle.Oracle(w.obs, 0, threads=host_threads())
        mid = orc.step_one(d_in.cpu().numpy(), flags, shifts=t1)
        wants.append((mid, orc.step_two(mid.ravel(), shifts=t2)))
        d_mid = torch.full((z.intermediate_elements,), -1.0, dtype=torch.float32, device="cuda")
        d_out = torch.full((z.sub_output_elements,), -1.0, dtype=torch.float32, device="cuda")
        plans.append(plan)
        bufs.append((d_in, d_mid, d_out))
    sw = workloads.tiny(channels=64, samples=800, dms=40, step=STEP, first=FIRST, min_freq=1300.0, bandwidth=0.5,
                        seed=84)
    sw.obs.nr_samples_per_dispersed_batch = 800 + MAX_SHIFT + 8
    st = fused.hazard_table(64, FIRST, STEP, 40, MAX_SHIFT, seed=84)
    splan = Plan(sw.obs, 0)
    splan.set_shifts(shifts=st)
    sz = splan.sizes
    s_in = make_input(torch, splan, sw, sz.in_pitch, sz.nr_dispersed_samples, sw.seed)
    s_out = torch.full((sz.output_elements,), -1.0, dtype=torch.float32, device="cuda")
    s_want = oracle.Oracle(sw.obs, 0, threads=host_threads()).dedispersion(s_in.cpu().numpy(), shifts=st)
    with Batch() as b:
        b.set_conf(DD_SINGLE_STEP, Conf(64, 4, 4, 2, 4, 1))
        b.set_conf(DD_STEP_ONE, Conf(64, 4, 4, 2, 4, 1))
        b.set_conf(DD_STEP_TWO, Conf(64, 2, 8, 4, 3, 1))
        b.set_pipeline(chunks)
        for plan, (d_in, d_mid, d_out) in zip(plans, bufs):
            b.add_two_step(plan, d_in.data_ptr(), d_mid.data_ptr(), d_out.data_ptr())
        b.add_single_step(splan, s_in.data_ptr(), s_out.data_ptr())
        info = b.info()
        assert info["compact_intermediate"] == 1 and info["solo_jobs"] == 0
        for use_graph in (False, True):
            for _, d_mid, d_out in bufs:
                d_mid.fill_(-1.0)
                d_out.fill_(-1.0)
            s_out.fill_(-1.0)
            b.run(use_graph=use_graph, stream=s)
            torch.cuda.synchronize()
            for j, (plan, (_, d_mid, d_out), (want_mid, want)) in enumerate(zip(plans, bufs, wants)):
                z = plan.sizes
                raw = d_mid.cpu().numpy()
                mid = raw.view(np.uint16)[: raw.size].astype(np.float32).reshape(want_mid.shape)
                compare_rows(mid, want_mid, z.nr_samples_subbanding, f"job {j} step one, graph {use_graph}")
                compare_rows(d_out.cpu().numpy().reshape(want.shape), want, plan.obs.nr_samples_per_batch,
                             f"job {j} step two, graph {use_graph}")
            compare_rows(s_out.cpu().numpy().reshape(s_want.shape), s_want, sw.obs.nr_samples_per_batch,
                         f"single-step job, graph {use_graph}")


def test_single_step_arts_band_first_dm(torch_cuda):
    """The ARTS band and sizes (1536 channels, 24576 samples, 77 zapped
    channels, Gaussian noise and pulses) with the observation's own table,
    512 DMs from 7.7 in steps of 1.3: 36 live (channel, DM) shifts differ
    between the fused and the unfused rule.  Whole array against the oracle."""
    w = workloads.arts_single()
    w.obs.first_dm, w.obs.dm_step, w.obs.nr_dms = 7.7, 1.3, 512
    w.pulses = []
    plan = Plan(w.obs, 128)
    s, _ = plan.get_shifts()
    live = [c for c in range(1536) if c not in set(w.zapped)]
    assert fused.differing(s[:1536], 7.7, 1.3, 512, rows=live) >= 30
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    z = plan.sizes
    d_in = make_input(torch_cuda, plan, w, z.in_pitch, z.nr_dispersed_samples, w.seed)
    d_out = torch_cuda.full((z.output_elements,), -1.0, dtype=torch_cuda.float32, device="cuda")
    plan.dedisperse(d_in.data_ptr(), d_out.data_ptr(), stream=stream_of(torch_cuda))
    torch_cuda.cuda.synchronize()
    want = oracle.Oracle(w.obs, 128, threads=host_threads()).dedispersion(d_in.cpu().numpy(), flags)
    compare_rows(d_out.cpu().numpy().reshape(want.shape), want, w.obs.nr_samples_per_batch, "ARTS band, first DM 7.7")
    plan.close()
