"""The allocation-free batch run (dd_batch_run_static) against the batch run
that allocates its work queues (dd_batch_run), through the C ABI: two calls
in a row on one stream, the second with a smaller batch, each bit-identical
to dd_batch_run of the same batch; and the refusals (queues beyond the
static array, jobs that would run beside the batch kernels)."""
import numpy as np
import pytest

from dedispersion_amd import DD_SINGLE_STEP, DD_STEP_ONE, DD_STEP_TWO, Batch, Conf, DDError, Plan, workloads
from dedispersion_amd._lib import check, load
from tests.gpu_helpers import make_input, stream_of, zapped_flags

pytestmark = pytest.mark.gpu

STATIC_BYTES = 4 << 20  # DD_BATCH_STATIC_BYTES


def run_static(batch: Batch, stream: int) -> None:
    check(load().dd_batch_run_static(batch.handle, stream or None))


def make_jobs(torch, specs):
    """Plans, inputs and output buffers of two-step (subbands > 1) and
    single-step jobs."""
    jobs = []
    for spec in specs:
        w = workloads.tiny(min_freq=1300.0, bandwidth=1.0, **spec)
        plan = Plan(w.obs, 128)
        plan.set_zapped_channels(zapped_flags(plan, w.zapped))
        z = plan.sizes
        if w.obs.nr_dms_subbanding:
            d_in = make_input(torch, plan, w, z.sub_in_pitch, z.nr_dispersed_samples_subbanding, w.seed)
            mid = torch.zeros(z.intermediate_elements, dtype=torch.float32, device="cuda")
            out = torch.zeros(z.sub_output_elements, dtype=torch.float32, device="cuda")
        else:
            d_in = make_input(torch, plan, w, z.in_pitch, z.nr_dispersed_samples, w.seed)
            mid = None
            out = torch.zeros(z.output_elements, dtype=torch.float32, device="cuda")
        jobs.append((plan, d_in, mid, out))
    return jobs


def batch_of(jobs):
    b = Batch()
    b.set_conf(DD_SINGLE_STEP, Conf(64, 4, 4, 2, 4, 1))
    b.set_conf(DD_STEP_ONE, Conf(64, 4, 8, 2, 4, 1))
    b.set_conf(DD_STEP_TWO, Conf(64, 2, 8, 4, 3, 1))
    b.set_pipeline(0)
    for plan, d_in, mid, out in jobs:
        if mid is None:
            b.add_single_step(plan, d_in.data_ptr(), out.data_ptr())
        else:
            b.add_two_step(plan, d_in.data_ptr(), mid.data_ptr(), out.data_ptr())
    return b


def outputs(torch, jobs):
    torch.cuda.synchronize()
    return [(None if mid is None else mid.cpu().numpy().view(np.uint16)[: mid.numel()].copy(), out.cpu().numpy())
            for _, _, mid, out in jobs]


def clear(torch, jobs):
    for _, _, mid, out in jobs:
        if mid is not None:
            mid.fill_(-3.0)
        out.fill_(-3.0)
    torch.cuda.synchronize()


BIG = [dict(channels=96, samples=2000, dms=24, step=0.6, subbands=8, coarse=12, coarse_step=4.0, beams=2, sbeams=3,
            zapped=(0, 50), seed=201, first=1.7, coarse_first=2.9),
       dict(channels=64, samples=1500, dms=16, step=0.8, subbands=4, coarse=6, coarse_step=5.0, seed=202),
       dict(channels=128, samples=1800, dms=40, step=0.5, zapped=(3,), seed=203, first=4.4)]
SMALL = [dict(channels=48, samples=400, dms=8, step=1.0, subbands=4, coarse=3, coarse_step=3.0, seed=204),
         dict(channels=32, samples=300, dms=5, step=1.2, seed=205)]


def test_static_run_equals_batch_run_twice_in_a_row(torch_cuda):
    torch = torch_cuda
    s = stream_of(torch)
    results = []
    for specs in (BIG, SMALL):
        jobs = make_jobs(torch, specs)
        with batch_of(jobs) as b:
            # both runs start from the same fill: row padding is not written
            clear(torch, jobs)
            b.run(use_graph=False, stream=s)
            want = outputs(torch, jobs)
            clear(torch, jobs)
            run_static(b, s)
            got = outputs(torch, jobs)
            info = b.info()
        results.append((jobs, want, got, info))
    # the second call left the smaller batch's queues in the shared array;
    # each call's outputs equal the allocating run of its own batch
    for jobs, want, got, info in results:
        assert info["solo_jobs"] == 0
        for j, ((wm, wo), (gm, go)) in enumerate(zip(want, got)):
            np.testing.assert_array_equal(go, wo, err_msg=f"job {j} output")
            if wm is not None:
                np.testing.assert_array_equal(gm, wm, err_msg=f"job {j} intermediate")
            assert np.all(go[: 16] != -3.0)


def test_static_run_refusals(torch_cuda):
    torch = torch_cuda
    s = stream_of(torch)
    # 2 channels, 65536 samples, 256 DMs in tiles of 256 samples x 1 DM:
    # 65536 tiles (512 KiB of descriptors) per job; nine jobs exceed the array
    w = workloads.tiny(channels=2, samples=65536, dms=256, step=0.01, min_freq=1300.0, bandwidth=1.0, seed=206)
    plan = Plan(w.obs, 128)
    buf = torch.zeros(1 << 20, dtype=torch.float32, device="cuda")
    with Batch() as b:
        b.set_conf(DD_SINGLE_STEP, Conf(64, 1, 4, 1, 1, 1))
        b.set_pipeline(0)
        for _ in range(9):
            b.add_single_step(plan, buf.data_ptr(), buf.data_ptr())
        assert b.info()["tiles"][DD_SINGLE_STEP] == 9 * 65536
        with pytest.raises(DDError, match="bytes"):
            run_static(b, s)
    # a job too wide for the batch kernels (it would run beside them)
    wide = workloads.tiny(channels=64, samples=2000, dms=12, step=20.0, min_freq=300.0, bandwidth=2.0, seed=207)
    jobs = make_jobs(torch, [dict(channels=64, samples=700, dms=12, step=0.9, seed=208)])
    wplan = Plan(wide.obs, 128)
    with batch_of(jobs) as b:
        b.set_conf(DD_SINGLE_STEP, Conf(64, 4, 4, 2, 4, 1))
        b.add_single_step(wplan, buf.data_ptr(), buf.data_ptr())
        assert b.info()["solo_jobs"] == 1
        with pytest.raises(DDError, match="do not fit"):
            run_static(b, s)
