"""Degenerate inputs on the device against the oracle, bit for bit: every
channel zapped (every sum is zero, as the reference's loops over no channel
leave it), a subband whose channels are all zapped, minimal shapes (one DM,
one output sample, two channels), a batch with no jobs, and shift spreads
too wide for any LDS window."""
import numpy as np
import pytest

import oracle
from dedispersion_amd import DD_SINGLE_STEP, DD_STEP_ONE, DD_STEP_TWO, Batch, Conf, Plan, workloads
from tests.conftest import host_threads
from tests.gpu_helpers import compare_rows, make_input, single_step, stream_of, zapped_flags

pytestmark = pytest.mark.gpu

CONFS = [None, Conf(64, 4, 4, 2, 4, 1), Conf(64, 2, 16, 4, 8, 0), Conf(64, 2, 8, 2, 2, 2)]
CONF_IDS = ["default", "staged", "local0", "local2"]


def two_step(torch, w, padding=128, conf_one=None, conf_two=None):
    """Step one then step two on the device; returns (plan, flags, host
    input, host intermediate, host output, oracle)."""
    plan = Plan(w.obs, padding)
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    if conf_one is not None:
        plan.set_conf(DD_STEP_ONE, conf_one)
    if conf_two is not None:
        plan.set_conf(DD_STEP_TWO, conf_two)
    z = plan.sizes
    d_in = make_input(torch, plan, w, z.sub_in_pitch, z.nr_dispersed_samples_subbanding, w.seed)
    d_mid = torch.full((z.intermediate_elements,), -1.0, dtype=torch.float32, device="cuda")
    d_out = torch.full((z.sub_output_elements,), -1.0, dtype=torch.float32, device="cuda")
    s = stream_of(torch)
    plan.step_one(d_in.data_ptr(), d_mid.data_ptr(), stream=s)
    plan.step_two(d_mid.data_ptr(), d_out.data_ptr(), stream=s)
    torch.cuda.synchronize()
    orc = oracle.Oracle(w.obs, padding, threads=host_threads())
    return plan, flags, d_in.cpu().numpy(), d_mid.cpu().numpy(), d_out.cpu().numpy(), orc


@pytest.mark.parametrize("conf", CONFS, ids=CONF_IDS)
def test_single_step_all_channels_zapped(torch_cuda, conf):
    w = workloads.tiny(channels=40, samples=300, dms=5, step=1.0, min_freq=1300.0, bandwidth=1.0,
                       zapped=tuple(range(40)), seed=3)
    plan, inp, out, orc, flags = single_step(torch_cuda, w, conf=conf)
    want = orc.dedispersion(inp, flags)
    z = plan.sizes
    assert not want[..., :z.nr_samples].any()
    compare_rows(out.reshape(want.shape), want, z.nr_samples, "all channels zapped")
    plan.close()


@pytest.mark.parametrize("conf", CONFS[:3], ids=CONF_IDS[:3])
def test_two_step_with_a_zapped_subband(torch_cuda, conf):
    """Subband 1 (channels 16-31) has no live channel: its step-one rows are
    zero and step two adds them like any other subband."""
    w = workloads.tiny(channels=64, samples=400, dms=6, step=0.5, subbands=4, coarse=3, coarse_step=3.0,
                       min_freq=1300.0, bandwidth=1.0, zapped=tuple(range(16, 32)) + (40,), seed=9)
    plan, flags, inp, mid, out, orc = two_step(torch_cuda, w, conf_one=conf)
    z = plan.sizes
    want_mid = orc.step_one(inp, flags)
    mid = mid.reshape(want_mid.shape)
    compare_rows(mid, want_mid, z.nr_samples_subbanding, "step one")
    assert not want_mid[:, :, 1, :z.nr_samples_subbanding].any()
    want = orc.step_two(mid.ravel())
    compare_rows(out.reshape(want.shape), want, w.obs.nr_samples_per_batch, "step two")
    plan.close()


@pytest.mark.parametrize("samples", [1, 3, 64])
def test_single_step_minimal_shapes(torch_cuda, samples):
    """Two channels, one DM, a handful of output samples."""
    w = workloads.tiny(channels=2, samples=samples, dms=1, step=0.0, seed=4)
    plan, inp, out, orc, flags = single_step(torch_cuda, w)
    want = orc.dedispersion(inp, flags)
    compare_rows(out.reshape(want.shape), want, plan.sizes.nr_samples, f"{samples} samples")
    plan.close()


@pytest.mark.parametrize("samples", [1, 5])
def test_two_step_minimal_shapes(torch_cuda, samples):
    """Two channels in one subband, one coarse and one fine DM."""
    w = workloads.tiny(channels=2, samples=samples, dms=1, step=0.0, subbands=1, coarse=1, coarse_step=0.0, seed=6)
    plan, flags, inp, mid, out, orc = two_step(torch_cuda, w)
    z = plan.sizes
    want_mid = orc.step_one(inp, flags)
    compare_rows(mid.reshape(want_mid.shape), want_mid, z.nr_samples_subbanding, "step one")
    want = orc.step_two(mid)
    compare_rows(out.reshape(want.shape), want, samples, "step two")
    plan.close()


def test_empty_batch(torch_cuda):
    with Batch() as b:
        assert b.info()["jobs"] == 0
        b.run()
        b.run(use_graph=True)
        b.run_per_job()
        for mode in (DD_SINGLE_STEP, DD_STEP_ONE, DD_STEP_TWO):
            b.run_mode(mode)
    torch_cuda.cuda.synchronize()


def test_single_step_wide_shift_spread(torch_cuda):
    """300 MHz at the bottom of a 128 MHz band and DM steps of 20: the shifts
    of neighbouring DMs are thousands of samples apart, far wider than any
    LDS window.  Every configuration runs, the LDS ones as the direct kernel
    of the same shape (rows read through the caches), and is exact."""
    w = workloads.tiny(channels=64, samples=2000, dms=12, step=20.0, min_freq=300.0, bandwidth=2.0,
                       zapped=(5,), seed=12)
    for conf in (None, Conf(64, 2, 8, 2, 4, 0), Conf(64, 4, 16, 8, 4, 1), Conf(64, 2, 8, 4, 2, 2)):
        plan, inp, out, orc, flags = single_step(torch_cuda, w, conf=conf)
        assert plan.sizes.nr_dispersed_samples > 30000
        want = orc.dedispersion(inp, flags)
        compare_rows(out.reshape(want.shape), want, plan.sizes.nr_samples, f"wide spread {conf}")
        plan.close()


def test_two_step_wide_shift_spread(torch_cuda):
    """The same band in the two-step path: coarse DM steps of 40 spread step
    one's shifts beyond its LDS window, fine steps of 10 spread step two's."""
    w = workloads.tiny(channels=64, samples=1000, dms=4, step=10.0, subbands=4, coarse=4, coarse_step=40.0,
                       min_freq=300.0, bandwidth=2.0, zapped=(5,), seed=13)
    plan, flags, inp, mid, out, orc = two_step(torch_cuda, w)
    z = plan.sizes
    want_mid = orc.step_one(inp, flags)
    compare_rows(mid.reshape(want_mid.shape), want_mid, z.nr_samples_subbanding, "step one")
    want = orc.step_two(mid)
    compare_rows(out.reshape(want.shape), want, w.obs.nr_samples_per_batch, "step two")
    plan.close()
