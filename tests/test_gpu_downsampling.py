"""Downsampled observations (getDownsampling() > 1) against the oracle.

The single step computes S / downsampling output samples per DM row with the
shift table scaled by S / downsampling (include/Dedispersion.hpp:76,
src/Shifts.cpp:26), reading the input rows sized from S (the test driver's
S_disp, src/DedispersionTest.cpp:129).  In the subband path each step
follows its own CPU function: step one writes S1 / downsampling samples into
rows of pad(S1) (include/Dedispersion.hpp:119, 148), step two reads rows of
pad(S1 / downsampling) (:163, 169).  The two layouts only agree at
downsampling 1, so each step is checked on its own input; batches and
streams, which chain them, refuse downsampling > 1."""
import numpy as np
import pytest

import oracle
from dedispersion_amd import DD_STEP_ONE, DD_STEP_TWO, Batch, Conf, DDError, Plan, workloads
from tests.conftest import host_threads
from tests.gpu_helpers import compare_rows, make_input, single_step, stream_of, zapped_flags

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("conf", [None, Conf(64, 4, 4, 2, 4, 1), Conf(64, 2, 8, 4, 4, 0)],
                         ids=["default", "staged", "local0"])
@pytest.mark.parametrize("ds", [2, 3, 8])
def test_single_step_downsampled(torch_cuda, ds, conf):
    w = workloads.tiny(channels=80, samples=1200, dms=21, step=0.7, beams=2, sbeams=2, min_freq=1300.0,
                       bandwidth=1.0, zapped=(1, 40), seed=61)
    w.obs.downsampling = ds
    plan, inp, out, orc, flags = single_step(torch_cuda, w, padding=128, conf=conf, seed=61 + ds)
    z = plan.sizes
    assert z.nr_samples == 1200 // ds
    assert z.nr_dispersed_samples == orc.o.nr_samples_per_dispersed_batch
    want = orc.dedispersion(inp, flags)
    compare_rows(out.reshape(want.shape), want, z.nr_samples, f"downsampling {ds}")
    plan.close()


def downsampled_two_step(ds, first=0.0, coarse_first=0.0):
    """A two-step observation at downsampling ds whose S1 leaves room for
    step two's reads: S / ds output samples plus the largest step-two shift
    within S1 / ds (the derived S1 adds the delay in downsampled samples to
    S, which is short by a factor ds)."""
    w = workloads.tiny(channels=96, samples=1100, dms=14, step=0.8, subbands=8, coarse=6, coarse_step=5.0, beams=2,
                       sbeams=3, min_freq=1300.0, bandwidth=1.0, zapped=(2, 50, 51), seed=63, first=first,
                       coarse_first=coarse_first)
    w.obs.downsampling = ds
    probe = oracle.Oracle(w.obs, 0)
    top = int(probe.shifts2[0] * (first + 13 * 0.8)) + 2
    w.obs.nr_samples_per_batch_subbanding = ds * (1100 // ds + top)
    return w


@pytest.mark.parametrize("confs", [(None, None), (Conf(64, 4, 4, 2, 4, 1), Conf(64, 2, 8, 4, 3, 1)),
                                   (Conf(64, 2, 8, 4, 3, 2), Conf(64, 1, 16, 2, 2, 0))],
                         ids=["default", "staged", "local2-local0"])
@pytest.mark.parametrize("ds", [2, 3])
def test_two_step_downsampled(torch_cuda, ds, confs):
    """Step one against the oracle on the device input; step two against the
    oracle on general floats laid out at pad(S1 / ds)."""
    torch = torch_cuda
    w = downsampled_two_step(ds, first=2.3, coarse_first=1.7)
    plan = Plan(w.obs, 128)
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    if confs[0] is not None:
        plan.set_conf(DD_STEP_ONE, confs[0])
        plan.set_conf(DD_STEP_TWO, confs[1])
    z = plan.sizes
    orc = oracle.Oracle(w.obs, 128, threads=host_threads())
    assert z.nr_samples == 1100 // ds and z.nr_samples_step_one == orc.o.nr_samples_per_batch_sub // ds
    assert z.sub_out_pitch == orc.sub_out_pitch and z.step_two_in_pitch == orc.sub_two_in_pitch
    assert z.step_two_in_pitch < z.sub_out_pitch
    d_in = make_input(torch, plan, w, z.sub_in_pitch, z.nr_dispersed_samples_subbanding, w.seed + ds)
    d_mid = torch.full((z.intermediate_elements,), -1.0, dtype=torch.float32, device="cuda")
    s = stream_of(torch)
    plan.step_one(d_in.data_ptr(), d_mid.data_ptr(), stream=s)
    torch.cuda.synchronize()
    want = orc.step_one(d_in.cpu().numpy(), flags)
    got = d_mid.cpu().numpy().reshape(want.shape)
    compare_rows(got, want, z.nr_samples_step_one, f"step one, downsampling {ds}")
    assert np.all(got[..., z.nr_samples_step_one:] == -1.0)
    rng = np.random.default_rng(ds)
    x = (rng.standard_normal(z.step_two_input_elements) * 300.0).astype(np.float32)
    d_x = torch.from_numpy(x).cuda()
    d_out = torch.full((z.sub_output_elements,), -1.0, dtype=torch.float32, device="cuda")
    plan.step_two(d_x.data_ptr(), d_out.data_ptr(), stream=s)
    torch.cuda.synchronize()
    want = orc.step_two(x)
    compare_rows(d_out.cpu().numpy().reshape(want.shape), want, z.nr_samples, f"step two, downsampling {ds}")
    plan.close()


def test_two_step_downsampled_not_batched(torch_cuda):
    """A batch chains step one into step two: it refuses a two-step plan at
    downsampling > 1, whose two intermediate pitches differ."""
    w = downsampled_two_step(2)
    plan = Plan(w.obs, 128)
    buf = torch_cuda.zeros(16, dtype=torch_cuda.float32, device="cuda")
    with Batch() as b:
        with pytest.raises(DDError, match="downsampling"):
            b.add_two_step(plan, buf.data_ptr(), buf.data_ptr(), buf.data_ptr())
    # a derived S1 too short for step two's reads is refused by the launch
    w.obs.nr_samples_per_batch_subbanding = 0
    short = Plan(w.obs, 128)
    with pytest.raises(DDError, match="shorter"):
        short.step_two(buf.data_ptr(), buf.data_ptr())
