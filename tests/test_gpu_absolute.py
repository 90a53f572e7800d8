"""Step one with each channel's own delay (dd_plan_set_step_one_shifts
DD_STEP_ONE_ABSOLUTE), the delay of the reference's OpenCL step-one kernel
(include/Dedispersion.hpp:595-597, 731, 733), against the oracle's absolute
branch; and the physics it restores: with absolute delays step one plus step
two dedisperses at DM coarse + fine, so the ARTS two-step beam's injected
pulses are found at their DMs, which the CPU function's relative delays
(:124) miss beyond coarse DM 0."""
import numpy as np
import pytest

import oracle
from dedispersion_amd import DD_STEP_ONE, DD_STEP_TWO, Conf, Plan, workloads
from tests.conftest import host_threads
from tests.gpu_helpers import compare_rows, make_input, stream_of, zapped_flags
from tests.test_oracle import int_shifts

pytestmark = pytest.mark.gpu


def test_absolute_integer_table_and_back(torch_cuda):
    w = workloads.tiny(channels=96, samples=500, dms=12, step=0.9, subbands=8, coarse=9, coarse_step=4.5,
                       coarse_first=3.1, min_freq=1300.0, bandwidth=1.0)
    plan = Plan(w.obs, 128)
    s, _ = plan.get_shifts()
    rel = plan.get_integer_shifts(DD_STEP_ONE)
    plan.set_step_one_shifts(True)
    np.testing.assert_array_equal(plan.get_integer_shifts(DD_STEP_ONE), int_shifts(s[:96], 3.1, 4.5, 9))
    assert plan.sizes.max_shift_step_one > rel.max()
    plan.set_step_one_shifts(False)
    np.testing.assert_array_equal(plan.get_integer_shifts(DD_STEP_ONE), rel)


@pytest.mark.parametrize("conf", [None, Conf(64, 4, 4, 2, 4, 1), Conf(64, 2, 8, 4, 3, 2), Conf(64, 1, 16, 8, 2, 0)],
                         ids=["default", "staged", "local2", "local0"])
@pytest.mark.parametrize("bits", [8, 4], ids=["8bit", "4bit"])
def test_absolute_step_one(torch_cuda, conf, bits):
    torch = torch_cuda
    w = workloads.tiny(channels=96, samples=700, dms=12, step=0.9, first=1.3, subbands=8, coarse=9, coarse_step=4.5,
                       coarse_first=2.1, beams=2, sbeams=3, min_freq=1300.0, bandwidth=1.0, zapped=(0, 11, 50),
                       seed=97)
    plan = Plan(w.obs, 128, input_bits=bits)
    plan.set_step_one_shifts(True)
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    if conf is not None:
        plan.set_conf(DD_STEP_ONE, conf)
    z = plan.sizes
    orc = oracle.Oracle(w.obs, 128, threads=host_threads(), input_bits=bits, step_one_absolute=True)
    if bits == 8:
        d_in = make_input(torch, plan, w, z.sub_in_pitch, z.nr_dispersed_samples_subbanding, w.seed)
        host = d_in.cpu().numpy()
    else:
        from tests.test_gpu_packed import packed_input
        host, d_in = packed_input(torch, orc, bits, z.nr_dispersed_samples_subbanding, z.sub_in_pitch, seed=w.seed)
    d_mid = torch.full((z.intermediate_elements,), -1.0, dtype=torch.float32, device="cuda")
    d_out = torch.full((z.sub_output_elements,), -1.0, dtype=torch.float32, device="cuda")
    s = stream_of(torch)
    plan.step_one(d_in.data_ptr(), d_mid.data_ptr(), stream=s)
    plan.step_two(d_mid.data_ptr(), d_out.data_ptr(), stream=s)
    torch.cuda.synchronize()
    want_mid = orc.step_one(host, flags)
    mid = d_mid.cpu().numpy().reshape(want_mid.shape)
    compare_rows(mid, want_mid, z.nr_samples_subbanding, f"absolute step one, {bits} bits, {conf}")
    want = orc.step_two(mid.ravel())
    compare_rows(d_out.cpu().numpy().reshape(want.shape), want, w.obs.nr_samples_per_batch, "step two after it")
    plan.close()


def recovered(out, pulses, live, S):
    """Pulses whose window (widened by the two-step rounding, a few samples)
    stands out of its DM row by a fifth of the full pulse sum."""
    found = []
    for p in pulses:
        row = out[p.dm_index, :S]
        peak = row[max(0, p.t0 - 8):p.t0 + p.width + 8].max()
        found.append(bool(peak > np.median(row) + 0.2 * live * p.amplitude))
    return found


def test_two_step_absolute_recovers_the_arts_pulses(torch_cuda):
    """The ARTS two-step beam (1536 channels, 64 coarse DMs of 6.4 x 32 fine
    DMs of 0.2, 8 dispersed pulses at seeded DMs): with absolute step-one
    delays the whole output matches the oracle and every pulse is found at
    its DM row coarse * 32 + fine; with the relative delays the pulses at
    coarse DM 2 and above are not."""
    torch = torch_cuda
    w = workloads.arts_two_step()
    o = w.obs
    S = o.nr_samples_per_batch
    live = o.nr_channels - len(w.zapped)
    outs = {}
    for absolute in (True, False):
        plan = Plan(o, 128)
        plan.set_step_one_shifts(absolute)
        flags = zapped_flags(plan, w.zapped)
        plan.set_zapped_channels(flags)
        z = plan.sizes
        d_in = make_input(torch, plan, w, z.sub_in_pitch, z.nr_dispersed_samples_subbanding, w.seed)
        d_mid = torch.full((z.intermediate_elements,), -1.0, dtype=torch.float32, device="cuda")
        d_out = torch.full((z.sub_output_elements,), -1.0, dtype=torch.float32, device="cuda")
        s = stream_of(torch)
        plan.step_one(d_in.data_ptr(), d_mid.data_ptr(), stream=s)
        plan.step_two(d_mid.data_ptr(), d_out.data_ptr(), stream=s)
        torch.cuda.synchronize()
        out = d_out.cpu().numpy().reshape(o.nr_dms_subbanding * o.nr_dms, z.out_pitch)
        if absolute:
            orc = oracle.Oracle(o, 128, threads=host_threads(), step_one_absolute=True)
            want_mid = orc.step_one(d_in.cpu().numpy(), flags)
            compare_rows(d_mid.cpu().numpy().reshape(want_mid.shape), want_mid, z.nr_samples_subbanding,
                         "ARTS absolute step one")
            want = orc.step_two(want_mid.ravel())
            compare_rows(out[None], want, S, "ARTS absolute two-step output")
        outs[absolute] = out
        plan.close()
    assert all(recovered(outs[True], w.pulses, live, S)), [(p.dm_index, p.t0, p.width) for p in w.pulses]
    far = [p for p in w.pulses if p.dm_index // o.nr_dms >= 2]
    assert far and not any(recovered(outs[False], far, live, S))
