"""Plans configured from the packaged tuned-configuration files (the
reference's tunedDedispersionConf lookup by device and DM count) run and
match the oracle."""
import pytest

import oracle
from dedispersion_amd import DD_SINGLE_STEP, DD_STEP_ONE, DD_STEP_TWO, Plan, tuner, workloads
from tests.conftest import host_threads
from tests.gpu_helpers import compare_rows, make_input, single_step, stream_of, zapped_flags

pytestmark = pytest.mark.gpu


def test_two_step_with_tuned_files(torch_cuda):
    torch = torch_cuda
    w = workloads.arts_two_step(16, 8192, seed=8)
    plan = Plan(w.obs, 128)
    one = plan.apply_tuned(DD_STEP_ONE, tuner.read_tuned_file(tuner.tuned_file_path("step_one")))
    two = plan.apply_tuned(DD_STEP_TWO, tuner.read_tuned_file(tuner.tuned_file_path("step_two")))
    assert plan.get_conf(DD_STEP_ONE) == one and plan.get_conf(DD_STEP_TWO) == two
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
    orc = oracle.Oracle(w.obs, 128, threads=host_threads())
    want_mid = orc.step_one(d_in.cpu().numpy(), flags)
    compare_rows(d_mid.cpu().numpy().reshape(want_mid.shape), want_mid, z.nr_samples_subbanding, "tuned step one")
    rows = (0, 2 * w.obs.nr_dms)
    want = orc.step_two(want_mid.ravel(), rows=rows)
    got = d_out.cpu().numpy().reshape(want.shape)
    compare_rows(got[:, : rows[1]], want[:, : rows[1]], w.obs.nr_samples_per_batch, "tuned step two")
    plan.close()


def test_single_step_with_tuned_file(torch_cuda):
    w = workloads.plain_128()
    tuned = tuner.read_tuned_file(tuner.tuned_file_path("single_step"))
    conf = tuner.select_tuned_conf(tuned, "MI355X", w.obs.nr_dms)
    plan, inp, out, orc, flags = single_step(torch_cuda, w, padding=128, conf=conf)
    want = orc.dedispersion(inp, flags)
    compare_rows(out.reshape(want.shape), want, w.obs.nr_samples_per_batch, "tuned single step")
    plan.close()
