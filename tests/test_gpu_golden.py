"""The device path against the committed golden fixtures (tests/golden/)."""
import json
import os

import numpy as np
import pytest

from dedispersion_amd import Plan
from tests.golden import make_fixtures as mf
from tests.gpu_helpers import compare_rows, stream_of

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_tiny_single_step_matches_fixture(torch_cuda):
    torch = torch_cuda
    z = np.load(os.path.join(HERE, "tiny_single.npz"))
    w, _ = mf.tiny_single()
    plan = Plan(w.obs, 0)
    plan.set_zapped_channels(z["zapped"])
    plan.set_beam_mapping(single=z["mapping"])
    d_in = torch.from_numpy(z["input"]).cuda()
    d_out = torch.full((plan.sizes.output_elements,), -1.0, dtype=torch.float32, device="cuda")
    plan.dedisperse(d_in.data_ptr(), d_out.data_ptr(), stream=stream_of(torch))
    want = z["output"]
    compare_rows(d_out.cpu().numpy().reshape(want.shape), want, w.obs.nr_samples_per_batch, "tiny single fixture")


def test_tiny_two_step_matches_fixture(torch_cuda):
    torch = torch_cuda
    z = np.load(os.path.join(HERE, "tiny_two_step.npz"))
    w, _ = mf.tiny_two_step()
    plan = Plan(w.obs, 0)
    plan.set_zapped_channels(z["zapped"])
    s = plan.sizes
    d_in = torch.from_numpy(z["input"]).cuda()
    d_mid = torch.full((s.intermediate_elements,), -1.0, dtype=torch.float32, device="cuda")
    d_out = torch.full((s.sub_output_elements,), -1.0, dtype=torch.float32, device="cuda")
    plan.step_one(d_in.data_ptr(), d_mid.data_ptr(), stream=stream_of(torch))
    plan.step_two(d_mid.data_ptr(), d_out.data_ptr(), stream=stream_of(torch))
    mid, out = z["intermediate"], z["output"]
    compare_rows(d_mid.cpu().numpy().reshape(mid.shape), mid, s.nr_samples_subbanding, "tiny step one fixture")
    compare_rows(d_out.cpu().numpy().reshape(out.shape), out, w.obs.nr_samples_per_batch, "tiny step two fixture")


def test_plain_128_matches_digest(torch_cuda):
    torch = torch_cuda
    want = json.load(open(os.path.join(HERE, "plain_128.json")))
    w, inp, flags, _ = mf.plain_128()
    plan = Plan(w.obs, 128)
    plan.set_zapped_channels(flags)
    d_in = torch.from_numpy(inp).cuda()
    d_out = torch.empty(plan.sizes.output_elements, dtype=torch.float32, device="cuda")
    plan.dedisperse(d_in.data_ptr(), d_out.data_ptr(), stream=stream_of(torch))
    got = mf.plain_128_digest(d_out.cpu().numpy().reshape(w.obs.nr_dms, -1), w.obs.nr_samples_per_batch)
    assert got["slices"] == want["slices"]
    assert got["sha256"] == want["sha256"]
