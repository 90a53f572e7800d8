"""Packed input (the reference's inputBits 1, 2 and 4) on the device against
the oracle, bit for bit.

Inputs are random values over the whole field (top bit included), packed on
the host by tests/packing.py; the device unpacks them while staging rows.
Cases: several beams with a mixed beam mapping and zapped channels, padded
and unpadded rows (odd byte offsets), every kernel selection (local 0, 1, 3
all run the pipelined kernel for packed rows), more than 257 live channels
(the kernels with 32-bit flush sums), the two-step path, an ARTS-sized
two-step beam, and batches mixing packed and 8-bit jobs."""
import numpy as np
import pytest

import oracle
from dedispersion_amd import DD_SINGLE_STEP, DD_STEP_ONE, DD_STEP_TWO, Batch, Conf, DDError, Plan, workloads
from tests.conftest import host_threads
from tests.gpu_helpers import compare_rows, stream_of, zapped_flags
from tests.packing import pack_rows, random_values

pytestmark = pytest.mark.gpu

BITS = [1, 2, 4]
STAGED_SMALL = Conf(64, 4, 4, 2, 4, 1)
STAGED_STEP_TWO = Conf(64, 2, 8, 4, 3, 1)


def packed_input(torch, orc, bits, rows_len, pitch, seed):
    rng = np.random.default_rng(seed)
    rows = orc.o.nr_beams * orc.o.nr_channels
    host = pack_rows(random_values(rng, rows, rows_len, bits), bits, pitch).ravel()
    return host, torch.from_numpy(host).cuda()


def single_workload():
    return workloads.tiny(channels=96, samples=1000, dms=24, step=0.9, beams=2, sbeams=3, min_freq=1300.0,
                          bandwidth=1.0, zapped=(0, 5, 50), seed=41)


def two_step_workload(seed=31):
    return workloads.tiny(channels=96, samples=700, dms=12, step=0.9, subbands=8, coarse=9, coarse_step=4.5, beams=2,
                          sbeams=3, zapped=(0, 11, 12, 50, 95), min_freq=1300.0, bandwidth=1.0, seed=seed)


@pytest.mark.parametrize("conf", [None, Conf(64, 4, 4, 2, 4, 1), Conf(64, 2, 16, 4, 8, 0), Conf(64, 4, 8, 8, 4, 3)],
                         ids=["default", "staged", "local0", "local3"])
@pytest.mark.parametrize("padding", [128, 0], ids=["pad128", "unpadded"])
@pytest.mark.parametrize("bits", BITS, ids=lambda b: f"{b}bit")
def test_single_step_packed(torch_cuda, bits, padding, conf):
    torch = torch_cuda
    w = single_workload()
    plan = Plan(w.obs, padding, input_bits=bits)
    orc = oracle.Oracle(w.obs, padding, threads=host_threads(), input_bits=bits)
    z = plan.sizes
    assert z.input_bits == bits and z.in_pitch == orc.in_pitch
    assert z.input_elements == w.obs.nr_beams * w.obs.nr_channels * orc.in_pitch
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    mapping = orc.default_map(96)
    mapping[2, :40] = 0
    mapping[0, 60:] = 1
    plan.set_beam_mapping(single=mapping)
    if conf is not None:
        plan.set_conf(DD_SINGLE_STEP, conf)
    host, d_in = packed_input(torch, orc, bits, z.nr_dispersed_samples, z.in_pitch, seed=bits)
    d_out = torch.full((z.output_elements,), -1.0, dtype=torch.float32, device="cuda")
    plan.dedisperse(d_in.data_ptr(), d_out.data_ptr(), stream=stream_of(torch))
    torch.cuda.synchronize()
    want = orc.dedispersion(host, flags, mapping)
    compare_rows(d_out.cpu().numpy().reshape(want.shape), want, z.nr_samples, f"{bits}-bit single step")
    plan.close()


@pytest.mark.parametrize("conf", [Conf(64, 2, 16, 8, 4, 1), Conf(64, 4, 8, 2, 4, 1)], ids=["16x8", "8x2"])
@pytest.mark.parametrize("bits", BITS, ids=lambda b: f"{b}bit")
def test_single_step_packed_many_channels(torch_cuda, bits, conf):
    """More than 257 live channels: the kernels that flush their 16-bit pair
    sums into 32-bit sums every 256 rows, over packed rows."""
    torch = torch_cuda
    w = workloads.tiny(channels=320, samples=1100, dms=16, step=0.4, beams=1, sbeams=1, min_freq=1300.0,
                       bandwidth=0.5, zapped=(3, 200), seed=43)
    plan = Plan(w.obs, 128, input_bits=bits)
    orc = oracle.Oracle(w.obs, 128, threads=host_threads(), input_bits=bits)
    z = plan.sizes
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    plan.set_conf(DD_SINGLE_STEP, conf)
    host, d_in = packed_input(torch, orc, bits, z.nr_dispersed_samples, z.in_pitch, seed=10 + bits)
    d_out = torch.full((z.output_elements,), -1.0, dtype=torch.float32, device="cuda")
    plan.dedisperse(d_in.data_ptr(), d_out.data_ptr(), stream=stream_of(torch))
    torch.cuda.synchronize()
    want = orc.dedispersion(host, flags)
    compare_rows(d_out.cpu().numpy().reshape(want.shape), want, z.nr_samples, f"{bits}-bit single step, 320 channels")
    plan.close()


@pytest.mark.parametrize("bits", BITS, ids=lambda b: f"{b}bit")
def test_two_step_packed(torch_cuda, bits):
    torch = torch_cuda
    w = two_step_workload()
    plan = Plan(w.obs, 128, input_bits=bits)
    orc = oracle.Oracle(w.obs, 128, threads=host_threads(), input_bits=bits)
    z = plan.sizes
    assert z.sub_in_pitch == orc.sub_in_pitch and z.sub_out_pitch == orc.sub_out_pitch
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    host, d_in = packed_input(torch, orc, bits, z.nr_dispersed_samples_subbanding, z.sub_in_pitch, seed=10 + bits)
    d_mid = torch.full((z.intermediate_elements,), -1.0, dtype=torch.float32, device="cuda")
    d_out = torch.full((z.sub_output_elements,), -1.0, dtype=torch.float32, device="cuda")
    s = stream_of(torch)
    plan.step_one(d_in.data_ptr(), d_mid.data_ptr(), stream=s)
    plan.step_two(d_mid.data_ptr(), d_out.data_ptr(), stream=s)
    torch.cuda.synchronize()
    want_mid = orc.step_one(host, flags)
    compare_rows(d_mid.cpu().numpy().reshape(want_mid.shape), want_mid, z.nr_samples_subbanding,
                 f"{bits}-bit step one")
    want = orc.step_two(want_mid.ravel())
    compare_rows(d_out.cpu().numpy().reshape(want.shape), want, w.obs.nr_samples_per_batch, f"{bits}-bit step two")
    plan.close()


@pytest.mark.parametrize("bits", [2, 4], ids=lambda b: f"{b}bit")
def test_arts_two_step_packed(torch_cuda, bits):
    """1536 channels, 32 subbands, 16 x 32 DMs, 8192 samples, default
    configurations."""
    torch = torch_cuda
    w = workloads.arts_two_step(16, 8192, seed=7)
    plan = Plan(w.obs, 128, input_bits=bits)
    orc = oracle.Oracle(w.obs, 128, threads=host_threads(), input_bits=bits)
    z = plan.sizes
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    host, d_in = packed_input(torch, orc, bits, z.nr_dispersed_samples_subbanding, z.sub_in_pitch, seed=20 + bits)
    d_mid = torch.full((z.intermediate_elements,), -1.0, dtype=torch.float32, device="cuda")
    d_out = torch.full((z.sub_output_elements,), -1.0, dtype=torch.float32, device="cuda")
    s = stream_of(torch)
    plan.step_one(d_in.data_ptr(), d_mid.data_ptr(), stream=s)
    plan.step_two(d_mid.data_ptr(), d_out.data_ptr(), stream=s)
    torch.cuda.synchronize()
    want_mid = orc.step_one(host, flags)
    compare_rows(d_mid.cpu().numpy().reshape(want_mid.shape), want_mid, z.nr_samples_subbanding, "ARTS step one")
    rows = (0, 3 * w.obs.nr_dms)  # three coarse DMs of step two
    want = orc.step_two(want_mid.ravel(), rows=rows)
    got = d_out.cpu().numpy().reshape(want.shape)
    compare_rows(got[:, : rows[1]], want[:, : rows[1]], w.obs.nr_samples_per_batch, "ARTS step two")
    plan.close()


@pytest.mark.parametrize("chunks", [0, 3], ids=["persistent", "pipelined"])
@pytest.mark.parametrize("compact", [True, False], ids=["uint16-intermediate", "float-intermediate"])
def test_batch_mixes_packed_and_byte_jobs(torch_cuda, compact, chunks):
    torch = torch_cuda
    s = stream_of(torch)
    jobs = []
    for j, bits in enumerate((2, 8, 1, 4)):
        w = two_step_workload(seed=50 + j)
        plan = Plan(w.obs, 0, input_bits=bits)
        orc = oracle.Oracle(w.obs, 0, threads=host_threads(), input_bits=bits)
        flags = zapped_flags(plan, w.zapped)
        plan.set_zapped_channels(flags)
        z = plan.sizes
        host, d_in = packed_input(torch, orc, bits, z.nr_dispersed_samples_subbanding, z.sub_in_pitch, seed=60 + j)
        want_mid = orc.step_one(host, flags)
        jobs.append(dict(plan=plan, d_in=d_in, want_mid=want_mid, want=orc.step_two(want_mid.ravel()),
                         mid=torch.full((z.intermediate_elements,), -1.0, dtype=torch.float32, device="cuda"),
                         out=torch.full((z.sub_output_elements,), -1.0, dtype=torch.float32, device="cuda")))
    w1 = single_workload()
    plan1 = Plan(w1.obs, 128, input_bits=4)
    orc1 = oracle.Oracle(w1.obs, 128, threads=host_threads(), input_bits=4)
    host1, d_in1 = packed_input(torch, orc1, 4, plan1.sizes.nr_dispersed_samples, plan1.sizes.in_pitch, seed=70)
    want1 = orc1.dedispersion(host1)
    out1 = torch.full((plan1.sizes.output_elements,), -1.0, dtype=torch.float32, device="cuda")
    with Batch() as b:
        b.set_conf(DD_SINGLE_STEP, STAGED_SMALL)
        b.set_conf(DD_STEP_ONE, STAGED_SMALL)
        b.set_conf(DD_STEP_TWO, STAGED_STEP_TWO)
        b.set_compact_intermediate(compact)
        b.set_pipeline(chunks)
        for j in jobs:
            b.add_two_step(j["plan"], j["d_in"].data_ptr(), j["mid"].data_ptr(), j["out"].data_ptr())
        b.add_single_step(plan1, d_in1.data_ptr(), out1.data_ptr())
        for replay in range(2):
            b.run(use_graph=True, stream=s)
            torch.cuda.synchronize()
            for k, j in enumerate(jobs):
                z = j["plan"].sizes
                raw = j["mid"].cpu().numpy()
                if compact:
                    raw = raw.view(np.uint16)[: raw.size].astype(np.float32)
                compare_rows(raw.reshape(j["want_mid"].shape), j["want_mid"], z.nr_samples_subbanding,
                             f"replay {replay} job {k} step one")
                compare_rows(j["out"].cpu().numpy().reshape(j["want"].shape), j["want"],
                             j["plan"].obs.nr_samples_per_batch, f"replay {replay} job {k} step two")
                j["out"].fill_(-1.0)
            compare_rows(out1.cpu().numpy().reshape(want1.shape), want1, plan1.sizes.nr_samples, "single-step job")
            out1.fill_(-1.0)
            torch.cuda.synchronize()


def test_packed_input_rejects_other_widths(torch_cuda):
    w = single_workload()
    with pytest.raises(DDError, match="input_bits"):
        Plan(w.obs, 128, input_bits=3)
