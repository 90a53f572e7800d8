"""The streaming engine (dd_stream_*): device history, admission under an
HBM budget, overlap on several HIP streams.  Every request's output must be
bit-identical whether it runs alone, in a mix of concurrent requests, or as
a plain two-step launch over the contiguous input it stands for."""
import numpy as np
import pytest

import oracle
from dedispersion_amd import Plan
from dedispersion_amd.streaming import ChunkSource, StreamEngine, chunk_observation, poisson_arrivals, run_loop
from tests.conftest import host_threads
from tests.gpu_helpers import compare_rows, device_to_numpy

pytestmark = pytest.mark.gpu

SMALL = dict(samples=1500, coarse=8, fine=8, channels=256, subbands=16, bandwidth=1.0, min_freq=1300.0,
             coarse_step=3.0, fine_step=0.5)
BEAMS, CHUNKS = 3, 4


@pytest.fixture(scope="module")
def setup(torch_cuda):
    obs = chunk_observation(**SMALL)
    plan = Plan(obs, 128)
    source = ChunkSource(torch_cuda, obs.nr_channels, obs.nr_samples_per_batch, seed=5)
    return obs, plan, source


def collect(engine, arrivals, source, plan, realtime=False):
    z = plan.sizes
    got = {}

    def consume(ev):
        got[(ev["beam"], ev["chunk"])] = (ev["first_sample"],
                                          device_to_numpy(ev["output"], z.sub_output_elements, np.float32))

    stats = run_loop(engine, arrivals, source, consume, realtime=realtime)
    return got, stats


def stream_input(source, plan, beam, chunk):
    """The dispersed batch of (beam, chunk) cut from the beam's contiguous
    sample stream: [chunk S - H, chunk S + S), zeros before sample 0."""
    o, z = plan.obs, plan.sizes
    S = o.nr_samples_per_batch
    H = z.nr_dispersed_samples_subbanding - S
    full = np.concatenate([source.get(beam, k).numpy().reshape(o.nr_channels, S) for k in range(chunk + 1)], axis=1)
    full = np.concatenate([np.zeros((o.nr_channels, H), np.uint8), full], axis=1)
    batch = full[:, chunk * S: chunk * S + H + S]
    inp = np.zeros((o.nr_channels, z.sub_in_pitch), np.uint8)
    inp[:, : H + S] = batch
    return inp.ravel()


def test_mix_alone_and_direct_are_identical(torch_cuda, setup):
    torch = torch_cuda
    obs, plan, source = setup
    z = plan.sizes
    arrivals = poisson_arrivals(BEAMS, CHUNKS, 0.1, 1.5, seed=3)
    # mix: every chunk submitted at once, 2 slots, 3 streams
    budget = 2 * (z.sub_input_elements + 4 * (z.intermediate_elements + z.sub_output_elements)) + (64 << 20)
    mix_engine = StreamEngine(plan, BEAMS, budget, max_slots=2, hip_streams=3)
    assert mix_engine.info()["slots"] == 2
    mix, _ = collect(mix_engine, arrivals, source, plan)
    mix_engine.close()
    assert len(mix) == BEAMS * CHUNKS
    # alone: one request in flight at a time
    alone = {}
    solo = StreamEngine(plan, BEAMS, 1 << 32, max_slots=1, hip_streams=1)
    for a in arrivals:
        got, _ = collect(solo, [a], source, plan)
        alone.update(got)
    solo.close()
    H = z.nr_dispersed_samples_subbanding - obs.nr_samples_per_batch
    rows = obs.nr_dms_subbanding * obs.nr_dms
    for key in sorted(mix):
        first, out = mix[key]
        assert first == key[1] * obs.nr_samples_per_batch - H
        a_first, a_out = alone[key]
        assert a_first == first
        compare_rows(out.reshape(rows, -1), a_out.reshape(rows, -1), obs.nr_samples_per_batch, f"mix vs alone {key}")
    # direct: a plain two-step launch over the contiguous input
    d_mid = torch.empty(z.intermediate_elements, dtype=torch.float32, device="cuda")
    d_out = torch.empty(z.sub_output_elements, dtype=torch.float32, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    for key in [(0, 0), (0, 3), (1, 2), (2, 1)]:
        d_in = torch.from_numpy(stream_input(source, plan, *key)).cuda()
        plan.step_one(d_in.data_ptr(), d_mid.data_ptr(), stream=s)
        plan.step_two(d_mid.data_ptr(), d_out.data_ptr(), stream=s)
        want = d_out.cpu().numpy().reshape(rows, -1)
        compare_rows(mix[key][1].reshape(rows, -1), want, obs.nr_samples_per_batch, f"mix vs direct {key}")


def test_streamed_chunk_matches_oracle(torch_cuda, setup):
    obs, plan, source = setup
    z = plan.sizes
    engine = StreamEngine(plan, BEAMS, 1 << 31, max_slots=4, hip_streams=2)
    arrivals = [a for a in poisson_arrivals(BEAMS, 2, 0.1, 1.5, seed=4) if a.beam == 1]
    got, _ = collect(engine, arrivals, source, plan)
    engine.close()
    orc = oracle.Oracle(obs, 128, threads=host_threads())
    flags = np.zeros(oracle.pad(obs.nr_channels, 32), dtype=np.uint32)
    inp = stream_input(source, plan, 1, 1)
    mid = orc.step_one(inp, flags)
    want = orc.step_two(mid.ravel())
    rows = obs.nr_dms_subbanding * obs.nr_dms
    compare_rows(got[(1, 1)][1].reshape(rows, -1), want.reshape(rows, -1), obs.nr_samples_per_batch, "oracle")


def test_admission_respects_the_budget(torch_cuda, setup):
    obs, plan, source = setup
    z = plan.sizes
    slot = z.sub_input_elements + 4 * (z.intermediate_elements + z.sub_output_elements)
    probe = StreamEngine(plan, BEAMS, 1 << 31, max_slots=1, hip_streams=1)
    history = probe.info()["history_bytes"]
    probe.close()
    budget = history + 3 * slot + slot // 2
    engine = StreamEngine(plan, BEAMS, budget, max_slots=64, hip_streams=2)
    info = engine.info()
    assert info["history_bytes"] == history and info["slots"] == 3
    for k, a in enumerate(poisson_arrivals(BEAMS, 2, 0.1, 1.5)):
        engine.submit(a.beam, source.get(a.beam, a.chunk).data_ptr(), obs.nr_samples_per_batch, k)
    info = engine.info()
    assert info["in_flight"] == 3 and info["queued"] == 3
    done = 0
    while done < 6:
        for ev in engine.progress():
            engine.release(ev["slot"])
            done += 1
    assert engine.info()["completed"] == 6
    engine.close()


def test_realtime_loop_reports_latency(torch_cuda, setup):
    obs, plan, source = setup
    engine = StreamEngine(plan, BEAMS, 1 << 31, max_slots=8, hip_streams=4)
    arrivals = poisson_arrivals(BEAMS, 6, 0.05, 1.5, seed=6)
    stats = run_loop(engine, arrivals, source)
    engine.close()
    assert stats["requests"] == BEAMS * 6
    assert 0 < stats["p50_ms"] <= stats["p99_ms"] <= stats["max_ms"] < 2000
