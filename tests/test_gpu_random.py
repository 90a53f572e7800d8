"""Seeded random shapes on the device against the oracle, bit for bit: the
single step over 8-bit and packed rows, the two-step path and batches of
uneven two-step jobs, each over shapes drawn from a splitmix64 stream
(channels, samples, DMs, first DMs and DM steps, band, zapped channels,
beams and their mapping, padding) and a configuration drawn from the valid
ones.  The fixed-shape tests sweep configurations; these sweep shapes."""
import numpy as np
import pytest

import oracle
from dedispersion_amd import DD_SINGLE_STEP, DD_STEP_ONE, DD_STEP_TWO, Batch, Conf, Plan, workloads
from dedispersion_amd.workloads import SplitMix64
from tests.conftest import host_threads
from tests.gpu_helpers import compare_rows, single_step, stream_of, zapped_flags
from tests.test_gpu_batch import Jobs
from tests.test_gpu_packed import packed_input
from tests.test_gpu_parity import STEP_ONE_CONFS, STEP_TWO_CONFS, SWEEP_CONFS, two_step_case

pytestmark = pytest.mark.gpu

PADDINGS = (0, 32, 128)


def uniform(rng: SplitMix64, lo: float, hi: float) -> float:
    return lo + (hi - lo) * rng.below(1 << 20) / float(1 << 20)


def first_dm(rng: SplitMix64) -> float:
    """0 for one shape in four, else a first DM in [0, 60): a nonzero first
    DM is where a fused multiply-add in the shift rule would show."""
    return 0.0 if rng.below(4) == 0 else uniform(rng, 0.0, 60.0)


def zapped_of(rng: SplitMix64, channels: int):
    n = rng.below(channels // 5 + 1)
    return tuple(sorted(rng.sample(channels, n))) if n else ()


def single_shape(seed: int):
    rng = SplitMix64(9000 + seed)
    channels = 2 + rng.below(399)
    beams = 1 + rng.below(2)
    return dict(channels=channels, samples=1 + rng.below(3000), dms=1 + rng.below(80),
                step=uniform(rng, 0.0, 2.0), min_freq=uniform(rng, 1100.0, 1500.0),
                bandwidth=uniform(rng, 0.2, 1.5), beams=beams, sbeams=beams + rng.below(2),
                zapped=zapped_of(rng, channels), seed=seed, first=first_dm(rng)), \
        PADDINGS[rng.below(3)], rng.below(len(SWEEP_CONFS) + 1)


def two_step_shape(seed: int):
    rng = SplitMix64(7000 + seed)
    subbands = (1, 2, 4, 8, 16)[rng.below(5)]
    channels = subbands * (1 + rng.below(40))
    beams = 1 + rng.below(2)
    return dict(channels=channels, samples=1 + rng.below(2000), dms=1 + rng.below(40), step=uniform(rng, 0.0, 1.5),
                subbands=subbands, coarse=1 + rng.below(12), coarse_step=uniform(rng, 0.0, 8.0),
                min_freq=uniform(rng, 1100.0, 1500.0), bandwidth=uniform(rng, 0.2, 1.5), beams=beams,
                sbeams=beams + rng.below(2), zapped=zapped_of(rng, channels), seed=seed, first=first_dm(rng),
                coarse_first=first_dm(rng)), \
        PADDINGS[rng.below(3)], rng.below(len(STEP_ONE_CONFS) + 1), rng.below(len(STEP_TWO_CONFS) + 1)


def pick(confs, i):
    """Index len(confs) stands for the plan default."""
    return None if i == len(confs) else confs[i]


@pytest.mark.parametrize("seed", range(24))
def test_single_step_random_shape(torch_cuda, seed):
    shape, padding, ci = single_shape(seed)
    w = workloads.tiny(**shape)
    conf = pick(SWEEP_CONFS, ci)
    plan, inp, out, orc, flags = single_step(torch_cuda, w, padding=padding, conf=conf)
    want = orc.dedispersion(inp, flags)
    got = out.reshape(want.shape)
    compare_rows(got, want, plan.sizes.nr_samples, f"seed {seed} {shape} padding {padding} conf {conf}")
    assert np.all(got[..., w.obs.nr_samples_per_batch:] == -1.0)
    plan.close()


@pytest.mark.parametrize("seed", range(12))
def test_single_step_packed_random_shape(torch_cuda, seed):
    """Packed rows (1, 2 or 4 bits per sample, random over the whole field)
    over random shapes whose shift spreads fit the pipelined kernel's row
    window, which every packed launch uses."""
    rng = SplitMix64(3000 + seed)
    channels = 2 + rng.below(399)
    beams = 1 + rng.below(2)
    w = workloads.tiny(channels=channels, samples=1 + rng.below(3000), dms=1 + rng.below(40),
                       step=uniform(rng, 0.0, 0.5), min_freq=uniform(rng, 1300.0, 1500.0),
                       bandwidth=uniform(rng, 0.2, 1.0), beams=beams, sbeams=beams + rng.below(2),
                       zapped=zapped_of(rng, channels), seed=seed, first=first_dm(rng) / 4)
    bits = (1, 2, 4)[rng.below(3)]
    padding = PADDINGS[rng.below(3)]
    conf = pick(SWEEP_CONFS, rng.below(len(SWEEP_CONFS) + 1))
    plan = Plan(w.obs, padding, input_bits=bits)
    orc = oracle.Oracle(w.obs, padding, threads=host_threads(), input_bits=bits)
    z = plan.sizes
    flags = zapped_flags(plan, w.zapped)
    plan.set_zapped_channels(flags)
    if conf is not None:
        plan.set_conf(DD_SINGLE_STEP, conf)
    host, d_in = packed_input(torch_cuda, orc, bits, z.nr_dispersed_samples, z.in_pitch, seed=seed)
    d_out = torch_cuda.full((z.output_elements,), -1.0, dtype=torch_cuda.float32, device="cuda")
    plan.dedisperse(d_in.data_ptr(), d_out.data_ptr(), stream=stream_of(torch_cuda))
    torch_cuda.cuda.synchronize()
    want = orc.dedispersion(host, flags)
    compare_rows(d_out.cpu().numpy().reshape(want.shape), want, z.nr_samples,
                 f"seed {seed} {bits}-bit {w.name} padding {padding} conf {conf}")
    plan.close()


@pytest.mark.parametrize("seed", range(24))
def test_two_step_random_shape(torch_cuda, seed):
    shape, padding, c1, c2 = two_step_shape(seed)
    w = workloads.tiny(**shape)
    conf1, conf2 = pick(STEP_ONE_CONFS, c1), pick(STEP_TWO_CONFS, c2)
    plan, flags, inp, mid, out, orc = two_step_case(torch_cuda, w, padding, conf1, conf2)
    z = plan.sizes
    what = f"seed {seed} {shape} padding {padding} confs {conf1} / {conf2}"
    want_mid = orc.step_one(inp, flags)
    compare_rows(mid.reshape(want_mid.shape), want_mid, z.nr_samples_subbanding, "step one, " + what)
    want = orc.step_two(mid)
    compare_rows(out.reshape(want.shape), want, w.obs.nr_samples_per_batch, "step two, " + what)
    plan.close()


@pytest.mark.parametrize("seed", range(8))
def test_batch_random_jobs(torch_cuda, seed):
    """3 to 6 random two-step jobs in one batch, uint16 intermediate, a random
    pipeline depth and staged configurations, against the oracle and against
    per-job launches, persistent or pipelined, with and without a graph.
    Jobs too wide for the uint16 step-two window run beside the batch
    kernels."""
    rng = SplitMix64(5000 + seed)
    ws = [workloads.tiny(**two_step_shape(100 * seed + j)[0]) for j in range(3 + rng.below(4))]
    chunks = (0, 2, 3)[rng.below(3)]

    def staged_conf():
        t1 = (1, 2, 4)[rng.below(3)]
        # a batch stages at most two rows per wavefront per group: unroll <= 2 threads_d1
        return Conf(64, t1, (8, 16)[rng.below(2)], (2, 4, 8)[rng.below(3)], 1 + rng.below(min(4, 2 * t1)), 1)

    one, two = staged_conf(), staged_conf()
    jobs = Jobs(torch_cuda, ws)
    what = f"seed {seed} chunks {chunks} confs {one} / {two}"
    with Batch() as b:
        b.set_conf(DD_STEP_ONE, one)
        b.set_conf(DD_STEP_TWO, two)
        b.set_pipeline(chunks)
        for plan, d_in, d_mid, d_out in zip(jobs.plans, jobs.inputs, jobs.mids, jobs.outs):
            b.add_two_step(plan, d_in.data_ptr(), d_mid.data_ptr(), d_out.data_ptr())
        info = b.info()
        assert info["compact_intermediate"] == 1
        # a job too wide for the batch kernels' windows runs beside them with
        # its plan's own configuration and the float intermediate
        place = [b.job_placement(j) for j in range(len(ws))]
        assert info["solo_jobs"] == sum(p["beside"] for p in place), what
        assert all(p["compact"] != p["beside"] for p in place), what
        compact = [p["compact"] for p in place]
        for use_graph in (False, True):
            b.run(use_graph=use_graph, stream=stream_of(torch_cuda))
            jobs.check(f"batch (graph {use_graph}), " + what, compact=compact)
            jobs.reset()
        b.run_per_job(stream=stream_of(torch_cuda))
        jobs.check("per job, " + what)
