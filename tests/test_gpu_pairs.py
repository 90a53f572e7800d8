"""The paired step two of batches (kernels_u16_pairs.hip) against the oracle
and against the unpaired uint16 step two.

With the compact intermediate, step two first adds the uint16 sums of
subbands 2k and 2k + 1 at every shift difference the pair takes over a DM
tile, then adds one row per pair for each DM.  A batch asked for it
(dd_batch_set_paired_step_two) takes it when every job allows it (a pair sum
below 2^16, at most four shift differences per pair and DM tile, 8-element
aligned rows) and otherwise adds every subband on its own.  Both must equal the oracle's step two (the reference's sequential float
sum, include/Dedispersion.hpp:155-176) bit for bit.
"""
import numpy as np
import pytest

import oracle
from dedispersion_amd import DD_STEP_ONE, DD_STEP_TWO, Batch, Conf, Plan, workloads
from tests.conftest import host_threads
from tests.gpu_helpers import compare_rows, make_input, stream_of, zapped_flags

pytestmark = pytest.mark.gpu

PADDING = 128
STEP_ONE = Conf(64, 4, 4, 2, 4, 1)
PAIR_SLOTS = 4


def pair_range(plan, tile):
    """Most values a pair's shift difference takes over a DM tile (the
    width of its range), from the device's integer step-two table."""
    t = plan.get_integer_shifts(DD_STEP_TWO).astype(np.int64)  # [subbands][dm pitch]
    D = plan.obs.nr_dms
    worst = 1
    for r in range(0, t.shape[0] - 1, 2):
        delta = t[r + 1, :D] - t[r, :D]
        for d0 in range(0, D, tile):
            seg = delta[d0:d0 + tile]
            worst = max(worst, int(seg.max() - seg.min() + 1))
    return worst


def pairs_allowed(plan, conf, zapped):
    o = plan.obs
    per = o.nr_channels // o.nr_subbands
    live = [per - sum(1 for c in zapped if s * per <= c < (s + 1) * per) for s in range(o.nr_subbands)]
    pitch = plan.sizes.sub_out_pitch
    return (2 * 255 * max(max(live), 1) <= 65535 and pair_range(plan, conf.threads_d1 * conf.items_d1) <= PAIR_SLOTS
            and pitch % 8 == 0)


class PairJobs:
    """Two-step jobs with their device buffers and oracle results."""

    def __init__(self, torch, ws, mappings=None, padding=PADDING, inputs=None):
        self.torch, self.ws = torch, ws
        self.plans, self.inputs, self.mids, self.outs, self.want = [], [], [], [], []
        for j, w in enumerate(ws):
            plan = Plan(w.obs, padding)
            flags = zapped_flags(plan, w.zapped)
            plan.set_zapped_channels(flags)
            mapping = None if mappings is None else mappings[j]
            if mapping is not None:
                plan.set_beam_mapping(step_two=mapping)
            z = plan.sizes
            if inputs is None:
                d_in = make_input(torch, plan, w, z.sub_in_pitch, z.nr_dispersed_samples_subbanding, w.seed)
            else:
                d_in = inputs(plan, w)
            orc = oracle.Oracle(w.obs, padding, threads=host_threads())
            mid = orc.step_one(d_in.cpu().numpy(), flags)
            self.want.append(orc.step_two(mid.ravel(), beam_mapping=mapping))
            self.plans.append(plan)
            self.inputs.append(d_in)
            self.mids.append(torch.full((z.intermediate_elements,), -1.0, dtype=torch.float32, device="cuda"))
            self.outs.append(torch.full((z.sub_output_elements,), -1.0, dtype=torch.float32, device="cuda"))

    def batch(self, conf_two, paired=True, chunks=0):
        b = Batch()
        b.set_conf(DD_STEP_ONE, STEP_ONE)
        b.set_conf(DD_STEP_TWO, conf_two)
        b.set_paired_step_two(paired)
        b.set_pipeline(chunks)
        for plan, d_in, d_mid, d_out in zip(self.plans, self.inputs, self.mids, self.outs):
            b.add_two_step(plan, d_in.data_ptr(), d_mid.data_ptr(), d_out.data_ptr())
        return b

    def reset(self):
        for t in self.mids + self.outs:
            t.fill_(-1.0)
        self.torch.cuda.synchronize()

    def check(self, what):
        self.torch.cuda.synchronize()
        for j, (w, want) in enumerate(zip(self.ws, self.want)):
            got = self.outs[j].cpu().numpy().reshape(want.shape)
            compare_rows(got, want, w.obs.nr_samples_per_batch, f"{what} job {j}")


def shapes():
    """Uneven jobs whose fine DM ranges per tile are narrow, as in a real
    subband search (a tile's fine DMs span one coarse DM step), so that
    every pair takes at most four shift differences per DM tile."""
    return [
        dict(channels=96, samples=700, dms=12, step=0.2, subbands=8, coarse=9, coarse_step=2.4, beams=2, sbeams=3,
             zapped=(0, 11, 12, 50, 95), seed=41),
        dict(channels=96, samples=333, dms=5, step=0.3, subbands=8, coarse=3, coarse_step=1.5, seed=42),
        dict(channels=64, samples=1100, dms=20, step=0.1, subbands=4, coarse=6, coarse_step=2.0, sbeams=2,
             zapped=tuple(range(16, 32)), seed=43),
        dict(channels=192, samples=900, dms=17, step=0.05, subbands=4, coarse=5, coarse_step=0.85, zapped=(5, 77),
             seed=44),
        dict(channels=96, samples=256, dms=8, step=0.25, subbands=8, coarse=1, coarse_step=0.0, zapped=(47,), seed=45),
        dict(channels=256, samples=1500, dms=64, step=0.05, subbands=16, coarse=4, coarse_step=3.2, seed=46),
    ]


CONFS = [Conf(64, t1, i0, i1, u, 1) for (t1, i0, i1, u) in
         [(4, 16, 8, 4), (1, 4, 1, 1), (1, 16, 8, 2), (2, 8, 2, 3), (4, 4, 8, 8), (8, 16, 1, 16), (8, 8, 4, 5),
          (4, 16, 2, 7), (2, 16, 4, 1)]]


@pytest.mark.parametrize("conf", CONFS, ids=lambda c: f"t{c.threads_d1}i{c.items_d0}x{c.items_d1}u{c.unroll}")
def test_paired_step_two_configurations(torch_cuda, conf):
    """Every register shape, waves per workgroup and pairs per barrier (an
    unroll of one or two rows stages one pair), paired against the oracle,
    then unpaired."""
    ws = [workloads.tiny(min_freq=1300.0, bandwidth=1.0, **s) for s in shapes()]
    jobs = PairJobs(torch_cuda, ws)
    expect = all(pairs_allowed(p, conf, w.zapped) for p, w in zip(jobs.plans, ws))
    assert expect, "these shapes are meant to take the paired kernel"
    s = stream_of(torch_cuda)
    with jobs.batch(conf) as b:
        info = b.info()
        assert info["compact_intermediate"] == 1 and info["paired_step_two"] == 1 and info["solo_jobs"] == 0
        b.run(stream=s)
        jobs.check(f"paired {conf}")
    jobs.reset()
    with jobs.batch(conf, paired=False) as b:
        assert b.info()["paired_step_two"] == 0
        b.run(stream=s)
        jobs.check(f"unpaired {conf}")


@pytest.mark.parametrize("subbands", [1, 2, 3, 5, 7])
def test_paired_odd_subband_counts(torch_cuda, subbands):
    """An odd number of subbands leaves the last one without a partner (its
    window is the row itself); one subband is a single lone row."""
    channels = 16 * subbands
    ws = [workloads.tiny(channels=channels, samples=600, dms=11, step=0.2, subbands=subbands, coarse=4,
                         coarse_step=2.2, min_freq=1300.0, bandwidth=1.0, zapped=(3,), seed=50 + subbands),
          workloads.tiny(channels=channels, samples=1500, dms=24, step=0.1, subbands=subbands, coarse=2,
                         coarse_step=2.4, min_freq=1250.0, bandwidth=1.5, seed=60 + subbands)]
    jobs = PairJobs(torch_cuda, ws)
    conf = Conf(64, 4, 8, 4, 4, 1)
    with jobs.batch(conf) as b:
        assert b.info()["paired_step_two"] == 1
        b.run(stream=stream_of(torch_cuda))
        jobs.check(f"{subbands} subbands")


@pytest.mark.parametrize("chunks", [0, 2, 3])
def test_paired_first_dm_mappings_and_graph_replays(torch_cuda, chunks):
    """Fine and coarse first DMs other than zero, synthesized beams whose
    adjacent subbands come from different input beams, persistent and
    pipelined launches, graph replays."""
    ws = [workloads.tiny(channels=128, samples=900, dms=20, step=0.15, subbands=8, coarse=6, coarse_step=3.0,
                         beams=3, sbeams=4, first=3.3, coarse_first=12.7, min_freq=1300.0, bandwidth=1.0,
                         zapped=(2, 64), seed=71),
          workloads.tiny(channels=64, samples=1300, dms=16, step=0.2, subbands=8, coarse=5, coarse_step=3.2,
                         beams=2, sbeams=2, first=1.1, min_freq=1350.0, bandwidth=1.0, seed=72)]
    rng = np.random.default_rng(73)
    mappings = []
    for w in ws:
        o = w.obs
        m = np.zeros((o.nr_synthesized_beams, oracle.pad(o.nr_subbands, PADDING // 4)), dtype=np.uint32)
        m[:, : o.nr_subbands] = rng.integers(0, o.nr_beams, size=(o.nr_synthesized_beams, o.nr_subbands))
        m[0, : o.nr_subbands] = np.arange(o.nr_subbands) % o.nr_beams  # neighbours always differ
        mappings.append(m)
    jobs = PairJobs(torch_cuda, ws, mappings)
    s = stream_of(torch_cuda)
    with jobs.batch(Conf(64, 4, 16, 8, 4, 1), chunks=chunks) as b:
        assert b.info()["paired_step_two"] == 1
        b.run(use_graph=False, stream=s)
        jobs.check("eager")
        for replay in range(3):
            jobs.reset()
            b.run(use_graph=True, stream=s)
            jobs.check(f"graph replay {replay}")


def test_paired_sums_at_the_pair_bound(torch_cuda):
    """All-255 input and 128 live channels per subband: every pair sum is
    2 x 255 x 128 = 65280, the largest the paired kernel adds in 16 bits
    (one pair per 16-bit add).  With 129 live channels a pair could pass
    2^16, so the batch adds subbands one by one."""
    torch = torch_cuda

    def full(plan, w):
        return torch.full((plan.sizes.sub_input_elements,), 255, dtype=torch.uint8, device="cuda")

    s = stream_of(torch)
    for per, paired in ((128, 1), (129, 0)):
        w = workloads.tiny(channels=4 * per, samples=700, dms=9, step=0.3, subbands=4, coarse=3, coarse_step=1.0,
                           min_freq=1300.0, bandwidth=0.25, seed=80 + per)
        jobs = PairJobs(torch, [w], inputs=full)
        with jobs.batch(Conf(64, 4, 8, 4, 4, 1)) as b:
            info = b.info()
            assert info["compact_intermediate"] == 1 and info["paired_step_two"] == paired, (per, info)
            b.run(stream=s)
            jobs.check(f"{per} live channels per subband")
        out = jobs.outs[0].cpu().numpy().reshape(-1, jobs.plans[0].sizes.out_pitch)[:, : w.obs.nr_samples_per_batch]
        assert np.all(out == 255.0 * 4 * per)


def test_unpaired_when_a_job_does_not_allow_pairs(torch_cuda):
    """A pair whose shift difference takes more than four values over a DM
    tile (wide fine DM steps), or rows that are not 8-element aligned (no
    padding): the whole batch adds subbands one by one, with the same
    results."""
    conf = Conf(64, 4, 8, 8, 4, 1)
    wide = workloads.tiny(channels=64, samples=1200, dms=32, step=0.5, subbands=8, coarse=2, coarse_step=16.0,
                          min_freq=1000.0, bandwidth=1.0, seed=91)
    plain = workloads.tiny(channels=96, samples=800, dms=12, step=0.2, subbands=8, coarse=3, coarse_step=2.4,
                           min_freq=1300.0, bandwidth=1.0, seed=92)
    s = stream_of(torch_cuda)
    jobs = PairJobs(torch_cuda, [plain, wide])
    assert pairs_allowed(jobs.plans[0], conf, ())
    assert not pairs_allowed(jobs.plans[1], conf, ()), pair_range(jobs.plans[1], 32)
    with jobs.batch(conf) as b:
        info = b.info()
        assert info["solo_jobs"] == 0 and info["compact_intermediate"] == 1 and info["paired_step_two"] == 0
        b.run(stream=s)
        jobs.check("wide shift differences")
    odd = workloads.tiny(channels=96, samples=803, dms=12, step=0.2, subbands=8, coarse=3, coarse_step=2.4,
                         min_freq=1300.0, bandwidth=1.0, seed=93)
    unaligned = PairJobs(torch_cuda, [plain, odd], padding=0)
    assert unaligned.plans[1].sizes.sub_out_pitch % 8 != 0
    with unaligned.batch(conf) as b:
        assert b.info()["paired_step_two"] == 0
        b.run(stream=s)
        unaligned.check("unaligned rows")
