"""Batched persistent launches (dd_batch_*) against the oracle and against
one launch per job.

A batch runs every tile of every job with the device function and the
parameters of that job's own launch, so its results must be bit-identical to
the per-job launches, and, through them, to the oracle.
"""
import numpy as np
import pytest

import oracle
from dedispersion_amd import DD_SINGLE_STEP, DD_STEP_ONE, DD_STEP_TWO, Batch, Conf, DDError, Plan, workloads
from dedispersion_amd.beams import BeamSet
from tests.conftest import host_threads
from tests.gpu_helpers import compare_rows, make_input, stream_of, zapped_flags

pytestmark = pytest.mark.gpu

STAGED_SMALL = Conf(64, 4, 4, 2, 4, 1)  # 4 waves, 256 samples x 8 DMs per tile, staged
STAGED_STEP_TWO = Conf(64, 2, 8, 4, 3, 1)


def tiny_two_step_workloads():
    """Uneven jobs: DMs, samples, coarse DMs, beams and zapped channels differ."""
    shapes = [
        dict(channels=96, samples=700, dms=12, step=0.9, subbands=8, coarse=9, coarse_step=4.5, beams=2, sbeams=3,
             zapped=(0, 11, 12, 50, 95), seed=31),
        dict(channels=96, samples=333, dms=5, step=1.3, subbands=8, coarse=3, coarse_step=6.0, beams=1, sbeams=1,
             zapped=(), seed=32),
        dict(channels=64, samples=1100, dms=20, step=0.5, subbands=4, coarse=6, coarse_step=9.0, beams=1, sbeams=2,
             zapped=tuple(range(16, 32)), seed=33),
        dict(channels=96, samples=256, dms=8, step=0.7, subbands=8, coarse=1, coarse_step=0.0, beams=1, sbeams=1,
             zapped=(47,), seed=34),
    ]
    return [workloads.tiny(min_freq=1300.0, bandwidth=1.0, **s) for s in shapes]


class Jobs:
    """Plans, device inputs and outputs, and oracle results of a job list."""

    def __init__(self, torch, ws, two_step=True, padding=0):
        self.torch = torch
        self.ws = ws
        self.plans, self.flags, self.inputs, self.mids, self.outs, self.want_mid, self.want_out = ([] for _ in range(7))
        for w in ws:
            plan = Plan(w.obs, padding)
            flags = zapped_flags(plan, w.zapped)
            plan.set_zapped_channels(flags)
            z = plan.sizes
            orc = oracle.Oracle(w.obs, padding, threads=host_threads())
            if two_step:
                d_in = make_input(torch, plan, w, z.sub_in_pitch, z.nr_dispersed_samples_subbanding, w.seed)
                mid = orc.step_one(d_in.cpu().numpy(), flags)
                self.want_mid.append(mid)
                self.want_out.append(orc.step_two(mid.ravel()))
                self.mids.append(torch.full((z.intermediate_elements,), -1.0, dtype=torch.float32, device="cuda"))
                self.outs.append(torch.full((z.sub_output_elements,), -1.0, dtype=torch.float32, device="cuda"))
            else:
                d_in = make_input(torch, plan, w, z.in_pitch, z.nr_dispersed_samples, w.seed)
                self.want_out.append(orc.dedispersion(d_in.cpu().numpy(), flags))
                self.mids.append(None)
                self.outs.append(torch.full((z.output_elements,), -1.0, dtype=torch.float32, device="cuda"))
            self.plans.append(plan)
            self.flags.append(flags)
            self.inputs.append(d_in)

    def reset(self):
        for m in self.mids:
            if m is not None:
                m.fill_(-1.0)
        for o in self.outs:
            o.fill_(-1.0)
        self.torch.cuda.synchronize()

    def check(self, what, compact=False):
        """compact: the intermediate buffers hold uint16 sums in their first
        half (a bool for every job, or one per job)."""
        self.torch.cuda.synchronize()
        for j, (w, plan) in enumerate(zip(self.ws, self.plans)):
            z = plan.sizes
            if self.mids[j] is not None:
                want = self.want_mid[j]
                raw = self.mids[j].cpu().numpy()
                if compact[j] if isinstance(compact, (list, tuple)) else compact:
                    raw = raw.view(np.uint16)[: raw.size].astype(np.float32)
                got = raw.reshape(want.shape)
                compare_rows(got, want, z.nr_samples_subbanding, f"{what} job {j} step one")
                want = self.want_out[j]
                got = self.outs[j].cpu().numpy().reshape(want.shape)
                compare_rows(got, want, w.obs.nr_samples_per_batch, f"{what} job {j} step two")
            else:
                want = self.want_out[j]
                got = self.outs[j].cpu().numpy().reshape(want.shape)
                compare_rows(got, want, z.nr_samples, f"{what} job {j} single step")


@pytest.mark.parametrize("chunks", [0, 2, 3, 9], ids=lambda c: f"pipeline{c}")
@pytest.mark.parametrize("compact", [True, False], ids=["uint16-intermediate", "float-intermediate"])
def test_batch_two_step_uneven_jobs(torch_cuda, compact, chunks):
    jobs = Jobs(torch_cuda, tiny_two_step_workloads())
    with Batch() as b:
        b.set_conf(DD_STEP_ONE, STAGED_SMALL)
        b.set_conf(DD_STEP_TWO, STAGED_STEP_TWO)
        b.set_compact_intermediate(compact)
        b.set_pipeline(chunks)
        for plan, d_in, d_mid, d_out in zip(jobs.plans, jobs.inputs, jobs.mids, jobs.outs):
            b.add_two_step(plan, d_in.data_ptr(), d_mid.data_ptr(), d_out.data_ptr())
        info = b.info()
        assert info["jobs"] == 4 and info["tiles"][0] == 0 and info["tiles"][1] > 0 and info["tiles"][2] > 0
        assert info["compact_intermediate"] == int(compact)
        assert (info["pipeline_chunks"] == 0) == (chunks < 2) and info["pipeline_chunks"] <= min(chunks, 4)
        want_units = sum(w.obs.nr_synthesized_beams * w.obs.nr_dms_subbanding * w.obs.nr_dms *
                         w.obs.nr_samples_per_batch for w in jobs.ws)
        assert info["units"] == want_units
        b.run(use_graph=False, stream=stream_of(torch_cuda))
        jobs.check("persistent", compact=compact)
        jobs.reset()
        b.run_per_job(stream=stream_of(torch_cuda))
        jobs.check("per job")


@pytest.mark.parametrize("conf", [Conf(64, t1, i0, i1, u, 1) for (t1, i0, i1, u) in
                                  [(1, 4, 1, 1), (1, 16, 8, 2), (2, 8, 2, 3), (4, 4, 8, 8), (8, 16, 1, 16),
                                   (8, 8, 4, 5), (4, 16, 2, 7)]],
                         ids=lambda c: f"t{c.threads_d1}i{c.items_d0}x{c.items_d1}u{c.unroll}")
def test_batch_uint16_step_two_configurations(torch_cuda, conf):
    """The uint16 step two over every register shape and unroll, including
    unrolls that do not divide the number of subbands."""
    ws = tiny_two_step_workloads() + [workloads.tiny(channels=192, samples=900, dms=17, step=0.6, subbands=4,
                                                      coarse=5, coarse_step=5.0, min_freq=1300.0, bandwidth=1.0,
                                                      seed=35, zapped=(5, 77))]
    jobs = Jobs(torch_cuda, ws)
    with Batch() as b:
        b.set_conf(DD_STEP_ONE, STAGED_SMALL)
        b.set_conf(DD_STEP_TWO, conf)
        for plan, d_in, d_mid, d_out in zip(jobs.plans, jobs.inputs, jobs.mids, jobs.outs):
            b.add_two_step(plan, d_in.data_ptr(), d_mid.data_ptr(), d_out.data_ptr())
        assert b.info()["compact_intermediate"] == 1
        b.run(stream=stream_of(torch_cuda))
        jobs.check(f"uint16 step two {conf}", compact=True)


def test_batch_falls_back_to_float_for_wide_subbands(torch_cuda):
    """300 channels per subband: 255 x 300 > 2^16, so the batch keeps the
    float intermediate (and its results)."""
    w = workloads.tiny(channels=600, samples=400, dms=6, step=0.4, subbands=2, coarse=3, coarse_step=2.0,
                       min_freq=1300.0, bandwidth=1.0, seed=36)
    jobs = Jobs(torch_cuda, [w])
    with Batch() as b:
        b.set_conf(DD_STEP_ONE, STAGED_SMALL)
        b.set_conf(DD_STEP_TWO, STAGED_STEP_TWO)
        b.add_two_step(jobs.plans[0], jobs.inputs[0].data_ptr(), jobs.mids[0].data_ptr(), jobs.outs[0].data_ptr())
        assert b.info()["compact_intermediate"] == 0
        b.run(stream=stream_of(torch_cuda))
        jobs.check("wide subbands")


@pytest.mark.parametrize("chunks", [0, 2], ids=["persistent", "pipelined"])
def test_batch_graph_replays(torch_cuda, chunks):
    """The captured graph (queue reset kernel + persistent launches, or the
    two-stream pipeline) gives the eager result on every replay, with other
    device work between replays."""
    torch = torch_cuda
    jobs = Jobs(torch, tiny_two_step_workloads()[:3])
    with Batch() as b:
        b.set_conf(DD_STEP_ONE, STAGED_SMALL)
        b.set_conf(DD_STEP_TWO, STAGED_STEP_TWO)
        b.set_pipeline(chunks)
        for plan, d_in, d_mid, d_out in zip(jobs.plans, jobs.inputs, jobs.mids, jobs.outs):
            b.add_two_step(plan, d_in.data_ptr(), d_mid.data_ptr(), d_out.data_ptr())
        scratch = torch.arange(1 << 20, device="cuda", dtype=torch.float32)
        for replay in range(4):
            jobs.reset()
            b.run(use_graph=True, stream=stream_of(torch))
            _ = (scratch * 3.0).sum().item()  # eager work and a synchronization between replays
            jobs.check(f"graph replay {replay}", compact=True)


def test_batch_single_step_jobs(torch_cuda):
    ws = [
        workloads.tiny(channels=300, samples=500, dms=9, step=1.1, zapped=(3, 299), seed=41, min_freq=1300.0,
                       bandwidth=1.0),
        workloads.tiny(channels=40, samples=1000, dms=30, step=0.3, beams=2, sbeams=3, seed=42, min_freq=1300.0,
                       bandwidth=1.0),
        workloads.tiny(channels=16, samples=64, dms=8, step=2.0, seed=43),
    ]
    jobs = Jobs(torch_cuda, ws, two_step=False)
    with Batch() as b:
        b.set_conf(DD_SINGLE_STEP, Conf(64, 2, 8, 2, 4, 1))
        for plan, d_in, d_out in zip(jobs.plans, jobs.inputs, jobs.outs):
            b.add_single_step(plan, d_in.data_ptr(), d_out.data_ptr())
        b.run(use_graph=True, stream=stream_of(torch_cuda))
        jobs.check("single-step batch")


def test_batch_mixed_single_and_two_step(torch_cuda):
    two = Jobs(torch_cuda, tiny_two_step_workloads()[1:3])
    one = Jobs(torch_cuda, [workloads.tiny(channels=48, samples=300, dms=10, step=0.8, seed=44, min_freq=1300.0,
                                           bandwidth=1.0)], two_step=False)
    with Batch() as b:
        b.set_conf(DD_SINGLE_STEP, STAGED_SMALL)
        b.set_conf(DD_STEP_ONE, STAGED_SMALL)
        b.set_conf(DD_STEP_TWO, STAGED_STEP_TWO)
        b.add_two_step(two.plans[0], two.inputs[0].data_ptr(), two.mids[0].data_ptr(), two.outs[0].data_ptr())
        b.add_single_step(one.plans[0], one.inputs[0].data_ptr(), one.outs[0].data_ptr())
        b.add_two_step(two.plans[1], two.inputs[1].data_ptr(), two.mids[1].data_ptr(), two.outs[1].data_ptr())
        b.run(stream=stream_of(torch_cuda))
        two.check("mixed batch", compact=True)
        one.check("mixed batch")


def test_batch_rejects_unstaged_configuration(torch_cuda):
    jobs = Jobs(torch_cuda, tiny_two_step_workloads()[:1])
    with Batch() as b:
        b.add_two_step(jobs.plans[0], jobs.inputs[0].data_ptr(), jobs.mids[0].data_ptr(), jobs.outs[0].data_ptr())
        b.set_conf(DD_STEP_ONE, Conf(64, 4, 4, 2, 4, 0))  # global-memory kernel: not batchable
        with pytest.raises(DDError, match="staged"):
            b.run(stream=stream_of(torch_cuda))
        b.set_conf(DD_STEP_ONE, Conf(64, 2, 4, 2, 8, 1))  # unroll > 2 * threads_d1
        with pytest.raises(DDError, match="staged"):
            b.run(stream=stream_of(torch_cuda))
        b.set_conf(DD_STEP_ONE, STAGED_SMALL)
        b.run(stream=stream_of(torch_cuda))
        jobs.check("after rejected configurations", compact=True)


@pytest.mark.parametrize("chunks", [0, 5], ids=["persistent", "pipelined"])
def test_batched_beams_40_match_per_beam_launches(torch_cuda, chunks):
    """The 40-beam uneven workload: the graph-replayed batch (persistent, or
    pipelined over two streams) equals one launch per beam and step, bit for
    bit, on every beam; one whole coarse-DM band of every beam (its step-one
    rows over all subbands and its step-two rows over all fine DMs) is
    checked against the oracle directly."""
    torch = torch_cuda
    ws = workloads.batched_beams()
    bs = BeamSet(torch, ws)
    bs.batch.set_pipeline(chunks)
    try:
        info = bs.info()
        assert info["jobs"] == 40
        bs.run_per_job()
        torch.cuda.synchronize()
        ref_mid = [m.clone() for m in bs.mids]
        ref_out = [o.clone() for o in bs.outs]
        for m in bs.mids:
            m.fill_(-7.0)
        for o in bs.outs:
            o.fill_(-7.0)
        bs.run(use_graph=True)
        bs.run(use_graph=True)
        torch.cuda.synchronize()
        for j, plan in enumerate(bs.plans):
            z = plan.sizes
            o = plan.obs
            got = bs.outs[j].view(o.nr_dms_subbanding * o.nr_dms, z.out_pitch)[:, : z.nr_samples]
            want = ref_out[j].view(o.nr_dms_subbanding * o.nr_dms, z.out_pitch)[:, : z.nr_samples]
            assert torch.equal(got, want), f"beam {j} ({ws[j].name}): batch differs from per-beam launches"
            # the batch keeps its intermediate as uint16 sums in the buffer's first half
            n = bs.mids[j].numel()
            gm = bs.mids[j].view(torch.int16)[:n].view(-1, z.sub_out_pitch)[:, : z.nr_samples_subbanding]
            gm = gm.to(torch.int32) & 0xFFFF
            wm = ref_mid[j].view(-1, z.sub_out_pitch)[:, : z.nr_samples_subbanding]
            assert torch.equal(gm.to(torch.float32), wm), f"beam {j}: intermediate differs"
        del ref_mid, ref_out
        # one full coarse-DM band of every beam against the oracle: step one
        # (all subbands) and step two (all fine DMs) of coarse DM c
        for j, plan in enumerate(bs.plans):
            w = ws[j]
            o, z = w.obs, plan.sizes
            c = (7 * j + 3) % o.nr_dms_subbanding
            orc = oracle.Oracle(o, bs.padding, threads=host_threads())
            flags = zapped_flags(plan, w.zapped)
            inp = bs.inputs[j].cpu().numpy()
            n = bs.mids[j].numel()
            band = bs.mids[j].view(torch.int16)[:n].view(o.nr_dms_subbanding, o.nr_subbands, z.sub_out_pitch)[c]
            band = (band.to(torch.int32) & 0xFFFF).to(torch.float32).cpu().numpy()
            want = orc.step_one(inp, flags, rows=(c, c + 1))
            compare_rows(band[None, None], want[:, c:c + 1], z.nr_samples_subbanding, f"beam {j} step one, coarse DM {c}")
            mid = np.zeros(n, dtype=np.float32)  # only band c is read by the rows below
            mid.reshape(o.nr_dms_subbanding, o.nr_subbands, z.sub_out_pitch)[c] = band
            r0, r1 = c * o.nr_dms, (c + 1) * o.nr_dms
            want = orc.step_two(mid, rows=(r0, r1))
            got = bs.outs[j].view(o.nr_dms_subbanding * o.nr_dms, z.out_pitch)[r0:r1].cpu().numpy()
            compare_rows(got[None], want[:, r0:r1], o.nr_samples_per_batch, f"beam {j} step two, coarse DM {c}")
    finally:
        bs.close()
        torch.cuda.empty_cache()


def test_batch_uint16_sums_at_the_16_bit_bound(torch_cuda):
    """All-255 input: every step-one sum is 255 x the subband's live channels
    (12240 for 48 live channels), the largest uint16 sums the batch writes,
    and step two adds 8 of them per output sample beyond 2^16; the output
    must be 255 x all live channels."""
    torch = torch_cuda
    w = workloads.tiny(channels=384, samples=777, dms=9, step=0.7, subbands=8, coarse=4, coarse_step=3.0,
                       min_freq=1300.0, bandwidth=1.0, zapped=(50, 51, 200), seed=37)
    plan = Plan(w.obs, 128)
    plan.set_zapped_channels(zapped_flags(plan, w.zapped))
    z = plan.sizes
    d_in = torch.full((z.sub_input_elements,), 255, dtype=torch.uint8, device="cuda")
    d_mid = torch.full((z.intermediate_elements,), -1.0, dtype=torch.float32, device="cuda")
    d_out = torch.full((z.sub_output_elements,), -1.0, dtype=torch.float32, device="cuda")
    live = np.full(8, 48)
    live[1] -= 2  # channels 50, 51
    live[4] -= 1  # channel 200
    for conf in (Conf(64, 4, 8, 4, 8, 1), Conf(64, 8, 16, 2, 16, 1), Conf(64, 1, 4, 1, 2, 1)):
        with Batch() as b:
            b.set_conf(DD_STEP_ONE, STAGED_SMALL)
            b.set_conf(DD_STEP_TWO, conf)
            b.add_two_step(plan, d_in.data_ptr(), d_mid.data_ptr(), d_out.data_ptr())
            assert b.info()["compact_intermediate"] == 1
            b.run(stream=stream_of(torch))
            torch.cuda.synchronize()
        raw = d_mid.cpu().numpy()
        mid = raw.view(np.uint16)[: raw.size].reshape(w.obs.nr_dms_subbanding, 8, z.sub_out_pitch)
        assert np.all(mid[:, :, : z.nr_samples_subbanding] == (255 * live)[None, :, None]), conf
        out = d_out.cpu().numpy().reshape(-1, z.out_pitch)[:, : w.obs.nr_samples_per_batch]
        assert np.all(out == 255.0 * live.sum()), conf


def test_batch_runs_wide_jobs_beside_the_kernels(torch_cuda):
    """A job whose row windows do not fit the staged kernels (300 MHz at the
    bottom of the band, wide DM steps) runs beside the batch kernels with its
    plan's own configuration and the float intermediate, in every run mode;
    the other jobs stay in the batch with the uint16 intermediate."""
    ws = tiny_two_step_workloads()[:2] + [
        workloads.tiny(channels=64, samples=1000, dms=4, step=10.0, subbands=4, coarse=4, coarse_step=40.0,
                       min_freq=300.0, bandwidth=2.0, zapped=(5,), seed=13)]
    jobs = Jobs(torch_cuda, ws)
    compact = [True, True, False]
    with Batch() as b:
        for plan, d_in, d_mid, d_out in zip(jobs.plans, jobs.inputs, jobs.mids, jobs.outs):
            b.add_two_step(plan, d_in.data_ptr(), d_mid.data_ptr(), d_out.data_ptr())
        for chunks in (0, 2):
            b.set_pipeline(chunks)
            info = b.info()
            assert info["solo_jobs"] == 1 and info["compact_intermediate"] == 1
            assert [b.job_placement(j) for j in range(3)] == [
                {"beside": False, "compact": True}, {"beside": False, "compact": True},
                {"beside": True, "compact": False}]
            want_units = sum(w.obs.nr_synthesized_beams * w.obs.nr_dms_subbanding * w.obs.nr_dms *
                             w.obs.nr_samples_per_batch for w in ws)
            assert info["units"] == want_units
            for use_graph in (False, True):
                b.run(use_graph=use_graph, stream=stream_of(torch_cuda))
                jobs.check(f"pipeline {chunks} graph {use_graph}", compact=compact)
                jobs.reset()
        b.run_mode(DD_STEP_ONE, stream=stream_of(torch_cuda))
        b.run_mode(DD_STEP_TWO, stream=stream_of(torch_cuda))
        jobs.check("per kind", compact=compact)


def test_batch_configuration_whose_float_windows_overflow_lds(torch_cuda):
    """A step-two configuration of 8 wavefronts x 16 samples and unroll 16
    over a job whose step-two shifts spread about 200 samples across a
    64-DM tile: 16 float row windows per buffer would not fit in LDS.  The
    batch does not fail: the staged kernels stage fewer rows per barrier (or
    the job runs beside them), with the float intermediate and with uint16,
    and the results equal the oracle's."""
    w = workloads.tiny(channels=256, samples=1500, dms=64, step=0.05, subbands=16, coarse=4, coarse_step=0.5,
                       min_freq=600.0, bandwidth=1.0, seed=38)
    orc = oracle.Oracle(w.obs, 0)
    s2 = orc.get_shifts_step_two()
    spread = orc.shift(0.0, 0.05, 63, float(s2[0])) - orc.shift(0.0, 0.05, 0, float(s2[0]))
    assert 150 <= spread <= 260, spread
    conf_two = Conf(64, 8, 16, 8, 16, 1)
    for compact in (False, True):
        jobs = Jobs(torch_cuda, [w, tiny_two_step_workloads()[1]])
        with Batch() as b:
            b.set_conf(DD_STEP_ONE, STAGED_SMALL)
            b.set_conf(DD_STEP_TWO, conf_two)
            b.set_compact_intermediate(compact)
            for plan, d_in, d_mid, d_out in zip(jobs.plans, jobs.inputs, jobs.mids, jobs.outs):
                b.add_two_step(plan, d_in.data_ptr(), d_mid.data_ptr(), d_out.data_ptr())
            place = [b.job_placement(j) for j in range(2)]
            b.run(stream=stream_of(torch_cuda))
            jobs.check(f"wide float windows, compact {compact}", compact=[p["compact"] for p in place])


def test_batch_runs_wide_single_step_jobs_beside_the_kernels(torch_cuda):
    ws = [workloads.tiny(channels=64, samples=700, dms=12, step=0.9, min_freq=1300.0, bandwidth=1.0, seed=41),
          workloads.tiny(channels=64, samples=2000, dms=12, step=20.0, min_freq=300.0, bandwidth=2.0, zapped=(5,),
                         seed=12)]
    jobs = Jobs(torch_cuda, ws, two_step=False)
    with Batch() as b:
        b.set_conf(DD_SINGLE_STEP, STAGED_SMALL)
        for plan, d_in, d_out in zip(jobs.plans, jobs.inputs, jobs.outs):
            b.add_single_step(plan, d_in.data_ptr(), d_out.data_ptr())
        assert b.info()["solo_jobs"] == 1
        for use_graph in (False, True):
            b.run(use_graph=use_graph, stream=stream_of(torch_cuda))
            jobs.check(f"graph {use_graph}")
            jobs.reset()
        b.run_mode(DD_SINGLE_STEP, stream=stream_of(torch_cuda))
        jobs.check("per kind")
