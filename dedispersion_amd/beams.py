"""Device-resident sets of beams for batched two-step dedispersion.

A ``BeamSet`` holds, for every workload of a list (one beam each, sizes free
to differ), a plan with the workload's zapped channels, the seeded uint8
input synthesized on the device (noise plus the workload's pulses), and the
intermediate and output buffers; ``batch`` runs all of them through one
``Batch`` (one persistent launch per step), ``run_per_job`` through one
launch per beam and step.  Device memory comes from torch.
"""
from __future__ import annotations

from typing import List, Optional

from . import DD_STEP_ONE, DD_STEP_TWO, Batch, Conf, Plan, synth_gaussian_u8, synth_uniform_u8
from .workloads import PADDING, Workload


class BeamSet:
    def __init__(self, torch, workloads: List[Workload], padding: int = PADDING, confs: Optional[dict] = None,
                 pulses: bool = True, input_bits: int = 8, tuned_plans: bool = True, chunks: Optional[int] = None):
        """confs: {mode: Conf} for the batch; None = the tuned batch
        configurations packaged with the library (the tuned pipelined pair
        and chunk count when there is one, else the best of each kind, else
        the defaults).
        input_bits 1, 2 or 4: packed input rows of uniform random fields
        (the workloads' noise model and pulses apply to 8-bit input only).
        tuned_plans: each plan takes its step-one and step-two configurations
        from the packaged tuned-configuration files by its DM counts (the
        reference's lookup), which the one-launch-per-beam path runs.
        chunks: the batch's pipeline depth (None = the packaged tuning's;
        0 = one persistent launch per kernel kind, in order)."""
        self.torch = torch
        self.workloads = workloads
        self.padding = padding
        self.plans: List[Plan] = []
        self.inputs, self.mids, self.outs = [], [], []
        s = torch.cuda.current_stream().cuda_stream
        self.stream = s
        tuned = {}
        if tuned_plans:
            from .tuner import read_tuned_file, tuned_file_path

            tuned = {name: read_tuned_file(tuned_file_path(name)) for name in ("step_one", "step_two")}
        for w in workloads:
            o = w.obs
            if o.nr_dms_subbanding == 0:
                raise ValueError(f"{w.name}: a beam set runs the two-step path")
            plan = Plan(o, padding, input_bits=input_bits)
            if tuned_plans:
                for mode, name in ((DD_STEP_ONE, "step_one"), (DD_STEP_TWO, "step_two")):
                    try:
                        plan.apply_tuned(mode, tuned[name])
                    except KeyError:
                        pass
            flags = [0] * plan.pad(o.nr_channels, 4)
            for c in w.zapped:
                flags[c] = 1
            plan.set_zapped_channels(flags)
            z = plan.sizes
            d_in = torch.zeros(z.sub_input_elements, dtype=torch.uint8, device="cuda")
            rows = o.nr_beams * o.nr_channels
            if input_bits != 8:
                per = 8 // input_bits
                synth_uniform_u8(d_in.data_ptr(), rows, -(-z.nr_dispersed_samples_subbanding // per), z.sub_in_pitch,
                                 w.seed, s)
            elif w.noise == "uniform":
                synth_uniform_u8(d_in.data_ptr(), rows, z.nr_dispersed_samples_subbanding, z.sub_in_pitch, w.seed, s)
            else:
                synth_gaussian_u8(d_in.data_ptr(), rows, z.nr_dispersed_samples_subbanding, z.sub_in_pitch, 128.0,
                                  16.0, w.seed, s)
            if pulses and input_bits == 8:
                for p in w.pulses:
                    plan.inject_pulse(d_in.data_ptr(), z.sub_in_pitch, z.nr_dispersed_samples_subbanding, p.dm_value,
                                      p.t0, p.width, p.amplitude, s)
            self.plans.append(plan)
            self.inputs.append(d_in)
            self.mids.append(torch.empty(z.intermediate_elements, dtype=torch.float32, device="cuda"))
            self.outs.append(torch.empty(z.sub_output_elements, dtype=torch.float32, device="cuda"))
        self.batch = Batch()
        if confs is None:
            from .tuner import load_batch_confs, load_batch_pipeline

            pipe = load_batch_pipeline()
            if pipe is not None:
                confs = pipe["confs"]
                self.batch.set_pipeline(pipe["chunks"])
                for mode, nbytes in pipe["lds_reserve"].items():
                    self.batch.set_lds_reserve(mode, nbytes)
            else:
                confs = load_batch_confs()
        if confs:
            for mode, conf in confs.items():
                self.batch.set_conf(mode, conf)
        if chunks is not None:
            self.batch.set_pipeline(chunks)
        for plan, d_in, d_mid, d_out in zip(self.plans, self.inputs, self.mids, self.outs):
            self.batch.add_two_step(plan, d_in.data_ptr(), d_mid.data_ptr(), d_out.data_ptr())
        torch.cuda.synchronize()

    def set_conf(self, mode: int, conf: Conf) -> None:
        self.batch.set_conf(mode, conf)

    def run(self, use_graph: bool = True) -> None:
        self.batch.run(use_graph, self.stream)

    def run_mode(self, mode: int) -> None:
        self.batch.run_mode(mode, self.stream)

    def step_bytes(self, intermediate_bytes: Optional[int] = None) -> dict:
        """Compulsory HBM bytes of each step over the whole set (padded rows):
        step one reads the input and writes the intermediate, step two reads
        the intermediate and writes the output.  The intermediate element is
        2 bytes when the batch keeps it compact (uint16), else 4."""
        if intermediate_bytes is None:
            intermediate_bytes = 2 if self.info()["compact_intermediate"] else 4
        one = two = 0
        for plan in self.plans:
            z = plan.sizes
            one += z.sub_input_elements + intermediate_bytes * z.intermediate_elements
            two += intermediate_bytes * z.intermediate_elements + 4 * z.sub_output_elements
        return {DD_STEP_ONE: one, DD_STEP_TWO: two}

    def run_per_job(self) -> None:
        self.batch.run_per_job(self.stream)

    def run_per_beam_same_format(self) -> None:
        """The batch's own kernels, configurations and intermediate format,
        launched one beam at a time: for every beam, step one then step two
        as persistent launches over that beam's tiles alone (no graph).  The
        time against the batch separates what batching gains from what the
        uint16 intermediate gains."""
        if not hasattr(self, "_single"):
            self._single = []
            for plan, d_in, d_mid, d_out in zip(self.plans, self.inputs, self.mids, self.outs):
                b = Batch()
                for mode in (DD_STEP_ONE, DD_STEP_TWO):
                    b.set_conf(mode, self.batch.get_conf(mode))
                b.set_pipeline(0)
                b.add_two_step(plan, d_in.data_ptr(), d_mid.data_ptr(), d_out.data_ptr())
                self._single.append(b)
        for b in self._single:
            b.run_mode(DD_STEP_ONE, self.stream)
            b.run_mode(DD_STEP_TWO, self.stream)

    def info(self) -> dict:
        return self.batch.info()

    def units(self) -> int:
        """DM trials x samples x beams of the final outputs."""
        return self.info()["units"]

    def bytes_moved(self, intermediate_bytes: Optional[int] = None) -> int:
        """Compulsory HBM bytes of one pass: input read once, intermediate
        written and read once, output written once (padded rows)."""
        b = self.step_bytes(intermediate_bytes)
        return b[DD_STEP_ONE] + b[DD_STEP_TWO]

    def close(self) -> None:
        for b in getattr(self, "_single", []):
            b.close()
        self._single = []
        self.batch.close()
        for p in self.plans:
            p.close()
        self.plans, self.inputs, self.mids, self.outs = [], [], [], []
