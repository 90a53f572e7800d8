"""Benchmark of the MI355X dedispersion engine (one JSON line on stdout).

A step is one pass of the hot path over one batch with the input already
resident in HBM.  The metric is DM-trial·samples per second: one unit is one
element of the final dedispersed output (beam x DM trial x sample).

Default workload: 40 synthesized beams of 1536 channels with seeded uneven
DM counts {512..4096} and sample counts {8192..32768}, two-step subband
dedispersion (32 subbands, coarse DMs of 6.4, 32 fine DMs of 0.2), run as a
batch replayed from a HIP graph: one persistent launch of step one over the
tiles of all beams, then one of step two (--chunks N >= 2 runs the pipelined
form instead: the beams split into N chunks, step one of a chunk beside step
two of the previous one on two streams).  The per-kernel durations are those
of the two persistent kernels of the timed graph, launched the same way.
The same work with one launch per beam and step, each beam with its plan's
tuned configuration and the reference's float intermediate (the reference's
launch pattern), is timed beside it.

Each persistent kernel's average duration is measured with HIP events on the
stream it runs on and turned into an HBM-roofline fraction from its
algorithmic bytes (input read once, intermediate written and read once,
output written once, padded rows).  The batch keeps its intermediate as
uint16 sums (2 bytes per element instead of the reference's 4); the
fraction is given for the executed data flow and, beside it, for the
reference's float-intermediate data flow.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # bytes/s, MI355X HBM3E


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="batched-beams",
                    choices=["batched-beams", "arts-two-step", "arts-single", "detect-peaks"])
    ap.add_argument("--input-bits", type=int, default=8, choices=[1, 2, 4, 8],
                    help="bits per input sample (1, 2, 4 = packed rows of uniform random fields)")
    ap.add_argument("--chunks", type=int, default=0,
                    help="batch pipeline depth (0 = one persistent launch per kernel kind, in order; "
                         ">= 2 = the pipelined form; -1 = the packaged tuning's)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline work")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run --steps plain device-to-device copies of BYTES (for scaling HBM counters)")
    ap.add_argument("--read-calibration", type=int, default=0, metavar="BYTES",
                    help="only time --steps plain streaming reads of BYTES (dd_read_probe): the read bandwidth a "
                         "streaming kernel reaches on this device")
    return ap.parse_args()


def read_calibration(args, torch):
    """The read bandwidth of a plain streaming read of a given size
    (dd_read_probe: coalesced 16-byte loads, one partial sum per thread), the
    practical ceiling of a read-bound kernel beside the 8 TB/s peak."""
    from dedispersion_amd import check, load

    n = args.read_calibration // 16 * 16
    x = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    sink = torch.empty(2097152, dtype=torch.float32, device="cuda")
    s = torch.cuda.current_stream().cuda_stream

    def read():
        check(load().dd_read_probe(x.data_ptr(), n, sink.data_ptr(), s))

    for _ in range(2):
        read()
    torch.cuda.synchronize()
    a, b = events(torch, 2)
    a.record()
    for _ in range(args.steps):
        read()
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / args.steps
    print(json.dumps({"read_bytes": n, "steps": args.steps, "ms_per_read": ms,
                      "GB_per_s_read": n / (ms / 1e3) / 1e9}), flush=True)


def copy_calibration(args, torch):
    """Plain device copies of a known size: the FETCH_SIZE / WRITE_SIZE a
    profiler reports for them calibrate the counters against bytes moved."""
    n = args.copy_calibration
    src = torch.randint(0, 255, (n,), dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    a, b = events(torch, 2)
    a.record()
    for _ in range(args.steps):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / args.steps
    print(json.dumps({"copy_bytes": n, "steps": args.steps, "ms_per_copy": ms,
                      "GB_per_s_read_plus_write": 2 * n / (ms / 1e3) / 1e9}), flush=True)


class EventTimer:
    """Per-kernel durations from HIP events on one stream."""

    def __init__(self, torch):
        self.torch = torch
        self.pairs = {}

    def around(self, name, fn):
        t = self.torch
        a, b = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        self.pairs.setdefault(name, []).append((a, b))

    def averages_ms(self):
        return {k: sum(a.elapsed_time(b) for a, b in v) / len(v) for k, v in self.pairs.items()}


def setup_two_step(torch, w, bits=8):
    from dedispersion_amd import Plan, synth_gaussian_u8, synth_uniform_u8

    plan = Plan(w.obs, 128, input_bits=bits)
    flags = w.zapped_flags() + [0] * (plan.pad(w.obs.nr_channels, 4) - w.obs.nr_channels)
    plan.set_zapped_channels(flags)
    z = plan.sizes
    s = torch.cuda.current_stream().cuda_stream
    d_in = torch.zeros(z.sub_input_elements, dtype=torch.uint8, device="cuda")
    rows = w.obs.nr_beams * w.obs.nr_channels
    if bits == 8:
        synth_gaussian_u8(d_in.data_ptr(), rows, z.nr_dispersed_samples_subbanding, z.sub_in_pitch, 128.0, 16.0,
                          w.seed, s)
        for p in w.pulses:
            plan.inject_pulse(d_in.data_ptr(), z.sub_in_pitch, z.nr_dispersed_samples_subbanding, p.dm_value, p.t0,
                              p.width, p.amplitude, s)
    else:
        per = 8 // bits
        synth_uniform_u8(d_in.data_ptr(), rows, -(-z.nr_dispersed_samples_subbanding // per), z.sub_in_pitch,
                         w.seed, s)
    d_mid = torch.empty(z.intermediate_elements, dtype=torch.float32, device="cuda")
    d_out = torch.empty(z.sub_output_elements, dtype=torch.float32, device="cuda")
    o = w.obs
    units = o.nr_synthesized_beams * o.nr_dms_subbanding * o.nr_dms * z.nr_samples
    C_live = o.nr_channels - len(w.zapped)
    bytes_one = o.nr_beams * o.nr_channels * z.sub_in_pitch + 4 * o.nr_beams * o.nr_dms_subbanding * o.nr_subbands * z.nr_samples_subbanding
    bytes_two = 4 * o.nr_beams * o.nr_dms_subbanding * o.nr_subbands * z.nr_samples_subbanding + 4 * units
    kernels = {
        "step_one": (lambda: plan.step_one(d_in.data_ptr(), d_mid.data_ptr(), stream=s), bytes_one),
        "step_two": (lambda: plan.step_two(d_mid.data_ptr(), d_out.data_ptr(), stream=s), bytes_two),
    }
    return plan, kernels, units, (d_in, d_mid, d_out)


def setup_single(torch, w, bits=8):
    from dedispersion_amd import Plan, synth_gaussian_u8, synth_uniform_u8

    plan = Plan(w.obs, 128, input_bits=bits)
    flags = w.zapped_flags() + [0] * (plan.pad(w.obs.nr_channels, 4) - w.obs.nr_channels)
    plan.set_zapped_channels(flags)
    z = plan.sizes
    s = torch.cuda.current_stream().cuda_stream
    d_in = torch.zeros(z.input_elements, dtype=torch.uint8, device="cuda")
    if bits == 8:
        synth_gaussian_u8(d_in.data_ptr(), w.obs.nr_channels, z.nr_dispersed_samples, z.in_pitch, 128.0, 16.0,
                          w.seed, s)
    else:
        synth_uniform_u8(d_in.data_ptr(), w.obs.nr_channels, -(-z.nr_dispersed_samples // (8 // bits)), z.in_pitch,
                         w.seed, s)
    d_out = torch.empty(z.output_elements, dtype=torch.float32, device="cuda")
    o = w.obs
    units = o.nr_synthesized_beams * o.nr_dms * z.nr_samples
    nbytes = o.nr_beams * o.nr_channels * z.in_pitch + 4 * units
    kernels = {"single_step": (lambda: plan.dedisperse(d_in.data_ptr(), d_out.data_ptr(), stream=s), nbytes)}
    return plan, kernels, units, (d_in, d_out)


def repeat_for(fn, seconds: float, max_passes: int = 100):
    """Runs fn until about `seconds` have passed (at least once); returns
    (passes, elapsed seconds)."""
    passes, dt = 0, 0.0
    while passes < max_passes and (passes == 0 or dt < 0.9 * seconds):
        t = time.perf_counter()
        fn()
        dt += time.perf_counter() - t
        passes += 1
    return passes, dt


def host_cpu() -> dict:
    """The host's CPU: model (/proc/cpuinfo), logical CPUs, the CPUs this
    process may run on, and the threads the baseline uses: all of those, or
    the CPU share the environment grants one process (OMP_NUM_THREADS) when
    that is smaller."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    try:
        share = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        share = 0
    threads = max(1, min(affinity, share) if share > 0 else affinity)
    return {"model": model, "logical_cpus": os.cpu_count(), "affinity_cpus": affinity,
            "omp_num_threads": share or None, "threads": threads}


def cpu_baseline(ws, mode: str, seconds: float, bits: int = 8):
    """The oracle on the host cores over a bounded sample of the workload,
    and, beside it, on one thread (as the reference's CPU functions run)
    over a shorter sample.  For the two-step mix the sample takes the same
    fraction of the coarse DMs of every beam (step one over all subbands,
    then step two over all fine DMs), so it has the mix's shapes."""
    import numpy as np

    import oracle

    rng = np.random.default_rng(0)
    inputs = {}

    def beam_input(w, orc):
        if w.name not in inputs:
            o = w.obs
            pitch = orc.sub_in_pitch if mode == "two-step" else orc.in_pitch
            inputs.clear()  # one beam's input at a time
            inputs[w.name] = rng.integers(0, 256, size=o.nr_beams * o.nr_channels * pitch, dtype=np.uint8)
        return inputs[w.name]

    def flags_of(w):
        f = np.zeros(oracle.pad(w.obs.nr_channels, 32), dtype=np.uint32)
        f[w.zapped] = 1
        return f

    def timed(threads: int, budget: float):
        if mode == "two-step":
            def run(w, n):
                o = w.obs
                orc = oracle.Oracle(o, 128, threads=threads, input_bits=bits)
                inp, flags = beam_input(w, orc), flags_of(w)
                t = time.perf_counter()
                mid = orc.step_one(inp, flags, rows=(0, n))
                orc.step_two(mid.ravel(), rows=(0, n * o.nr_dms))
                return time.perf_counter() - t, n * o.nr_dms * (o.nr_samples_per_batch // o.downsampling)

            # calibrate on one coarse DM of the first beam, then take the same
            # fraction of every beam's coarse DMs
            dt, units = run(ws[0], 1)
            rate = units / max(dt, 1e-4)
            total = sum(w.obs.nr_dms_subbanding * w.obs.nr_dms * w.obs.nr_samples_per_batch for w in ws)
            frac = min(1.0, budget * rate / total)
            spent, done = 0.0, 0
            for w in ws:
                n = max(1, round(frac * w.obs.nr_dms_subbanding))
                dt, units = run(w, n)
                spent += dt
                done += units
            sample = (f"{frac:.3f} of the coarse DMs of each of the {len(ws)} beams of the mix (at least one; step "
                      f"one over all subbands, then step two over all fine DMs)")
            return done / spent, sample, spent
        w = ws[0]
        o = w.obs
        orc = oracle.Oracle(o, 128, threads=threads, input_bits=bits)
        inp, flags = beam_input(w, orc), flags_of(w)
        t = time.perf_counter()
        orc.dedispersion(inp, flags, rows=(0, threads))
        per = (time.perf_counter() - t) / threads
        n = max(threads, min(o.nr_dms, int(budget / max(per, 1e-4) * threads)))
        passes, dt = repeat_for(lambda: orc.dedispersion(inp, flags, rows=(0, n)), budget)
        units = passes * n * (o.nr_samples_per_batch // o.downsampling)
        return units / dt, f"{passes} pass(es) over {n} of {o.nr_dms} DM rows of {w.name}", dt

    host = host_cpu()
    threads = host["threads"]
    value, sample, dt = timed(threads, seconds)
    value1, sample1, dt1 = timed(1, seconds / 3)
    return {"value": value, "unit": "DM-trial*samples/s", "cores": threads, "kind": "port", "sample": sample,
            "seconds": round(dt, 3), "host": host,
            "single_thread": {"value": value1, "cores": 1, "sample": sample1, "seconds": round(dt1, 3)}}


def measured_traffic(workload: str, kernel: str, conf: str):
    """HBM bytes per launch of `kernel` under `conf` on `workload`, from the
    calibrated rocprofv3 FETCH_SIZE / WRITE_SIZE runs committed under
    profiles/ (None when no run matches)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(workload, {}).get(kernel)
    except (OSError, ValueError):
        return None
    if not rec or rec.get("conf") != conf:
        return None
    return rec["traffic_bytes"]


def events(torch, n):
    return [torch.cuda.Event(enable_timing=True) for _ in range(n)]


def run_batched(args, torch):
    from dedispersion_amd import DD_STEP_ONE, DD_STEP_TWO, workloads
    from dedispersion_amd.beams import BeamSet

    ws = workloads.batched_beams()
    bs = BeamSet(torch, ws, input_bits=args.input_bits, chunks=None if args.chunks < 0 else args.chunks)
    info = bs.info()
    units = info["units"]
    log(f"batched beams: {info}")
    for _ in range(args.warmup):
        bs.run(use_graph=True)
        bs.run_per_job()
        bs.run_per_beam_same_format()
    torch.cuda.synchronize()

    # timed region: K graph replays of the whole batch
    start, stop = events(torch, 2)
    t0 = time.perf_counter()
    start.record()
    for _ in range(args.steps):
        bs.run(use_graph=True)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(stop) / args.steps

    # per-kernel durations: each kernel kind alone (its persistent launch over
    # all beams), events around each
    names = {DD_STEP_ONE: "step_one_batch", DD_STEP_TWO: "step_two_batch"}
    pairs = {m: [] for m in names}
    for _ in range(args.steps):
        for m in names:
            a, b = events(torch, 2)
            a.record()
            bs.run_mode(m)
            b.record()
            pairs[m].append((a, b))
    torch.cuda.synchronize()
    avg = {m: sum(a.elapsed_time(b) for a, b in v) / len(v) for m, v in pairs.items()}

    # the batch's kernels and uint16 intermediate, one beam at a time
    a, b = events(torch, 2)
    a.record()
    for _ in range(args.steps):
        bs.run_per_beam_same_format()
    b.record()
    torch.cuda.synchronize()
    ms_same = a.elapsed_time(b) / args.steps

    # the reference's launch pattern: one launch per beam and step
    a, b = events(torch, 2)
    a.record()
    for _ in range(args.steps):
        bs.run_per_job()
    b.record()
    torch.cuda.synchronize()
    ms_plain = a.elapsed_time(b) / args.steps

    step_bytes = bs.step_bytes()
    dominant = max(avg, key=avg.get)
    achieved = step_bytes[dominant] / (avg[dominant] / 1e3)
    total_bytes = bs.bytes_moved()

    # a plain device copy (torch copy_) moving as many bytes as the dominant
    # kernel, half read and half written: the practical ceiling of a kernel
    # that reads and writes HBM, beside the 8 TB/s peak
    src = torch.empty(step_bytes[dominant] // 2, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    a, b = events(torch, 2)
    a.record()
    for _ in range(args.steps):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    copy_rate = 2 * src.numel() / (a.elapsed_time(b) / args.steps / 1e3)
    del src, dst
    reference_bytes = bs.bytes_moved(intermediate_bytes=4)  # float intermediate, the reference data flow
    value = units / (ms / 1e3)
    dms = sorted({w.obs.nr_dms_subbanding * w.obs.nr_dms for w in ws})
    workload = f"batched-beams-{len(ws)}x{ws[0].obs.nr_channels}ch-two-step"
    if args.input_bits != 8:
        workload += f"-{args.input_bits}bit"
    confs = {names[m]: bs.batch.get_conf(m).print() for m in names}
    samples = sorted({w.obs.nr_samples_per_batch for w in ws})
    line = {
        "metric": "DM-trial*samples/s",
        "value": value,
        "unit": "DM-trial*samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/f32",
        "data": ("synthetic (seeded Gaussian noise N(128,16) clipped to uint8 + 8 dispersed pulses per beam, "
                 "5% zapped channels per beam)" if args.input_bits == 8 else
                 f"synthetic (seeded uniform random {args.input_bits}-bit samples, packed, 5% zapped channels per beam)"),
        "config": {"workload": workload, "beams": len(ws), "input_bits": args.input_bits,
                   "channels": ws[0].obs.nr_channels, "subbands": ws[0].obs.nr_subbands, "dms_per_beam": dms,
                   "samples_per_beam": samples, "units_per_step": units, "padding": 128,
                   "launch": (f"pipelined: {info['pipeline_chunks']} chunks on 2 streams, tile-list grids, HIP graph"
                              if info["pipeline_chunks"] else "persistent work queues per kernel kind, HIP graph"),
                   "tiles": info["tiles"], "workgroups": info["workgroups"],
                   "conf": confs},
        "plain_per_beam": {"value": units / (ms_plain / 1e3), "ms_per_step": ms_plain,
                           "launch": "one launch per beam and step, each plan configured from the tuned files "
                                     "by its DM counts, float intermediate",
                           "launches_per_step": 2 * len(ws)},
        "batched_vs_plain": ms_plain / ms,
        "per_beam_same_format": {"value": units / (ms_same / 1e3), "ms_per_step": ms_same,
                                 "launch": "the batch's kernels, configurations and uint16 intermediate, one "
                                           "persistent launch per beam and step",
                                 "launches_per_step": 2 * len(ws)},
        "batched_vs_per_beam_same_format": ms_same / ms,
        "intermediate": "uint16" if info["compact_intermediate"] else "float32",
        "bytes_per_unit": total_bytes / units,
        "hbm_roofline_frac_path": total_bytes / (ms / 1e3) / HBM_PEAK,
        "reference_flow_bytes_per_unit": reference_bytes / units,
        "hbm_roofline_frac_reference_flow": reference_bytes / (ms / 1e3) / HBM_PEAK,
        "kernels_ms": {names[m]: v for m, v in avg.items()},
        "roofline": {"bound": "hbm", "kernel": names[dominant], "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": measured_traffic(workload, names[dominant], confs[names[dominant]]),
                     "algorithmic_bytes": step_bytes[dominant],
                     "copy_ceiling": {"GB_per_s": copy_rate / 1e9, "frac": achieved / copy_rate,
                                      "path_frac": total_bytes / (ms / 1e3) / copy_rate,
                                      "what": "a plain device copy (torch copy_) moving the dominant kernel's bytes"}},
        "wall_s": wall,
    }
    return line, ws, "two-step"


def run_detect(args, torch):
    """The step after the path: the boxcar peak search (widths 1..64) over
    every dedispersed row of the 40-beam batch's outputs, all 40 beams' rows
    in one persistent launch (dd_detect_peaks_batch); one launch per beam
    (dd_detect_peaks) is timed beside it.  The batch runs once beforehand to
    fill the outputs."""
    from dedispersion_amd import PEAK_DTYPE, detect_peaks, detect_peaks_batch, workloads
    from dedispersion_amd.beams import BeamSet

    ws = workloads.batched_beams()
    bs = BeamSet(torch, ws, input_bits=args.input_bits)
    bs.run(use_graph=False)
    s = torch.cuda.current_stream().cuda_stream
    jobs = []
    for plan, out in zip(bs.plans, bs.outs):
        o, z = plan.obs, plan.sizes
        rows = o.nr_synthesized_beams * o.nr_dms_subbanding * o.nr_dms
        peaks = torch.empty(rows * PEAK_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
        jobs.append((out.data_ptr(), rows, z.nr_samples, z.out_pitch, peaks))
    units = sum(r * n for _, r, n, _, _ in jobs)
    nbytes = sum(r * (4 * n + PEAK_DTYPE.itemsize) for _, r, n, _, _ in jobs)

    batch_jobs = [(ptr, r, n, pitch, peaks.data_ptr()) for ptr, r, n, pitch, peaks in jobs]

    def step():
        detect_peaks_batch(batch_jobs, 7, s)

    def per_beam():
        for ptr, r, n, pitch, peaks in jobs:
            detect_peaks(ptr, r, n, pitch, 7, peaks.data_ptr(), s)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        start, stop = events(torch, 2)
        t0 = time.perf_counter()
        start.record()
        for _ in range(args.steps):
            fn()
        stop.record()
        torch.cuda.synchronize()
        return start.elapsed_time(stop) / args.steps, time.perf_counter() - t0

    ms, wall = timed(step)
    per_beam_ms, _ = timed(per_beam)
    cpu = None if args.no_cpu_baseline else detect_cpu_baseline(torch, bs, jobs, args.cpu_seconds / 2)
    # the same bytes read by a plain streaming kernel: the practical ceiling
    from dedispersion_amd import check, load

    buf = torch.ones(nbytes // 16 * 4, dtype=torch.float32, device="cuda")
    sink = torch.empty(2097152, dtype=torch.float32, device="cuda")
    read_ms, _ = timed(lambda: check(load().dd_read_probe(buf.data_ptr(), buf.numel() * 4, sink.data_ptr(), s)))
    read_ceiling = buf.numel() * 4 / (read_ms / 1e3)
    del buf, sink
    achieved = nbytes / (ms / 1e3)
    line = {
        "metric": "DM-trial*samples/s",
        "value": units / (ms / 1e3),
        "unit": "DM-trial*samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "i32/i64/f64",
        "data": "the dedispersed outputs of the 40-beam batch (synthetic input)",
        "config": {"workload": "detect-peaks-40-beams", "rows": sum(r for _, r, _, _, _ in jobs),
                   "widths": [1 << j for j in range(7)], "launches_per_step": 1,
                   "per_beam_launches_ms": per_beam_ms},
        "roofline": {"bound": "hbm", "kernel": "detect_peaks", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": measured_traffic("detect-peaks-40-beams", "detect_peaks_batch",
                                                 "68-sample runs per lane, one wavefront per row, one launch"),
                     "algorithmic_bytes": nbytes,
                     "read_ceiling": {"GB_per_s": read_ceiling / 1e9, "frac": achieved / read_ceiling,
                                      "what": "dd_read_probe streaming the same number of bytes"}},
        "wall_s": wall,
    }
    if cpu is not None:
        line["cpu_baseline"] = cpu
    return line, None, "detect"


def detect_cpu_baseline(torch, bs, jobs, seconds: float):
    """The oracle's straight-loop peak search (one thread) over a bounded
    sample of the same rows: the first rows of every beam, the same number
    from each, sized to about `seconds` of CPU work.  The sampled rows' peaks
    are also compared with the device's, record for record."""
    import numpy as np

    import oracle
    from dedispersion_amd import PEAK_DTYPE

    def rows_of(b, k):
        ptr, r, n, pitch, peaks = jobs[b]
        host = bs.outs[b][: k * pitch].view(k, pitch).cpu().numpy()
        dev = peaks[: k * PEAK_DTYPE.itemsize].cpu().numpy().view(PEAK_DTYPE)
        return host, n, dev

    host, n, _ = rows_of(0, 2)
    t = time.perf_counter()
    oracle.detect_peaks(host, n, 7)
    rate = 2 * n / max(time.perf_counter() - t, 1e-4)  # samples per second
    mean_samples = sum(n for _, _, n, _, _ in jobs) / len(jobs)
    k = int(max(1, min(512, seconds * rate / (mean_samples * len(jobs)))))
    spent, done, mismatched = 0.0, 0, 0
    for b in range(len(jobs)):
        k_b = min(k, jobs[b][1])
        host, n, dev = rows_of(b, k_b)
        t = time.perf_counter()
        want = oracle.detect_peaks(host, n, 7)
        spent += time.perf_counter() - t
        done += k_b * n
        mismatched += int(np.count_nonzero((want["sample"] != dev["sample"]) | (want["width"] != dev["width"])))
    return {"value": done / spent, "unit": "DM-trial*samples/s", "cores": 1, "kind": "port",
            "sample": f"the first {k} rows of each of the {len(jobs)} beams' outputs (the oracle's straight loops)",
            "seconds": round(spent, 3), "sampled_rows_differing_from_device": mismatched}


def run_single_plan(args, torch):
    from dedispersion_amd import workloads

    if args.workload == "arts-two-step":
        w = workloads.arts_two_step()
        plan, kernels, units, keep = setup_two_step(torch, w, args.input_bits)
        mode = "two-step"
    else:
        w = workloads.arts_single()
        plan, kernels, units, keep = setup_single(torch, w, args.input_bits)
        mode = "single-step"
    torch.cuda.synchronize()

    timer = EventTimer(torch)
    for _ in range(args.warmup):
        for fn, _ in kernels.values():
            fn()
    torch.cuda.synchronize()
    start, stop = events(torch, 2)
    t0 = time.perf_counter()
    start.record()
    for _ in range(args.steps):
        for name, (fn, _) in kernels.items():
            timer.around(name, fn)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(stop) / args.steps
    value = units / (ms / 1e3)
    avg = timer.averages_ms()
    dominant = max(avg, key=avg.get)
    bytes_dom = kernels[dominant][1]
    achieved = bytes_dom / (avg[dominant] / 1e3)
    total_bytes = sum(b for _, b in kernels.values())
    line = {
        "metric": "DM-trial*samples/s",
        "value": value,
        "unit": "DM-trial*samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/f32" if mode == "two-step" else "u8",
        "data": ("synthetic (seeded Gaussian noise N(128,16) + 8 dispersed pulses, 5% zapped channels)"
                 if args.input_bits == 8 else
                 f"synthetic (seeded uniform random {args.input_bits}-bit samples, packed, 5% zapped channels)"),
        "config": {"workload": w.name if args.input_bits == 8 else f"{w.name}-{args.input_bits}bit", "mode": mode,
                   "input_bits": args.input_bits, "beams": w.obs.nr_synthesized_beams,
                   "channels": w.obs.nr_channels, "samples": w.obs.nr_samples_per_batch,
                   "dms": units // w.obs.nr_samples_per_batch // w.obs.nr_synthesized_beams, "padding": 128},
        "hbm_roofline_frac_path": total_bytes / (ms / 1e3) / HBM_PEAK,
        "kernels_ms": avg,
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None},
        "wall_s": wall,
    }
    return line, w, mode


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU per process; runs over several GPUs are not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    if args.copy_calibration:
        copy_calibration(args, torch)
        return
    if args.read_calibration:
        read_calibration(args, torch)
        return
    if args.workload == "batched-beams":
        line, w, mode = run_batched(args, torch)
    elif args.workload == "detect-peaks":
        line, w, mode = run_detect(args, torch)
    else:
        line, w, mode = run_single_plan(args, torch)
    if not args.no_cpu_baseline and w is not None:
        log("cpu baseline ...")
        line["cpu_baseline"] = cpu_baseline(w if isinstance(w, list) else [w], mode, args.cpu_seconds,
                                            args.input_bits)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
