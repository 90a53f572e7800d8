"""Exhaustive tuning harness for the HIP kernels.

Replaces the reference's DedispersionTuning program (src/DedispersionTuning.cpp):
it enumerates kernel configurations under the same kinds of constraints
(:201-294, here with HIP limits), times each one with HIP events through
``dd_tune_time`` (:384-395), reports the reference's metric
GFLOP/s = giga(adds)/seconds (:327-333, :412) next to DM-trial·samples/s,
and writes the best configuration per (device, DMs) in the format read by
readTunedDedispersionConf (src/Dedispersion.cpp:20-81):
``device nrDMs split local unroll t0 t1 t2 i0 i1 i2``.

Records are JSON lines; a sweep resumes by skipping configurations already
recorded for the same layout key.  Invalid configurations are recorded with
``"valid": false`` and the sweep continues (the reference skips a
configuration that fails to compile or runs out of resources, :343-347,
:396-409).
"""
from __future__ import annotations

import argparse
import ctypes
import itertools
import json
import os
import sys
import time
from dataclasses import asdict
from typing import Dict, Iterable, List, Optional

from . import DD_SINGLE_STEP, DD_STEP_ONE, DD_STEP_TWO, Conf, DDError, Plan, _lib, load, synth_gaussian_u8
from . import workloads as wl

MODES = {"single_step": DD_SINGLE_STEP, "step_one": DD_STEP_ONE, "step_two": DD_STEP_TWO}

SPACE = {
    "threads_d1": (1, 2, 4, 8),
    "items_d0": (4, 8, 16),
    "items_d1": (1, 2, 4, 8),
    "unroll": (1, 2, 4, 8, 16),
    "local": (0, 1, 2),
}


def enumerate_space(space: Dict = SPACE) -> List[Conf]:
    confs = []
    for t1, i0, i1, u, loc in itertools.product(space["threads_d1"], space["items_d0"], space["items_d1"],
                                                space["unroll"], space["local"]):
        if t1 * i1 > 64:
            continue
        confs.append(Conf(64, t1, i0, i1, u, loc))
    return confs


def layout_of(plan: Plan, mode: int, live_channels: int) -> Dict:
    """Features of the problem a configuration runs on."""
    o, z = plan.obs, plan.sizes
    if mode == DD_SINGLE_STEP:
        rows, dms, samples, step = live_channels, o.nr_dms, z.nr_samples, o.dm_step
        beams, groups = o.nr_synthesized_beams, 1
    elif mode == DD_STEP_ONE:
        rows, dms, samples, step = live_channels, o.nr_dms_subbanding, z.nr_samples_subbanding, o.dm_step_subbanding
        beams, groups = o.nr_beams, o.nr_subbands
    else:
        rows, dms, samples, step = o.nr_subbands, o.nr_dms, z.nr_samples, o.dm_step
        beams, groups = o.nr_synthesized_beams * o.nr_dms_subbanding, 1
    return {
        "mode": [k for k, v in MODES.items() if v == mode][0],
        "channels": o.nr_channels,
        "subbands": o.nr_subbands,
        "rows": rows,
        "dms": dms,
        "samples": samples,
        "beams": beams,
        "groups": groups,
        "dm_step": step,
        "max_shift": [z.max_shift, z.max_shift_step_one, z.max_shift_step_two][mode],
    }


def adds_of(layout: Dict) -> int:
    """Additions of one launch: the reference's op count (Tuning:327-333)."""
    if layout["mode"] == "step_one":
        return layout["beams"] * layout["dms"] * layout["rows"] * layout["samples"]
    return layout["beams"] * layout["dms"] * layout["rows"] * layout["samples"] * layout["groups"]


class Bench:
    """Device buffers and plan of one (workload, mode) layout."""

    def __init__(self, workload: wl.Workload, mode: int, torch, padding: int = 128):
        self.torch = torch
        self.workload = workload
        self.mode = mode
        self.plan = Plan(workload.obs, padding)
        flags = [0] * self.plan.pad(workload.obs.nr_channels, 4)
        for c in workload.zapped:
            flags[c] = 1
        self.plan.set_zapped_channels(flags)
        z = self.plan.sizes
        o = workload.obs
        s = torch.cuda.current_stream().cuda_stream
        self.stream = s
        if mode == DD_SINGLE_STEP:
            self.d_in = torch.zeros(z.input_elements, dtype=torch.uint8, device="cuda")
            synth_gaussian_u8(self.d_in.data_ptr(), o.nr_beams * o.nr_channels, z.nr_dispersed_samples, z.in_pitch,
                              128.0, 16.0, workload.seed, s)
            self.d_out = torch.empty(z.output_elements, dtype=torch.float32, device="cuda")
            units = o.nr_synthesized_beams * o.nr_dms * z.nr_samples
        elif mode == DD_STEP_ONE:
            self.d_in = torch.zeros(z.sub_input_elements, dtype=torch.uint8, device="cuda")
            synth_gaussian_u8(self.d_in.data_ptr(), o.nr_beams * o.nr_channels, z.nr_dispersed_samples_subbanding,
                              z.sub_in_pitch, 128.0, 16.0, workload.seed, s)
            self.d_out = torch.empty(z.intermediate_elements, dtype=torch.float32, device="cuda")
            units = o.nr_beams * o.nr_dms_subbanding * o.nr_subbands * z.nr_samples_subbanding
        else:
            self.d_in = torch.rand(z.intermediate_elements, dtype=torch.float32, device="cuda") * 6000.0
            self.d_out = torch.empty(z.sub_output_elements, dtype=torch.float32, device="cuda")
            units = o.nr_synthesized_beams * o.nr_dms_subbanding * o.nr_dms * z.nr_samples
        self.units = units
        self.layout = layout_of(self.plan, mode, o.nr_channels - len(workload.zapped))
        torch.cuda.synchronize()

    def time(self, conf: Conf, warmup: int = 2, iterations: int = 5) -> Optional[Dict]:
        lib = load()
        c = conf.to_c()
        mean, std = ctypes.c_float(), ctypes.c_float()
        status = lib.dd_tune_time(self.plan.handle, self.mode, ctypes.byref(c), self.d_in.data_ptr(),
                                  self.d_out.data_ptr(), warmup, iterations, self.stream, ctypes.byref(mean),
                                  ctypes.byref(std))
        if status != 0:
            return None
        return {"ms": mean.value, "std_ms": std.value}


def key_of(layout: Dict) -> str:
    return json.dumps(layout, sort_keys=True)


def load_records(path: str) -> List[Dict]:
    if not path or not os.path.exists(path):
        return []
    with open(path) as f:
        return [json.loads(line) for line in f if line.strip()]


def sweep(bench: Bench, confs: Iterable[Conf], records_path: Optional[str] = None, warmup: int = 2,
          iterations: int = 5, device: str = "MI355X", progress=None) -> List[Dict]:
    """Times every configuration (skipping those already recorded)."""
    done = {(r["layout_key"], tuple(r["conf"].values())) for r in load_records(records_path)}
    lkey = key_of(bench.layout)
    out = []
    f = open(records_path, "a") if records_path else None
    try:
        for conf in confs:
            ck = tuple(asdict(conf).values())
            if (lkey, ck) in done:
                continue
            t = bench.time(conf, warmup, iterations)
            rec = {"device": device, "workload": bench.workload.name, "layout": bench.layout, "layout_key": lkey,
                   "conf": asdict(conf), "valid": t is not None}
            if t is not None:
                rec.update(t)
                rec["units_per_s"] = bench.units / (t["ms"] / 1e3)
                rec["gflops"] = adds_of(bench.layout) / (t["ms"] / 1e3) / 1e9
            out.append(rec)
            if f:
                f.write(json.dumps(rec) + "\n")
                f.flush()
            if progress:
                progress(rec)
    finally:
        if f:
            f.close()
    return out


def best_of(records: List[Dict]) -> Optional[Dict]:
    valid = [r for r in records if r.get("valid")]
    return min(valid, key=lambda r: r["ms"]) if valid else None


def tuned_conf_line(device: str, nr_dms: int, conf: Dict) -> str:
    """One line of the reference's tuned-configuration file."""
    c = Conf(**conf)
    return f"{device} {nr_dms} {c.print()}"


def write_tuned_file(path: str, entries: List[str]) -> None:
    with open(path, "w") as f:
        f.write("# device nrDMs split local unroll threadsD0 threadsD1 threadsD2 itemsD0 itemsD1 itemsD2\n")
        for e in entries:
            f.write(e + "\n")


def read_tuned_file(path: str) -> Dict[str, Dict[int, Conf]]:
    """The reference's reader (src/Dedispersion.cpp:20-81): lines starting
    with a letter are ``device nrDMs split local unroll t0 t1 t2 i0 i1 i2``;
    others (comments, blank lines) are skipped; the first line for a
    (device, nrDMs) is kept (std::map::insert, :71-78).  local is the local
    mode 0-3.  Raises OSError if the file cannot be opened (the reference
    throws AstroData::FileError)."""
    out: Dict[str, Dict[int, Conf]] = {}
    with open(path) as f:
        for line in f:
            if not line[:1].isalpha():
                continue
            v = line.split()
            split, local, unroll, t0, t1, _t2, i0, i1, _i2 = (int(x) for x in v[2:11])
            out.setdefault(v[0], {}).setdefault(int(v[1]), Conf(t0, t1, i0, i1, unroll, local, split))
    return out


def tuned_lines_from_records(records: List[Dict], device: str = "MI355X") -> Dict[str, List[str]]:
    """Reference-format lines per kernel kind from sweep records: for each
    DM count, the configuration with the least summed time relative to each
    layout's best, over the layouts with that DM count that timed it."""
    from . import learned

    out: Dict[str, List[str]] = {}
    groups: Dict[tuple, Dict[str, List[Dict]]] = {}
    for key, recs in learned.by_layout(records).items():
        lay = recs[0]["layout"]
        groups.setdefault((lay["mode"], lay["dms"]), {})[key] = recs
    for (mode, dms), layouts in sorted(groups.items()):
        score: Dict[tuple, float] = {}
        count: Dict[tuple, int] = {}
        for recs in layouts.values():
            best = min(r["ms"] for r in recs)
            for r in recs:
                c = tuple(r["conf"])
                score[c] = score.get(c, 0.0) + r["ms"] / best
                count[c] = count.get(c, 0) + 1
        full = [c for c in score if count[c] == len(layouts)]
        pick = min(full, key=lambda c: score[c])
        out.setdefault(mode, []).append(tuned_conf_line(device, dms, asdict(Conf(*pick))))
    return out


TUNED_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning")
# Exhaustive sweeps of the plan configurations (every kernel of the library)
# over the 128-channel, ARTS and 40-beam layouts: the source of the
# reference-format tuned files and the learned tuner's training and
# evaluation records.
PLAN_RECORDS_PATH = os.path.join(TUNED_DIR, "sweep_mi355x.jsonl.gz")


def tuned_file_path(mode: str, device: str = "MI355X") -> str:
    """Packaged reference-format file of one kernel kind (mode name as in MODES)."""
    return os.path.join(TUNED_DIR, f"{mode}_{device.lower()}.conf")


def select_tuned_conf(tuned: Dict[str, Dict[int, Conf]], device: str, nr_dms: int) -> Conf:
    """tuned[device][nrDMs], the lookup a pipeline makes in the reader's map;
    KeyError when either is absent."""
    return tuned[device][nr_dms]


def _beam_shapes():
    """Layouts of the 40-beam workload (two-step, coarse DMs x 32 fine DMs of
    0.2, samples per beam), held out of the learned tuner's training set."""
    return [wl.arts_two_step(dc, s, seed=500 + i) for i, (dc, s) in
            enumerate(((16, 8192), (32, 16384), (128, 32768), (64, 8192)))]


WORKLOADS = {
    "plain-128": (lambda: [wl.plain_128()], ["single_step"]),
    "arts-single": (lambda: [wl.arts_single()], ["single_step"]),
    "arts-two-step": (lambda: [wl.arts_two_step()], ["step_one", "step_two"]),
    "beam-shapes": (_beam_shapes, ["step_one", "step_two"]),
}


def staged_space(space: Dict = SPACE) -> List[Conf]:
    """Configurations that select the staged pipeline (what a batch runs)."""
    return [c for c in enumerate_space(space) if c.local == 1 and c.unroll <= 2 * c.threads_d1]


def sweep_batch(beams, mode: int, confs: Iterable[Conf], iterations: int = 3, progress=None) -> List[Dict]:
    """Times one kernel kind of a batch (beams.BeamSet) under each
    configuration with HIP events around its persistent launch."""
    import torch

    out = []
    for conf in confs:
        try:
            beams.set_conf(mode, conf)
            beams.run_mode(mode)  # builds the queues, warms up
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(iterations):
                beams.run_mode(mode)
            b.record()
            torch.cuda.synchronize()
            rec = {"conf": asdict(conf), "valid": True, "ms": a.elapsed_time(b) / iterations}
        except DDError as e:
            rec = {"conf": asdict(conf), "valid": False, "error": str(e)}
        out.append(rec)
        if progress:
            progress(rec)
    return out


BATCH_TUNED_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "batch_mi355x.json")


def load_batch_confs(path: str = BATCH_TUNED_PATH) -> Dict[int, Conf]:
    """Tuned batch configurations per kernel kind ({} when there are none)."""
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return {}
    return {MODES[k]: Conf(**v["conf"]) for k, v in d.get("confs", {}).items()}


def time_pipelined(beams, iterations: int = 5) -> float:
    """Milliseconds per graph replay of the whole batch (pipelined or not,
    as the batch is set)."""
    import torch

    beams.run(use_graph=True)  # builds queues and graph, warms up
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(iterations):
        beams.run(use_graph=True)
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / iterations


def pipeline_candidates(recs: List[Dict], top: int, small_lds: int) -> List[Dict]:
    """The `top` fastest configurations of one kind alone, plus the
    `small_lds` fastest among those staging at most 4 rows per barrier (a
    smaller LDS footprint lets the other kind's workgroups share a CU)."""
    valid = sorted((r for r in recs if r.get("valid")), key=lambda r: r["ms"])
    out = valid[:top]
    out += [r for r in valid if r["conf"]["unroll"] <= 4 and r not in out][:small_lds]
    return out


def tune_batch(path: Optional[str], iterations: int = 3, top: int = 6, small_lds: int = 4,
               chunk_counts=(2, 3, 4, 6), reserves=(0, 24576, 32768, 40960, 53248, 65536, 81920)) -> Dict:
    """Tunes the 40-beam batch: an exhaustive sweep of the staged
    configurations of each kernel kind alone (persistent launches), then the
    pipelined graph over pairs of the best candidates of each kind and chunk
    counts.  The best of each kind alone and the best pipelined pair are
    written to `path`."""
    import torch

    from .beams import BeamSet

    bs = BeamSet(torch, wl.batched_beams(), confs={})
    best, cands = {}, {}
    try:
        bs.batch.set_pipeline(0)
        bs.run(use_graph=False)  # a valid intermediate for step two
        for name in ("step_one", "step_two"):
            t0 = time.time()
            recs = sweep_batch(bs, MODES[name], staged_space(), iterations)
            b = best_of(recs)
            best[name] = {"conf": {k: v for k, v in b["conf"].items()}, "ms": b["ms"], "swept": len(recs)}
            cands[name] = pipeline_candidates(recs, top, small_lds)
            bs.set_conf(MODES[name], Conf(**b["conf"]))
            print(f"batch {name}: {len(recs)} configurations in {time.time() - t0:.1f}s; best {b['conf']} "
                  f"{b['ms']:.3f} ms", file=sys.stderr, flush=True)
        t0 = time.time()
        pipe = None
        n = 0
        for r1 in cands["step_one"]:
            for r2 in cands["step_two"]:
                bs.set_conf(DD_STEP_ONE, Conf(**r1["conf"]))
                bs.set_conf(DD_STEP_TWO, Conf(**r2["conf"]))
                for chunks in chunk_counts:
                    bs.batch.set_pipeline(chunks)
                    ms = time_pipelined(bs, iterations + 2)
                    n += 1
                    if pipe is None or ms < pipe["ms"]:
                        pipe = {"step_one": r1["conf"], "step_two": r2["conf"], "chunks": chunks, "ms": ms}
        pipe["swept"] = n
        print(f"pipelined: {n} pairs x chunk counts in {time.time() - t0:.1f}s; best {pipe}", file=sys.stderr,
              flush=True)
        # LDS reserves: fewer workgroups of one kind per compute unit leave
        # room for the other kind's beside them
        t0 = time.time()
        bs.set_conf(DD_STEP_ONE, Conf(**pipe["step_one"]))
        bs.set_conf(DD_STEP_TWO, Conf(**pipe["step_two"]))
        pipe["lds_reserve"] = {"step_one": 0, "step_two": 0}
        n = 0
        for r1 in reserves:
            for r2 in reserves:
                bs.batch.set_lds_reserve(DD_STEP_ONE, r1)
                bs.batch.set_lds_reserve(DD_STEP_TWO, r2)
                for chunks in chunk_counts:
                    bs.batch.set_pipeline(chunks)
                    try:
                        ms = time_pipelined(bs, iterations + 2)
                    except DDError:
                        continue
                    n += 1
                    if ms < pipe["ms"]:
                        pipe.update(ms=ms, chunks=chunks, lds_reserve={"step_one": r1, "step_two": r2})
        print(f"LDS reserves: {n} settings in {time.time() - t0:.1f}s; best {pipe}", file=sys.stderr, flush=True)
    finally:
        bs.close()
    out = {"workload": "batched-beams-40x1536ch-two-step", "device": "MI355X", "confs": best, "pipeline": pipe}
    if path:
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
    return out


def load_batch_pipeline(path: str = BATCH_TUNED_PATH) -> Optional[Dict]:
    """The tuned pipelined setting: {"confs": {mode: Conf}, "chunks": n,
    "lds_reserve": {mode: bytes}}, or None when the file holds none."""
    try:
        with open(path) as f:
            p = json.load(f).get("pipeline")
    except (OSError, ValueError):
        return None
    if not p:
        return None
    r = p.get("lds_reserve", {})
    return {"confs": {DD_STEP_ONE: Conf(**p["step_one"]), DD_STEP_TWO: Conf(**p["step_two"])},
            "chunks": int(p["chunks"]),
            "lds_reserve": {DD_STEP_ONE: int(r.get("step_one", 0)), DD_STEP_TWO: int(r.get("step_two", 0))}}


def learned_tune(bench: Bench, model, budget: float, iterations: int) -> Dict:
    """Times the model's best-predicted fraction `budget` of the space."""
    from . import learned

    def time_fn(c):
        t = bench.time(Conf(**c), iterations=iterations)
        return None if t is None else t["ms"]

    return learned.pick(model, bench.layout, [asdict(c) for c in enumerate_space()], time_fn, budget)


def main(argv=None):
    ap = argparse.ArgumentParser(description="HIP kernel tuning: exhaustive sweep or learned pruning")
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + ["batched-beams"], default="arts-two-step")
    ap.add_argument("--modes", default=None, help="comma list of single_step,step_one,step_two")
    ap.add_argument("--records", default=None, help="JSON-lines file to append to (resumable)")
    ap.add_argument("--tuned", default=None, help="write the best configurations in the reference format")
    ap.add_argument("--iterations", type=int, default=5)
    ap.add_argument("--device", default="MI355X")
    ap.add_argument("--learned", action="store_true",
                    help="train the cost model on the recorded sweeps and time only the best-predicted fraction")
    ap.add_argument("--budget", type=float, default=0.1, help="fraction of the space timed with --learned")
    ap.add_argument("--train-records", default=None, help="sweep records for --learned (default: the packaged ones)")
    ap.add_argument("--from-records", default=None, metavar="RECORDS",
                    help="write the reference-format files of every kernel kind from sweep records (no device)")
    args = ap.parse_args(argv)
    if args.from_records:
        from . import learned

        for mode, lines in tuned_lines_from_records(learned.load_records(args.from_records), args.device).items():
            path = tuned_file_path(mode, args.device)
            write_tuned_file(path, lines)
            print(f"{path}: {len(lines)} entries", file=sys.stderr)
        return
    import torch

    if args.workload == "batched-beams":
        tune_batch(args.tuned or BATCH_TUNED_PATH, args.iterations)
        return
    make, modes = WORKLOADS[args.workload]
    if args.modes:
        modes = args.modes.split(",")
    model = None
    if args.learned:
        from . import learned

        records = learned.load_records(args.train_records or learned.RECORDS_PATH)
        model = learned.CostModel()
        losses = model.fit(records)
        print(f"cost model trained on {len(records)} records: loss {losses[0]:.3f} -> {losses[-1]:.4f}",
              file=sys.stderr, flush=True)
    lines = []
    for w in make():
        for m in modes:
            bench = Bench(w, MODES[m], torch)
            t0 = time.time()
            if model is not None:
                b = learned_tune(bench, model, args.budget, args.iterations)
                print(f"{w.name} {m}: timed {b['timed']} of {b['candidates']} in {time.time() - t0:.1f}s; best "
                      f"{b['conf']} {b['ms']} ms", file=sys.stderr, flush=True)
                conf = b["conf"]
            else:
                recs = sweep(bench, enumerate_space(), args.records, iterations=args.iterations, device=args.device)
                b = best_of(recs)
                print(f"{w.name} {m}: {len(recs)} configurations in {time.time() - t0:.1f}s; best "
                      f"{b['conf'] if b else None} {b['ms'] if b else None} ms", file=sys.stderr, flush=True)
                conf = b["conf"] if b else None
            if conf:
                conf = {k: v for k, v in conf.items() if k in Conf.__dataclass_fields__}
                lines.append(tuned_conf_line(args.device, bench.layout["dms"], conf))
            del bench
            torch.cuda.empty_cache()
    if args.tuned:
        write_tuned_file(args.tuned, lines)


if __name__ == "__main__":
    main()
