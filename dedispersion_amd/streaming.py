"""Streaming request loop over the native streaming engine (dd_stream_*).

Beams send chunks of S samples per channel ([C][S] uint8); chunk arrivals of
each beam are a Poisson process whose mean interval is the chunk's duration
(S x sampling time) divided by a speed-up factor (1.5 = data arrives 1.5x
faster than real time).  The loop submits each chunk when it arrives, keeps
polling the engine, and records every request's latency from arrival to
observed completion on the device.  The engine keeps each beam's history on
the device, admits requests into slots under an HBM budget and overlaps
requests on several HIP streams.

    python -m dedispersion_amd.streaming --beams 16 --chunks 8 --speedup 1.5
"""
from __future__ import annotations

import argparse
import ctypes
import json
import sys
import time
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from . import Plan, synth_gaussian_u8
from ._lib import DDStreamEvent, DDStreamInfo, check, load
from .workloads import ARTS_BANDWIDTH, ARTS_CHANNELS, ARTS_MIN_FREQ, PADDING, SAMPLING_TIME, SplitMix64
from . import Observation


def chunk_observation(samples: int = 12500, coarse: int = 64, fine: int = 32, channels: int = ARTS_CHANNELS,
                      subbands: int = 32, min_freq: float = ARTS_MIN_FREQ, bandwidth: float = ARTS_BANDWIDTH,
                      coarse_step: float = 6.4, fine_step: float = 0.2) -> Observation:
    """One beam, one chunk: 1536 channels, 12500 samples (1.024 s), two-step
    64 coarse DMs of 6.4 x 32 fine DMs of 0.2 (2048 DMs) by default."""
    return Observation(
        nr_channels=channels, min_freq=min_freq, channel_bandwidth=bandwidth, sampling_time=SAMPLING_TIME,
        nr_samples_per_batch=samples, nr_dms=fine, first_dm=0.0, dm_step=fine_step, nr_subbands=subbands,
        nr_dms_subbanding=coarse, first_dm_subbanding=0.0, dm_step_subbanding=coarse_step,
    )


@dataclass
class Arrival:
    time: float  # seconds after the start
    beam: int
    chunk: int


def poisson_arrivals(beams: int, chunks: int, chunk_seconds: float, speedup: float, seed: int = 9) -> List[Arrival]:
    """`chunks` arrivals per beam with exponential gaps of mean
    chunk_seconds / speedup (splitmix64 uniforms, inverse transform)."""
    rng = SplitMix64(seed)
    out = []
    mean = chunk_seconds / speedup
    for b in range(beams):
        t = 0.0
        for k in range(chunks):
            u = (rng.next() >> 11) * (1.0 / (1 << 53))
            t += -mean * np.log1p(-u)
            out.append(Arrival(t, b, k))
    out.sort(key=lambda a: (a.time, a.beam))
    return out


class StreamEngine:
    def __init__(self, plan: Plan, beams: int, hbm_budget: int, max_slots: int = 64, hip_streams: int = 4):
        self.lib = load()
        self.plan = plan
        h = ctypes.c_void_p()
        check(self.lib.dd_stream_create(ctypes.byref(h), plan.handle, beams, hbm_budget, max_slots, hip_streams))
        self.handle = h
        self._events = (DDStreamEvent * 64)()

    def info(self) -> Dict:
        i = DDStreamInfo()
        check(self.lib.dd_stream_get_info(self.handle, ctypes.byref(i)))
        return {f: getattr(i, f) for f, _ in i._fields_}

    def submit(self, beam: int, host_ptr: int, host_pitch: int, tag: int) -> None:
        check(self.lib.dd_stream_submit(self.handle, beam, host_ptr, host_pitch, tag))

    def progress(self) -> List[Dict]:
        n = ctypes.c_uint32()
        check(self.lib.dd_stream_progress(self.handle, self._events, len(self._events), ctypes.byref(n)))
        return [{f: getattr(self._events[i], f) for f, _ in DDStreamEvent._fields_} for i in range(n.value)]

    def release(self, slot: int) -> None:
        check(self.lib.dd_stream_release(self.handle, slot))

    def drain(self) -> None:
        check(self.lib.dd_stream_drain(self.handle))

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.dd_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ChunkSource:
    """Pinned host chunks [C][S] uint8, Gaussian noise N(128, 16) seeded by
    (seed, beam, chunk), generated on the device and copied to the host."""

    def __init__(self, torch, channels: int, samples: int, seed: int = 77):
        self.torch, self.channels, self.samples, self.seed = torch, channels, samples, seed
        self.scratch = torch.empty(channels * samples, dtype=torch.uint8, device="cuda")
        self.cache: Dict = {}

    def chunk_seed(self, beam: int, chunk: int) -> int:
        return (self.seed * 1_000_003 + beam * 65_537 + chunk) & ((1 << 63) - 1)

    def get(self, beam: int, chunk: int):
        key = (beam, chunk)
        if key not in self.cache:
            s = self.torch.cuda.current_stream().cuda_stream
            synth_gaussian_u8(self.scratch.data_ptr(), self.channels, self.samples, self.samples, 128.0, 16.0,
                              self.chunk_seed(beam, chunk), s)
            host = self.torch.empty(self.channels * self.samples, dtype=self.torch.uint8, pin_memory=True)
            host.copy_(self.scratch)
            self.cache[key] = host
        return self.cache[key]


def run_loop(engine: StreamEngine, arrivals: Sequence[Arrival], source: ChunkSource,
             consume: Optional[Callable[[Dict], None]] = None, realtime: bool = True, poll_s: float = 2e-4) -> Dict:
    """Submits each arrival at its time (or at once when realtime is False),
    polls completions, releases slots after `consume`; returns latencies."""
    for a in arrivals:  # chunks are prepared before the clock starts
        source.get(a.beam, a.chunk)
    source.torch.cuda.synchronize()
    pending = list(arrivals)
    t_arrival: Dict[int, float] = {}
    lat: Dict[int, float] = {}
    i = 0
    start = time.perf_counter()
    while len(lat) < len(arrivals):
        now = time.perf_counter() - start
        while i < len(pending) and (not realtime or pending[i].time <= now):
            a = pending[i]
            tag = i
            t_arrival[tag] = a.time if realtime else now
            engine.submit(a.beam, source.get(a.beam, a.chunk).data_ptr(), source.samples, tag)
            i += 1
        for ev in engine.progress():
            done = time.perf_counter() - start
            lat[ev["tag"]] = done - t_arrival[ev["tag"]]
            if consume:
                consume(ev)
            engine.release(ev["slot"])
        if poll_s and i < len(pending) and realtime:
            time.sleep(min(poll_s, max(0.0, pending[i].time - (time.perf_counter() - start))))
        elif poll_s:
            time.sleep(poll_s / 4)
    wall = time.perf_counter() - start
    v = np.array([lat[k] for k in sorted(lat)])
    return {"requests": len(v), "wall_s": wall, "p50_ms": float(np.percentile(v, 50) * 1e3),
            "p99_ms": float(np.percentile(v, 99) * 1e3), "max_ms": float(v.max() * 1e3),
            "mean_ms": float(v.mean() * 1e3)}


def main(argv=None):
    ap = argparse.ArgumentParser(description="streaming two-step dedispersion with Poisson chunk arrivals")
    ap.add_argument("--beams", type=int, default=16)
    ap.add_argument("--chunks", type=int, default=8, help="chunks per beam")
    ap.add_argument("--samples", type=int, default=12500)
    ap.add_argument("--speedup", type=float, default=1.5, help="arrival rate over real time")
    ap.add_argument("--budget-gb", type=float, default=8.0, help="HBM budget of the engine")
    ap.add_argument("--streams", type=int, default=4)
    ap.add_argument("--seed", type=int, default=9)
    args = ap.parse_args(argv)
    import torch

    obs = chunk_observation(args.samples)
    plan = Plan(obs, PADDING)
    engine = StreamEngine(plan, args.beams, int(args.budget_gb * 2**30), 64, args.streams)
    info = engine.info()
    chunk_seconds = args.samples * obs.sampling_time
    arrivals = poisson_arrivals(args.beams, args.chunks, chunk_seconds, args.speedup, args.seed)
    source = ChunkSource(torch, obs.nr_channels, args.samples)
    stats = run_loop(engine, arrivals, source)
    units = len(arrivals) * obs.nr_dms_subbanding * obs.nr_dms * args.samples
    data_seconds = len(arrivals) * chunk_seconds
    out = {"workload": f"streaming-{args.beams}beams-{obs.nr_channels}ch-{args.samples}s-chunks-"
                       f"{obs.nr_dms_subbanding * obs.nr_dms}dm-two-step",
           "arrival_speedup": args.speedup, "engine": info, **stats,
           "throughput_units_per_s": units / stats["wall_s"], "data_seconds": data_seconds,
           "realtime_factor": data_seconds / stats["wall_s"]}
    engine.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
