"""The synthetic workloads the engine is measured and tested on.

Each workload is an observation plus the seeded choices that define its input
(zapped channels, injected pulses, per-beam sizes).  All random choices use a
splitmix64 stream so they are identical on every machine; the bulk input
bytes are generated on the device by ``dd_synth_*`` from (seed, row, column).
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace
from typing import List, Tuple

import numpy as np

from . import Observation

MASK64 = (1 << 64) - 1
PADDING = 128  # bytes; one MI355X cache line


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014): a seeded 64-bit integer stream."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def below(self, n: int) -> int:
        return self.next() % n

    def sample(self, n: int, k: int) -> List[int]:
        """k distinct values of range(n), by a partial Fisher-Yates shuffle."""
        pool = list(range(n))
        for i in range(k):
            j = i + self.below(n - i)
            pool[i], pool[j] = pool[j], pool[i]
        return sorted(pool[:k])


@dataclass
class Pulse:
    dm_index: int
    dm_value: float
    t0: int
    width: int
    amplitude: int = 40


@dataclass
class Workload:
    name: str
    obs: Observation
    zapped: List[int] = field(default_factory=list)
    pulses: List[Pulse] = field(default_factory=list)
    noise: str = "gaussian"  # or "uniform"
    seed: int = 0

    def zapped_flags(self) -> List[int]:
        flags = [0] * self.obs.nr_channels
        for c in self.zapped:
            flags[c] = 1
        return flags


ARTS_CHANNELS = 1536
ARTS_MIN_FREQ = 1250.0
ARTS_BANDWIDTH = 0.1953125
SAMPLING_TIME = 0.00008192


def plain_128() -> Workload:
    """128 channels from 1400 MHz at 0.78125 MHz, 4096 samples, 128 DMs of
    0.5 pc/cc, channels 17 and 90 zapped, uniform uint8 input."""
    obs = Observation(
        nr_channels=128, min_freq=1400.0, channel_bandwidth=0.78125, sampling_time=SAMPLING_TIME,
        nr_samples_per_batch=4096, nr_dms=128, first_dm=0.0, dm_step=0.5,
    )
    return Workload("plain-128ch-128dm-4096s", obs, zapped=[17, 90], noise="uniform", seed=1)


def _arts_choices(seed: int, nr_dms: int, dm_step: float, samples: int) -> Tuple[List[int], List[Pulse]]:
    rng = SplitMix64(seed)
    zapped = rng.sample(ARTS_CHANNELS, round(0.05 * ARTS_CHANNELS))
    pulses = []
    for _ in range(8):
        dm_index = rng.below(nr_dms)
        width = 1 + rng.below(16)
        t0 = rng.below(samples - width)
        # the DM trial's value in the reference's float arithmetic: first + dm * step
        dm_value = float(np.float32(0.0) + np.float32(dm_index) * np.float32(dm_step))
        pulses.append(Pulse(dm_index, dm_value, t0, width))
    return zapped, pulses


def arts_single() -> Workload:
    """ARTS-sized single step: 1536 channels, 24576 samples, 2048 DMs of 0.2."""
    obs = Observation(
        nr_channels=ARTS_CHANNELS, min_freq=ARTS_MIN_FREQ, channel_bandwidth=ARTS_BANDWIDTH,
        sampling_time=SAMPLING_TIME, nr_samples_per_batch=24576, nr_dms=2048, first_dm=0.0, dm_step=0.2,
    )
    zapped, pulses = _arts_choices(2, 2048, 0.2, 24576)
    return Workload("arts-1536ch-2048dm-24576s", obs, zapped=zapped, pulses=pulses, seed=2)


def arts_two_step(nr_dms_coarse: int = 64, samples: int = 24576, seed: int = 3, beams: int = 1) -> Workload:
    """ARTS two-step: 32 subbands, Dc coarse DMs of 6.4, 32 fine DMs of 0.2."""
    obs = Observation(
        nr_channels=ARTS_CHANNELS, min_freq=ARTS_MIN_FREQ, channel_bandwidth=ARTS_BANDWIDTH,
        sampling_time=SAMPLING_TIME, nr_samples_per_batch=samples, nr_dms=32, first_dm=0.0, dm_step=0.2,
        nr_subbands=32, nr_dms_subbanding=nr_dms_coarse, first_dm_subbanding=0.0, dm_step_subbanding=6.4,
        nr_beams=beams, nr_synthesized_beams=beams,
    )
    zapped, pulses = _arts_choices(seed, 32 * nr_dms_coarse, 0.2, samples)
    name = f"arts-two-step-32sb-{nr_dms_coarse}x32dm-{samples}s"
    return Workload(name, obs, zapped=zapped, pulses=pulses, seed=seed)


def batched_beams(nr_beams: int = 40, seed: int = 4) -> List[Workload]:
    """40 synthesized beams, 1536 channels each, DMs from {512, 1024, 2048,
    4096} and samples from {8192, 16384, 24576, 32768}, two-step with 32
    subbands, Dc = D/32 coarse DMs of 6.4 and 32 fine DMs of 0.2."""
    rng = SplitMix64(seed)
    out = []
    for b in range(nr_beams):
        dms = (512, 1024, 2048, 4096)[rng.below(4)]
        samples = (8192, 16384, 24576, 32768)[rng.below(4)]
        w = arts_two_step(dms // 32, samples, seed=1000 + b)
        w.name = f"beam{b}-{dms}dm-{samples}s"
        out.append(w)
    return out


def tiny(channels=16, samples=64, dms=8, step=2.0, subbands=1, coarse=0, coarse_step=0.0, beams=1, sbeams=1,
         min_freq=1400.0, bandwidth=2.0, zapped=(), seed=5, first=0.0, coarse_first=0.0) -> Workload:
    """Small shapes for exhaustive parity checks (first, coarse_first: the
    first DM of the fine / single-step and of the coarse DM range)."""
    obs = Observation(
        nr_channels=channels, min_freq=min_freq, channel_bandwidth=bandwidth, sampling_time=SAMPLING_TIME,
        nr_samples_per_batch=samples, nr_dms=dms, first_dm=first, dm_step=step, nr_subbands=subbands,
        nr_dms_subbanding=coarse, first_dm_subbanding=coarse_first, dm_step_subbanding=coarse_step,
        nr_beams=beams, nr_synthesized_beams=sbeams,
    )
    return Workload(f"tiny-{channels}ch-{dms}dm-{samples}s", obs, zapped=list(zapped), noise="uniform", seed=seed)
