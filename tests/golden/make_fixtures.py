"""Writes the golden fixtures of the parity tests (run from the repo root:
``python tests/golden/make_fixtures.py``).

The reference commits no golden vectors; its one known answer is a constant
input (src/DedispersionTest.cpp:213, 260, 304), which the tests check in
closed form.  These fixtures freeze the oracle's outputs -- after the oracle
has been pinned by the closed-form, impulse, invariant and brute-force tests
of tests/test_oracle.py -- on seeded inputs, so that the device path and any
later oracle change are compared with the same stored vectors:

  tiny_single.npz    16 channels x 64 samples, 8 DMs, channel 3 zapped,
                     2 beams / 3 synthesized beams with a mixed mapping
  tiny_two_step.npz  32 channels, 4 subbands, 3 coarse x 4 fine DMs,
                     channels 0 and 9 zapped
  plain_128.json     the 128-channel, 4096-sample, 128-DM single-step
                     configuration (channels 17 and 90 zapped): SHA-256 of
                     the output rows, row sums and sample slices

Inputs are bytes of a splitmix64 stream (fixture_bytes), so they are
regenerated identically on any machine.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from dedispersion_amd import workloads  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def fixture_bytes(n: int, seed: int) -> np.ndarray:
    """Low byte of splitmix64(seed + i + 1) for i < n (Steele et al. 2014)."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z & np.uint64(0xFF)).astype(np.uint8)


def tiny_single():
    w = workloads.tiny(channels=16, samples=64, dms=8, step=2.0, beams=2, sbeams=3, zapped=(3,))
    orc = oracle.Oracle(w.obs, 0)
    flags = np.zeros(16, np.uint32)
    flags[3] = 1
    mapping = orc.default_map(16)
    mapping[1, :8] = 0
    mapping[2, ::3] = 1
    inp = fixture_bytes(2 * 16 * orc.in_pitch, 11)
    out = orc.dedispersion(inp, flags, mapping)
    return w, dict(input=inp, zapped=flags, mapping=mapping, output=out)


def tiny_two_step():
    w = workloads.tiny(channels=32, samples=96, dms=4, step=1.5, subbands=4, coarse=3, coarse_step=6.0,
                       zapped=(0, 9))
    orc = oracle.Oracle(w.obs, 0)
    flags = np.zeros(32, np.uint32)
    flags[[0, 9]] = 1
    inp = fixture_bytes(32 * orc.sub_in_pitch, 12)
    mid = orc.step_one(inp, flags)
    out = orc.step_two(mid.ravel())
    return w, dict(input=inp, zapped=flags, intermediate=mid, output=out)


def plain_128_digest(out: np.ndarray, samples: int) -> dict:
    rows = np.ascontiguousarray(out.reshape(out.shape[-2], -1)[:, :samples])
    return {
        "sha256": hashlib.sha256(rows.astype("<f4").tobytes()).hexdigest(),
        "row_sums": [float(x) for x in rows.astype(np.float64).sum(axis=1)],
        "slices": {str(r): [float(x) for x in rows[r, :16]] for r in (0, 1, 64, 127)},
    }


def plain_128():
    w = workloads.plain_128()
    orc = oracle.Oracle(w.obs, 128)
    flags = np.zeros(oracle.pad(128, 32), np.uint32)
    flags[[17, 90]] = 1
    inp = fixture_bytes(128 * orc.in_pitch, 13)
    out = orc.dedispersion(inp, flags)
    return w, inp, flags, out


def main():
    _, a = tiny_single()
    np.savez_compressed(os.path.join(HERE, "tiny_single.npz"), **a)
    _, b = tiny_two_step()
    np.savez_compressed(os.path.join(HERE, "tiny_two_step.npz"), **b)
    w, inp, flags, out = plain_128()
    d = plain_128_digest(out, w.obs.nr_samples_per_batch)
    d.update({"input": "fixture_bytes(128 * in_pitch, 13), in_pitch = 4352 (S_disp 4306 padded to 128 bytes)",
              "zapped": [17, 90], "padding": 128})
    with open(os.path.join(HERE, "plain_128.json"), "w") as f:
        json.dump(d, f, indent=1)


if __name__ == "__main__":
    main()
