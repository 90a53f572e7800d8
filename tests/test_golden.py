"""The oracle against the committed golden fixtures (tests/golden/)."""
import json
import os

import numpy as np

from tests.golden import make_fixtures as mf

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_fixture_bytes_are_splitmix64():
    # splitmix64 with state 0: first outputs 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4
    b = mf.fixture_bytes(2, 0)
    assert b.tolist() == [0xAF, 0xF4]


def test_tiny_single_step_fixture():
    z = np.load(os.path.join(HERE, "tiny_single.npz"))
    _, a = mf.tiny_single()
    np.testing.assert_array_equal(a["input"], z["input"])
    np.testing.assert_array_equal(a["output"], z["output"])


def test_tiny_two_step_fixture():
    z = np.load(os.path.join(HERE, "tiny_two_step.npz"))
    _, a = mf.tiny_two_step()
    np.testing.assert_array_equal(a["intermediate"], z["intermediate"])
    np.testing.assert_array_equal(a["output"], z["output"])


def test_plain_128_digest():
    want = json.load(open(os.path.join(HERE, "plain_128.json")))
    w, _, _, out = mf.plain_128()
    got = mf.plain_128_digest(out, w.obs.nr_samples_per_batch)
    assert got["sha256"] == want["sha256"]
    assert got["slices"] == want["slices"]
