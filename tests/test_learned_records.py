"""Host logic of the learned tuner: features, packaged sweep records and the
offline pick (no device)."""
import math

import numpy as np

from dedispersion_amd import learned


def test_packaged_records_cover_every_configuration_of_12_layouts():
    recs = learned.load_records()
    groups = learned.by_layout(recs)
    assert len(groups) == 12
    assert all(len(v) == 720 for v in groups.values())
    assert all(r["ms"] > 0 for r in recs)


def test_feature_row_has_12_finite_features():
    r = learned.load_records()[0]
    f = learned.feature_row(r["layout"], learned.conf_dict(r["conf"]))
    assert len(f) == len(learned.FEATURES) == 12
    assert all(math.isfinite(x) for x in f)


class Perfect:
    """A model that knows the recorded timings."""

    def __init__(self, recs):
        self.t = {tuple(r["conf"]): r["ms"] for r in recs}

    def predict(self, layout, confs):
        return np.array([self.t[tuple(c[k] for k in learned.CONF_KEYS)] for c in confs])


class Reversed(Perfect):
    def predict(self, layout, confs):
        return -super().predict(layout, confs)


def test_offline_pick_bounds():
    recs = next(iter(learned.by_layout(learned.load_records()).values()))
    best = learned.evaluate_offline(Perfect(recs), recs, 0.1)
    assert best["ratio"] == 1.0 and best["timed"] == 72
    worst = learned.evaluate_offline(Reversed(recs), recs, 0.1)
    assert worst["ratio"] > 1.0
    exhaustive = learned.evaluate_offline(None, recs, 0.1)
    assert exhaustive["ratio"] == 1.0 and exhaustive["timed"] == 720 and abs(exhaustive["time_fraction"] - 1) < 1e-9
    r = learned.random_pick_ratio(recs, 0.1, trials=200)
    assert 1.0 <= r <= worst["ratio"]
