"""Learned tuner: a cost model that ranks kernel configurations so that only
the most promising ones are timed.

The reference tunes by timing every configuration
(src/DedispersionTuning.cpp:201-427).  Here the exhaustive sweep
(``tuner.sweep``) records timings on the device; small networks -- 12
features, three ReLU layers of 64 units, one output, FP32 -- are trained on
those records by the library's own HIP kernels (``dd_mlp_*``), predict the
run time of every configuration of a new layout, and the tuner times only
the best-predicted fraction.  The exhaustive sweep stays the fallback (no
model, or a budget of 1) and the yardstick (tests compare with it).

Features of a (layout, configuration) pair:
  log2(1 + the shift spread of a DM tile), log2 threads_d1, log2 items_d0,
  log2 items_d1, log2 unroll, local (0 cache / 1 pipelined LDS / 2 LDS),
  log2 rows summed per output
  ("channels"), log2 independent row groups per launch ("subbands"),
  log2 DMs and log2 samples -- both counted in tiles of the configuration
  (ceil(DMs / (threads_d1 items_d1)), ceil(samples / (64 items_d0))), which
  is how the grid sees them --, log2 DM step and log2(1 + largest shift).
The first slot would hold threads_d0, which is 64 (one wavefront along the
samples) in every configuration of these kernels and so carries nothing; it
holds instead the spread of the shifts across one DM tile of the
configuration, largest shift x DMs per tile / DMs, which sets how much wider
than the sample tile every staged row window is.

Target: log(run time / additions of the launch), the time per addition, so
that what is learned on one problem size carries to another; the ranking
inside a layout is unchanged by the per-layout constant.  The model is the
mean of a few networks from different initial weights, which steadies the
ranking on layouts far from the training ones.
"""
from __future__ import annotations

import ctypes
import gzip
import json
import math
import os
from typing import Callable, Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

from ._lib import check, load

FEATURES = ("log2_1p_tile_spread", "log2_threads_d1", "log2_items_d0", "log2_items_d1", "log2_unroll", "local",
            "log2_rows", "log2_groups", "log2_dm_tiles", "log2_sample_tiles", "log2_dm_step", "log2_1p_max_shift")
CONF_KEYS = ("threads_d0", "threads_d1", "items_d0", "items_d1", "unroll", "local")
RECORDS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "sweep_mi355x.jsonl.gz")


def feature_row(layout: Dict, conf: Dict) -> List[float]:
    mode = layout["mode"]
    rows = layout["rows"] / layout["groups"] if mode == "step_one" else layout["rows"]
    groups = layout["beams"] * layout["groups"]
    tile_dms = conf["threads_d1"] * conf["items_d1"]
    return [
        math.log2(1.0 + layout["max_shift"] * min(tile_dms, layout["dms"]) / max(layout["dms"], 1)),
        math.log2(conf["threads_d1"]),
        math.log2(conf["items_d0"]),
        math.log2(conf["items_d1"]),
        math.log2(conf["unroll"]),
        float(conf["local"]),
        math.log2(max(rows, 1.0)),
        math.log2(max(groups, 1)),
        math.log2(math.ceil(max(layout["dms"], 1) / (conf["threads_d1"] * conf["items_d1"]))),
        math.log2(math.ceil(max(layout["samples"], 1) / (64 * conf["items_d0"]))),
        math.log2(max(layout["dm_step"], 1e-3)),
        math.log2(1.0 + layout["max_shift"]),
    ]


def additions(layout: Dict) -> float:
    """Row additions of one launch of the layout (the reference's op count,
    src/DedispersionTuning.cpp:327-333)."""
    return float(layout["rows"]) * layout["dms"] * layout["samples"] * layout["beams"] * (
        layout["groups"] if layout["mode"] == "single_step" else 1)


def layout_key(layout: Dict) -> str:
    return json.dumps(layout, sort_keys=True)


def compact(record: Dict) -> Dict:
    """The fields of a sweep record the tuner learns from."""
    return {"workload": record["workload"], "layout": record["layout"],
            "conf": [record["conf"][k] for k in CONF_KEYS], "ms": record["ms"]}


def conf_dict(c) -> Dict:
    return dict(zip(CONF_KEYS, c)) if isinstance(c, (list, tuple)) else dict(c)


def load_records(path: str = RECORDS_PATH) -> List[Dict]:
    """Valid timed records (compact or full sweep records, .jsonl or .jsonl.gz)."""
    opener = gzip.open if path.endswith(".gz") else open
    out = []
    with opener(path, "rt") as f:
        for line in f:
            if not line.strip():
                continue
            r = json.loads(line)
            if r.get("valid", True) and r.get("ms"):
                out.append(r if isinstance(r["conf"], list) else compact(r))
    return out


def save_records(records: Iterable[Dict], path: str) -> None:
    with gzip.open(path, "wt") as f:
        for r in records:
            f.write(json.dumps(compact(r) if isinstance(r["conf"], dict) else r, sort_keys=True) + "\n")


def by_layout(records: Sequence[Dict]) -> Dict[str, List[Dict]]:
    out: Dict[str, List[Dict]] = {}
    for r in records:
        out.setdefault(layout_key(r["layout"]), []).append(r)
    return out


class Network:
    """One device network (dd_mlp) of the cost model."""

    def __init__(self, hidden: int, layers: int, seed: int):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.dd_mlp_create(ctypes.byref(h), len(FEATURES), hidden, layers))
        self.handle = h
        check(self.lib.dd_mlp_init(self.handle, seed))
        n = ctypes.c_uint32()
        check(self.lib.dd_mlp_param_count(self.handle, ctypes.byref(n)))
        self.n_params = n.value

    def close(self):
        if getattr(self, "handle", None):
            self.lib.dd_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        check(self.lib.dd_mlp_get_params(self.handle, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        check(self.lib.dd_mlp_set_params(self.handle, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))

    def train(self, dx, dt, n: int, steps: int, lr: float, stream: int) -> np.ndarray:
        losses = np.zeros(steps, dtype=np.float32)
        check(self.lib.dd_mlp_train(self.handle, dx.data_ptr(), dt.data_ptr(), n, steps, lr,
                                    losses.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), stream or None))
        return losses

    def forward(self, dx, n: int, dy, stream: int) -> None:
        check(self.lib.dd_mlp_forward(self.handle, dx.data_ptr(), n, dy.data_ptr(), stream or None))


class CostModel:
    """`members` networks (12 -> hidden x layers -> 1) from seeds seed,
    seed + 1, ...; the prediction is the mean of their outputs."""

    def __init__(self, hidden: int = 64, layers: int = 3, seed: int = 0, members: int = 3):
        self.hidden, self.layers = hidden, layers
        self.nets = [Network(hidden, layers, seed + k) for k in range(members)]
        self.n_params = self.nets[0].n_params
        self.x_mean = np.zeros(len(FEATURES))
        self.x_std = np.ones(len(FEATURES))
        self.y_mean, self.y_std = 0.0, 1.0

    def close(self):
        for n in self.nets:
            n.close()
        self.nets = []

    def save(self, path: str) -> None:
        np.savez(path, params=np.stack([n.get_params() for n in self.nets]), x_mean=self.x_mean, x_std=self.x_std,
                 y=np.array([self.y_mean, self.y_std]), shape=np.array([self.hidden, self.layers]))

    @classmethod
    def load_file(cls, path: str) -> "CostModel":
        z = np.load(path)  # plain arrays only (allow_pickle stays False)
        params = z["params"]
        m = cls(int(z["shape"][0]), int(z["shape"][1]), members=len(params))
        for net, p in zip(m.nets, params):
            net.set_params(p)
        m.x_mean, m.x_std = z["x_mean"], z["x_std"]
        m.y_mean, m.y_std = float(z["y"][0]), float(z["y"][1])
        return m

    def _x(self, rows: np.ndarray) -> np.ndarray:
        return ((rows - self.x_mean) / self.x_std).astype(np.float32)

    @staticmethod
    def rows_of(records: Sequence[Dict]) -> np.ndarray:
        return np.array([feature_row(r["layout"], conf_dict(r["conf"])) for r in records], dtype=np.float64)

    def fit(self, records: Sequence[Dict], steps: int = 3000, lr: float = 2e-3) -> np.ndarray:
        """Full-batch Adam on the device for every member; returns the losses
        (members x steps) before each step."""
        import torch

        rows = self.rows_of(records)
        self.x_mean = rows.mean(axis=0)
        self.x_std = np.where(rows.std(axis=0) > 0, rows.std(axis=0), 1.0)
        y = np.log(np.array([r["ms"] / additions(r["layout"]) for r in records], dtype=np.float64))
        self.y_mean, self.y_std = float(y.mean()), float(y.std() or 1.0)
        t = ((y - self.y_mean) / self.y_std).astype(np.float32)
        dx = torch.from_numpy(self._x(rows)).cuda()
        dt = torch.from_numpy(t).cuda()
        stream = torch.cuda.current_stream().cuda_stream
        return np.stack([n.train(dx, dt, len(records), steps, lr, stream) for n in self.nets])

    def predict_rows(self, rows: np.ndarray, layouts: Sequence[Dict]) -> np.ndarray:
        """Predicted run time (ms) of feature rows of the given layouts."""
        import torch

        if len(rows) == 0:
            return np.zeros(0)
        dx = torch.from_numpy(self._x(np.asarray(rows, dtype=np.float64))).cuda()
        dy = torch.empty(len(rows), dtype=torch.float32, device="cuda")
        stream = torch.cuda.current_stream().cuda_stream
        acc = np.zeros(len(rows))
        for n in self.nets:
            n.forward(dx, len(rows), dy, stream)
            acc += dy.cpu().numpy().astype(np.float64)
        y = acc / len(self.nets)
        work = np.array([additions(l) for l in layouts])
        return np.exp(y * self.y_std + self.y_mean) * work

    def predict(self, layout: Dict, confs: Sequence) -> np.ndarray:
        rows = np.array([feature_row(layout, conf_dict(c)) for c in confs])
        return self.predict_rows(rows, [layout] * len(rows))


def pick(model: Optional[CostModel], layout: Dict, confs: Sequence, time_fn: Callable[[Dict], Optional[float]],
         budget: float = 0.1) -> Dict:
    """Times the best-predicted ceil(budget * len(confs)) configurations with
    time_fn (returns ms, or None for a configuration that cannot run) and
    returns the fastest.  No model or budget >= 1: the exhaustive sweep."""
    confs = [conf_dict(c) for c in confs]
    if model is None or budget >= 1.0:
        order = list(range(len(confs)))
        k = len(confs)
    else:
        order = list(np.argsort(model.predict(layout, confs), kind="stable"))
        k = max(1, math.ceil(budget * len(confs)))
    timed, best, best_ms, spent = 0, None, math.inf, 0.0
    for i in order:
        if timed >= k:
            break
        ms = time_fn(confs[i])
        if ms is None:
            continue  # not runnable: does not count against the budget
        timed += 1
        spent += ms
        if ms < best_ms:
            best, best_ms = confs[i], ms
    return {"conf": best, "ms": best_ms, "timed": timed, "candidates": len(confs), "timed_ms": spent}


def evaluate_offline(model: Optional[CostModel], records: Sequence[Dict], budget: float = 0.1) -> Dict:
    """Runs ``pick`` on one layout's exhaustive records, using the recorded
    timings as the measurements; reports the pick against the optimum and
    the fraction of the sweep's timing cost it spent."""
    table = {tuple(r["conf"]): r["ms"] for r in records}
    layout = records[0]["layout"]
    confs = [list(c) for c in table]
    res = pick(model, layout, confs, lambda c: table[tuple(c[k] for k in CONF_KEYS)], budget)
    opt = min(table.values())
    res.update({"optimum_ms": opt, "ratio": res["ms"] / opt, "time_fraction": res["timed_ms"] / sum(table.values()),
                "count_fraction": res["timed"] / len(table)})
    return res


def random_pick_ratio(records: Sequence[Dict], budget: float = 0.1, trials: int = 2000, seed: int = 0) -> float:
    """Mean ratio to the optimum of the best of ceil(budget * n) configurations
    drawn at random: the same timing budget without the model."""
    ms = np.array([r["ms"] for r in records])
    k = max(1, math.ceil(budget * len(ms)))
    rng = np.random.default_rng(seed)
    best = [ms[rng.choice(len(ms), k, replace=False)].min() for _ in range(trials)]
    return float(np.mean(best) / ms.min())
