"""The batch autotuner's search and cache on the CPU, with a stand-in batch
whose solve time is a known function of the shape: coordinate descent finds
the fastest shape knob by knob, every shape must give the same costs, and a
cached choice is reused without timing."""
import json
import time

import numpy as np
import pytest

import autotune


class FakeBatch:
    """set_params/solve like emv.Batch; a solve sleeps for a time that is a
    function of the shape (best: 32 workgroups, 4 rows, 5 dense rounds, 64
    split items) and returns fixed costs."""

    def __init__(self, costs, wrong_at=None):
        self.costs = np.asarray(costs, dtype=np.int64)
        self.params = {}
        self.solves = 0
        self.wrong_at = wrong_at

    def set_params(self, use_graph=1, **shape):
        self.params = dict(shape, use_graph=use_graph)

    def cost_ms(self):
        p = self.params
        return (abs(p["relax_blocks_per_cu"] - 32) / 32 + abs(p["merge_rows_per_wave"] - 4) / 4
                + abs(p["dense_rounds"] - 5) / 5 + abs(p["split_items"] - 64) / 64)

    def solve(self, flags=0):
        self.solves += 1
        time.sleep(0.002 + 0.02 * self.cost_ms())
        shape = {k: v for k, v in self.params.items() if k != "use_graph"}
        if self.wrong_at is not None and shape == self.wrong_at:
            return self.costs + 1, None
        return self.costs.copy(), None


INSTANCES = [(100, np.zeros((0, 3)), [0, 1, 2]), (200, np.zeros((0, 3)), [3, 4])]


def test_coordinate_descent_finds_each_knob(tmp_path, monkeypatch):
    monkeypatch.setenv("EMV_TUNE_CACHE", str(tmp_path / "tune.json"))
    b = FakeBatch([7, 9])
    shape, table = autotune.tune_batch(b, INSTANCES, num_cu=256, reps=2)
    assert shape == dict(relax_blocks_per_cu=32, merge_rows_per_wave=4, dense_rounds=5,
                         split_items=64)
    # one pass: each knob's candidates once, shapes already timed are reused
    assert len(table) == sum(len(c) for _, c in autotune.KNOBS) - (len(autotune.KNOBS) - 1)
    assert b.params == dict(shape, use_graph=1)
    cache = json.loads((tmp_path / "tune.json").read_text())
    key = autotune.signature(INSTANCES, 256)
    assert cache[key]["shape"] == shape


def test_cached_shape_is_reused(tmp_path, monkeypatch):
    monkeypatch.setenv("EMV_TUNE_CACHE", str(tmp_path / "tune.json"))
    autotune.tune_batch(FakeBatch([7, 9]), INSTANCES, num_cu=256)
    b = FakeBatch([7, 9])
    shape, _ = autotune.tune_batch(b, INSTANCES, num_cu=256)
    assert b.solves == 1  # one eager solve that records the chosen shape's rounds
    assert b.params["relax_blocks_per_cu"] == 32 and b.params["use_graph"] == 1
    # another device size or another batch is another key
    assert autotune.signature(INSTANCES, 256) != autotune.signature(INSTANCES, 304)
    assert autotune.signature(INSTANCES, 256) != autotune.signature(INSTANCES[:1], 256)


def test_shapes_with_different_costs_are_rejected(tmp_path, monkeypatch):
    monkeypatch.setenv("EMV_TUNE_CACHE", str(tmp_path / "tune.json"))
    bad = dict(autotune.DEFAULTS, relax_blocks_per_cu=128)
    with pytest.raises(RuntimeError, match="differ"):
        autotune.tune_batch(FakeBatch([7, 9], wrong_at=bad), INSTANCES, num_cu=256,
                            use_cache=False)
