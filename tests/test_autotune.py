"""Launch-shape autotuner: bucketing, planner filtering and the JSON cache
(host only: the timer is replaced by a deterministic cost model)."""
import json

import numpy as np
import pytest
import torch

from quantom_loits_amd import autotune
from quantom_loits_amd.batched import Tuning


def test_bucket_keys_round_up_to_powers_of_two():
    assert autotune.bucket_key(np.full(256, 2048), np.full(256, 1 << 18), torch.float32) == \
        "f32/N2048/E262144/B256"
    assert autotune.bucket_key([100, 3000, 17], [5, 7000, 1], torch.float64) == \
        "f64/N4096/E4096/B4"
    assert autotune.bucket_shape("f32/N32768/E1024/B8") == (8, 32767, 1024)


def _cost(calls):
    def timer(n_nodes, events, dtype, t):
        calls.append((int(n_nodes.max()), t.threads, t.chunk_events, t.guide_shift))
        # fastest at threads 512, chunk 2^17, guide_shift 3
        return (abs(t.threads - 512) / 256 + abs(np.log2(t.chunk_events) - 17)
                + (3 - t.guide_shift) * 0.1 + 1.0)
    return timer


def test_tuner_picks_fastest_and_caches(tmp_path):
    calls = []
    path = tmp_path / "tune.json"
    tuner = autotune.Autotuner(path=path, timer=_cost(calls))
    t = tuner.tuning_for(np.full(16, 1000), np.full(16, 5000), torch.float32)
    assert (t.threads, t.chunk_events, t.guide_shift) == (512, 1 << 17, 3)
    assert len(calls) == len(autotune.candidates())  # every shape plans at N=1024
    entry = json.loads(path.read_text())["f32/N1024/E8192/B16"]
    assert entry["threads"] == 512 and entry["candidates"] == len(calls)

    again = autotune.Autotuner(path=path, timer=_cost(calls))
    n_calls = len(calls)
    t2 = again.tuning_for(np.full(9, 700), np.full(9, 4100), torch.float32, scan_mode=1)
    assert len(calls) == n_calls  # same bucket: served from the cache
    assert (t2.threads, t2.chunk_events, t2.guide_shift, t2.scan_mode) == (512, 1 << 17, 3, 1)


def test_tuner_skips_shapes_the_planner_rejects(tmp_path):
    calls = []
    tuner = autotune.Autotuner(path=tmp_path / "t.json", timer=_cost(calls))
    t = tuner.tuning_for([20000, 30000], [100, 100], torch.float32)
    # 32767 nodes need at least 1024 threads (32 nodes per thread at most)
    assert calls and all(c[1] == 1024 for c in calls)
    assert t.threads == 1024 and t.chunk_events == 1 << 17


def test_unplannable_bucket_raises(tmp_path):
    tuner = autotune.Autotuner(path=tmp_path / "t.json", timer=_cost([]),
                               candidates_fn=lambda scan: [Tuning(threads=256)])
    with pytest.raises(ValueError, match="no launch shape"):
        tuner.tuning_for([30000], [10], torch.float64)


def test_ragged_batches_are_tuned_per_grid_size_bucket(tmp_path):
    """Curves in several grid-size buckets: per-kind or persistent launches, then
    each bucket's chunk size (and guide shift where its curves take guide tables)
    by coordinate descent on the whole batch; the result is cached."""
    calls = []
    # cost model of the whole batch: persistent saves 10%; bucket 7 is fastest at
    # 2^16, bucket 14 at 2^17 with guide shift 1, the others at the default 2^18
    want = {7: (1 << 16, 2), 14: (1 << 17, 1)}

    def timer(n_nodes, events, dtype, t, window=None):
        calls.append((dict(t.buckets), t.persistent, window))
        ms = 10.0 - t.persistent
        for b in (7, 10, 12, 14):
            ce, gs = t.buckets.get(b, (1 << 18, 2))
            goal = want.get(b, (1 << 18, 2))
            ms += abs(np.log2(ce) - np.log2(goal[0])) * 0.5 + abs(gs - goal[1]) * 0.2
        return ms

    nodes = np.array([100, 120, 900, 1000, 3000, 9000, 12000])
    events = np.array([5000, 7000, 3000, 9000, 40000, 20000, 1000])
    tuner = autotune.Autotuner(path=tmp_path / "t.json", timer=timer)
    t = tuner.tuning_for(nodes, events, torch.float32, window=(0.05, 0.95))
    # buckets: 7 (100, 120), 10 (900, 1000), 12 (3000), 14 (9000, 12000)
    assert t.persistent == 1 and t.threads == 1024
    assert t.buckets == {7: (1 << 16, 2), 14: (1 << 17, 1)}
    n_guide = len(autotune.BUCKET_CHUNKS) * len(autotune.GUIDE_SHIFTS) - 1
    n_cell = len(autotune.BUCKET_CHUNKS) - 1
    assert len(calls) == 2 + 3 * n_cell + n_guide  # the defaults are not timed twice
    assert all(c[2] == (0.05, 0.95) for c in calls)
    cache = json.loads((tmp_path / "t.json").read_text())
    (entry,) = cache.values()
    assert entry["persistent"] == 1 and entry["buckets"] == {"7": [65536, 2], "14": [131072, 1]}
    n_calls = len(calls)
    again = autotune.Autotuner(path=tmp_path / "t.json", timer=timer)
    t2 = again.tuning_for(nodes, events, torch.float32, window=(0.05, 0.95))
    assert len(calls) == n_calls and t2.buckets == t.buckets and t2.persistent == 1
    ce, shift = t.bucket_arrays()
    assert ce[7] == 16 and ce[14] == 17 and shift[14] == 2 and ce[12] == 0 and shift[12] == 0
