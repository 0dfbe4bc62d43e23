"""CPU checks of the bench contract and of the learned filter's host logic."""
import json
import os
import subprocess
import sys

import numpy as np

from mdet_amd import learned as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def _records(n_rec, n, seed):
    dt = np.dtype([("i", np.uint32), ("j", np.uint32), ("pre", np.float32), ("flags", np.uint32)])
    rng = np.random.default_rng(seed)
    keys = rng.choice(n * (n - 1) // 2, n_rec, replace=False)
    # unordered pair index -> (i, j), i < j
    i = np.floor((2 * n - 1 - np.sqrt((2 * n - 1) ** 2 - 8 * keys)) / 2).astype(np.int64)
    j = keys - i * (2 * n - i - 1) // 2 + i + 1
    r = np.zeros(n_rec, dt)
    swap = rng.random(n_rec) < 0.5  # the device lists a pair in either orientation
    r["i"] = np.where(swap, j, i)
    r["j"] = np.where(swap, i, j)
    r["pre"] = rng.random(n_rec)
    return r


def test_training_pairs_do_not_depend_on_record_order():
    n = 3000
    rec = _records(20000, n, 1)
    perm = np.random.default_rng(2).permutation(len(rec))
    for m in (0, 1, 5000, 20000, 30000):  # below, equal to and above the record count
        a = L._sample_pairs(rec, m, n, 7, np.random.default_rng(3))
        b = L._sample_pairs(rec[perm], m, n, 7, np.random.default_rng(3))
        assert len(a) == m and np.array_equal(a, b), m
    # distinct seeds give distinct samples; a sample has no repeated pair when m <= len(rec)
    a = L._sample_pairs(rec, 5000, n, 7, np.random.default_rng(3))
    c = L._sample_pairs(rec, 5000, n, 8, np.random.default_rng(3))
    assert not np.array_equal(a, c)
    key, _ = L._pair_keys(a, n, 7)
    assert len(np.unique(key)) == len(a)
    # the key is symmetric in (i, j)
    flipped = a.copy()
    flipped["i"], flipped["j"] = a["j"], a["i"]
    assert np.array_equal(L._pair_keys(flipped, n, 7)[0], key)
