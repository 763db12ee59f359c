"""CPU tests of the tuner's JSON cache: keys bucket workloads by size and theta, and a
stored result reads back as the same knobs (no device work)."""
from __future__ import annotations

import json

from galaxy_amd.tune import KNOBS, TuneCache


def test_keys_bucket_by_size_theta_and_build():
    k = TuneCache.key([1024, 65536, 3000], 0.5)
    assert k == TuneCache.key([1100, 65000, 3100], 0.5)      # same log2 buckets
    assert k != TuneCache.key([1024, 65536, 3000], 0.6)      # theta
    assert k != TuneCache.key([1024] * 3, 0.5)               # size bucket
    assert k != TuneCache.key([1024, 65536, 3000], 0.5, dict(morton_bits=30, walk=0,
                                                             multipole=0))


def test_store_and_lookup_round_trip(tmp_path):
    path = str(tmp_path / "tune.json")
    result = dict(leaf_size=1, far_theta=0.55, mono_ratio=0.8, morton_bits=30, group_size=64,
                  group_lanes=4, radix_bits=9, table=[{"rate": 1.0}])
    c = TuneCache(path)
    assert c.lookup("absent") is None
    key = TuneCache.key([4096] * 8, 0.5)
    c.store(key, result)
    got = TuneCache(path).lookup(key)
    assert got == {k: result[k] for k in KNOBS}
    with open(path) as f:
        assert set(json.load(f)[key]) == set(KNOBS)  # knobs only, no timing tables


def test_entries_without_a_key_width_read_as_63_bits(tmp_path):
    path = str(tmp_path / "old.json")
    old = dict(leaf_size=1, far_theta=0.55, mono_ratio=0.8, group_size=64, group_lanes=4,
               radix_bits=8)
    with open(path, "w") as f:
        json.dump({"k": old}, f)
    assert TuneCache(path).lookup("k") == dict(old, morton_bits=63)
