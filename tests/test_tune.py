"""Autotuner logic on the CPU: the choice database, the search (fastest
accepted candidate; a candidate whose decoded bytes differ is rejected),
size binning and the candidate space.  No GPU calls."""
import json

import numpy as np
import pytest

from compression_amd import tune as T


def test_db_round_trip(tmp_path):
    p = tmp_path / "sub" / "tune.json"
    db = T.TuneDB(str(p))
    assert db.get("k") is None
    db.put("k", {"choice": {"order": "size"}, "ms": 1.5})
    db2 = T.TuneDB(str(p))
    assert db2.get("k") == {"choice": {"order": "size"}, "ms": 1.5}
    assert json.load(open(p))["k"]["ms"] == 1.5


def test_signature_is_canonical():
    a = T.signature(kind="batch", rlev2={"n": 4, "chunk": 8}, brle=None)
    b = T.signature(brle=None, rlev2={"chunk": 8, "n": 4}, kind="batch")
    assert a == b and '"arch":"gfx950"' in a
    assert a != T.signature(kind="batch", rlev2={"n": 5, "chunk": 8}, brle=None)


def test_search_picks_fastest_accepted_and_rejects_changed_output():
    cands = [{"i": 0}, {"i": 1}, {"i": 2}, {"i": 3}]
    digests = {0: "A", 1: "A", 2: "B", 3: "A"}          # candidate 2 decodes other bytes
    times = {0: 3.0, 1: 2.0, 2: 0.5, 3: 2.5}
    ran = []
    best, ms, table = T.search(cands, lambda c: ran.append(c["i"]), lambda c: digests[c["i"]],
                               lambda c: times[c["i"]])
    assert best == {"i": 1} and ms == 2.0
    assert ran == [0, 1, 2, 3]
    assert table[2] == ({"i": 2}, None)                   # rejected, never chosen


def test_search_needs_an_accepted_candidate():
    with pytest.raises(RuntimeError):
        T.search([], lambda c: None, lambda c: "A", lambda c: 1.0)


def test_bin_order():
    rng = np.random.default_rng(3)
    cost = rng.integers(1, 1000, 200).astype(np.float64)
    items = np.arange(200, dtype=np.uint32) + 1000
    o = T.bin_order(items, cost, bins=8)
    assert sorted(o.tolist()) == items.tolist()           # a permutation
    idx = o - 1000
    edges = np.quantile(cost, np.linspace(0, 1, 9)[1:-1])
    b = np.searchsorted(edges, cost[idx], side="right")
    assert (np.diff(b) <= 0).all()                        # largest bin first
    for k in np.unique(b):                                 # index order inside a bin
        assert (np.diff(idx[b == k]) > 0).all()


def test_batch_candidate_space():
    c = list(T.batch_candidates(True, True, 256))
    assert len(c) == 2 * 3 * 3 * 2
    assert {x["window"] for x in c} == {8, 16, 32}
    assert {x["blocks"] for x in c} == {0, 256}
    assert {x["order"] for x in c} == {"size", "bins", "index"}
    assert len(list(T.batch_candidates(True, False, 256))) == 18     # no quarters
    assert {x["window"] for x in T.batch_candidates(False, False, 256)} == {16}
