"""CPU models of the two cross-wave protocols in the HIP decoders, run over
many schedules (no GPU).

1. Inflate's producer/consumer waves (compression_amd/csrc/inflate.hip): one
   workgroup barrier per step; at step s the producer writes item s into
   buffer s % 2 unless it has produced the stream's last item or the consumer
   has halted, and the consumer processes item s - 1 from buffer (s - 1) % 2;
   after the barrier both leave when the consumer has set `done`.  Checked:
   the consumer sees every item in order exactly once up to the first that
   ends the stream (last item, or an error of its own), never reads a buffer
   the producer is writing in the same step, and both waves pass the same
   number of barriers.

2. RLE v1's token walk and unit decoders (compression_amd/csrc/rlev1.hip):
   the walker appends tokens to a lane's list, then publishes the list length
   and then the values the list covers; a decoder waits until the published
   values cover its unit or the walk is done, then reads the list length and
   searches the list.  Checked, over random interleavings of one walker and
   several decoders: a decoder never reads a list entry the walker has not
   written, finds the token covering its unit, and every wait ends."""
import random

import pytest


def run_stream(items, consumer_error_at=None, rng=None):
    """items: list of (kind, last, status) the producer would emit in order."""
    buf = [None, None]
    halt = done = False
    pdone = False
    pi = 0
    seen = []
    barriers = [0, 0]
    writing = None
    for step in range(len(items) + 3):
        # both roles act within the step; their order inside it is arbitrary
        order = [0, 1]
        if rng:
            rng.shuffle(order)
        for role in order:
            if role == 0:                        # producer
                if not pdone and not halt:
                    b = step & 1
                    writing = b
                    buf[b] = (pi, items[pi])
                    if items[pi][1]:
                        pdone = True
                    pi += 1
            else:                                # consumer
                if step > 0 and not done:
                    b = (step - 1) & 1
                    assert writing != b or role_wrote_other(step, b)
                    idx, (kind, last, status) = buf[b]
                    assert idx == step - 1       # the item produced last step
                    seen.append(idx)
                    err = consumer_error_at is not None and idx == consumer_error_at
                    if err or status or last:
                        halt = done = True
        barriers[0] += 1
        barriers[1] += 1
        writing = None
        if done:
            break
    return seen, barriers, pi


def role_wrote_other(step, b):
    return (step & 1) != b


@pytest.mark.parametrize("seed", range(200))
def test_inflate_step_protocol(seed):
    rng = random.Random(seed)
    n = rng.randint(1, 12)
    items = [("round", False, 0) for _ in range(n)]
    end = rng.randrange(n)
    kind = rng.choice(["last", "error"])
    # an item that carries an error status is also the stream's last
    items[end] = ("round", True, -5 if kind == "error" else 0)
    cerr = rng.choice([None, rng.randrange(end + 1)])
    seen, barriers, produced = run_stream(items[:end + 1], cerr, rng)
    stop = end if cerr is None else min(end, cerr)
    assert seen == list(range(stop + 1))            # in order, once each, up to the end
    assert barriers[0] == barriers[1]               # both waves pass the same barriers
    assert produced <= stop + 2                     # at most one item decoded past the end


def walker(lst, pub, tokens, covers):
    """Appends tokens; after each, publishes the length, then the values covered."""
    v = 0
    for i, (first, count) in enumerate(tokens):
        lst.append((first, count))
        yield
        pub["ntok"] = i + 1
        yield
        v = first + count
        pub["vpub"] = v
        yield
    pub["done"] = True


def decoder(lst, pub, unit_v, unit_n, out, spin_max=10_000):
    it = 0
    while True:
        if pub["done"]:
            break
        if pub["vpub"] >= unit_v + unit_n:
            break
        it += 1
        assert it < spin_max, "wait did not end"
        yield
    nt = pub["ntok"]
    assert nt <= len(lst)                            # only written entries are read
    lo, hi = 0, nt - 1
    while lo < hi:                                   # the last token with first value <= unit_v
        mid = (lo + hi + 1) // 2
        if lst[mid][0] <= unit_v:
            lo = mid
        else:
            hi = mid - 1
        yield
    out.append((unit_v, lst[lo]))


@pytest.mark.parametrize("seed", range(100))
def test_rlev1_walk_publication(seed):
    rng = random.Random(seed)
    tokens, v = [], 0
    while v < 1024:
        c = rng.randint(1, 130)
        tokens.append((v, c))
        v += c
    lst, pub, out = [], {"ntok": 0, "vpub": 0, "done": False}, []
    gens = [walker(lst, pub, tokens, v)]
    units = [(32 * k, 32) for k in range(32)]
    gens += [decoder(lst, pub, u, n, out) for u, n in units]
    live = list(gens)
    while live:                                      # random interleaving of the waves
        g = rng.choice(live)
        try:
            next(g)
        except StopIteration:
            live.remove(g)
    assert len(out) == len(units)
    for unit_v, (first, count) in out:
        assert first <= unit_v < first + count        # the token that covers the unit
