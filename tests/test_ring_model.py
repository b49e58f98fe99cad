"""CPU models of the two cross-wave protocols in the HIP decoders, run over
many schedules (no GPU).

1. Inflate's producer/consumer waves (compression_amd/csrc/inflate.hip): an
   item queue of KSLOTS slots in LDS with two counts.  The producer writes
   item k into slot k % KSLOTS once the consumer has finished with item
   k - KSLOTS, then publishes n_pub = k + 1; the consumer waits for
   n_pub > k, processes item k, and publishes n_fin = k + 1.  The consumer
   sets `halt` when it stops (the stream's last item, or an error); the
   producer checks `halt` before each item and in its waits, and a wave whose
   wait runs out sets `halt` itself.  A round's symbols go to a ring of RING
   entries shared by the items in flight: the producer lists a round of n
   symbols once the symbols listed minus those of finished items (sym_fin,
   published by the consumer) leave room for n.  Checked over random interleavings: the
   consumer sees every item in order exactly once up to the one that ends the
   stream, no slot is written while the consumer reads it, no ring entry is
   reused before the consumer has finished with its item, the producer runs
   at most KSLOTS items past the consumer's stop, every wait ends, and a
   producer that gives up makes the consumer stop with an internal error.

2. RLE v1's token walk and unit decoders (compression_amd/csrc/rlev1.hip):
   the walker appends tokens to a lane's list, then publishes the list length
   and then the values the list covers; a decoder waits until the published
   values cover its unit or the walk is done, then reads the list length and
   searches the list.  Checked, over random interleavings of one walker and
   several decoders: a decoder never reads a list entry the walker has not
   written, finds the token covering its unit, and every wait ends."""
import random

import pytest


KSLOTS = 3
RING = 6144                      # symbol ring entries; a round lists at most CAP
CAP = 4096


def run_stream(items, consumer_error_at=None, rng=None, producer_gives_up_at=None):
    """items: list of (kind, last, status) the producer would emit in order.
    Returns (seen, produced, consumer_status)."""
    st = {"n_pub": 0, "n_fin": 0, "sym_fin": 0, "halt": False}
    slot = [None] * KSLOTS
    ring_owner = [None] * RING       # item whose symbol occupies each ring entry
    ring_reading = set()             # items the consumer is reading
    writing = [False] * KSLOTS
    reading = [False] * KSLOTS
    seen = []
    out = {"produced": 0, "cstatus": 0}
    sym_pos = [0, 0]                 # producer: symbols listed, ring end
    sym_done = [0]                   # consumer: symbols of finished items
    finished = {}

    def producer():
        ik = 0
        while True:
            if st["halt"]:
                return
            want = ik + 1 - KSLOTS if ik >= KSLOTS else 0
            polls = 0
            while st["n_fin"] < want:
                if st["halt"]:
                    return
                if producer_gives_up_at is not None and polls >= producer_gives_up_at:
                    st["halt"] = True            # the wait ran out
                    return
                polls += 1
                yield
            b = ik % KSLOTS
            assert not reading[b]
            writing[b] = True
            yield
            # the round's symbols go to the ring once the consumer freed room
            nsym = items[ik][3] if len(items[ik]) > 3 else 0
            polls = 0
            while sym_pos[0] + nsym - st["sym_fin"] > RING:
                if st["halt"]:
                    writing[b] = False
                    return
                polls += 1
                yield
            for j in range(nsym):
                e = (sym_pos[1] + j) % RING
                # an entry is reused only after the consumer finished its item
                assert ring_owner[e] is None or finished.get(ring_owner[e], False)
                ring_owner[e] = ik
            sym0 = sym_pos[1]
            sym_pos[0] += nsym
            sym_pos[1] = (sym_pos[1] + nsym) % RING
            assert not reading[b]
            slot[b] = (ik, items[ik], sym0, nsym)
            writing[b] = False
            out["produced"] = ik + 1
            yield
            st["n_pub"] = ik + 1
            yield
            if items[ik][1]:
                return
            ik += 1

    def consumer():
        ik = 0
        while True:
            while st["n_pub"] < ik + 1:
                if st["halt"]:
                    out["cstatus"] = "internal"
                    return
                yield
            b = ik % KSLOTS
            assert not writing[b]
            reading[b] = True
            idx, item, sym0, nsym = slot[b]
            kind, last, status = item[:3]
            assert idx == ik
            ring_reading.add(ik)
            assert all(ring_owner[(sym0 + j) % RING] == ik for j in range(nsym))
            yield
            assert not writing[b]
            assert all(ring_owner[(sym0 + j) % RING] == ik for j in range(nsym))
            ring_reading.discard(ik)
            finished[ik] = True
            reading[b] = False
            seen.append(idx)
            err = consumer_error_at is not None and idx == consumer_error_at
            if err or status or last:
                out["cstatus"] = status or ("error" if err else 0)
                st["halt"] = True
                return
            sym_done[0] += nsym
            st["sym_fin"] = sym_done[0]
            st["n_fin"] = ik + 1
            yield
            ik += 1

    live = [producer(), consumer()]
    for n in range(100_000):
        if not live:
            break
        g = rng.choice(live) if rng else live[n % len(live)]
        try:
            next(g)
        except StopIteration:
            live.remove(g)
    assert not live, "a wait never ended"
    return seen, out["produced"], out["cstatus"]


@pytest.mark.parametrize("seed", range(300))
def test_inflate_queue_protocol(seed):
    rng = random.Random(seed)
    n = rng.randint(1, 12)
    items = [("round", False, 0, rng.choice([0, 1, 2048, CAP, rng.randint(0, CAP)])) for _ in range(n)]
    end = rng.randrange(n)
    kind = rng.choice(["last", "error"])
    # an item that carries an error status is also the stream's last
    items[end] = ("round", True, -5 if kind == "error" else 0, rng.randint(0, CAP))
    cerr = rng.choice([None, rng.randrange(end + 1)])
    seen, produced, cstatus = run_stream(items[:end + 1], cerr, rng)
    stop = end if cerr is None else min(end, cerr)
    assert seen == list(range(stop + 1))            # in order, once each, up to the end
    assert produced <= stop + 1 + KSLOTS            # bounded run-ahead past the stop
    assert cstatus != "internal"


@pytest.mark.parametrize("seed", range(100))
def test_inflate_queue_producer_gives_up(seed):
    rng = random.Random(1000 + seed)
    n = rng.randint(KSLOTS + 1, 12)
    items = [("round", False, 0, rng.randint(0, CAP)) for _ in range(n)]
    items[-1] = ("round", True, 0, rng.randint(0, CAP))
    seen, produced, cstatus = run_stream(items, None, rng, producer_gives_up_at=rng.randint(0, 3))
    assert seen == list(range(len(seen)))
    # the consumer ends: with the last item, or with an internal error
    assert cstatus == "internal" or seen[-1] == n - 1


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
