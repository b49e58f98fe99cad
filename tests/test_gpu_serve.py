"""GPU: the serving loop (continuous batching over HIP streams, HBM-budget
admission) returns bit-exact results, and each request's result is the same
alone as in the mix."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(n, seed, arrivals=None):
    from compression_amd import serve
    sizes = serve.request_sizes(n, seed, 16 << 10, 2 << 20)
    rng = np.random.default_rng(seed)
    codecs = [serve.CODECS[i] for i in rng.integers(0, 4, n)]
    arr = arrivals if arrivals is not None else np.zeros(n)
    return [serve.make_request(i, codecs[i], int(sizes[i]), 1000 + i, float(arr[i])) for i in range(n)]


def test_mix_bit_exact_and_same_alone(gpu):
    from compression_amd import serve
    reqs = _requests(24, 7, serve.poisson_arrivals(24, 400.0, 8))
    srv = serve.Server(hbm_budget=256 << 20, n_streams=4, verify=True)
    res = srv.run(reqs, time_limit=120)
    assert res["completed"] == 24 and res["bit_exact"]
    mixed = dict(srv.outputs)
    for r in reqs:
        alone = serve.Server(hbm_budget=1 << 30, n_streams=1, verify=True)
        r2 = serve.make_request(r.rid, r.codec, r.out_bytes, 1000 + r.rid, 0.0)
        out = alone.run([r2], time_limit=60)
        assert out["bit_exact"] and np.array_equal(alone.outputs[r.rid], mixed[r.rid])


def test_budget_admission_limits_inflight(gpu):
    from compression_amd import serve
    reqs = _requests(12, 9)
    budget = max(r.hbm_bytes() for r in reqs)         # about one request at a time
    srv = serve.Server(hbm_budget=budget, n_streams=4, verify=True)
    res = srv.run(reqs, time_limit=120)
    assert res["completed"] == 12 and res["bit_exact"]
    assert srv.in_use == 0


def test_full_request_range_and_largest_alone(gpu):
    """Config 5's whole size range (16 KiB..64 MiB, log-uniform) under its
    64 GB budget, plus one 64 MiB request of every codec: every result is
    bit-exact, and the 64 MiB ones equal their decode alone."""
    from compression_amd import serve
    n = 12
    sizes = serve.request_sizes(n, 0x5E, 16 << 10, 64 << 20)
    rng = np.random.default_rng(0x5E)
    codecs = [serve.CODECS[i] for i in rng.integers(0, 4, n)]
    arr = serve.poisson_arrivals(n + 4, 200.0, 0x5F)
    reqs = [serve.make_request(i, codecs[i], int(sizes[i]), 2000 + i, float(arr[i])) for i in range(n)]
    big = [serve.make_request(n + k, c, 64 << 20, 3000 + k, float(arr[n + k]))
           for k, c in enumerate(serve.CODECS)]
    srv = serve.Server(hbm_budget=64 << 30, n_streams=4, verify=True)
    res = srv.run(reqs + big, time_limit=300)
    assert res["completed"] == n + 4 and res["bit_exact"]
    assert srv.in_use == 0
    for r in big:
        alone = serve.Server(hbm_budget=64 << 30, n_streams=1, verify=True)
        out = alone.run([serve.make_request(r.rid, r.codec, r.out_bytes, 3000 + r.rid - n, 0.0)],
                        time_limit=120)
        assert out["bit_exact"] and np.array_equal(alone.outputs[r.rid], srv.outputs[r.rid])


def test_live_queue_expiry_republishes_and_matches_launch_mode(gpu):
    """Deflate requests arriving farther apart than the live queue's idle
    limit: expired queues are replaced and their requests published again;
    every output equals the per-tick launch mode's."""
    from compression_amd import serve
    n = 10
    sizes = serve.request_sizes(n, 0x61, 16 << 10, 1 << 20)
    codecs = ["deflate" if i % 2 == 0 else serve.CODECS[i % 3] for i in range(n)]
    arr = np.arange(n) * 0.08
    mk = lambda: [serve.make_request(i, codecs[i], int(sizes[i]), 4000 + i, float(arr[i])) for i in range(n)]
    live = serve.Server(hbm_budget=1 << 30, n_streams=2, verify=True, live_idle_ms=20)
    res = live.run(mk(), time_limit=120)
    assert res["completed"] == n and res["bit_exact"]
    launch = serve.Server(hbm_budget=1 << 30, n_streams=2, verify=True, deflate_mode="launch")
    res2 = launch.run(mk(), time_limit=120)
    assert res2["completed"] == n and res2["bit_exact"]
    for i in range(n):
        assert np.array_equal(live.outputs[i], launch.outputs[i])


def test_native_engine_matches_python_loop(gpu):
    """The native serving engine (mdz_serve_*): every codec over the full
    request range, bit-exact, each output equal to the Python loop's, the
    budget released; a second run reuses its pools."""
    from compression_amd import serve
    n = 12
    sizes = serve.request_sizes(n, 0x71, 16 << 10, 64 << 20)
    codecs = [serve.CODECS[i % 4] for i in range(n)]
    arr = serve.poisson_arrivals(n, 300.0, 0x72)
    mk = lambda: [serve.make_request(i, codecs[i], int(sizes[i]), 5000 + i, float(arr[i])) for i in range(n)]
    nat = serve.Server(hbm_budget=64 << 30, n_streams=2, verify=True, native=True)
    reqs = mk()
    res = nat.run(reqs, time_limit=120)
    assert res["completed"] == n and res["bit_exact"]
    for r in reqs:
        r.arrival = 0.0
    res2 = nat.run(reqs, time_limit=120)                 # all at once
    assert res2["completed"] == n and res2["bit_exact"]
    py = serve.Server(hbm_budget=64 << 30, n_streams=2, verify=True)
    res3 = py.run(mk(), time_limit=120)
    assert res3["completed"] == n and res3["bit_exact"]
    for i in range(n):
        assert np.array_equal(nat.outputs[i], py.outputs[i])


def test_native_engine_budget_and_damaged_request(gpu):
    """A budget of about one request admits them one by one; a request with
    damaged input reports its failing status and the others stay exact."""
    from compression_amd import serve
    reqs = _requests(10, 13)
    budget = max(r.hbm_bytes() for r in reqs)
    bad = next(r for r in reqs if r.codec == "deflate")
    pay = np.array(bad.payload, copy=True)
    o = int(bad.meta["src_off"][0])
    pay[o + 2:o + 40] ^= 0x5A                               # inside the first zlib stream
    bad.payload = pay
    srv = serve.Server(hbm_budget=budget, n_streams=2, verify=True, native=True)
    res = srv.run(reqs, time_limit=120)
    assert res["completed"] == 10 and res["bit_exact"] is False
    for r in reqs:
        if r is not bad:
            assert np.array_equal(srv.outputs[r.rid], r.raw)
