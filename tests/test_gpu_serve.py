"""Serving loop (ihe_server_*): encrypt -> cloud -> decrypt requests on
resident model bands, admitted by HBM budget. Each request's scores are the
same alone (one slot, one request at a time) and in the mix (eight slots, all
requests in flight), equal the per-request device path chunk by chunk, and
match the oracle on a small request."""
import numpy as np
import pytest

import oracle
from idash2019_2_amd import Server, _abi, synth
from idash2019_2_amd._abi import IheError

pytestmark = pytest.mark.gpu

GIB = 1 << 30


@pytest.fixture(scope="module")
def models(engine):
    out = {}
    for n in (5, 20):
        c = synth.generate(50, 120, 1024, n, seed=30 + n)
        band = engine.band(1024, c.num_in_cts, *c.model_csr())
        out[n] = (c, band)
    yield out
    for _, band in out.values():
        band.free()


def requests(models, sample_counts, seed=0):
    rng = np.random.default_rng(seed)
    reqs = []
    for i, ns in enumerate(sample_counts):
        n = (5, 20)[i % 2]
        c, _ = models[n]
        geno = rng.integers(-1, 3, (c.num_tags, ns)).astype(np.int8)
        key = rng.integers(0, 2, 1024).astype(np.int32)
        reqs.append((n, geno, key, 1000 + 17 * i))
    return reqs


def per_request_path(engine, models, n, geno, key, seed):
    """The plain device path chunk by chunk: encrypt (seed + k), cloud, decrypt."""
    c, band = models[n]
    ns_total = geno.shape[1]
    out = np.zeros((c.num_outputs, ns_total), np.float32)
    k_dev = engine.key(key)
    for k, col0 in enumerate(range(0, ns_total, 1024)):
        ns = min(1024, ns_total - col0)
        R = 1024 // ns
        num_in = (3 * c.num_tags + R - 1) // R
        d_geno = engine.to_device(np.ascontiguousarray(geno[:, col0:col0 + ns]))
        d_ct = engine.empty((num_in, 2, 1024), np.int32)
        d_var = engine.empty((num_in,), np.float64)
        engine.encrypt(k_dev, ns, d_geno, seed + k, d_ct, d_var)
        d_out = engine.empty((c.num_outputs, 2, 1024), np.int32)
        engine.cloud_layout(band, ns, num_in, d_ct, None, d_out, None)
        d_dec = engine.empty((c.num_outputs, ns), np.float32)
        engine.decrypt(k_dev, ns, d_out, c.num_outputs, d_dec)
        out[:, col0:col0 + ns] = d_dec.numpy()
        for d in (d_geno, d_ct, d_var, d_out, d_dec):
            d.free()
    k_dev.free()
    return out


def serve(engine, models, reqs, slots, budget, one_at_a_time):
    srv = engine.server(budget, num_slots=slots)
    ids = {n: srv.add_model(models[n][1], models[n][0].num_tags) for n in models}
    outs, lat = [], []
    tickets = []
    for n, geno, key, seed in reqs:
        rows, ns = models[n][0].num_outputs, geno.shape[1]
        out = np.full(rows * ns, np.nan, np.float32)
        outs.append((out, rows, ns))
        t = srv.submit(ids[n], Server.to_chunks(geno), key, seed, out, num_samples=ns)
        if one_at_a_time:
            lat.append(srv.wait(t))
        else:
            tickets.append(t)
    for t in tickets:
        lat.append(srv.wait(t))
    st = srv.stats()
    srv.close()
    return [Server.from_chunks(o, rows, ns) for o, rows, ns in outs], lat, st


def test_results_same_alone_in_mix_and_per_request(engine, models):
    counts = [1, 16, 100, 1024, 1500, 3000, 64, 333, 1024, 7]
    reqs = requests(models, counts)
    alone, lat1, _ = serve(engine, models, reqs, 1, 4 * GIB, True)
    mixed, lat8, st = serve(engine, models, reqs, 8, 4 * GIB, False)
    assert st["completed"] == len(reqs) and st["arena_peak"] <= st["arena_bytes"]
    assert all(x > 0 for x in lat1 + lat8)
    for i, (n, geno, key, seed) in enumerate(reqs):
        assert np.array_equal(alone[i], mixed[i]), i
        if i in (0, 2, 4, 7):
            want = per_request_path(engine, models, n, geno, key, seed)
            assert np.array_equal(alone[i], want), i
    # the oracle on one request: encrypt, cloud, decrypt (NS = 100, R = 10)
    n, geno, key, seed = reqs[2]
    c, _ = models[n]
    ns, R = 100, 10
    cts, var = oracle.encrypt(geno, key, ns, R, 1024 // R, seed=seed)
    rp, ib, w = c.model_csr()
    ref, _ = oracle.cloud(cts, var, rp, ib, w, ns, R, 1024 // R)
    assert np.array_equal(alone[2], oracle.decrypt(ref, key, ns))


def test_admission_waits_for_memory(engine, models):
    reqs = requests(models, [1024, 1024, 2048, 512, 1024], seed=3)
    # room for one full 1024-sample chunk at a time
    one = max(c.num_outputs * (8192 + 4 * 1024) + c.num_tags * 1024 + c.num_in_cts * (8192 + 8)
              for c, _ in models.values()) + 5 * 256
    outs, _, st = serve(engine, models, reqs, 8, one, False)
    assert st["arena_peak"] <= one
    alone, _, _ = serve(engine, models, reqs, 1, 4 * GIB, True)
    for a, b in zip(outs, alone):
        assert np.array_equal(a, b)


def test_oversized_request_fails_cleanly(engine, models):
    c, band = models[5]
    srv = engine.server(1 << 20, num_slots=2)       # 1 MiB: no chunk fits
    m = srv.add_model(band, c.num_tags)
    geno = np.zeros((c.num_tags, 8), np.int8)
    out = np.zeros((c.num_outputs, 8), np.float32)
    t = srv.submit(m, geno, np.zeros(1024, np.int32), 1, out)
    with pytest.raises(IheError) as e:
        srv.wait(t)
    assert e.value.code == _abi.IHE_ERR_NOMEM
    with pytest.raises(IheError) as e:
        srv.submit(m, geno, np.full(1024, 2, np.int32), 1, out)   # key not binary
    assert e.value.code == _abi.IHE_ERR_ARG
    srv.close()


def test_latency_counts_from_the_scheduled_arrival(engine, models):
    """A request whose arrival lies 80 ms before its submit reports a latency of
    at least 80 ms; without one, latency counts from the submit."""
    import time
    c, band = models[5]
    srv = engine.server(4 * GIB, num_slots=2)
    m = srv.add_model(band, c.num_tags)
    geno = np.zeros((c.num_tags, 16), np.int8)
    out = np.zeros((c.num_outputs, 16), np.float32)
    key = np.zeros(1024, np.int32)
    t = srv.submit(m, geno, key, 1, out, arrival_ns=time.monotonic_ns() - 80_000_000)
    assert srv.wait(t) >= 80.0
    t = srv.submit(m, geno, key, 1, out)
    assert srv.wait(t) < 80.0
    srv.close()


def test_a_band_outlives_its_server(engine, models):
    c, band = models[20]
    b2 = engine.band(1024, c.num_in_cts, *c.model_csr())
    srv = engine.server(1 * GIB, num_slots=1)
    srv.add_model(b2, c.num_tags)
    with pytest.raises(IheError) as e:
        b2.free()
    assert e.value.code == _abi.IHE_ERR_BUSY
    srv.close()
    b2.free()


def test_two_waiters_on_one_ticket(engine, models):
    """Two threads wait on one ticket: one collects it, the other gets
    IHE_ERR_ARG (not a crash), whichever of them the worker wakes first."""
    import threading
    c, band = models[5]
    srv = engine.server(4 * GIB, num_slots=2)
    m = srv.add_model(band, c.num_tags)
    geno = np.zeros((c.num_tags, 1024), np.int8)
    out = np.zeros((c.num_outputs, 1024), np.float32)
    t = srv.submit(m, geno, np.zeros(1024, np.int32), 1, out)
    got = []

    def waiter():
        try:
            got.append(("ok", srv.wait(t)))
        except IheError as e:
            got.append(("err", e.code))

    th = [threading.Thread(target=waiter) for _ in range(2)]
    for x in th:
        x.start()
    for x in th:
        x.join(timeout=60)
    assert not any(x.is_alive() for x in th)
    assert sorted(k for k, _ in got) == ["err", "ok"]
    assert [v for k, v in got if k == "err"] == [_abi.IHE_ERR_ARG]
    with pytest.raises(IheError) as e:
        srv.poll(t)
    assert e.value.code == _abi.IHE_ERR_ARG
    srv.close()
