"""Runtime utilities on the GPU: the device checksum (ihe_checksum) against its
closed form computed on the host, the HBM streaming probe (ihe_stream_probe):
timed, bounded to the bytes it is given, argument errors; and launches on a
caller's stream (ihe_ctx_set_stream / ihe_ctx_get_stream)."""
import ctypes

import numpy as np
import pytest

from idash2019_2_amd import _abi
from idash2019_2_amd._abi import IheError

pytestmark = pytest.mark.gpu


def host_checksum(words: np.ndarray) -> int:
    """Sum over words i of mix((i << 32) ^ w_i) mod 2^64, mix being splitmix64's
    output mix (shifts 30/27/31, multipliers 0xBF58476D1CE4E5B9 and
    0x94D049BB133111EB)."""
    x = (np.arange(words.size, dtype=np.uint64) << np.uint64(32)) ^ words.astype(np.uint64)
    with np.errstate(over="ignore"):
        x ^= x >> np.uint64(30)
        x *= np.uint64(0xBF58476D1CE4E5B9)
        x ^= x >> np.uint64(27)
        x *= np.uint64(0x94D049BB133111EB)
        x ^= x >> np.uint64(31)
        return int(x.sum(dtype=np.uint64))


def test_checksum_matches_closed_form_and_sees_order(engine):
    rng = np.random.default_rng(5)
    w = rng.integers(0, 1 << 32, 1_000_003, dtype=np.uint64).astype(np.uint32)
    d = engine.to_device(w)
    assert engine.checksum(d) == host_checksum(w)
    assert engine.checksum(d, 4 * 1000) == host_checksum(w[:1000])
    swapped = w.copy()
    swapped[[10, 11]] = swapped[[11, 10]]
    assert w[10] != w[11]
    assert engine.checksum(engine.to_device(swapped)) != host_checksum(w)
    assert engine.checksum(d, 0) == 0
    with pytest.raises(IheError) as e:
        engine.checksum(d, 6)
    assert e.value.code == _abi.IHE_ERR_ARG


def test_stream_probe_is_timed_and_stays_in_bounds(engine):
    in_b, out_b = 3 << 20, (5 << 20) + 65536
    d_in = engine.to_device(np.arange(in_b // 4, dtype=np.uint32))
    guard = 1 << 16
    d_out = engine.to_device(np.full(out_b + guard, 0x5A, np.uint8))
    engine.timing(True)
    engine.timing_reset()
    engine.stream_probe(d_in, in_b, d_out, out_b)
    ms, n = engine.timing_read(_abi.TIMER_STREAM)
    engine.timing(False)
    assert n == 1 and ms > 0
    out = d_out.numpy()
    assert (out[out_b:] == 0x5A).all()              # nothing written past out_bytes
    assert (out[:out_b] != 0x5A).any()
    for args in ((in_b + 1, out_b), (in_b, 1000), (in_b, out_b + 512), (in_b, 1 << 31)):
        with pytest.raises((IheError, ValueError)):
            engine.stream_probe(d_in, args[0], d_out, args[1])


@pytest.mark.parametrize("policies", [(0, 0), (0, 1), (0, 2), (0, 16), (0, 18), (2, 2), (2, 18)])
def test_stream_probe_patterns_stay_in_bounds(engine, policies):
    """ihe_stream_probe_ex: reads alone (1 KiB written per wave), writes alone
    and a mix, under each cache policy pair; nothing lands past out_bytes."""
    in_b, guard = 4 << 20, 1 << 16
    d_in = engine.to_device(np.arange(in_b // 4, dtype=np.uint32))
    for ib, ob, waves in ((in_b, 1024 * 300, 300), (0, 3 << 20, 700), (in_b, 5 << 20, 80)):
        d_out = engine.to_device(np.full(ob + guard, 0x5A, np.uint8))
        engine.stream_probe_ex(d_in, ib, d_out, ob, waves, *policies)
        out = d_out.numpy()
        assert (out[ob:] == 0x5A).all()
        assert (out[:ob] != 0x5A).any()
        d_out.free()
    d_out = engine.to_device(np.zeros(4096, np.uint8))
    for bad in ((in_b, 4096, 5, 0, 18), (in_b, 4096, 4, 1, 18), (in_b, 4096, 0, 0, 18)):
        with pytest.raises(IheError) as e:
            engine.stream_probe_ex(d_in, bad[0], d_out, bad[1], bad[2], bad[3], bad[4])
        assert e.value.code == _abi.IHE_ERR_ARG


def test_user_stream_switch(engine):
    """ihe_ctx_set_stream: launches go to a caller's stream (a decrypt and a
    cloud step, equal to the oracle), ihe_ctx_get_stream reports it, and NULL
    returns the context to its own stream (the switch waits for the old one)."""
    import oracle
    from idash2019_2_amd import synth
    hip = ctypes.CDLL("libamdhip64.so.7")         # the runtime the library links
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    s = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(s)) == 0
    L = _abi.lib()
    try:
        _abi.check(L.ihe_ctx_set_stream(engine.ctx, s))
        got = ctypes.c_void_p()
        _abi.check(L.ihe_ctx_get_stream(engine.ctx, ctypes.byref(got)))
        assert got.value == s.value
        c = synth.generate(30, 60, 128, 5, seed=8)
        rng = np.random.default_rng(8)
        ct = rng.integers(-(1 << 31), 1 << 31, (c.num_in_cts, 2, 1024)).astype(np.int32)
        var = np.full(c.num_in_cts, 2.0 ** -50)
        rp, ib, w = c.model_csr()
        for phase in ("user stream", "own stream"):
            band = engine.band(128, c.num_in_cts, rp, ib, w)
            d_in, d_var = engine.to_device(ct), engine.to_device(var)
            d_out = engine.empty((c.num_outputs, 2, 1024), np.int32)
            d_ov = engine.empty((c.num_outputs,), np.float64)
            engine.cloud(band, d_in, d_var, d_out, d_ov)
            key = engine.key(c.key)
            d_ph = engine.empty((c.num_outputs, 128), np.float32)
            engine.decrypt(key, 128, d_out, c.num_outputs, d_ph)
            engine.sync()
            ref, rv = oracle.cloud(ct, var, rp, ib, w, *c.layout)
            assert (d_out.numpy() == ref).all() and (d_ov.numpy() == rv).all(), phase
            assert (d_ph.numpy() == oracle.decrypt(ref, c.key, 128)).all(), phase
            key.free()
            band.free()
            _abi.check(L.ihe_ctx_set_stream(engine.ctx, None))
        _abi.check(L.ihe_ctx_get_stream(engine.ctx, ctypes.byref(got)))
        assert got.value != s.value
    finally:
        L.ihe_ctx_set_stream(engine.ctx, None)
        hip.hipStreamDestroy(s)
