"""ihe_decrypt_noalloc, the decryption entry point that allocates nothing
(its ring kernel's work counters live in the library's per-GPU device array),
against ihe_decrypt (eval/idash.cpp:711-719 restated on the same ring kernel
with a context's counters): bit-exact on two calls in a row, the second with a
smaller input, and against the oracle on a few ciphertexts."""
import ctypes

import numpy as np
import pytest

import oracle
from idash2019_2_amd import _abi
from idash2019_2_amd._abi import IheError, check, lib

pytestmark = pytest.mark.gpu


def test_decrypt_noalloc_matches_decrypt(engine):
    rng = np.random.default_rng(2026)
    key = rng.integers(0, 2, 1024).astype(np.int32)
    k = engine.key(key)
    stream = ctypes.c_void_p()
    check(lib().ihe_ctx_get_stream(engine.ctx, ctypes.byref(stream)))
    for num_cts, ns in ((5000, 1024), (777, 100)):
        ct = rng.integers(-(1 << 31), 1 << 31, (num_cts, 2, 1024), dtype=np.int64).astype(np.int32)
        d_in = engine.to_device(ct)
        d_ref = engine.empty((num_cts, ns), np.float32)
        d_out = engine.empty((num_cts, ns), np.float32)
        check(lib().ihe_memset(engine.ctx, d_out.ptr, 0xFF, d_out.nbytes))   # NaN everywhere
        engine.decrypt(k, ns, d_in, num_cts, d_ref)
        check(lib().ihe_decrypt_noalloc(key.ctypes.data, ns, d_in.ptr, num_cts, d_out.ptr,
                                        stream))
        engine.sync()
        got, want = d_out.numpy(), d_ref.numpy()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (num_cts, ns)
        rows = [0, num_cts // 2, num_cts - 1]
        assert np.array_equal(got[rows], oracle.decrypt(ct[rows], key, ns))
        for d in (d_in, d_ref, d_out):
            d.free()
    k.free()


def test_decrypt_noalloc_rejects_bad_arguments(engine):
    key = np.zeros(1024, np.int32)
    for args in [(None, 16, 1, 1, 1, None), (key.ctypes.data, 0, 1, 1, 1, None),
                 (key.ctypes.data, 1025, 1, 1, 1, None),
                 (key.ctypes.data, 16, 1, (1 << 26) + 1, 1, None),
                 (key.ctypes.data, 16, None, 4, None, None)]:
        with pytest.raises(IheError) as e:
            check(lib().ihe_decrypt_noalloc(*args))
        assert e.value.code == _abi.IHE_ERR_ARG
    bad = np.full(1024, 2, np.int32)
    with pytest.raises(IheError) as e:
        check(lib().ihe_decrypt_noalloc(bad.ctypes.data, 16, 1, 1, 1, None))
    assert e.value.code == _abi.IHE_ERR_ARG
