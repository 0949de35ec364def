"""GPU parity of encrypt (encrypt_data_ph1/ph2, eval/idash.cpp:615-679) and
decrypt (decrypt_predictions, eval/idash.cpp:681-761): exact negacyclic a*s,
bit-exact against the CPU oracle, and the encrypt -> cloud -> decrypt round
trip against the plaintext scorer (eval/idash.cpp:888-927)."""
import numpy as np
import pytest

import oracle
from idash2019_2_amd import synth

pytestmark = pytest.mark.gpu


def ring_kernel(request, monkeypatch, var):
    """mfma: the int8-MFMA Toeplitz product (default: two column tiles per
    wave, items from the context's counters); mfma1: one column tile per wave
    (IHE_RING_TILES=1); mfma_fixed: fixed items per workgroup (IHE_RING_QUEUE=0);
    valu: the one-wave-per-ciphertext VALU kernel."""
    for v in ("IHE_RING_TILES", "IHE_RING_QUEUE", var):
        monkeypatch.delenv(v, raising=False)
    if request.param == "valu":
        monkeypatch.setenv(var, "valu")
    elif request.param == "mfma1":
        monkeypatch.setenv("IHE_RING_TILES", "1")
    elif request.param == "mfma_fixed":
        monkeypatch.setenv("IHE_RING_QUEUE", "0")
    return request.param


# encrypt kernels: a halves by Philox then b on the MFMA product, or the VALU
# kernel (IHE_ENCRYPT_KERNEL=valu)
@pytest.fixture(params=["mfma", "mfma1", "mfma_fixed", "valu"])
def enc_kernel(request, monkeypatch):
    return ring_kernel(request, monkeypatch, "IHE_ENCRYPT_KERNEL")


def messages(geno, ns, R, RS, C):
    """setScore's one-hot and missing-value terms (eval/idash.cpp:642-679) as an
    array added to b, computed here independently of the oracle."""
    one, nan = 1 << 18, (131072, 87381, 43690)
    add = np.zeros((C, 1024), np.int64)
    tags = geno.shape[0]
    for c in range(C):
        for r in range(R):
            bidx = c * R + r
            if bidx >= 3 * tags:
                break
            tag, v = divmod(bidx, 3)
            g = geno[tag].astype(np.int64)
            add[c, r * RS:r * RS + ns] = np.where(g == v, one, 0) + np.where(g < 0, nan[v], 0)
    return add


@pytest.mark.parametrize("ns", [128, 1024, 100, 1, 333])
def test_encrypt_matches_oracle(engine, enc_kernel, ns):
    rng = np.random.default_rng(ns)
    tags = 23
    geno = rng.integers(0, 3, (tags, ns)).astype(np.int8)
    geno[rng.random((tags, ns)) < 0.05] = -1
    key = rng.integers(0, 2, 1024).astype(np.int32)
    k = engine.key(key)
    R = 1024 // ns
    C = (3 * tags + R - 1) // R
    d_geno = engine.to_device(geno)
    d_out = engine.empty((C, 2, 1024), np.int32)
    d_var = engine.empty((C,), np.float64)
    assert engine.encrypt(k, ns, d_geno, 0xDEADBEEF12345, d_out, d_var) == C
    ref, rv = oracle.encrypt(geno, key, ns, R, 1024 // R, seed=0xDEADBEEF12345)
    out = d_out.numpy()
    assert (out == ref).all()
    assert (d_var.numpy() == rv).all()
    k.free()


@pytest.mark.parametrize("ns,tags", [(1024, 150), (64, 150), (2, 40)])
def test_encrypt_many_ciphertexts_and_extreme_keys(engine, enc_kernel, ns, tags):
    """Several groups of 32 ciphertexts (a partial last one) with a random, an
    all-zero and an all-one key."""
    rng = np.random.default_rng(tags + ns)
    geno = rng.integers(0, 3, (tags, ns)).astype(np.int8)
    geno[rng.random((tags, ns)) < 0.02] = -1
    R = 1024 // ns
    C = (3 * tags + R - 1) // R
    d_geno = engine.to_device(geno)
    for key in (rng.integers(0, 2, 1024), np.zeros(1024), np.ones(1024)):
        key = key.astype(np.int32)
        k = engine.key(key)
        d_out = engine.empty((C, 2, 1024), np.int32)
        d_var = engine.empty((C,), np.float64)
        engine.encrypt(k, ns, d_geno, 77 + int(key.sum()), d_out, d_var)
        ref, rv = oracle.encrypt(geno, key, ns, R, 1024 // R, seed=77 + int(key.sum()))
        assert (d_out.numpy() == ref).all()
        assert (d_var.numpy() == rv).all()
        k.free()


def test_encrypt_without_genotypes_is_ph1(engine, enc_kernel):
    """A null genotype pointer gives encrypt_data_ph1 alone (eval/idash.cpp:615-640):
    the oracle's ciphertexts minus their messages."""
    import ctypes
    from idash2019_2_amd import _abi
    ns, tags = 128, 30
    R, RS = 8, 128
    C = (3 * tags + R - 1) // R
    rng = np.random.default_rng(3)
    geno = rng.integers(0, 3, (tags, ns)).astype(np.int8)
    geno[rng.random((tags, ns)) < 0.1] = -1
    key = rng.integers(0, 2, 1024).astype(np.int32)
    k = engine.key(key)
    d_out = engine.empty((C, 2, 1024), np.int32)
    d_var = engine.empty((C,), np.float64)
    lay = _abi.Layout(ns, R, RS)
    _abi.check(_abi.lib().ihe_encrypt(engine.ctx, k.handle, ctypes.byref(lay), tags, None, 99,
                                      d_out.ptr, d_var.ptr))
    ref, rv = oracle.encrypt(geno, key, ns, R, RS, seed=99)
    want = ref.astype(np.int64)
    want[:, 1, :] -= messages(geno, ns, R, RS, C)
    assert (d_out.numpy() == want.astype(np.uint32).astype(np.int32)).all()
    assert (d_var.numpy() == rv).all()
    k.free()


# decrypt kernels: the MFMA product, or the VALU sliding-window kernel
# (IHE_DECRYPT_KERNEL=valu)
@pytest.fixture(params=["mfma", "mfma1", "mfma_fixed", "valu"])
def dec_kernel(request, monkeypatch):
    return ring_kernel(request, monkeypatch, "IHE_DECRYPT_KERNEL")


@pytest.mark.parametrize("ns", [1, 16, 64, 65, 128, 200, 256, 512, 513, 1000, 1024])
def test_decrypt_matches_oracle(engine, dec_kernel, ns):
    rng = np.random.default_rng(ns)
    key = rng.integers(0, 2, 1024).astype(np.int32)
    n = 75                       # two full groups of 32 ciphertexts and a partial one
    cts = rng.integers(-(1 << 31), 1 << 31, (n, 2, 1024)).astype(np.int32)
    k = engine.key(key)
    d_in = engine.to_device(cts)
    d_out = engine.empty((n, ns), np.float32)
    engine.decrypt(k, ns, d_in, n, d_out)
    assert (d_out.numpy() == oracle.decrypt(cts, key, ns)).all()
    k.free()


def test_decrypt_launches_of_different_sizes_in_a_row(engine):
    """The MFMA kernel's work counters are cleared by the last workgroup of each
    launch: launches with many items (a full grid), with fewer items than
    workgroups, and many again, on one context, each equal to the oracle."""
    rng = np.random.default_rng(11)
    key = rng.integers(0, 2, 1024).astype(np.int32)
    k = engine.key(key)
    for n, ns in ((4100, 1024), (3, 64), (1, 1), (300, 700), (40, 1024)):
        cts = rng.integers(-(1 << 31), 1 << 31, (n, 2, 1024)).astype(np.int32)
        d_in = engine.to_device(cts)
        d_out = engine.empty((n, ns), np.float32)
        engine.decrypt(k, ns, d_in, n, d_out)
        assert (d_out.numpy() == oracle.decrypt(cts, key, ns)).all(), (n, ns)
        d_in.free()
        d_out.free()
    k.free()


def test_decrypt_extreme_keys(engine, dec_kernel):
    for key in (np.zeros(1024, np.int32), np.ones(1024, np.int32)):
        rng = np.random.default_rng(int(key[0]))
        cts = rng.integers(-(1 << 31), 1 << 31, (5, 2, 1024)).astype(np.int32)
        k = engine.key(key)
        d_in = engine.to_device(cts)
        d_out = engine.empty((5, 1024), np.float32)
        engine.decrypt(k, 1024, d_in, 5, d_out)
        assert (d_out.numpy() == oracle.decrypt(cts, key, 1024)).all()
        k.free()


@pytest.mark.parametrize("ns", [128, 1024, 77])
def test_round_trip_matches_oracle_and_plaintext(engine, ns):
    c = synth.generate(50, 120, ns, 5, seed=ns + 1)
    c.geno[2, ::7] = -1
    nsv, R, RS = c.layout
    k = engine.key(c.key)
    d_geno = engine.to_device(c.geno)
    C = c.num_in_cts
    d_ct = engine.empty((C, 2, 1024), np.int32)
    d_var = engine.empty((C,), np.float64)
    engine.encrypt(k, ns, d_geno, 77, d_ct, d_var)
    rp, ib, w = c.model_csr()
    band = engine.band(ns, C, rp, ib, w)
    d_out = engine.empty((c.num_outputs, 2, 1024), np.int32)
    d_ov = engine.empty((c.num_outputs,), np.float64)
    engine.cloud(band, d_ct, d_var, d_out, d_ov)
    d_dec = engine.empty((c.num_outputs, ns), np.float32)
    engine.decrypt(k, ns, d_out, c.num_outputs, d_dec)
    dec = d_dec.numpy()
    # exact against the oracle pipeline
    cts, var = oracle.encrypt(c.geno, c.key, ns, R, RS, seed=77)
    out, ov = oracle.cloud(cts, var, rp, ib, w, ns, R, RS)
    assert (d_out.numpy() == out).all() and (d_ov.numpy() == ov).all()
    assert (dec == oracle.decrypt(out, c.key, ns)).all()
    # and within the noise bound of the plaintext scorer
    for o in range(0, c.num_outputs, 7):
        s, e = rp[o], rp[o + 1]
        score = oracle.plaintext_score(c.geno, ib[s:e], w[s:e], ns)
        bound = (np.abs(w[s:e]).sum() * 1280) / 2.0 ** 32 + 1e-6
        assert np.abs(dec[o] - score).max() <= bound
    band.free()
    k.free()
