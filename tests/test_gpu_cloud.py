"""GPU parity of the cloud step (cloud_compute_score, eval/idash.cpp:763-848)
against the CPU oracle, bit-exact on ciphertext words and variances."""
import numpy as np
import pytest

import oracle
from idash2019_2_amd import _abi, synth
from idash2019_2_amd._abi import IheError

pytestmark = pytest.mark.gpu

# The cloud kernels. MFMA shapes (IHE_CLOUD_SHAPE = row blocks, K-blocks):
# 64-row single-K-block tiles (the default at 5 neighbours), 32-row tiles,
# and 2 or 3 K-blocks (forcing padded K-blocks on narrow windows); NTW = 32-word
# column tiles per wave. 128-row tiles ("w") run the LDS-staged kernel up to 6
# K-blocks (16-byte staging loads where unrotated or rotated with RS % 4 == 0,
# as at NS = 128 and 512; IHE_MFMAW_LDS=1 forces dword loads) and the
# register-staged one above (IHE_MFMAW_LDS=0 forces it). Then the LDS-staged
# dot2 VALU and direct dot2 VALU.
KERNELS = {
    "mfma": {},
    "mfma-rb1": {"IHE_CLOUD_SHAPE": "1,1"},
    "mfma-k2": {"IHE_CLOUD_SHAPE": "1,2", "IHE_CLOUD_NTW": "8"},
    "mfma-k3": {"IHE_CLOUD_SHAPE": "1,3"},
    "mfma-ntw16": {"IHE_CLOUD_NTW": "16"},
    "mfma-rb2": {"IHE_CLOUD_SHAPE": "2,1"},
    "mfma-w1": {"IHE_CLOUD_SHAPE": "4,1", "IHE_CLOUD_NTW": "8"},
    "mfma-w2": {"IHE_CLOUD_SHAPE": "4,2"},
    "mfma-w2-regs": {"IHE_CLOUD_SHAPE": "4,2", "IHE_MFMAW_LDS": "0"},
    "mfma-w4-dword": {"IHE_CLOUD_SHAPE": "4,4", "IHE_MFMAW_LDS": "1"},
    "mfma-w6": {"IHE_CLOUD_SHAPE": "4,6", "IHE_CLOUD_NTW": "16"},
    "mfma-w7": {"IHE_CLOUD_SHAPE": "4,7", "IHE_CLOUD_NTW": "32"},
    "staged": {"IHE_CLOUD_KERNEL": "staged"},
    "direct": {"IHE_CLOUD_KERNEL": "direct"},
}


@pytest.fixture(params=list(KERNELS))
def kernel(request, monkeypatch):
    for k in ("IHE_CLOUD_KERNEL", "IHE_CLOUD_SHAPE", "IHE_CLOUD_NTW", "IHE_MFMAW_LDS"):
        monkeypatch.delenv(k, raising=False)
    for k, v in KERNELS[request.param].items():
        monkeypatch.setenv(k, v)
    return request.param


def expected_variant(kernel):
    return _abi.VARIANT_MFMA if kernel.startswith("mfma") else _abi.VARIANT_VALU


def run_cloud(engine, ct, var, row_ptr, in_bidx, weight, ns, variant=_abi.VARIANT_AUTO):
    num_in = ct.shape[0]
    band = engine.band(ns, num_in, row_ptr, in_bidx, weight, variant)
    d_in = engine.to_device(ct)
    d_var = engine.to_device(var)
    num_out = len(row_ptr) - 1
    d_out = engine.empty((num_out, 2, 1024), np.int32)
    d_out_var = engine.empty((num_out,), np.float64)
    engine.cloud(band, d_in, d_var, d_out, d_out_var)
    out, ov = d_out.numpy(), d_out_var.numpy()
    info = band.info()
    band.free()
    return out, ov, info


def random_cts(rng, num_in):
    ct = rng.integers(-(1 << 31), 1 << 31, (num_in, 2, 1024)).astype(np.int32)
    ct[0, :, :8] = [-(1 << 31), (1 << 31) - 1, -1, 0, 1, 0x7FFF, 0x8000, -0x8000]
    return ct


@pytest.mark.parametrize("ns", [128, 1024, 100, 333, 512, 1000, 1, 7])
def test_banded_model_matches_oracle(engine, kernel, ns):
    c = synth.generate(40, 150, ns, 5, seed=ns)
    rng = np.random.default_rng(ns)
    ct = random_cts(rng, c.num_in_cts)
    var = np.full(c.num_in_cts, 2.0 ** -50)
    var[::3] = 2.0 ** -40
    rp, ib, w = c.model_csr()
    out, ov, info = run_cloud(engine, ct, var, rp, ib, w, ns)
    assert info["variant"] == expected_variant(kernel)
    ref, rv = oracle.cloud(ct, var, rp, ib, w, *c.layout)
    assert (out == ref).all()
    assert (ov == rv).all()


@pytest.mark.parametrize("ns", [128, 1024])
def test_zero_weights_dropped_matches_oracle(engine, kernel, ns):
    c = synth.generate(30, 90, ns, 5, seed=7)
    c.weights[::4, 3] = 0                            # .hr files omit zero coefficients
    rng = np.random.default_rng(1)
    ct = random_cts(rng, c.num_in_cts)
    var = np.full(c.num_in_cts, 2.0 ** -50)
    m = c.model_csr(with_zeros=False)
    out, ov, _ = run_cloud(engine, ct, var, *m, ns)
    ref, rv = oracle.cloud(ct, var, *m, *c.layout)
    assert (out == ref).all() and (ov == rv).all()


@pytest.mark.parametrize("ns,variant,wmax", [(1024, _abi.VARIANT_VALU, 32767),
                                             (64, _abi.VARIANT_VALU, 32767),
                                             (1024, _abi.VARIANT_MFMA, 32639),
                                             (77, _abi.VARIANT_MFMA, 32639),
                                             (1024, _abi.VARIANT_WIDE, (1 << 31) - 1),
                                             (50, _abi.VARIANT_WIDE, (1 << 31) - 1),
                                             (1024, _abi.VARIANT_AUTO, 1 << 20)])
def test_random_sparse_model_matches_oracle(engine, ns, variant, wmax):
    """Arbitrary (non-window) column sets, widths up to 40, unsorted entries."""
    rng = np.random.default_rng(ns + wmax % 1000)
    R = 1024 // ns
    num_cols = 300
    num_in = (num_cols + R - 1) // R
    ct = random_cts(rng, num_in)
    var = np.full(num_in, 2.0 ** -50)
    rows = [0]
    ib, w = [], []
    for o in range(97):
        k = int(rng.integers(0, 40))
        cols = rng.choice(num_cols, size=k, replace=False)
        ib += [int(x) for x in cols]
        w += [int(rng.integers(-wmax, wmax + 1)) for _ in cols]
        if o % 5:
            ib.append(_abi.IHE_CONSTANT_BIDX)
            w.append(int(rng.integers(-wmax, wmax + 1)))
        rows.append(len(ib))
    rp = np.array(rows, np.uint32)
    ib = np.array(ib, np.uint32)
    w = np.array(w, np.int32)
    out, ov, info = run_cloud(engine, ct, var, rp, ib, w, ns, variant)
    if wmax > 32767:
        assert info["variant"] == _abi.VARIANT_WIDE
    else:
        assert info["variant"] == (variant or _abi.VARIANT_MFMA)
    ref, rv = oracle.cloud(ct, var, rp, ib, w, ns, R, 1024 // R)
    assert (out == ref).all()
    assert (ov == rv).all()


@pytest.mark.parametrize("span", [300, 70000])
def test_variance_entry_layouts_match_oracle(engine, span):
    """Output variances (tLweAddMulTo's w*w * var, region-0 columns only) with
    the band's packed entries (a row's big indices within 65534 of each other)
    and, for rows spanning 70000 big indices, its unpacked ones; NS = 16, so
    R = 64 and only every 64th column reaches the variance."""
    ns, R = 16, 64
    rng = np.random.default_rng(span)
    num_in = (span + R - 1) // R
    ct = random_cts(rng, num_in)
    var = rng.integers(1, 1 << 20, num_in) * 2.0 ** -60
    rows, ib, w = [0], [], []
    for o in range(150):
        cols = rng.choice(span, size=int(rng.integers(1, 30)), replace=False)
        if o % 3 == 0:
            cols = np.concatenate([cols[:1] * 0, cols, [span - 1]])    # the full span
            cols = np.unique(cols)
        cols = cols[rng.permutation(cols.size)]                      # row order, unsorted
        ib += [int(x) for x in cols]
        w += [int(rng.integers(-32768, 32640)) for _ in cols]
        ib.append(_abi.IHE_CONSTANT_BIDX)
        w.append(int(rng.integers(-2048, 2048)))
        rows.append(len(ib))
    rp, ib, w = np.array(rows, np.uint32), np.array(ib, np.uint32), np.array(w, np.int32)
    out, ov, _ = run_cloud(engine, ct, var, rp, ib, w, ns)
    ref, rv = oracle.cloud(ct, var, rp, ib, w, ns, R, 1024 // R)
    assert (out == ref).all()
    assert (ov == rv).all()


def test_wide_window_matches_oracle(engine, kernel):
    """50 neighbours (150 columns, several register chunks per group)."""
    c = synth.generate(120, 200, 1024, 50, seed=50)
    rng = np.random.default_rng(50)
    ct = random_cts(rng, c.num_in_cts)
    var = np.full(c.num_in_cts, 2.0 ** -50)
    rp, ib, w = c.model_csr()
    out, ov, info = run_cloud(engine, ct, var, rp, ib, w, 1024)
    assert info["max_window"] == 150
    if kernel in ("mfma", "mfma-ntw16", "mfma-w7"):
        assert info["variant"] == _abi.VARIANT_MFMA
    ref, rv = oracle.cloud(ct, var, rp, ib, w, *c.layout)
    assert (out == ref).all() and (ov == rv).all()


def test_missing_ciphertext_raises(engine):
    with pytest.raises(IheError) as e:
        engine.band(1024, 3, [0, 1], [5], [1])
    assert e.value.code == _abi.IHE_ERR_RANGE


def test_output_without_variance_buffer(engine):
    c = synth.generate(20, 40, 256, 5, seed=2)
    rng = np.random.default_rng(2)
    ct = random_cts(rng, c.num_in_cts)
    rp, ib, w = c.model_csr()
    band = engine.band(256, c.num_in_cts, rp, ib, w)
    d_in = engine.to_device(ct)
    d_out = engine.empty((c.num_outputs, 2, 1024), np.int32)
    engine.cloud(band, d_in, None, d_out, None)
    ref, _ = oracle.cloud(ct, np.zeros(c.num_in_cts), rp, ib, w, *c.layout)
    assert (d_out.numpy() == ref).all()
    band.free()


@pytest.mark.parametrize("kernel", ["mfma", "mfma-rb1", "staged"], indirect=True)
def test_full_size_rows_match_oracle(engine, kernel):
    """The full 16,184-tag / 80,882-target / 1,024-sample workload on the GPU;
    a random subset of 600 output rows (plus the first and last) is recomputed
    by the oracle, and every output's b-polynomial tail and variance are checked
    by closed form."""
    c = synth.workload("full")
    rng = np.random.default_rng(99)
    num_in = c.num_in_cts
    ct = rng.integers(-(1 << 31), 1 << 31, (num_in, 2, 1024), dtype=np.int64).astype(np.int32)
    var = np.full(num_in, 2.0 ** -50)
    rp, ib, w = c.model_csr()
    out, ov, info = run_cloud(engine, ct, var, rp, ib, w, 1024)
    rows = np.unique(np.r_[0, c.num_outputs - 1, rng.choice(c.num_outputs, 600, replace=False)])
    sub_rp = np.r_[0, np.cumsum(rp[rows + 1] - rp[rows])].astype(np.uint32)
    sub_ib = np.concatenate([ib[rp[r]:rp[r + 1]] for r in rows])
    sub_w = np.concatenate([w[rp[r]:rp[r + 1]] for r in rows])
    ref, rv = oracle.cloud(ct, var, sub_rp, sub_ib, sub_w, 1024, 1, 1024)
    assert (out[rows] == ref).all()
    assert (ov[rows] == rv).all()
    w2 = (c.weights.astype(np.int64) ** 2).sum(axis=1).astype(np.float64) * 2.0 ** -50
    assert (ov == w2).all()


@pytest.mark.parametrize("shape", ["2,1", "4,2", "1,2"])
@pytest.mark.parametrize("ns", [1024, 128, 341, 16, 1])
def test_one_band_serves_every_layout(engine, ns, shape, monkeypatch):
    """An MFMA band compiled once (its columns are big indices) runs cohorts of
    any sample count: ciphertext c / R, rotation (c % R) * RS per launch."""
    monkeypatch.setenv("IHE_CLOUD_SHAPE", shape)
    c = synth.generate(40, 120, 1024, 5, seed=3)
    rp, ib, w = c.model_csr()
    band = engine.band(1024, c.num_in_cts, rp, ib, w)
    info = band.info()
    assert info["variant"] == _abi.VARIANT_MFMA
    assert (info["tile_rows"], info["tile_kblocks"]) == (32 * int(shape[0]), int(shape[2]))
    R = 1024 // ns
    num_in = (3 * c.num_tags + R - 1) // R
    rng = np.random.default_rng(ns)
    ct = random_cts(rng, num_in)
    var = rng.integers(1, 1 << 20, num_in) * 2.0 ** -60
    d_in, d_var = engine.to_device(ct), engine.to_device(var)
    d_out = engine.empty((c.num_outputs, 2, 1024), np.int32)
    d_ov = engine.empty((c.num_outputs,), np.float64)
    engine.cloud_layout(band, ns, num_in, d_in, d_var, d_out, d_ov)
    ref, rv = oracle.cloud(ct, var, rp, ib, w, ns, R, 1024 // R)
    assert (d_out.numpy() == ref).all()
    assert (d_ov.numpy() == rv).all()
    with pytest.raises(IheError) as e:          # a ciphertext the model needs is missing
        engine.cloud_layout(band, ns, num_in - 1, d_in, d_var, d_out, d_ov)
    assert e.value.code == _abi.IHE_ERR_RANGE
    band.free()
