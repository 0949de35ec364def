"""CPU checks of the C ABI: the library loads, exports every entry point the
header declares, and its host-only logic (band compiler, noise table, argument
checks) behaves; no kernel is launched here."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle
from idash2019_2_amd import _abi, band_plan, synth
from idash2019_2_amd._abi import IheError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "idash_hip.h")).read()
    return sorted(set(re.findall(r"\b(ihe_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (ihe_[a-z0-9_]+)$", out, re.M))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    L = _abi.lib()
    for n in names:
        assert hasattr(L, n)


def test_python_binding_covers_header():
    assert set(header_functions()) == set(_abi.SIGNATURES)


def test_noise_table_matches_oracle():
    from idash2019_2_amd import noise_table
    assert (noise_table() == oracle.noise_table()).all()


def test_ctx_create_without_device_reports_error():
    n = ctypes.c_int()
    _abi.check(_abi.lib().ihe_device_count(ctypes.byref(n)))
    if n.value > 0:
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    rc = _abi.lib().ihe_ctx_create(0, ctypes.byref(h))
    assert rc == _abi.IHE_ERR_HIP and h.value is None
    assert b"device" in _abi.lib().ihe_last_error()


def test_null_arguments_are_rejected():
    L = _abi.lib()
    assert L.ihe_sync(None) == _abi.IHE_ERR_ARG
    assert L.ihe_cloud_score(None, None, None, None, None, None) == _abi.IHE_ERR_ARG
    assert L.ihe_band_plan(None, 0, None, 0, None) == _abi.IHE_ERR_ARG


def test_band_plan_groups_shared_windows(monkeypatch):
    c = synth.generate(60, 300, 128, 5, seed=3)
    monkeypatch.setenv("IHE_CLOUD_KERNEL", "direct")
    info = band_plan(c.num_samples, c.num_in_cts, *c.model_csr())
    assert info["num_outputs"] == c.num_outputs
    assert info["variant"] == _abi.VARIANT_VALU
    assert info["max_window"] == 15
    # the three classes of a target always share a group; consecutive targets
    # with the same window merge until 12 outputs
    windows = np.unique(c.lo).size
    assert windows <= info["num_groups"] <= c.num_targets
    assert info["num_cols"] == 16 * info["num_groups"]   # direct: 15 columns padded to 16
    # staged: consecutive windows share their staged columns within a tile
    monkeypatch.setenv("IHE_CLOUD_KERNEL", "staged")
    staged = band_plan(c.num_samples, c.num_in_cts, *c.model_csr())
    assert staged["num_groups"] == info["num_groups"]
    used = np.unique(3 * c.lo[:, None] + np.arange(15)[None, :]).size
    assert used <= staged["num_cols"] < 16 * info["num_groups"] / 2
    # MFMA with wide K budget: tiles of 32 consecutive outputs over the union of
    # their windows
    monkeypatch.setenv("IHE_CLOUD_KERNEL", "mfma")
    monkeypatch.setenv("IHE_CLOUD_SHAPE", "1,4")
    mfma = band_plan(c.num_samples, c.num_in_cts, *c.model_csr())
    assert mfma["variant"] == _abi.VARIANT_MFMA
    assert mfma["num_tiles"] == -(-c.num_outputs // 32)
    assert used <= mfma["num_cols"] <= 32 * mfma["num_tiles"]
    # default: single-K-block tiles of up to 64 rows, cut short where the
    # union passes 32 columns
    assert (mfma["tile_rows"], mfma["tile_kblocks"], mfma["col_tiles"]) == (32, 4, 4)
    # four row blocks per workgroup: tiles of 128 rows; 256 columns hold the
    # union of 128 consecutive outputs here
    monkeypatch.setenv("IHE_CLOUD_SHAPE", "4,8")
    monkeypatch.setenv("IHE_CLOUD_NTW", "32")
    w = band_plan(c.num_samples, c.num_in_cts, *c.model_csr())
    assert (w["variant"], w["tile_rows"], w["tile_kblocks"], w["col_tiles"]) == \
        (_abi.VARIANT_MFMA, 128, 8, 32)
    assert w["num_tiles"] == -(-c.num_outputs // 128)
    monkeypatch.setenv("IHE_CLOUD_NTW", "4")            # not a column-tile count of that kernel:
    assert band_plan(c.num_samples, c.num_in_cts, *c.model_csr())["col_tiles"] == 32  # the default
    monkeypatch.delenv("IHE_CLOUD_NTW")
    # one K-block: tiles are cut where the union passes 32 columns
    monkeypatch.setenv("IHE_CLOUD_SHAPE", "4,1")
    w1 = band_plan(c.num_samples, c.num_in_cts, *c.model_csr())
    assert w1["tile_rows"] == 128 and w1["num_tiles"] > w["num_tiles"]
    monkeypatch.delenv("IHE_CLOUD_KERNEL")
    monkeypatch.delenv("IHE_CLOUD_SHAPE")
    k1 = band_plan(c.num_samples, c.num_in_cts, *c.model_csr())
    assert k1["variant"] == _abi.VARIANT_MFMA
    assert -(-c.num_outputs // 128) <= k1["num_tiles"] <= mfma["num_tiles"]
    assert used <= k1["num_cols"] <= 32 * k1["tile_kblocks"] * k1["num_tiles"]


def test_band_plan_wide_window_picks_multi_kblock_tiles(monkeypatch):
    c = synth.generate(200, 400, 1024, 50, seed=8)      # 150 columns per output
    info = band_plan(c.num_samples, c.num_in_cts, *c.model_csr())
    assert info["variant"] == _abi.VARIANT_MFMA and info["max_window"] == 150
    # tiles over >= 5 K-blocks
    assert info["num_tiles"] >= -(-c.num_outputs // info["tile_rows"])
    assert info["tile_kblocks"] >= 5
    for shape in ("1,4", "4,4"):                       # 128 columns cannot hold a row
        monkeypatch.setenv("IHE_CLOUD_SHAPE", shape)
        assert band_plan(c.num_samples, c.num_in_cts, *c.model_csr())["variant"] == \
            _abi.VARIANT_VALU
        with pytest.raises(IheError) as e:
            band_plan(c.num_samples, c.num_in_cts, *c.model_csr(), variant=_abi.VARIANT_MFMA)
        assert e.value.code == _abi.IHE_ERR_UNSUPPORTED
    monkeypatch.setenv("IHE_CLOUD_SHAPE", "4,6")
    w = band_plan(c.num_samples, c.num_in_cts, *c.model_csr())
    assert (w["variant"], w["tile_rows"], w["tile_kblocks"]) == (_abi.VARIANT_MFMA, 128, 6)
    assert w["num_tiles"] >= -(-c.num_outputs // 128)


def test_band_plan_wide_weights_select_wide_variant():
    rp = np.array([0, 2], np.uint32)
    info = band_plan(1024, 2, rp, np.array([0, 1], np.uint32), np.array([1 << 20, 3], np.int32))
    assert info["variant"] == _abi.VARIANT_WIDE
    for v in (_abi.VARIANT_VALU, _abi.VARIANT_MFMA):
        with pytest.raises(IheError) as e:
            band_plan(1024, 2, rp, np.array([0, 1], np.uint32), np.array([1 << 20, 3], np.int32),
                      variant=v)
        assert e.value.code == _abi.IHE_ERR_UNSUPPORTED
    # 32640..32767 fit int16 but not the balanced two-byte split of the MFMA kernel
    info = band_plan(1024, 2, rp, np.array([0, 1], np.uint32), np.array([32700, 3], np.int32))
    assert info["variant"] == _abi.VARIANT_VALU


def test_band_plan_missing_ciphertext_is_range_error():
    # the reference aborts in getTLWE (eval/idash.h:164)
    with pytest.raises(IheError) as e:
        band_plan(128, 4, [0, 1], [8 * 4], [1])      # R = 8: big index 32 -> FeatIndex 4
    assert e.value.code == _abi.IHE_ERR_RANGE
    band_plan(128, 5, [0, 1], [8 * 4], [1])


def test_band_plan_rejects_duplicates_and_bad_layout():
    with pytest.raises(IheError) as e:
        band_plan(1024, 4, [0, 2], [1, 1], [1, 2])
    assert e.value.code == _abi.IHE_ERR_ARG
    with pytest.raises(IheError) as e:
        band_plan(1024, 4, [0, 2], [_abi.IHE_CONSTANT_BIDX, _abi.IHE_CONSTANT_BIDX], [1, 2])
    assert e.value.code == _abi.IHE_ERR_ARG
    lay = _abi.Layout(100, 9, 113)                   # not R = N/NS
    m = _abi.ModelCSR(0, None, None, None)
    bi = _abi.BandInfo()
    assert _abi.lib().ihe_band_plan(ctypes.byref(lay), 0, ctypes.byref(m), 0,
                                    ctypes.byref(bi)) == _abi.IHE_ERR_ARG


def test_band_plan_empty_and_bias_only_outputs():
    info = band_plan(64, 0, [0, 1, 1], [_abi.IHE_CONSTANT_BIDX], [7])
    # one group without columns
    assert info["num_outputs"] == 2 and info["num_groups"] == 1 and info["max_window"] == 0


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with a message (before any
    device work), as multi-GPU runs are not part of this build."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stderr and r.stdout == ""


class _Buf:
    """Stands in for a DeviceArray: the checks read nbytes only."""

    def __init__(self, nbytes):
        self.nbytes = nbytes
        self.ptr = 1


def test_batch_checks_buffer_sizes_before_the_library():
    from idash2019_2_amd.engine import Band, Batch
    band = Band(engine=None, handle=0, num_in=10, num_out=20, layout=(1024, 1, 1024))
    ok = (band, 1024, 10, _Buf(10 * 8192), _Buf(80), _Buf(20 * 8192), _Buf(160))
    bad_cases = [
        (band, 1024, 10, _Buf(10 * 8192 - 4), _Buf(80), _Buf(20 * 8192), _Buf(160)),
        (band, 1024, 10, _Buf(10 * 8192), _Buf(72), _Buf(20 * 8192), _Buf(160)),
        (band, 1024, 10, _Buf(10 * 8192), _Buf(80), _Buf(19 * 8192), _Buf(160)),
        (band, 1024, 10, _Buf(10 * 8192), _Buf(80), _Buf(20 * 8192), _Buf(152)),
        (band, 1024, 10, _Buf(10 * 8192), None, _Buf(20 * 8192), _Buf(160)),
        (band, 0, 10, _Buf(10 * 8192), _Buf(80), _Buf(20 * 8192), _Buf(160)),
    ]
    for bad in bad_cases:
        with pytest.raises(ValueError):
            Batch(None, [ok, bad])          # engine None: any library call would fail first


def test_server_checks_request_sizes_before_the_library():
    from idash2019_2_amd.engine import Server
    srv = Server.__new__(Server)            # no device: only the host-side checks run
    srv.handle, srv._live, srv._models = 0, {}, {0: (50, 360)}
    geno = np.zeros(50 * 7, np.int8)
    out = np.zeros(360 * 7, np.float32)
    key = np.zeros(1024, np.int32)
    srv.check_request(0, geno, key, out, 7)
    for args in [(1, geno, key, out, 7), (0, geno[:-1], key, out, 7), (0, geno, key, out[:-1], 7),
                 (0, geno, key[:-1], out, 7), (0, geno, key, out, 0),
                 (0, geno.astype(np.int16), key, out, 7), (0, geno, key, out.astype(np.float64), 7)]:
        with pytest.raises(ValueError):
            srv.check_request(*args)
    with pytest.raises(ValueError):
        srv.submit(0, geno, key, 1, out[:-7], num_samples=7)


def test_decrypt_noalloc_checks_arguments_before_the_device():
    """ihe_decrypt_noalloc validates its arguments before any HIP call: these
    return IHE_ERR_ARG on a machine without a GPU too."""
    L = _abi.lib()
    key = np.zeros(1024, np.int32)
    bad = np.full(1024, 2, np.int32)
    for args in [(None, 16, 1, 1, 1, None), (key.ctypes.data, 0, 1, 1, 1, None),
                 (key.ctypes.data, 1025, 1, 1, 1, None),
                 (key.ctypes.data, 16, 1, (1 << 26) + 1, 1, None),
                 (bad.ctypes.data, 16, 1, 1, 1, None),
                 (key.ctypes.data, 16, None, 4, None, None)]:
        assert L.ihe_decrypt_noalloc(*args) == _abi.IHE_ERR_ARG
