"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py. The product (idash2019_2_amd) never imports it.
Each wrapper names the reference function it restates; oracle.h lists how each
one is pinned.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

N = 1024
CONSTANT_BIDX = 0xFFFFFFFF
NOISE_TABLE_LEN = 1280

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run make -C oracle)")
        _LIB = ctypes.CDLL(path)
        _declare(_LIB)
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _declare(L):
    vp, u32, i32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint64
    L.orc_dtot32.argtypes = [ctypes.c_double]
    L.orc_dtot32.restype = i32
    L.orc_t32tod.argtypes = [i32]
    L.orc_t32tod.restype = ctypes.c_double
    L.orc_encoding_constants.argtypes = [vp]
    L.orc_mul_by_xai.argtypes = [vp, i32, vp, i32]
    L.orc_negacyclic_mul.argtypes = [vp, vp, vp, i32]
    cloud_args = [i32, u32, u32, u32, vp, vp, u32, u32, vp, vp, vp, vp, vp]
    L.orc_cloud_reference.argtypes = cloud_args
    L.orc_cloud_reference.restype = ctypes.c_int
    L.orc_cloud_closed_form.argtypes = cloud_args
    L.orc_cloud_closed_form.restype = ctypes.c_int
    L.orc_phase.argtypes = [i32, vp, vp, vp]
    L.orc_decrypt.argtypes = [i32, u32, vp, u32, vp, vp]
    L.orc_philox4x32_10.argtypes = [vp, vp, vp]
    L.orc_noise_table.argtypes = [vp]
    L.orc_encrypt.argtypes = [i32, u32, u32, u32, u32, vp, vp, u64, vp, vp]
    L.orc_encrypt_cts.argtypes = [i32, u32, u32, u32, u32, vp, vp, u64, vp, u32, vp, vp]
    L.orc_plaintext_score.argtypes = [u32, u32, vp, u32, vp, vp, vp]
    L.orc_cpu_baseline_cloud.argtypes = [ctypes.c_int, u32, u32, u32, vp, vp, u32, u32, vp, vp,
                                         vp, vp, vp]
    L.orc_cpu_baseline_cloud.restype = ctypes.c_double


def dtot32(d: float) -> int:
    return lib().orc_dtot32(d)


def t32tod(x: int) -> float:
    return lib().orc_t32tod(x)


def encoding_constants() -> np.ndarray:
    out = np.zeros(4, np.int32)
    lib().orc_encoding_constants(_p(out))
    return out


def mul_by_xai(a: int, poly: np.ndarray) -> np.ndarray:
    poly = np.ascontiguousarray(poly, np.int32)
    out = np.empty_like(poly)
    lib().orc_mul_by_xai(_p(out), a, _p(poly), poly.size)
    return out


def negacyclic_mul(a: np.ndarray, s: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, np.int32)
    s = np.ascontiguousarray(s, np.int32)
    out = np.empty_like(a)
    lib().orc_negacyclic_mul(_p(out), _p(a), _p(s), a.size)
    return out


def _csr(row_ptr, in_bidx, weight):
    return (np.ascontiguousarray(row_ptr, np.uint32), np.ascontiguousarray(in_bidx, np.uint32),
            np.ascontiguousarray(weight, np.int32))


def cloud(ct, var_in, row_ptr, in_bidx, weight, ns, R, RS, n=N, form="reference"):
    """cloud_compute_score (eval/idash.cpp:763-848). ct: [num_in, 2, n] int32."""
    ct = np.ascontiguousarray(ct, np.int32)
    var_in = np.ascontiguousarray(var_in, np.float64)
    row_ptr, in_bidx, weight = _csr(row_ptr, in_bidx, weight)
    num_out = row_ptr.size - 1
    out = np.zeros((num_out, 2, n), np.int32)
    var_out = np.zeros(num_out, np.float64)
    fn = lib().orc_cloud_reference if form == "reference" else lib().orc_cloud_closed_form
    rc = fn(n, ns, R, RS, _p(ct), _p(var_in), ct.shape[0], num_out, _p(row_ptr), _p(in_bidx),
            _p(weight), _p(out), _p(var_out))
    if rc != 0:
        raise ValueError("model references a ciphertext that is not present")
    return out, var_out


def cpu_baseline_cloud(ct, var_in, row_ptr, in_bidx, weight, ns, R, RS, threads):
    """The reference-structured multithreaded cloud step; returns (out, var, seconds)."""
    ct = np.ascontiguousarray(ct, np.int32)
    var_in = np.ascontiguousarray(var_in, np.float64)
    row_ptr, in_bidx, weight = _csr(row_ptr, in_bidx, weight)
    num_out = row_ptr.size - 1
    out = np.zeros((num_out, 2, N), np.int32)
    var_out = np.zeros(num_out, np.float64)
    secs = lib().orc_cpu_baseline_cloud(threads, ns, R, RS, _p(ct), _p(var_in), ct.shape[0],
                                        num_out, _p(row_ptr), _p(in_bidx), _p(weight), _p(out),
                                        _p(var_out))
    return out, var_out, secs


def phase(ct, key, n=N):
    ct = np.ascontiguousarray(ct, np.int32)
    key = np.ascontiguousarray(key, np.int32)
    out = np.empty(n, np.int32)
    lib().orc_phase(n, _p(ct), _p(key), _p(out))
    return out


def decrypt(cts, key, ns, n=N):
    """decrypt_predictions (eval/idash.cpp:681-761), per ciphertext. cts: [C, 2, n]."""
    cts = np.ascontiguousarray(cts, np.int32)
    key = np.ascontiguousarray(key, np.int32)
    out = np.empty((cts.shape[0], ns), np.float32)
    lib().orc_decrypt(n, ns, _p(cts), cts.shape[0], _p(key), _p(out))
    return out


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.empty(4, np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def noise_table():
    t = np.empty(NOISE_TABLE_LEN, np.uint64)
    lib().orc_noise_table(_p(t))
    return t


def encrypt(geno, key, ns, R, RS, seed, n=N):
    """encrypt_data_ph1 + ph2 (eval/idash.cpp:615-679) with this build's random streams."""
    geno = np.ascontiguousarray(geno, np.int8)
    key = np.ascontiguousarray(key, np.int32)
    tags = geno.shape[0]
    C = (3 * tags + R - 1) // R
    out = np.empty((C, 2, n), np.int32)
    var = np.empty(C, np.float64)
    lib().orc_encrypt(n, ns, R, RS, tags, _p(geno), _p(key), seed, _p(out), _p(var))
    return out, var


def encrypt_cts(geno, key, ns, R, RS, seed, cts, n=N):
    """Ciphertexts cts[i] of encrypt()'s output only: ([len(cts), 2, n], [len(cts)])."""
    geno = np.ascontiguousarray(geno, np.int8)
    key = np.ascontiguousarray(key, np.int32)
    cts = np.ascontiguousarray(cts, np.uint32)
    out = np.empty((cts.size, 2, n), np.int32)
    var = np.empty(cts.size, np.float64)
    lib().orc_encrypt_cts(n, ns, R, RS, geno.shape[0], _p(geno), _p(key), seed, _p(cts), cts.size,
                          _p(out), _p(var))
    return out, var


def plaintext_score(geno, in_bidx, weight, ns):
    geno = np.ascontiguousarray(geno, np.int8)
    in_bidx = np.ascontiguousarray(in_bidx, np.uint32)
    weight = np.ascontiguousarray(weight, np.int32)
    out = np.empty(ns, np.float64)
    lib().orc_plaintext_score(ns, geno.shape[0], _p(geno), in_bidx.size, _p(in_bidx), _p(weight),
                              _p(out))
    return out
