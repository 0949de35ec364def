"""Readers of the binary files the idash-compatible host writes, in the
reference's formats (eval/idash.cpp:128-246 params, :476-502 keys,
:513-613 ciphertext records):
  params.bin   7 uint32 (NUM_SAMPLES, NUM_INPUT_POSITIONS, NUM_OUTPUT_POSITIONS,
               NUM_INPUT_FEATURES, NUM_OUTPUT_FEATURES, NUM_REGIONS, REGION_SIZE),
               then per tag position uint64 pos + 3 uint32 big indices, then per
               target (file order) uint64 pos + NUL-terminated name + 3 uint32
  keys.bin     params stream, then int32 N, int32 k, double alpha_min,
               double alpha_max, N int32 key coefficients (this build's key stream)
  *.bin data   uint64 count, then per record uint32 index + one 8204-byte sample
               stream: int32 framing word, a[N], b[N] (int32), double variance
Also a writer of ciphertext record files (test inputs in the reference format).
"""
from __future__ import annotations

import struct

import numpy as np

N = 1024
SAMPLE_STREAM_BYTES = 4 + 8 * N + 8
RECORD_BYTES = 4 + SAMPLE_STREAM_BYTES


def _params(buf: bytes, off: int = 0):
    (ns, nin, nout, nfeat_in, nfeat_out, R, RS) = struct.unpack_from("<7I", buf, off)
    off += 28
    in_index = {}
    for _ in range(nin):
        pos, b0, b1, b2 = struct.unpack_from("<Q3I", buf, off)
        off += 20
        in_index[pos] = (b0, b1, b2)
    out_names, out_index = [], {}
    for _ in range(nout):
        (pos,) = struct.unpack_from("<Q", buf, off)
        off += 8
        end = buf.index(b"\0", off)
        name = buf[off:end].decode()
        off = end + 1
        b = struct.unpack_from("<3I", buf, off)
        off += 12
        out_names.append((pos, name))
        out_index[pos] = b
    p = dict(NUM_SAMPLES=ns, NUM_INPUT_POSITIONS=nin, NUM_OUTPUT_POSITIONS=nout,
             NUM_INPUT_FEATURES=nfeat_in, NUM_OUTPUT_FEATURES=nfeat_out, NUM_REGIONS=R,
             REGION_SIZE=RS, in_features_index=in_index, out_features_index=out_index,
             out_position_names=out_names)
    return p, off


def read_params(path: str) -> dict:
    with open(path, "rb") as f:
        return _params(f.read())[0]


def read_key(path: str):
    with open(path, "rb") as f:
        buf = f.read()
    p, off = _params(buf)
    n, k, amin, amax = struct.unpack_from("<iidd", buf, off)
    off += 24
    key = np.frombuffer(buf, np.int32, n, off).copy()
    return p, key, dict(N=n, k=k, alpha_min=amin, alpha_max=amax)


def read_records(path: str, slots: int | None = None):
    """Returns (indices [count], dense ct [slots][2][N] int32, var [slots]) with
    records placed at their index (slots defaults to max index + 1)."""
    with open(path, "rb") as f:
        (count,) = struct.unpack("<Q", f.read(8))
        raw = np.frombuffer(f.read(count * RECORD_BYTES), np.uint8).reshape(count, RECORD_BYTES)
    idx = raw[:, 0:4].copy().view(np.uint32).ravel()
    words = raw[:, 8:8 + 8 * N].copy().view(np.int32).reshape(count, 2, N)
    var = raw[:, 8 + 8 * N:].copy().view(np.float64).ravel()
    if slots is None:
        slots = int(idx.max()) + 1 if count else 0
    ct = np.zeros((slots, 2, N), np.int32)
    v = np.zeros(slots, np.float64)
    ct[idx] = words
    v[idx] = var
    return idx, ct, v


def read_stream_words(path: str) -> np.ndarray:
    """The framing word that opens each record's sample stream, in file order."""
    with open(path, "rb") as f:
        (count,) = struct.unpack("<Q", f.read(8))
        raw = np.frombuffer(f.read(count * RECORD_BYTES), np.uint8).reshape(count, RECORD_BYTES)
    return raw[:, 4:8].copy().view(np.int32).ravel()


def write_records(path: str, idx, ct, var, stream_word: int = 0) -> None:
    """Writes records (index idx[i], sample ct[idx[i]], variance var[idx[i]]) in
    the order of idx, each sample stream opened with stream_word."""
    idx = np.asarray(idx, np.uint32)
    ct = np.ascontiguousarray(ct, np.int32)
    var = np.ascontiguousarray(var, np.float64)
    raw = np.empty((idx.size, RECORD_BYTES), np.uint8)
    raw[:, 0:4] = idx.view(np.uint8).reshape(-1, 4)
    raw[:, 4:8] = np.full(idx.size, stream_word, np.int32).view(np.uint8).reshape(-1, 4)
    raw[:, 8:8 + 8 * N] = ct[idx].reshape(idx.size, -1).view(np.uint8)
    raw[:, 8 + 8 * N:] = var[idx].view(np.uint8).reshape(-1, 8)
    with open(path, "wb") as f:
        f.write(struct.pack("<Q", idx.size))
        f.write(raw.tobytes())
