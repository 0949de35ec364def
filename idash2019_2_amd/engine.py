"""Python front end of the C ABI: device buffers, bands, keys and the three
operations of the secure-imputation path (encrypt, cloud, decrypt).

Mirrors the reference's operation split (eval/idash.h:248-254): encrypt_data
(ph1 + ph2), cloud_compute_score and decrypt_predictions, with the data held
in device memory between them.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._abi import check, lib

N = _abi.IHE_N
CT_WORDS = 2 * N


def layout_for(num_samples: int) -> tuple[int, int, int]:
    """(NS, R, RS) exactly as keygen_ph1 derives them (eval/idash.cpp:409-412)."""
    if not 1 <= num_samples <= N:
        raise ValueError("NUM_SAMPLES must be in [1, 1024] (one ciphertext packs N slots)")
    R = N // num_samples
    return num_samples, R, N // R


class DeviceArray:
    """A device allocation with a numpy shape/dtype attached."""

    def __init__(self, engine: "Engine", shape, dtype):
        self.engine = engine
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().ihe_malloc(engine.ctx, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value or 0

    def upload(self, host: np.ndarray) -> "DeviceArray":
        host = np.ascontiguousarray(host, dtype=self.dtype)
        if host.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        if self.nbytes:
            check(lib().ihe_memcpy_h2d(self.engine.ctx, self.ptr, host.ctypes.data, self.nbytes))
        return self

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().ihe_memcpy_d2h(self.engine.ctx, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def take(self, rows) -> np.ndarray:
        """Rows of the leading axis (any order, repeats allowed) copied to the host,
        one copy per run of consecutive rows: a few rows of a multi-GB array."""
        rows = np.asarray(rows, np.int64).ravel()
        lead = self.shape[0] if self.shape else 0
        if rows.size and (rows.min() < 0 or rows.max() >= lead):
            raise IndexError("row index out of range")
        row_bytes = self.nbytes // lead if lead else 0
        uniq = np.unique(rows)
        got = np.empty((uniq.size,) + self.shape[1:], self.dtype)
        i = 0
        while i < uniq.size:
            j = i + 1
            while j < uniq.size and uniq[j] == uniq[j - 1] + 1:
                j += 1
            check(lib().ihe_memcpy_d2h(self.engine.ctx, got[i:j].ctypes.data,
                                       self.ptr + int(uniq[i]) * row_bytes, (j - i) * row_bytes))
            i = j
        return got[np.searchsorted(uniq, rows)]

    def fill_zero(self) -> None:
        if self.nbytes:
            check(lib().ihe_memset(self.engine.ctx, self.ptr, 0, self.nbytes))

    def free(self) -> None:
        if self.ptr:
            check(lib().ihe_free(self.engine.ctx, self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            if self.ptr and self.engine.ctx:
                lib().ihe_free(self.engine.ctx, self.ptr)
                self.ptr = 0
        except Exception:
            pass


@dataclass
class Band:
    engine: "Engine"
    handle: int
    num_in: int
    num_out: int
    layout: tuple

    def info(self) -> dict:
        bi = _abi.BandInfo()
        check(lib().ihe_band_get_info(self.handle, ctypes.byref(bi)))
        return {f: getattr(bi, f) for f, _ in _abi.BandInfo._fields_}

    def free(self) -> None:
        if self.handle:
            check(lib().ihe_band_destroy(self.handle))
            self.handle = 0


def check_cloud_buffers(band: "Band", num_in: int, d_in, d_in_var, d_out, d_out_var,
                        what: str = "") -> None:
    """The sizes the cloud kernels address: num_in input ciphertexts and
    variances, band.num_out output ciphertexts and variances (ValueError)."""
    need = ((d_in, num_in * CT_WORDS * 4, "input ciphertext"),
            (d_in_var, num_in * 8, "input variance"),
            (d_out, band.num_out * CT_WORDS * 4, "output ciphertext"),
            (d_out_var, band.num_out * 8, "output variance"))
    for buf, size, name in need:
        if buf is not None and buf.nbytes < size:
            raise ValueError(f"{what}{name} buffer too small ({buf.nbytes} < {size} bytes)")
    if d_out is None or (num_in > 0 and d_in is None):
        raise ValueError(f"{what}null ciphertext buffer")
    if d_out_var is not None and num_in > 0 and d_in_var is None:
        raise ValueError(f"{what}output variances need input variances")


class Batch:
    """Many cohorts' cloud steps in one kernel launch (ihe_batch_*), or one launch
    per kernel family when a family has no body in that kernel."""

    def __init__(self, engine: "Engine", cohorts):
        self.engine = engine
        for i, c in enumerate(cohorts):
            band, ns, num_in, d_in, d_in_var, d_out, d_out_var = c
            layout_for(ns)
            check_cloud_buffers(band, num_in, d_in, d_in_var, d_out, d_out_var, f"cohort {i}: ")
        arr = (_abi.Cohort * max(1, len(cohorts)))()
        self._keep = []
        self.handle = 0
        for i, c in enumerate(cohorts):
            band, ns, num_in, d_in, d_in_var, d_out, d_out_var = c
            _, R, RS = layout_for(ns)
            arr[i].band = band.handle
            arr[i].layout = _abi.Layout(ns, R, RS)
            arr[i].num_in_cts = num_in
            arr[i].d_in = d_in.ptr
            arr[i].d_in_var = d_in_var.ptr if d_in_var is not None else None
            arr[i].d_out = d_out.ptr
            arr[i].d_out_var = d_out_var.ptr if d_out_var is not None else None
            self._keep.append(c)
        h = ctypes.c_void_p()
        check(lib().ihe_batch_create(engine.ctx, arr, len(cohorts), ctypes.byref(h)))
        self.handle = h.value

    def capture(self) -> None:
        check(lib().ihe_batch_capture(self.handle))

    def run(self) -> None:
        check(lib().ihe_batch_run(self.handle))

    def info(self) -> dict:
        bi = _abi.BatchInfo()
        check(lib().ihe_batch_get_info(self.handle, ctypes.byref(bi)))
        return {f: getattr(bi, f) for f, _ in _abi.BatchInfo._fields_}

    def free(self) -> None:
        if self.handle:
            check(lib().ihe_batch_destroy(self.handle))
            self.handle = 0


class PinnedArray:
    """Page-locked host memory (ihe_host_alloc) viewed as a numpy array, so that
    the server's 2-D copies run asynchronously."""

    def __init__(self, engine: "Engine", shape, dtype):
        self.engine = engine
        dtype = np.dtype(dtype)
        nbytes = max(1, int(np.prod(shape)) * dtype.itemsize)
        p = ctypes.c_void_p()
        check(lib().ihe_host_alloc(engine.ctx, nbytes, ctypes.byref(p)))
        self.ptr = p.value
        buf = (ctypes.c_char * nbytes).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)

    def free(self) -> None:
        if self.ptr:
            self.array = None
            check(lib().ihe_host_free(self.engine.ctx, self.ptr))
            self.ptr = 0


class Server:
    """The request loop of ihe_server_* (encrypt -> cloud -> decrypt per request)."""

    def __init__(self, engine: "Engine", hbm_budget: int, num_slots: int = 8,
                 queue_capacity: int = 0):
        self.engine = engine
        cfg = _abi.ServerConfig(hbm_budget, num_slots, queue_capacity)
        h = ctypes.c_void_p()
        check(lib().ihe_server_create(engine.ctx, ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h.value
        self._live = {}
        self._models = {}   # model id -> (tags, outputs)

    def add_model(self, band: Band, num_tags: int) -> int:
        m = ctypes.c_uint32()
        check(lib().ihe_server_add_model(self.handle, band.handle, num_tags, ctypes.byref(m)))
        self._models[m.value] = (int(num_tags), int(band.num_out))
        return m.value

    def check_request(self, model: int, geno: np.ndarray, key: np.ndarray, out: np.ndarray,
                      ns: int) -> None:
        """The buffer sizes a request's copies use (ValueError), checked before
        anything reaches the library."""
        if model not in self._models:
            raise ValueError(f"unknown model {model}")
        tags, outputs = self._models[model]
        if ns < 1:
            raise ValueError("a request needs at least one sample")
        if geno.dtype != np.int8 or not geno.flags.c_contiguous:
            raise ValueError("genotypes must be C-contiguous int8")
        if out.dtype != np.float32 or not out.flags.c_contiguous:
            raise ValueError("scores must be C-contiguous float32")
        if geno.size != tags * ns:
            raise ValueError(f"genotypes hold {geno.size} values, the model needs {tags} x {ns}")
        if out.size != outputs * ns:
            raise ValueError(f"scores hold {out.size} values, the model writes {outputs} x {ns}")
        if key.shape != (N,):
            raise ValueError("key must have N = 1024 coefficients")

    @staticmethod
    def to_chunks(a: np.ndarray) -> np.ndarray:
        """[rows][NS] -> the chunk-major layout of requests (1-D): chunk k holds
        columns 1024k .. 1024k + 1023 as a contiguous [rows][ns_k] block."""
        return np.concatenate([np.ascontiguousarray(a[:, k:k + 1024]).ravel()
                               for k in range(0, a.shape[1], 1024)])

    @staticmethod
    def from_chunks(flat: np.ndarray, rows: int, ns: int) -> np.ndarray:
        """Inverse of to_chunks."""
        out = np.empty((rows, ns), flat.dtype)
        off = 0
        for k in range(0, ns, 1024):
            w = min(1024, ns - k)
            out[:, k:k + w] = flat[off:off + rows * w].reshape(rows, w)
            off += rows * w
        return out

    def submit(self, model: int, geno: np.ndarray, key: np.ndarray, seed: int,
               out: np.ndarray, num_samples: int | None = None,
               arrival_ns: int = 0) -> int:
        """geno int8 and out float32 in the chunk-major request layout (for NS <= 1024
        simply [num_tags][NS] and [num_outputs][NS]; see to_chunks), key [N] int32;
        all three must stay alive (and unmoved) until the request completes.
        arrival_ns: the request's scheduled arrival on time.monotonic_ns()'s clock
        (its latency is counted from there); 0 counts from the submit."""
        ns = int(num_samples if num_samples is not None else geno.shape[-1])
        key = np.ascontiguousarray(key, np.int32)
        self.check_request(model, geno, key, out, ns)
        r = _abi.Request(model, ns, geno.ctypes.data, key.ctypes.data,
                         seed & 0xFFFFFFFFFFFFFFFF, out.ctypes.data, int(arrival_ns))
        t = ctypes.c_uint64()
        check(lib().ihe_server_submit(self.handle, ctypes.byref(r), ctypes.byref(t)))
        self._live[t.value] = (geno, key, out)
        return t.value

    def poll(self, ticket: int):
        """None while running, else the latency in ms."""
        done = ctypes.c_int()
        ms = ctypes.c_double()
        check(lib().ihe_server_poll(self.handle, ticket, ctypes.byref(done), ctypes.byref(ms)))
        if not done.value:
            return None
        self._live.pop(ticket, None)
        return ms.value

    def wait(self, ticket: int) -> float:
        ms = ctypes.c_double()
        try:
            check(lib().ihe_server_wait(self.handle, ticket, ctypes.byref(ms)))
        finally:
            self._live.pop(ticket, None)
        return ms.value

    def stats(self) -> dict:
        st = _abi.ServerStats()
        check(lib().ihe_server_get_stats(self.handle, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in _abi.ServerStats._fields_}

    def close(self) -> None:
        if self.handle:
            check(lib().ihe_server_destroy(self.handle))
            self.handle = 0
            self._live.clear()


class Key:
    def __init__(self, engine: "Engine", key: np.ndarray):
        key = np.ascontiguousarray(key, np.int32)
        if key.shape != (N,):
            raise ValueError("key must have N = 1024 coefficients")
        self.engine = engine
        h = ctypes.c_void_p()
        check(lib().ihe_key_create(engine.ctx, key.ctypes.data, ctypes.byref(h)))
        self.handle = h.value

    def free(self) -> None:
        if self.handle:
            check(lib().ihe_key_destroy(self.handle))
            self.handle = 0


class Engine:
    """One context on one GPU (a HIP stream plus the library's device state)."""

    def __init__(self, device: int = 0):
        n = ctypes.c_int()
        check(lib().ihe_device_count(ctypes.byref(n)))
        if n.value == 0:
            raise RuntimeError("no HIP device visible")
        h = ctypes.c_void_p()
        check(lib().ihe_ctx_create(device, ctypes.byref(h)))
        self.ctx = h.value

    def close(self) -> None:
        if self.ctx:
            check(lib().ihe_ctx_destroy(self.ctx))
            self.ctx = 0

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- memory
    def empty(self, shape, dtype) -> DeviceArray:
        return DeviceArray(self, shape, dtype)

    def to_device(self, host: np.ndarray) -> DeviceArray:
        host = np.ascontiguousarray(host)
        return DeviceArray(self, host.shape, host.dtype).upload(host)

    def sync(self) -> None:
        check(lib().ihe_sync(self.ctx))

    def checksum(self, d: DeviceArray, nbytes: int | None = None) -> int:
        """ihe_checksum of the first nbytes (default: all) of a device array."""
        out = ctypes.c_uint64()
        check(lib().ihe_checksum(self.ctx, d.ptr, d.nbytes if nbytes is None else nbytes,
                                 ctypes.byref(out)))
        return out.value

    def stream_probe(self, d_in: DeviceArray, in_bytes: int, d_out: DeviceArray,
                     out_bytes: int) -> None:
        """ihe_stream_probe: reads in_bytes of d_in and writes out_bytes of d_out
        (HBM bandwidth of a read:write mix); timed under TIMER_STREAM."""
        if in_bytes > d_in.nbytes or out_bytes > d_out.nbytes:
            raise ValueError("probe sizes exceed the buffers")
        check(lib().ihe_stream_probe(self.ctx, d_in.ptr, in_bytes, d_out.ptr, out_bytes))

    def stream_probe_ex(self, d_in: DeviceArray | None, in_bytes: int, d_out: DeviceArray,
                        out_bytes: int, num_waves: int, load_policy: int = 0,
                        store_policy: int = 18) -> None:
        """ihe_stream_probe_ex: the streaming probe with num_waves waves and the
        cache-policy bits (sc0 1, nt 2, sc1 16) of its loads and stores."""
        if d_out.nbytes < out_bytes or (in_bytes and (d_in is None or d_in.nbytes < in_bytes)):
            raise ValueError("probe sizes exceed the buffers")
        check(lib().ihe_stream_probe_ex(self.ctx, d_in.ptr if d_in is not None else None, in_bytes,
                                        d_out.ptr, out_bytes, num_waves, load_policy,
                                        store_policy))

    def copy(self, dst: DeviceArray, src: DeviceArray, nbytes: int | None = None) -> None:
        n = min(dst.nbytes, src.nbytes) if nbytes is None else nbytes
        check(lib().ihe_memcpy_d2d(self.ctx, dst.ptr, src.ptr, n))

    # -- timing
    def timing(self, on: bool) -> None:
        check(lib().ihe_timing_enable(self.ctx, int(on)))

    def timing_reset(self) -> None:
        check(lib().ihe_timing_reset(self.ctx))

    def timing_region(self, begin: bool) -> None:
        """Start or end event of a region of this context's stream (TIMER_REGION)."""
        check(lib().ihe_timing_region(self.ctx, int(begin)))

    def timing_read(self, timer: int) -> tuple[float, int]:
        ms = ctypes.c_double()
        cnt = ctypes.c_uint32()
        check(lib().ihe_timing_read(self.ctx, timer, ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value

    # -- cloud
    def band(self, num_samples: int, num_in: int, row_ptr, in_bidx, weight,
             variant: int = _abi.VARIANT_AUTO) -> Band:
        ns, R, RS = layout_for(num_samples)
        lay = _abi.Layout(ns, R, RS)
        row_ptr = np.ascontiguousarray(row_ptr, np.uint32)
        in_bidx = np.ascontiguousarray(in_bidx, np.uint32)
        weight = np.ascontiguousarray(weight, np.int32)
        m = _abi.ModelCSR(row_ptr.size - 1, row_ptr.ctypes.data, in_bidx.ctypes.data,
                          weight.ctypes.data)
        h = ctypes.c_void_p()
        check(lib().ihe_band_create(self.ctx, ctypes.byref(lay), num_in, ctypes.byref(m), variant,
                                    ctypes.byref(h)))
        return Band(self, h.value, num_in, row_ptr.size - 1, (ns, R, RS))

    def band_tuned(self, num_samples: int, num_in: int, row_ptr, in_bidx, weight,
                   d_in: DeviceArray, d_out: DeviceArray, reps: int = 3) -> tuple[Band, list]:
        """ihe_band_create_tuned: the fastest MFMA tile shape on these buffers, and
        the timings of every candidate as (rows per tile, K-blocks, column tiles
        per wave, ms)."""
        ns, R, RS = layout_for(num_samples)
        lay = _abi.Layout(ns, R, RS)
        row_ptr = np.ascontiguousarray(row_ptr, np.uint32)
        in_bidx = np.ascontiguousarray(in_bidx, np.uint32)
        weight = np.ascontiguousarray(weight, np.int32)
        m = _abi.ModelCSR(row_ptr.size - 1, row_ptr.ctypes.data, in_bidx.ctypes.data,
                          weight.ctypes.data)
        res = (_abi.TuneEntry * 64)()
        n = ctypes.c_uint32(64)
        h = ctypes.c_void_p()
        check(lib().ihe_band_create_tuned(self.ctx, ctypes.byref(lay), num_in, ctypes.byref(m),
                                          d_in.ptr, d_out.ptr, reps, ctypes.byref(h), res,
                                          ctypes.byref(n)))
        table = [(r.rows_per_tile, r.kblocks, r.col_tiles_per_wave, r.ms) for r in res[:n.value]]
        return Band(self, h.value, num_in, row_ptr.size - 1, (ns, R, RS)), table

    def pinned(self, shape, dtype) -> PinnedArray:
        return PinnedArray(self, shape, dtype)

    def server(self, hbm_budget: int, num_slots: int = 8, queue_capacity: int = 0) -> Server:
        return Server(self, hbm_budget, num_slots, queue_capacity)

    def batch(self, cohorts) -> Batch:
        """cohorts: (band, num_samples, num_in, d_in, d_in_var, d_out, d_out_var) tuples."""
        return Batch(self, cohorts)

    def cloud(self, band: Band, d_in: DeviceArray, d_in_var: DeviceArray | None,
              d_out: DeviceArray, d_out_var: DeviceArray | None) -> None:
        """cloud_compute_score (eval/idash.cpp:763-848) on device buffers (async)."""
        check_cloud_buffers(band, band.num_in, d_in, d_in_var, d_out, d_out_var)
        check(lib().ihe_cloud_score(self.ctx, band.handle, d_in.ptr,
                                    d_in_var.ptr if d_in_var is not None else None, d_out.ptr,
                                    d_out_var.ptr if d_out_var is not None else None))

    def cloud_layout(self, band: Band, num_samples: int, num_in: int, d_in: DeviceArray,
                     d_in_var: DeviceArray | None, d_out: DeviceArray,
                     d_out_var: DeviceArray | None) -> None:
        """The band's model on a cohort with another sample count (MFMA bands)."""
        ns, R, RS = layout_for(num_samples)
        lay = _abi.Layout(ns, R, RS)
        check_cloud_buffers(band, num_in, d_in, d_in_var, d_out, d_out_var)
        check(lib().ihe_cloud_score_layout(self.ctx, band.handle, ctypes.byref(lay), num_in,
                                           d_in.ptr, d_in_var.ptr if d_in_var is not None else None,
                                           d_out.ptr,
                                           d_out_var.ptr if d_out_var is not None else None))

    # -- keys, encrypt, decrypt
    def key(self, key: np.ndarray) -> Key:
        return Key(self, key)

    def encrypt(self, key: Key, num_samples: int, d_geno: DeviceArray, seed: int,
                d_out: DeviceArray, d_out_var: DeviceArray) -> int:
        """encrypt_data_ph1 + ph2 (eval/idash.cpp:615-679); returns the ciphertext count."""
        ns, R, RS = layout_for(num_samples)
        tags = d_geno.shape[0]
        if d_geno.shape != (tags, ns) or d_geno.dtype != np.int8:
            raise ValueError("genotypes must be int8 [tags][NUM_SAMPLES]")
        C = (3 * tags + R - 1) // R
        if d_out.nbytes < C * CT_WORDS * 4 or d_out_var.nbytes < C * 8:
            raise ValueError("output buffer too small")
        lay = _abi.Layout(ns, R, RS)
        check(lib().ihe_encrypt(self.ctx, key.handle, ctypes.byref(lay), tags, d_geno.ptr,
                                seed & 0xFFFFFFFFFFFFFFFF, d_out.ptr, d_out_var.ptr))
        return C

    def decrypt(self, key: Key, num_samples: int, d_in: DeviceArray, num_cts: int,
                d_out: DeviceArray) -> None:
        """decrypt_predictions' per-ciphertext phase and rescale (eval/idash.cpp:711-719)."""
        if d_in.nbytes < num_cts * CT_WORDS * 4 or d_out.nbytes < num_cts * num_samples * 4:
            raise ValueError("buffer too small")
        check(lib().ihe_decrypt(self.ctx, key.handle, num_samples, d_in.ptr, num_cts, d_out.ptr))


def band_plan(num_samples: int, num_in: int, row_ptr, in_bidx, weight,
              variant: int = _abi.VARIANT_AUTO) -> dict:
    """Runs the band compiler on the host only; raises IheError like ihe_band_create."""
    ns, R, RS = layout_for(num_samples)
    lay = _abi.Layout(ns, R, RS)
    row_ptr = np.ascontiguousarray(row_ptr, np.uint32)
    in_bidx = np.ascontiguousarray(in_bidx, np.uint32)
    weight = np.ascontiguousarray(weight, np.int32)
    m = _abi.ModelCSR(row_ptr.size - 1, row_ptr.ctypes.data, in_bidx.ctypes.data,
                      weight.ctypes.data)
    bi = _abi.BandInfo()
    check(lib().ihe_band_plan(ctypes.byref(lay), num_in, ctypes.byref(m), variant,
                              ctypes.byref(bi)))
    return {f: getattr(bi, f) for f, _ in _abi.BandInfo._fields_}


def noise_table() -> np.ndarray:
    t = np.empty(_abi.IHE_NOISE_TABLE_LEN, np.uint64)
    check(lib().ihe_noise_table(t.ctypes.data, t.size))
    return t
