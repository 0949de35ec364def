"""Throughput of the cloud step (cloud_compute_score, eval/idash.cpp:763-848)
on one MI355X.

A step is one cloud pass over the whole synthetic workload with its input
ciphertexts already resident in HBM: 16,184 tag SNPs, 80,882 target SNPs,
1,024 samples, 5 neighbours (48,552 input and 242,646 output ciphertexts).
The inputs are produced on the device by the library's own encrypt kernel.

Prints one JSON line. `value` is imputed target SNPs per second
(targets x samples / step time). A step is one launch: the cloud kernel with
its variance blocks in the same grid. `roofline` prices that launch against
HBM bandwidth with its algorithmic bytes (every input ciphertext read once,
every output written once, plus the band); its duration is the time between
two HIP events on the launch stream around the timed steps, divided by the
steps, and `roofline.kernel_only` gives the same from a separate pass with
an event pair around each launch. `cpu_baseline` times the CPU restatement of
the reference's cloud step (hash-map model, OpenMP over outputs) on this
host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from idash2019_2_amd import Engine, Server, _abi, synth  # noqa: E402

METRIC = ("encrypted target SNPs imputed/s at 1 GPU; % HBM roofline; "
          "p50/p99 cohort latency")
HBM_PEAK_GBS = 8000.0
CPU_SHARE = 16          # host cores that go with one GPU of the box
VARIANTS = {"auto": _abi.VARIANT_AUTO, "valu": _abi.VARIANT_VALU, "wide": _abi.VARIANT_WIDE,
            "mfma": _abi.VARIANT_MFMA}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def source_hash() -> str:
    """SHA-256 (first 16 hex digits) of the GPU library's sources: the kernels,
    their headers, the build flags and the C ABI header."""
    import hashlib
    h = hashlib.sha256()
    csrc = os.path.join(ROOT, "idash2019_2_amd", "csrc")
    files = sorted(f for f in os.listdir(csrc) if f.endswith((".hip", ".h")) or f == "Makefile")
    for f in files + ["../../include/idash_hip.h"]:
        h.update(f.encode())
        with open(os.path.join(csrc, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def read_traffic(workload: str, variant: str):
    """HBM bytes per launch measured with rocprofv3 counters, from the committed
    summary, when it was taken on the library sources of this run (its
    source_hash line); None otherwise."""
    path = os.path.join(ROOT, "profiles", f"cloud_{variant}_{workload}_traffic.txt")
    if not os.path.exists(path):
        return None
    fields = {}
    for line in open(path):
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            fields[k.strip()] = v.strip()
    if fields.get("source_hash") != source_hash() or "traffic_bytes_per_launch" not in fields:
        return None
    return float(fields["traffic_bytes_per_launch"])


def build_bands(eng, models, keys, num_in, shapes=None):
    """Bands of several model sets; shapes {n: (rb, nkb, ntw)} pins the tile shape
    (IHE_CLOUD_SHAPE / IHE_CLOUD_NTW are read when a band is compiled)."""
    bands = {}
    for k in keys:
        env = {}
        if shapes and k[1] in shapes:
            rb, nkb, ntw = shapes[k[1]]
            env = {"IHE_CLOUD_SHAPE": f"{rb},{nkb}", "IHE_CLOUD_NTW": str(ntw)}
        os.environ.update(env)
        try:
            bands[k] = eng.band(1024, num_in, *models[k].model_csr())
        finally:
            for kk in env:
                os.environ.pop(kk, None)
    return bands


def forms_batched(eng, base, log_fn, seed=4):
    """64 cohorts, NS ~ U[16, 1024], 3 model sets x neighbours {5, 10, 20, 50} on the
    full SNP layout: 64 per-cohort launches vs one launch per kernel family vs the one
    batch launch (k_cloud_batch) vs its hipGraph replay; every cohort's output checksum
    must agree."""
    rng = np.random.default_rng(seed)
    models = synth.model_sets(base, (5, 10, 20, 50), 3, seed)
    specs = [(int(rng.integers(16, 1025)), int(rng.integers(0, 3)), int(rng.choice([5, 10, 20, 50])))
             for _ in range(64)]
    # autotune one band per neighbour count on a full-size cohort, pin that shape for the others
    key = eng.key(base.key)
    d_geno = eng.to_device(base.geno)
    d_ct = eng.empty((base.num_in_cts, 2, 1024), np.int32)
    d_var = eng.empty((base.num_in_cts,), np.float64)
    eng.encrypt(key, 1024, d_geno, 7, d_ct, d_var)
    d_scratch = eng.empty((base.num_outputs, 2, 1024), np.int32)
    shapes, tune_tables = {}, {}
    for n in (5, 10, 20, 50):
        b, table = eng.band_tuned(1024, base.num_in_cts, *models[(0, n)].model_csr(), d_ct,
                                  d_scratch, reps=2)
        # the fastest tiling among those the persistent batch kernel runs
        # (one K-block over 32 or 64 rows, 128 rows over at most 6 K-blocks)
        runs = [t for t in table if (t[0] in (32, 64) and t[1] == 1) or (t[0] == 128 and t[1] <= 6)]
        best = min(runs or table, key=lambda t: t[3])
        shapes[n] = (best[0] // 32, best[1], best[2])
        tune_tables[n] = table
        b.free()
    d_scratch.free()
    d_ct.free()
    d_var.free()
    d_geno.free()
    log_fn(f"autotuned shapes (rows/32, K-blocks, column tiles per wave): {shapes}")
    bands = build_bands(eng, models, sorted({(s, n) for _, s, n in specs}), base.num_in_cts, shapes)
    cohorts, work = [], 0
    for i, (ns, sset, n) in enumerate(specs):
        cols = rng.choice(1024, ns, replace=False)
        R = 1024 // ns
        num_in = (3 * base.num_tags + R - 1) // R
        d_g = eng.to_device(np.ascontiguousarray(base.geno[:, cols]))
        d_in = eng.empty((num_in, 2, 1024), np.int32)
        d_iv = eng.empty((num_in,), np.float64)
        eng.encrypt(key, ns, d_g, 1000 + i, d_in, d_iv)
        d_g.free()
        d_out = eng.empty((base.num_outputs, 2, 1024), np.int32)
        d_ov = eng.empty((base.num_outputs,), np.float64)
        cohorts.append((bands[(sset, n)], ns, num_in, d_in, d_iv, d_out, d_ov))
        work += base.num_targets * ns
    eng.sync()

    def sums():
        return [(eng.checksum(c[5]), eng.checksum(c[6])) for c in cohorts]

    def sequential():
        for band, ns, num_in, d_in, d_iv, d_out, d_ov in cohorts:
            eng.cloud_layout(band, ns, num_in, d_in, d_iv, d_out, d_ov)

    def make_batch(mode):
        os.environ["IHE_BATCH_PERSISTENT"] = mode
        try:
            return eng.batch(cohorts)
        finally:
            os.environ.pop("IHE_BATCH_PERSISTENT", None)

    fam = make_batch("0")                                # one launch per kernel family
    fam_launches = fam.info()["num_launches"]
    fam_graph = make_batch("0")                          # the same, replayed from a hipGraph
    fam_graph.capture()
    batch = make_batch("1")                              # one persistent launch
    info = batch.info()
    launches, persistent = info["num_launches"], bool(info["persistent"])
    graph = make_batch("1")                              # the same, replayed from a hipGraph
    graph.capture()
    default = eng.batch(cohorts)                         # the library's default form
    default_form = "one_launch" if default.info()["persistent"] else "per_family"
    default.free()
    # each form's outputs against the per-cohort launches', then the forms
    # timed in turn, round after round (so that no form always runs on a
    # warmer chip), median of each
    sequential()
    eng.sync()
    ref = sums()
    same = {}
    for name, fn in (("family", fam.run), ("family_graph", fam_graph.run), ("batch", batch.run),
                     ("graph", graph.run)):
        for c in cohorts:
            c[5].fill_zero()
            c[6].fill_zero()
        fn()
        eng.sync()
        same[name] = sums() == ref
    times = {"seq": [], "family": [], "family_graph": [], "batch": [], "graph": []}
    for _ in range(5):
        for name, fn in (("seq", sequential), ("family", fam.run), ("family_graph", fam_graph.run),
                         ("batch", batch.run), ("graph", graph.run)):
            eng.sync()
            t0 = time.perf_counter()
            fn()
            eng.sync()
            times[name].append(time.perf_counter() - t0)
    t_seq, t_fam, t_fam_graph, t_batch, t_graph = (
        float(np.median(times[k])) for k in ("seq", "family", "family_graph", "batch", "graph"))
    for b in (fam, fam_graph, batch, graph):
        b.free()
    for c in cohorts:
        for d in c[3:]:
            d.free()
    for b in bands.values():
        b.free()
    key.free()
    if not all(same.values()):
        raise SystemExit("batched outputs differ from the per-cohort launches")
    out = {"cohorts": 64, "imputed_target_snps_per_step": work, "kernel_launches": launches,
           "persistent": persistent, "default_form": default_form,
           "sequential_ms": round(t_seq * 1e3, 3),
           "per_family_ms": round(t_fam * 1e3, 3), "per_family_launches": fam_launches,
           "per_family_graph_ms": round(t_fam_graph * 1e3, 3),
           "batched_ms": round(t_batch * 1e3, 3), "graph_ms": round(t_graph * 1e3, 3),
           "sequential_value": round(work / t_seq, 1), "batched_value": round(work / t_batch, 1),
           "graph_value": round(work / t_graph, 1), "bit_exact": True,
           "tile_shapes": {str(n): list(v) for n, v in shapes.items()}}
    log_fn(f"batched form: 64 cohorts, sequential {t_seq * 1e3:.2f} ms, per family {t_fam * 1e3:.2f} "
           f"ms ({fam_launches} launches; graph {t_fam_graph * 1e3:.2f} ms), one kernel "
           f"{t_batch * 1e3:.2f} ms ({launches} launch{'es' if launches != 1 else ''}"
           f"{', persistent' if persistent else ''}; graph {t_graph * 1e3:.2f} ms), default form "
           f"{default_form}; outputs identical")
    return out


def forms_served(eng, base, log_fn, seed=5, num_requests=64, budget=64 << 30, slots=8):
    """Poisson arrivals of cohort requests, NS log-uniform in [1, 4096], each on one
    of 3 resident 5-neighbour model sets; encrypt -> cloud -> decrypt on the device.
    The served loop (slots in flight, admission by HBM budget) against the plain
    per-request path (one request at a time) on the same arrival times; every
    request's scores must agree."""
    rng = np.random.default_rng(seed)
    models = synth.model_sets(base, (5,), 3, seed)
    bands = build_bands(eng, models, sorted(models), base.num_in_cts)
    ns_list = np.exp(rng.uniform(0, np.log(4096), num_requests)).astype(int).clip(1, 4096)
    reqs = []
    for i, ns in enumerate(ns_list.tolist()):
        cols = rng.integers(0, base.num_samples, ns)
        geno = Server.to_chunks(base.geno[:, cols]) if ns > 1024 else \
            np.ascontiguousarray(base.geno[:, cols])
        geno_ns = ns
        key = rng.integers(0, 2, 1024).astype(np.int32)
        reqs.append((int(rng.integers(0, 3)), geno, key, 10_000 + 7 * i, geno_ns))
    work = sum(base.num_targets * int(ns) for ns in ns_list)
    # page-locked score buffers (allocated once, reused by every run) so that the
    # device-to-host copies of the scores run asynchronously
    t_pin = time.perf_counter()
    pinned = [eng.pinned((base.num_outputs * int(ns),), np.float32) for ns in ns_list]
    log_fn(f"served form: {sum(p.array.nbytes for p in pinned) / 2**30:.1f} GiB of pinned score "
           f"buffers in {time.perf_counter() - t_pin:.1f} s")

    def digest(a):
        """Sum and xor of the 64-bit words (no copies of multi-GB score arrays)."""
        raw = a.reshape(-1).view(np.uint8)
        head = raw[:raw.size // 8 * 8].view(np.uint64)
        tail = bytes(raw[raw.size // 8 * 8:])
        return (int(head.sum(dtype=np.uint64)), int(np.bitwise_xor.reduce(head)) if head.size else 0,
                tail, a.shape)

    def run(nslots, arrivals):
        """Submits every request at its scheduled arrival (back to back when
        arrivals is None) and waits for all; each latency runs from the request's
        scheduled arrival (its submit, back to back) to its scores on the host, as
        the server clocks it. The scores are digested after the timed run."""
        srv = eng.server(budget, num_slots=nslots)
        mids = [srv.add_model(bands[(s, 5)], base.num_tags) for s in range(3)]
        tickets = []
        t0 = time.monotonic_ns()
        for i, (m, geno, key, seed_i, ns_i) in enumerate(reqs):
            due = 0
            if arrivals is not None:
                due = t0 + int(arrivals[i] * 1e9)
                while time.monotonic_ns() < due:
                    time.sleep(0.0002)
            tickets.append(srv.submit(mids[m], geno, key, seed_i, pinned[i].array,
                                      num_samples=ns_i, arrival_ns=due))
        lat = []
        for k, t in enumerate(tickets):
            lat.append(srv.wait(t))
            if (k + 1) % 16 == 0:
                log_fn(f"served: {k + 1} of {len(reqs)} requests done")
        elapsed = (time.monotonic_ns() - t0) / 1e9
        peak = srv.stats()["arena_peak"]
        srv.close()
        dig = [digest(p.array) for p in pinned]
        return elapsed, lat, dig, peak

    log_fn(f"served form: {len(reqs)} requests, {sum(int(x) for x in ns_list)} samples in all")
    run(slots, None)                                       # warm-up: first DMA to the buffers
    t_srv_bb, _, dig_bb, _ = run(slots, None)              # back to back: service capacity
    log_fn(f"served back to back: {t_srv_bb:.2f} s")
    t_plain_bb, _, dig_ref, _ = run(1, None)
    log_fn(f"plain path back to back: {t_plain_bb:.2f} s")
    rate = 0.8 * len(reqs) / t_plain_bb                     # offered load 0.8 of the plain path
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(reqs)))
    arrivals -= arrivals[0]
    t_srv, lat_srv, dig_srv, peak = run(slots, arrivals)
    t_pl, lat_pl, dig_pl, _ = run(1, arrivals)
    for b in bands.values():
        b.free()
    for pa in pinned:
        pa.free()
    if dig_srv != dig_ref or dig_pl != dig_ref or dig_bb != dig_ref:
        raise SystemExit("served scores differ from the per-request path")
    p = lambda v, q: round(float(np.percentile(v, q)), 2)
    out = {"requests": len(reqs), "arrival_rate_per_s": round(rate, 2), "slots": slots,
           "hbm_budget_gib": budget >> 30, "arena_peak_gib": round(peak / (1 << 30), 2),
           "served_value": round(work / t_srv, 1), "served_p50_ms": p(lat_srv, 50),
           "served_p99_ms": p(lat_srv, 99), "plain_value": round(work / t_pl, 1),
           "plain_p50_ms": p(lat_pl, 50), "plain_p99_ms": p(lat_pl, 99),
           "served_capacity_value": round(work / t_srv_bb, 1),
           "plain_capacity_value": round(work / t_plain_bb, 1),
           "imputed_target_snps": work, "bit_exact": True}
    log_fn(f"served form: {len(reqs)} requests at {rate:.1f}/s: served p50 {out['served_p50_ms']} ms "
           f"p99 {out['served_p99_ms']} ms, plain p50 {out['plain_p50_ms']} ms p99 "
           f"{out['plain_p99_ms']} ms; scores identical")
    return out


def cli_pipeline(workload, log_fn, workdir=None):
    """The idash-compatible executables on the workload's files in the reference's
    formats: keygen, encrypt, cloud (twice: parsing the .hr files, then from the
    model cache), decrypt; returns their own BENCHMARK figures."""
    import re
    import shutil
    import subprocess
    import tempfile
    host = os.path.join(ROOT, "idash2019_2_amd", "host")
    d = workdir or tempfile.mkdtemp(prefix="ihe_cli_")
    t0 = time.time()
    c = synth.workload(workload)
    synth.write_reference_files(c, d)
    log_fn(f"cli: wrote {workload} input files ({3 * c.num_targets} .hr files) in "
           f"{time.time() - t0:.1f} s")
    figures = {}

    def run(name, *args):
        t = time.time()
        r = subprocess.run([os.path.join(host, name), *args], cwd=d, capture_output=True, text=True,
                           env=dict(os.environ, IHE_SEED="1"))
        if r.returncode:
            raise SystemExit(f"{name} failed: {r.stdout[-2000:]}{r.stderr[-2000:]}")
        vals = {k.strip(" ."): float(v) for k, v in re.findall(r"^(.*?):\s*([0-9.e+-]+)\s*$",
                                                                   r.stdout, re.M)}
        vals["process wall time (s)"] = round(time.time() - t, 3)
        for ln in r.stderr.splitlines():  # IHE_HOST_PHASES=1: the host functions' phases
            if " phase " in ln:
                log_fn(f"cli: {name}: {ln.strip()}")
        log_fn(f"cli: {name} {' '.join(args)}: {vals}")
        return vals

    try:
        figures["keygen"] = run("keygen", "targets.txt", "tags.txt", "1")
        figures["encrypt"] = run("encrypt", "tags.txt")
        figures["cloud_parse"] = run("cloud", "model")
        figures["cloud_cached"] = run("cloud", "model")
        figures["decrypt"] = run("decrypt")
        sizes = {f: os.path.getsize(os.path.join(d, f)) for f in
                 ("encrypted_data.bin", "encrypted_prediction.bin", "result.csv")}
    finally:
        if workdir is None:
            shutil.rmtree(d, ignore_errors=True)
    return {"workload": workload, "figures": figures, "file_bytes": sizes}


def staging_ab(eng, layouts=(1024, 512, 200, 100), rounds=3, launches=20) -> str:
    """The 128-row cloud kernel's staging forms side by side on the full SNP
    layout (16,184 tags, 80,882 targets): input words staged in registers
    (IHE_MFMAW_LDS=0), staged in LDS by dword loads (1) and by 16-byte loads
    where a column's words come in runs of four (the default), on 128 x 2, 4 and
    6 K-block tiles at 10, 20 and 50 neighbours, for several sample counts.
    Forms alternate launch by launch group, `rounds` times; each figure is the
    best round's mean over `launches` launches (HIP events), and every form's
    outputs must agree."""
    base = synth.workload("full")
    models = synth.model_sets(base, (10, 20, 50), 1, 4)
    forms = (("0", "registers"), ("1", "LDS, dword"), ("2", "LDS, 16-byte"))
    lines = ["# The 128-row cloud kernel's staging forms (IHE_MFMAW_LDS), full SNP layout,",
             f"# best of {rounds} interleaved rounds of {launches} launches; outputs identical",
             "#", "# Command (from the repository root):", "#   python3 bench.py --staging-ab",
             f"# source_hash of the library sources measured: {source_hash()}", "#",
             f"{'NS':>5} {'n':>3} {'tile':>8}  " + "  ".join(f"{name:>13}" for _, name in forms)]
    key = eng.key(base.key)
    for ns in layouts:
        R = 1024 // ns
        num_in = (3 * base.num_tags + R - 1) // R
        cols = np.random.default_rng(ns).choice(1024, ns, replace=False)
        d_g = eng.to_device(np.ascontiguousarray(base.geno[:, cols]))
        d_in, d_iv = eng.empty((num_in, 2, 1024), np.int32), eng.empty((num_in,), np.float64)
        eng.encrypt(key, ns, d_g, 11, d_in, d_iv)
        d_g.free()
        d_out = eng.empty((base.num_outputs, 2, 1024), np.int32)
        d_ov = eng.empty((base.num_outputs,), np.float64)
        for n, shape in ((10, "4,2"), (20, "4,4"), (50, "4,6")):
            os.environ["IHE_CLOUD_SHAPE"] = shape
            try:
                band = eng.band(1024, base.num_in_cts, *models[(0, n)].model_csr())
            finally:
                os.environ.pop("IHE_CLOUD_SHAPE", None)
            best, sums = {}, {}
            try:
                for _rnd in range(rounds):
                    for mode, _name in forms:
                        os.environ["IHE_MFMAW_LDS"] = mode
                        for _ in range(3):
                            eng.cloud_layout(band, ns, num_in, d_in, d_iv, d_out, d_ov)
                        eng.sync()
                        eng.timing(True)
                        eng.timing_reset()
                        for _ in range(launches):
                            eng.cloud_layout(band, ns, num_in, d_in, d_iv, d_out, d_ov)
                        ms, cnt = eng.timing_read(_abi.TIMER_CLOUD)
                        eng.timing(False)
                        best[mode] = min(best.get(mode, float("inf")), ms / cnt)
                        sums[mode] = (eng.checksum(d_out), eng.checksum(d_ov))
            finally:
                os.environ.pop("IHE_MFMAW_LDS", None)
            band.free()
            if len(set(sums.values())) != 1:
                raise SystemExit(f"staging forms disagree at NS={ns}, n={n}")
            lines.append(f"{ns:>5} {n:>3} {'128 x ' + shape[2]:>8}  " +
                         "  ".join(f"{best[m]:>10.4f} ms" for m, _ in forms))
        for d in (d_in, d_iv, d_out, d_ov):
            d.free()
    key.free()
    return "\n".join(lines) + "\n"


def hbm_probes(eng) -> str:
    """What HBM sustains on this device for the streaming probe's access
    patterns (contiguous 16-byte loads and stores, 1 KiB per wave instruction
    pair) and under each cache policy pair of its loads and stores: the median
    of 20 launches (HIP events on the launch stream) after 3 warm-ups."""
    tags, targets = synth.WORKLOADS["full"][:2]
    cin = 3 * tags * 8192 // 1024 * 1024          # the cloud step's reads
    cout = 3 * targets * 8192 // 65536 * 65536    # and writes
    big = ((1 << 31) - (1 << 20)) // 65536 * 65536
    gib = 1 << 30
    a, b = eng.empty((big,), np.uint8), eng.empty((big,), np.uint8)
    a.fill_zero()

    def median_ms(ib, ob, waves, lp, sp):
        for _ in range(3):
            eng.stream_probe_ex(a, ib, b, ob, waves, lp, sp)
        eng.sync()
        eng.timing(True)
        times = []
        for _ in range(20):
            eng.timing_reset()
            eng.stream_probe_ex(a, ib, b, ob, waves, lp, sp)
            times.append(eng.timing_read(_abi.TIMER_STREAM)[0])
        eng.timing(False)
        return float(np.median(times))

    lines = ["pattern                                   read MB  write MB   loads/stores   ms       TB/s"]
    patterns = [("contiguous 16-byte reads", big, 8192 * 1024, 8192),
                ("contiguous 16-byte writes", 0, big, big // 65536),
                ("copy (1:1 read:write)", gib, gib, gib // 65536),
                ("1:5 read:write, the cloud step's volumes", cin, cout, cout // 65536)]
    for name, ib, ob, waves in patterns:
        for lp, sp in ((0, 0), (0, 18)):
            ms = median_ms(ib, ob, waves, lp, sp)
            lines.append(f"{name:42s}{ib / 1e6:8.1f}{ob / 1e6:10.1f}   {lp:>5}/{sp:<6}  "
                         f"{ms:8.4f}  {(ib + ob) / ms / 1e9:6.2f}")
    lines.append("")
    lines.append("cache policy bits on the cloud step's volumes (sc0 = 1, nt = 2, sc1 = 16)")
    lines.append("loads/stores   ms")
    for lp, sp in ((0, 0), (0, 2), (0, 1), (0, 16), (0, 18), (2, 2), (2, 18)):
        lines.append(f"{lp:>5}/{sp:<6}  {median_ms(cin, cout, cout // 65536, lp, sp):8.4f}")
    a.free()
    b.free()
    return "\n".join(lines) + "\n"


def _csv_rows(directory: str, suffix: str):
    """Rows of the rocprofv3 CSV files under `directory` whose names end in `suffix`."""
    import csv
    import glob
    rows = []
    for path in sorted(glob.glob(os.path.join(directory, "**", f"*{suffix}"), recursive=True)):
        with open(path, newline="") as f:
            rows += list(csv.DictReader(f))
    return rows


def _counter_per_dispatch(directory: str, counter: str, kernel: str):
    """{kernel name: [value per dispatch]} of one counter (KiB for FETCH_SIZE and
    WRITE_SIZE), for the kernels whose name contains `kernel`."""
    per = {}
    for r in _csv_rows(directory, "counter_collection.csv"):
        if r.get("Counter_Name") != counter or kernel not in r.get("Kernel_Name", ""):
            continue
        key = (r["Kernel_Name"], r.get("Dispatch_Id", len(per)))
        per[key] = per.get(key, 0.0) + float(r["Counter_Value"])
    out = {}
    for (name, _), v in sorted(per.items(), key=lambda kv: int(kv[0][1]) if str(kv[0][1]).isdigit()
                               else 0):
        out.setdefault(name, []).append(v)
    return out


def profile_summary(args) -> str:
    """Text summaries of rocprofv3 runs for profiles/: `trace` (per-kernel call
    counts and durations from --kernel-trace --stats) or `traffic` (HBM bytes per
    launch of one kernel from FETCH_SIZE and WRITE_SIZE runs, each scaled by what
    the counter reads for the 1 GiB device copies of --calibrate-copy)."""
    lines = [f"# {args.title}", "#", "# Commands (from /tmp with TMPDIR=/tmp; REPO = the repository root):"]
    lines += [f"#   {c}" for c in args.cmd]
    lines.append(f"# source_hash of the library sources profiled: {source_hash()}")
    if args.profile_summary == "counters":
        # every counter of the listed runs, averaged over the kernel's dispatches
        vals = {}
        for d in args.dirs:
            for r in _csv_rows(d, "counter_collection.csv"):
                if args.kernel not in r.get("Kernel_Name", ""):
                    continue
                key = (r["Kernel_Name"], r["Counter_Name"])
                vals.setdefault(key, {}).setdefault(r.get("Dispatch_Id"), 0.0)
                vals[key][r.get("Dispatch_Id")] += float(r["Counter_Value"])
        lines += ["#", f"{'kernel':<56} {'counter':<30} {'dispatches':>10} {'per dispatch':>16}"]
        for (k, c), per in sorted(vals.items()):
            name = k.replace("ihe::", "").replace("(anonymous namespace)::", "")
            lines.append(f"{name[:56]:<56} {c:<30} {len(per):>10} {sum(per.values()) / len(per):>16.1f}")
        return "\n".join(lines) + "\n"
    if args.profile_summary == "trace":
        rows = _csv_rows(args.trace, "kernel_stats.csv")
        total = sum(float(r["TotalDurationNs"]) for r in rows) or 1.0
        lines += ["#", f"{'kernel':<72} {'calls':>6} {'avg_us':>9} {'min_us':>9} {'max_us':>9} {'pct':>6}"]
        for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
            name = r["Name"].replace("ihe::", "").replace("(anonymous namespace)::", "")
            lines.append(f"{name[:72]:<72} {int(r['Calls']):>6} {float(r['AverageNs']) / 1e3:>9.1f} "
                         f"{float(r['MinNs']) / 1e3:>9.1f} {float(r['MaxNs']) / 1e3:>9.1f} "
                         f"{100 * float(r['TotalDurationNs']) / total:>6.2f}")
        return "\n".join(lines) + "\n"
    # traffic
    copy_r = _counter_per_dispatch(args.copy_fetch, "FETCH_SIZE", "copy")
    copy_w = _counter_per_dispatch(args.copy_write, "WRITE_SIZE", "copy")
    (cname_r, cr), = copy_r.items()
    (cname_w, cw), = copy_w.items()
    gib_kib = float(1 << 20)                  # one 1 GiB copy reads and writes 2^20 KiB
    scale_r = round(gib_kib / (sum(cr) / len(cr)), 3)
    scale_w = round(gib_kib / (sum(cw) / len(cw)), 3)
    kr = _counter_per_dispatch(args.fetch, "FETCH_SIZE", args.kernel)
    kw = _counter_per_dispatch(args.write, "WRITE_SIZE", args.kernel)
    (kname, fr), = kr.items()
    (_, fw), = kw.items()
    read_b = sum(fr) / len(fr) * 1024 * scale_r
    write_b = sum(fw) / len(fw) * 1024 * scale_w
    alg = int(args.algorithmic_bytes)
    lines += ["#",
              f"# calibration: {cname_r.strip()} FETCH_SIZE per dispatch (KiB): "
              + ", ".join(f"{v:.1f}" for v in cr) + f" -> read scale {scale_r}",
              f"# calibration: {cname_w.strip()} WRITE_SIZE per dispatch (KiB): "
              + ", ".join(f"{v:.1f}" for v in cw) + f" -> write scale {scale_w}",
              f"# kernel: {kname.strip()}",
              "# FETCH_SIZE per dispatch (KiB): " + ", ".join(f"{v:.2f}" for v in fr),
              "# WRITE_SIZE per dispatch (KiB): " + ", ".join(f"{v:.2f}" for v in fw),
              f"# HBM bytes per launch: read {read_b:.0f}, write {write_b:.0f}",
              f"# algorithmic bytes per launch: {alg}",
              f"# traffic / algorithmic = {(read_b + write_b) / alg:.4f}",
              f"source_hash: {source_hash()}",
              f"traffic_bytes_per_launch: {read_b + write_b:.0f}"]
    return "\n".join(lines) + "\n"


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(eng, d_ct, c, rp, ib, w, d_out, threads):
    """The oracle's restatement of the reference's cloud loop, timed at `threads`
    and at the README's 4 threads; the GPU rows must equal the CPU rows."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    ct = d_ct.numpy()
    var = np.full(c.num_in_cts, 2.0 ** -50)
    out, _, secs = oracle.cpu_baseline_cloud(ct, var, rp, ib, w, *c.layout, threads)
    gpu = d_out.numpy()
    same = bool((gpu == out).all())
    del out
    secs4 = None
    if threads != 4:
        _, _, secs4 = oracle.cpu_baseline_cloud(ct, var, rp, ib, w, *c.layout, 4)
    return secs, same, secs4


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="full", choices=sorted(synth.WORKLOADS))
    ap.add_argument("--variant", default="auto", choices=sorted(VARIANTS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--forms", action="store_true",
                    help="measure the batched form (64 cohorts) and the served form (64 Poisson "
                         "requests, p50/p99 latency) for any workload (the default does so "
                         "for the full workload)")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the batched and served forms")
    ap.add_argument("--cli", action="store_true",
                    help="only run the idash-compatible executables end to end on the workload's "
                         "files and print their timings as JSON")
    ap.add_argument("--probes", action="store_true",
                    help="HBM access patterns and cache policies of the streaming probe (a text "
                         "table on stdout)")
    ap.add_argument("--staging-ab", action="store_true",
                    help="the 128-row cloud kernel's staging forms side by side on the full "
                         "layout (a text table on stdout)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run three 1 GiB device-to-device copies (a kernel of known "
                         "traffic for scaling HBM counters) and exit")
    ap.add_argument("--profile-summary", choices=["trace", "traffic", "counters"],
                    help="only print a profiles/ summary of rocprofv3 CSV output (no GPU use)")
    ap.add_argument("--title", default="rocprofv3 summary")
    ap.add_argument("--cmd", action="append", default=[], help="command that made the data")
    ap.add_argument("--trace", help="directory of a --kernel-trace --stats run")
    ap.add_argument("--dirs", nargs="*", default=[], help="directories of --pmc runs (counters)")
    ap.add_argument("--fetch", help="directory of the FETCH_SIZE run")
    ap.add_argument("--write", help="directory of the WRITE_SIZE run")
    ap.add_argument("--copy-fetch", help="directory of the FETCH_SIZE run of --calibrate-copy")
    ap.add_argument("--copy-write", help="directory of the WRITE_SIZE run of --calibrate-copy")
    ap.add_argument("--kernel", default="k_cloud", help="kernel name part to sum counters of")
    ap.add_argument("--algorithmic-bytes", type=int, default=0)
    args = ap.parse_args(argv)

    if args.profile_summary:
        sys.stdout.write(profile_summary(args))
        return 0

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the cloud step "
              "shards by target range with no exchange, and multi-GPU runs are not part of "
              "this build", file=sys.stderr)
        return 2

    if args.cli:
        print(json.dumps(cli_pipeline(args.workload, log)), flush=True)
        return 0

    if args.probes:
        with Engine(0) as eng:
            sys.stdout.write(hbm_probes(eng))
        return 0

    if args.staging_ab:
        with Engine(0) as eng:
            sys.stdout.write(staging_ab(eng))
        return 0

    if args.calibrate_copy:
        with Engine(0) as eng:
            n = 1 << 30
            a, b = eng.empty((n,), np.uint8), eng.empty((n,), np.uint8)
            a.fill_zero()
            for _ in range(3):
                eng.copy(b, a)
            eng.sync()
        log("copied 3 x 1 GiB device to device")
        return 0

    t0 = time.time()
    c = synth.workload(args.workload)
    ns, R, RS = c.layout
    log(f"workload {args.workload}: {c.num_tags} tags, {c.num_targets} targets, NS={ns}, "
        f"n={c.neighbours}, {c.num_in_cts} input / {c.num_outputs} output ciphertexts "
        f"({time.time() - t0:.1f}s)")

    with Engine(0) as eng:
        key = eng.key(c.key)
        d_geno = eng.to_device(c.geno)
        C = c.num_in_cts
        d_ct = eng.empty((C, 2, 1024), np.int32)
        d_var = eng.empty((C,), np.float64)
        eng.encrypt(key, ns, d_geno, 0x5EED, d_ct, d_var)
        rp, ib, w = c.model_csr()
        band = eng.band(ns, C, rp, ib, w, VARIANTS[args.variant])
        info = band.info()
        variant_name = {v: k for k, v in VARIANTS.items()}[info["variant"]]
        d_out = eng.empty((c.num_outputs, 2, 1024), np.int32)
        d_ov = eng.empty((c.num_outputs,), np.float64)
        eng.sync()
        log(f"band: {info['num_groups']} groups, {info['num_tiles']} tiles of {info['tile_rows']} "
            f"rows x {info['tile_kblocks']} K-blocks ({info['col_tiles']} column tiles per wave or "
            f"workgroup), variant {variant_name}")

        for _ in range(args.warmup):
            eng.cloud(band, d_ct, d_var, d_out, d_ov)
        eng.sync()
        eng.timing(False)
        eng.timing_reset()
        # the timed steps: synchronized on both sides, bracketed on the launch
        # stream by two events and nothing between the launches
        t_start = time.perf_counter()
        eng.timing_region(True)
        for _ in range(args.steps):
            eng.cloud(band, d_ct, d_var, d_out, d_ov)
        eng.timing_region(False)
        eng.sync()
        elapsed = time.perf_counter() - t_start
        region_ms, _ = eng.timing_read(_abi.TIMER_REGION)
        # a separate pass of the same steps with an event pair around each launch
        eng.timing(True)
        eng.timing_reset()
        for _ in range(args.steps):
            eng.cloud(band, d_ct, d_var, d_out, d_ov)
        kern_ms, kern_n = eng.timing_read(_abi.TIMER_CLOUD)
        var_ms, var_n = eng.timing_read(_abi.TIMER_CLOUD_VAR)
        eng.timing(False)
        launches = (kern_n + var_n) / args.steps
        log(f"{args.steps} steps in {elapsed * 1e3:.2f} ms wall, {region_ms:.3f} ms between the "
            f"region's events; {launches:g} launch(es) per step; cloud launch avg "
            f"{kern_ms / kern_n:.4f} ms" + (f", variance launch avg {var_ms / var_n:.4f} ms"
                                           if var_n else " (variances in the same launch)"))

        step_s = elapsed / args.steps
        value = c.num_targets * ns / step_s
        alg_bytes = synth.cloud_bytes(c)
        step_event_s = region_ms / args.steps / 1e3
        avg_kernel_s = kern_ms / kern_n / 1e3
        achieved = alg_bytes / step_event_s / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": read_traffic(args.workload, variant_name),
                    "duration": "average step between two HIP events on the launch stream around "
                                "the timed steps (one launch per step: the cloud kernel with its "
                                "variance blocks)" if launches == 1 else
                                "average step between two HIP events on the launch stream around "
                                "the timed steps",
                    "launches_per_step": launches,
                    "kernel_only": {"ms": round(avg_kernel_s * 1e3, 4),
                                    "achieved": round(alg_bytes / avg_kernel_s / 1e9, 1),
                                    "frac": round(alg_bytes / avg_kernel_s / 1e9 / HBM_PEAK_GBS, 4),
                                    "timed": "a separate pass of the same steps, an event pair "
                                             "around each cloud launch"},
                    "source_hash": source_hash()}

        # the bandwidth HBM sustains for the same read and write volumes, from a
        # contiguous streaming kernel (the attainable ceiling, beside the spec peak)
        in_b = min(C * 8192, (1 << 31) - 1024) // 1024 * 1024
        out_b = min(c.num_outputs * 8192, (1 << 31) - 1024) // 65536 * 65536
        p_in, p_out = eng.empty((in_b,), np.uint8), eng.empty((out_b,), np.uint8)
        p_in.fill_zero()
        for _ in range(2):
            eng.stream_probe(p_in, in_b, p_out, out_b)
        eng.sync()
        eng.timing(True)
        eng.timing_reset()
        for _ in range(10):
            eng.stream_probe(p_in, in_b, p_out, out_b)
        probe_ms, probe_n = eng.timing_read(_abi.TIMER_STREAM)
        eng.timing(False)
        p_in.free()
        p_out.free()
        probe_ms /= probe_n
        ceiling = {"probe": "contiguous 16-byte reads and writes interleaved per wave, the "
                            "volumes one cloud step reads and writes; streaming (nt) "
                            "device-scope (sc1) stores, as the cloud kernel's",
                   "read_bytes": in_b, "write_bytes": out_b, "ms": round(probe_ms, 4),
                   "GB_per_s": round((in_b + out_b) / probe_ms / 1e6, 1),
                   "cloud_kernel_frac": round(probe_ms / (avg_kernel_s * 1e3), 4)}
        log(f"streaming probe over the same volumes: {probe_ms:.4f} ms "
            f"({ceiling['GB_per_s']} GB/s); cloud kernel at {ceiling['cloud_kernel_frac']:.1%} "
            f"of it")

        # the step with the host hand-over of the C++ boundary: inputs copied from
        # pinned host memory, outputs back to it (ciphertext words and variances)
        from idash2019_2_amd import _abi as abi
        L = abi.lib()
        h_in = eng.pinned((C, 2, 1024), np.int32)
        h_out = eng.pinned((c.num_outputs, 2, 1024), np.int32)
        h_iv, h_ov = eng.pinned((C,), np.float64), eng.pinned((c.num_outputs,), np.float64)
        h_in.array[...] = d_ct.numpy()
        h_iv.array[...] = 2.0 ** -50
        pcie = []
        for _ in range(3):
            eng.sync()
            t0 = time.perf_counter()
            abi.check(L.ihe_memcpy_h2d_async(eng.ctx, d_ct.ptr, h_in.ptr, d_ct.nbytes))
            abi.check(L.ihe_memcpy_h2d_async(eng.ctx, d_var.ptr, h_iv.ptr, d_var.nbytes))
            eng.cloud(band, d_ct, d_var, d_out, d_ov)
            abi.check(L.ihe_memcpy_d2h_async(eng.ctx, h_out.ptr, d_out.ptr, d_out.nbytes))
            abi.check(L.ihe_memcpy_d2h_async(eng.ctx, h_ov.ptr, d_ov.ptr, d_ov.nbytes))
            eng.sync()
            pcie.append(time.perf_counter() - t0)
        for h in (h_in, h_out, h_iv, h_ov):
            h.free()
        pcie_s = min(pcie)
        log(f"step with host copies (pinned, {(d_ct.nbytes + d_out.nbytes) / 1e9:.2f} GB): "
            f"{pcie_s * 1e3:.1f} ms")

        # encrypt and decrypt kernels on the same workload (outside the timed steps)
        eng.timing(True)
        eng.timing_reset()
        d_dec = eng.empty((c.num_outputs, ns), np.float32)
        for _ in range(10):
            eng.encrypt(key, ns, d_geno, 0x5EED, d_ct, d_var)
            eng.decrypt(key, ns, d_out, c.num_outputs, d_dec)
        enc_ms, enc_n = eng.timing_read(_abi.TIMER_ENCRYPT)
        dec_ms, dec_n = eng.timing_read(_abi.TIMER_DECRYPT)
        eng.timing(False)
        d_dec.free()
        log(f"encrypt kernel avg {enc_ms / enc_n:.4f} ms ({C} ciphertexts), decrypt kernel avg "
            f"{dec_ms / dec_n:.4f} ms ({c.num_outputs} ciphertexts)")

        cpu = None
        if not args.no_cpu_baseline:
            # NB_THREADS = the CPUs this process may run on (eval/idash.h:36-38 makes it a
            # build constant), held to the host's share for one GPU (CPU_SHARE cores)
            visible = len(os.sched_getaffinity(0))
            threads = args.cpu_threads or min(visible, CPU_SHARE)
            secs, same, secs4 = cpu_baseline(eng, d_ct, c, rp, ib, w, d_out, threads)
            log(f"cpu baseline: {secs:.3f} s with {threads} threads ({visible} CPUs visible)"
                + (f", {secs4:.3f} s with 4" if secs4 else "")
                + f"; GPU output {'equals' if same else 'DIFFERS FROM'} the CPU output")
            if not same:
                raise SystemExit("GPU output differs from the CPU baseline")
            cpu = {"value": round(c.num_targets * ns / secs, 1), "unit": "imputed target SNPs/s",
                   "cores": threads, "kind": "port",
                   "sample": f"one full cloud step of the same workload ({c.num_targets} targets "
                             f"x {ns} samples), compute only, as eval/cloud.cpp:16-18 times it",
                   "cpus_visible": visible,
                   "threads_rule": f"OpenMP threads = min(CPUs visible, {CPU_SHARE}), the host's "
                                   "CPU share for one GPU",
                   "cpu_model": cpu_model()}
            if secs4:
                cpu["value_at_4_threads"] = round(c.num_targets * ns / secs4, 1)
        band.free()
        key.free()
        forms, forms_error = None, None
        if not args.no_forms and (args.forms or args.workload == "full"):
            # the cohort latency of the metric comes from the served form; a failure
            # there is reported beside the cloud-step line, which stays valid
            try:
                forms = {"batched": forms_batched(eng, c, log), "served": forms_served(eng, c, log)}
            except (Exception, SystemExit) as e:  # noqa: BLE001
                forms_error = f"{type(e).__name__}: {e}"
                log(f"forms failed: {forms_error}")

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "imputed target SNPs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"{args.workload}: {c.num_tags} tag SNPs, {c.num_targets} target "
                               f"SNPs, {ns} samples, {c.neighbours} neighbours, plain cloud step",
                   "num_in_cts": c.num_in_cts, "num_out_cts": c.num_outputs,
                   "variant": variant_name,
                   "tile": [info["tile_rows"], info["tile_kblocks"], info["col_tiles"]],
                   "algorithmic_bytes": alg_bytes,
                   "target_positions_per_s": round(c.num_targets / step_s, 1),
                   "cloud_kernel_ms": round(avg_kernel_s * 1e3, 4),
                   "step_event_ms": round(step_event_s * 1e3, 4),
                   "step_with_host_copies_ms": round(pcie_s * 1e3, 3),
                   "value_with_host_copies": round(c.num_targets * ns / pcie_s, 1),
                   "variance_kernel_ms": round(var_ms / var_n, 4) if var_n else None,
                   "encrypt_kernel_ms": round(enc_ms / enc_n, 4),
                   "decrypt_kernel_ms": round(dec_ms / dec_n, 4)},
        "roofline": roofline,
        "hbm_ceiling": ceiling,
        "cpu_baseline": cpu,
    }
    if forms is not None:
        sv = forms["served"]
        line["latency"] = {"form": "served", "requests": sv["requests"],
                           "arrival_rate_per_s": sv["arrival_rate_per_s"],
                           "p50_ms": sv["served_p50_ms"], "p99_ms": sv["served_p99_ms"],
                           "plain_p50_ms": sv["plain_p50_ms"], "plain_p99_ms": sv["plain_p99_ms"]}
        line["forms"] = forms
    if forms_error is not None:
        line["forms_error"] = forms_error
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
