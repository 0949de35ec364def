"""Throughput of the cloud step (cloud_compute_score, eval/idash.cpp:763-848)
on one MI355X.

A step is one cloud pass over the whole synthetic workload with its input
ciphertexts already resident in HBM: 16,184 tag SNPs, 80,882 target SNPs,
1,024 samples, 5 neighbours (48,552 input and 242,646 output ciphertexts).
The inputs are produced on the device by the library's own encrypt kernel.

Prints one JSON line. `value` is imputed target SNPs per second
(targets x samples / step time). `roofline` prices the main cloud kernel
against HBM bandwidth with its algorithmic bytes (every input ciphertext read
once, every output written once, plus the band); its launch durations come
from HIP events on the launch stream. `cpu_baseline` times the CPU
restatement of the reference's cloud step (hash-map model, OpenMP over
outputs) on this host.
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

from idash2019_2_amd import Engine, _abi, synth  # noqa: E402

METRIC = ("encrypted target SNPs imputed/s at 1 GPU; % HBM roofline; "
          "p50/p99 cohort latency")
HBM_PEAK_GBS = 8000.0
VARIANTS = {"auto": _abi.VARIANT_AUTO, "valu": _abi.VARIANT_VALU, "wide": _abi.VARIANT_WIDE,
            "mfma": _abi.VARIANT_MFMA}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_traffic(workload: str, variant: str):
    """HBM bytes per launch measured with rocprofv3 counters, if a summary is committed."""
    path = os.path.join(ROOT, "profiles", f"cloud_{variant}_{workload}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(eng, d_ct, c, rp, ib, w, d_out, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    ct = d_ct.numpy()
    var = np.full(c.num_in_cts, 2.0 ** -50)
    out, _, secs = oracle.cpu_baseline_cloud(ct, var, rp, ib, w, *c.layout, threads)
    # the GPU rows must equal the CPU rows
    gpu = d_out.numpy()
    same = bool((gpu == out).all())
    return secs, same


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="full", choices=sorted(synth.WORKLOADS))
    ap.add_argument("--variant", default="auto", choices=sorted(VARIANTS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run three 1 GiB device-to-device copies (a kernel of known "
                         "traffic for scaling HBM counters) and exit")
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the cloud step "
              "shards by target range with no exchange, and multi-GPU runs are not part of "
              "this build", file=sys.stderr)
        return 2

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
        log(f"band: {info['num_groups']} groups, {info['num_tiles']} tiles, variant "
            f"{variant_name}")

        for _ in range(args.warmup):
            eng.cloud(band, d_ct, d_var, d_out, d_ov)
        eng.sync()
        eng.timing(True)
        eng.timing_reset()
        eng.sync()
        t_start = time.perf_counter()
        for _ in range(args.steps):
            eng.cloud(band, d_ct, d_var, d_out, d_ov)
        eng.sync()
        elapsed = time.perf_counter() - t_start
        kern_ms, kern_n = eng.timing_read(_abi.TIMER_CLOUD)
        var_ms, _ = eng.timing_read(_abi.TIMER_CLOUD_VAR)
        eng.timing(False)
        log(f"{args.steps} steps in {elapsed * 1e3:.2f} ms; cloud kernel avg "
            f"{kern_ms / kern_n:.4f} ms, variance kernel avg {var_ms / kern_n:.4f} ms")

        step_s = elapsed / args.steps
        value = c.num_targets * ns / step_s
        alg_bytes = synth.cloud_bytes(c)
        avg_kernel_s = kern_ms / kern_n / 1e3
        achieved = alg_bytes / avg_kernel_s / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": read_traffic(args.workload, variant_name)}

        # encrypt and decrypt kernels on the same workload (outside the timed steps)
        eng.timing(True)
        eng.timing_reset()
        d_dec = eng.empty((c.num_outputs, ns), np.float32)
        for _ in range(3):
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
            threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) \
                or min(16, os.cpu_count() or 1)
            secs, same = cpu_baseline(eng, d_ct, c, rp, ib, w, d_out, threads)
            log(f"cpu baseline: {secs:.3f} s with {threads} threads; GPU output "
                f"{'equals' if same else 'DIFFERS FROM'} the CPU output")
            if not same:
                raise SystemExit("GPU output differs from the CPU baseline")
            cpu = {"value": round(c.num_targets * ns / secs, 1), "unit": "imputed target SNPs/s",
                   "cores": threads, "kind": "port",
                   "sample": f"one full cloud step of the same workload ({c.num_targets} targets "
                             f"x {ns} samples), compute only, as eval/cloud.cpp:16-18 times it"}
        band.free()
        key.free()

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "imputed target SNPs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"{args.workload}: {c.num_tags} tag SNPs, {c.num_targets} target "
                               f"SNPs, {ns} samples, {c.neighbours} neighbours, plain cloud step",
                   "num_in_cts": c.num_in_cts, "num_out_cts": c.num_outputs,
                   "variant": variant_name, "algorithmic_bytes": alg_bytes,
                   "target_positions_per_s": round(c.num_targets / step_s, 1),
                   "cloud_kernel_ms": round(avg_kernel_s * 1e3, 4),
                   "variance_kernel_ms": round(var_ms / kern_n, 4),
                   "encrypt_kernel_ms": round(enc_ms / enc_n, 4),
                   "decrypt_kernel_ms": round(dec_ms / dec_n, 4)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
