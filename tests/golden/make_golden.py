"""Writes the golden fixtures of the hot path: small cohorts (inputs stored in
full) with the oracle's encrypted inputs, cloud outputs and decrypted phases.

The reference ships no vectors for this path (its arithmetic lives in the
absent TFHE submodule), so these fixtures are this build's own: the oracle
that makes them is pinned by tests/test_oracle.py (closed forms, textbook
negacyclic products, the plaintext round trip). They anchor the oracle
against drift (tests/test_golden.py) and give the device path known answers
that need no oracle at run time (tests/test_gpu_golden.py).

Run from the repository root after building the oracle (make -C oracle):
    python3 tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
from idash2019_2_amd import synth  # noqa: E402

N = 1024
FULL_PHASE_BYTES = 64 << 10   # phases stored in full up to this size, else hash + head rows

# name: (tags, targets, NS, neighbours, cohort seed, encryption seed, missing-genotype
# fraction, drop zero weights). R = N / NS: 8, 1, 64 and 3 (RS = 341, a hidden tail).
CASES = {
    "r8_ns128": (60, 120, 128, 5, 101, 0x1D45_2019_0001, 0.03, False),
    "r1_ns1024": (40, 60, 1024, 5, 102, 0x1D45_2019_0002, 0.0, False),
    "r64_ns16": (200, 80, 16, 5, 103, 0x1D45_2019_0003, 0.02, False),
    "r3_ns333": (50, 70, 333, 10, 104, 0x1D45_2019_0004, 0.0, True),
}


def sha(a: np.ndarray) -> str:
    """SHA-256 of the array's little-endian bytes, C order."""
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def make(name: str) -> dict:
    tags, targets, ns, n, seed, enc_seed, nan_frac, drop_zeros = CASES[name]
    c = synth.generate(tags, targets, ns, n, seed=seed)
    geno = c.geno.copy()
    if nan_frac:
        rng = np.random.default_rng(seed)
        geno[rng.random(geno.shape) < nan_frac] = -1          # NaN genotypes (setScore)
    if drop_zeros:
        c.weights[::5, 2] = 0                                # .hr files omit zero terms
    rp, ib, w = c.model_csr(with_zeros=not drop_zeros)
    _, R, RS = c.layout
    ct, var = oracle.encrypt(geno, c.key, ns, R, RS, enc_seed)
    out, ov = oracle.cloud(ct, var, rp, ib, w, ns, R, RS)
    ph = oracle.decrypt(out, c.key, ns).astype(np.float32)
    f = dict(ns=np.uint32(ns), R=np.uint32(R), RS=np.uint32(RS), enc_seed=np.uint64(enc_seed),
             geno=geno, key=c.key.astype(np.int32), row_ptr=rp.astype(np.uint32),
             in_bidx=ib.astype(np.uint32), weight=w.astype(np.int32),
             enc_sha=np.array(sha(ct)), enc_var=var, enc_head=ct[:2],
             cloud_sha=np.array(sha(out)), cloud_var=ov, cloud_head=out[:2],
             phase_sha=np.array(sha(ph)))
    if ph.nbytes <= FULL_PHASE_BYTES:
        f["phase"] = ph
    else:
        f["phase_head"] = ph[:4]
    return f


def main():
    for name in CASES:
        f = make(name)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **f)
        print(f"{path}: {os.path.getsize(path)} bytes, {f['row_ptr'].size - 1} outputs, "
              f"{f['enc_var'].size} input ciphertexts")


if __name__ == "__main__":
    main()
