"""Deterministic synthetic imputation workloads.

The real iDASH data is not available, so workloads are generated with the
shapes of the reference's run (README.md:220-224) and the semantics of its
training pipeline:
  * tag and target SNP positions: distinct integers in a chr22-like range,
    each list sorted; keygen keeps at most 50 tags past the last target
    (eval/idash.cpp:383-386);
  * genotypes over the merged sorted SNP list: p ~ Beta(0.5, 0.5) per SNP,
    g ~ Binomial(2, p), and with probability 0.9 a SNP copies its left
    neighbour's genotype for that sample;
  * each target's model uses its n nearest tags (train/to_vw.py:33), which is
    a contiguous interval of the sorted tag list; ties go to the lower tag;
  * weights: uniform integers in [-2048, 2047] per (target, class, tag,
    value), bias uniform in [-2048, 2047]; if B = |bias| + sum over tags of
    max_v |w| exceeds 8191, every weight and the bias of that output are
    scaled by 8191/B with truncation (the discretisation keeps predictions
    within +-8191, README.md:135-141); stored as int16;
  * key: uniform binary, N = 1024.
One numpy Generator per sub-stream (positions, genotypes, weights, key)
spawned from a single master seed per workload.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

N = 1024
CONSTANT_BIDX = 0xFFFFFFFF
POS_LO, POS_HI = 16_050_000, 51_250_000

# name -> (tags, targets, samples, neighbours, master seed)
WORKLOADS = {
    "small": (500, 2000, 128, 5, 1),
    "full": (16184, 80882, 1024, 5, 2),
    "wide": (16184, 80882, 1024, 50, 3),
}


@dataclass
class Cohort:
    num_samples: int
    neighbours: int
    tag_pos: np.ndarray      # [tags] sorted uint64
    target_pos: np.ndarray   # [targets] sorted uint64
    geno: np.ndarray         # [tags][NS] int8 (tag file order)
    lo: np.ndarray           # [targets] first tag of each window
    weights: np.ndarray      # [3*targets][3*n] int16, column c -> big index 3*lo + c
    bias: np.ndarray         # [3*targets] int32
    key: np.ndarray          # [N] int32 in {0,1}
    seed: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def num_tags(self) -> int:
        return int(self.tag_pos.size)

    @property
    def num_targets(self) -> int:
        return int(self.target_pos.size)

    @property
    def layout(self) -> tuple[int, int, int]:
        R = N // self.num_samples
        return self.num_samples, R, N // R

    @property
    def num_in_cts(self) -> int:
        R = self.layout[1]
        return (3 * self.num_tags + R - 1) // R

    @property
    def num_outputs(self) -> int:
        return 3 * self.num_targets

    def model_csr(self, with_zeros: bool = True):
        """The Model of read_model (eval/idash.cpp:66-114) as CSR over output big
        indices 3t+k: window columns in increasing big index, then the constant."""
        T, n3 = self.num_targets, 3 * self.neighbours
        cols = (3 * self.lo.astype(np.int64))[:, None] + np.arange(n3)[None, :]
        cols = np.repeat(cols, 3, axis=0)                       # [3T][3n]
        w = self.weights.astype(np.int32)
        in_bidx = np.concatenate([cols.astype(np.uint32),
                                  np.full((3 * T, 1), CONSTANT_BIDX, np.uint32)], axis=1)
        weight = np.concatenate([w, self.bias[:, None].astype(np.int32)], axis=1)
        if with_zeros:
            row_ptr = np.arange(3 * T + 1, dtype=np.uint32) * (n3 + 1)
            return row_ptr, in_bidx.ravel(), weight.ravel()
        keep = weight != 0
        keep[:, -1] = True
        counts = keep.sum(axis=1)
        row_ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
        return row_ptr, in_bidx[keep], weight[keep]


def _positions(rng, count):
    got = np.empty(0, np.int64)
    while got.size < count:
        draw = rng.integers(POS_LO, POS_HI, size=int(count * 1.1) + 16)
        got = np.unique(np.concatenate([got, draw]))
    return np.sort(rng.choice(got, size=count, replace=False))


def _windows(tags: np.ndarray, targets: np.ndarray, n: int) -> np.ndarray:
    T = tags.size
    n = min(n, T)
    idx = np.searchsorted(tags, targets)
    lo = np.clip(idx - n, 0, T - n)
    t = targets.astype(np.int64)
    tg = tags.astype(np.int64)
    for _ in range(n + 1):
        nxt = np.minimum(lo + n, T - 1)
        move = (lo + n < T) & ((t - tg[lo]) > (tg[nxt] - t))
        if not move.any():
            break
        lo = lo + move
    return lo.astype(np.int64)


def _weights(r_w, T: int, n: int):
    """Per (target, class, tag, value) weights and per-output biases, scaled so
    that |bias| + sum over tags of max_v |w| <= 8191, stored as int16/int32."""
    w = r_w.integers(-2048, 2048, size=(T, 3, n, 3), dtype=np.int64)
    b = r_w.integers(-2048, 2048, size=(T, 3), dtype=np.int64)
    B = np.abs(b) + np.abs(w).max(axis=3).sum(axis=2)          # [T][3]
    scale = B > 8191
    Bs = np.where(scale, B, 1)
    w = np.where(scale[:, :, None, None], np.sign(w) * ((np.abs(w) * 8191) // Bs[:, :, None, None]), w)
    b = np.where(scale, np.sign(b) * ((np.abs(b) * 8191) // Bs), b)
    return w.reshape(3 * T, 3 * n).astype(np.int16), b.reshape(3 * T).astype(np.int32)


def generate(num_tags: int, num_targets: int, num_samples: int, neighbours: int,
             seed: int) -> Cohort:
    ss = np.random.SeedSequence(seed)
    r_pos, r_geno, r_w, r_key = (np.random.default_rng(s) for s in ss.spawn(4))

    pos = _positions(r_pos, num_tags + num_targets)
    is_tag = np.zeros(pos.size, bool)
    is_tag[r_pos.choice(pos.size, size=num_tags, replace=False)] = True
    tag_pos, target_pos = pos[is_tag], pos[~is_tag]

    # keygen_ph1 stops reading tags 50 tags past the last target (eval/idash.cpp:383-386)
    after = np.nonzero(tag_pos > target_pos[-1])[0]
    if after.size > 50:
        keep = np.ones(tag_pos.size, bool)
        keep[after[50:]] = False
        tag_pos = tag_pos[keep]
        is_tag_kept = np.zeros(pos.size, bool)
        is_tag_kept[np.searchsorted(pos, tag_pos)] = True
    else:
        is_tag_kept = is_tag

    # genotypes over the merged sorted list
    p = r_geno.beta(0.5, 0.5, size=pos.size)
    geno_all = np.empty((pos.size, num_samples), np.int8)
    prev = None
    for c0 in range(0, pos.size, 4096):
        c1 = min(pos.size, c0 + 4096)
        fresh = r_geno.binomial(2, p[c0:c1, None], size=(c1 - c0, num_samples)).astype(np.int8)
        copy = r_geno.random((c1 - c0, num_samples), dtype=np.float32) < 0.9
        for i in range(c1 - c0):
            if prev is not None:
                np.copyto(fresh[i], prev, where=copy[i])
            prev = fresh[i]
        geno_all[c0:c1] = fresh
    geno = np.ascontiguousarray(geno_all[is_tag_kept])

    lo = _windows(tag_pos, target_pos, neighbours)
    n = min(neighbours, tag_pos.size)
    weights, bias = _weights(r_w, target_pos.size, n)
    key = r_key.integers(0, 2, size=N, dtype=np.int32)
    return Cohort(num_samples, n, tag_pos.astype(np.uint64), target_pos.astype(np.uint64), geno,
                  lo, weights, bias, key, seed)


def workload(name: str, **overrides) -> Cohort:
    tags, targets, ns, n, seed = WORKLOADS[name]
    args = dict(num_tags=tags, num_targets=targets, num_samples=ns, neighbours=n, seed=seed)
    args.update(overrides)
    return generate(**args)


def cloud_bytes(c: Cohort) -> int:
    """Algorithmic HBM bytes of one cloud step: every input ciphertext read once,
    every output ciphertext written once, int16 weights, int32 bias and window
    start per target."""
    return (8192 * (c.num_in_cts + c.num_outputs) + 2 * 3 * c.neighbours * c.num_outputs
            + 4 * c.num_outputs + 4 * c.num_targets)


def write_reference_files(c: Cohort, directory: str, nan_fraction: float = 0.0,
                          seed: int = 0) -> dict:
    """Writes the cohort in the reference's input formats:
      tags.txt     "22 pos pos+1 name g_0 .. g_{NS-1}" per tag, "NA" when missing
                   (read by keygen_ph1 and read_plaintext_data, eval/idash.cpp:252-311, 374-405)
      targets.txt  one target position per line (the positions-only target file,
                   eval/idash.cpp:327-343)
      model/       "<pos>_<k>.hr" per target and class with "<tagpos>_<v> <w>" lines and
                   a "Constant <bias>" line (eval/parse_vw.cpp:7-29, eval/idash.cpp:66-90)
    Returns the paths and the genotypes written (with -1 where missing)."""
    import os
    os.makedirs(os.path.join(directory, "model"), exist_ok=True)
    geno = c.geno.copy()
    if nan_fraction > 0:
        rng = np.random.default_rng(seed)
        geno[rng.random(geno.shape) < nan_fraction] = -1
    tags = os.path.join(directory, "tags.txt")
    with open(tags, "w") as f:
        for q, pos in enumerate(c.tag_pos.tolist()):
            g = " ".join("NA" if v < 0 else str(int(v)) for v in geno[q])
            f.write(f"22 {pos} {pos + 1} rs{pos} {g}\n")
    targets = os.path.join(directory, "targets.txt")
    with open(targets, "w") as f:
        f.write("".join(f"{p}\n" for p in c.target_pos.tolist()))
    n = c.neighbours
    tag_pos = c.tag_pos.tolist()
    for t, pos in enumerate(c.target_pos.tolist()):
        lo = int(c.lo[t])
        for k in range(3):
            o = 3 * t + k
            lines = [f"{tag_pos[lo + j // 3]}_{j % 3} {int(c.weights[o, j])}\n" for j in range(3 * n)]
            lines.append(f"Constant {int(c.bias[o])}\n")
            with open(os.path.join(directory, "model", f"{pos}_{k}.hr"), "w") as f:
                f.write("".join(lines))
    return {"tags": tags, "targets": targets, "model": os.path.join(directory, "model"),
            "geno": geno}


def model_sets(base: Cohort, neighbours, sets: int, seed: int) -> dict:
    """Model sets sharing base's SNP layout (positions, genotypes): for every
    neighbour count a window per target, and for every set independently drawn
    weights. Returns {(set, n): Cohort} (the batched and served workloads)."""
    out = {}
    for n in neighbours:
        lo = _windows(base.tag_pos, base.target_pos, n)
        for s in range(sets):
            r_w = np.random.default_rng(np.random.SeedSequence([seed, n, s]))
            weights, bias = _weights(r_w, base.num_targets, n)
            out[(s, n)] = Cohort(base.num_samples, n, base.tag_pos, base.target_pos, base.geno,
                                 lo, weights, bias, base.key, seed)
    return out
