// Byte-split int8 MFMA cloud kernels: the interface between the band compiler
// (ihe_cloud.hip), the batched launcher (ihe_batch.hip) and the kernels
// (ihe_mfma.hip).
#pragma once

#include "ihe_internal.h"

namespace ihe {

constexpr uint32_t kMfmaRows = 32;      // outputs per MFMA row block
constexpr uint32_t kMfmaMaxNkb = 8;     // K-blocks (of 32 columns) per tile
constexpr uint32_t kSegCts = (1u << 18) - 1;  // ciphertexts per segment: byte offsets < 2^31
constexpr uint32_t kSegPadRows = 128;          // rows per tile at most

typedef int int4v __attribute__((ext_vector_type(4)));

// One tile: up to 32*RB consecutive output rows over nkb*32 columns. Column
// lists hold big indices, so one tiling serves every packing layout; the
// kernels derive the ciphertext (c / R) and rotation ((c % R) * RS) per launch.
struct MTileDev {
  uint32_t kcol_off;  // first column (nkb * 32 big indices)
  uint32_t nkb;       // K-blocks actually used (<= the launch's NKB)
  uint32_t a_off;     // first A operand block (64 lanes x 16 bytes): [row block][kb][weight byte]
  uint32_t row_off;   // 32 * RB row biases (bias * 2^18)
  uint32_t slot0;     // first output slot, relative to the segment's out_base
  uint32_t nrows;     // rows in use
  uint32_t pad0, pad1;
};

// One cohort segment as the kernels see it. Passed by value for a single
// cohort, or read from a device table in a batch.
struct alignas(16) MfmaEntry {
  const MTileDev *tiles;
  const uint32_t *kcols;
  const int4v *adata;
  const uint32_t *rowc;
  const int32_t *in;   // ciphertext ct_base of the cohort
  int32_t *out;        // output slot out_base of the cohort
  uint32_t in_bytes, out_bytes;
  uint32_t ns, rs, R, ct_base;  // layout; ct_base = first ciphertext of `in`
  uint32_t tile_base;           // first tile of the segment
  uint32_t pad0;
  uint64_t rmagic;  // c / R = umul64hi(c, rmagic) for every 32-bit c (R > 1)
  uint32_t pad[10];
};
static_assert(sizeof(MfmaEntry) == 128, "entry is two 64-byte scalar loads");

// Kernel family of a tiling.
//   rb 1 or 2: each wave owns column tiles of the tile's rows (rb 2 only with
//              nkb 1); ntw = column tiles per wave (4, 8 or 16).
//   rb 4:      the workgroup's four waves own one row block each and share the
//              column tiles; ntw = column tiles per workgroup (8, 16 or 32).
constexpr int kMfmaWideRb = 4;
struct MfmaShape {
  int rb = 2;   // row blocks per tile
  int nkb = 1;  // K-blocks per tile
  int ntw = 4;  // 32-word column tiles per wave (per workgroup for rb 4)
};

// Workgroups per tile of a shape (a ciphertext has kCtWords / 32 column tiles).
inline uint32_t mfma_wg_per_tile(const MfmaShape &s) {
  const uint32_t waves_per_tile = s.rb == kMfmaWideRb ? 1 : 4;
  return kCtWords / (waves_per_tile * (uint32_t)s.ntw * 32);
}


// Output variances: var_out[o] = sum over the row's entries with c % R == 0 of
// w2 * var_in[c / R] (int32 w*w, summed in row order). Entries are stored
// interleaved per block of 64 outputs: entry j of output 64*b + l sits at
// (blk_off[b] + j) * 64 + l, rows padded to the block's longest with
// kVarPad columns, so a wave reads entry j of its 64 outputs in one load.
constexpr uint32_t kVarPad = 0xFFFFFFFFu;
struct VarEntry {
  const uint32_t *blk_off;  // [ceil(num_out / 64) + 1]
  const uint32_t *col;      // big indices, or packed entries when `base` is set
  const int32_t *w2;        // int32 w*w (unpacked entries)
  const uint32_t *base;     // packed: each output's smallest big index, else null
  const double *var_in;
  double *var_out;
  uint32_t num_out, R;
  uint64_t rmagic;          // c / R = high 64 bits of c * rmagic (R > 1)
};

// The variance entry of one cohort: rmagic = floor((2^64 - 1) / R) + 1 makes
// the high half of c * rmagic equal floor(c / R) for every 32-bit c and
// 1 < R <= 2^31 (the error c * (rmagic - 2^64 / R) / 2^64 stays below 2^-32,
// under the 1 / R gap of a non-integral quotient to the next integer).
inline VarEntry make_var_entry(const uint32_t *blk_off, const uint32_t *col, const int32_t *w2,
                               const uint32_t *base, const double *var_in, double *var_out,
                               uint32_t num_out, uint32_t R) {
  return {blk_off, col, w2, base, var_in, var_out, num_out, R, R > 1 ? UINT64_MAX / R + 1 : 0};
}
// A packed entry's offset field for padding.
constexpr uint32_t kVarPackedPad = 0xFFFFu;

// Variance work of a launch: one cohort (`entries` null: `single`) or a table
// of cohorts (first[e] = entry e's first block, first[num_entries] =
// num_blocks), in blocks of 256 outputs.
struct VarArgs {
  VarEntry single;
  const VarEntry *entries;
  const uint32_t *first;
  uint32_t num_entries;
  uint32_t num_blocks;
};

void launch_var(hipStream_t stream, const VarArgs &var);

// Arguments of one family launch: a single cohort (`entries` null: `single`)
// or a table of cohort segments (first[e] = entry e's first tile, and
// optionally item_entry[t] = the entry of tile t); blocks at or past
// cloud_blocks run the variance blocks of `var`.
struct MfmaArgs {
  MfmaEntry single;
  const MfmaEntry *entries;
  const uint32_t *first;
  const uint32_t *item_entry;
  uint32_t num_entries;
  uint32_t cloud_blocks;
  VarArgs var;
};

// Persistent batch (k_cloud_batch): the families whose bodies it runs, with
// its own column-tile counts (NTW does not change a band's tiling).
constexpr int kPersistNtw1 = 4;          // per wave, one-K-block tiles of 32 or 64 rows
constexpr int kPersistNtwW = 16;         // per workgroup, 128-row tiles of up to 4 K-blocks
constexpr int kPersistNtwW6 = 32;        // per workgroup, 128-row tiles of 5 or 6 K-blocks
constexpr uint32_t kPersistMaxNkb = 6;   // 128-row tiles of up to 6 K-blocks
constexpr uint32_t kPersistQueueWords = 16;  // 8 XCD counters, finished workgroups
constexpr uint32_t kPersistBlocks = 2;       // workgroup blocks per item

struct alignas(16) PersistFamily {
  MfmaArgs args;  // no variance work (var.num_blocks = 0)
  uint32_t body;  // persist_body of the family
};

struct PersistArgs {
  const PersistFamily *fams;  // the num_once one-K-block families, then the num_fams others
  const uint32_t *once_first; // [num_once + 1]: first workgroup block of each one-K-block family
  const uint32_t *fam_first;  // [num_fams + 1]: first item of each other family (kPersistBlocks
                              // blocks per item); then the variance items (one block each)
  uint32_t num_once, once_blocks;  // once_blocks = once_first[num_once]: the grid's first blocks
  uint32_t num_fams, total;   // total = items of the resident workgroups
  uint32_t *queue;            // kPersistQueueWords words, zero before the first launch
  uint32_t bounds[9];         // XCD x's own items: [bounds[x], bounds[x + 1])
  VarArgs var;
};

// The body of a family in k_cloud_batch, or -1 when it has none (per-wave
// tiles over several K-blocks, 128-row tiles over 7 or 8).
int persist_body(const MfmaShape &shape, bool rot);
// Workgroup blocks per tile of a family in k_cloud_batch.
uint32_t persist_wg_per_tile(const MfmaShape &shape);
// Resident workgroups of a k_cloud_batch launch on `device` (all resident at
// once); the grid is once_blocks + this.
int persist_grid(int device, uint32_t &grid);
void launch_persistent(hipStream_t stream, const PersistArgs &P, uint32_t grid);

// Enqueues the MFMA cloud kernel. Single cohort: entries == nullptr and
// `single` describes it, with num_items = its tiles. Batch: entries[e] for
// e < num_entries, first[e] = index of entry e's first tile in the launch
// (first[num_entries] = num_items). `var` (num_blocks > 0) appends its
// variance blocks to the same grid: the step is one launch, and the variance
// blocks fill the cloud blocks' drain tail.
// item_entry (batches, optional): the entry of each of the num_items tiles,
// in place of a search over `first`. rot_runs4 (batches): every entry's
// region size is a multiple of 4, so rotated columns keep their words in runs
// of four (a single cohort's is read from `single`).
void launch_mfma(hipStream_t stream, const MfmaShape &shape, bool rot, const MfmaEntry &single,
                 const MfmaEntry *entries, const uint32_t *first, uint32_t num_entries,
                 uint32_t num_items, const VarArgs &var, const uint32_t *item_entry = nullptr,
                 bool rot_runs4 = false);

}  // namespace ihe
