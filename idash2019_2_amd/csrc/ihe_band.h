// Device band of a compiled model: the interface between the band compiler
// (ihe_cloud.hip) and the batched launcher (ihe_batch.hip).
#pragma once

#include <atomic>
#include <vector>

#include "ihe_mfma.h"

namespace ihe {

struct GroupDev;  // direct dot2 layout (ihe_cloud.hip)
struct TileDev;   // LDS-staged dot2 layout (ihe_cloud.hip)

struct Segment {
  uint32_t tile_begin, tile_end;  // tiles [begin, end)
  uint32_t in_base;               // first input ciphertext (VALU) or big index (MFMA) addressed
  uint32_t out_base;              // first output slot addressed
};

enum { kModeDirect = 0, kModeStaged = 1, kModeMfma = 2 };

}  // namespace ihe

struct ihe_band {
  ihe_ctx *ctx = nullptr;
  // batches and servers holding the band (ihe_band_destroy refuses while > 0)
  mutable std::atomic<uint32_t> users{0};
  ihe_layout layout{};
  uint32_t num_in = 0, num_out = 0;
  int variant = IHE_VARIANT_VALU;
  int mode = 0;  // kModeDirect, kModeStaged or kModeMfma
  int wpl = 1;
  ihe::MfmaShape shape;        // MFMA: tile rows, K-blocks, column tiles per wave
  uint32_t max_bidx = 0;  // largest input big index the model references
  bool has_cols = false;
  uint32_t num_groups = 0, num_tiles = 0, num_cols = 0, nnz = 0, max_window = 0;
  uint32_t lds_bytes = 0, meta_lds = 0;
  std::vector<ihe::Segment> segments;
  void *dmem = nullptr;
  size_t dbytes = 0;
  // direct layout
  ihe::GroupDev *d_grp = nullptr;
  uint32_t *d_tile_off = nullptr, *d_cols = nullptr, *d_oslot = nullptr, *d_obias = nullptr;
  int32_t *d_wts = nullptr;
  // staged layout
  ihe::TileDev *d_tiles = nullptr;
  uint32_t *d_stage = nullptr, *d_meta = nullptr;
  // MFMA layout
  ihe::MTileDev *d_mtiles = nullptr;
  uint32_t *d_kcols = nullptr, *d_rowc = nullptr;
  ihe::int4v *d_adata = nullptr;
  // variances: every (big index, int32 w*w) entry of each row, in row order,
  // interleaved per block of 64 outputs; packed (d_var_base set): entries are
  // (big index - row base) << 16 | |w|, and d_var_w2 is unused
  uint32_t *d_var_ptr = nullptr, *d_var_col = nullptr, *d_var_base = nullptr;
  int32_t *d_var_w2 = nullptr;
};
