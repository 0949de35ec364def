// Batched cloud step: many cohorts (each with its own sample count, input
// ciphertexts and resident model band) in one persistent launch, optionally
// captured into a hipGraph so that a step is a single graph launch.
//
// The reference runs one cloud_compute_score per cohort (eval/cloud.cpp:17);
// here the per-cohort work lists are concatenated. Cohorts whose bands share
// a kernel family (tile shape and rotation) form one family: its work is a
// list of workgroup blocks, each finding its (cohort segment, tile) by a
// lookup in a tile -> entry table. The one kernel (k_cloud_batch) runs the
// one-K-block families' blocks first, one per workgroup as the dispatcher
// hands them out, and then its resident workgroups take the other families'
// blocks and the variance blocks from per-XCD counters; XCD x starts on its
// own contiguous share of that list (shares cut by estimated time), so
// consecutive blocks of one cohort meet in one L2. A batch with a family the
// kernel has no body for (or with IHE_BATCH_PERSISTENT=0) runs one launch
// per family instead, the variance blocks joining the last. By default a
// batch with 128-row families runs per family too: their own kernel stages
// input words in LDS at three workgroups per CU, where the one kernel's
// register budget (its largest body's) holds two workgroups per CU and its
// 128-row bodies stage in registers; IHE_BATCH_PERSISTENT=1 asks for the one
// kernel whenever it has a body for every family.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ihe_band.h"

using namespace ihe;

struct ihe_batch {
  ihe_ctx *ctx = nullptr;
  struct Launch {
    MfmaShape shape;
    bool rot;
    uint32_t entry0, num_entries, first0, num_items;
    uint32_t item0;  // first tile in the tile -> entry table
    bool runs4;      // rotated 128-row family whose region sizes are multiples of 4
  };
  std::vector<Launch> launches;
  bool persistent = false;
  PersistArgs P{};
  uint32_t pgrid = 0;
  std::vector<const ihe_band *> bands;  // each distinct band once; its use count is held
  uint32_t num_cohorts = 0, num_items = 0;
  uint32_t num_ventries = 0, num_vblocks = 0;
  void *dmem = nullptr;
  MfmaEntry *d_entries = nullptr;
  uint32_t *d_first = nullptr;
  uint32_t *d_item_entry = nullptr;  // per family launch: entry of each tile
  VarEntry *d_ventries = nullptr;
  uint32_t *d_vfirst = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // per-family launches alternate between the context's stream and `side`,
  // forked and joined by events, so that one family's drain overlaps the
  // next one's start (null: every launch on the context's stream)
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};

namespace {

struct Keyed {
  MfmaShape shape;
  bool rot;
  uint32_t cohort;
  MfmaEntry e;
  uint32_t tiles;
  // rotated 128-row tiles whose region size is a multiple of 4: their own
  // family, launched with 16-byte staging loads
  bool runs4;
};

bool same_family(const Keyed &a, const Keyed &b) {
  return a.shape.rb == b.shape.rb && a.shape.nkb == b.shape.nkb && a.shape.ntw == b.shape.ntw &&
         a.rot == b.rot && a.runs4 == b.runs4;
}

VarArgs var_args(const ihe_batch *b) {
  VarArgs V{};
  V.entries = b->d_ventries;
  V.first = b->d_vfirst;
  V.num_entries = b->num_ventries;
  V.num_blocks = b->num_ventries ? b->num_vblocks : 0;
  return V;
}

// Relative time per output word of a persistent body (one-K-block tiles run
// the full 16,184 / 80,882 layout in about 0.47 ms, 128-row tiles of 2, 4 and
// 6 K-blocks in about 0.55, 0.68 and 0.88 ms), and the words of its blocks.
double body_cost(uint32_t body) {
  if (body < 4) return 0.47;
  const uint32_t k = (body - 4) % 3;
  return k == 0 ? 0.55 : k == 1 ? 0.68 : 0.88;
}
double block_words(const MfmaShape &shape) {
  return (double)kCtWords * kMfmaRows * shape.rb / persist_wg_per_tile(shape);
}

void release_side(ihe_batch *b) {
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->join) (void)hipEventDestroy(b->join);
  if (b->side) (void)hipStreamDestroy(b->side);
  b->fork = b->join = nullptr;
  b->side = nullptr;
}

void enqueue(const ihe_batch *b, hipStream_t s) {
  if (b->persistent) return launch_persistent(s, b->P, b->pgrid);
  MfmaEntry none = {};
  const VarArgs V = var_args(b);
  if (b->side) {
    (void)hipEventRecord(b->fork, s);
    (void)hipStreamWaitEvent(b->side, b->fork, 0);
  }
  // the variance blocks join the last family's launch
  for (size_t i = 0; i < b->launches.size(); ++i) {
    const ihe_batch::Launch &L = b->launches[i];
    launch_mfma(b->side && i % 2 ? b->side : s, L.shape, L.rot, none, b->d_entries + L.entry0,
                b->d_first + L.first0, L.num_entries, L.num_items,
                i + 1 == b->launches.size() ? V : VarArgs{}, b->d_item_entry + L.item0, L.runs4);
  }
  if (b->side) {
    (void)hipEventRecord(b->join, b->side);
    (void)hipStreamWaitEvent(s, b->join, 0);
  }
  if (b->launches.empty()) launch_var(s, V);
}

}  // namespace

extern "C" {

int ihe_batch_create(ihe_ctx *ctx, const ihe_cohort *cohorts, uint32_t num_cohorts,
                     ihe_batch **out) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!out || (num_cohorts && !cohorts)) return fail(IHE_ERR_ARG, "null argument");
  *out = nullptr;
  std::vector<Keyed> keyed;
  std::vector<VarEntry> vent;
  std::vector<uint32_t> vfirst(1, 0);
  for (uint32_t i = 0; i < num_cohorts; ++i) {
    const ihe_cohort &c = cohorts[i];
    const std::string which = "cohort " + std::to_string(i) + ": ";
    const ihe_band *band = c.band;
    if (!band || band->ctx->device != ctx->device)
      return fail(IHE_ERR_ARG, which + "band is not on this context's device");
    if (band->mode != kModeMfma)
      return fail(IHE_ERR_UNSUPPORTED, which + "batches run MFMA bands only");
    const ihe_layout &L = c.layout;
    if (L.num_samples == 0 || L.num_samples > kN || L.num_regions != kN / L.num_samples ||
        L.region_size != kN / L.num_regions)
      return fail(IHE_ERR_ARG, which + "layout must satisfy R = N / NS and RS = N / R");
    if (band->has_cols && band->max_bidx / L.num_regions >= c.num_in_cts)
      return fail(IHE_ERR_RANGE, which + "the model references FeatIndex " +
                                     std::to_string(band->max_bidx / L.num_regions) + " >= " +
                                     std::to_string(c.num_in_cts));
    if (band->num_out == 0) continue;
    if (!c.d_out || (c.num_in_cts && !c.d_in)) return fail(IHE_ERR_ARG, which + "null ciphertext buffer");
    if (c.d_out_var && c.num_in_cts && !c.d_in_var) return fail(IHE_ERR_ARG, which + "null input variance");
    for (const Segment &s : band->segments) {
      Keyed k;
      k.shape = band->shape;
      k.rot = L.num_regions > 1;
      k.cohort = i;
      k.tiles = s.tile_end - s.tile_begin;
      MfmaEntry &E = k.e;
      E = {};
      E.tiles = band->d_mtiles;
      E.kcols = band->d_kcols;
      E.adata = band->d_adata;
      E.rowc = band->d_rowc;
      E.ns = L.num_samples;
      E.rs = L.region_size;
      E.R = L.num_regions;
      E.rmagic = L.num_regions > 1 ? UINT64_MAX / L.num_regions + 1 : 0;
      E.ct_base = s.in_base / L.num_regions;
      const uint32_t in_cts = c.num_in_cts > E.ct_base ? std::min(c.num_in_cts - E.ct_base, kSegCts) : 0u;
      E.in = in_cts ? c.d_in + (size_t)E.ct_base * kCtWords : nullptr;
      E.in_bytes = in_cts * kCtWords * 4u;
      E.out = c.d_out + (size_t)s.out_base * kCtWords;
      E.out_bytes = std::min(band->num_out - s.out_base, kSegCts) * kCtWords * 4u;
      E.tile_base = s.tile_begin;
      k.runs4 = k.rot && k.shape.rb == kMfmaWideRb && E.rs % 4 == 0;
      if (k.tiles) keyed.push_back(k);
    }
    if (c.d_out_var) {
      vent.push_back(make_var_entry(band->d_var_ptr, band->d_var_col, band->d_var_w2,
                                    band->d_var_base, c.d_in_var,
                                    c.d_out_var, band->num_out, L.num_regions));
      vfirst.push_back(vfirst.back() + (band->num_out + 255) / 256);
    }
  }
  // kernel families (tile shape x rotation), cohorts in order within each
  std::stable_sort(keyed.begin(), keyed.end(), [](const Keyed &a, const Keyed &b) {
    if (a.rot != b.rot) return a.rot < b.rot;
    if (a.runs4 != b.runs4) return a.runs4 < b.runs4;
    if (a.shape.nkb != b.shape.nkb) return a.shape.nkb < b.shape.nkb;
    if (a.shape.rb != b.shape.rb) return a.shape.rb < b.shape.rb;
    return a.shape.ntw < b.shape.ntw;
  });
  ihe_batch *b = new ihe_batch();
  b->ctx = ctx;
  b->num_cohorts = num_cohorts;
  std::vector<MfmaEntry> entries;
  std::vector<uint32_t> first, item_entry;
  for (size_t i = 0; i < keyed.size();) {
    size_t j = i;
    ihe_batch::Launch L{keyed[i].shape, keyed[i].rot, (uint32_t)entries.size(), 0,
                        (uint32_t)first.size(), 0, (uint32_t)item_entry.size(), keyed[i].runs4};
    uint32_t items = 0;
    for (; j < keyed.size() && same_family(keyed[i], keyed[j]); ++j) {
      entries.push_back(keyed[j].e);
      first.push_back(items);
      items += keyed[j].tiles;
      item_entry.insert(item_entry.end(), keyed[j].tiles, (uint32_t)(j - i));
    }
    first.push_back(items);
    L.num_entries = (uint32_t)(j - i);
    L.num_items = items;
    b->num_items += items;
    b->launches.push_back(L);
    i = j;
  }
  b->num_ventries = (uint32_t)vent.size();
  b->num_vblocks = vfirst.back();
  // one persistent launch when every family has a body there; by default
  // only when every family is a one-K-block one
  const char *pe = std::getenv("IHE_BATCH_PERSISTENT");
  const bool ask_off = pe && std::strcmp(pe, "0") == 0, ask_on = pe && std::strcmp(pe, "1") == 0;
  b->persistent = !b->launches.empty() && !ask_off;
  for (const ihe_batch::Launch &L : b->launches)
    b->persistent = b->persistent && persist_body(L.shape, L.rot) >= 0 &&
                    (ask_on || L.shape.rb != kMfmaWideRb);
  if (b->persistent) {
    if (int rc = persist_grid(ctx->device, b->pgrid)) {
      delete b;
      return rc;
    }
  }
  const size_t num_fams = b->persistent ? b->launches.size() : 0;  // one-shot and resident
  auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t o_first = align(entries.size() * sizeof(MfmaEntry));
  const size_t o_vent = o_first + align(first.size() * 4);
  const size_t o_vfirst = o_vent + align(vent.size() * sizeof(VarEntry));
  const size_t o_items = o_vfirst + align(vfirst.size() * 4);
  const size_t o_fams = o_items + align(item_entry.size() * 4);
  const size_t o_ffirst = o_fams + align(num_fams * sizeof(PersistFamily));
  const size_t o_queue = o_ffirst + align((num_fams + 2) * 4);
  const size_t bytes = o_queue + (b->persistent ? align(kPersistQueueWords * 4) : 0);
  if (int rc = ihe_malloc(ctx, bytes, &b->dmem)) {
    delete b;
    return rc;
  }
  std::vector<char> h(bytes, 0);
  if (!entries.empty()) std::memcpy(h.data(), entries.data(), entries.size() * sizeof(MfmaEntry));
  if (!first.empty()) std::memcpy(h.data() + o_first, first.data(), first.size() * 4);
  if (!vent.empty()) std::memcpy(h.data() + o_vent, vent.data(), vent.size() * sizeof(VarEntry));
  std::memcpy(h.data() + o_vfirst, vfirst.data(), vfirst.size() * 4);
  if (!item_entry.empty())
    std::memcpy(h.data() + o_items, item_entry.data(), item_entry.size() * 4);
  char *d = static_cast<char *>(b->dmem);
  b->d_entries = reinterpret_cast<MfmaEntry *>(d);
  b->d_first = reinterpret_cast<uint32_t *>(d + o_first);
  b->d_ventries = reinterpret_cast<VarEntry *>(d + o_vent);
  b->d_vfirst = reinterpret_cast<uint32_t *>(d + o_vfirst);
  b->d_item_entry = reinterpret_cast<uint32_t *>(d + o_items);
  if (b->persistent) {
    // the one-K-block families first (their blocks lead the grid), then the
    // others in launch order (kPersistBlocks workgroup blocks per item) and
    // the variance blocks, cut into XCD shares by estimated time (a variance
    // block taken as 1/16 of a 64-row block)
    std::vector<PersistFamily> fams;
    std::vector<const ihe_batch::Launch *> order;
    for (int pass = 0; pass < 2; ++pass)
      for (const ihe_batch::Launch &L : b->launches)
        if ((persist_body(L.shape, L.rot) < 4) == (pass == 0)) order.push_back(&L);
    std::vector<uint32_t> once_first(1, 0), ffirst(1, 0);
    std::vector<double> cost;
    double total_cost = 0;
    for (const ihe_batch::Launch *L : order) {
      PersistFamily F{};
      F.args.entries = b->d_entries + L->entry0;
      F.args.first = b->d_first + L->first0;
      F.args.item_entry = b->d_item_entry + L->item0;
      F.args.num_entries = L->num_entries;
      F.args.cloud_blocks = L->num_items * persist_wg_per_tile(L->shape);
      F.body = (uint32_t)persist_body(L->shape, L->rot);
      if (F.body < 4) {
        once_first.push_back(once_first.back() + F.args.cloud_blocks);
      } else {
        const uint32_t items = (F.args.cloud_blocks + kPersistBlocks - 1) / kPersistBlocks;
        ffirst.push_back(ffirst.back() + items);
        cost.push_back(body_cost(F.body) * block_words(L->shape) * kPersistBlocks);
        total_cost += cost.back() * items;
      }
      fams.push_back(F);
    }
    const size_t num_res = ffirst.size() - 1;
    const double var_cost = 0.47 * kCtWords * kMfmaRows * 2 / 4 / 16;
    total_cost += var_cost * b->num_vblocks;
    PersistArgs &P = b->P;
    P.fams = reinterpret_cast<const PersistFamily *>(d + o_fams);
    P.once_first = reinterpret_cast<const uint32_t *>(d + o_ffirst);
    P.fam_first = P.once_first + once_first.size();
    P.num_once = (uint32_t)(once_first.size() - 1);
    P.once_blocks = once_first.back();
    P.num_fams = (uint32_t)num_res;
    P.total = ffirst.back() + b->num_vblocks;
    P.queue = reinterpret_cast<uint32_t *>(d + o_queue);
    P.var = var_args(b);
    // bounds[x] = first item whose running cost reaches x / 8 of the total
    P.bounds[0] = 0;
    P.bounds[8] = P.total;
    double run = 0;
    uint32_t x = 1;
    for (size_t f = 0; f <= num_res && x < 8; ++f) {
      const uint32_t n = f < num_res ? ffirst[f + 1] - ffirst[f] : b->num_vblocks;
      const double c = f < num_res ? cost[f] : var_cost;
      const uint32_t base = ffirst[f];
      while (x < 8 && run + c * n >= total_cost * x / 8) {
        const double need = std::max(0.0, total_cost * x / 8 - run);
        P.bounds[x] = base + std::min<uint32_t>(n, (uint32_t)std::ceil(need / c));
        ++x;
      }
      run += c * n;
    }
    for (; x < 8; ++x) P.bounds[x] = P.total;
    for (x = 1; x <= 8; ++x) P.bounds[x] = std::max(P.bounds[x], P.bounds[x - 1]);
    std::memcpy(h.data() + o_fams, fams.data(), fams.size() * sizeof(PersistFamily));
    std::memcpy(h.data() + o_ffirst, once_first.data(), once_first.size() * 4);
    std::memcpy(h.data() + o_ffirst + once_first.size() * 4, ffirst.data(), ffirst.size() * 4);
  }
  if (int rc = ihe_memcpy_h2d(ctx, b->dmem, h.data(), bytes)) {
    (void)ihe_free(ctx, b->dmem);
    delete b;
    return rc;
  }
  for (uint32_t i = 0; i < num_cohorts; ++i)
    if (std::find(b->bands.begin(), b->bands.end(), cohorts[i].band) == b->bands.end())
      b->bands.push_back(cohorts[i].band);
  for (const ihe_band *band : b->bands) band->users++;
  // a second stream for per-family launches (IHE_BATCH_STREAMS=1: one stream)
  const char *se = std::getenv("IHE_BATCH_STREAMS");
  if (!b->persistent && b->launches.size() > 1 && !(se && std::strcmp(se, "1") == 0)) {
    if (hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      release_side(b);
    }
  }
  *out = b;
  return IHE_OK;
}

int ihe_batch_run(ihe_batch *b) {
  if (!b) return fail(IHE_ERR_ARG, "null batch");
  if (int rc = make_current(b->ctx)) return rc;
  ScopedTimer t(b->ctx, IHE_TIMER_BATCH);
  if (b->exec) {
    IHE_HIP(hipGraphLaunch(b->exec, b->ctx->stream));
  } else {
    enqueue(b, b->ctx->stream);
    IHE_HIP(hipGetLastError());
  }
  return IHE_OK;
}

int ihe_batch_capture(ihe_batch *b) {
  if (!b) return fail(IHE_ERR_ARG, "null batch");
  if (int rc = make_current(b->ctx)) return rc;
  if (b->exec) return IHE_OK;
  IHE_HIP(hipStreamBeginCapture(b->ctx->stream, hipStreamCaptureModeThreadLocal));
  enqueue(b, b->ctx->stream);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->ctx->stream, &g);
  if (e != hipSuccess) return fail(IHE_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  b->graph = g;
  IHE_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return IHE_OK;
}

int ihe_batch_get_info(const ihe_batch *b, ihe_batch_info *info) {
  if (!b || !info) return fail(IHE_ERR_ARG, "null argument");
  info->num_cohorts = b->num_cohorts;
  info->num_launches = b->persistent ? 1u
                                      : (uint32_t)b->launches.size() +
                                            (b->launches.empty() && b->num_ventries ? 1u : 0u);
  info->persistent = b->persistent ? 1u : 0u;
  info->num_tiles = b->num_items;
  info->captured = b->exec ? 1u : 0u;
  return IHE_OK;
}

int ihe_batch_destroy(ihe_batch *b) {
  if (!b) return IHE_OK;
  (void)ihe_sync(b->ctx);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  release_side(b);
  (void)ihe_free(b->ctx, b->dmem);
  for (const ihe_band *band : b->bands) band->users--;
  delete b;
  return IHE_OK;
}

}  // extern "C"
