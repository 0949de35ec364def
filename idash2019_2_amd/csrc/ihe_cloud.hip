// Cloud step: band compiler (host) and the VALU weighted-sum kernels.
//
// Replaces cloud_compute_score (eval/idash.cpp:763-848). For every output slot
// o the step computes, mod 2^32,
//   out_o = sum_c w_{o,c} * rho_{(c%R)*RS}(CT[c/R])      (both polynomials)
//   out_o.b[j] += bias_o * 2^18  for j < NS              (:805-810)
//   out_o.b[j]  = 0              for j in [RS, N)         (:838-841)
//   var_o = sum_{c : c%R == 0} (int32)(w*w) * var(CT[c/R])
// where rho_m(P)[i] = P[i+m] if i+m < N else -P[i+m-N] is multiplication by
// X^(2N-m) (torusPolynomialMulByXai, :826-836). The reference builds per-region
// temporaries and rotates them after the sum (:795-836); rotation is linear
// mod 2^32, so rotating each input column on load gives the same words.
//
// Device band format. Outputs with identical column lists are grouped (the
// three classes of a target and the consecutive targets that share the same
// nearest-tag window). A group lists its columns as packed (ct << 10 | shift)
// words, its outputs, and a dense [nout][ncol] weight block. Groups are cut
// into tiles of roughly equal work; one workgroup handles one tile over one
// slice of the 2048 coefficient words.
//
// Exact int32 arithmetic on the VALU: with x = hi*2^16 + lo (lo, hi read as
// signed 16-bit after the recode hi = (x + 0x8000) >> 16), and a weight w
// within int16,
//   w*x = w*lo + 2^16 * w*hi   (mod 2^32)
// so two columns share one v_dot2_i32_i16 per half: a column pair costs two
// dot2 instructions per output word.
#include <algorithm>
#include <cstring>

#include "ihe_internal.h"

namespace ihe {

constexpr int kOmax = 12;        // outputs per group (register-blocked accumulators)
constexpr int kPairChunk = 8;    // column pairs held in registers at a time
constexpr int kColChunk = 8;     // columns per chunk in the wide kernel
constexpr uint32_t kMaxCt = 1u << 22;

struct GroupDev {
  uint32_t col_off;  // first packed column of the group
  uint32_t nunits;   // column pairs (dot2) or columns (wide)
  uint32_t out_off;  // first entry in outs[]
  uint32_t nout;     // outputs in the group (<= kOmax)
  uint32_t w_off;    // first weight word of the [nout][nunits] block
  uint32_t pad0, pad1, pad2;
};

struct CloudParams {
  const GroupDev *__restrict__ grp;
  const uint32_t *__restrict__ tile_off;
  const uint32_t *__restrict__ cols;
  const uint32_t *__restrict__ outs;
  const int32_t *__restrict__ wts;
  const uint32_t *__restrict__ bias;
  const int32_t *__restrict__ in;
  int32_t *__restrict__ out;
  uint32_t ns, rs, nslices;
};

typedef short short2v __attribute__((ext_vector_type(2)));

// Word `word` (0..2047: a then b) of column cw = (ct << 10 | shift), rotated.
template <bool ROT>
__device__ __forceinline__ int32_t load_col(const int32_t *__restrict__ in, uint32_t cw,
                                            uint32_t word) {
  const int32_t *base = in + (size_t)(cw >> 10) * kCtWords;
  if constexpr (!ROT) {
    return base[word];
  } else {
    const uint32_t i = (word & (kN - 1)) + (cw & (kN - 1));
    const int32_t x = base[(word & kN) + (i & (kN - 1))];
    return i >= kN ? (int32_t)(0u - (uint32_t)x) : x;
  }
}

__device__ __forceinline__ int32_t dot2(uint32_t a, uint32_t b, int32_t c) {
  return __builtin_amdgcn_sdot2(__builtin_bit_cast(short2v, a), __builtin_bit_cast(short2v, b), c,
                                false);
}

__device__ __forceinline__ void store_out(int32_t *p, uint32_t v) {
  __builtin_nontemporal_store((int32_t)v, p);
}

template <int WPL, bool ROT>
__global__ __launch_bounds__(256) void k_cloud_dot2(CloudParams P) {
  const uint32_t slice = blockIdx.x % P.nslices;
  const uint32_t tile = blockIdx.x / P.nslices;
  const uint32_t wbase = slice * (256u * WPL);
  const bool is_b = wbase >= kN;
  uint32_t word[WPL];
#pragma unroll
  for (int u = 0; u < WPL; ++u) word[u] = wbase + (uint32_t)u * 256u + threadIdx.x;

  const uint32_t g_end = P.tile_off[tile + 1];
  for (uint32_t g = P.tile_off[tile]; g < g_end; ++g) {
    const GroupDev G = P.grp[g];
    int32_t alo[kOmax][WPL], ahi[kOmax][WPL];
#pragma unroll
    for (int o = 0; o < kOmax; ++o)
#pragma unroll
      for (int u = 0; u < WPL; ++u) alo[o][u] = ahi[o][u] = 0;

    for (uint32_t p0 = 0; p0 < G.nunits; p0 += kPairChunk) {
      uint32_t lo[kPairChunk][WPL], hi[kPairChunk][WPL];
#pragma unroll
      for (int pp = 0; pp < kPairChunk; ++pp) {
        if (p0 + pp < G.nunits) {
          const uint32_t c0 = P.cols[G.col_off + 2 * (p0 + pp)];
          const uint32_t c1 = P.cols[G.col_off + 2 * (p0 + pp) + 1];
#pragma unroll
          for (int u = 0; u < WPL; ++u) {
            const uint32_t x0 = (uint32_t)load_col<ROT>(P.in, c0, word[u]);
            const uint32_t x1 = (uint32_t)load_col<ROT>(P.in, c1, word[u]);
            lo[pp][u] = __builtin_amdgcn_perm(x1, x0, 0x05040100u);
            hi[pp][u] = __builtin_amdgcn_perm(x1 + 0x8000u, x0 + 0x8000u, 0x07060302u);
          }
        } else {
#pragma unroll
          for (int u = 0; u < WPL; ++u) lo[pp][u] = hi[pp][u] = 0;
        }
      }
#pragma unroll
      for (int o = 0; o < kOmax; ++o) {
        if (o < (int)G.nout) {
          const int32_t *wrow = P.wts + G.w_off + (uint32_t)o * G.nunits + p0;
#pragma unroll
          for (int pp = 0; pp < kPairChunk; ++pp) {
            if (p0 + pp < G.nunits) {
              const uint32_t w = (uint32_t)wrow[pp];
#pragma unroll
              for (int u = 0; u < WPL; ++u) {
                alo[o][u] = dot2(lo[pp][u], w, alo[o][u]);
                ahi[o][u] = dot2(hi[pp][u], w, ahi[o][u]);
              }
            }
          }
        }
      }
    }
#pragma unroll
    for (int o = 0; o < kOmax; ++o) {
      if (o < (int)G.nout) {
        const uint32_t slot = P.outs[G.out_off + o];
        const uint32_t bias = is_b ? P.bias[slot] : 0u;
        int32_t *dst = P.out + (size_t)slot * kCtWords;
#pragma unroll
        for (int u = 0; u < WPL; ++u) {
          uint32_t v = (uint32_t)alo[o][u] + ((uint32_t)ahi[o][u] << 16);
          if (is_b) {
            const uint32_t i = word[u] - kN;
            if (i < P.ns) v += bias;
            if (i >= P.rs) v = 0;
          }
          store_out(dst + word[u], v);
        }
      }
    }
  }
}

// Weights outside int16: one 32-bit multiply per column and output word.
template <int WPL, bool ROT>
__global__ __launch_bounds__(256) void k_cloud_wide(CloudParams P) {
  const uint32_t slice = blockIdx.x % P.nslices;
  const uint32_t tile = blockIdx.x / P.nslices;
  const uint32_t wbase = slice * (256u * WPL);
  const bool is_b = wbase >= kN;
  uint32_t word[WPL];
#pragma unroll
  for (int u = 0; u < WPL; ++u) word[u] = wbase + (uint32_t)u * 256u + threadIdx.x;

  const uint32_t g_end = P.tile_off[tile + 1];
  for (uint32_t g = P.tile_off[tile]; g < g_end; ++g) {
    const GroupDev G = P.grp[g];
    uint32_t acc[kOmax][WPL];
#pragma unroll
    for (int o = 0; o < kOmax; ++o)
#pragma unroll
      for (int u = 0; u < WPL; ++u) acc[o][u] = 0;
    for (uint32_t c0 = 0; c0 < G.nunits; c0 += kColChunk) {
      uint32_t x[kColChunk][WPL];
#pragma unroll
      for (int cc = 0; cc < kColChunk; ++cc) {
        if (c0 + cc < G.nunits) {
          const uint32_t cw = P.cols[G.col_off + c0 + cc];
#pragma unroll
          for (int u = 0; u < WPL; ++u) x[cc][u] = (uint32_t)load_col<ROT>(P.in, cw, word[u]);
        } else {
#pragma unroll
          for (int u = 0; u < WPL; ++u) x[cc][u] = 0;
        }
      }
#pragma unroll
      for (int o = 0; o < kOmax; ++o) {
        if (o < (int)G.nout) {
          const int32_t *wrow = P.wts + G.w_off + (uint32_t)o * G.nunits + c0;
#pragma unroll
          for (int cc = 0; cc < kColChunk; ++cc) {
            if (c0 + cc < G.nunits) {
              const uint32_t w = (uint32_t)wrow[cc];
#pragma unroll
              for (int u = 0; u < WPL; ++u) acc[o][u] += w * x[cc][u];
            }
          }
        }
      }
    }
#pragma unroll
    for (int o = 0; o < kOmax; ++o) {
      if (o < (int)G.nout) {
        const uint32_t slot = P.outs[G.out_off + o];
        const uint32_t bias = is_b ? P.bias[slot] : 0u;
        int32_t *dst = P.out + (size_t)slot * kCtWords;
#pragma unroll
        for (int u = 0; u < WPL; ++u) {
          uint32_t v = acc[o][u];
          if (is_b) {
            const uint32_t i = word[u] - kN;
            if (i < P.ns) v += bias;
            if (i >= P.rs) v = 0;
          }
          store_out(dst + word[u], v);
        }
      }
    }
  }
}

// var_o = sum over region-0 entries, in the model's visiting order, of
// (double)(int32)(w*w) * var_in (tLweAddMulTo's variance update; the rotated
// regions never carry a variance into the output, eval/idash.cpp:786, :835).
__global__ __launch_bounds__(256) void k_cloud_var(const uint32_t *__restrict__ var_ptr,
                                                   const uint32_t *__restrict__ var_ct,
                                                   const int32_t *__restrict__ var_w2,
                                                   const double *__restrict__ var_in,
                                                   double *__restrict__ var_out, uint32_t num_out) {
  const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= num_out) return;
  double v = 0.0;
  for (uint32_t e = var_ptr[o]; e < var_ptr[o + 1]; ++e)
    v = __dadd_rn(v, __dmul_rn((double)var_w2[e], var_in[var_ct[e]]));
  var_out[o] = v;
}

}  // namespace ihe

using namespace ihe;

struct ihe_band {
  ihe_ctx *ctx = nullptr;
  ihe_layout layout{};
  uint32_t num_in = 0, num_out = 0;
  int variant = IHE_VARIANT_VALU;
  int wpl = 1;
  uint32_t num_groups = 0, num_tiles = 0, num_cols = 0, nnz = 0, max_window = 0;
  void *dmem = nullptr;
  size_t dbytes = 0;
  GroupDev *d_grp = nullptr;
  uint32_t *d_tile_off = nullptr, *d_cols = nullptr, *d_outs = nullptr, *d_bias = nullptr;
  int32_t *d_wts = nullptr;
  uint32_t *d_var_ptr = nullptr, *d_var_ct = nullptr;
  int32_t *d_var_w2 = nullptr;
};

namespace {

struct HostBand {
  std::vector<GroupDev> grp;
  std::vector<uint32_t> tile_off, cols, outs, bias, var_ptr, var_ct;
  std::vector<int32_t> wts, var_w2;
  uint32_t max_window = 0;
};

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

int compile_band(const ihe_layout &L, uint32_t num_in, const ihe_model_csr &M, bool wide,
                 uint32_t tile_work, HostBand &H) {
  const uint32_t R = L.num_regions, RS = L.region_size;
  const uint32_t num_out = M.num_outputs;
  H.bias.assign(num_out, 0);
  H.var_ptr.assign(num_out + 1, 0);

  std::vector<std::pair<uint32_t, int32_t>> row;      // sorted (bidx, w) of the current output
  std::vector<uint32_t> gcols;                        // column list of the open group
  std::vector<std::vector<int32_t>> gw;               // weight rows of the open group
  std::vector<uint32_t> gouts;
  const uint32_t unit = wide ? 1 : 2;

  auto flush = [&]() {
    if (gouts.empty()) return;
    GroupDev G{};
    const uint32_t ncol = (uint32_t)gcols.size();
    const uint32_t nunits = (ncol + unit - 1) / unit;
    G.col_off = (uint32_t)H.cols.size();
    G.nunits = nunits;
    G.out_off = (uint32_t)H.outs.size();
    G.nout = (uint32_t)gouts.size();
    G.w_off = (uint32_t)H.wts.size();
    for (uint32_t k = 0; k < nunits * unit; ++k) {
      const uint32_t c = gcols[std::min(k, ncol - 1)];
      H.cols.push_back(((c / R) << 10) | ((c % R) * RS));
    }
    for (uint32_t o = 0; o < G.nout; ++o) {
      const auto &w = gw[o];
      for (uint32_t p = 0; p < nunits; ++p) {
        if (wide) {
          H.wts.push_back(w[p]);
        } else {
          const uint32_t lo = (uint16_t)(int16_t)w[2 * p];
          const uint32_t hi = (2 * p + 1 < ncol) ? (uint16_t)(int16_t)w[2 * p + 1] : 0u;
          H.wts.push_back((int32_t)(lo | (hi << 16)));
        }
      }
    }
    H.outs.insert(H.outs.end(), gouts.begin(), gouts.end());
    H.max_window = std::max(H.max_window, ncol);
    H.grp.push_back(G);
    gouts.clear();
    gw.clear();
  };

  for (uint32_t o = 0; o < num_out; ++o) {
    row.clear();
    bool has_const = false;
    for (uint32_t e = M.row_ptr[o]; e < M.row_ptr[o + 1]; ++e) {
      const uint32_t c = M.in_bidx[e];
      const int32_t w = M.weight[e];
      if (c == IHE_CONSTANT_BIDX) {
        if (has_const) return fail(IHE_ERR_ARG, "duplicate constant entry in output " + std::to_string(o));
        has_const = true;
        H.bias[o] = (uint32_t)w * kOne;
        continue;
      }
      if (c / R >= num_in)
        return fail(IHE_ERR_RANGE, "output " + std::to_string(o) + " references FeatIndex " +
                                       std::to_string(c / R) + " >= " + std::to_string(num_in));
      if (c / R >= kMaxCt) return fail(IHE_ERR_UNSUPPORTED, "FeatIndex beyond 2^22");
      if (!wide && (w < -32768 || w > 32767))
        return fail(IHE_ERR_UNSUPPORTED, "weight outside int16 for the dot2 variant");
      row.emplace_back(c, w);
      if (c % R == 0) {
        H.var_ct.push_back(c / R);
        H.var_w2.push_back((int32_t)((uint32_t)w * (uint32_t)w));
      }
    }
    H.var_ptr[o + 1] = (uint32_t)H.var_ct.size();
    std::sort(row.begin(), row.end());
    for (size_t k = 1; k < row.size(); ++k)
      if (row[k].first == row[k - 1].first)
        return fail(IHE_ERR_ARG, "duplicate input index in output " + std::to_string(o));
    bool same = !gouts.empty() && gouts.size() < (size_t)kOmax && gcols.size() == row.size();
    for (size_t k = 0; same && k < row.size(); ++k) same = gcols[k] == row[k].first;
    if (!same) {
      flush();
      gcols.clear();
      for (auto &p : row) gcols.push_back(p.first);
    }
    gouts.push_back(o);
    std::vector<int32_t> w(row.size());
    for (size_t k = 0; k < row.size(); ++k) w[k] = row[k].second;
    gw.push_back(std::move(w));
  }
  flush();

  // Tiles of roughly tile_work (outputs x units) each.
  H.tile_off.push_back(0);
  uint64_t acc = 0;
  for (uint32_t g = 0; g < H.grp.size(); ++g) {
    acc += (uint64_t)H.grp[g].nout * (H.grp[g].nunits + 1);
    if (acc >= tile_work) {
      H.tile_off.push_back(g + 1);
      acc = 0;
    }
  }
  if (H.tile_off.back() != H.grp.size()) H.tile_off.push_back((uint32_t)H.grp.size());
  return IHE_OK;
}

template <typename T>
T *carve(char *base, size_t &off, size_t count) {
  T *p = reinterpret_cast<T *>(base + off);
  off += align_up(sizeof(T) * std::max<size_t>(count, 1));
  return p;
}

int check_layout(const ihe_layout *L) {
  if (!L) return fail(IHE_ERR_ARG, "null layout");
  if (L->num_samples == 0 || L->num_samples > kN) return fail(IHE_ERR_ARG, "NUM_SAMPLES must be in [1, 1024]");
  if (L->num_regions != kN / L->num_samples || L->region_size != kN / L->num_regions)
    return fail(IHE_ERR_ARG, "layout must satisfy R = N / NS and RS = N / R (eval/idash.cpp:409-410)");
  return IHE_OK;
}

}  // namespace

namespace {

// Validates arguments, picks the kernel family and compiles the band on the host.
int plan_band(const ihe_layout *layout, uint32_t num_in_cts, const ihe_model_csr *model,
              int variant, HostBand &H, bool &wide) {
  if (!model) return fail(IHE_ERR_ARG, "null model");
  if (int rc = check_layout(layout)) return rc;
  if (model->num_outputs > 0 && (!model->row_ptr || (model->row_ptr[model->num_outputs] > 0 &&
                                                     (!model->in_bidx || !model->weight))))
    return fail(IHE_ERR_ARG, "null model arrays");
  if (variant != IHE_VARIANT_AUTO && variant != IHE_VARIANT_VALU && variant != IHE_VARIANT_WIDE &&
      variant != IHE_VARIANT_MFMA)
    return fail(IHE_ERR_ARG, "unknown variant");
  wide = variant == IHE_VARIANT_WIDE;
  if (variant == IHE_VARIANT_AUTO || variant == IHE_VARIANT_MFMA) {
    // Weights beyond int16 need the 32-bit multiply kernel.
    const uint32_t nnz = model->num_outputs ? model->row_ptr[model->num_outputs] : 0;
    for (uint32_t e = 0; e < nnz; ++e)
      if (model->in_bidx[e] != IHE_CONSTANT_BIDX &&
          (model->weight[e] < -32768 || model->weight[e] > 32767)) {
        wide = true;
        break;
      }
    if (variant == IHE_VARIANT_MFMA && wide)
      return fail(IHE_ERR_UNSUPPORTED, "MFMA variant needs weights within int16");
  }
  if (variant == IHE_VARIANT_MFMA) return fail(IHE_ERR_UNSUPPORTED, "MFMA variant not built");
  return compile_band(*layout, num_in_cts, *model, wide, 512, H);
}

void fill_info(const HostBand &H, const ihe_model_csr *model, bool wide, ihe_band_info *info) {
  info->num_outputs = model->num_outputs;
  info->num_groups = (uint32_t)H.grp.size();
  info->num_tiles = (uint32_t)H.tile_off.size() - 1;
  info->num_cols = (uint32_t)H.cols.size();
  info->nnz = model->num_outputs ? model->row_ptr[model->num_outputs] : 0;
  info->variant = wide ? IHE_VARIANT_WIDE : IHE_VARIANT_VALU;
  info->max_window = H.max_window;
  info->device_bytes = 0;
}

}  // namespace

extern "C" {

int ihe_band_plan(const ihe_layout *layout, uint32_t num_in_cts, const ihe_model_csr *model,
                  int variant, ihe_band_info *info) {
  if (!info) return fail(IHE_ERR_ARG, "null info");
  HostBand H;
  bool wide = false;
  if (int rc = plan_band(layout, num_in_cts, model, variant, H, wide)) return rc;
  fill_info(H, model, wide, info);
  return IHE_OK;
}

int ihe_band_create(ihe_ctx *ctx, const ihe_layout *layout, uint32_t num_in_cts,
                    const ihe_model_csr *model, int variant, ihe_band **out) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!out) return fail(IHE_ERR_ARG, "null argument");
  *out = nullptr;
  HostBand H;
  bool wide = false;
  if (int rc = plan_band(layout, num_in_cts, model, variant, H, wide)) return rc;

  ihe_band *b = new ihe_band();
  b->ctx = ctx;
  b->layout = *layout;
  b->num_in = num_in_cts;
  b->num_out = model->num_outputs;
  b->variant = wide ? IHE_VARIANT_WIDE : IHE_VARIANT_VALU;
  b->num_groups = (uint32_t)H.grp.size();
  b->num_tiles = (uint32_t)H.tile_off.size() - 1;
  b->num_cols = (uint32_t)H.cols.size();
  b->nnz = model->num_outputs ? model->row_ptr[model->num_outputs] : 0;
  b->max_window = H.max_window;

  size_t bytes = 0;
  bytes += align_up(sizeof(GroupDev) * std::max<size_t>(H.grp.size(), 1));
  bytes += align_up(4 * std::max<size_t>(H.tile_off.size(), 1));
  bytes += align_up(4 * std::max<size_t>(H.cols.size(), 1));
  bytes += align_up(4 * std::max<size_t>(H.outs.size(), 1));
  bytes += align_up(4 * std::max<size_t>(H.bias.size(), 1));
  bytes += align_up(4 * std::max<size_t>(H.wts.size(), 1));
  bytes += align_up(4 * std::max<size_t>(H.var_ptr.size(), 1));
  bytes += align_up(4 * std::max<size_t>(H.var_ct.size(), 1));
  bytes += align_up(4 * std::max<size_t>(H.var_w2.size(), 1));
  if (int rc = ihe_malloc(ctx, bytes, &b->dmem)) {
    delete b;
    return rc;
  }
  b->dbytes = bytes;
  std::vector<char> staging(bytes, 0);
  size_t off = 0;
  char *hb = staging.data();
  auto put = [&](auto *dst_dev_field, const auto &vec) {
    using T = typename std::decay_t<decltype(vec)>::value_type;
    T *h = carve<T>(hb, off, vec.size());
    if (!vec.empty()) std::memcpy(h, vec.data(), sizeof(T) * vec.size());
    *dst_dev_field = reinterpret_cast<T *>(static_cast<char *>(b->dmem) + (reinterpret_cast<char *>(h) - hb));
  };
  put(&b->d_grp, H.grp);
  put(&b->d_tile_off, H.tile_off);
  put(&b->d_cols, H.cols);
  put(&b->d_outs, H.outs);
  put(&b->d_bias, H.bias);
  put(&b->d_wts, H.wts);
  put(&b->d_var_ptr, H.var_ptr);
  put(&b->d_var_ct, H.var_ct);
  put(&b->d_var_w2, H.var_w2);
  if (int rc = ihe_memcpy_h2d(ctx, b->dmem, staging.data(), bytes)) {
    (void)ihe_free(ctx, b->dmem);
    delete b;
    return rc;
  }
  *out = b;
  return IHE_OK;
}

int ihe_band_destroy(ihe_band *band) {
  if (!band) return IHE_OK;
  (void)ihe_sync(band->ctx);
  (void)ihe_free(band->ctx, band->dmem);
  delete band;
  return IHE_OK;
}

int ihe_band_get_info(const ihe_band *band, ihe_band_info *info) {
  if (!band || !info) return fail(IHE_ERR_ARG, "null argument");
  info->num_outputs = band->num_out;
  info->num_groups = band->num_groups;
  info->num_tiles = band->num_tiles;
  info->num_cols = band->num_cols;
  info->nnz = band->nnz;
  info->variant = (uint32_t)band->variant;
  info->max_window = band->max_window;
  info->device_bytes = band->dbytes;
  return IHE_OK;
}

int ihe_cloud_score(ihe_ctx *ctx, const ihe_band *band, const int32_t *d_in,
                    const double *d_in_var, int32_t *d_out, double *d_out_var) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!band || band->ctx != ctx) return fail(IHE_ERR_ARG, "band does not belong to this context");
  if (band->num_out == 0) return IHE_OK;
  if (!d_out || (band->num_in > 0 && !d_in)) return fail(IHE_ERR_ARG, "null ciphertext buffer");
  if (d_out_var && band->num_in > 0 && !d_in_var) return fail(IHE_ERR_ARG, "null input variance");
  if (int rc = make_current(ctx)) return rc;

  CloudParams P;
  P.grp = band->d_grp;
  P.tile_off = band->d_tile_off;
  P.cols = band->d_cols;
  P.outs = band->d_outs;
  P.wts = band->d_wts;
  P.bias = band->d_bias;
  P.in = d_in;
  P.out = d_out;
  P.ns = band->layout.num_samples;
  P.rs = band->layout.region_size;
  const int wpl = band->wpl;
  P.nslices = kCtWords / (256u * (uint32_t)wpl);
  const bool rot = band->layout.num_regions > 1;
  const dim3 grid(band->num_tiles * P.nslices), block(256);
  {
    ScopedTimer t(ctx, IHE_TIMER_CLOUD);
    if (band->variant == IHE_VARIANT_WIDE) {
      if (rot) hipLaunchKernelGGL((k_cloud_wide<1, true>), grid, block, 0, ctx->stream, P);
      else hipLaunchKernelGGL((k_cloud_wide<1, false>), grid, block, 0, ctx->stream, P);
    } else {
      if (rot) hipLaunchKernelGGL((k_cloud_dot2<1, true>), grid, block, 0, ctx->stream, P);
      else hipLaunchKernelGGL((k_cloud_dot2<1, false>), grid, block, 0, ctx->stream, P);
    }
  }
  IHE_HIP(hipGetLastError());
  if (d_out_var) {
    ScopedTimer t(ctx, IHE_TIMER_CLOUD_VAR);
    const dim3 vgrid((band->num_out + 255) / 256);
    hipLaunchKernelGGL(k_cloud_var, vgrid, block, 0, ctx->stream, band->d_var_ptr, band->d_var_ct,
                       band->d_var_w2, d_in_var, d_out_var, band->num_out);
    IHE_HIP(hipGetLastError());
  }
  return IHE_OK;
}

}  // extern "C"
