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
// nearest-tag window), up to 12 outputs. A group lists its columns as packed
// words (byte offset of the ciphertext, plus the rotation shift for R > 1)
// padded to whole chunks, a chunk-major weight block and a 16-entry table of
// its output slots and biases. Groups are cut into tiles of roughly equal
// work, and tiles into segments whose input and output offsets fit 32 bits;
// one launch per segment, one workgroup per (tile, slice of the 2048 words).
//
// Exact int32 arithmetic on the VALU: with x = hi*2^16 + lo (lo, hi read as
// signed 16-bit after the recode hi = (x + 0x8000) >> 16), and a weight w
// within int16,
//   w*x = w*lo + 2^16 * w*hi   (mod 2^32)
// so two columns share one v_dot2_i32_i16 per half: a column pair costs two
// dot2 instructions per output word. All per-column and per-output addressing
// is scalar: one buffer descriptor per launch, the column's byte offset as the
// SGPR offset, the lane's word as the VGPR offset.
#include <algorithm>
#include <array>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <string>
#include <thread>
#include <unordered_map>

#include "ihe_internal.h"
#include "ihe_band.h"

namespace ihe {

constexpr int kOmax = 12;                    // outputs per group (register-blocked)
constexpr int kOtab = 16;                    // output-table entries per group
constexpr int kDot2ChunkCols = 16;           // 8 column pairs per chunk (dot2)
constexpr int kWideChunkCols = 8;            // 8 columns per chunk (32-bit weights)
constexpr int kRsrcWord3 = 0x00020000;       // gfx9 raw buffer descriptor, dword 3

#define IHE_CONSTANT __attribute__((address_space(4)))
// Cache policy of the output stores: nt (2, streaming) with sc1 (16, device
// scope), as for the MFMA kernels.
#ifndef IHE_STORE_POLICY
#define IHE_STORE_POLICY 18
#endif

// Uniform metadata is read through the constant address space so the loads
// become scalar (SMEM) loads.
template <typename T>
__device__ __forceinline__ const IHE_CONSTANT T *cptr(const T *p) {
  return (const IHE_CONSTANT T *)p;
}

struct GroupDev {
  uint32_t col_off;  // first column word (multiple of 16)
  uint32_t nchunks;  // chunks of columns
  uint32_t nout;     // outputs in the group (<= kOmax)
  uint32_t w_off;    // weights of chunk c, output o at w_off + (c * nout + o) * 8
  uint32_t otab;     // first entry of the group's output tables (multiple of 16)
  uint32_t pad0, pad1, pad2;
};

struct CloudParams {
  const GroupDev *grp;
  const uint32_t *tile_off;
  const uint32_t *cols;
  const int32_t *wts;
  const uint32_t *oslot;  // output byte offset (slot * 8192) relative to out
  const uint32_t *obias;  // bias * 2^18
  const int32_t *in;      // segment input base
  int32_t *out;           // segment output base
  uint32_t in_bytes, out_bytes;
  uint32_t ns, rs, nslices, tile_base;
};

typedef short short2v __attribute__((ext_vector_type(2)));
typedef uint32_t u32x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ GroupDev load_group(const GroupDev *grp, uint32_t g) {
  return __builtin_bit_cast(GroupDev, *(const IHE_CONSTANT u32x8 *)(grp + g));
}

__device__ __forceinline__ int32_t dot2(uint32_t a, uint32_t b, int32_t c) {
  return __builtin_amdgcn_sdot2(__builtin_bit_cast(short2v, a), __builtin_bit_cast(short2v, b), c,
                                false);
}

// Per-lane addressing of the words a lane owns (word = 0..2047, a then b).
template <int WPL>
struct Lane {
  uint32_t voff[WPL];    // word * 4
  uint32_t wlo[WPL];     // word & (N-1)
  uint32_t wpoly4[WPL];  // (word & N) * 4
};

// Column word cw = byte offset of the ciphertext | shift << 2 (shift < N).
// rho_shift(P)[i] = P[i + shift], negated past N.
template <int WPL, bool ROT, int NC>
__device__ __forceinline__ void load_chunk(__amdgpu_buffer_rsrc_t in_r,
                                           const IHE_CONSTANT uint32_t *cols, uint32_t off,
                                           const Lane<WPL> &L, uint32_t (&x)[NC][WPL]) {
  uint32_t cw[NC];
  if constexpr (NC == 16) {
    const u32x16 v = *(const IHE_CONSTANT u32x16 *)(cols + off);
#pragma unroll
    for (int k = 0; k < NC; ++k) cw[k] = v[k];
  } else {
    const u32x8 v = *(const IHE_CONSTANT u32x8 *)(cols + off);
#pragma unroll
    for (int k = 0; k < NC; ++k) cw[k] = v[k];
  }
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    if constexpr (!ROT) {
#pragma unroll
      for (int u = 0; u < WPL; ++u)
        x[k][u] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(in_r, L.voff[u], (int)cw[k], 0);
    } else {
      const uint32_t soff = cw[k] & ~0x1FFFu;
      const uint32_t sh = (cw[k] >> 2) & (kN - 1);
#pragma unroll
      for (int u = 0; u < WPL; ++u) {
        const uint32_t t = L.wlo[u] + sh;
        const uint32_t v = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(
            in_r, ((t & (kN - 1)) << 2) + L.wpoly4[u], (int)soff, 0);
        x[k][u] = t >= kN ? 0u - v : v;
      }
    }
  }
}

template <int WPL, bool ROT, bool WIDE>
__global__ __launch_bounds__(256) void k_cloud(CloudParams P) {
  constexpr int NC = WIDE ? kWideChunkCols : kDot2ChunkCols;
  constexpr int NP = NC / 2;  // column pairs per chunk (dot2)
  const uint32_t slice = blockIdx.x % P.nslices;
  const uint32_t tile = P.tile_base + blockIdx.x / P.nslices;
  const uint32_t wbase = slice * (256u * WPL);
  const bool is_b = wbase >= kN;
  const __amdgpu_buffer_rsrc_t in_r =
      __builtin_amdgcn_make_buffer_rsrc((void *)P.in, (short)0, (int)P.in_bytes, kRsrcWord3);
  const __amdgpu_buffer_rsrc_t out_r =
      __builtin_amdgcn_make_buffer_rsrc((void *)P.out, (short)0, (int)P.out_bytes, kRsrcWord3);
  Lane<WPL> L;
  uint32_t bias_mask[WPL], keep_mask[WPL];
#pragma unroll
  for (int u = 0; u < WPL; ++u) {
    const uint32_t word = wbase + (uint32_t)u * 256u + threadIdx.x;
    L.voff[u] = word * 4u;
    L.wlo[u] = word & (kN - 1);
    L.wpoly4[u] = (word & kN) * 4u;
    // b polynomial: samples that get the bias, words that stay visible
    const uint32_t i = word - kN;
    bias_mask[u] = i < P.ns ? 0xFFFFFFFFu : 0u;
    keep_mask[u] = i >= P.rs ? 0u : 0xFFFFFFFFu;
  }

  const IHE_CONSTANT uint32_t *cols = cptr(P.cols);
  const IHE_CONSTANT uint32_t *wts = cptr((const uint32_t *)P.wts);
  const IHE_CONSTANT uint32_t *toff = cptr(P.tile_off);

  uint32_t g = toff[tile];
  const uint32_t g_end = toff[tile + 1];
  if (g >= g_end) return;

  // The chunk being computed was loaded into xn one step earlier. A finished
  // group's words stay in alo[] and are stored during the next iteration,
  // before that iteration issues its loads: stores share the in-order vector
  // memory counter with loads, so storing late keeps the wait for the next
  // chunk from also waiting on fresh stores.
  uint32_t xn[NC][WPL];
  GroupDev G = load_group(P.grp, g);
  uint32_t c = 0;
  load_chunk<WPL, ROT, NC>(in_r, cols, G.col_off, L, xn);

  uint32_t alo[kOmax][WPL], ahi[kOmax][WPL];
#pragma unroll
  for (int o = 0; o < kOmax; ++o)
#pragma unroll
    for (int u = 0; u < WPL; ++u) alo[o][u] = ahi[o][u] = 0;
  uint32_t pend_nout = 0;
  u32x16 pend_slots = {};

  while (true) {
    // operands of this chunk
    uint32_t lo[WIDE ? NC : NP][WPL], hi[WIDE ? 1 : NP][WPL];
    if constexpr (WIDE) {
#pragma unroll
      for (int k = 0; k < NC; ++k)
#pragma unroll
        for (int u = 0; u < WPL; ++u) lo[k][u] = xn[k][u];
    } else {
#pragma unroll
      for (int pp = 0; pp < NP; ++pp)
#pragma unroll
        for (int u = 0; u < WPL; ++u) {
          const uint32_t x0 = xn[2 * pp][u], x1 = xn[2 * pp + 1][u];
          lo[pp][u] = __builtin_amdgcn_perm(x1, x0, 0x05040100u);
          hi[pp][u] = __builtin_amdgcn_perm(x1 + 0x8000u, x0 + 0x8000u, 0x07060302u);
        }
    }
    // stores of the group finished last iteration
    if (pend_nout) {
#pragma unroll
      for (int o = 0; o < kOmax; ++o) {
        if (o < (int)pend_nout) {
#pragma unroll
          for (int u = 0; u < WPL; ++u) {
            __builtin_amdgcn_raw_buffer_store_b32((int)alo[o][u], out_r, L.voff[u],
                                                  (int)pend_slots[o], IHE_STORE_POLICY);
            alo[o][u] = ahi[o][u] = 0;
          }
        }
      }
      pend_nout = 0;
    }
    // advance and start the next chunk's loads before computing this one
    const bool last_chunk = c + 1 >= G.nchunks;
    const GroupDev Gc = G;
    const uint32_t cc = c;
    bool more = true;
    if (!last_chunk) {
      ++c;
    } else if (g + 1 < g_end) {
      ++g;
      G = load_group(P.grp, g);
      c = 0;
    } else {
      more = false;
    }
    if (more) load_chunk<WPL, ROT, NC>(in_r, cols, G.col_off + (uint32_t)NC * c, L, xn);

    // outputs of this chunk; the next output's weights load while this one computes
    const IHE_CONSTANT uint32_t *wc = wts + Gc.w_off + cc * Gc.nout * 8u;
    u32x8 wnext = *(const IHE_CONSTANT u32x8 *)wc;
#pragma unroll
    for (int o = 0; o < kOmax; ++o) {
      if (o < (int)Gc.nout) {
        const u32x8 w = wnext;
        if (o + 1 < (int)Gc.nout) wnext = *(const IHE_CONSTANT u32x8 *)(wc + 8 * (o + 1));
#pragma unroll
        for (int k = 0; k < 8; ++k)
#pragma unroll
          for (int u = 0; u < WPL; ++u) {
            if constexpr (WIDE) {
              alo[o][u] += w[k] * lo[k][u];
            } else {
              alo[o][u] = (uint32_t)dot2(lo[k][u], w[k], (int32_t)alo[o][u]);
              ahi[o][u] = (uint32_t)dot2(hi[k][u], w[k], (int32_t)ahi[o][u]);
            }
          }
      }
    }
    if (last_chunk) {
      const u32x16 bias = *(const IHE_CONSTANT u32x16 *)(cptr(P.obias) + Gc.otab);
      pend_slots = *(const IHE_CONSTANT u32x16 *)(cptr(P.oslot) + Gc.otab);
      pend_nout = Gc.nout;
#pragma unroll
      for (int o = 0; o < kOmax; ++o) {
        if (o < (int)Gc.nout) {
#pragma unroll
          for (int u = 0; u < WPL; ++u) {
            uint32_t v = WIDE ? alo[o][u] : alo[o][u] + (ahi[o][u] << 16);
            if (is_b) v = (v + (bias[o] & bias_mask[u])) & keep_mask[u];
            alo[o][u] = v;
          }
        }
      }
    }
    if (!more) break;
  }
  // the last group's stores
#pragma unroll
  for (int o = 0; o < kOmax; ++o) {
    if (o < (int)pend_nout) {
#pragma unroll
      for (int u = 0; u < WPL; ++u)
        __builtin_amdgcn_raw_buffer_store_b32((int)alo[o][u], out_r, L.voff[u], (int)pend_slots[o],
                                              IHE_STORE_POLICY);
    }
  }
}

// ---------------------------------------------------------------------------
// LDS-staged kernel for narrow windows (<= 32 columns per group, int16
// weights). A workgroup owns one tile (consecutive groups) and one 256-word
// slice. It first stages the tile's sliding window of input columns in LDS,
// each column once (consecutive targets share most of their window), and
// copies the tile's metadata (group headers, output slots, biases, weights)
// into LDS. The compute phase then reads only LDS: column words are per-lane
// LDS reads, weights are broadcast LDS reads that complete in order, so the
// next output's weights are fetched while the current one computes, and the
// output stores never queue behind vector loads.
constexpr uint32_t kStageSlice = 256;    // words per workgroup slice
constexpr uint32_t kStageMaxCols = 48;   // staged columns per tile
constexpr uint32_t kStageMaxMeta = 2048; // metadata words per tile

struct TileDev {
  uint32_t stage_off;  // first column word in stage[] (multiple of 16)
  uint32_t nstage;     // columns staged into LDS
  uint32_t meta_off;   // first metadata word in meta[]
  uint32_t nmeta;      // metadata words
  uint32_t ngrp;       // groups
  uint32_t pad0, pad1, pad2;
};

// Metadata blob of a tile (32-bit words):
//   header of group i at 8*i: {first staged column, chunks, outputs,
//                              weights offset, slot/bias offset}
//   slots[nout] then biases[nout] at the slot/bias offset
//   weights [chunk][output][8 packed int16 pairs] at the weights offset
struct StagedParams {
  const TileDev *tiles;
  const uint32_t *stage;
  const uint32_t *meta;
  const int32_t *in;
  int32_t *out;
  uint32_t in_bytes, out_bytes;
  uint32_t ns, rs, tile_base, meta_lds;  // meta_lds: word offset of the metadata in LDS
};

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <bool ROT>
__device__ __forceinline__ uint32_t load_word(__amdgpu_buffer_rsrc_t in_r, uint32_t cw,
                                              uint32_t voff, uint32_t wlo, uint32_t wpoly4) {
  if constexpr (!ROT) {
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(in_r, voff, (int)cw, 0);
  } else {
    const uint32_t t = wlo + ((cw >> 2) & (kN - 1));
    const uint32_t v = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(
        in_r, ((t & (kN - 1)) << 2) + wpoly4, (int)(cw & ~0x1FFFu), 0);
    return t >= kN ? 0u - v : v;
  }
}

__device__ __forceinline__ uint32_t rfl(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

template <bool ROT>
__global__ __launch_bounds__(256) void k_cloud_staged(StagedParams P) {
  extern __shared__ uint32_t lds[];
  const uint32_t slice = blockIdx.x & 7u;
  const uint32_t tile = P.tile_base + (blockIdx.x >> 3);
  const uint32_t tid = threadIdx.x;
  const uint32_t word = slice * kStageSlice + tid;
  const bool is_b = slice >= 4;
  const __amdgpu_buffer_rsrc_t in_r =
      __builtin_amdgcn_make_buffer_rsrc((void *)P.in, (short)0, (int)P.in_bytes, kRsrcWord3);
  const __amdgpu_buffer_rsrc_t out_r =
      __builtin_amdgcn_make_buffer_rsrc((void *)P.out, (short)0, (int)P.out_bytes, kRsrcWord3);
  const TileDev T = __builtin_bit_cast(TileDev, *(const IHE_CONSTANT u32x8 *)(P.tiles + tile));
  const uint32_t voff = word * 4u, wlo = word & (kN - 1), wpoly4 = (word & kN) * 4u;
  const uint32_t i_b = word - kN;
  const uint32_t bias_mask = i_b < P.ns ? 0xFFFFFFFFu : 0u;
  const uint32_t keep_mask = i_b >= P.rs ? 0u : 0xFFFFFFFFu;
  uint32_t *meta = lds + P.meta_lds;

  // phase 1: stage the tile's columns (all loads in flight together) and metadata
  {
    uint32_t x[kStageMaxCols / 16][16];
#pragma unroll
    for (int b = 0; b < (int)(kStageMaxCols / 16); ++b) {
      if ((uint32_t)b * 16u < T.nstage) {
        const u32x16 cw = *(const IHE_CONSTANT u32x16 *)(cptr(P.stage) + T.stage_off + 16 * b);
#pragma unroll
        for (int k = 0; k < 16; ++k) x[b][k] = load_word<ROT>(in_r, cw[k], voff, wlo, wpoly4);
      }
    }
    for (uint32_t i = tid; i < T.nmeta; i += kStageSlice) meta[i] = P.meta[T.meta_off + i];
#pragma unroll
    for (int b = 0; b < (int)(kStageMaxCols / 16); ++b) {
      if ((uint32_t)b * 16u < T.nstage) {
#pragma unroll
        for (int k = 0; k < 16; ++k)
          if ((uint32_t)(b * 16 + k) < T.nstage) lds[(b * 16 + k) * kStageSlice + tid] = x[b][k];
      }
    }
  }
  __syncthreads();

  // phase 2: groups of the tile, from LDS only
  for (uint32_t gi = 0; gi < T.ngrp; ++gi) {
    const uint32_t s0 = rfl(meta[8 * gi]), nchunks = rfl(meta[8 * gi + 1]);
    const uint32_t nout = rfl(meta[8 * gi + 2]), wo = rfl(meta[8 * gi + 3]);
    const uint32_t oo = rfl(meta[8 * gi + 4]);
    uint32_t alo[kOmax], ahi[kOmax];
#pragma unroll
    for (int o = 0; o < kOmax; ++o) alo[o] = ahi[o] = 0;
    for (uint32_t c = 0; c < nchunks; ++c) {
      const uint32_t *col = lds + (s0 + 16 * c) * kStageSlice + tid;
      uint32_t lo[8], hi[8];
#pragma unroll
      for (int pp = 0; pp < 8; ++pp) {
        const uint32_t x0 = col[(2 * pp) * kStageSlice], x1 = col[(2 * pp + 1) * kStageSlice];
        lo[pp] = __builtin_amdgcn_perm(x1, x0, 0x05040100u);
        hi[pp] = __builtin_amdgcn_perm(x1 + 0x8000u, x0 + 0x8000u, 0x07060302u);
      }
      const uint32_t *wc = meta + wo + c * nout * 8u;
#pragma unroll
      for (int o = 0; o < kOmax; ++o) {
        if (o < (int)nout) {
          const u32x4 w0 = *(const u32x4 *)(wc + 8 * o);
          const u32x4 w1 = *(const u32x4 *)(wc + 8 * o + 4);
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            alo[o] = (uint32_t)dot2(lo[k], w0[k], (int32_t)alo[o]);
            ahi[o] = (uint32_t)dot2(hi[k], w0[k], (int32_t)ahi[o]);
          }
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            alo[o] = (uint32_t)dot2(lo[4 + k], w1[k], (int32_t)alo[o]);
            ahi[o] = (uint32_t)dot2(hi[4 + k], w1[k], (int32_t)ahi[o]);
          }
        }
      }
    }
#pragma unroll
    for (int o = 0; o < kOmax; ++o) {
      if (o < (int)nout) {
        const uint32_t slot = rfl(meta[oo + o]);
        const uint32_t bias = rfl(meta[oo + nout + o]);
        uint32_t v = alo[o] + (ahi[o] << 16);
        if (is_b) v = (v + (bias & bias_mask)) & keep_mask;
        __builtin_amdgcn_raw_buffer_store_b32((int)v, out_r, voff, (int)slot, IHE_STORE_POLICY);
      }
    }
  }
}

}  // namespace ihe

using namespace ihe;


namespace {

struct HostGroup {
  uint32_t ncol;                 // real columns
  uint32_t col_begin;            // into HostBand::gcol (absolute big indices)
  uint32_t out_begin, nout;      // into HostBand::gout
  uint32_t w_begin;              // into HostBand::gw (nout rows of ncol weights)
  uint32_t ct_min, ct_max;       // FeatIndex range of the real columns
};

struct HostBand {
  bool wide = false;
  int mode = kModeDirect;
  // direct layout
  std::vector<GroupDev> grp;
  std::vector<uint32_t> tile_off, cols, oslot, obias;
  std::vector<int32_t> wts;
  // staged layout
  std::vector<TileDev> tiles;
  std::vector<uint32_t> stage, meta;
  uint32_t max_stage = 0, max_meta = 0;
  // MFMA layout
  std::vector<MTileDev> mtiles;
  std::vector<uint32_t> kcols, rowc;
  std::vector<int4v> adata;
  MfmaShape shape;
  std::vector<MfmaShape> candidates;  // feasible MFMA shapes, for the autotuner
  // all layouts
  std::vector<Segment> segments;
  std::vector<uint32_t> var_ptr, var_col, var_base, bias;
  std::vector<int32_t> var_w2;
  bool var_wide_w = false;  // some |w| > 32768: variance entries stay unpacked
  uint32_t max_window = 0, max_bidx = 0;
  bool has_cols = false;
  // grouping
  std::vector<HostGroup> hg;
  std::vector<uint32_t> gcol, gout;
  std::vector<int32_t> gw;
};

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Outputs (in slot order) with identical sorted column lists form groups of
// up to kOmax outputs. Validates the model like the reference's lookups do.
int build_groups(const ihe_layout &L, uint32_t num_in, const ihe_model_csr &M, bool wide,
                 HostBand &H) {
  const uint32_t R = L.num_regions;
  const uint32_t num_out = M.num_outputs;
  H.bias.assign(num_out, 0);
  H.var_ptr.assign(num_out + 1, 0);
  std::vector<std::pair<uint32_t, int32_t>> row;
  for (uint32_t o = 0; o < num_out; ++o) {
    row.clear();
    bool has_const = false;
    for (uint32_t e = M.row_ptr[o]; e < M.row_ptr[o + 1]; ++e) {
      const uint32_t c = M.in_bidx[e];
      const int32_t w = M.weight[e];
      if (c == IHE_CONSTANT_BIDX) {
        if (has_const) return fail(IHE_ERR_ARG, "duplicate constant entry in output " + std::to_string(o));
        has_const = true;
        H.bias[o] = (uint32_t)w * kOne;  // coeff * ONE_IN_T32, an int32 product that may wrap
        continue;
      }
      if (c / R >= num_in)
        return fail(IHE_ERR_RANGE, "output " + std::to_string(o) + " references FeatIndex " +
                                       std::to_string(c / R) + " >= " + std::to_string(num_in));
      if (!wide && (w < -32768 || w > 32767))
        return fail(IHE_ERR_UNSUPPORTED, "weight outside int16 for the dot2 variant");
      row.emplace_back(c, w);
      // tLweAddMulTo's variance update, (int32)(w*w) * var; only region-0 columns
      // reach the output (the rotated regions carry none, eval/idash.cpp:786, :835)
      H.var_col.push_back(c);
      H.var_w2.push_back((int32_t)((uint32_t)w * (uint32_t)w));
      if (w < -32768 || w > 32768) H.var_wide_w = true;
      H.max_bidx = std::max(H.max_bidx, c);
      H.has_cols = true;
    }
    H.var_ptr[o + 1] = (uint32_t)H.var_col.size();
    std::sort(row.begin(), row.end());
    for (size_t k = 1; k < row.size(); ++k)
      if (row[k].first == row[k - 1].first)
        return fail(IHE_ERR_ARG, "duplicate input index in output " + std::to_string(o));
    bool same = !H.hg.empty() && H.hg.back().nout < (uint32_t)kOmax && H.hg.back().ncol == row.size();
    if (same) {
      const uint32_t *gc = H.gcol.data() + H.hg.back().col_begin;
      for (size_t k = 0; same && k < row.size(); ++k) same = gc[k] == row[k].first;
    }
    if (!same) {
      HostGroup hg{};
      hg.ncol = (uint32_t)row.size();
      hg.col_begin = (uint32_t)H.gcol.size();
      hg.out_begin = (uint32_t)H.gout.size();
      hg.w_begin = (uint32_t)H.gw.size();
      hg.ct_min = row.empty() ? UINT32_MAX : row.front().first / R;
      hg.ct_max = row.empty() ? 0 : row.back().first / R;
      for (auto &p : row) H.gcol.push_back(p.first);
      H.hg.push_back(hg);
      if (hg.ncol > 0 && hg.ct_max - hg.ct_min >= kSegCts)
        return fail(IHE_ERR_UNSUPPORTED, "an output spans more than 4 GiB of input ciphertexts");
    }
    HostGroup &hg = H.hg.back();
    H.gout.push_back(o);
    hg.nout++;
    for (auto &p : row) H.gw.push_back(p.second);
    H.max_window = std::max(H.max_window, (uint32_t)row.size());
  }
  return IHE_OK;
}

// Packs column big index c as (byte offset of its ciphertext relative to the
// segment's first one) | (rotation shift << 2).
uint32_t col_word(uint32_t c, uint32_t R, uint32_t RS, uint32_t in_base) {
  return (((c / R) - in_base) << 13) | (((c % R) * RS) << 2);
}

// Grows segments over consecutive units (tiles) while their input and output
// ranges stay addressable with 32-bit byte offsets.
struct SegmentBuilder {
  std::vector<Segment> &out;
  uint32_t first = 0, in_lo = UINT32_MAX, in_hi = 0, out_lo = UINT32_MAX, out_hi = 0;
  // returns the segment bases the unit was placed in
  void add(uint32_t unit, uint32_t ilo, uint32_t ihi, uint32_t olo, uint32_t ohi) {
    const uint32_t nil = std::min(in_lo, ilo), nih = std::max(in_hi, ihi);
    const uint32_t nol = std::min(out_lo, olo), noh = std::max(out_hi, ohi);
    const bool in_ok = nil == UINT32_MAX || nih - nil < kSegCts;
    // outputs stay kSegPadRows below kSegCts: a tile's padding rows store at
    // the drop offset plus their row offset, which must stay below 2^32
    if (unit > first && (!in_ok || noh - nol >= kSegCts - kSegPadRows)) {
      close(unit);
      in_lo = ilo, in_hi = ihi, out_lo = olo, out_hi = ohi;
    } else {
      in_lo = nil, in_hi = nih, out_lo = nol, out_hi = noh;
    }
  }
  void close(uint32_t unit_end) {
    if (unit_end > first) out.push_back({first, unit_end, in_lo == UINT32_MAX ? 0 : in_lo, out_lo});
    first = unit_end;
    in_lo = out_lo = UINT32_MAX;
    in_hi = out_hi = 0;
  }
};

// Direct layout: tiles of roughly tile_work units, columns read from HBM/L2 per group.
void layout_direct(const ihe_layout &L, uint32_t tile_work, HostBand &H) {
  const uint32_t R = L.num_regions, RS = L.region_size;
  const uint32_t chunk_cols = H.wide ? kWideChunkCols : kDot2ChunkCols;
  H.tile_off.assign(1, 0);
  std::vector<std::array<uint32_t, 4>> tile_rng;  // in lo/hi, out lo/hi per tile
  uint64_t acc = 0;
  std::array<uint32_t, 4> rng = {UINT32_MAX, 0, UINT32_MAX, 0};
  for (uint32_t g = 0; g < H.hg.size(); ++g) {
    const HostGroup &hg = H.hg[g];
    rng[0] = std::min(rng[0], hg.ct_min);
    rng[1] = std::max(rng[1], hg.ct_max);
    rng[2] = std::min(rng[2], H.gout[hg.out_begin]);
    rng[3] = std::max(rng[3], H.gout[hg.out_begin + hg.nout - 1]);
    const uint32_t nchunks = std::max<uint32_t>(1, (hg.ncol + chunk_cols - 1) / chunk_cols);
    acc += (uint64_t)hg.nout * (nchunks * 8 + 1);
    // a tile must fit one segment; a large spread starts a new tile
    const bool spread = rng[1] != 0 && rng[0] != UINT32_MAX && rng[1] - rng[0] >= kSegCts / 2;
    if (acc >= tile_work || g + 1 == H.hg.size() || spread) {
      H.tile_off.push_back(g + 1);
      tile_rng.push_back(rng);
      acc = 0;
      rng = {UINT32_MAX, 0, UINT32_MAX, 0};
    }
  }
  SegmentBuilder sb{H.segments};
  for (uint32_t t = 0; t + 1 < H.tile_off.size(); ++t)
    sb.add(t, tile_rng[t][0], tile_rng[t][1], tile_rng[t][2], tile_rng[t][3]);
  sb.close((uint32_t)H.tile_off.size() - 1);

  for (const Segment &s : H.segments) {
    for (uint32_t t = s.tile_begin; t < s.tile_end; ++t) {
      for (uint32_t g = H.tile_off[t]; g < H.tile_off[t + 1]; ++g) {
        const HostGroup &hg = H.hg[g];
        GroupDev G{};
        const uint32_t ncol_pad = std::max<uint32_t>(16, (hg.ncol + 15) / 16 * 16);
        G.col_off = (uint32_t)H.cols.size();
        G.nchunks = ncol_pad / chunk_cols;
        G.nout = hg.nout;
        G.w_off = (uint32_t)H.wts.size();
        G.otab = (uint32_t)H.oslot.size();
        for (uint32_t k = 0; k < ncol_pad; ++k) {
          // padding repeats the last real column (weight 0) or reads the
          // segment's first ciphertext
          const uint32_t c =
              hg.ncol ? H.gcol[hg.col_begin + std::min(k, hg.ncol - 1)] : s.in_base * R;
          H.cols.push_back(col_word(c, R, RS, s.in_base));
        }
        for (uint32_t ch = 0; ch < G.nchunks; ++ch) {
          for (uint32_t o = 0; o < hg.nout; ++o) {
            const int32_t *w = H.gw.data() + hg.w_begin + (size_t)o * hg.ncol;
            for (uint32_t k = 0; k < 8; ++k) {
              if (H.wide) {
                const uint32_t col = ch * 8 + k;
                H.wts.push_back(col < hg.ncol ? w[col] : 0);
              } else {
                const uint32_t c0 = ch * 16 + 2 * k, c1 = c0 + 1;
                const uint32_t lo = c0 < hg.ncol ? (uint16_t)(int16_t)w[c0] : 0u;
                const uint32_t hi = c1 < hg.ncol ? (uint16_t)(int16_t)w[c1] : 0u;
                H.wts.push_back((int32_t)(lo | (hi << 16)));
              }
            }
          }
        }
        for (uint32_t o = 0; o < (uint32_t)kOtab; ++o) {
          const uint32_t slot = o < hg.nout ? H.gout[hg.out_begin + o] : s.out_base;
          H.oslot.push_back((slot - s.out_base) * kCtWords * 4u);
          H.obias.push_back(o < hg.nout ? H.bias[slot] : 0u);
        }
        H.grp.push_back(G);
      }
    }
  }
}

// Staged layout: tiles whose staged column list (the union of their groups'
// windows, shared where consecutive windows overlap) fits the LDS budget.
void layout_staged(const ihe_layout &L, uint32_t tile_outputs, HostBand &H) {
  const uint32_t R = L.num_regions, RS = L.region_size;
  struct TileHost {
    std::vector<uint32_t> S;                       // staged column big indices
    std::vector<std::pair<uint32_t, uint32_t>> gs; // (group, first staged column)
    uint32_t nstage_pad = 0, nmeta = 0, nout = 0;
    uint32_t in_lo = UINT32_MAX, in_hi = 0, out_lo = UINT32_MAX, out_hi = 0;
  };
  std::vector<TileHost> tiles;
  TileHost cur;
  std::unordered_map<uint32_t, uint32_t> pos;  // big index -> position in cur.S
  auto meta_of = [](const HostGroup &hg, uint32_t nchunks) {
    return 8u + (2 * hg.nout + 3) / 4 * 4 + nchunks * hg.nout * 8u;
  };
  // Where a group's sorted columns go in the staged list: inside it, or
  // overlapping its tail (a sliding window), or appended whole. Returns the
  // first staged column and how many columns would be appended.
  auto probe = [&](const HostGroup &hg, const TileHost &T) -> std::pair<uint32_t, uint32_t> {
    if (hg.ncol == 0) return {0u, 0u};
    const uint32_t *C = H.gcol.data() + hg.col_begin;
    auto it = pos.find(C[0]);
    if (it != pos.end()) {
      const uint32_t p0 = it->second;
      uint32_t m = 0;
      while (m < hg.ncol && p0 + m < T.S.size() && T.S[p0 + m] == C[m]) ++m;
      if (m == hg.ncol || p0 + m == T.S.size()) return {p0, hg.ncol - m};
    }
    return {(uint32_t)T.S.size(), hg.ncol};
  };
  auto fits = [&](const HostGroup &hg, const TileHost &T) {
    const uint32_t nchunks = std::max<uint32_t>(1, (hg.ncol + 15) / 16);
    const auto [s0, added] = probe(hg, T);
    const uint32_t nstage = std::max({T.nstage_pad, (uint32_t)T.S.size() + added, s0 + 16 * nchunks});
    return nstage <= kStageMaxCols && T.nmeta + meta_of(hg, nchunks) <= kStageMaxMeta &&
           T.nout + hg.nout <= tile_outputs;
  };
  auto place = [&](const HostGroup &hg, TileHost &T) {
    const auto [s0, added] = probe(hg, T);
    const uint32_t *C = H.gcol.data() + hg.col_begin;
    for (uint32_t k = hg.ncol - added; k < hg.ncol; ++k) {
      pos[C[k]] = (uint32_t)T.S.size();
      T.S.push_back(C[k]);
    }
    return s0;
  };
  for (uint32_t g = 0; g < H.hg.size(); ++g) {
    const HostGroup &hg = H.hg[g];
    if (!cur.gs.empty() && !fits(hg, cur)) {
      tiles.push_back(std::move(cur));
      cur = TileHost();
      pos.clear();
    }
    const uint32_t nchunks = std::max<uint32_t>(1, (hg.ncol + 15) / 16);
    const uint32_t s0 = place(hg, cur);
    cur.gs.emplace_back(g, s0);
    cur.nstage_pad = std::max(cur.nstage_pad, std::max<uint32_t>((uint32_t)cur.S.size(), s0 + 16 * nchunks));
    cur.nmeta += meta_of(hg, nchunks);
    cur.nout += hg.nout;
    cur.in_lo = std::min(cur.in_lo, hg.ct_min);
    cur.in_hi = std::max(cur.in_hi, hg.ct_max);
    cur.out_lo = std::min(cur.out_lo, H.gout[hg.out_begin]);
    cur.out_hi = std::max(cur.out_hi, H.gout[hg.out_begin + hg.nout - 1]);
  }
  if (!cur.gs.empty()) tiles.push_back(std::move(cur));

  SegmentBuilder sb{H.segments};
  for (uint32_t t = 0; t < tiles.size(); ++t)
    sb.add(t, tiles[t].in_lo, tiles[t].in_hi, tiles[t].out_lo, tiles[t].out_hi);
  sb.close((uint32_t)tiles.size());

  for (const Segment &seg : H.segments) {
    for (uint32_t t = seg.tile_begin; t < seg.tile_end; ++t) {
      TileHost &T = tiles[t];
      TileDev D{};
      // staged columns, padded by repeats of the last one to whole chunks and
      // to a multiple of 16 entries
      const uint32_t nstage = std::max<uint32_t>(T.nstage_pad, 16);
      D.stage_off = (uint32_t)H.stage.size();
      D.nstage = nstage;
      const uint32_t fill = T.S.empty() ? seg.in_base * R : T.S.back();
      for (uint32_t k = 0; k < (nstage + 15) / 16 * 16; ++k)
        H.stage.push_back(col_word(k < T.S.size() ? T.S[k] : fill, R, RS, seg.in_base));
      // metadata blob
      D.meta_off = (uint32_t)H.meta.size();
      D.ngrp = (uint32_t)T.gs.size();
      std::vector<uint32_t> blob(8 * T.gs.size(), 0);
      for (uint32_t i = 0; i < T.gs.size(); ++i) {
        const HostGroup &hg = H.hg[T.gs[i].first];
        blob[8 * i] = T.gs[i].second;
        blob[8 * i + 1] = std::max<uint32_t>(1, (hg.ncol + 15) / 16);
        blob[8 * i + 2] = hg.nout;
        blob[8 * i + 4] = (uint32_t)blob.size();
        for (uint32_t o = 0; o < hg.nout; ++o)
          blob.push_back((H.gout[hg.out_begin + o] - seg.out_base) * kCtWords * 4u);
        for (uint32_t o = 0; o < hg.nout; ++o) blob.push_back(H.bias[H.gout[hg.out_begin + o]]);
        while (blob.size() % 4) blob.push_back(0);
      }
      for (uint32_t i = 0; i < T.gs.size(); ++i) {
        const HostGroup &hg = H.hg[T.gs[i].first];
        const uint32_t nchunks = blob[8 * i + 1];
        blob[8 * i + 3] = (uint32_t)blob.size();
        for (uint32_t ch = 0; ch < nchunks; ++ch)
          for (uint32_t o = 0; o < hg.nout; ++o) {
            const int32_t *w = H.gw.data() + hg.w_begin + (size_t)o * hg.ncol;
            for (uint32_t k = 0; k < 8; ++k) {
              const uint32_t c0 = ch * 16 + 2 * k, c1 = c0 + 1;
              const uint32_t lo = c0 < hg.ncol ? (uint16_t)(int16_t)w[c0] : 0u;
              const uint32_t hi = c1 < hg.ncol ? (uint16_t)(int16_t)w[c1] : 0u;
              blob.push_back(lo | (hi << 16));
            }
          }
      }
      D.nmeta = (uint32_t)blob.size();
      H.meta.insert(H.meta.end(), blob.begin(), blob.end());
      H.max_stage = std::max(H.max_stage, nstage);
      H.max_meta = std::max(H.max_meta, D.nmeta);
      H.tiles.push_back(D);
      H.cols.resize(H.cols.size() + T.S.size());  // counts staged columns for the band info
    }
  }
}


// Runs fn(begin, end) over [0, n) in contiguous ranges on up to 16 threads
// (one when n is small).
template <typename F>
void parallel_ranges(size_t n, F fn) {
  const size_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const size_t T = std::min(hw, (n + 63) / 64);
  if (T <= 1) return fn((size_t)0, n);
  std::vector<std::thread> th;
  for (size_t t = 1; t < T; ++t) th.emplace_back(fn, n * t / T, n * (t + 1) / T);
  fn((size_t)0, n / T);
  for (std::thread &x : th) x.join();
}

// MFMA layout: tiles of 32 consecutive output slots over the union of their
// columns. Returns false when a tile's union exceeds kMfmaMaxKb blocks.
// MFMA layout: tiles of up to 32*RB consecutive rows whose column union fits
// nkb K-blocks of 32 columns. Column lists hold big indices, so the tiling is
// independent of the packing layout; segments bound the big-index and output
// ranges so that every launch addresses its buffers with 32-bit offsets.
// Returns false when a single row needs more than nkb K-blocks.
bool layout_mfma(uint32_t num_out, const MfmaShape &shape, HostBand &H) {
  const uint32_t rows = kMfmaRows * (uint32_t)shape.rb;
  const uint32_t kb_budget = (uint32_t)shape.nkb;
  // output slot -> (group, row within the group)
  std::vector<uint32_t> og(num_out), orow(num_out);
  for (uint32_t g = 0; g < H.hg.size(); ++g)
    for (uint32_t i = 0; i < H.hg[g].nout; ++i) {
      og[H.gout[H.hg[g].out_begin + i]] = g;
      orow[H.gout[H.hg[g].out_begin + i]] = i;
    }
  std::vector<std::vector<uint32_t>> U;
  std::vector<uint32_t> t_o0, t_rows;
  std::vector<uint32_t> merged;
  for (uint32_t o = 0; o < num_out;) {
    const uint32_t o0 = o;
    std::vector<uint32_t> u;
    uint32_t last_g = UINT32_MAX;
    while (o < num_out && o - o0 < rows) {
      const uint32_t g = og[o];
      if (g != last_g) {
        const HostGroup &hg = H.hg[g];
        merged.clear();
        std::set_union(u.begin(), u.end(), H.gcol.begin() + hg.col_begin,
                       H.gcol.begin() + hg.col_begin + hg.ncol, std::back_inserter(merged));
        if (merged.size() > 32 * kb_budget) {
          if (o == o0) return false;  // one row alone is wider than the budget
          break;
        }
        u.swap(merged);
        last_g = g;
      }
      ++o;
    }
    U.push_back(std::move(u));
    t_o0.push_back(o0);
    t_rows.push_back(o - o0);
  }
  const uint32_t ntiles = (uint32_t)U.size();
  SegmentBuilder sb{H.segments};
  for (uint32_t t = 0; t < ntiles; ++t) {
    const std::vector<uint32_t> &u = U[t];
    if (!u.empty() && u.back() - u.front() >= kSegCts) return false;
    sb.add(t, u.empty() ? UINT32_MAX : u.front(), u.empty() ? 0 : u.back(), t_o0[t],
           t_o0[t] + t_rows[t] - 1);
  }
  sb.close(ntiles);

  H.shape = shape;
  // each tile's slices of the device arrays, then the tiles filled in parallel
  const uint32_t rb_count = (uint32_t)shape.rb;
  std::vector<uint32_t> tile_nkb(ntiles), tile_seg(ntiles);
  std::vector<size_t> kcol_at(ntiles + 1), a_at(ntiles + 1);
  const size_t kcol0 = H.kcols.size(), a0 = H.adata.size(), row0 = H.rowc.size(), tile0 = H.mtiles.size();
  size_t ncols = 0;
  for (uint32_t i = 0; i < (uint32_t)H.segments.size(); ++i)
    for (uint32_t t = H.segments[i].tile_begin; t < H.segments[i].tile_end; ++t) tile_seg[t] = i;
  kcol_at[0] = kcol0, a_at[0] = a0;
  for (uint32_t t = 0; t < ntiles; ++t) {
    tile_nkb[t] = std::max<uint32_t>(1, (uint32_t)(U[t].size() + 31) / 32);
    kcol_at[t + 1] = kcol_at[t] + 32 * tile_nkb[t];
    a_at[t + 1] = a_at[t] + (size_t)rb_count * tile_nkb[t] * 2 * 64;
    ncols += U[t].size();
  }
  H.kcols.resize(kcol_at[ntiles]);
  H.adata.resize(a_at[ntiles]);
  H.rowc.resize(row0 + (size_t)rows * ntiles);
  H.mtiles.resize(tile0 + ntiles);
  H.cols.resize(H.cols.size() + ncols);  // counts tile columns for the band info
  parallel_ranges(ntiles, [&](size_t t_begin, size_t t_end) {
    std::vector<int32_t> W;
    for (uint32_t t = (uint32_t)t_begin; t < (uint32_t)t_end; ++t) {
      const Segment &seg = H.segments[tile_seg[t]];
      const std::vector<uint32_t> &u = U[t];
      const uint32_t o0 = t_o0[t], nrows = t_rows[t], nkb = tile_nkb[t];
      MTileDev D{};
      D.kcol_off = (uint32_t)kcol_at[t];
      D.nkb = nkb;
      D.a_off = (uint32_t)(a_at[t] / 64);
      D.row_off = (uint32_t)(row0 + (size_t)rows * t);
      D.slot0 = o0 - seg.out_base;
      D.nrows = nrows;
      // padding columns repeat a real one (or the segment base) with weight 0
      const uint32_t fill = u.empty() ? seg.in_base : u.back();
      for (uint32_t k = 0; k < 32 * nkb; ++k) H.kcols[kcol_at[t] + k] = k < u.size() ? u[k] : fill;
      // dense weights of the tile: W[m][k]
      W.assign((size_t)rows * 32 * nkb, 0);
      for (uint32_t m = 0; m < nrows; ++m) {
        const HostGroup &hg = H.hg[og[o0 + m]];
        const uint32_t *gc = H.gcol.data() + hg.col_begin;
        const int32_t *gw = H.gw.data() + hg.w_begin + (size_t)orow[o0 + m] * hg.ncol;
        for (uint32_t i = 0; i < hg.ncol; ++i) {
          const uint32_t k = (uint32_t)(std::lower_bound(u.begin(), u.end(), gc[i]) - u.begin());
          W[(size_t)m * 32 * nkb + k] = gw[i];
        }
      }
      // A operand blocks in (row block, K-block, weight byte) order
      size_t a = a_at[t];
      for (uint32_t rb = 0; rb < rb_count; ++rb)
        for (uint32_t kb = 0; kb < nkb; ++kb)
          for (uint32_t hp = 0; hp < 2; ++hp)
            for (uint32_t lane = 0; lane < 64; ++lane) {
              int8_t bytes[16];
              for (uint32_t j = 0; j < 16; ++j) {
                const uint32_t m = 32 * rb + lane % 32;
                const int32_t w = W[(size_t)m * 32 * nkb + 32 * kb + 16 * (lane / 32) + j];
                const int32_t w0 = (int8_t)(w & 0xFF);
                bytes[j] = (int8_t)(hp == 0 ? w0 : (w - w0) / 256);
              }
              std::memcpy(&H.adata[a++], bytes, 16);
            }
      for (uint32_t m = 0; m < rows; ++m) H.rowc[D.row_off + m] = m < nrows ? H.bias[o0 + m] : 0u;
      H.mtiles[tile0 + t] = D;
    }
  });
  return true;
}

int check_layout(const ihe_layout *L) {
  if (!L) return fail(IHE_ERR_ARG, "null layout");
  if (L->num_samples == 0 || L->num_samples > kN) return fail(IHE_ERR_ARG, "NUM_SAMPLES must be in [1, 1024]");
  if (L->num_regions != kN / L->num_samples || L->region_size != kN / L->num_regions)
    return fail(IHE_ERR_ARG, "layout must satisfy R = N / NS and RS = N / R (eval/idash.cpp:409-410)");
  return IHE_OK;
}

uint32_t env_u32(const char *name, uint32_t dflt, uint32_t lo, uint32_t hi) {
  if (const char *e = std::getenv(name)) {
    const long v = std::atol(e);
    if (v >= (long)lo && v <= (long)hi) return (uint32_t)v;
  }
  return dflt;
}

// Number of tiles `shape` cuts the outputs into (0 when a row does not fit).
uint32_t count_tiles(uint32_t num_out, const MfmaShape &shape, const HostBand &H) {
  const uint32_t rows = kMfmaRows * (uint32_t)shape.rb, budget = 32 * (uint32_t)shape.nkb;
  std::vector<uint32_t> og(num_out);
  for (uint32_t g = 0; g < H.hg.size(); ++g)
    for (uint32_t i = 0; i < H.hg[g].nout; ++i) og[H.gout[H.hg[g].out_begin + i]] = g;
  std::vector<uint32_t> u, merged;
  uint32_t tiles = 0;
  for (uint32_t o = 0; o < num_out;) {
    const uint32_t o0 = o;
    u.clear();
    uint32_t last_g = UINT32_MAX;
    while (o < num_out && o - o0 < rows) {
      const uint32_t g = og[o];
      if (g != last_g) {
        const HostGroup &hg = H.hg[g];
        merged.clear();
        std::set_union(u.begin(), u.end(), H.gcol.begin() + hg.col_begin,
                       H.gcol.begin() + hg.col_begin + hg.ncol, std::back_inserter(merged));
        if (merged.size() > budget) {
          if (o == o0) return 0;
          break;
        }
        u.swap(merged);
        last_g = g;
      }
      ++o;
    }
    ++tiles;
  }
  return tiles;
}

// Column tiles per wave (per workgroup for the four-row-block kernel):
// IHE_CLOUD_NTW if it names one the kernel family has, else its default: 4
// per wave; 16 per workgroup up to 4 K-blocks, 32 above (n = 50 on the full
// layout, 128 x 6 tiles: 0.887 ms at 32, 0.911 at 16, 0.951 at 8).
int pick_ntw(int rb, int nkb) {
  const int v = (int)env_u32("IHE_CLOUD_NTW", 0, 4, 32);
  if (rb == kMfmaWideRb) return v == 8 || v == 16 || v == 32 ? v : nkb > 4 ? 32 : 16;
  return v == 8 || v == 16 ? v : 4;
}

// MFMA tile shape: IHE_CLOUD_SHAPE="rb,nkb" if given, else the candidate with
// the least estimated time. Per-wave kernels (rb 1, 2): tiles * (nkb + 1.5 rb),
// their per-tile and column-tile loads (16 per K-block) and stores (16 per row
// block). Four-row-block kernel: tiles * (3.2 nkb + 4.5), fitted on one
// MI355X to the full 80,882-target layout (n = 5: 0.48 ms for 64-row tiles
// against 0.61 ms for 128 rows x 2 K-blocks; n = 50: 0.88 ms for 128 rows x 6
// K-blocks, 0.98 ms for 128 x 8, against 1.01 ms for 32 x 5). The kernel
// autotuner (ihe_band_create_tuned) refines this choice by measurement.
bool pick_mfma_shape(uint32_t num_out, HostBand &H, MfmaShape &shape) {
  if (const char *e = std::getenv("IHE_CLOUD_SHAPE")) {
    int rb = 0, nkb = 0;
    if (std::sscanf(e, "%d,%d", &rb, &nkb) == 2 &&
        (rb == 1 || (rb == 2 && nkb == 1) || rb == kMfmaWideRb) && nkb >= 1 &&
        nkb <= (int)kMfmaMaxNkb) {
      shape = {rb, nkb, pick_ntw(rb, nkb)};
      return count_tiles(num_out, shape, H) > 0;
    }
  }
  const int kmin = std::max<int>(1, (int)((H.max_window + 31) / 32));
  std::vector<MfmaShape> cands;
  if (kmin == 1) cands.push_back({2, 1, pick_ntw(2, 1)}), cands.push_back({1, 1, pick_ntw(1, 1)});
  for (int k = std::max(2, kmin); k <= std::min<int>(kmin + 2, (int)kMfmaMaxNkb); ++k)
    cands.push_back({1, k, pick_ntw(1, k)});
  // the four-row-block kernel runs an even number of K-blocks
  for (int k = kmin + kmin % 2; k <= std::min<int>(kmin + 3, (int)kMfmaMaxNkb); k += 2)
    cands.push_back({kMfmaWideRb, k, pick_ntw(kMfmaWideRb, k)});
  double best = 0;
  bool found = false;
  for (const MfmaShape &c : cands) {
    const uint32_t t = count_tiles(num_out, c, H);
    if (t == 0) continue;
    H.candidates.push_back(c);
    const double cost = c.rb == kMfmaWideRb ? (double)t * (3.2 * c.nkb + 4.5)
                                            : (double)t * (c.nkb + 1.5 * c.rb);
    if (!found || cost < best) best = cost, shape = c, found = true;
  }
  return found || num_out == 0;
}

// Rewrites the per-row variance entries (var_ptr / var_col / var_w2, row
// order) into the kernel's layout: per block of 64 outputs, entry j of every
// row side by side, rows padded to the block's longest row; var_ptr becomes
// the per-block offsets in units of 64 entries. When every |w| <= 32768 and
// every row's big indices lie within 65534 of its smallest, an entry packs
// into 32 bits, (c - base) << 16 | |w| (w*w recomputed on the device), with
// the row bases in var_base; otherwise (c, w*w) pairs, padded with kVarPad.
void interleave_variances(HostBand &H) {
  const uint32_t num_out = (uint32_t)H.var_ptr.size() - 1;
  const uint32_t nblk = (num_out + 63) / 64;
  std::vector<uint32_t> base(num_out, 0);
  bool packed = !H.var_wide_w;
  for (uint32_t o = 0; o < num_out && packed; ++o) {
    if (H.var_ptr[o] == H.var_ptr[o + 1]) continue;
    uint32_t lo = UINT32_MAX, hi = 0;
    for (uint32_t e = H.var_ptr[o]; e < H.var_ptr[o + 1]; ++e)
      lo = std::min(lo, H.var_col[e]), hi = std::max(hi, H.var_col[e]);
    base[o] = lo;
    packed = hi - lo < kVarPackedPad;
  }
  std::vector<uint32_t> off(nblk + 1, 0), col;
  std::vector<int32_t> w2;
  for (uint32_t b = 0; b < nblk; ++b) {
    uint32_t len = 0;
    for (uint32_t o = 64 * b; o < std::min(num_out, 64 * b + 64); ++o)
      len = std::max(len, H.var_ptr[o + 1] - H.var_ptr[o]);
    off[b + 1] = off[b] + len;
  }
  if (packed) {
    col.assign((size_t)off[nblk] * 64, kVarPackedPad << 16);
    for (uint32_t o = 0; o < num_out; ++o) {
      const uint32_t b = o / 64, l = o % 64;
      for (uint32_t e = H.var_ptr[o], j = 0; e < H.var_ptr[o + 1]; ++e, ++j) {
        // |w| from w*w (exact below 2^31)
        const uint32_t aw = (uint32_t)std::llround(std::sqrt((double)(uint32_t)H.var_w2[e]));
        col[((size_t)off[b] + j) * 64 + l] = (H.var_col[e] - base[o]) << 16 | aw;
      }
    }
    H.var_base.swap(base);
  } else {
    col.assign((size_t)off[nblk] * 64, kVarPad);
    w2.assign((size_t)off[nblk] * 64, 0);
    for (uint32_t o = 0; o < num_out; ++o) {
      const uint32_t b = o / 64, l = o % 64;
      for (uint32_t e = H.var_ptr[o], j = 0; e < H.var_ptr[o + 1]; ++e, ++j) {
        col[((size_t)off[b] + j) * 64 + l] = H.var_col[e];
        w2[((size_t)off[b] + j) * 64 + l] = H.var_w2[e];
      }
    }
    H.var_base.clear();
  }
  H.var_ptr.swap(off);
  H.var_col.swap(col);
  H.var_w2.swap(w2);
}

// The grouped model (build_groups' results) of an earlier plan of the same
// model and layout.
void copy_groups(const HostBand &G, HostBand &H) {
  H.bias = G.bias;
  H.var_ptr = G.var_ptr;
  H.var_col = G.var_col;
  H.var_w2 = G.var_w2;
  H.var_wide_w = G.var_wide_w;
  H.max_bidx = G.max_bidx;
  H.has_cols = G.has_cols;
  H.hg = G.hg;
  H.gcol = G.gcol;
  H.gout = G.gout;
  H.gw = G.gw;
  H.max_window = G.max_window;
}

// Validates arguments, picks the kernel family and compiles the band on the host.
// `grouped`: an earlier plan of the same model and layout whose grouping is
// reused (the autotuner's candidates differ only in their tiling).
int plan_band(const ihe_layout *layout, uint32_t num_in_cts, const ihe_model_csr *model,
              int variant, HostBand &H, const MfmaShape *force_shape = nullptr,
              const HostBand *grouped = nullptr) {
  if (!model) return fail(IHE_ERR_ARG, "null model");
  if (int rc = check_layout(layout)) return rc;
  if (model->num_outputs > 0 && (!model->row_ptr || (model->row_ptr[model->num_outputs] > 0 &&
                                                     (!model->in_bidx || !model->weight))))
    return fail(IHE_ERR_ARG, "null model arrays");
  if (variant != IHE_VARIANT_AUTO && variant != IHE_VARIANT_VALU && variant != IHE_VARIANT_WIDE &&
      variant != IHE_VARIANT_MFMA)
    return fail(IHE_ERR_ARG, "unknown variant");
  bool wide = variant == IHE_VARIANT_WIDE;
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
  H.wide = wide;
  if (grouped) copy_groups(*grouped, H);
  else if (int rc = build_groups(*layout, num_in_cts, *model, wide, H)) return rc;
  // MFMA needs weights that split into two signed bytes
  bool mfma_ok = !wide;
  if (mfma_ok) {
    const uint32_t nnz = model->num_outputs ? model->row_ptr[model->num_outputs] : 0;
    for (uint32_t e = 0; e < nnz && mfma_ok; ++e)
      if (model->in_bidx[e] != IHE_CONSTANT_BIDX && model->weight[e] > 32639) mfma_ok = false;
  }
  // kernel family; IHE_CLOUD_KERNEL=direct|staged|valu|mfma overrides it for
  // experiments within what the variant allows
  const char *force_env = std::getenv("IHE_CLOUD_KERNEL");
  const std::string force = force_env ? force_env : "";
  const bool fits_stage = H.max_window <= kStageMaxCols - 16;
  int mode;
  if (variant == IHE_VARIANT_MFMA || (variant == IHE_VARIANT_AUTO && mfma_ok &&
                                      force != "direct" && force != "staged" && force != "valu")) {
    if (!mfma_ok) return fail(IHE_ERR_UNSUPPORTED, "MFMA variant needs weights in [-32768, 32639]");
    mode = kModeMfma;
  } else if (wide || force == "direct") {
    mode = kModeDirect;
  } else {
    mode = fits_stage ? kModeStaged : kModeDirect;
  }
  if (mode == kModeMfma) {
    MfmaShape shape;
    const bool ok = force_shape ? (shape = *force_shape, count_tiles(model->num_outputs, shape, H) > 0)
                          : pick_mfma_shape(model->num_outputs, H, shape);
    if (!ok || !layout_mfma(model->num_outputs, shape, H)) {
      if (variant == IHE_VARIANT_MFMA)
        return fail(IHE_ERR_UNSUPPORTED, "an MFMA tile needs more than 256 distinct columns");
      H.segments.clear();
      H.mtiles.clear(), H.kcols.clear(), H.rowc.clear(), H.adata.clear(), H.cols.clear();
      mode = fits_stage ? kModeStaged : kModeDirect;
    }
  }
  H.mode = mode;
  if (mode == kModeStaged) layout_staged(*layout, env_u32("IHE_CLOUD_TILE_OUTPUTS", 48, 1, 4096), H);
  if (mode == kModeDirect) layout_direct(*layout, env_u32("IHE_CLOUD_TILE_WORK", 512, 8, 1 << 20), H);
  return IHE_OK;
}

void fill_info(const HostBand &H, const ihe_model_csr *model, ihe_band_info *info) {
  info->num_outputs = model->num_outputs;
  info->num_groups = (uint32_t)H.hg.size();
  info->num_tiles = H.mode == kModeStaged                          ? (uint32_t)H.tiles.size()
                   : H.mode == kModeMfma ? (uint32_t)H.mtiles.size()
                                                                     : (uint32_t)H.tile_off.size() - 1;
  info->num_cols = (uint32_t)H.cols.size();
  info->nnz = model->num_outputs ? model->row_ptr[model->num_outputs] : 0;
  info->variant = H.wide ? IHE_VARIANT_WIDE
                : H.mode == kModeMfma ? IHE_VARIANT_MFMA
                                                                 : IHE_VARIANT_VALU;
  info->max_window = H.max_window;
  info->device_bytes = 0;
  const bool mf = H.mode == kModeMfma;
  info->tile_rows = mf ? kMfmaRows * (uint32_t)H.shape.rb : 0;
  info->tile_kblocks = mf ? (uint32_t)H.shape.nkb : 0;
  info->col_tiles = mf ? (uint32_t)H.shape.ntw : 0;
}

template <typename T>
T *carve(char *base, size_t &off, size_t count) {
  T *p = reinterpret_cast<T *>(base + off);
  off += align_up(sizeof(T) * std::max<size_t>(count, 1));
  return p;
}

}  // namespace

extern "C" {

int ihe_band_plan(const ihe_layout *layout, uint32_t num_in_cts, const ihe_model_csr *model,
                  int variant, ihe_band_info *info) {
  if (!info) return fail(IHE_ERR_ARG, "null info");
  HostBand H;
  if (int rc = plan_band(layout, num_in_cts, model, variant, H)) return rc;
  fill_info(H, model, info);
  return IHE_OK;
}

}  // extern "C"

namespace {

void launch_mfma_band(ihe_ctx *ctx, const ihe_band *band, const ihe_layout &L, uint32_t num_in,
                      const int32_t *d_in, int32_t *d_out, const VarArgs &var = VarArgs{});

// Uploads a compiled band.
int upload_band(ihe_ctx *ctx, const ihe_layout *layout, uint32_t num_in_cts,
                const ihe_model_csr *model, HostBand &H, ihe_band **out) {
  ihe_band *b = new ihe_band();
  b->wpl = (int)env_u32("IHE_CLOUD_WPL", 1, 1, 4);
  if (b->wpl == 3) b->wpl = 2;
  b->ctx = ctx;
  b->layout = *layout;
  b->num_in = num_in_cts;
  b->num_out = model->num_outputs;
  ihe_band_info info;
  fill_info(H, model, &info);
  b->variant = (int)info.variant;
  b->mode = H.mode;
  b->num_groups = info.num_groups;
  b->num_tiles = info.num_tiles;
  b->num_cols = info.num_cols;
  b->nnz = info.nnz;
  b->max_window = info.max_window;
  b->segments = H.segments;
  b->meta_lds = H.max_stage * kStageSlice;
  b->shape = H.shape;
  b->max_bidx = H.max_bidx;
  b->has_cols = H.has_cols;
  b->lds_bytes = (b->meta_lds + H.max_meta) * 4u;

  interleave_variances(H);
  size_t bytes = 0;
  for (size_t n : {H.grp.size() * 8, H.tile_off.size(), H.cols.size(), H.oslot.size(),
                   H.obias.size(), H.wts.size(), H.tiles.size() * 8, H.stage.size(), H.meta.size(),
                   H.mtiles.size() * 8, H.kcols.size(), H.rowc.size(), H.adata.size() * 4,
                   H.var_ptr.size(), H.var_col.size(), H.var_w2.size(), H.var_base.size()})
    bytes += align_up(4 * std::max<size_t>(n, 1));
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
  put(&b->d_cols, H.mode == kModeDirect ? H.cols : std::vector<uint32_t>());
  put(&b->d_oslot, H.oslot);
  put(&b->d_obias, H.obias);
  put(&b->d_wts, H.wts);
  put(&b->d_tiles, H.tiles);
  put(&b->d_stage, H.stage);
  put(&b->d_meta, H.meta);
  put(&b->d_mtiles, H.mtiles);
  put(&b->d_kcols, H.kcols);
  put(&b->d_rowc, H.rowc);
  put(&b->d_adata, H.adata);
  put(&b->d_var_ptr, H.var_ptr);
  put(&b->d_var_col, H.var_col);
  put(&b->d_var_w2, H.var_w2);
  put(&b->d_var_base, H.var_base);
  if (H.var_base.empty()) b->d_var_base = nullptr;
  if (int rc = ihe_memcpy_h2d(ctx, b->dmem, staging.data(), bytes)) {
    (void)ihe_free(ctx, b->dmem);
    delete b;
    return rc;
  }
  *out = b;
  return IHE_OK;
}

}  // namespace

extern "C" {

int ihe_band_create(ihe_ctx *ctx, const ihe_layout *layout, uint32_t num_in_cts,
                    const ihe_model_csr *model, int variant, ihe_band **out) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!out) return fail(IHE_ERR_ARG, "null argument");
  *out = nullptr;
  HostBand H;
  if (int rc = plan_band(layout, num_in_cts, model, variant, H)) return rc;
  return upload_band(ctx, layout, num_in_cts, model, H, out);
}

int ihe_band_create_tuned(ihe_ctx *ctx, const ihe_layout *layout, uint32_t num_in_cts,
                          const ihe_model_csr *model, const int32_t *d_in, int32_t *d_out,
                          uint32_t reps, ihe_band **out, ihe_tune_entry *results,
                          uint32_t *num_results) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!out) return fail(IHE_ERR_ARG, "null argument");
  *out = nullptr;
  const uint32_t cap = num_results ? *num_results : 0;
  if (num_results) *num_results = 0;
  HostBand P;
  if (int rc = plan_band(layout, num_in_cts, model, IHE_VARIANT_MFMA, P)) return rc;
  if (model->num_outputs == 0 || P.candidates.empty())
    return upload_band(ctx, layout, num_in_cts, model, P, out);
  if (!d_out || (num_in_cts > 0 && !d_in)) return fail(IHE_ERR_ARG, "null ciphertext buffer");
  if (int rc = make_current(ctx)) return rc;
  reps = std::max<uint32_t>(reps, 1);
  hipEvent_t e0, e1;
  IHE_HIP(hipEventCreate(&e0));
  IHE_HIP(hipEventCreate(&e1));
  ihe_band *best = nullptr;
  float best_ms = 0;
  uint32_t n = 0;
  int rc = IHE_OK;
  for (const MfmaShape &cand : std::vector<MfmaShape>(P.candidates)) {
    HostBand H;
    if ((rc = plan_band(layout, num_in_cts, model, IHE_VARIANT_MFMA, H, &cand, &P))) break;
    ihe_band *b = nullptr;
    if ((rc = upload_band(ctx, layout, num_in_cts, model, H, &b))) break;
    int best_ntw = 0;
    float band_ms = 0;
    const std::array<int, 3> ntws =
        cand.rb == kMfmaWideRb ? std::array<int, 3>{8, 16, 32} : std::array<int, 3>{4, 8, 16};
    for (int ntw : ntws) {
      b->shape.ntw = ntw;
      launch_mfma_band(ctx, b, *layout, num_in_cts, d_in, d_out);  // warm-up
      (void)hipEventRecord(e0, ctx->stream);
      for (uint32_t i = 0; i < reps; ++i) launch_mfma_band(ctx, b, *layout, num_in_cts, d_in, d_out);
      (void)hipEventRecord(e1, ctx->stream);
      if (hipEventSynchronize(e1) != hipSuccess) {
        rc = fail(IHE_ERR_HIP, "autotuner launch failed");
        break;
      }
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      ms /= (float)reps;
      if (n < cap) results[n] = {32u * (uint32_t)cand.rb, (uint32_t)cand.nkb, (uint32_t)ntw, ms};
      ++n;
      if (best_ntw == 0 || ms < band_ms) band_ms = ms, best_ntw = ntw;
    }
    b->shape.ntw = best_ntw;
    if (rc == IHE_OK && (!best || band_ms < best_ms)) {
      if (best) ihe_band_destroy(best);
      best = b, best_ms = band_ms;
    } else {
      ihe_band_destroy(b);
    }
    if (rc) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (num_results) *num_results = std::min(n, cap);
  if (rc) {
    if (best) ihe_band_destroy(best);
    return rc;
  }
  *out = best;
  return IHE_OK;
}

int ihe_band_destroy(ihe_band *band) {
  if (!band) return IHE_OK;
  if (band->users.load() > 0)
    return fail(IHE_ERR_BUSY, "the band is used by " + std::to_string(band->users.load()) +
                                  " live batch or server holder(s)");
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
  const bool mf = band->mode == kModeMfma;
  info->tile_rows = mf ? kMfmaRows * (uint32_t)band->shape.rb : 0;
  info->tile_kblocks = mf ? (uint32_t)band->shape.nkb : 0;
  info->col_tiles = mf ? (uint32_t)band->shape.ntw : 0;
  return IHE_OK;
}

namespace {

void launch_direct(ihe_ctx *ctx, const ihe_band *band, const int32_t *d_in, int32_t *d_out) {
  CloudParams P;
  P.grp = band->d_grp;
  P.tile_off = band->d_tile_off;
  P.cols = band->d_cols;
  P.wts = band->d_wts;
  P.oslot = band->d_oslot;
  P.obias = band->d_obias;
  P.ns = band->layout.num_samples;
  P.rs = band->layout.region_size;
  const int wpl = band->wpl;
  P.nslices = kCtWords / (256u * (uint32_t)wpl);
  const bool rot = band->layout.num_regions > 1;
  const bool wide = band->variant == IHE_VARIANT_WIDE;
  for (const Segment &s : band->segments) {
    P.tile_base = s.tile_begin;
    const uint32_t in_cts = band->num_in > s.in_base ? std::min(band->num_in - s.in_base, kSegCts) : 0u;
    P.in = in_cts ? d_in + (size_t)s.in_base * kCtWords : nullptr;
    P.in_bytes = in_cts * kCtWords * 4u;
    P.out = d_out + (size_t)s.out_base * kCtWords;
    P.out_bytes = std::min(band->num_out - s.out_base, kSegCts) * kCtWords * 4u;
    const dim3 grid((s.tile_end - s.tile_begin) * P.nslices), block(256);
#define IHE_LAUNCH(W, RT, WD) hipLaunchKernelGGL((k_cloud<W, RT, WD>), grid, block, 0, ctx->stream, P)
#define IHE_LAUNCH_W(RT, WD)                  \
  do {                                        \
    if (wpl == 4) IHE_LAUNCH(4, RT, WD);      \
    else if (wpl == 2) IHE_LAUNCH(2, RT, WD); \
    else IHE_LAUNCH(1, RT, WD);               \
  } while (0)
    if (wide) {
      if (rot) IHE_LAUNCH_W(true, true);
      else IHE_LAUNCH_W(false, true);
    } else {
      if (rot) IHE_LAUNCH_W(true, false);
      else IHE_LAUNCH_W(false, false);
    }
#undef IHE_LAUNCH_W
#undef IHE_LAUNCH
  }
}

void launch_staged(ihe_ctx *ctx, const ihe_band *band, const int32_t *d_in, int32_t *d_out) {
  StagedParams P;
  P.tiles = band->d_tiles;
  P.stage = band->d_stage;
  P.meta = band->d_meta;
  P.ns = band->layout.num_samples;
  P.rs = band->layout.region_size;
  P.meta_lds = band->meta_lds;
  const bool rot = band->layout.num_regions > 1;
  for (const Segment &s : band->segments) {
    P.tile_base = s.tile_begin;
    const uint32_t in_cts = band->num_in > s.in_base ? std::min(band->num_in - s.in_base, kSegCts) : 0u;
    P.in = in_cts ? d_in + (size_t)s.in_base * kCtWords : nullptr;
    P.in_bytes = in_cts * kCtWords * 4u;
    P.out = d_out + (size_t)s.out_base * kCtWords;
    P.out_bytes = std::min(band->num_out - s.out_base, kSegCts) * kCtWords * 4u;
    const dim3 grid((s.tile_end - s.tile_begin) * (kCtWords / kStageSlice)), block(kStageSlice);
    if (rot) hipLaunchKernelGGL(k_cloud_staged<true>, grid, block, band->lds_bytes, ctx->stream, P);
    else hipLaunchKernelGGL(k_cloud_staged<false>, grid, block, band->lds_bytes, ctx->stream, P);
  }
}

// One launch per segment; the tiling is shared by every layout. The variance
// blocks (if any) join the last segment's grid.
void launch_mfma_band(ihe_ctx *ctx, const ihe_band *band, const ihe_layout &L, uint32_t num_in,
                      const int32_t *d_in, int32_t *d_out, const VarArgs &var) {
  for (const Segment &s : band->segments) {
    MfmaEntry E = {};
    E.tiles = band->d_mtiles;
    E.kcols = band->d_kcols;
    E.adata = band->d_adata;
    E.rowc = band->d_rowc;
    E.ns = L.num_samples;
    E.rs = L.region_size;
    E.R = L.num_regions;
    E.rmagic = L.num_regions > 1 ? UINT64_MAX / L.num_regions + 1 : 0;
    E.ct_base = s.in_base / L.num_regions;
    const uint32_t in_cts = num_in > E.ct_base ? std::min(num_in - E.ct_base, kSegCts) : 0u;
    E.in = in_cts ? d_in + (size_t)E.ct_base * kCtWords : nullptr;
    E.in_bytes = in_cts * kCtWords * 4u;
    E.out = d_out + (size_t)s.out_base * kCtWords;
    E.out_bytes = std::min(band->num_out - s.out_base, kSegCts) * kCtWords * 4u;
    E.tile_base = s.tile_begin;
    const bool last = &s == &band->segments.back();
    launch_mfma(ctx->stream, band->shape, L.num_regions > 1, E, nullptr, nullptr, 0,
                s.tile_end - s.tile_begin, last ? var : VarArgs{});
  }
}

// The variance work of one cohort on a band.
VarArgs band_variances(const ihe_band *band, const ihe_layout &L, const double *d_in_var,
                       double *d_out_var) {
  VarArgs V{};
  V.single = make_var_entry(band->d_var_ptr, band->d_var_col, band->d_var_w2, band->d_var_base,
                            d_in_var, d_out_var, band->num_out, L.num_regions);
  V.num_blocks = (band->num_out + 255) / 256;
  return V;
}

}  // namespace

int ihe_cloud_score_layout(ihe_ctx *ctx, const ihe_band *band, const ihe_layout *layout,
                           uint32_t num_in_cts, const int32_t *d_in, const double *d_in_var,
                           int32_t *d_out, double *d_out_var) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!band || band->ctx->device != ctx->device)
    return fail(IHE_ERR_ARG, "band is not on this context's device");
  if (int rc = check_layout(layout)) return rc;
  const ihe_layout &L = *layout;
  if (band->mode != kModeMfma &&
      (L.num_samples != band->layout.num_samples || num_in_cts != band->num_in))
    return fail(IHE_ERR_UNSUPPORTED, "only MFMA bands run on another layout than their own");
  // every referenced ciphertext must exist (getTLWE, eval/idash.h:162-166)
  if (band->has_cols && band->max_bidx / L.num_regions >= num_in_cts)
    return fail(IHE_ERR_RANGE, "the model references FeatIndex " +
                                   std::to_string(band->max_bidx / L.num_regions) + " >= " +
                                   std::to_string(num_in_cts));
  if (band->num_out == 0) return IHE_OK;
  if (!d_out || (num_in_cts > 0 && !d_in)) return fail(IHE_ERR_ARG, "null ciphertext buffer");
  if (d_out_var && num_in_cts > 0 && !d_in_var) return fail(IHE_ERR_ARG, "null input variance");
  if (int rc = make_current(ctx)) return rc;
  // MFMA bands compute the variances in the cloud launch itself (its trailing
  // blocks); IHE_CLOUD_FUSED_VAR=0 gives them a launch of their own
  const bool fused = band->mode == kModeMfma && env_u32("IHE_CLOUD_FUSED_VAR", 1, 0, 1) == 1;
  const VarArgs V = d_out_var ? band_variances(band, L, d_in_var, d_out_var) : VarArgs{};
  {
    ScopedTimer t(ctx, IHE_TIMER_CLOUD);
    if (band->mode == kModeMfma)
      launch_mfma_band(ctx, band, L, num_in_cts, d_in, d_out, fused ? V : VarArgs{});
    else if (band->mode == kModeStaged) launch_staged(ctx, band, d_in, d_out);
    else launch_direct(ctx, band, d_in, d_out);
  }
  IHE_HIP(hipGetLastError());
  if (d_out_var && !fused) {
    ScopedTimer t(ctx, IHE_TIMER_CLOUD_VAR);
    launch_var(ctx->stream, V);
    IHE_HIP(hipGetLastError());
  }
  return IHE_OK;
}

int ihe_cloud_score(ihe_ctx *ctx, const ihe_band *band, const int32_t *d_in,
                    const double *d_in_var, int32_t *d_out, double *d_out_var) {
  if (!band) return fail(IHE_ERR_ARG, "null band");
  return ihe_cloud_score_layout(ctx, band, &band->layout, band->num_in, d_in, d_in_var, d_out,
                                d_out_var);
}

}  // extern "C"
