// Byte-split int8 MFMA cloud kernels (gfx950, v_mfma_i32_32x32x32_i8).
//
// The cloud step of cloud_compute_score (eval/idash.cpp:763-848) for int16
// weights, as a matrix product per tile: out[m][word] = sum_k W[m][k] *
// rho_k(CT_k)[word] (mod 2^32), where k runs over the tile's columns (big
// indices), CT_k = ciphertext k / R rotated by (k % R) * RS (the reference
// rotates the per-region sums instead; rotation is linear mod 2^32).
//
// Exact on int8 MFMA: a ciphertext word x is recoded into balanced signed
// bytes, x = sum_b 2^(8b) d_b with d_b = byte_b(x + 0x80808080) - 128 in
// [-128, 127], and a weight into w = w_0 + 2^8 w_1 with w_0 the sign-extended
// low byte (w in [-32768, 32639]). Then
//   sum_k w x = sum_{b+h<=3} 2^(8(b+h)) sum_k w_h d_b   (mod 2^32)
// because the terms with b+h >= 4 vanish mod 2^32. A K-block of 32 columns
// costs seven MFMAs into four int32 accumulators (shifts 0, 8, 16, 24); every
// partial sum stays below 2^31 in magnitude, and the accumulators are
// recombined with shifts.
//
// Operand layout of v_mfma_i32_32x32x32_i8: A lane l holds row l%32, bytes
// k = 16*(l/32) + j; B lane l holds column (word) l%32 with the same k; value v
// of the result in lane l is row 8*(v/4) + 4*(l/32) + v%4, word l%32.
#include "ihe_mfma.h"

#include <cstdlib>

namespace ihe {
namespace {

typedef int int16v __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4v __attribute__((ext_vector_type(4)));
typedef uint32_t u32x16v __attribute__((ext_vector_type(16)));
#define IHE_CONSTANT __attribute__((address_space(4)))
constexpr int kRsrcWord3 = 0x00020000;          // gfx9 raw buffer descriptor, dword 3
constexpr uint32_t kDropOffset = 0x80000000u;   // >= every segment's byte size: stores dropped
constexpr uint32_t kWaves = 4;                  // waves per workgroup
// Output stores: nt (2, streaming) with sc1 (16, device scope: written through
// the XCD's L2). Outputs are never re-read by the step; of the documented
// cache-policy bits this pair streams the step's volumes fastest.
constexpr int kStorePolicy = 18;

template <typename T>
__device__ __forceinline__ const IHE_CONSTANT T *cptr(const T *p) {
  return (const IHE_CONSTANT T *)p;
}

__device__ __forceinline__ void var_block(const VarArgs &A, uint32_t chunk,
                                          uint32_t t = threadIdx.x);

struct EntryWords {
  u32x16v lo, hi;
};

// The workgroup's entry and tile: blocks are renumbered so that each XCD (the
// dispatcher deals blocks round-robin over 8 XCDs) walks a contiguous run of
// items, whose column windows overlap in that XCD's L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t n) {
  const uint32_t per = n / 8;
  return b < 8 * per ? (b % 8) * per + b / 8 : b;
}

// The thread's index in its workgroup, opaque to the optimizer. The persistent
// kernel hands it to the bodies per item, so that what they derive from it is
// computed where it is used, never hoisted out of the item loop into
// registers that would live (and be spilled) across every body.
__device__ __forceinline__ uint32_t tid() {
  uint32_t t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

__device__ __forceinline__ MfmaEntry entry_of(const MfmaArgs &A, uint32_t item, uint32_t &tile) {
  if (!A.entries) {
    tile = A.single.tile_base + item;
    return A.single;
  }
  const IHE_CONSTANT uint32_t *first = cptr(A.first);
  uint32_t lo = 0, hi = A.num_entries;  // first[lo] <= item < first[hi]
  if (A.item_entry) {
    lo = cptr(A.item_entry)[item];
  } else {
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (first[mid] <= item) lo = mid;
      else hi = mid;
    }
  }
  const IHE_CONSTANT u32x16v *e = (const IHE_CONSTANT u32x16v *)(A.entries + lo);
  const EntryWords w = {e[0], e[1]};
  const MfmaEntry E = __builtin_bit_cast(MfmaEntry, w);
  tile = E.tile_base + (item - first[lo]);
  return E;
}

__device__ __forceinline__ MTileDev tile_of(const MfmaEntry &E, uint32_t tile) {
  typedef uint32_t u32x8v __attribute__((ext_vector_type(8)));
  return __builtin_bit_cast(MTileDev, *(const IHE_CONSTANT u32x8v *)(E.tiles + tile));
}

// Column word of big index c: ciphertext byte offset relative to the
// segment (ct - ct_base) << 13, plus the rotation shift (c % R) * RS << 2.
template <bool ROT>
__device__ __forceinline__ uint32_t col_word(uint32_t c, const MfmaEntry &E) {
  if constexpr (!ROT) {
    return (c - E.ct_base) << 13;
  } else {
    const uint32_t ct = (uint32_t)__umul64hi((uint64_t)c, E.rmagic);
    return ((ct - E.ct_base) << 13) | (((c - ct * E.R) * E.rs) << 2);
  }
}

// Loads word `word` (0..2047: a then b) of column word cw; a rotated column
// reads P[i + m] for i + m < N and -P[i + m - N] past the end.
template <bool ROT>
__device__ __forceinline__ uint32_t load_col(__amdgpu_buffer_rsrc_t in_r, uint32_t cw,
                                             uint32_t word) {
  if constexpr (!ROT) {
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(in_r, cw + word * 4u, 0, 0);
  } else {
    const uint32_t t = (word & (kN - 1)) + ((cw >> 2) & (kN - 1));
    const uint32_t v = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(
        in_r, (cw & ~0x1FFFu) + ((t & (kN - 1)) << 2) + (word & kN) * 4u, 0, 0);
    return t >= kN ? 0u - v : v;
  }
}

// Word `word` of the lane's 16 columns.
template <bool ROT>
__device__ __forceinline__ void load_cols(__amdgpu_buffer_rsrc_t in_r, const uint32_t (&cw)[16],
                                          uint32_t word, uint32_t (&x)[16]) {
#pragma unroll
  for (int j = 0; j < 16; ++j) x[j] = load_col<ROT>(in_r, cw[j], word);
}

// 4 words -> 4 dwords of balanced signed bytes: dword b holds digit b of
// words 0..3, in word order.
__device__ __forceinline__ void byte_quad(const uint32_t *x, uint32_t (&d)[4]) {
  uint32_t y[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) y[i] = (x[i] + 0x80808080u) ^ 0x80808080u;
  const uint32_t t0 = __builtin_amdgcn_perm(y[1], y[0], 0x05010400u);
  const uint32_t t1 = __builtin_amdgcn_perm(y[1], y[0], 0x07030602u);
  const uint32_t t2 = __builtin_amdgcn_perm(y[3], y[2], 0x05010400u);
  const uint32_t t3 = __builtin_amdgcn_perm(y[3], y[2], 0x07030602u);
  d[0] = __builtin_amdgcn_perm(t2, t0, 0x05040100u);
  d[1] = __builtin_amdgcn_perm(t2, t0, 0x07060302u);
  d[2] = __builtin_amdgcn_perm(t3, t1, 0x05040100u);
  d[3] = __builtin_amdgcn_perm(t3, t1, 0x07060302u);
}

// 16 words -> 4 planes of balanced signed bytes (plane b = digit b of each word).
__device__ __forceinline__ void byte_planes(const uint32_t (&x)[16], int4v (&p)[4]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    uint32_t d[4];
    byte_quad(x + 4 * q, d);
#pragma unroll
    for (int b = 0; b < 4; ++b) p[b][q] = (int)d[b];
  }
}

struct Acc {
  int16v s0, s8, s16, s24;  // partial sums at shifts 0, 8, 16, 24
};

__device__ __forceinline__ void mfma_kblock(const int4v &a0, const int4v &a1, const int4v (&pl)[4],
                                            Acc &c) {
  c.s0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, pl[0], c.s0, 0, 0, 0);
  c.s8 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, pl[1], c.s8, 0, 0, 0);
  c.s16 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, pl[2], c.s16, 0, 0, 0);
  c.s24 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, pl[3], c.s24, 0, 0, 0);
  c.s8 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, pl[0], c.s8, 0, 0, 0);
  c.s16 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, pl[1], c.s16, 0, 0, 0);
  c.s24 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, pl[2], c.s24, 0, 0, 0);
}

// (a << 8) + b, kept as one v_lshl_add_u32: the empty asm makes the result
// opaque, so the recombination below stays a chain of three per word (the
// compiler otherwise reassociates it into separate shifts of each accumulator
// and an add3, four per word).
__device__ __forceinline__ uint32_t shl8_add(uint32_t a, uint32_t b) {
  uint32_t d = (a << 8) + b;
  asm("" : "+v"(d));
  return d;
}

// Recombines, applies the b-polynomial bias (samples < NS) and hiding (words
// >= RS), and stores 32 rows x 32 words: row m of lanes 0..31 (and m + 4 of
// lanes 32..63) for value v, m = 8*(v/4) + v%4. Rows at or past nrows store
// out of range. Always 16 stores. n0 (the column tile's first word) is
// wave-uniform, so every case below is a scalar branch: whole column tiles
// inside NS take plain adds, and full tiles store without selects, when SPLIT
// (the per-wave kernels; the 128-row kernel keeps the selects, a single store
// sequence and its row base as computed: its register budget needs them).
template <bool SPLIT>
__device__ __forceinline__ void finish_rows(const Acc &c, const MfmaEntry &E,
                                            __amdgpu_buffer_rsrc_t out_r, const uint32_t *bias,
                                            uint32_t n0, uint32_t r, uint32_t h, uint32_t row0,
                                            uint32_t nrows, bool full, uint32_t rbase) {
  const uint32_t word = n0 + r;
  uint32_t o[16];
#pragma unroll
  for (int v = 0; v < 16; ++v)  // s0 + 2^8 s8 + 2^16 s16 + 2^24 s24, Horner form
    o[v] = shl8_add(shl8_add(shl8_add((uint32_t)c.s24[v], (uint32_t)c.s16[v]), (uint32_t)c.s8[v]),
                    (uint32_t)c.s0[v]);
  if (n0 >= kN && n0 - kN < E.ns) {
    if (SPLIT && n0 + 31 - kN < E.ns) {  // every word of the tile takes the bias
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const u32x4v bv = *(const u32x4v *)(bias + row0 + 8 * g + 4 * h);
#pragma unroll
        for (int q = 0; q < 4; ++q) o[4 * g + q] += bv[q];
      }
    } else {
      const bool lane_bias = word - kN < E.ns;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const u32x4v bv = *(const u32x4v *)(bias + row0 + 8 * g + 4 * h);
#pragma unroll
        for (int q = 0; q < 4; ++q) o[4 * g + q] += lane_bias ? bv[q] : 0u;
      }
    }
  }
  if (n0 >= kN && n0 + 31 - kN >= E.rs) {
    const bool keep = word - kN < E.rs;
#pragma unroll
    for (int v = 0; v < 16; ++v) o[v] = keep ? o[v] : 0u;
  }
  const uint32_t voff = word * 4u + 4 * h * kCtWords * 4u;
  // the row block's byte offset, kept scalar: each store's offset is then this
  // SGPR plus a constant
  uint32_t sbase = rbase + row0 * kCtWords * 4u;
  if constexpr (SPLIT) asm volatile("" : "+s"(sbase));
  if (SPLIT && full) {
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const uint32_t m = 8 * (v / 4) + (v % 4);
      __builtin_amdgcn_raw_buffer_store_b32((int)o[v], out_r, voff,
                                            (int)(sbase + m * kCtWords * 4u), kStorePolicy);
    }
  } else {
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const uint32_t m = 8 * (v / 4) + (v % 4);
      __builtin_amdgcn_raw_buffer_store_b32((int)o[v], out_r,
                                            full || row0 + m + 4 * h < nrows ? voff : kDropOffset,
                                            (int)(sbase + m * kCtWords * 4u), kStorePolicy);
    }
  }
}

// ---------------------------------------------------------------------------
// One K-block (<= 32 columns), RB row blocks per tile. A wave keeps the A
// operands and column words in registers and walks NTW column tiles of 32
// words; each step's B operand feeds RB row blocks, so input words read per
// output word fall to 1/RB. Every step issues 16 loads (the next step's
// columns; the last step reloads its own) followed by 16*RB stores, so the
// loads are waited with the stores still in flight; the first step is peeled
// so the loop is entered with the same pending counts as it is re-entered.
//
// Workgroup block `blk` of the launch (its tile and column part), for thread t;
// s_bias holds 32 RB words of LDS.
template <bool ROT, int NTW, int RB>
__device__ __forceinline__ void mfma1_body(const MfmaArgs &A, uint32_t blk, uint32_t *s_bias,
                                           uint32_t t) {
  constexpr uint32_t wg_per_tile = kCtWords / (kWaves * NTW * 32);
  constexpr uint32_t rows = kMfmaRows * RB;
  uint32_t tile;
  const MfmaEntry E = entry_of(A, blk / wg_per_tile, tile);
  const uint32_t part = blk % wg_per_tile;
  // the wave index is uniform: column-tile positions and their branches stay scalar
  const uint32_t wave = __builtin_amdgcn_readfirstlane(t / 64), lane = t % 64;
  const uint32_t h = lane / 32, r = lane % 32;
  const MTileDev T = tile_of(E, tile);
  const __amdgpu_buffer_rsrc_t in_r =
      __builtin_amdgcn_make_buffer_rsrc((void *)E.in, (short)0, (int)E.in_bytes, kRsrcWord3);
  const __amdgpu_buffer_rsrc_t out_r =
      __builtin_amdgcn_make_buffer_rsrc((void *)E.out, (short)0, (int)E.out_bytes, kRsrcWord3);
  if (t < rows) s_bias[t] = E.rowc[T.row_off + t];
  uint32_t cw[16];
  {
    const IHE_CONSTANT uint32_t *kc = cptr(E.kcols) + T.kcol_off;
    const u32x16v cl = *(const IHE_CONSTANT u32x16v *)kc;
    const u32x16v chh = *(const IHE_CONSTANT u32x16v *)(kc + 16);
#pragma unroll
    for (int j = 0; j < 16; ++j) cw[j] = col_word<ROT>(h ? chh[j] : cl[j], E);
  }
  int4v a[RB][2];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
    a[rb][0] = E.adata[(size_t)(T.a_off + 2 * rb) * 64 + lane];
    a[rb][1] = E.adata[(size_t)(T.a_off + 2 * rb + 1) * 64 + lane];
  }
  // a wave walks NTW consecutive column tiles (interleaving them over the
  // tile's waves measured no faster)
  const uint32_t nt0 = (part * kWaves + wave) * NTW;
  auto col = [&](uint32_t nt) { return nt0 + nt; };
  const uint32_t row_base = T.slot0 * kCtWords * 4u;
  const bool full = T.nrows == rows;
  __syncthreads();

  auto step = [&](const uint32_t (&x)[16], uint32_t nt) {
    int4v pl[4];
    byte_planes(x, pl);
    // the row offsets are formed per step, not hoisted into 16*RB live SGPRs
    uint32_t rbase = row_base;
    asm volatile("" : "+s"(rbase));
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      Acc c = {};
      mfma_kblock(a[rb][0], a[rb][1], pl, c);
      finish_rows<true>(c, E, out_r, s_bias, col(nt) * 32, r, h, 32 * rb, T.nrows, full, rbase);
    }
  };

  uint32_t x[16], xn[16];
  load_cols<ROT>(in_r, cw, col(0) * 32 + r, x);
  load_cols<ROT>(in_r, cw, col(NTW > 1 ? 1 : 0) * 32 + r, xn);
  step(x, 0);
  for (uint32_t nt = 1; nt < (uint32_t)NTW; ++nt) {
#pragma unroll
    for (int j = 0; j < 16; ++j) x[j] = xn[j];
    const uint32_t ntn = nt + 1 < (uint32_t)NTW ? nt + 1 : nt;
    load_cols<ROT>(in_r, cw, col(ntn) * 32 + r, xn);
    step(x, nt);
  }
}

template <bool ROT, int NTW, int RB>
__global__ __launch_bounds__(64 * kWaves) void k_cloud_mfma1(MfmaArgs A) {
  __shared__ uint32_t s_bias[kMfmaRows * RB];
  if (blockIdx.x >= A.cloud_blocks) return var_block(A.var, blockIdx.x - A.cloud_blocks);
  mfma1_body<ROT, NTW, RB>(A, xcd_block(blockIdx.x, A.cloud_blocks), s_bias, threadIdx.x);
}

// ---------------------------------------------------------------------------
// NKB K-blocks (up to NKB*32 columns, wider windows), one row block per tile.
// The tile's A operands and column words live in LDS. Within a column tile the
// K-blocks are unrolled and each issues the next K-block's 16 loads (or the
// next column tile's first) before consuming its own, so its wait leaves 16
// loads, or 16 loads and 16 stores, in flight.
template <bool ROT, int NTW, int NKB>
__global__ __launch_bounds__(64 * kWaves) void k_cloud_mfmak(MfmaArgs A) {
  constexpr uint32_t wg_per_tile = kCtWords / (kWaves * NTW * 32);
  __shared__ uint32_t s_cw[NKB * 32];
  __shared__ int4v s_a[NKB * 2 * 64];
  __shared__ uint32_t s_bias[kMfmaRows];
  if (blockIdx.x >= A.cloud_blocks) return var_block(A.var, blockIdx.x - A.cloud_blocks);
  const uint32_t blk = xcd_block(blockIdx.x, A.cloud_blocks);
  uint32_t tile;
  const MfmaEntry E = entry_of(A, blk / wg_per_tile, tile);
  const uint32_t part = blk % wg_per_tile;
  // the wave index is uniform: column-tile positions and their branches stay scalar
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
  const uint32_t h = lane / 32, r = lane % 32;
  const MTileDev T = tile_of(E, tile);
  const __amdgpu_buffer_rsrc_t in_r =
      __builtin_amdgcn_make_buffer_rsrc((void *)E.in, (short)0, (int)E.in_bytes, kRsrcWord3);
  const __amdgpu_buffer_rsrc_t out_r =
      __builtin_amdgcn_make_buffer_rsrc((void *)E.out, (short)0, (int)E.out_bytes, kRsrcWord3);
  // tiles using fewer K-blocks than NKB are padded with zero-weight blocks
  for (uint32_t i = threadIdx.x; i < NKB * 32; i += 64 * kWaves) {
    const uint32_t kb = i / 32;
    const uint32_t c = E.kcols[T.kcol_off + (kb < T.nkb ? i : (T.nkb - 1) * 32)];
    s_cw[i] = col_word<ROT>(c, E);
  }
  for (uint32_t i = threadIdx.x; i < NKB * 2 * 64; i += 64 * kWaves)
    s_a[i] = i < T.nkb * 2 * 64 ? E.adata[(size_t)T.a_off * 64 + i] : int4v{0, 0, 0, 0};
  if (threadIdx.x < kMfmaRows) s_bias[threadIdx.x] = E.rowc[T.row_off + threadIdx.x];
  const uint32_t nt0 = (part * kWaves + wave) * NTW;
  auto col = [&](uint32_t nt) { return nt0 + nt; };
  const uint32_t row_base = T.slot0 * kCtWords * 4u;
  const bool full = T.nrows == kMfmaRows;
  __syncthreads();

  auto cols_of = [&](uint32_t kb, uint32_t (&cw)[16]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const u32x4v v = *(const u32x4v *)(s_cw + 32 * kb + 16 * h + 4 * q);
#pragma unroll
      for (int i = 0; i < 4; ++i) cw[4 * q + i] = v[i];
    }
  };

  // One column tile. Scheduling barriers keep each K-block's loads ahead of
  // the previous K-block's compute; the next column tile's first K-block is
  // loaded before this tile's stores.
  uint32_t xa[16];
  auto column_tile = [&](uint32_t nt) {
    Acc c = {};
    uint32_t xnext[16];
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb) {
      uint32_t xb[16];
      {
        uint32_t cw[16];
        cols_of(kb + 1 < NKB ? kb + 1 : 0, cw);
        if (kb + 1 < NKB) {
          load_cols<ROT>(in_r, cw, col(nt) * 32 + r, xb);
        } else {
          const uint32_t ntn = nt + 1 < (uint32_t)NTW ? nt + 1 : nt;
          load_cols<ROT>(in_r, cw, col(ntn) * 32 + r, xnext);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
      int4v pl[4];
      byte_planes(xa, pl);
      mfma_kblock(s_a[(2 * kb) * 64 + lane], s_a[(2 * kb + 1) * 64 + lane], pl, c);
      __builtin_amdgcn_sched_barrier(0);
      if (kb + 1 < NKB) {
#pragma unroll
        for (int j = 0; j < 16; ++j) xa[j] = xb[j];
      }
    }
    uint32_t rbase = row_base;
    asm volatile("" : "+s"(rbase));
    finish_rows<true>(c, E, out_r, s_bias, col(nt) * 32, r, h, 0, T.nrows, full, rbase);
#pragma unroll
    for (int j = 0; j < 16; ++j) xa[j] = xnext[j];
  };
  {
    uint32_t cw[16];
    cols_of(0, cw);
    load_cols<ROT>(in_r, cw, col(0) * 32 + r, xa);
  }
  // the first column tile is peeled so that the loop is entered, like it is
  // re-entered, with 16 loads followed by 16 stores outstanding
  column_tile(0);
  for (uint32_t nt = 1; nt < (uint32_t)NTW; ++nt) column_tile(nt);
}

// ---------------------------------------------------------------------------
// Four row blocks per tile (128 rows), one per wave, over NKB K-blocks (wide
// windows; NKB even). The workgroup recodes each 32-word column tile of its
// columns into byte planes once, in LDS, and the four waves take their B
// operands from there, so an input word is loaded once per 128 output rows.
//
// Loading is split in half K-blocks (16 columns x 32 words): a lane loads the
// 16 columns of one word, so each load instruction reads two columns' 128-byte
// lines, and writes the 16-byte B operand of one lane per plane, lane-linear.
// A column tile has 2 NKB half K-blocks: in each of NKB / 4 full rounds a wave
// takes two (one per half-wave); when NKB / 2 is odd a last round gives each
// wave one, which both its halves load (the second half's copy is not written).
// LDS holds two column tiles: while the waves multiply one, the next is recoded
// and the one after is being loaded. Per column tile a lane issues 16 loads per
// round and 16 stores, in a fixed order, and the workgroup meets at one barrier.
// Waves per SIMD asked of the register allocator for the shapes that would
// otherwise land just above an occupancy step and fit below it without
// spilling: rotated with 6 K-blocks (261-264 VGPRs -> 247, 2 waves instead of
// 1), unrotated with 4 (169 -> 168, 3 waves instead of 2). 1 = no request.
template <bool ROT, int NKB>
constexpr int mfmaw_waves() { return ROT ? (NKB == 6 ? 2 : 1) : (NKB == 4 ? 3 : 1); }

//
// Workgroup block `blk` of the launch, for thread t; s_b is [buffer][(kb * 4 +
// plane) * 64 + B lane] and s_bias 128 words of LDS.
template <bool ROT, int NTW, int NKB>
__device__ __forceinline__ void mfmaw_body(const MfmaArgs &A, uint32_t blk,
                                           int4v (*s_b)[NKB * 4 * 64], uint32_t *s_bias,
                                           uint32_t t) {
  static_assert(NKB % 2 == 0, "the four waves split NKB / 2 K-block pairs");
  constexpr uint32_t wg_per_tile = kCtWords / (NTW * 32);
  constexpr uint32_t rows = kMfmaRows * kWaves;
  constexpr int kFull = NKB / 4;                    // rounds of two half K-blocks per wave
  constexpr int kRounds = kFull + (NKB % 4 ? 1 : 0);
  uint32_t tile;
  const MfmaEntry E = entry_of(A, blk / wg_per_tile, tile);
  const uint32_t nt0 = (blk % wg_per_tile) * NTW;
  auto col = [&](uint32_t nt) { return nt0 + nt; };
  // the wave index is uniform: row offsets stay scalar
  const uint32_t wave = __builtin_amdgcn_readfirstlane(t / 64), lane = t % 64;
  const uint32_t h = lane / 32, r = lane % 32;
  const MTileDev T = tile_of(E, tile);
  const __amdgpu_buffer_rsrc_t in_r =
      __builtin_amdgcn_make_buffer_rsrc((void *)E.in, (short)0, (int)E.in_bytes, kRsrcWord3);
  const __amdgpu_buffer_rsrc_t out_r =
      __builtin_amdgcn_make_buffer_rsrc((void *)E.out, (short)0, (int)E.out_bytes, kRsrcWord3);
  if (t < rows) s_bias[t] = E.rowc[T.row_off + t];
  // this wave's A operands; K-blocks past the tile's own are zero
  int4v a[NKB][2];
#pragma unroll
  for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
    for (int hp = 0; hp < 2; ++hp)
      a[kb][hp] = (uint32_t)kb < T.nkb
                      ? E.adata[(size_t)(T.a_off + (wave * T.nkb + kb) * 2 + hp) * 64 + lane]
                      : int4v{0, 0, 0, 0};
  // round rr: K-block and half (16 columns) this lane loads, for word r
  auto kb_of = [&](int rr) -> uint32_t { return rr < kFull ? 4 * rr + wave : 4 * kFull + wave / 2; };
  auto half_of = [&](int rr) -> uint32_t { return rr < kFull ? h : wave % 2; };
  // column words; K-blocks past the tile's own repeat its last one (zero weights)
  uint32_t cw[kRounds][16];
  {
    const IHE_CONSTANT uint32_t *kc = cptr(E.kcols) + T.kcol_off;
#pragma unroll
    for (int rr = 0; rr < kRounds; ++rr) {
      const uint32_t kb = kb_of(rr), kbs = kb < T.nkb ? kb : T.nkb - 1;
      const IHE_CONSTANT u32x4v *c4 = (const IHE_CONSTANT u32x4v *)(kc + 32 * kbs + 16 * half_of(rr));
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const u32x4v v = c4[g];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          cw[rr][4 * g + i] = col_word<ROT>(v[i], E);
          // unrotated columns: the lane's word joins the column word, and the
          // column tile's offset goes in the scalar offset of the load
          if constexpr (!ROT) cw[rr][4 * g + i] += r * 4u;
        }
      }
    }
  }
  const uint32_t row_base = T.slot0 * kCtWords * 4u;
  const bool full = T.nrows == rows;

  uint32_t x[kRounds][16];
  auto load_round = [&](int rr, uint32_t nt) {
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      if constexpr (!ROT)
        x[rr][j] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(in_r, cw[rr][j],
                                                                  (int)(col(nt) * 128u), 0);
      else
        x[rr][j] = load_col<ROT>(in_r, cw[rr][j], col(nt) * 32 + r);
    }
  };
  auto stage_round = [&](int rr, uint32_t buf) {
    int4v pl[4];
    byte_planes(x[rr], pl);
    const uint32_t at = kb_of(rr) * 4 * 64 + 32 * half_of(rr) + r;
    if (rr < kFull || h == 0) {
#pragma unroll
      for (int b = 0; b < 4; ++b) s_b[buf][at + b * 64] = pl[b];
    }
  };

#pragma unroll
  for (int rr = 0; rr < kRounds; ++rr) load_round(rr, 0);
#pragma unroll
  for (int rr = 0; rr < kRounds; ++rr) stage_round(rr, 0);
#pragma unroll
  for (int rr = 0; rr < kRounds; ++rr) load_round(rr, NTW > 1 ? 1 : 0);
  __syncthreads();
  // Column tile nt: the K-blocks' MFMAs read LDS buffer nt % 2, each K-block's
  // B operands read while the previous K-block multiplies; round rr of tile
  // nt + 1 (in registers since the previous tile) is recoded into the other
  // buffer after K-block (rr + 1) NKB / kRounds - 1, and its registers are
  // refilled with tile nt + 2, so the recoding runs beside the MFMAs.
  auto column_tile = [&](uint32_t nt) {
    const uint32_t buf = nt & 1;
    const uint32_t nt2 = nt + 2 < (uint32_t)NTW ? nt + 2 : NTW - 1;
    Acc c = {};
    int4v pl[2][4];
#pragma unroll
    for (int b = 0; b < 4; ++b) pl[0][b] = s_b[buf][b * 64 + lane];
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb) {
      if (kb + 1 < NKB) {
#pragma unroll
        for (int b = 0; b < 4; ++b) pl[(kb + 1) & 1][b] = s_b[buf][((kb + 1) * 4 + b) * 64 + lane];
      }
      mfma_kblock(a[kb][0], a[kb][1], pl[kb & 1], c);
#pragma unroll
      for (int rr = 0; rr < kRounds; ++rr)
        if (kb == (rr + 1) * NKB / kRounds - 1) {
          stage_round(rr, buf ^ 1);
          load_round(rr, nt2);
        }
      __builtin_amdgcn_sched_barrier(0);
    }
    uint32_t rbase = row_base;
    asm volatile("" : "+s"(rbase));
    finish_rows<false>(c, E, out_r, s_bias, col(nt) * 32, r, h, 32 * wave, T.nrows, full, rbase);
    __syncthreads();
  };
  // the first column tile is peeled so that the loop is entered, like it is
  // re-entered, with each round's loads followed by the tile's 16 stores
  column_tile(0);
  for (uint32_t nt = 1; nt < (uint32_t)NTW; ++nt) column_tile(nt);
}

template <bool ROT, int NTW, int NKB>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(mfmaw_waves<ROT, NKB>())))
void k_cloud_mfmaw(MfmaArgs A) {
  __shared__ int4v s_b[2][NKB * 4 * 64];
  __shared__ uint32_t s_bias[kMfmaRows * kWaves];
  if (blockIdx.x >= A.cloud_blocks) return var_block(A.var, blockIdx.x - A.cloud_blocks);
  mfmaw_body<ROT, NTW, NKB>(A, xcd_block(blockIdx.x, A.cloud_blocks), s_b, s_bias, threadIdx.x);
}

// ---------------------------------------------------------------------------
// 128-row tiles at three workgroups per CU. The product is k_cloud_mfmaw's;
// the input words reach LDS by LDS-direct buffer loads (no load registers) and
// the column words stay in LDS, so that a wave fits the 168 registers of three
// waves per SIMD. A rotated column loads its words from their rotated places,
// and the recoding negates those that wrapped past N. Per column tile: the
// workgroup recodes the staged words into the B planes (one buffer), then
// issues the next column tile's loads straight into the staging buffer and
// multiplies while they arrive; the other workgroups of the CU fill each one's
// barriers.
// s_raw: [NKB * 32 columns][32 words]. With dword loads, load instruction q of
// the tile writes columns 2q (lanes 0..31) and 2q + 1 (lanes 32..63); with
// 16-byte loads (W16, where a column's words come in runs of four: unrotated,
// or rotated by multiples of 4) it writes columns 8q..8q + 7, four words per
// lane.
template <bool ROT, bool W16, int NTW, int NKB>
__device__ __forceinline__ void mfmaw_lds_body(const MfmaArgs &A, uint32_t blk, int4v *s_b,
                                               uint32_t *s_raw, uint32_t *s_cw, uint32_t *s_bias,
                                               uint32_t t) {
  constexpr uint32_t wg_per_tile = kCtWords / (NTW * 32);
  constexpr uint32_t rows = kMfmaRows * kWaves;
  uint32_t tile;
  const MfmaEntry E = entry_of(A, blk / wg_per_tile, tile);
  const uint32_t nt0 = (blk % wg_per_tile) * NTW;
  auto col = [&](uint32_t nt) { return nt0 + nt; };
  const uint32_t wave = __builtin_amdgcn_readfirstlane(t / 64), lane = t % 64;
  const uint32_t h = lane / 32, r = lane % 32;
  const MTileDev T = tile_of(E, tile);
  const __amdgpu_buffer_rsrc_t in_r =
      __builtin_amdgcn_make_buffer_rsrc((void *)E.in, (short)0, (int)E.in_bytes, kRsrcWord3);
  const __amdgpu_buffer_rsrc_t out_r =
      __builtin_amdgcn_make_buffer_rsrc((void *)E.out, (short)0, (int)E.out_bytes, kRsrcWord3);
  if (t < rows) s_bias[t] = E.rowc[T.row_off + t];
  // column words; K-blocks past the tile's own repeat its last one (zero weights)
  for (uint32_t i = t; i < NKB * 32; i += 64 * kWaves) {
    const uint32_t kb = i / 32, kbs = kb < T.nkb ? kb : T.nkb - 1;
    s_cw[i] = col_word<ROT>(E.kcols[T.kcol_off + 32 * kbs + i % 32], E);
  }
  int4v a[NKB][2];
#pragma unroll
  for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
    for (int hp = 0; hp < 2; ++hp)
      a[kb][hp] = (uint32_t)kb < T.nkb
                      ? E.adata[(size_t)(T.a_off + (wave * T.nkb + kb) * 2 + hp) * 64 + lane]
                      : int4v{0, 0, 0, 0};
  const uint32_t row_base = T.slot0 * kCtWords * 4u;
  const bool full = T.nrows == rows;
  typedef __attribute__((address_space(3))) void *lds_ptr;
  static_assert((NKB * 4) % kWaves == 0, "16-byte staging loads split evenly over the waves");
  // byte offset of word `word` (0..2047: a then b) of column word cw: a
  // rotated column reads word i + m of its ciphertext for i + m < N, else
  // word i + m - N (negated in the recoding)
  auto word_off = [&](uint32_t cw, uint32_t word) {
    if constexpr (!ROT) {
      return cw + word * 4u;
    } else {
      const uint32_t tt = (word & (kN - 1)) + ((cw >> 2) & (kN - 1));
      return (cw & ~0x1FFFu) + ((tt & (kN - 1)) << 2) + (word & kN) * 4u;
    }
  };
  // W16 with ROT: every column's shift m = (c mod R) * RS is a multiple of 4
  // (RS is), so the 16-byte runs stay whole and none wraps past N
  auto issue_raw = [&](uint32_t nt) {
    if constexpr (W16) {
#pragma unroll
      for (int i = 0; i < NKB * 4 / kWaves; ++i) {
        const uint32_t q = (uint32_t)i * kWaves + wave;
        const uint32_t off = word_off(s_cw[8 * q + lane / 8], col(nt) * 32 + (lane % 8) * 4);
        __builtin_amdgcn_raw_ptr_buffer_load_lds(in_r, (lds_ptr)(s_raw + q * 256), 16, off, 0, 0,
                                                 0);
      }
      return;
    }
#pragma unroll
    for (int i = 0; i < NKB * 4; ++i) {
      const uint32_t q = (uint32_t)i * kWaves + wave;
      const uint32_t off = word_off(s_cw[2 * q + h], col(nt) * 32 + r);
      __builtin_amdgcn_raw_ptr_buffer_load_lds(in_r, (lds_ptr)(s_raw + q * 64), 4, off, 0, 0, 0);
    }
  };
  // task k: K-block k / 64 and B lane k % 64 (half, word) of every plane
  auto recode = [&](uint32_t nt) {
    for (uint32_t k = t; k < NKB * 64; k += 64 * kWaves) {
      const uint32_t kb = k / 64, bl = k % 64;
      const uint32_t c0 = kb * 32 + 16 * (bl / 32);
      const uint32_t *src = s_raw + c0 * 32 + bl % 32;
      uint32_t x[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) x[j] = src[32 * j];
      if constexpr (ROT) {  // words that wrapped past N enter negated
        const uint32_t wn = (col(nt) * 32 + bl % 32) & (kN - 1);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const u32x4v cv = *(const u32x4v *)(s_cw + c0 + 4 * q);
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (wn + ((cv[i] >> 2) & (kN - 1)) >= kN) x[4 * q + i] = 0u - x[4 * q + i];
        }
      }
      int4v pl[4];
      byte_planes(x, pl);
#pragma unroll
      for (int b = 0; b < 4; ++b) s_b[(kb * 4 + b) * 64 + bl] = pl[b];
    }
  };
  __syncthreads();
  issue_raw(0);
  for (uint32_t nt = 0; nt < (uint32_t)NTW; ++nt) {
    // this wave's staging loads have landed (the previous tile's 16 stores,
    // issued after them, may still be in flight), then every wave's have
    if (nt == 0) __builtin_amdgcn_s_waitcnt(0x0F70);        // vmcnt(0)
    else __builtin_amdgcn_s_waitcnt(0x4F70);                // vmcnt(16)
    __syncthreads();
    recode(nt);
    __syncthreads();
    if (nt + 1 < (uint32_t)NTW) issue_raw(nt + 1);
    Acc c = {};
    int4v pl[2][4];
#pragma unroll
    for (int b = 0; b < 4; ++b) pl[0][b] = s_b[b * 64 + lane];
#pragma unroll
    for (int kb = 0; kb < NKB; ++kb) {
      if (kb + 1 < NKB) {
#pragma unroll
        for (int b = 0; b < 4; ++b) pl[(kb + 1) & 1][b] = s_b[((kb + 1) * 4 + b) * 64 + lane];
      }
      mfma_kblock(a[kb][0], a[kb][1], pl[kb & 1], c);
    }
    uint32_t rbase = row_base;
    asm volatile("" : "+s"(rbase));
    finish_rows<false>(c, E, out_r, s_bias, col(nt) * 32, r, h, 32 * wave, T.nrows, full, rbase);
  }
}

template <bool ROT, bool W16, int NTW, int NKB>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(3)))
void k_cloud_mfmaw_lds(MfmaArgs A) {
  __shared__ int4v s_b[NKB * 4 * 64];
  __shared__ __attribute__((aligned(16))) uint32_t s_raw[NKB * 32 * 32];
  __shared__ __attribute__((aligned(16))) uint32_t s_cw[NKB * 32];
  __shared__ uint32_t s_bias[kMfmaRows * kWaves];
  if (blockIdx.x >= A.cloud_blocks) return var_block(A.var, blockIdx.x - A.cloud_blocks);
  mfmaw_lds_body<ROT, W16, NTW, NKB>(A, xcd_block(blockIdx.x, A.cloud_blocks), s_b, s_raw, s_cw,
                                     s_bias, threadIdx.x);
}

// ---------------------------------------------------------------------------
// A whole batch in one launch. The grid's first workgroups each run one block
// of the one-K-block families (32- or 64-row tiles): those blocks are short
// and store-bound, and run as in their own kernel when the dispatcher hands
// them out (taken from a work counter instead, they measured a third slower).
// The workgroups after them stay resident and take items (kPersistBlocks
// consecutive workgroup blocks of one 128-row family, or a block of 256
// output variances) until none are left: XCD x first takes the items of its
// own contiguous range, from its own counter, so neighbouring tiles
// (overlapping column windows) meet in that XCD's L2; then it takes from the
// other XCDs' ranges. They are dispatched as the first ones drain, so the two
// kinds of work overlap at the switch. Every body runs as in its own kernel;
// the register and LDS budgets are the largest body's (two workgroups per CU).
template <int NKB>
__device__ __forceinline__ int4v (*as_planes(int4v *p))[NKB * 4 * 64] {
  return reinterpret_cast<int4v (*)[NKB * 4 * 64]>(p);
}

// The next item from the counter value `got` (lane 0's) of range `victim`:
// that range's item if it has one left, else the first left in the ranges
// after it (more increments, by lane 0), else P.total. Run by a whole wave
// with wave-uniform values, so the bounds are scalar loads: vector loads
// here, issued behind the item's stores, would wait for all of them.
__device__ __forceinline__ uint32_t resolve_item(const PersistArgs &P, uint32_t xcd,
                                                 uint32_t &victim, uint32_t got) {
  uint32_t g = __builtin_amdgcn_readfirstlane(got);
  for (;;) {
    const uint32_t v = P.bounds[victim] + g;
    if (v < P.bounds[victim + 1]) return v;
    victim = (victim + 1) % 8;
    if (victim == xcd) return P.total;
    uint32_t a = 0;
    if (tid() % 64 == 0) a = atomicAdd(&P.queue[victim], 1u);
    g = __builtin_amdgcn_readfirstlane(a);
  }
}

// A family's arguments read through the constant address space: scalar
// loads, so the buffer descriptors the bodies build from them stay scalar
// (read as generic memory, every buffer access of a body took a loop over the
// lanes' descriptors).
__device__ __forceinline__ MfmaArgs family_args(const IHE_CONSTANT PersistFamily *F) {
  MfmaArgs a{};
  a.entries = F->args.entries;
  a.first = F->args.first;
  a.item_entry = F->args.item_entry;
  a.num_entries = F->args.num_entries;
  a.cloud_blocks = F->args.cloud_blocks;
  return a;
}

__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(2)))
void k_cloud_batch(PersistArgs P) {
  __shared__ int4v s_b[2 * kPersistMaxNkb * 4 * 64];
  __shared__ uint32_t s_bias[kMfmaRows * kWaves];
  // one-K-block families: one workgroup block per workgroup, in the
  // dispatcher's order (XCD-contiguous, as in their own kernels), ahead of
  // the resident workgroups
  const IHE_CONSTANT PersistFamily *fams = cptr(P.fams);
  if (blockIdx.x < P.once_blocks) {
    const IHE_CONSTANT uint32_t *once_first = cptr(P.once_first);
    const uint32_t b = xcd_block(blockIdx.x, P.once_blocks);
    uint32_t f = 0;
    while (f + 1 < P.num_once && once_first[f + 1] <= b) ++f;
    const MfmaArgs args = family_args(fams + f);
    const uint32_t blk = b - once_first[f];
    switch (fams[f].body) {
      case 0: return mfma1_body<false, kPersistNtw1, 1>(args, blk, s_bias, threadIdx.x);
      case 1: return mfma1_body<false, kPersistNtw1, 2>(args, blk, s_bias, threadIdx.x);
      case 2: return mfma1_body<true, kPersistNtw1, 1>(args, blk, s_bias, threadIdx.x);
      default: return mfma1_body<true, kPersistNtw1, 2>(args, blk, s_bias, threadIdx.x);
    }
  }
  // items k and k + 1, and the range taken from. The next item is asked for
  // as an item starts: its counter value returns behind this item's loads
  // (the memory counter is in order, stores included), where asking at the
  // end would wait for all of the item's stores. That value is the one
  // vector register kept across the bodies; the rest is derived per body
  // (tid()), as a value hoisted across them would be spilled, and its reloads
  // too would wait for every store in flight.
  __shared__ uint32_t s_item[2], s_victim;
  const uint32_t xcd = blockIdx.x % 8;
  const bool wave0 = __builtin_amdgcn_readfirstlane(threadIdx.x / 64) == 0;
  if (wave0) {
    uint32_t victim = xcd, got = 0;
    if (threadIdx.x == 0) got = atomicAdd(&P.queue[victim], 1u);
    const uint32_t item = resolve_item(P, xcd, victim, got);
    if (threadIdx.x == 0) {
      s_item[0] = item;
      s_victim = victim;
    }
  }
  const IHE_CONSTANT uint32_t *fam_first = cptr(P.fam_first);
  const uint32_t var0 = fam_first[P.num_fams];
  for (uint32_t k = 0;; k ^= 1) {
    // the barrier also ends the previous item's reads of LDS
    __syncthreads();
    const uint32_t item = __builtin_amdgcn_readfirstlane(s_item[k]);
    if (item >= P.total) break;
    uint32_t victim = __builtin_amdgcn_readfirstlane(s_victim), got = 0;
    if (threadIdx.x == 0) got = atomicAdd(&P.queue[victim], 1u);
    if (item >= var0) {
      var_block(P.var, item - var0, tid());
    } else {
      uint32_t f = 0;
      while (f + 1 < P.num_fams && fam_first[f + 1] <= item) ++f;
      const MfmaArgs args = family_args(fams + P.num_once + f);
      const uint32_t body = fams[P.num_once + f].body;
#pragma unroll 1
      for (uint32_t g = 0; g < kPersistBlocks; ++g) {
        const uint32_t blk = (item - fam_first[f]) * kPersistBlocks + g;
        if (blk >= args.cloud_blocks) break;
        if (g) __syncthreads();
        switch (body) {
          case 4: mfmaw_body<false, kPersistNtwW, 2>(args, blk, as_planes<2>(s_b), s_bias, tid()); break;
          case 5: mfmaw_body<false, kPersistNtwW, 4>(args, blk, as_planes<4>(s_b), s_bias, tid()); break;
          case 6: mfmaw_body<false, kPersistNtwW6, 6>(args, blk, as_planes<6>(s_b), s_bias, tid()); break;
          case 7: mfmaw_body<true, kPersistNtwW, 2>(args, blk, as_planes<2>(s_b), s_bias, tid()); break;
          case 8: mfmaw_body<true, kPersistNtwW, 4>(args, blk, as_planes<4>(s_b), s_bias, tid()); break;
          default: mfmaw_body<true, kPersistNtwW6, 6>(args, blk, as_planes<6>(s_b), s_bias, tid()); break;
        }
      }
    }
    if (wave0) {
      const uint32_t next = resolve_item(P, xcd, victim, got);
      if (threadIdx.x == 0) {
        s_item[k ^ 1] = next;
        s_victim = victim;
      }
    }
  }
  // every resident workgroup has found no item left: the last one to get here
  // clears the counters for the next launch (no memset node in a captured graph)
  if (threadIdx.x == 0 && atomicAdd(&P.queue[8], 1u) == gridDim.x - P.once_blocks - 1)
    for (uint32_t i = 0; i <= 8; ++i) atomicExch(&P.queue[i], 0u);
}

// ---------------------------------------------------------------------------
// Output variances, 256 outputs per block.
__device__ __forceinline__ uint32_t div_r(uint32_t c, uint32_t R, uint64_t rmagic) {
  return R == 1 ? c : (uint32_t)__umul64hi((uint64_t)c, rmagic);
}

// One wave per 64 outputs, lane = output. Entries are taken 8 at a time: their
// column, weight and input-variance loads are all issued before the first
// add, and the adds then run in entry order (the order of the model's row).
// Block `chunk` of the launch's variance work; run by k_cloud_var and by the
// cloud kernels' trailing blocks.
__device__ __forceinline__ void var_block(const VarArgs &A, uint32_t chunk, uint32_t t) {
  VarEntry E = A.single;
  if (A.entries) {
    uint32_t lo = 0, hi = A.num_entries;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (A.first[mid] <= chunk) lo = mid;
      else hi = mid;
    }
    E = A.entries[lo];
    chunk -= A.first[lo];
  }
  const uint32_t blk = chunk * 4 + t / 64, lane = t % 64;
  const uint32_t o = blk * 64 + lane;
  if (blk * 64 >= E.num_out) return;
  constexpr int kU = 8;
  double v = 0.0;
  const uint32_t e1 = E.blk_off[blk + 1];
  uint32_t e = E.blk_off[blk];
  // packed entries: (c - base) << 16 | |w|; padding has offset kVarPackedPad
  const bool packed = E.base != nullptr;
  const uint32_t cb = packed && o < E.num_out ? E.base[o] : 0;
  auto take = [&](uint32_t raw, int32_t w2in, bool &keep, double &x, int32_t &w2) {
    uint32_t c = raw;
    bool pad = c == kVarPad;
    w2 = w2in;
    if (packed) {
      pad = (raw >> 16) == kVarPackedPad;
      c = cb + (raw >> 16);
      const uint32_t aw = raw & 0xFFFFu;
      w2 = (int32_t)(aw * aw);
    }
    const uint32_t ct = div_r(c, E.R, E.rmagic);
    keep = !pad && c == ct * E.R;
    x = keep ? E.var_in[ct] : 0.0;
  };
  for (; e + kU <= e1; e += kU) {
    uint32_t c[kU];
    int32_t w2[kU];
#pragma unroll
    for (int j = 0; j < kU; ++j) {
      c[j] = E.col[(size_t)(e + j) * 64 + lane];
      w2[j] = packed ? 0 : E.w2[(size_t)(e + j) * 64 + lane];
    }
    bool keep[kU];
    double x[kU];
#pragma unroll
    for (int j = 0; j < kU; ++j) take(c[j], w2[j], keep[j], x[j], w2[j]);
#pragma unroll
    for (int j = 0; j < kU; ++j)
      if (keep[j]) v = __dadd_rn(v, __dmul_rn((double)w2[j], x[j]));
  }
  for (; e < e1; ++e) {
    const uint32_t c = E.col[(size_t)e * 64 + lane];
    int32_t w2 = packed ? 0 : E.w2[(size_t)e * 64 + lane];
    bool keep;
    double x;
    take(c, w2, keep, x, w2);
    if (keep) v = __dadd_rn(v, __dmul_rn((double)w2, x));
  }
  if (o < E.num_out) E.var_out[o] = v;
}

__global__ __launch_bounds__(256) void k_cloud_var(VarArgs A) { var_block(A, blockIdx.x); }

template <bool ROT, int NTW>
void launch_k1(hipStream_t s, int rb, dim3 grid, const MfmaArgs &A) {
  if (rb == 2) hipLaunchKernelGGL((k_cloud_mfma1<ROT, NTW, 2>), grid, dim3(64 * kWaves), 0, s, A);
  else hipLaunchKernelGGL((k_cloud_mfma1<ROT, NTW, 1>), grid, dim3(64 * kWaves), 0, s, A);
}

template <bool ROT, int NTW>
void launch_kk(hipStream_t s, int nkb, dim3 grid, const MfmaArgs &A) {
  const dim3 block(64 * kWaves);
  switch (nkb) {
    case 2: hipLaunchKernelGGL((k_cloud_mfmak<ROT, NTW, 2>), grid, block, 0, s, A); break;
    case 3: hipLaunchKernelGGL((k_cloud_mfmak<ROT, NTW, 3>), grid, block, 0, s, A); break;
    case 4: hipLaunchKernelGGL((k_cloud_mfmak<ROT, NTW, 4>), grid, block, 0, s, A); break;
    case 5: hipLaunchKernelGGL((k_cloud_mfmak<ROT, NTW, 5>), grid, block, 0, s, A); break;
    case 6: hipLaunchKernelGGL((k_cloud_mfmak<ROT, NTW, 6>), grid, block, 0, s, A); break;
    case 7: hipLaunchKernelGGL((k_cloud_mfmak<ROT, NTW, 7>), grid, block, 0, s, A); break;
    default: hipLaunchKernelGGL((k_cloud_mfmak<ROT, NTW, 8>), grid, block, 0, s, A); break;
  }
}

template <bool ROT, bool W16, int NTW>
void launch_kw_lds(hipStream_t s, int nkb, dim3 grid, const MfmaArgs &A) {
  const dim3 block(64 * kWaves);
  switch ((nkb + 1) / 2) {
    case 1: hipLaunchKernelGGL((k_cloud_mfmaw_lds<ROT, W16, NTW, 2>), grid, block, 0, s, A); break;
    case 2: hipLaunchKernelGGL((k_cloud_mfmaw_lds<ROT, W16, NTW, 4>), grid, block, 0, s, A); break;
    default: hipLaunchKernelGGL((k_cloud_mfmaw_lds<ROT, W16, NTW, 6>), grid, block, 0, s, A); break;
  }
}

// tiles of nkb K-blocks run on the kernel with the next even NKB (padding
// K-blocks carry zero weights). IHE_MFMAW_LDS: 0 = register-staged kernel,
// 1 = LDS-staged with dword staging loads, else LDS-staged with 16-byte
// staging loads where a column's words come in runs of four.
template <bool ROT, int NTW>
void launch_kw(hipStream_t s, int nkb, dim3 grid, const MfmaArgs &A, bool runs4) {
  const dim3 block(64 * kWaves);
  const char *e = std::getenv("IHE_MFMAW_LDS");
  const char mode = e && (e[0] == '0' || e[0] == '1') ? e[0] : '2';
  if (mode != '0' && nkb <= 6) {
    if (mode == '2' && runs4) launch_kw_lds<ROT, true, NTW>(s, nkb, grid, A);
    else launch_kw_lds<ROT, false, NTW>(s, nkb, grid, A);
    return;
  }
  switch ((nkb + 1) / 2) {
    case 1: hipLaunchKernelGGL((k_cloud_mfmaw<ROT, NTW, 2>), grid, block, 0, s, A); break;
    case 2: hipLaunchKernelGGL((k_cloud_mfmaw<ROT, NTW, 4>), grid, block, 0, s, A); break;
    case 3: hipLaunchKernelGGL((k_cloud_mfmaw<ROT, NTW, 6>), grid, block, 0, s, A); break;
    default: hipLaunchKernelGGL((k_cloud_mfmaw<ROT, NTW, 8>), grid, block, 0, s, A); break;
  }
}

template <bool ROT>
void launch_rot(hipStream_t s, const MfmaShape &sh, dim3 grid, const MfmaArgs &A, bool runs4) {
  if (sh.rb == kMfmaWideRb) {
    if (sh.ntw == 32) launch_kw<ROT, 32>(s, sh.nkb, grid, A, runs4);
    else if (sh.ntw == 16) launch_kw<ROT, 16>(s, sh.nkb, grid, A, runs4);
    else launch_kw<ROT, 8>(s, sh.nkb, grid, A, runs4);
  } else if (sh.nkb == 1) {
    if (sh.ntw == 16) launch_k1<ROT, 16>(s, sh.rb, grid, A);
    else if (sh.ntw == 8) launch_k1<ROT, 8>(s, sh.rb, grid, A);
    else launch_k1<ROT, 4>(s, sh.rb, grid, A);
  } else {
    if (sh.ntw == 16) launch_kk<ROT, 16>(s, sh.nkb, grid, A);
    else if (sh.ntw == 8) launch_kk<ROT, 8>(s, sh.nkb, grid, A);
    else launch_kk<ROT, 4>(s, sh.nkb, grid, A);
  }
}

}  // namespace

void launch_mfma(hipStream_t stream, const MfmaShape &shape, bool rot, const MfmaEntry &single,
                 const MfmaEntry *entries, const uint32_t *first, uint32_t num_entries,
                 uint32_t num_items, const VarArgs &var, const uint32_t *item_entry,
                 bool rot_runs4) {
  if (num_items == 0) return launch_var(stream, var);
  const uint32_t cloud_blocks = num_items * mfma_wg_per_tile(shape);
  const MfmaArgs A = {single, entries, first, item_entry, num_entries, cloud_blocks, var};
  const dim3 grid(cloud_blocks + var.num_blocks);
  if (rot) launch_rot<true>(stream, shape, grid, A, entries ? rot_runs4 : single.rs % 4 == 0);
  else launch_rot<false>(stream, shape, grid, A, true);
}

int persist_body(const MfmaShape &shape, bool rot) {
  static_assert(kPersistNtw1 > 0 && kPersistNtwW > 0 && kPersistNtwW6 > 0, "column tiles");
  if (shape.nkb == 1 && (shape.rb == 1 || shape.rb == 2)) return (rot ? 2 : 0) + shape.rb - 1;
  if (shape.rb == kMfmaWideRb && shape.nkb >= 1 && shape.nkb <= (int)kPersistMaxNkb)
    return (rot ? 7 : 4) + (shape.nkb + 1) / 2 - 1;
  return -1;
}

uint32_t persist_wg_per_tile(const MfmaShape &shape) {
  MfmaShape s = shape;
  s.ntw = shape.rb != kMfmaWideRb ? kPersistNtw1 : shape.nkb > 4 ? kPersistNtwW6 : kPersistNtwW;
  return mfma_wg_per_tile(s);
}

int persist_grid(int device, uint32_t &grid) {
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      cus <= 0)
    return fail(IHE_ERR_HIP, "CU count query failed");
  grid = 8 * ((2u * (uint32_t)cus + 7) / 8);  // two workgroups per CU, a multiple of the 8 XCDs
  return IHE_OK;
}

void launch_persistent(hipStream_t stream, const PersistArgs &P, uint32_t grid) {
  hipLaunchKernelGGL(k_cloud_batch, dim3(P.once_blocks + grid), dim3(64 * kWaves), 0, stream, P);
}

void launch_var(hipStream_t stream, const VarArgs &var) {
  if (var.num_blocks == 0) return;
  hipLaunchKernelGGL(k_cloud_var, dim3(var.num_blocks), dim3(256), 0, stream, var);
}

}  // namespace ihe
