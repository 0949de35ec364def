// Exact negacyclic products a*s mod (X^N + 1, 2^32) for encrypt and decrypt.
//
// The reference computes b = a*s + e (tLweSymEncryptZero, eval/idash.cpp:625)
// and phase = b - a*s (tLwePhase, :714) through TFHE's floating-point FFT.
// Here the key s is binary, and a*s is computed exactly two ways:
//  - k_ring_mfma (the default for both directions): a product of the a-rows
//    with the key's negacyclic Toeplitz matrix on int8 MFMA, described there;
//    encryption first runs k_encrypt_ae for a and b = e + messages.
//  - k_encrypt / k_decrypt<J> (IHE_ENCRYPT_KERNEL / IHE_DECRYPT_KERNEL =
//    valu): (a*s)[j] = sum_{k : s_k = 1} A2[j - k + N], A2 = [-a[0..N), a[0..N)],
//    a sum of signed coefficients. One wavefront owns one ciphertext: a is
//    staged in LDS as A2 (padded one word per J so lanes reading J-strided
//    words hit distinct banks), lane l owns the J consecutive outputs
//    j = l*J + u and slides a J-register window down A2 as k advances, so each
//    key bit costs one LDS read per lane plus, when the bit is set, J adds. The
//    key bits are kernel arguments, hence wave-uniform branches.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "ihe_internal.h"

// Work counters of ihe_decrypt_noalloc: the ring kernel's per-XCD item
// counters and its count of finished workgroups, in a file-scope device array
// instead of a context's allocation. Each GPU holds its own copy for the life
// of the process; host code uses the address hipGetSymbolAddress gives on
// that device, never this name.
static __device__ __align__(256) uint32_t scratch[ihe::kRingQueueWords];

namespace ihe {

constexpr int kWavesPerBlock = 4;

template <int J>
__device__ __forceinline__ uint32_t lds_pad(uint32_t t) {
  if constexpr (J == 1) return t;
  else return t + t / J;
}

template <int J>
constexpr uint32_t lds_words() {
  return J == 1 ? 2 * kN : 2 * kN + 2 * kN / J;
}

// acc[u] = (a*s)[jb + u] for u < J, A2 staged (padded) in lds.
template <int J>
__device__ __forceinline__ void negacyclic_window(const int32_t *lds, const KeyBits &key,
                                                  uint32_t jb, uint32_t acc[J]) {
  uint32_t win[J];
#pragma unroll
  for (int u = 0; u < J; ++u) {
    acc[u] = 0;
    win[u] = (uint32_t)lds[lds_pad<J>(jb + kN + u)];
  }
  // At step k the logical element u (= A2[jb + N + u - k]) sits in
  // win[(u - k) mod J]; after the step the slot of u = J-1 receives the new
  // logical u = 0 of step k+1.
  for (uint32_t kb = 0; kb < kN; kb += J) {
    const uint32_t bits = (key.w[kb / 32] >> (kb % 32)) & (J == 32 ? 0xFFFFFFFFu : ((1u << J) - 1u));
#pragma unroll
    for (int d = 0; d < J; ++d) {
      if ((bits >> d) & 1u) {
#pragma unroll
        for (int u = 0; u < J; ++u) acc[u] += win[(u - d) & (J - 1)];
      }
      const uint32_t t = jb + kN - (kb + (uint32_t)d) - 1u;
      win[(J - 1 - d) & (J - 1)] = (uint32_t)lds[lds_pad<J>(t)];
    }
  }
}

template <int J>
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_decrypt(const int32_t *__restrict__ in,
                                                                 uint32_t num_cts, uint32_t ns,
                                                                 KeyBits key,
                                                                 float *__restrict__ out) {
  __shared__ int32_t lds_all[kWavesPerBlock][lds_words<J>()];
  const uint32_t wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const uint32_t ct = blockIdx.x * kWavesPerBlock + wave;
  const bool active = ct < num_cts;
  int32_t *lds = lds_all[wave];
  if (active) {
    const int32_t *a = in + (size_t)ct * kCtWords;
    for (uint32_t t = lane; t < kN; t += 64) {
      const int32_t x = a[t];
      lds[lds_pad<J>(t)] = (int32_t)(0u - (uint32_t)x);
      lds[lds_pad<J>(t + kN)] = x;
    }
  }
  __syncthreads();
  if (!active) return;
  const uint32_t jb = lane * J;
  if (jb >= ns) return;
  uint32_t acc[J];
  negacyclic_window<J>(lds, key, jb, acc);
  const int32_t *b = in + (size_t)ct * kCtWords + kN;
#pragma unroll
  for (int u = 0; u < J; ++u) {
    const uint32_t j = jb + u;
    if (j < ns) {
      const int32_t phase = (int32_t)((uint32_t)b[j] - acc[u]);
      // t32tod then float: the int32 -> float conversion rounds to nearest
      // even; scaling by 2^-32 is exact.
      out[(size_t)ct * ns + j] = (float)phase * 0x1p-32f;
    }
  }
}

// Philox4x32-10 (Salmon et al., SC'11).
__device__ __forceinline__ void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                       uint32_t k0, uint32_t k1, uint32_t r[4]) {
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    // one 32 x 32 -> 64-bit product per multiplier
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  r[0] = c0;
  r[1] = c1;
  r[2] = c2;
  r[3] = c3;
}

// |e| = smallest k with (u >> 1) < table[k]; sign from bit 0. The table is
// non-decreasing and table[LEN - 1] = 2^63 exceeds every m, so k is the count
// of entries <= m among the first LEN - 1, found in a fixed 11 steps (no
// data-dependent loop, so the samples of a lane interleave).
__device__ __forceinline__ int32_t noise_sample(uint64_t u, const uint64_t *__restrict__ table) {
  static_assert(IHE_NOISE_TABLE_LEN - 1 <= 2047, "11 steps cover the table");
  const uint64_t m = u >> 1;
  uint32_t k = 0;
#pragma unroll
  for (uint32_t step = 1024; step > 0; step >>= 1) {
    const uint32_t nk = k + step;
    const uint64_t tv = table[min(nk, IHE_NOISE_TABLE_LEN - 1) - 1];
    k = nk <= IHE_NOISE_TABLE_LEN - 1 && tv <= m ? nk : k;
  }
  return (u & 1) ? -(int32_t)k : (int32_t)k;
}

// The same sample, the search started from the bucket index (build_noise_lut):
// three steps where the bucket spans at most 7 entries (all but the tails
// beyond about 3.5 sigma), the full eleven from the bucket's low end otherwise.
__device__ __forceinline__ int32_t noise_sample_lut(uint64_t u, const uint64_t *table,
                                                    const uint32_t *lut) {
  const uint64_t m = u >> 1;
  const uint32_t e = lut[m >> (63 - kNoiseLutBits)];
  uint32_t k = e & 0xFFFFu;
  auto step_to = [&](uint32_t step) {
    const uint32_t nk = k + step;
    const uint64_t tv = table[min(nk, IHE_NOISE_TABLE_LEN - 1) - 1];
    k = nk <= IHE_NOISE_TABLE_LEN - 1 && tv <= m ? nk : k;
  };
  if ((e >> 16) <= 7) {
#pragma unroll
    for (uint32_t step = 4; step > 0; step >>= 1) step_to(step);
  } else {
#pragma unroll
    for (uint32_t step = 1024; step > 0; step >>= 1) step_to(step);
  }
  return (u & 1) ? -(int32_t)k : (int32_t)k;
}

// ---------------------------------------------------------------------------
// Decryption and encryption on int8 MFMA. For every ciphertext c and word j,
//   (a_c * s)[j] = sum_i a_c[i] T[i][j],  T[i][j] = S[j - i],
// with S[m] = s[m] for m >= 0 and -s[m + N] for m < 0 (the negacyclic wrap),
// so decryption is the product of the ciphertexts' a-rows with one Toeplitz
// matrix of entries in {-1, 0, 1}. Each a word is recoded into balanced
// signed bytes (x = sum_p 2^(8p) d_p, d_p = byte_p(x + 0x80808080) - 128), one
// v_mfma_i32_32x32x32_i8 per byte plane and K-block of 32 coefficients, four
// int32 accumulators recombined with shifts mod 2^32 (|partial| <= 2^17).
//
// Work items are 32 ciphertexts x 256 samples; a workgroup of 4 waves owns one
// item at a time, each wave two 32-sample column tiles (so each A operand read
// feeds eight MFMAs). The workgroups are persistent: each builds the key's B
// operands once, then takes items from its XCD's counter, so the sample groups
// of one ciphertext group run side by side on one XCD (sharing its L2).
//  - B: T's rows 32 ks + 16 h + jj, column j, are the 16 consecutive bytes
//    u[N - j + 32 ks + 16 h ...] of u[N + x] = S[-x]. LDS holds 16 copies of u
//    shifted by 0..15 bytes, so every B operand is one aligned 16-byte read;
//    copies are 2064 bytes apart, which puts neighbouring lanes on different
//    banks.
//  - A: per stage of 128 coefficients the workgroup loads the 32 x 128 a-words
//    (each load instruction reads 8 rows x 128 bytes, one stage ahead of use),
//    writes their byte planes to LDS (double-buffered, one barrier per stage),
//    and every wave reads its A operands back lane-linear.
// Decryption (words j < NS) ends with phase = b - a*s as a float; encryption
// (all N words; k_encrypt_ae has written a and b = e + messages) adds a*s to
// b in place. An item's results are written in the next item's first stage.
constexpr uint32_t kDecCols = 256;              // samples per work item
constexpr uint32_t kDecUStride = 2 * kN + 16;   // bytes between shifted copies of u

struct DecryptParams {
  const int32_t *in;
  float *out;     // decryption: phases [num_cts][ns]
  int32_t *bout;  // encryption: the ciphertexts of `in` (same buffer), b halves updated
  uint32_t num_cts, ns, col_groups, num_items;
  // kRingQueueWords zero words: items are taken from per-XCD counters; null:
  // workgroup x takes items x, x + 8 ... of its XCD's range
  uint32_t *queue;
};


// u[y] = S[N - y]: s[N - y] for 1 <= y <= N, -s[2N - y] for N < y < 2N, else 0
// (key words in LDS).
__device__ __forceinline__ uint32_t u_byte(uint32_t y, const uint32_t *kw) {
  if (y >= 1 && y <= kN) {
    const uint32_t m = kN - y;
    return (kw[m / 32] >> (m % 32)) & 1u;
  }
  if (y > kN && y < 2 * kN) {
    const uint32_t m = 2 * kN - y;
    return ((kw[m / 32] >> (m % 32)) & 1u) ? 0xFFu : 0u;
  }
  return 0u;
}

// TW column tiles per wave, 8 / TW waves per workgroup: TW = 2 reads each A
// operand into 8 MFMAs at 2 waves per SIMD, TW = 1 into 4 at 4 waves per SIMD.
template <bool ENC, int TW>
__global__ __launch_bounds__(64 * (8 / TW)) __attribute__((amdgpu_waves_per_eu(4 / TW, 4 / TW)))
void k_ring_mfma(DecryptParams P, KeyBits key) {
  typedef int int4v __attribute__((ext_vector_type(4)));
  typedef int int16v __attribute__((ext_vector_type(16)));
  constexpr uint32_t kDecWaves = 8 / TW;
  // loader role: kLpr lanes per ciphertext row, each loading kNl 16-byte
  // pieces kStride words apart per stage (each load instruction reads whole
  // 128-byte lines of 64 * 16 / (16 * kLpr) rows)
  constexpr uint32_t kLpr = 2 * kDecWaves, kNl = 32 / kLpr, kStride = 4 * kLpr;
  __shared__ __attribute__((aligned(16))) uint8_t U[16 * kDecUStride];
  // byte planes of one stage: [buffer][kb][plane][k half][row][4 dwords]
  __shared__ __attribute__((aligned(16))) uint32_t planes[2][4][4][2][32][4];
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
  const uint32_t h = lane / 32, c = lane % 32;
  __shared__ uint32_t kw[kN / 32];
  if (threadIdx.x < kN / 32) kw[threadIdx.x] = key.w[threadIdx.x];
  __syncthreads();
  // U[s * kDecUStride + y] = u[y + s]
  for (uint32_t t = threadIdx.x; t < 16 * (2 * kN / 4); t += 64 * kDecWaves) {
    const uint32_t sh = t / (2 * kN / 4), d = t % (2 * kN / 4);
    uint32_t v = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) v |= u_byte(4 * d + sh + i, kw) << (8 * i);
    *(uint32_t *)(U + sh * kDecUStride + 4 * d) = v;
  }
  __syncthreads();
  // items of XCD x (blocks x, x + 8, ...): a contiguous range, walked in order.
  // From the counter, a workgroup takes the next item when it starts one, so
  // the sample groups of a ciphertext group run side by side all the way
  // (a fixed stride lets workgroups drift apart, and their shared a-words
  // then leave L2 between the reads).
  const uint32_t xcd = blockIdx.x % 8, per_xcd = gridDim.x / 8, slot = blockIdx.x / 8;
  const uint32_t chunk = (P.num_items + 7) / 8;
  const uint32_t it_beg = min(xcd * chunk, P.num_items);
  const uint32_t it_end = min((xcd + 1) * chunk, P.num_items);
  // the item after the k-th is announced in s_next[(k + 1) % 2]
  __shared__ uint32_t s_next[2];
  const uint32_t r = threadIdx.x / kLpr, cq = threadIdx.x % kLpr;
  // an item's ciphertext group starts at a wave-uniform address; a loader
  // thread adds its row (clamped to the last ciphertext) as a byte offset
  auto group_of = [&](uint32_t item) {
    return (const char *)(P.in + (size_t)(item / P.col_groups) * 32 * kCtWords);
  };
  auto row_off = [&](uint32_t item) {
    return (min(r, P.num_cts - 1 - (item / P.col_groups) * 32) * kCtWords + 4 * cq) * 4;
  };
  int4v xr[kNl];
  uint32_t item = it_beg + slot;
  if (P.queue) {
    if (threadIdx.x == 0) s_next[0] = it_beg + atomicAdd(&P.queue[xcd], 1u);
    __syncthreads();
    item = __builtin_amdgcn_readfirstlane(s_next[0]);
  }
  if (item < it_end) {
    const char *g = group_of(item);
    const uint32_t o = row_off(item);
#pragma unroll
    for (uint32_t i = 0; i < kNl; ++i) xr[i] = *(const int4v *)(g + o + 4 * kStride * i);
  }
  uint32_t next = 0;
  // An item's results are written in the next item's first stage, after that
  // stage's loads are issued: the memory counter covers stores too, so waits
  // for loads issued after the stores would also wait for the stores.
  int16v acc[TW][4];
  uint32_t bw[16];  // first column tile's b words, loaded in the last stage
  bool pend = false;
  uint32_t p_ct0 = 0, p_jw = 0;
  // decryption needs the words j < NS, encryption all N
  const uint32_t lim = ENC ? kN : P.ns;
  auto whole = [&](uint32_t ct0, uint32_t jw) {
    return ct0 + 32 <= P.num_cts && jw + 32 * TW <= lim;
  };
  // whole tiles: one wave-uniform base per item, 32-bit byte offsets (row v of
  // a tile is ciphertext ct0 + 8 (v / 4) + 4 h + v % 4)
  auto in_off = [&](uint32_t jw, int v) {
    return (4 * h * kCtWords + kN + jw + c + (8 * (v / 4) + (v % 4)) * kCtWords) * 4;
  };
  auto load_bw1 = [&](uint32_t ct0, uint32_t jw, uint32_t (&bw1)[16]) {
    const char *in_base = (const char *)(P.in + (size_t)ct0 * kCtWords);
#pragma unroll
    for (int v = 0; v < 16; ++v) bw1[v] = *(const uint32_t *)(in_base + in_off(jw, v) + 128);
  };
  // decryption: phase = b - a*s; encryption: b += a*s in place
  auto finish = [&](uint32_t ct0, uint32_t jw, const uint32_t (&bw1)[16]) {
    auto as_of = [&](int t, int v) {
      return ((((((uint32_t)acc[t][3][v] << 8) + (uint32_t)acc[t][2][v]) << 8) +
               (uint32_t)acc[t][1][v]) << 8) + (uint32_t)acc[t][0][v];
    };
    if (whole(ct0, jw)) {  // streaming (nt) stores: outputs are not re-read here
      char *out_base = ENC ? (char *)(P.bout + (size_t)ct0 * kCtWords)
                           : (char *)(P.out + (size_t)ct0 * P.ns);
      const uint32_t out_lane = (4 * h * P.ns + jw + c) * 4;
#pragma unroll
      for (int t = 0; t < TW; ++t) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const uint32_t as = as_of(t, v), b = t == 0 ? bw[v] : bw1[v];
          if constexpr (ENC)
            __builtin_nontemporal_store((int32_t)(b + as),
                                        (int32_t *)(out_base + in_off(jw, v) + 128 * t));
          else
            __builtin_nontemporal_store(
                (float)(int32_t)(b - as) * 0x1p-32f,
                (float *)(out_base + out_lane + (8 * (v / 4) + (v % 4)) * P.ns * 4 + 128 * t));
        }
      }
      return;
    }
    const char *ag = (const char *)(P.in + (size_t)ct0 * kCtWords);
    const uint32_t jc1 = min(jw + 32 + c, lim - 1);
#pragma unroll
    for (int t = 0; t < TW; ++t) {
      const uint32_t j = jw + 32 * t + c;
      if (jw + 32 * t >= lim) continue;
      if (t == 1) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const uint32_t ct = min(ct0 + 8 * (v / 4) + 4 * h + (v % 4), P.num_cts - 1);
          bw[v] = *(const uint32_t *)(ag + ((ct - ct0) * kCtWords + kN + jc1) * 4);
        }
      }
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const uint32_t ct = ct0 + 8 * (v / 4) + 4 * h + (v % 4);
        const uint32_t as = as_of(t, v);
        if (ct >= P.num_cts || j >= lim) continue;
        if constexpr (ENC)
          P.bout[(size_t)ct * kCtWords + kN + j] = (int32_t)(bw[v] + as);
        else
          P.out[(size_t)ct * P.ns + j] = (float)(int32_t)(bw[v] - as) * 0x1p-32f;
      }
    }
  };
  // byte planes of the loaded stage into planes[buf]
  auto recode = [&](uint32_t buf) {
#pragma unroll
    for (uint32_t i = 0; i < kNl; ++i) {
      // coefficients w .. w + 3 of the stage: K-block w / 32, bytes k0 .. k0 + 3
      const uint32_t w = kStride * i + 4 * cq, kb = w / 32, k0 = w % 32;
      uint32_t y[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) y[q] = ((uint32_t)xr[i][q] + 0x80808080u) ^ 0x80808080u;
      const uint32_t t0 = __builtin_amdgcn_perm(y[1], y[0], 0x05010400u);
      const uint32_t t1 = __builtin_amdgcn_perm(y[1], y[0], 0x07030602u);
      const uint32_t t2 = __builtin_amdgcn_perm(y[3], y[2], 0x05010400u);
      const uint32_t t3 = __builtin_amdgcn_perm(y[3], y[2], 0x07030602u);
      planes[buf][kb][0][k0 / 16][r][(k0 % 16) / 4] = __builtin_amdgcn_perm(t2, t0, 0x05040100u);
      planes[buf][kb][1][k0 / 16][r][(k0 % 16) / 4] = __builtin_amdgcn_perm(t2, t0, 0x07060302u);
      planes[buf][kb][2][k0 / 16][r][(k0 % 16) / 4] = __builtin_amdgcn_perm(t3, t1, 0x05040100u);
      planes[buf][kb][3][k0 / 16][r][(k0 % 16) / 4] = __builtin_amdgcn_perm(t3, t1, 0x07060302u);
    }
  };
  // stage st's K-blocks. A wave with any active tile multiplies all of them
  // (a tile past NS costs only its multiplications: its B reads stay inside
  // u), so the K-block loop has no branches and one wait per K-block.
  auto multiply = [&](uint32_t st, uint32_t jw) {
    const uint32_t buf = st & 1;
#pragma unroll
    for (uint32_t kb = 0; kb < 4; ++kb) {
      const uint32_t ks = 4 * st + kb;
      int4v a[4], bop[TW];
#pragma unroll
      for (int p = 0; p < 4; ++p) a[p] = *(const int4v *)&planes[buf][kb][p][h][c][0];
#pragma unroll
      for (int t = 0; t < TW; ++t) {
        const uint32_t start = kN - (jw + 32 * t) + 32 * ks - c + 16 * h;  // in [1, 2N - 16]
        const uint32_t sh = start & 15u;
        bop[t] = *(const int4v *)(U + sh * kDecUStride + (start - sh));
      }
#pragma unroll
      for (int t = 0; t < TW; ++t)
#pragma unroll
        for (int p = 0; p < 4; ++p)
          acc[t][p] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[p], bop[t], acc[t][p], 0, 0, 0);
    }
  };
  for (uint32_t k = 1; item < it_end; item = next, k ^= 1) {
    const uint32_t ct0 = (item / P.col_groups) * 32;
    const uint32_t jw = (item % P.col_groups) * kDecCols + wave * 32 * TW;
    const char *ag = group_of(item);
    const uint32_t ao = row_off(item);
    const uint32_t jc0 = min(jw + c, lim - 1);
    {
      // first stage: take the next item from the counter (announced at this
      // stage's barrier), write the previous item's results after issuing
      // this stage's loads
      const bool grab = P.queue && threadIdx.x == 0;
      uint32_t taken = 0;
      if (grab) taken = atomicAdd(&P.queue[xcd], 1u);  // first used at the barrier
      recode(0);
      uint32_t bw1[16];
      if (TW == 2 && pend && whole(p_ct0, p_jw)) load_bw1(p_ct0, p_jw, bw1);
#pragma unroll
      for (uint32_t i = 0; i < kNl; ++i) xr[i] = *(const int4v *)(ag + ao + 4 * (128 + kStride * i));
      if (pend) finish(p_ct0, p_jw, bw1);
#pragma unroll
      for (int t = 0; t < TW; ++t)
#pragma unroll
        for (int p = 0; p < 4; ++p) acc[t][p] = int16v{};
      if (grab) s_next[k] = taken;
      __syncthreads();
      if (jw < lim) multiply(0, jw);
    }
    for (uint32_t st = 1; st < kN / 128; ++st) {
      recode(st & 1);
      if (st + 1 < kN / 128) {
#pragma unroll
        for (uint32_t i = 0; i < kNl; ++i)
          xr[i] = *(const int4v *)(ag + ao + 4 * (128 * (st + 1) + kStride * i));
      } else {
        // the first column tile's b words, then the next item's first stage:
        // both arrive while this stage multiplies
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const uint32_t ct = min(ct0 + 8 * (v / 4) + 4 * h + (v % 4), P.num_cts - 1);
          bw[v] = *(const uint32_t *)(ag + ((ct - ct0) * kCtWords + kN + jc0) * 4);
        }
        next = P.queue ? it_beg + __builtin_amdgcn_readfirstlane(s_next[k]) : item + per_xcd;
        if (next < it_end) {
          const char *ng = group_of(next);
          const uint32_t no = row_off(next);
#pragma unroll
          for (uint32_t i = 0; i < kNl; ++i) xr[i] = *(const int4v *)(ng + no + 4 * kStride * i);
        }
      }
      __syncthreads();
      if (jw < lim) multiply(st, jw);
    }
    pend = true;
    p_ct0 = ct0;
    p_jw = jw;
  }
  if (pend) {
    uint32_t bw1[16];
    if (TW == 2 && whole(p_ct0, p_jw)) load_bw1(p_ct0, p_jw, bw1);
    finish(p_ct0, p_jw, bw1);
  }
  // every workgroup has taken its last (empty) item: the last one to finish
  // clears the counters for the next launch
  if (P.queue && threadIdx.x == 0) {
    if (atomicAdd(&P.queue[8], 1u) == gridDim.x - 1)
      for (uint32_t i = 0; i <= 8; ++i) atomicExch(&P.queue[i], 0u);
  }
}

struct EncryptParams {
  const int8_t *__restrict__ geno;
  const uint64_t *__restrict__ table;
  const uint32_t *__restrict__ lut;
  int32_t *__restrict__ out;
  double *__restrict__ out_var;
  uint32_t num_cts, num_cols, ns, R, rs;
  uint32_t seed_lo, seed_hi;
  uint64_t rs_recip;  // ceil(2^32 / rs): j / rs = (j * rs_recip) >> 32 for every word j < N
};

// One wavefront per ciphertext c (FeatIndex):
//   a = Philox stream 0, b = a*s + e + one-hot messages (setScore); without
//   genotypes (geno == nullptr) the ciphertexts encrypt zero (tLweSymEncryptZero).
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_encrypt(EncryptParams P, KeyBits key) {
  constexpr int J = 16;
  __shared__ int32_t lds_all[kWavesPerBlock][lds_words<J>()];
  const uint32_t wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const uint32_t ct = blockIdx.x * kWavesPerBlock + wave;
  const bool active = ct < P.num_cts;
  int32_t *lds = lds_all[wave];
  const uint32_t jb = lane * J;
  if (active) {
    int32_t *a = P.out + (size_t)ct * kCtWords;
#pragma unroll
    for (int q = 0; q < J / 4; ++q) {
      uint32_t r[4];
      const uint32_t w0 = jb + 4 * q;
      philox(w0 / 4, ct, 0, 0, P.seed_lo, P.seed_hi, r);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        a[w0 + t] = (int32_t)r[t];
        lds[lds_pad<J>(w0 + t)] = (int32_t)(0u - r[t]);
        lds[lds_pad<J>(w0 + t + kN)] = (int32_t)r[t];
      }
    }
  }
  __syncthreads();
  if (!active) return;
  uint32_t acc[J];
  negacyclic_window<J>(lds, key, jb, acc);
  int32_t *b = P.out + (size_t)ct * kCtWords + kN;
#pragma unroll
  for (int q = 0; q < J / 2; ++q) {
    uint32_t r[4];
    const uint32_t j0 = jb + 2 * q;
    philox(j0 / 2, ct, 1, 0, P.seed_lo, P.seed_hi, r);
    const uint64_t u0 = (uint64_t)r[0] | ((uint64_t)r[1] << 32);
    const uint64_t u1 = (uint64_t)r[2] | ((uint64_t)r[3] << 32);
    acc[2 * q] += (uint32_t)noise_sample(u0, P.table);
    acc[2 * q + 1] += (uint32_t)noise_sample(u1, P.table);
  }
#pragma unroll
  for (int u = 0; u < J; ++u) {
    const uint32_t j = jb + u;
    const uint32_t r = (uint32_t)(((uint64_t)j * P.rs_recip) >> 32), s = j - r * P.rs;
    uint32_t v = acc[u];
    if (P.geno && r < P.R && s < P.ns) {
      const uint32_t bidx = ct * P.R + r;
      if (bidx < P.num_cols) {
        const uint32_t tag = bidx / 3, val = bidx - 3 * tag;
        const int8_t g = P.geno[(size_t)tag * P.ns + s];
        if (g == 0 || g == 1 || g == 2) {
          if ((uint32_t)g == val) v += kOne;
        } else {
          v += val == 0 ? kNan[0] : (val == 1 ? kNan[1] : kNan[2]);
        }
      }
    }
    b[j] = (int32_t)v;
  }
  if (lane == 0) P.out_var[ct] = 0x1p-50;
}

// Everything of encryption but a*s, with k_encrypt's streams and counters: a
// (Philox stream 0, one counter per four words), b = e + messages (stream 1,
// one counter per two words, e sampled from the CDF table), and the variance.
// k_ring_mfma<true> then adds a*s to b. Persistent workgroups copy the CDF
// table to LDS (its searches are scattered reads) and take ciphertexts in
// turn, a thread four words j..j+3 of each.
__global__ __launch_bounds__(256) void k_encrypt_ae(EncryptParams P) {
  typedef uint32_t u32x4v __attribute__((ext_vector_type(4)));
  __shared__ uint64_t s_table[IHE_NOISE_TABLE_LEN];
  __shared__ uint32_t s_lut[1u << kNoiseLutBits];
  for (uint32_t t = threadIdx.x; t < IHE_NOISE_TABLE_LEN; t += 256) s_table[t] = P.table[t];
  for (uint32_t t = threadIdx.x; t < (1u << kNoiseLutBits); t += 256) s_lut[t] = P.lut[t];
  __syncthreads();
  const uint32_t q = threadIdx.x;
  for (uint32_t ct = blockIdx.x; ct < P.num_cts; ct += gridDim.x) {
    int32_t *a = P.out + (size_t)ct * kCtWords;
    // The genotype bytes of this thread's four b words (setScore) are loaded
    // first, so that their latency runs under the Philox rounds below.
    int32_t g[4] = {0, 0, 0, 0};
    uint32_t val[4] = {0, 0, 0, 0};
    bool has[4] = {false, false, false, false};
    if (P.geno) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t j = 4 * q + u;
        const uint32_t rr = (uint32_t)(((uint64_t)j * P.rs_recip) >> 32), sj = j - rr * P.rs;
        const uint32_t bidx = ct * P.R + rr, tag = bidx / 3;
        has[u] = rr < P.R && sj < P.ns && bidx < P.num_cols;
        val[u] = bidx - 3 * tag;
        if (has[u]) g[u] = P.geno[(size_t)tag * P.ns + sj];
      }
    }
    uint32_t r[4];
    philox(q, ct, 0, 0, P.seed_lo, P.seed_hi, r);
    *(u32x4v *)(a + 4 * q) = u32x4v{r[0], r[1], r[2], r[3]};
    uint32_t b[4];
#pragma unroll
    for (int pr = 0; pr < 2; ++pr) {
      philox(2 * q + pr, ct, 1, 0, P.seed_lo, P.seed_hi, r);
      b[2 * pr] = (uint32_t)noise_sample_lut((uint64_t)r[0] | ((uint64_t)r[1] << 32), s_table, s_lut);
      b[2 * pr + 1] =
          (uint32_t)noise_sample_lut((uint64_t)r[2] | ((uint64_t)r[3] << 32), s_table, s_lut);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (!has[u]) continue;
      if (g[u] == 0 || g[u] == 1 || g[u] == 2) {
        if ((uint32_t)g[u] == val[u]) b[u] += kOne;
      } else {
        b[u] += val[u] == 0 ? kNan[0] : (val[u] == 1 ? kNan[1] : kNan[2]);
      }
    }
    *(u32x4v *)(a + kN + 4 * q) = u32x4v{b[0], b[1], b[2], b[3]};
    if (q == 0) P.out_var[ct] = 0x1p-50;
  }
}

// Largest decrypt of ihe_decrypt_noalloc: its item count (32-ciphertext groups
// x column groups of 1024 words) stays far inside the 32-bit counters.
constexpr uint32_t kNoallocMaxCts = 1u << 26;

// IHE_RING_TILES = 1 or 2 column tiles per wave (default 2); IHE_RING_QUEUE=0
// gives the workgroups fixed items instead of the counters
template <bool ENC>
void launch_ring_on(hipStream_t stream, DecryptParams P, const KeyBits &key, uint32_t blocks,
                    uint32_t *queue) {
  static_assert(kRingQueueWords >= 9, "eight XCD counters and a finished count");
  const char *q = std::getenv("IHE_RING_QUEUE");
  P.queue = (q && std::strcmp(q, "0") == 0) ? nullptr : queue;
  const char *e = std::getenv("IHE_RING_TILES");
  if (e && std::strcmp(e, "1") == 0)
    hipLaunchKernelGGL((k_ring_mfma<ENC, 1>), dim3(blocks), dim3(64 * 8), 0, stream, P, key);
  else
    hipLaunchKernelGGL((k_ring_mfma<ENC, 2>), dim3(blocks), dim3(64 * 4), 0, stream, P, key);
}

template <bool ENC>
void launch_ring(ihe_ctx *ctx, DecryptParams P, const KeyBits &key, uint32_t blocks) {
  launch_ring_on<ENC>(ctx->stream, P, key, blocks, ctx->d_ring_queue);
}

// Decrypt launch shape: items of 32 ciphertexts x 32 * kDecCols words, run by
// persistent workgroups, two per CU, a multiple of the 8 XCDs.
DecryptParams decrypt_params(const int32_t *d_in, uint32_t num_samples, uint32_t num_cts,
                             float *d_out, uint32_t num_cus, uint32_t &blocks) {
  const uint32_t col_groups = (num_samples + kDecCols - 1) / kDecCols;
  DecryptParams P{d_in, d_out, nullptr, num_cts, num_samples, col_groups,
                  ((num_cts + 31) / 32) * col_groups, nullptr};
  blocks = std::min<uint32_t>((num_cus * 2 + 7) / 8 * 8, (P.num_items + 7) / 8 * 8);
  return P;
}

__global__ void k_clear_words(uint32_t *p, uint32_t n) {
  if (threadIdx.x < n) p[threadIdx.x] = 0;
}

// Per device: the address of `scratch` and the CU count, looked up on the
// first ihe_decrypt_noalloc call on that device (with that device current).
constexpr int kMaxDevices = 64;
struct NoallocDevice {
  uint32_t *scratch = nullptr;
  uint32_t num_cus = 0;
};
std::mutex g_noalloc_mu;
NoallocDevice g_noalloc[kMaxDevices];

int noalloc_device(int dev, NoallocDevice &out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(IHE_ERR_UNSUPPORTED, "device index beyond 63");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  NoallocDevice &d = g_noalloc[dev];
  if (!d.scratch) {
    hipDeviceProp_t prop;
    IHE_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      return fail(IHE_ERR_UNSUPPORTED, std::string("kernels are built for gfx950, device is ") +
                                           prop.gcnArchName);
    void *p = nullptr;
    IHE_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(IHE_ERR_HIP, "hipGetSymbolAddress gave no address for the work counters");
    d.num_cus = (uint32_t)prop.multiProcessorCount;
    d.scratch = static_cast<uint32_t *>(p);
  }
  out = d;
  return IHE_OK;
}

int noalloc_setup(int device) {
  NoallocDevice d;
  return noalloc_device(device, d);
}

}  // namespace ihe

using namespace ihe;

extern "C" {

int ihe_key_create(ihe_ctx *ctx, const int32_t *key, ihe_key **out) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!key || !out) return fail(IHE_ERR_ARG, "null argument");
  *out = nullptr;
  ihe_key *k = new ihe_key();
  k->ctx = ctx;
  std::memset(&k->bits, 0, sizeof(k->bits));
  for (uint32_t i = 0; i < kN; ++i) {
    if (key[i] != 0 && key[i] != 1) {
      delete k;
      return fail(IHE_ERR_ARG, "TLWE key must be binary");
    }
    if (key[i]) k->bits.w[i / 32] |= 1u << (i % 32);
  }
  *out = k;
  return IHE_OK;
}

int ihe_key_destroy(ihe_key *key) {
  delete key;
  return IHE_OK;
}

int ihe_encrypt(ihe_ctx *ctx, const ihe_key *key, const ihe_layout *layout, uint32_t num_tags,
                const int8_t *d_geno, uint64_t seed, int32_t *d_out, double *d_out_var) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!key || key->ctx->device != ctx->device)
    return fail(IHE_ERR_ARG, "key is not on this context's device");
  if (!layout) return fail(IHE_ERR_ARG, "null layout");
  const uint32_t NS = layout->num_samples;
  if (NS == 0 || NS > kN || layout->num_regions != kN / NS ||
      layout->region_size != kN / layout->num_regions)
    return fail(IHE_ERR_ARG, "layout must satisfy R = N / NS and RS = N / R (eval/idash.cpp:409-410)");
  if (num_tags == 0) return IHE_OK;
  if (!d_out || !d_out_var) return fail(IHE_ERR_ARG, "null buffer");
  if (num_tags > 0x55555555u) return fail(IHE_ERR_ARG, "too many tags");
  if (int rc = make_current(ctx)) return rc;
  EncryptParams P;
  P.geno = d_geno;
  P.table = ctx->d_noise_table;
  P.lut = ctx->d_noise_lut;
  P.out = d_out;
  P.out_var = d_out_var;
  P.num_cols = 3 * num_tags;
  P.R = layout->num_regions;
  P.rs = layout->region_size;
  P.rs_recip = ((1ull << 32) + P.rs - 1) / P.rs;
  P.ns = NS;
  P.num_cts = (P.num_cols + P.R - 1) / P.R;
  P.seed_lo = (uint32_t)seed;
  P.seed_hi = (uint32_t)(seed >> 32);
  const char *force = std::getenv("IHE_ENCRYPT_KERNEL");
  if (!(force && std::strcmp(force, "valu") == 0)) {
    // a and b = e + messages, then b += a*s on the MFMA kernel
    const uint32_t col_groups = kN / kDecCols;
    DecryptParams D{d_out, nullptr, d_out, P.num_cts, NS, col_groups,
                    ((P.num_cts + 31) / 32) * col_groups, nullptr};
    const uint32_t blocks = std::min<uint32_t>(((uint32_t)ctx->num_cus * 2 + 7) / 8 * 8,
                                               (D.num_items + 7) / 8 * 8);
    ScopedTimer t(ctx, IHE_TIMER_ENCRYPT);
    static_assert(kN / 4 == 256, "a thread per four words of a ciphertext");
    hipLaunchKernelGGL(k_encrypt_ae, dim3(std::min<uint32_t>(P.num_cts, (uint32_t)ctx->num_cus * 8)),
                       dim3(256), 0, ctx->stream, P);
    launch_ring<true>(ctx, D, key->bits, blocks);
    IHE_HIP(hipGetLastError());
    return IHE_OK;
  }
  const dim3 grid((P.num_cts + kWavesPerBlock - 1) / kWavesPerBlock), block(64 * kWavesPerBlock);
  {
    ScopedTimer t(ctx, IHE_TIMER_ENCRYPT);
    hipLaunchKernelGGL(k_encrypt, grid, block, 0, ctx->stream, P, key->bits);
  }
  IHE_HIP(hipGetLastError());
  return IHE_OK;
}

int ihe_decrypt(ihe_ctx *ctx, const ihe_key *key, uint32_t num_samples, const int32_t *d_in,
                uint32_t num_cts, float *d_out) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!key || key->ctx->device != ctx->device)
    return fail(IHE_ERR_ARG, "key is not on this context's device");
  if (num_samples == 0 || num_samples > kN) return fail(IHE_ERR_ARG, "NUM_SAMPLES must be in [1, 1024]");
  if (num_cts == 0) return IHE_OK;
  if (!d_in || !d_out) return fail(IHE_ERR_ARG, "null buffer");
  if (int rc = make_current(ctx)) return rc;
  ScopedTimer t(ctx, IHE_TIMER_DECRYPT);
  const char *force = std::getenv("IHE_DECRYPT_KERNEL");
  if (!(force && std::strcmp(force, "valu") == 0)) {
    uint32_t blocks;
    const DecryptParams P =
        decrypt_params(d_in, num_samples, num_cts, d_out, (uint32_t)ctx->num_cus, blocks);
    launch_ring<false>(ctx, P, key->bits, blocks);
    IHE_HIP(hipGetLastError());
    return IHE_OK;
  }
  const dim3 grid((num_cts + kWavesPerBlock - 1) / kWavesPerBlock), block(64 * kWavesPerBlock);
  // J = outputs per lane, the smallest power of two with 64 * J >= NS.
  if (num_samples <= 64) hipLaunchKernelGGL(k_decrypt<1>, grid, block, 0, ctx->stream, d_in, num_cts, num_samples, key->bits, d_out);
  else if (num_samples <= 128) hipLaunchKernelGGL(k_decrypt<2>, grid, block, 0, ctx->stream, d_in, num_cts, num_samples, key->bits, d_out);
  else if (num_samples <= 256) hipLaunchKernelGGL(k_decrypt<4>, grid, block, 0, ctx->stream, d_in, num_cts, num_samples, key->bits, d_out);
  else if (num_samples <= 512) hipLaunchKernelGGL(k_decrypt<8>, grid, block, 0, ctx->stream, d_in, num_cts, num_samples, key->bits, d_out);
  else hipLaunchKernelGGL(k_decrypt<16>, grid, block, 0, ctx->stream, d_in, num_cts, num_samples, key->bits, d_out);
  IHE_HIP(hipGetLastError());
  return IHE_OK;
}

int ihe_decrypt_noalloc(const int32_t *key, uint32_t num_samples, const int32_t *d_in,
                        uint32_t num_cts, float *d_out, void *hip_stream) {
  if (!key) return fail(IHE_ERR_ARG, "null key");
  if (num_samples == 0 || num_samples > kN) return fail(IHE_ERR_ARG, "NUM_SAMPLES must be in [1, 1024]");
  if (num_cts > kNoallocMaxCts)
    return fail(IHE_ERR_ARG, "ihe_decrypt_noalloc takes at most 2^26 ciphertexts per call, got " +
                                 std::to_string(num_cts));
  KeyBits bits;
  std::memset(&bits, 0, sizeof(bits));
  for (uint32_t i = 0; i < kN; ++i) {
    if (key[i] != 0 && key[i] != 1) return fail(IHE_ERR_ARG, "TLWE key must be binary");
    if (key[i]) bits.w[i / 32] |= 1u << (i % 32);
  }
  if (num_cts == 0) return IHE_OK;
  if (!d_in || !d_out) return fail(IHE_ERR_ARG, "null buffer");
  int dev = 0;
  IHE_HIP(hipGetDevice(&dev));
  NoallocDevice D;
  if (int rc = noalloc_device(dev, D)) return rc;
  hipStream_t stream = static_cast<hipStream_t>(hip_stream);
  uint32_t blocks;
  const DecryptParams P = decrypt_params(d_in, num_samples, num_cts, d_out, D.num_cus, blocks);
  // the counters start at zero (a fill kernel: no memset node under capture)
  hipLaunchKernelGGL(k_clear_words, dim3(1), dim3(64), 0, stream, D.scratch, kRingQueueWords);
  launch_ring_on<false>(stream, P, bits, blocks, D.scratch);
  IHE_HIP(hipGetLastError());
  return IHE_OK;
}

}  // extern "C"
