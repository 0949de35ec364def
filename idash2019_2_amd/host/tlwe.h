// TLWE sample and key types of the idash-compatible host (k = 1, N = 1024).
//
// The reference takes these from TFHE (TLweSample, TLweKey, TLweParams,
// Torus32, dtot32/t32tod; used throughout eval/idash.h and eval/idash.cpp).
// Here they are the host view of the dense layout the GPU kernels use: a
// sample is 2N consecutive int32 words (mask a, then body b) inside a
// SampleBlock, with its variance beside it. Only what the idash entry points
// rely on is provided.
#ifndef IHE_HOST_TLWE_H
#define IHE_HOST_TLWE_H

#include <cstddef>
#include <cstdint>
#include <memory>

typedef int32_t Torus32;

// double in (-1, 1) -> Torus32: int32(int64((d - trunc(d)) * 2^32)) (TFHE dtot32)
Torus32 dtot32(double d);
// Torus32 -> double: x / 2^32 (TFHE t32tod)
double t32tod(Torus32 x);

struct TLweParams {
  int32_t N;         // polynomial size
  int32_t k;         // number of mask polynomials (always 1 here)
  double alpha_min;  // noise standard deviation of a fresh sample
  double alpha_max;
};

struct TLweKey {
  const TLweParams *params;
  int32_t key[1024];  // binary secret polynomial s
};

struct TLweSample {
  Torus32 *a;               // mask polynomial a[0..N)
  Torus32 *b;               // body polynomial b[0..N)
  double *current_variance; // variance slot of the sample
};

// C samples stored densely: words [C][2][N] and variances [C], in pinned host
// memory when a device is present (plain memory otherwise), plus one
// TLweSample view per slot.
class SampleBlock {
 public:
  explicit SampleBlock(size_t count);
  ~SampleBlock();
  SampleBlock(const SampleBlock &) = delete;
  SampleBlock &operator=(const SampleBlock &) = delete;

  size_t size() const { return count_; }
  TLweSample *sample(size_t i) { return &views_[i]; }
  int32_t *words() { return words_; }
  const int32_t *words() const { return words_; }
  double *variances() { return var_; }
  const double *variances() const { return var_; }

  // The 4-byte word that opens each sample's stream in a ciphertext file
  // (written by TFHE's export_tlweSample_toStream before a, b and the
  // variance). Files read keep the word their records carried, and samples
  // computed from them (the cloud step's outputs) are written with it; blocks
  // this build encrypts carry 0.
  int32_t stream_word = 0;

 private:
  size_t count_;
  int32_t *words_ = nullptr;
  double *var_ = nullptr;
  bool pinned_ = false;
  std::unique_ptr<TLweSample[]> views_;
};

#endif  // IHE_HOST_TLWE_H
