// CPU baseline for the cloud step. TEST/BENCH INFRASTRUCTURE ONLY: bench.py
// times it as cpu_baseline (and fails when the GPU output differs from it);
// tests/test_oracle.py holds it bit-exact to orc_cloud_reference at R = 1, 8
// and 64.
//
// It keeps the data structures and threading of cloud_compute_score
// (eval/idash.cpp:763-848): the model is an
// unordered_map<FeatBigIndex, unordered_map<FeatBigIndex, int32_t>>
// (eval/idash.h:129-134), the ciphertexts an unordered_map<FeatIndex, sample*>
// looked up twice per coefficient (getTLWE, eval/idash.h:162-166), outputs are
// allocated one by one into an unordered_map before the parallel loop
// (createAndGet, eval/idash.h:174-183), and an OpenMP team with per-thread
// region temporaries runs `omp for` over the outputs (:779-847).
#include <omp.h>
#include <sys/time.h>

#include <cstring>
#include <unordered_map>
#include <utility>
#include <vector>

#include "oracle.h"

namespace {

constexpr int kN = 1024;

struct Sample {
  int32_t *coefs;  // a[0..N) then b[0..N)
  double var;
};

double wall() {
  timeval tv;
  gettimeofday(&tv, nullptr);
  return (double)tv.tv_sec + (double)tv.tv_usec * 1e-6;
}

inline void mul_by_xai(int32_t *out, int32_t a, const int32_t *in) {
  // torusPolynomialMulByXai for a in [N, 2N): the cloud step rotates region
  // r in [1, R) by a = 2N - r*RS (eval/idash.cpp:832), and r*RS < R*RS <= N
  // puts every such a in (N, 2N).
  const int32_t aa = a - kN;
  for (int32_t i = 0; i < aa; ++i) out[i] = in[i - aa + kN];
  for (int32_t i = aa; i < kN; ++i) out[i] = (int32_t)(0u - (uint32_t)in[i - aa]);
}

}  // namespace

extern "C" double orc_cpu_baseline_cloud(int num_threads, uint32_t NS, uint32_t R, uint32_t RS,
                                         const int32_t *ct, const double *var_in, uint32_t num_in,
                                         uint32_t num_out, const uint32_t *row_ptr,
                                         const uint32_t *in_bidx, const int32_t *weight,
                                         int32_t *out, double *var_out) {
  int32_t C4[4];
  orc_encoding_constants(C4);
  const uint32_t ONE = (uint32_t)C4[0];

  // Inputs, as read_encrypted_data leaves them (eval/idash.cpp:513-535).
  std::vector<Sample> in_samples(num_in);
  std::unordered_map<uint32_t, Sample *> enc_data;
  for (uint32_t i = 0; i < num_in; ++i) {
    in_samples[i].coefs = const_cast<int32_t *>(ct + (size_t)i * 2 * kN);
    in_samples[i].var = var_in[i];
    enc_data.emplace(i, &in_samples[i]);
  }
  // Model, as read_model leaves it (eval/idash.cpp:66-114).
  std::unordered_map<uint32_t, std::unordered_map<uint32_t, int32_t>> model;
  for (uint32_t o = 0; o < num_out; ++o) {
    auto &m = model[o];
    for (uint32_t e = row_ptr[o]; e < row_ptr[o + 1]; ++e) m[in_bidx[e]] = weight[e];
  }
  std::unordered_map<uint32_t, Sample *> preds;

  const double t0 = wall();
  // ---- cloud_compute_score ----
  std::vector<std::pair<uint32_t, std::unordered_map<uint32_t, int32_t>>> model_vec(model.begin(),
                                                                                   model.end());
  for (const auto &it : model_vec) {
    Sample *s = new Sample;
    s->coefs = new int32_t[2 * kN];
    s->var = 0.0;
    preds.emplace(it.first, s);
  }
#pragma omp parallel num_threads(num_threads)
  {
    std::vector<int32_t> tmp_store((size_t)R * 2 * kN);
    std::vector<double> tmp_var(R);
    std::vector<int32_t> rot(2 * kN);
#pragma omp for
    for (unsigned i = 0; i < model_vec.size(); ++i) {
      const uint32_t out_bidx = model_vec[i].first;
      const auto &coeffs = model_vec[i].second;
      std::memset(tmp_store.data(), 0, sizeof(int32_t) * tmp_store.size());
      for (uint32_t r = 0; r < R; ++r) tmp_var[r] = 0.0;
      for (const auto &it2 : coeffs) {
        const uint32_t bidx = it2.first;
        const int32_t coeff = it2.second;
        if (bidx == ORC_CONSTANT_BIDX) {
          const uint32_t scaled = (uint32_t)coeff * ONE;
          uint32_t *b = reinterpret_cast<uint32_t *>(tmp_store.data()) + kN;
          for (uint32_t j = 0; j < NS; ++j) b[j] += scaled;
        } else {
          const uint32_t region = bidx % R, idx = bidx / R;
          if (enc_data.count(idx) == 0) continue;  // the reference aborts here
          const Sample *src = enc_data.at(idx);
          uint32_t *dst = reinterpret_cast<uint32_t *>(tmp_store.data()) + (size_t)region * 2 * kN;
          const uint32_t *s = reinterpret_cast<const uint32_t *>(src->coefs);
          for (int k = 0; k < 2 * kN; ++k) dst[k] += (uint32_t)coeff * s[k];
          tmp_var[region] += (double)(int32_t)((uint32_t)coeff * (uint32_t)coeff) * src->var;
        }
      }
      Sample *o = preds.at(out_bidx);
      std::memcpy(o->coefs, tmp_store.data(), sizeof(int32_t) * 2 * kN);
      o->var = tmp_var[0];
      for (uint32_t r = 1; r < R; ++r) {
        const int32_t amount = 2 * kN - (int32_t)(r * RS);
        for (int p = 0; p < 2; ++p)
          mul_by_xai(rot.data() + p * kN, amount, tmp_store.data() + (size_t)r * 2 * kN + p * kN);
        uint32_t *d = reinterpret_cast<uint32_t *>(o->coefs);
        for (int k = 0; k < 2 * kN; ++k) d[k] += (uint32_t)rot[k];
      }
      for (int j = (int)RS; j < kN; ++j) o->coefs[kN + j] = 0;
    }
  }
  const double t1 = wall();

  for (auto &it : preds) {
    std::memcpy(out + (size_t)it.first * 2 * kN, it.second->coefs, sizeof(int32_t) * 2 * kN);
    var_out[it.first] = it.second->var;
    delete[] it.second->coefs;
    delete it.second;
  }
  return t1 - t0;
}
