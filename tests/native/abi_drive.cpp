// abi_drive: the batched, served and tuned forms of the cloud step driven from
// C++ through the C ABI alone (include/idash_hip.h), each checked bit for bit
// against the single-cohort entry points it must equal:
//   batch   ihe_batch_run, eager and as a captured graph replayed twice, against
//           ihe_cloud_score_layout per cohort (one cloud_compute_score call per
//           cohort, eval/cloud.cpp:17)
//   server  ihe_server_* requests, one above 1024 samples (two chunks), against
//           ihe_encrypt (seed + chunk) -> ihe_cloud_score_layout -> ihe_decrypt
//           (encrypt_data_ph1/ph2, cloud_compute_score, decrypt_predictions,
//           eval/idash.cpp:615-761)
//   tuner   the band from ihe_band_create_tuned against ihe_band_create's
// Prints one line per check and exits 0 when all hold. Built plain and with the
// host-side sanitizers (make asan), so the host code of those forms also runs
// under AddressSanitizer and UBSan.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "idash_hip.h"

namespace {

int g_failures = 0;

void check(int status, const char *what) {
  if (status != IHE_OK) {
    std::printf("FAIL %s: status %d (%s)\n", what, status, ihe_last_error());
    std::fflush(stdout);
    std::exit(1);
  }
}

void expect(bool ok, const char *what) {
  std::printf("%s %s\n", ok ? "ok" : "FAIL", what);
  std::fflush(stdout);
  if (!ok) ++g_failures;
}

struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  int32_t range(int32_t lo, int32_t hi) { return lo + int32_t(next() % uint64_t(hi - lo + 1)); }
};

ihe_layout layout_of(uint32_t ns) {
  uint32_t r = IHE_N / ns;
  return {ns, r, IHE_N / r};
}

uint32_t num_in_cts(uint32_t tags, uint32_t ns) {
  uint32_t r = IHE_N / ns;
  return (3 * tags + r - 1) / r;
}

// A banded model like read_model's (eval/idash.cpp:66-114): target t reads the
// three one-hot columns of n consecutive tags starting at lo(t), non-decreasing
// in t, plus the constant; one row per output big index 3t+k.
struct Model {
  uint32_t tags, targets;
  std::vector<uint32_t> row_ptr, in_bidx;
  std::vector<int32_t> weight;
  ihe_model_csr csr() const {
    return {3 * targets, row_ptr.data(), in_bidx.data(), weight.data()};
  }
};

Model make_model(uint32_t tags, uint32_t targets, uint32_t n, uint64_t seed) {
  Model m{tags, targets, {0}, {}, {}};
  SplitMix rng{seed};
  for (uint32_t t = 0; t < targets; ++t) {
    uint32_t lo = uint32_t(uint64_t(t) * (tags - n) / (targets > 1 ? targets - 1 : 1));
    for (uint32_t k = 0; k < 3; ++k) {
      for (uint32_t c = 3 * lo; c < 3 * (lo + n); ++c) {
        m.in_bidx.push_back(c);
        m.weight.push_back(rng.range(-2048, 2047));
      }
      m.in_bidx.push_back(IHE_CONSTANT_BIDX);
      m.weight.push_back(rng.range(-2048, 2047));
      m.row_ptr.push_back(uint32_t(m.in_bidx.size()));
    }
  }
  return m;
}

// Device buffer of `count` elements of T, freed with the context's allocator.
template <class T>
struct Dev {
  ihe_ctx *ctx = nullptr;
  T *ptr = nullptr;
  size_t count = 0;
  Dev(ihe_ctx *c, size_t n) : ctx(c), count(n) {
    void *p = nullptr;
    check(ihe_malloc(ctx, n * sizeof(T), &p), "ihe_malloc");
    ptr = static_cast<T *>(p);
  }
  Dev(const Dev &) = delete;
  Dev(Dev &&o) noexcept : ctx(o.ctx), ptr(o.ptr), count(o.count) { o.ptr = nullptr; }
  ~Dev() {
    if (ptr) ihe_free(ctx, ptr);
  }
  void upload(const std::vector<T> &h) {
    check(ihe_memcpy_h2d(ctx, ptr, h.data(), count * sizeof(T)), "h2d");
  }
  std::vector<T> download() const {
    std::vector<T> h(count);
    check(ihe_memcpy_d2h(ctx, h.data(), ptr, count * sizeof(T)), "d2h");
    return h;
  }
};

// Random ciphertexts (every word) and the fresh-sample variance 2^-50.
void fill_inputs(Dev<int32_t> &ct, Dev<double> &var, uint64_t seed) {
  SplitMix rng{seed};
  std::vector<int32_t> h(ct.count);
  for (auto &w : h) w = int32_t(uint32_t(rng.next()));
  ct.upload(h);
  var.upload(std::vector<double>(var.count, 0x1p-50));
}

void drive_batch(ihe_ctx *ctx, const ihe_band *bands[2], const Model *models[2]) {
  const uint32_t sample_counts[] = {16, 100, 128, 333, 1024, 64, 7};
  const uint32_t num = sizeof(sample_counts) / sizeof(sample_counts[0]);
  std::vector<Dev<int32_t>> in, out, ref;
  std::vector<Dev<double>> in_var, out_var, ref_var;
  std::vector<ihe_cohort> cohorts;
  for (uint32_t i = 0; i < num; ++i) {
    uint32_t ns = sample_counts[i], b = i % 2;
    uint32_t cin = num_in_cts(models[b]->tags, ns), cout = 3 * models[b]->targets;
    in.emplace_back(ctx, size_t(cin) * 2 * IHE_N);
    in_var.emplace_back(ctx, cin);
    out.emplace_back(ctx, size_t(cout) * 2 * IHE_N);
    out_var.emplace_back(ctx, cout);
    ref.emplace_back(ctx, size_t(cout) * 2 * IHE_N);
    ref_var.emplace_back(ctx, cout);
    fill_inputs(in.back(), in_var.back(), 0xB47C0000u + i);
  }
  for (uint32_t i = 0; i < num; ++i) {
    uint32_t ns = sample_counts[i], b = i % 2;
    ihe_layout lay = layout_of(ns);
    uint32_t cin = num_in_cts(models[b]->tags, ns);
    check(ihe_cloud_score_layout(ctx, bands[b], &lay, cin, in[i].ptr, in_var[i].ptr, ref[i].ptr,
                                 ref_var[i].ptr),
          "ihe_cloud_score_layout");
    // the last cohort asks for no variances
    double *ov = i + 1 < num ? out_var[i].ptr : nullptr;
    cohorts.push_back({bands[b], lay, cin, in[i].ptr, ov ? in_var[i].ptr : nullptr, out[i].ptr, ov});
  }
  ihe_batch *batch = nullptr;
  check(ihe_batch_create(ctx, cohorts.data(), num, &batch), "ihe_batch_create");
  auto same = [&](const char *what) {
    check(ihe_sync(ctx), "ihe_sync");
    bool ok = true;
    for (uint32_t i = 0; i < num; ++i) {
      ok = ok && out[i].download() == ref[i].download();
      if (cohorts[i].d_out_var) ok = ok && out_var[i].download() == ref_var[i].download();
    }
    expect(ok, what);
  };
  auto clear = [&] {
    for (uint32_t i = 0; i < num; ++i) {
      check(ihe_memset(ctx, out[i].ptr, 0x5A, out[i].count * sizeof(int32_t)), "memset");
      check(ihe_memset(ctx, out_var[i].ptr, 0x5A, out_var[i].count * sizeof(double)), "memset");
    }
    check(ihe_sync(ctx), "ihe_sync");
  };
  clear();
  check(ihe_batch_run(batch), "ihe_batch_run");
  same("batch (eager) equals ihe_cloud_score_layout per cohort");
  check(ihe_batch_capture(batch), "ihe_batch_capture");
  for (int rep = 0; rep < 2; ++rep) {
    clear();
    check(ihe_batch_run(batch), "ihe_batch_run (graph)");
    same(rep ? "batch (graph replay 2) equals the per-cohort results"
             : "batch (graph replay 1) equals the per-cohort results");
  }
  ihe_batch_info info{};
  check(ihe_batch_get_info(batch, &info), "ihe_batch_get_info");
  expect(info.num_cohorts == num && info.captured, "batch info: cohort count, captured");
  check(ihe_batch_destroy(batch), "ihe_batch_destroy");
}

// The request's chunk-major host genotypes: chunk k is [tags][ns_k].
std::vector<int8_t> chunk_major_geno(uint32_t tags, uint32_t ns, uint64_t seed) {
  SplitMix rng{seed};
  std::vector<int8_t> g(size_t(tags) * ns);
  for (auto &v : g) v = int8_t(rng.range(-1, 2));
  return g;
}

void drive_server(ihe_ctx *ctx, const ihe_band *band, const Model &m) {
  ihe_server_config cfg{uint64_t(2) << 30, 2, 0};
  ihe_server *srv = nullptr;
  check(ihe_server_create(ctx, &cfg, &srv), "ihe_server_create");
  uint32_t model_id = 0;
  check(ihe_server_add_model(srv, band, m.tags, &model_id), "ihe_server_add_model");
  const uint32_t sample_counts[] = {50, 1500, 1024, 7, 300};
  const uint32_t num = sizeof(sample_counts) / sizeof(sample_counts[0]);
  const uint32_t outs = 3 * m.targets;
  std::vector<std::vector<int8_t>> geno(num);
  std::vector<std::vector<int32_t>> key(num);
  std::vector<std::vector<float>> scores(num);
  std::vector<uint64_t> ticket(num);
  for (uint32_t i = 0; i < num; ++i) {
    geno[i] = chunk_major_geno(m.tags, sample_counts[i], 0x5E4F0000u + i);
    SplitMix rng{0x4E700000u + i};
    key[i].resize(IHE_N);
    for (auto &s : key[i]) s = int32_t(rng.next() & 1);
    scores[i].assign(size_t(outs) * sample_counts[i], -1.0f);
    ihe_request req{model_id, sample_counts[i], geno[i].data(), key[i].data(), 1000u + 17u * i,
                    scores[i].data()};
    check(ihe_server_submit(srv, &req, &ticket[i]), "ihe_server_submit");
  }
  for (uint32_t i = 0; i < num; ++i) {
    double ms = 0;
    check(ihe_server_wait(srv, ticket[i], &ms), "ihe_server_wait");
  }
  ihe_server_stats st{};
  check(ihe_server_get_stats(srv, &st), "ihe_server_get_stats");
  expect(st.completed == num && st.pending == 0 && st.running == 0 && st.arena_used == 0,
         "server stats: all requests completed, arena empty");
  check(ihe_server_destroy(srv), "ihe_server_destroy");

  // The per-request path, chunk by chunk, on the context stream.
  bool ok = true;
  for (uint32_t i = 0; i < num; ++i) {
    ihe_key *k = nullptr;
    check(ihe_key_create(ctx, key[i].data(), &k), "ihe_key_create");
    size_t geno_off = 0, score_off = 0;
    for (uint32_t col0 = 0, chunk = 0; col0 < sample_counts[i]; col0 += IHE_N, ++chunk) {
      uint32_t ns = sample_counts[i] - col0 < IHE_N ? sample_counts[i] - col0 : IHE_N;
      ihe_layout lay = layout_of(ns);
      uint32_t cin = num_in_cts(m.tags, ns);
      Dev<int8_t> d_geno(ctx, size_t(m.tags) * ns);
      d_geno.upload(std::vector<int8_t>(geno[i].begin() + geno_off,
                                        geno[i].begin() + geno_off + d_geno.count));
      Dev<int32_t> d_ct(ctx, size_t(cin) * 2 * IHE_N), d_out(ctx, size_t(outs) * 2 * IHE_N);
      Dev<double> d_var(ctx, cin);
      Dev<float> d_dec(ctx, size_t(outs) * ns);
      check(ihe_encrypt(ctx, k, &lay, m.tags, d_geno.ptr, 1000u + 17u * i + chunk, d_ct.ptr,
                        d_var.ptr),
            "ihe_encrypt");
      check(ihe_cloud_score_layout(ctx, band, &lay, cin, d_ct.ptr, nullptr, d_out.ptr, nullptr),
            "ihe_cloud_score_layout");
      check(ihe_decrypt(ctx, k, ns, d_out.ptr, outs, d_dec.ptr), "ihe_decrypt");
      std::vector<float> want = d_dec.download();
      ok = ok && std::memcmp(want.data(), scores[i].data() + score_off,
                             want.size() * sizeof(float)) == 0;
      geno_off += d_geno.count;
      score_off += want.size();
    }
    check(ihe_key_destroy(k), "ihe_key_destroy");
  }
  expect(ok, "server results equal encrypt -> cloud -> decrypt per chunk");
}

void drive_tuner(ihe_ctx *ctx, const ihe_band *band, const Model &m) {
  ihe_layout lay = layout_of(IHE_N);
  uint32_t cin = num_in_cts(m.tags, IHE_N), outs = 3 * m.targets;
  Dev<int32_t> in(ctx, size_t(cin) * 2 * IHE_N), out(ctx, size_t(outs) * 2 * IHE_N),
      ref(ctx, size_t(outs) * 2 * IHE_N);
  Dev<double> in_var(ctx, cin), out_var(ctx, outs), ref_var(ctx, outs);
  fill_inputs(in, in_var, 0x7E4E);
  ihe_model_csr csr = m.csr();
  ihe_band *tuned = nullptr;
  ihe_tune_entry results[64];
  uint32_t num_results = 64;
  check(ihe_band_create_tuned(ctx, &lay, cin, &csr, in.ptr, out.ptr, 2, &tuned, results,
                              &num_results),
        "ihe_band_create_tuned");
  expect(num_results >= 1 && num_results <= 64, "tuner timed at least one tile shape");
  check(ihe_cloud_score(ctx, band, in.ptr, in_var.ptr, ref.ptr, ref_var.ptr), "ihe_cloud_score");
  check(ihe_cloud_score(ctx, tuned, in.ptr, in_var.ptr, out.ptr, out_var.ptr),
        "ihe_cloud_score (tuned)");
  check(ihe_sync(ctx), "ihe_sync");
  expect(out.download() == ref.download() && out_var.download() == ref_var.download(),
         "tuned band equals the default band");
  check(ihe_band_destroy(tuned), "ihe_band_destroy");
}

}  // namespace

int main() {
  int devices = 0;
  check(ihe_device_count(&devices), "ihe_device_count");
  if (devices == 0) {
    std::printf("FAIL no GPU visible\n");
    return 1;
  }
  ihe_ctx *ctx = nullptr;
  check(ihe_ctx_create(0, &ctx), "ihe_ctx_create");
  const Model m5 = make_model(60, 150, 5, 5), m20 = make_model(60, 90, 20, 20);
  const Model *models[2] = {&m5, &m20};
  const ihe_band *bands[2];
  std::vector<ihe_band *> owned;
  for (int b = 0; b < 2; ++b) {
    ihe_layout lay = layout_of(IHE_N);
    ihe_model_csr csr = models[b]->csr();
    ihe_band *band = nullptr;
    check(ihe_band_create(ctx, &lay, num_in_cts(models[b]->tags, IHE_N), &csr, IHE_VARIANT_MFMA,
                          &band),
          "ihe_band_create");
    owned.push_back(band);
    bands[b] = band;
  }
  drive_batch(ctx, bands, models);
  drive_server(ctx, bands[0], m5);
  drive_tuner(ctx, bands[1], m20);
  for (ihe_band *b : owned) check(ihe_band_destroy(b), "ihe_band_destroy");
  check(ihe_ctx_destroy(ctx), "ihe_ctx_destroy");
  std::printf("%s\n", g_failures ? "FAILED" : "all checks passed");
  return g_failures ? 1 : 0;
}
