// Record files through the idash-compatible host library, on the CPU: reads a
// ciphertext file with read_encrypted_predictions (or read_encrypted_data) and
// writes the samples back with write_encrypted_predictions (or
// write_encrypted_data), printing the framing word and the records' indices in
// the order written; or writes a result file with write_decrypted_predictions
// from float32 scores [positions][3][samples] read from a file (position i at
// 1000 + 7 i, named "rs<i>"). Usage:
//   records_drive pred <in> <out> <NUM_OUTPUT_FEATURES>
//   records_drive data <in> <out> <NUM_INPUT_POSITIONS>
//   records_drive csv <scores> <out> <positions> <samples> <names 0|1>
//   records_drive gfmt <stride>: the result file's %g formatter against
//     snprintf on every stride-th float bit pattern, every multiple of 5 from
//     10^6 to 2^24 (exact ties at six digits) and powers of ten +- 1 ulp; prints
//     the values checked and the mismatches.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "gfmt.h"
#include "idash.h"

static int csv(int argc, char **argv) {
  if (argc != 7) return 2;
  const uint32_t P = (uint32_t)std::strtoul(argv[4], nullptr, 0);
  const uint32_t ns = (uint32_t)std::strtoul(argv[5], nullptr, 0);
  std::vector<float> v((size_t)P * 3 * ns);
  std::ifstream in(argv[2], std::ios::binary);
  in.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(v.size() * sizeof(float)));
  if (!in) return 3;
  IdashParams params{};
  params.NUM_SAMPLES = ns;
  DecryptedPredictions pred;
  for (uint32_t i = 0; i < P; ++i) {
    const uint64_t pos = 1000 + 7 * (uint64_t)i;
    params.registerOutPositionName(pos, "rs" + std::to_string(i));
    for (int k = 0; k < 3; ++k)
      pred.score[pos][k].assign(v.begin() + ((size_t)i * 3 + k) * ns,
                                v.begin() + ((size_t)i * 3 + k + 1) * ns);
  }
  write_decrypted_predictions(pred, params, argv[3], argv[6][0] == '1');
  return 0;
}

static int gfmt(uint32_t stride) {
  uint64_t checked = 0, bad = 0;
  auto check = [&](float v) {
    if (!std::isfinite(v)) return;
    char a[64], b[64];
    const int nb = std::snprintf(b, sizeof(b), "%g", (double)v);
    const char *e = ihe_host::format_g6(a, v);
    const std::string got(static_cast<const char *>(a), e);
    ++checked;
    if (got != std::string(b, (size_t)nb) && bad++ < 10)
      std::printf("mismatch %a: %s vs %s\n", (double)v, got.c_str(), b);
  };
  for (uint64_t u = 0; u < (1ull << 32); u += stride) {
    float v;
    const uint32_t w = (uint32_t)u;
    std::memcpy(&v, &w, 4);
    check(v);
  }
  for (uint32_t i = 1000000; i <= (1u << 24); i += 5) check((float)i), check(-(float)i);
  for (int x = -45; x <= 38; ++x) {
    const float p = (float)std::pow(10.0, x);
    check(p), check(std::nextafter(p, 0.0f)), check(std::nextafter(p, 1e38f));
  }
  std::printf("checked %llu mismatches %llu\n", (unsigned long long)checked, (unsigned long long)bad);
  return bad ? 1 : 0;
}

int main(int argc, char **argv) {
  if (argc >= 2 && std::string(argv[1]) == "csv") return csv(argc, argv);
  if (argc == 3 && std::string(argv[1]) == "gfmt") return gfmt((uint32_t)std::strtoul(argv[2], nullptr, 0));
  if (argc != 5) {
    std::fprintf(stderr, "usage: records_drive pred|data <in> <out> <count>\n");
    return 2;
  }
  const std::string mode = argv[1];
  IdashParams params{};
  params.NUM_SAMPLES = IdashParams::N;
  params.NUM_REGIONS = 1;
  params.REGION_SIZE = IdashParams::N;
  const uint32_t n = (uint32_t)std::strtoul(argv[4], nullptr, 0);
  if (mode == "pred") {
    params.NUM_OUTPUT_FEATURES = n;
    EncryptedPredictions p;
    read_encrypted_predictions(p, params, argv[2]);
    write_encrypted_predictions(p, params, argv[3]);
    std::printf("word %d\n", p.block ? p.block->stream_word : 0);
    for (const auto &e : p.score) std::printf("%u\n", e.first);
  } else {
    params.NUM_INPUT_POSITIONS = n;
    EncryptedData d;
    read_encrypted_data(d, params, argv[2]);
    write_encrypted_data(d, params, argv[3]);
    std::printf("word %d\n", d.block ? d.block->stream_word : 0);
    for (const auto &e : d.enc_data) std::printf("%u\n", e.first);
  }
  return 0;
}
