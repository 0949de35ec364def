// idash-compatible host functions (eval/idash.cpp) over the gfx950 engine.
//
// The file formats, index assignment and error behaviour follow the
// reference function cited at each definition. The ciphertext arithmetic
// (encrypt, cloud_compute_score, decrypt) runs on the GPU through
// include/idash_hip.h; a missing GPU is an error, never a CPU fallback.
#include "idash.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/resource.h>
#include <sys/stat.h>
#include <unistd.h>
#include <sys/time.h>

#include <algorithm>
#include <atomic>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <memory>
#include <future>
#include <random>
#include <sstream>
#include <thread>

#include "gfmt.h"
#include "idash_hip.h"
#include "parse_vw.h"

using std::string;

// ---------------------------------------------------------------------------
// constants (eval/idash.cpp:20-45)

Torus32 dtot32(double d) {
  const double frac = d - (double)(int64_t)d;
  return (Torus32)(int64_t)(frac * 4294967296.0);
}
double t32tod(Torus32 x) { return (double)x / 4294967296.0; }

namespace {
const TLweParams kTlweParams = {1024, 1, 0x1p-25, 0.25};
}

const double IdashParams::alpha = 0x1p-25;
const TLweParams *IdashParams::tlweParams = &kTlweParams;
const double IdashParams::IN_SCALING_FACTOR = 1.0 / 16384;
const double IdashParams::ONE_IN_D = IdashParams::IN_SCALING_FACTOR;
const double IdashParams::NAN_0_IN_D = 3. / 6. * IdashParams::IN_SCALING_FACTOR;
const double IdashParams::NAN_1_IN_D = 2. / 6. * IdashParams::IN_SCALING_FACTOR;
const double IdashParams::NAN_2_IN_D = 1. / 6. * IdashParams::IN_SCALING_FACTOR;
const Torus32 IdashParams::ONE_IN_T32 = dtot32(IdashParams::ONE_IN_D);
const Torus32 IdashParams::NAN_0_IN_T32 = dtot32(IdashParams::NAN_0_IN_D);
const Torus32 IdashParams::NAN_1_IN_T32 = dtot32(IdashParams::NAN_1_IN_D);
const Torus32 IdashParams::NAN_2_IN_T32 = dtot32(IdashParams::NAN_2_IN_D);

IdashKey::IdashKey(const IdashParams *p, const TLweKey *k) : idashParams(p), tlweKey(k) {}

// ---------------------------------------------------------------------------
// engine access

namespace {

constexpr size_t kCtWords = 2 * IdashParams::N;
constexpr size_t kCtBytes = kCtWords * sizeof(int32_t);
// One sample in a file: 4-byte framing word, a, b, variance (8204 bytes).
constexpr size_t kSampleStreamBytes = 4 + kCtBytes + sizeof(double);

[[noreturn]] void die_engine(const char *what) {
  std::cout << "ERROR: " << what << ": " << ihe_last_error() << std::endl;
  abort();
}

void check(int rc, const char *what) {
  if (rc != IHE_OK) die_engine(what);
}

// Process-wide context on device IHE_DEVICE (default 0), created on first use.
ihe_ctx *engine() {
  static ihe_ctx *ctx = nullptr;
  if (!ctx) {
    const char *dev = std::getenv("IHE_DEVICE");
    check(ihe_ctx_create(dev ? std::atoi(dev) : 0, &ctx), "no usable gfx950 device");
  }
  return ctx;
}

// The context if a device is present, without failing when there is none.
ihe_ctx *engine_if_present() {
  static int present = -1;
  if (present < 0) {
    int n = 0;
    present = ihe_device_count(&n) == IHE_OK && n > 0 ? 1 : 0;
  }
  return present ? engine() : nullptr;
}

// Device buffer released at scope exit.
struct DeviceBuffer {
  void *p = nullptr;
  explicit DeviceBuffer(size_t bytes) {
    if (bytes) check(ihe_malloc(engine(), bytes, &p), "device allocation");
  }
  ~DeviceBuffer() {
    if (p) ihe_free(engine(), p);
  }
  template <typename T>
  T *as() const { return static_cast<T *>(p); }
};

struct DeviceKey {
  ihe_key *k = nullptr;
  explicit DeviceKey(const TLweKey *key) { check(ihe_key_create(engine(), key->key, &k), "key upload"); }
  ~DeviceKey() { ihe_key_destroy(k); }
};

ihe_layout layout_of(const IdashParams &p) { return {p.NUM_SAMPLES, p.NUM_REGIONS, p.REGION_SIZE}; }

// Wall-clock phases of a host function, printed to stderr when IHE_HOST_PHASES
// is set (the executables' stdout printouts stay as the reference's).
struct Phases {
  const char *fn;
  bool on = std::getenv("IHE_HOST_PHASES") != nullptr;
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  explicit Phases(const char *f) : fn(f) {}
  void mark(const char *what) {
    if (!on) return;
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "%s phase %-28s %9.3f ms\n", fn, what,
                 std::chrono::duration<double, std::milli>(now - t).count());
    t = now;
  }
  // a figure summed over a loop (not a span since the last mark)
  void note(const char *what, double ms) const {
    if (on) std::fprintf(stderr, "%s phase %-28s %9.3f ms (summed)\n", fn, what, ms);
  }
};

uint64_t fresh_seed() {
  if (const char *s = std::getenv("IHE_SEED")) {
    static uint64_t next = std::strtoull(s, nullptr, 0);
    return next++;
  }
  std::random_device rd;
  return ((uint64_t)rd() << 32) ^ rd();
}

}  // namespace

// ---------------------------------------------------------------------------
// sample storage

SampleBlock::SampleBlock(size_t count) : count_(count), views_(new TLweSample[count ? count : 1]) {
  const size_t wbytes = std::max<size_t>(count * kCtBytes, 64);
  const size_t vbytes = std::max<size_t>(count * sizeof(double), 64);
  if (ihe_ctx *ctx = engine_if_present()) {
    void *w = nullptr, *v = nullptr;
    check(ihe_host_alloc(ctx, wbytes, &w), "pinned host allocation");
    check(ihe_host_alloc(ctx, vbytes, &v), "pinned host allocation");
    words_ = static_cast<int32_t *>(w);
    var_ = static_cast<double *>(v);
    pinned_ = true;
  } else {
    words_ = static_cast<int32_t *>(std::aligned_alloc(4096, (wbytes + 4095) & ~size_t(4095)));
    var_ = static_cast<double *>(std::aligned_alloc(64, (vbytes + 63) & ~size_t(63)));
    REQUIRE_DRAMATICALLY(words_ && var_, "host allocation failed");
  }
  for (size_t i = 0; i < count; ++i) {
    views_[i].a = words_ + i * kCtWords;
    views_[i].b = words_ + i * kCtWords + IdashParams::N;
    views_[i].current_variance = var_ + i;
  }
}

SampleBlock::~SampleBlock() {
  if (pinned_) {
    ihe_host_free(engine(), words_);
    ihe_host_free(engine(), var_);
  } else {
    std::free(words_);
    std::free(var_);
  }
}

namespace {

// Dense device-ready copy of `map` (slot i = key i, i < count): the block the
// map already points into when it is laid out that way, else a gathered copy.
template <typename Map>
std::shared_ptr<SampleBlock> dense_view(const Map &map, const std::shared_ptr<SampleBlock> &block,
                                        size_t count) {
  if (block && block->size() == count) {
    bool dense = true;
    for (const auto &e : map)
      if (e.first >= count || e.second != block->sample(e.first)) {
        dense = false;
        break;
      }
    if (dense) return block;
  }
  auto out = std::make_shared<SampleBlock>(count);
  std::memset(out->words(), 0, count * kCtBytes);
  std::memset(out->variances(), 0, count * sizeof(double));
  for (const auto &e : map) {
    REQUIRE_DRAMATICALLY(e.first < count, "ciphertext index out of range");
    std::memcpy(out->sample(e.first)->a, e.second->a, IdashParams::N * sizeof(int32_t));
    std::memcpy(out->sample(e.first)->b, e.second->b, IdashParams::N * sizeof(int32_t));
    *out->sample(e.first)->current_variance = *e.second->current_variance;
  }
  return out;
}

}  // namespace

// ---------------------------------------------------------------------------
// parameters (eval/idash.cpp:128-246)

namespace {

template <typename T>
void put(std::ostream &out, const T &v) {
  out.write(reinterpret_cast<const char *>(&v), sizeof(T));
}
template <typename T>
void get(std::istream &in, T &v) {
  in.read(reinterpret_cast<char *>(&v), sizeof(T));
}

void write_params_stream(const IdashParams &p, std::ostream &out) {
  put(out, p.NUM_SAMPLES);
  put(out, p.NUM_INPUT_POSITIONS);
  put(out, p.NUM_OUTPUT_POSITIONS);
  put(out, p.NUM_INPUT_FEATURES);
  put(out, p.NUM_OUTPUT_FEATURES);
  put(out, p.NUM_REGIONS);
  put(out, p.REGION_SIZE);
  REQUIRE_DRAMATICALLY(p.NUM_INPUT_POSITIONS == p.in_features_index.size(),
                       "NUM_INPUT_POSITIONS != in_features_index.size()");
  for (const auto &e : p.in_features_index) {
    put(out, e.first);
    for (FeatBigIndex b : e.second) put(out, b);
  }
  REQUIRE_DRAMATICALLY(p.NUM_OUTPUT_POSITIONS == p.out_features_index.size(),
                       "NUM_OUTPUT_POSITIONS != out_features_index.size()");
  REQUIRE_DRAMATICALLY(p.NUM_OUTPUT_POSITIONS == p.out_position_names.size(),
                       "NUM_OUTPUT_POSITIONS != out_position_names.size()");
  for (const auto &e : p.out_position_names) {
    put(out, e.first);
    out.write(e.second.c_str(), (std::streamsize)e.second.size() + 1);
    for (FeatBigIndex b : p.out_features_index.at(e.first)) put(out, b);
  }
}

void read_params_stream(IdashParams &p, std::istream &in) {
  get(in, p.NUM_SAMPLES);
  get(in, p.NUM_INPUT_POSITIONS);
  get(in, p.NUM_OUTPUT_POSITIONS);
  get(in, p.NUM_INPUT_FEATURES);
  get(in, p.NUM_OUTPUT_FEATURES);
  get(in, p.NUM_REGIONS);
  get(in, p.REGION_SIZE);
  for (uint32_t i = 0; i < p.NUM_INPUT_POSITIONS; ++i) {
    uint64_t pos;
    get(in, pos);
    std::array<FeatBigIndex, 3> &b = p.in_features_index.emplace(pos, std::array<FeatBigIndex, 3>()).first->second;
    for (int j = 0; j < 3; ++j) get(in, b[j]);
  }
  REQUIRE_DRAMATICALLY(p.NUM_INPUT_POSITIONS == p.in_features_index.size(),
                       "NUM_INPUT_POSITIONS != in_features_index.size()");
  for (uint32_t i = 0; i < p.NUM_OUTPUT_POSITIONS; ++i) {
    uint64_t pos;
    get(in, pos);
    string name;
    for (int c = in.get(); c > 0; c = in.get()) name.push_back((char)c);
    p.out_position_names.push_back({pos, name});
    std::array<FeatBigIndex, 3> &b = p.out_features_index.emplace(pos, std::array<FeatBigIndex, 3>()).first->second;
    for (int j = 0; j < 3; ++j) get(in, b[j]);
  }
  REQUIRE_DRAMATICALLY(p.NUM_OUTPUT_POSITIONS == p.out_features_index.size(),
                       "NUM_OUTPUT_POSITIONS != out_features_index.size()");
  REQUIRE_DRAMATICALLY(p.NUM_OUTPUT_POSITIONS == p.out_position_names.size(),
                       "NUM_OUTPUT_POSITIONS != out_position_names.size()");
}

}  // namespace

void write_params(const IdashParams &params, const string &filename) {
  std::ofstream out(filename, std::ios::binary);
  REQUIRE_DRAMATICALLY(out.is_open(), "Cannot open parameters file for write");
  write_params_stream(params, out);
}

void read_params(IdashParams &params, const string &filename) {
  std::ifstream in(filename, std::ios::binary);
  REQUIRE_DRAMATICALLY(in.is_open(), "Cannot open parameters file for read");
  read_params_stream(params, in);
}

// ---------------------------------------------------------------------------
// keys (eval/idash.cpp:476-502). The key stream after the parameters is this
// build's own: N and k (int32), alpha_min and alpha_max (double), then the
// N binary key coefficients (int32).

void write_key(const IdashKey &key, const string &filename) {
  std::ofstream out(filename, std::ios::binary);
  REQUIRE_DRAMATICALLY(out.is_open(), "Cannot open key file for write");
  write_params_stream(*key.idashParams, out);
  const TLweParams *tp = key.tlweKey->params;
  put(out, tp->N);
  put(out, tp->k);
  put(out, tp->alpha_min);
  put(out, tp->alpha_max);
  out.write(reinterpret_cast<const char *>(key.tlweKey->key), sizeof(key.tlweKey->key));
}

void read_key(IdashKey &key, const string &filename) {
  std::ifstream in(filename, std::ios::binary);
  REQUIRE_DRAMATICALLY(in.is_open(), "Cannot open key file for read");
  IdashParams *params = new IdashParams();
  read_params_stream(*params, in);
  TLweParams tp;
  get(in, tp.N);
  get(in, tp.k);
  get(in, tp.alpha_min);
  get(in, tp.alpha_max);
  REQUIRE_DRAMATICALLY(in && tp.N == (int32_t)IdashParams::N && tp.k == 1, "unsupported key parameters");
  TLweKey *k = new TLweKey();
  k->params = IdashParams::tlweParams;
  in.read(reinterpret_cast<char *>(k->key), sizeof(k->key));
  REQUIRE_DRAMATICALLY(in, "truncated key file");
  for (int32_t c : k->key) REQUIRE_DRAMATICALLY(c == 0 || c == 1, "key is not binary");
  key.idashParams = params;
  key.tlweKey = k;
}

// ---------------------------------------------------------------------------
// model (eval/idash.cpp:47-114)

namespace {
std::pair<uint64_t, uint8_t> split_pos_snp(const string &s) {
  const size_t u = s.find('_');
  return {std::stoull(s.substr(0, u)), (uint8_t)std::stoi(s.substr(u + 1))};
}
}  // namespace

namespace {

// Model cache: <model dir>/.ihe_model_cache holds every .hr file's entries
// (already mapped to big indices) with the file's size and modification time,
// and a signature of the tag index; it is used when all of these still match.
constexpr char kCacheMagic[8] = {'I', 'H', 'E', 'M', 'O', 'D', 'L', '1'};

struct HrFile {
  uint64_t pos;
  uint32_t snp, out_bidx;
  string path;
  uint64_t size = 0;
  int64_t mtime_ns = -1;
  std::vector<std::pair<FeatBigIndex, int32_t>> entries;  // in read() order
};

// Order-independent signature of the tag index: every (position, class, big
// index) triple and the feature counts.
uint64_t index_signature(const IdashParams &p) {
  auto mix = [](uint64_t x) {
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    return x ^ (x >> 31);
  };
  uint64_t h = mix(0x9E3779B97F4A7C15ull ^ p.NUM_INPUT_POSITIONS) ^
               mix(0xD6E8FEB86659FD93ull ^ p.NUM_INPUT_FEATURES);
  for (const auto &e : p.in_features_index)
    for (uint64_t snp = 0; snp < 3; ++snp)
      h += mix(mix(e.first * 3 + snp) ^ e.second[snp]);
  return h;
}

bool stat_file(HrFile &f) {
  struct stat st;
  if (::stat(f.path.c_str(), &st)) return false;
  f.size = (uint64_t)st.st_size;
  f.mtime_ns = (int64_t)st.st_mtim.tv_sec * 1000000000ll + st.st_mtim.tv_nsec;
  return true;
}

bool load_cache(const string &file, uint64_t sig, std::vector<HrFile> &files) {
  FILE *f = std::fopen(file.c_str(), "rb");
  if (!f) return false;
  char magic[8];
  uint64_t csig = 0, n = 0;
  bool ok = std::fread(magic, 8, 1, f) == 1 && std::memcmp(magic, kCacheMagic, 8) == 0 &&
            std::fread(&csig, 8, 1, f) == 1 && csig == sig && std::fread(&n, 8, 1, f) == 1 &&
            n == files.size();
  for (size_t i = 0; ok && i < files.size(); ++i) {
    uint64_t pos, size, cnt;
    int64_t mt;
    uint32_t snp, bidx;
    ok = std::fread(&pos, 8, 1, f) == 1 && std::fread(&snp, 4, 1, f) == 1 &&
         std::fread(&bidx, 4, 1, f) == 1 && std::fread(&size, 8, 1, f) == 1 &&
         std::fread(&mt, 8, 1, f) == 1 && std::fread(&cnt, 8, 1, f) == 1 && pos == files[i].pos &&
         snp == files[i].snp && bidx == files[i].out_bidx && size == files[i].size &&
         mt == files[i].mtime_ns && cnt < (1u << 24);
    if (ok) {
      files[i].entries.resize(cnt);
      ok = cnt == 0 || std::fread(files[i].entries.data(), sizeof(files[i].entries[0]), cnt, f) == cnt;
    }
  }
  std::fclose(f);
  return ok;
}

void save_cache(const string &file, uint64_t sig, const std::vector<HrFile> &files) {
  const string tmp = file + ".tmp" + std::to_string(::getpid());
  FILE *f = std::fopen(tmp.c_str(), "wb");
  if (!f) return;  // read-only model directory: no cache
  const uint64_t n = files.size();
  bool ok = std::fwrite(kCacheMagic, 8, 1, f) == 1 && std::fwrite(&sig, 8, 1, f) == 1 &&
            std::fwrite(&n, 8, 1, f) == 1;
  for (const HrFile &h : files) {
    const uint64_t cnt = h.entries.size();
    ok = ok && std::fwrite(&h.pos, 8, 1, f) == 1 && std::fwrite(&h.snp, 4, 1, f) == 1 &&
         std::fwrite(&h.out_bidx, 4, 1, f) == 1 && std::fwrite(&h.size, 8, 1, f) == 1 &&
         std::fwrite(&h.mtime_ns, 8, 1, f) == 1 && std::fwrite(&cnt, 8, 1, f) == 1 &&
         (cnt == 0 || std::fwrite(h.entries.data(), sizeof(h.entries[0]), cnt, f) == cnt);
  }
  ok = std::fclose(f) == 0 && ok;
  if (ok) ok = std::rename(tmp.c_str(), file.c_str()) == 0;
  if (!ok) std::remove(tmp.c_str());
}

}  // namespace

// One .hr file per target position and class (eval/idash.cpp:66-90), read
// with read() (eval/parse_vw.cpp:7-29) and mapped to big indices; the files
// are parsed in parallel, and a cache in the model directory replaces the
// parsing while every file keeps its size and modification time.
void read_model(Model &model, const IdashParams &params, const string &path) {
  Phases ph("read_model");
  std::vector<HrFile> files;
  files.reserve(3 * params.out_features_index.size());
  for (const auto &e : params.out_features_index)
    for (uint32_t snp = 0; snp < 3; ++snp)
      files.push_back({e.first, snp, e.second[snp],
                       path + "/" + std::to_string(e.first) + "_" + std::to_string(snp) + ".hr"});
  const int64_t n = (int64_t)files.size();
  std::vector<char> present(n);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) present[i] = stat_file(files[i]);
  const uint64_t sig = index_signature(params);
  const string cache = path + "/.ihe_model_cache";
  bool all_present = true;
  for (char c : present) all_present = all_present && c;
  ph.mark("stat of the .hr files");
  if (!(all_present && load_cache(cache, sig, files))) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
      HrFile &h = files[i];
      h.entries.clear();
      for (const auto &c : read(h.path)) {  // aborts on a missing file, like the reference
        if (c.first == "Constant") {
          h.entries.emplace_back(params.constant_bigIndex(), c.second);
        } else {
          const auto ps = split_pos_snp(c.first);
          h.entries.emplace_back(params.inBigIdx(ps.first, ps.second), c.second);
        }
      }
    }
    save_cache(cache, sig, files);
  }
  ph.mark("model cache or .hr parse");
  // each output's map is filled by one thread, in its file's order (the same
  // insertions as a sequential reader makes, so the same iteration order),
  // then moved into the model in file order
  std::vector<std::unordered_map<FeatBigIndex, int32_t>> rows(files.size());
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i)
    for (const auto &c : files[i].entries) rows[i][c.first] = c.second;
  for (size_t i = 0; i < files.size(); ++i) model.model[files[i].out_bidx] = std::move(rows[i]);
  ph.mark("model maps");
}

// ---------------------------------------------------------------------------
// plaintext tag file (eval/idash.cpp:252-286)

void read_plaintext_data(PlaintextData &plaintext_data, const IdashParams &params,
                         const string &filename) {
  std::ifstream in(filename);
  REQUIRE_DRAMATICALLY(in, "file not found");
  string line;
  uint64_t read_positions = 0;
  while (read_positions < params.NUM_INPUT_POSITIONS && std::getline(in, line)) {
    std::istringstream iss(line);
    int chrom = 0;
    uint64_t pos, pos2;
    string name;
    iss >> chrom;
    REQUIRE_DRAMATICALLY(chrom == 22, "file format error");
    iss >> pos >> pos2 >> name;
    REQUIRE_DRAMATICALLY(iss, "file format error");
    std::vector<int8_t> &v = plaintext_data.data[pos];
    v.resize(params.NUM_SAMPLES);
    string g;
    for (uint32_t s = 0; s < params.NUM_SAMPLES; ++s) {
      iss >> g;
      v[s] = (g == "0" || g == "1" || g == "2") ? (int8_t)(g[0] - '0') : (int8_t)-1;
    }
    ++read_positions;
  }
  REQUIRE_DRAMATICALLY(read_positions == params.NUM_INPUT_POSITIONS,
                       "missing positions in plaintext file");
}

// ---------------------------------------------------------------------------
// key generation (eval/idash.cpp:315-434)

IdashParams *keygen_ph1(const string &targetFile, const string &challengeFile,
                        const bool targetFilePosOnly) {
  IdashParams *p = new IdashParams();
  p->NUM_OUTPUT_POSITIONS = 0;
  p->NUM_OUTPUT_FEATURES = 0;
  string line;
  uint64_t last_target = 0;
  {
    std::ifstream target(targetFile);
    while (std::getline(target, line)) {
      std::istringstream iss(line);
      uint64_t pos;
      string name;
      if (targetFilePosOnly) {
        iss >> pos;
        name = std::to_string(pos);
        REQUIRE_DRAMATICALLY(iss, "file format error");
      } else {
        int chrom = 0;
        uint64_t pos2;
        iss >> chrom;
        REQUIRE_DRAMATICALLY(chrom == 22, "file format error");
        iss >> pos >> pos2 >> name;
        REQUIRE_DRAMATICALLY(pos2 == pos + 1, "file format error");
        REQUIRE_DRAMATICALLY(iss, "file format error");
      }
      for (uint64_t snp = 0; snp < 3; ++snp) p->registerOutBigIdx(pos, snp, p->NUM_OUTPUT_FEATURES++);
      p->NUM_OUTPUT_POSITIONS++;
      p->registerOutPositionName(pos, name);
      last_target = pos;
    }
  }
  // tags: all of them up to the 50th one past the last target
  p->NUM_SAMPLES = 0;
  p->NUM_INPUT_POSITIONS = 0;
  p->NUM_INPUT_FEATURES = 0;
  {
    std::ifstream tags(challengeFile);
    uint64_t past_last = 0;
    while (std::getline(tags, line)) {
      std::istringstream iss(line);
      int chrom = 0;
      uint64_t pos;
      iss >> chrom;
      REQUIRE_DRAMATICALLY(chrom == 22, "file format error");
      iss >> pos;
      REQUIRE_DRAMATICALLY(iss, "file format error");
      if (pos > last_target && ++past_last > 50) break;
      for (uint64_t snp = 0; snp < 3; ++snp) p->registerInBigIdx(pos, snp, p->NUM_INPUT_FEATURES++);
      p->NUM_INPUT_POSITIONS++;
      if (p->NUM_SAMPLES == 0) {
        string pos2, name, g;
        iss >> pos2 >> name;
        REQUIRE_DRAMATICALLY(iss, "file format error");
        while (iss >> g) p->NUM_SAMPLES++;
      }
    }
  }
  p->NUM_INPUT_FEATURES++;  // the constant
  REQUIRE_DRAMATICALLY(p->NUM_SAMPLES >= 1 && p->NUM_SAMPLES <= IdashParams::N,
                       "NUM_SAMPLES must be in [1, N]");
  p->NUM_REGIONS = IdashParams::N / p->NUM_SAMPLES;
  p->REGION_SIZE = IdashParams::N / p->NUM_REGIONS;
  REQUIRE_DRAMATICALLY(p->REGION_SIZE >= p->NUM_SAMPLES, "REGION_SIZE must be >= NUM_SAMPLES ");
  std::cout << "target_file (headers): " << targetFile << std::endl;
  std::cout << "tag_file: " << challengeFile << std::endl;
  std::cout << "NUM_SAMPLES: " << p->NUM_SAMPLES << std::endl;
  std::cout << "NUM_REGIONS: " << p->NUM_REGIONS << std::endl;
  std::cout << "REGION_SIZE: " << p->REGION_SIZE << std::endl;
  std::cout << "NUM_TARGET_POSITIONS: " << p->NUM_OUTPUT_POSITIONS << std::endl;
  std::cout << "NUM_TAG_POSITIONS: " << p->NUM_INPUT_POSITIONS << std::endl;
  return p;
}

IdashKey *keygen_ph2(IdashParams *idashParams) {
  TLweKey *k = new TLweKey();
  k->params = IdashParams::tlweParams;
  std::mt19937_64 rng(fresh_seed());
  for (uint32_t i = 0; i < IdashParams::N; i += 64) {
    const uint64_t bits = rng();
    for (uint32_t j = 0; j < 64; ++j) k->key[i + j] = (int32_t)((bits >> j) & 1u);
  }
  return new IdashKey(idashParams, k);
}

// ---------------------------------------------------------------------------
// ciphertext files (eval/idash.cpp:513-613): uint64 count, then per record a
// uint32 index and one sample stream (framing word, a, b, variance).

namespace {

constexpr size_t kRecordBytes = 4 + kSampleStreamBytes;  // index + sample stream: 8212 bytes

// Writes the count and one record per (index, sample), in the given order
// (with the collection's framing word). Records have a fixed size, so threads
// pack contiguous ranges of them into buffers of their own and write each at
// its file offset: the file is the one a sequential writer makes.
void write_records(const string &filename,
                   const std::vector<std::pair<uint32_t, const TLweSample *>> &recs,
                   int32_t stream_word, const char *open_error) {
  const int fd = ::open(filename.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
  REQUIRE_DRAMATICALLY(fd >= 0, open_error);
  auto write_at = [fd](const char *p, size_t n, off_t off) {
    while (n) {
      const ssize_t w = ::pwrite(fd, p, n, off);
      if (w <= 0) return false;
      p += w, n -= (size_t)w, off += w;
    }
    return true;
  };
  const uint64_t n = recs.size();
  bool ok = write_at(reinterpret_cast<const char *>(&n), 8, 0);
  constexpr size_t kPer = 1024;  // records per buffer (8.4 MB)
  const int64_t blocks = (int64_t)((n + kPer - 1) / kPer);
#pragma omp parallel reduction(&& : ok)
  {
    std::vector<char> buf(kPer * kRecordBytes);
#pragma omp for schedule(dynamic, 1)
    for (int64_t blk = 0; blk < blocks; ++blk) {
      const size_t r0 = (size_t)blk * kPer, r1 = std::min<size_t>(n, r0 + kPer);
      char *p = buf.data();
      for (size_t r = r0; r < r1; ++r, p += kRecordBytes) {
        const TLweSample *s = recs[r].second;
        std::memcpy(p, &recs[r].first, 4);
        std::memcpy(p + 4, &stream_word, 4);
        std::memcpy(p + 8, s->a, IdashParams::N * 4);
        std::memcpy(p + 8 + IdashParams::N * 4, s->b, IdashParams::N * 4);
        std::memcpy(p + 8 + kCtBytes, s->current_variance, 8);
      }
      ok = ok && write_at(buf.data(), (r1 - r0) * kRecordBytes, (off_t)(8 + r0 * kRecordBytes));
    }
  }
  ok = ::close(fd) == 0 && ok;
  REQUIRE_DRAMATICALLY(ok, "write failed");
}

// Reads the records into a dense block of `slots` samples (slot = index) and
// binds them in `map`; slots no record fills are zeroed. Every record must
// open its sample stream with the same framing word, which the block keeps.
// The file is mapped a window of records at a time: the window's headers are
// checked in file order (the errors of a sequential reader), then its
// payloads are copied to their slots by several threads.
template <typename Map>
std::shared_ptr<SampleBlock> read_records(const string &filename, size_t slots, Map &map,
                                          const char *what) {
  Phases ph("read_records");
  const int fd = ::open(filename.c_str(), O_RDONLY);
  REQUIRE_DRAMATICALLY(fd >= 0, "Cannot open " << what << " file for read");
  struct stat st;
  REQUIRE_DRAMATICALLY(::fstat(fd, &st) == 0, "Cannot open " << what << " file for read");
  const size_t size = (size_t)st.st_size;
  uint64_t count = 0;
  REQUIRE_DRAMATICALLY(size >= 8 && ::pread(fd, &count, 8, 0) == 8, "truncated " << what << " file");
  REQUIRE_DRAMATICALLY(count <= (size - 8) / kRecordBytes, "truncated " << what << " file");
  auto block = std::make_shared<SampleBlock>(slots);
  std::vector<char> filled(slots, 0);
  const size_t page = (size_t)::sysconf(_SC_PAGESIZE);
  constexpr uint64_t kWindow = 8192;  // records per mapping (67 MB)
  std::vector<uint32_t> index(kWindow);
  for (uint64_t r0 = 0; r0 < count; r0 += kWindow) {
    const uint64_t r1 = std::min(count, r0 + kWindow);
    const size_t off = 8 + r0 * kRecordBytes, moff = off & ~(page - 1);
    const size_t len = 8 + r1 * kRecordBytes - moff;
    void *m = ::mmap(nullptr, len, PROT_READ, MAP_PRIVATE | MAP_POPULATE, fd, (off_t)moff);
    REQUIRE_DRAMATICALLY(m != MAP_FAILED, "Cannot read " << what << " file");
    const char *recs = static_cast<const char *>(m) + (off - moff);
    for (uint64_t i = r0; i < r1; ++i) {
      const char *rec = recs + (i - r0) * kRecordBytes;
      uint32_t ix;
      int32_t word;
      std::memcpy(&ix, rec, 4);
      std::memcpy(&word, rec + 4, 4);
      REQUIRE_DRAMATICALLY(ix < slots, what << " index out of range");
      if (i == 0) block->stream_word = word;
      REQUIRE_DRAMATICALLY(word == block->stream_word,
                           what << " records open with different framing words");
      REQUIRE_DRAMATICALLY(map.emplace(ix, block->sample(ix)).second, "duplicate " << what << " index");
      filled[ix] = 1;
      index[i - r0] = ix;
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)(r1 - r0); ++i) {
      const char *rec = recs + (size_t)i * kRecordBytes;
      TLweSample *s = block->sample(index[i]);
      std::memcpy(s->a, rec + 8, IdashParams::N * 4);
      std::memcpy(s->b, rec + 8 + IdashParams::N * 4, IdashParams::N * 4);
      std::memcpy(s->current_variance, rec + 8 + kCtBytes, 8);
    }
    ::munmap(m, len);
  }
  ::close(fd);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)slots; ++i)
    if (!filled[i]) {
      TLweSample *s = block->sample(i);
      std::memset(s->a, 0, kCtBytes);
      *s->current_variance = 0.0;
    }
  ph.mark(what);
  return block;
}

// The framing word a collection's samples are written with: the word of the
// block they were read into or computed from (0 for blocks encrypted here).
int32_t stream_word_of(const std::shared_ptr<SampleBlock> &block) {
  return block ? block->stream_word : 0;
}

size_t num_input_cts(const IdashParams &p) {
  return (3 * (size_t)p.NUM_INPUT_POSITIONS + p.NUM_REGIONS - 1) / p.NUM_REGIONS;
}

}  // namespace

void write_encrypted_data(const EncryptedData &encrypted_data, const IdashParams &,
                          const string &filename) {
  std::vector<std::pair<uint32_t, const TLweSample *>> recs(encrypted_data.enc_data.begin(),
                                                            encrypted_data.enc_data.end());
  write_records(filename, recs, stream_word_of(encrypted_data.block),
                "Cannot open encrypted data file for write");
}

void read_encrypted_data(EncryptedData &encrypted_data, const IdashParams &params,
                         const string &filename) {
  encrypted_data.block =
      read_records(filename, num_input_cts(params), encrypted_data.enc_data, "encrypted data");
}

void write_encrypted_predictions(const EncryptedPredictions &encrypted_preds, const IdashParams &,
                                 const string &filename) {
  Phases ph("write_encrypted_predictions");
  std::vector<std::pair<uint32_t, const TLweSample *>> recs(encrypted_preds.score.begin(),
                                                            encrypted_preds.score.end());
  write_records(filename, recs, stream_word_of(encrypted_preds.block),
                "Cannot open encrypted predictions file for write");
  ph.mark("records written");
}

void read_encrypted_predictions(EncryptedPredictions &encrypted_preds, const IdashParams &params,
                                const string &filename) {
  encrypted_preds.block = read_records(filename, params.NUM_OUTPUT_FEATURES, encrypted_preds.score,
                                       "encrypted predictions");
}

// ---------------------------------------------------------------------------
// result CSV (eval/idash.cpp:436-469): default ostream formatting of a float
// is printf's %g with 6 significant digits.

void write_decrypted_predictions(const DecryptedPredictions &predictions, const IdashParams &params,
                                 const string &filename, const bool PRINT_POS_NAME) {
  Phases ph("write_decrypted_predictions");
  const int fd = ::open(filename.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
  REQUIRE_DRAMATICALLY(fd >= 0, "Cannot open result file for write");
  std::atomic<bool> ok{true};
  auto write_at = [fd, &ok](const char *p, size_t n, off_t off) {
    while (n) {
      const ssize_t w = ::pwrite(fd, p, n, off);
      if (w <= 0) {
        ok = false;
        return;
      }
      p += w, n -= (size_t)w, off += w;
    }
  };
  const char header[] = "Subject ID,target SNP,0,1,2\n";
  write_at(header, sizeof(header) - 1, 0);
  off_t at = sizeof(header) - 1;
  std::vector<const std::array<std::vector<float>, 3> *> rows;
  std::vector<string> labels;
  rows.reserve(params.out_position_names.size());
  for (const auto &e : params.out_position_names) {
    rows.push_back(&predictions.score.at(e.first));
    for (const std::vector<float> &v : *rows.back())
      REQUIRE_DRAMATICALLY(v.size() >= params.NUM_SAMPLES, "fewer scores than samples for a position");
    labels.push_back(PRINT_POS_NAME ? e.second : std::to_string(e.first));
  }
  // Samples are formatted in parallel, one buffer each, a block of them at a
  // time; each buffer's file offset follows from the lengths before it, so
  // while the next block is formatted, writer threads write this block's
  // buffers at their offsets. A buffer holds a sample's lines at their longest
  // (20-digit sample number, three 16-character scores); its pages are touched
  // only as far as the lines reach.
  const int64_t ns = params.NUM_SAMPLES, block = 64;
  constexpr int kWriters = 4;
  size_t label_bytes = 0;
  for (const string &l : labels) label_bytes += l.size();
  const size_t cap = rows.size() * (20 + 1 + 3 * 17 + 1) + label_bytes;
  std::vector<std::unique_ptr<char[]>> text[2];
  for (auto &t : text)
    for (int64_t i = 0; i < block; ++i) t.emplace_back(new char[cap + 8]);  // + format_g6's scratch
  std::vector<size_t> length[2] = {std::vector<size_t>(block), std::vector<size_t>(block)};
  std::vector<off_t> offset[2] = {std::vector<off_t>(block), std::vector<off_t>(block)};
  std::vector<std::thread> writers;
  double t_format = 0, t_wait = 0;
  using clk = std::chrono::steady_clock;
  // the block's scores gathered position-major ([position][class][sample]),
  // so that a sample's lines read them at one stride instead of one vector
  // (and page) per score
  std::vector<float> tile(rows.size() * 3 * block);
  ph.mark("setup");
  for (int64_t s0 = 0, buf = 0; s0 < ns; s0 += block, buf ^= 1) {
    const int64_t s1 = std::min(ns, s0 + block);
    std::vector<std::unique_ptr<char[]>> &tb = text[buf];
    std::vector<size_t> &lb = length[buf];
    const clk::time_point t0 = clk::now();
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < (int64_t)rows.size(); ++t)
      for (int k = 0; k < 3; ++k)
        std::memcpy(&tile[((size_t)t * 3 + k) * block], (*rows[t])[k].data() + s0,
                    (size_t)(s1 - s0) * sizeof(float));
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t s = s0; s < s1; ++s) {
      // "%lld,%s,%g,%g,%g\n", the scores by format_g6 (printf's %g of a float)
      char *const begin = tb[s - s0].get();
      char *p = begin;
      char num[24];
      const std::to_chars_result sr = std::to_chars(num, num + sizeof(num), (long long)s);
      const size_t slen = (size_t)(sr.ptr - num);
      const float *sc = tile.data() + (s - s0);
      for (size_t t = 0; t < rows.size(); ++t, sc += 3 * block) {
        std::memcpy(p, num, slen);
        p += slen;
        *p++ = ',';
        std::memcpy(p, labels[t].data(), labels[t].size());
        p += labels[t].size();
        for (int k = 0; k < 3; ++k) {
          *p++ = ',';
          p = ihe_host::format_g6(p, sc[k * block]);
        }
        *p++ = '\n';
      }
      lb[s - s0] = (size_t)(p - begin);
    }
    std::vector<off_t> &ob = offset[buf];
    for (int64_t i = 0; i < s1 - s0; ++i) {
      ob[i] = at;
      at += (off_t)lb[i];
    }
    // the previous block's writers (the other buffers) end first
    const clk::time_point t1 = clk::now();
    for (std::thread &w : writers) w.join();
    writers.clear();
    t_format += std::chrono::duration<double, std::milli>(t1 - t0).count();
    t_wait += std::chrono::duration<double, std::milli>(clk::now() - t1).count();
    for (int w = 0; w < kWriters; ++w)
      writers.emplace_back([&tb, &lb, &ob, &write_at, w, n = s1 - s0] {
        for (int64_t i = w; i < n; i += kWriters) write_at(tb[i].get(), lb[i], ob[i]);
      });
  }
  ph.mark("blocks");
  for (std::thread &w : writers) w.join();
  REQUIRE_DRAMATICALLY(::close(fd) == 0 && ok, "result file write failed");
  ph.mark("last writes, close");
  ph.note("formatting", t_format);
  ph.note("waiting for the writes", t_wait);
}

// ---------------------------------------------------------------------------
// encryption (eval/idash.cpp:615-679)

void encrypt_zero_sample(TLweSample *s, const IdashKey &key) {
  const ihe_layout lay = layout_of(*key.idashParams);
  const size_t cts = (3 + lay.num_regions - 1) / lay.num_regions;
  DeviceBuffer d_ct(cts * kCtBytes), d_var(cts * sizeof(double));
  DeviceKey dk(key.tlweKey);
  check(ihe_encrypt(engine(), dk.k, &lay, 1, nullptr, fresh_seed(), d_ct.as<int32_t>(),
                    d_var.as<double>()),
        "encrypt");
  check(ihe_memcpy_d2h(engine(), s->a, d_ct.p, IdashParams::N * 4), "copy");
  check(ihe_memcpy_d2h(engine(), s->b, d_ct.as<int32_t>() + IdashParams::N, IdashParams::N * 4), "copy");
  check(ihe_memcpy_d2h(engine(), s->current_variance, d_var.p, 8), "copy");
}

// Phase 1: encryptions of zero, one per FeatIndex of the tag big indices, made
// on the GPU (a uniform, b = a*s + e) into a dense block.
void encrypt_data_ph1(EncryptedData &enc_data, const PlaintextData &plain_data, const IdashKey &key) {
  const IdashParams &p = *key.idashParams;
  REQUIRE_DRAMATICALLY(plain_data.data.size() == p.NUM_INPUT_POSITIONS, "Incomplete plaintext");
  for (const auto &e : plain_data.data)
    REQUIRE_DRAMATICALLY(e.second.size() == p.NUM_SAMPLES, "plaintext dimensions inconsistency");
  const size_t cts = num_input_cts(p);
  auto block = std::make_shared<SampleBlock>(cts);
  if (cts) {
    const ihe_layout lay = layout_of(p);
    DeviceBuffer d_ct(cts * kCtBytes), d_var(cts * sizeof(double));
    DeviceKey dk(key.tlweKey);
    check(ihe_encrypt(engine(), dk.k, &lay, p.NUM_INPUT_POSITIONS, nullptr, fresh_seed(),
                      d_ct.as<int32_t>(), d_var.as<double>()),
          "encrypt");
    check(ihe_memcpy_d2h(engine(), block->words(), d_ct.p, cts * kCtBytes), "copy");
    check(ihe_memcpy_d2h(engine(), block->variances(), d_var.p, cts * sizeof(double)), "copy");
  }
  // bind every FeatIndex the tags use (eval/idash.cpp:631-637)
  for (const auto &e : plain_data.data)
    for (uint64_t snp = 0; snp < 3; ++snp) {
      const FeatIndex f = p.feature_indexOf(p.inBigIdx(e.first, snp));
      REQUIRE_DRAMATICALLY(f < cts, "memory error");
      enc_data.enc_data.emplace(f, block->sample(f));
    }
  enc_data.block = block;
}

// Phase 2: add the one-hot scores (ONE at the genotype's column, the NAN
// triple when missing) at slot r*RS + sample.
void encrypt_data_ph2(EncryptedData &enc_data, const PlaintextData &plain_data, const IdashKey &key) {
  const IdashParams &p = *key.idashParams;
  REQUIRE_DRAMATICALLY(plain_data.data.size() == p.NUM_INPUT_POSITIONS, "Incomplete plaintext");
  std::vector<const std::pair<const uint64_t, std::vector<int8_t>> *> rows;
  for (const auto &e : plain_data.data) rows.push_back(&e);
#pragma omp parallel for num_threads(NB_THREADS)
  for (size_t i = 0; i < rows.size(); ++i) {
    const uint64_t pos = rows[i]->first;
    const std::vector<int8_t> &g = rows[i]->second;
    for (uint64_t s = 0; s < p.NUM_SAMPLES; ++s) {
      const int8_t v = g[s];
      if (v == 0 || v == 1 || v == 2) {
        enc_data.setScore(p.inBigIdx(pos, (uint64_t)v), s, IdashParams::ONE_IN_T32, p);
      } else {
        enc_data.setScore(p.inBigIdx(pos, 0), s, IdashParams::NAN_0_IN_T32, p);
        enc_data.setScore(p.inBigIdx(pos, 1), s, IdashParams::NAN_1_IN_T32, p);
        enc_data.setScore(p.inBigIdx(pos, 2), s, IdashParams::NAN_2_IN_T32, p);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// cloud step (eval/idash.cpp:763-848) on the GPU

void cloud_compute_score(EncryptedPredictions &enc_preds, const EncryptedData &enc_data,
                         const Model &model, const IdashParams &params) {
  REQUIRE_DRAMATICALLY(IdashParams::k == 1, "k must be 1");
  Phases ph("cloud_compute_score");
  // the model's rows in its own iteration order (which fixes the order of
  // each row's entries, and so the variances' summation order)
  using Row = std::unordered_map<FeatBigIndex, int32_t>;
  std::vector<std::pair<uint32_t, const Row *>> rows;
  rows.reserve(model.model.size());
  uint32_t num_out = 0;
  for (const auto &e : model.model) {
    rows.push_back({e.first, &e.second});
    num_out = std::max(num_out, e.first + 1);
  }
  // the output block (page-locked, 8 KB per output) is allocated on a second
  // thread while this one builds the band and runs the step; the context
  // exists before the thread starts
  (void)engine_if_present();
  std::future<std::shared_ptr<SampleBlock>> out_block = std::async(
      std::launch::async, [num_out] { return std::make_shared<SampleBlock>(num_out); });
  // model -> CSR over output big indices
  std::vector<uint32_t> row_ptr(num_out + 1, 0);
  for (const auto &r : rows) row_ptr[r.first + 1] = (uint32_t)r.second->size();
  for (uint32_t o = 0; o < num_out; ++o) row_ptr[o + 1] += row_ptr[o];
  std::vector<uint32_t> in_bidx(row_ptr[num_out]);
  std::vector<int32_t> weight(row_ptr[num_out]);
  // every ciphertext the model reads must be present (getTLWE,
  // eval/idash.h:162-166), checked before the device is touched
  const size_t cts = num_input_cts(params);
  std::vector<char> present(cts, 0);
  for (const auto &e : enc_data.enc_data)
    if (e.first < cts) present[e.first] = 1;
  bool all_present = true;
#pragma omp parallel for schedule(dynamic, 256) reduction(&& : all_present)
  for (int64_t i = 0; i < (int64_t)rows.size(); ++i) {
    uint32_t at = row_ptr[rows[i].first];
    for (const auto &c : *rows[i].second) {
      in_bidx[at] = c.first;
      weight[at++] = c.second;
      if (c.first != params.constant_bigIndex()) {
        const FeatIndex f = params.feature_indexOf(c.first);
        all_present = all_present && f < cts && present[f];
      }
    }
  }
  // outputs are created once each (eval/idash.cpp:777-779)
  for (const auto &r : rows)
    if (enc_preds.score.count(r.first)) DIE_DRAMATICALLY("output ciphertext created twice");
  REQUIRE_DRAMATICALLY(all_present, "missing input ciphertext");
  ph.mark("model CSR and input check");

  std::shared_ptr<SampleBlock> in = dense_view(enc_data.enc_data, enc_data.block, cts);
  ph.mark("dense input view");
  std::shared_ptr<SampleBlock> out;
  if (num_out) {
    ihe_ctx *ctx = engine();
    const ihe_layout lay = layout_of(params);
    const ihe_model_csr csr = {num_out, row_ptr.data(), in_bidx.data(), weight.data()};
    ihe_band *band = nullptr;
    check(ihe_band_create(ctx, &lay, (uint32_t)cts, &csr, IHE_VARIANT_AUTO, &band), "model band");
    ph.mark("band build and upload");
    {
      DeviceBuffer d_in(cts * kCtBytes), d_in_var(cts * sizeof(double));
      DeviceBuffer d_out(num_out * kCtBytes), d_out_var(num_out * sizeof(double));
      ph.mark("device allocation");
      if (cts) {
        check(ihe_memcpy_h2d(ctx, d_in.p, in->words(), cts * kCtBytes), "copy");
        check(ihe_memcpy_h2d(ctx, d_in_var.p, in->variances(), cts * sizeof(double)), "copy");
      }
      ph.mark("inputs to the device");
      check(ihe_cloud_score(ctx, band, d_in.as<int32_t>(), d_in_var.as<double>(), d_out.as<int32_t>(),
                            d_out_var.as<double>()),
            "cloud");
      check(ihe_sync(ctx), "cloud");
      ph.mark("cloud step");
      out = out_block.get();
      ph.mark("output block (wait)");
      check(ihe_memcpy_d2h(ctx, out->words(), d_out.p, num_out * kCtBytes), "copy");
      check(ihe_memcpy_d2h(ctx, out->variances(), d_out_var.p, num_out * sizeof(double)), "copy");
      ph.mark("outputs to the host");
    }
    ihe_band_destroy(band);
    ph.mark("device release");
  } else {
    out = out_block.get();
  }
  out->stream_word = stream_word_of(enc_data.block);
  enc_preds.score.reserve(enc_preds.score.size() + model.model.size());
  for (const auto &r : rows) enc_preds.score.emplace(r.first, out->sample(r.first));
  enc_preds.block = out;
  ph.mark("prediction map");
}

// ---------------------------------------------------------------------------
// decryption (eval/idash.cpp:681-761): float(t32tod(phase_j)), j < NUM_SAMPLES

void decrypt_predictions(DecryptedPredictions &predictions, const EncryptedPredictions &enc_preds,
                         const IdashKey &key) {
  const IdashParams &p = *key.idashParams;
  const uint32_t ns = p.NUM_SAMPLES;
  Phases ph("decrypt_predictions");
  uint32_t num = 0;
  for (const auto &e : enc_preds.score) num = std::max(num, e.first + 1);
  // (output big index, its score vector), by index; the map's nodes stay put
  std::vector<std::pair<FeatBigIndex, std::vector<float> *>> dest;
  dest.reserve(3 * p.out_features_index.size());
  for (const auto &e : p.out_features_index) {
    std::array<std::vector<float>, 3> &v = predictions.score[e.first];
    for (int snp = 0; snp < 3; ++snp) {
      const FeatBigIndex b = e.second[snp];
      REQUIRE_DRAMATICALLY(enc_preds.score.count(b), "missing prediction ciphertext");
      num = std::max(num, b + 1);
      dest.push_back({b, &v[snp]});
    }
  }
  std::sort(dest.begin(), dest.end());
  ph.mark("score vectors");
  if (num == 0) return;
  std::shared_ptr<SampleBlock> in = dense_view(enc_preds.score, enc_preds.block, num);
  ph.mark("dense input view");
  ihe_ctx *ctx = engine();
  DeviceBuffer d_in(num * kCtBytes), d_out((size_t)num * ns * sizeof(float));
  DeviceKey dk(key.tlweKey);
  check(ihe_memcpy_h2d(ctx, d_in.p, in->words(), num * kCtBytes), "copy");
  check(ihe_decrypt(ctx, dk.k, ns, d_in.as<int32_t>(), num, d_out.as<float>()), "decrypt");
  check(ihe_sync(ctx), "decrypt");
  ph.mark("inputs to the device, decrypt");
  // The scores come back through two page-locked staging blocks of about 32 MB:
  // while several threads copy one block's rows into their score vectors (their
  // storage reserved beforehand, by this thread), the next rows cross the link
  // into the other block.
  for (auto &e : dest) e.second->reserve(ns);
  const size_t row_bytes = (size_t)ns * sizeof(float);
  const uint32_t rows = (uint32_t)std::max<size_t>(1, std::min<size_t>(num, (32u << 20) / row_bytes));
  struct Staging {
    void *p = nullptr;
    explicit Staging(size_t bytes) { check(ihe_host_alloc(engine(), bytes, &p), "pinned host allocation"); }
    Staging(const Staging &) = delete;
    ~Staging() { ihe_host_free(engine(), p); }
  } stage[2] = {Staging(rows * row_bytes), Staging(rows * row_bytes)};
  auto fetch = [&](uint32_t r0, int k) {
    const uint32_t r1 = std::min(num, r0 + rows);
    check(ihe_memcpy_d2h_async(ctx, stage[k].p, static_cast<const char *>(d_out.p) + r0 * row_bytes,
                               (r1 - r0) * row_bytes),
          "copy");
  };
  fetch(0, 0);
  auto d = dest.begin();
  for (uint32_t r0 = 0, k = 0; r0 < num; r0 += rows, k ^= 1) {
    const uint32_t r1 = std::min(num, r0 + rows);
    check(ihe_sync(ctx), "copy");  // rows r0..r1 are in block k
    if (r1 < num) fetch(r1, k ^ 1);
    const auto lo = d;
    while (d != dest.end() && d->first < r1) ++d;
    const int64_t n = d - lo;
    const float *block = static_cast<const float *>(stage[k].p);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      const float *src = block + (size_t)(lo[i].first - r0) * ns;
      lo[i].second->assign(src, src + ns);
    }
  }
  ph.mark("scores to the host");
}

// ---------------------------------------------------------------------------
// plaintext model (eval/idash.cpp:850-927), on the CPU: the validation path

PlaintextOnehot compute_plaintext_onehot(const PlaintextData &X, const IdashParams &params) {
  PlaintextOnehot oh;
  for (const auto &e : params.in_features_index) {
    const uint64_t pos = e.first;
    for (int v = 0; v < 3; ++v) oh[e.second[v]].resize(params.NUM_SAMPLES);
    const std::vector<int8_t> &g = X.data.at(pos);
    for (uint64_t s = 0; s < params.NUM_SAMPLES; ++s) {
      double val[3];
      switch (g[s]) {
        case 0: val[0] = params.ONE_IN_D, val[1] = 0, val[2] = 0; break;
        case 1: val[0] = 0, val[1] = params.ONE_IN_D, val[2] = 0; break;
        case 2: val[0] = 0, val[1] = 0, val[2] = params.ONE_IN_D; break;
        default: val[0] = params.NAN_0_IN_D, val[1] = params.NAN_1_IN_D, val[2] = params.NAN_2_IN_D;
      }
      for (int v = 0; v < 3; ++v) oh[e.second[v]][s] = val[v];
    }
  }
  return oh;
}

void compute_score(DecryptedPredictions &predictions, const PlaintextData &X, const Model &M,
                   const IdashParams &params) {
  PlaintextOnehot oh = compute_plaintext_onehot(X, params);
  for (const auto &e : params.out_features_index) {
    for (int snp = 0; snp < 3; ++snp) {
      std::vector<float> &out = predictions.score[e.first][snp];
      out.assign(params.NUM_SAMPLES, 0.0f);
      for (const auto &c : M.model.at(e.second[snp])) {
        for (uint64_t s = 0; s < params.NUM_SAMPLES; ++s) {
          if (c.first == params.constant_bigIndex())
            out[s] += c.second * IdashParams::ONE_IN_D;
          else
            out[s] += c.second * oh[c.first][s];
        }
      }
    }
  }
}

bool DecryptedPredictions::testEquals(const DecryptedPredictions &val1,
                                      const DecryptedPredictions &val2, const IdashParams &params) {
  for (uint64_t s = 0; s < params.NUM_SAMPLES; ++s)
    for (const auto &e : params.out_features_index)
      for (int snp = 0; snp < 3; ++snp) {
        const float x = val1.score.at(e.first)[snp][s], y = val2.score.at(e.first)[snp][s];
        if (x != y) std::cout << x << " " << y << std::endl;
      }
  return true;
}

// ---------------------------------------------------------------------------
// profiler (eval/idash.cpp:929-967)

Profiler::Profiler() : tw0(universalWallTime()), tc0(universalClockTime()) {}

double Profiler::universalWallTime() {
  struct timeval tv;
  if (gettimeofday(&tv, nullptr)) return 0;
  return (double)tv.tv_sec + (double)tv.tv_usec * 1e-6;
}

double Profiler::universalClockTime() { return (double)clock() / CLOCKS_PER_SEC; }
double Profiler::walltime() const { return universalWallTime() - tw0; }
double Profiler::clocktime() const { return universalClockTime() - tc0; }

long int Profiler::maxrss() const {
  struct rusage u;
  if (getrusage(RUSAGE_SELF, &u)) return -1;
  return u.ru_maxrss * 1000;
}
