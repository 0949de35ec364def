// idash-compatible host API (the reference's eval/idash.h:1-287), served by
// the gfx950 engine through the C ABI of include/idash_hip.h.
//
// Same type names, members, constants and free-function signatures as the
// reference, so its keygen/encrypt/cloud/decrypt mains and the plaintext
// validation code compile against this header unchanged. Ciphertexts live in
// dense SampleBlocks (tlwe.h) so that whole files move to and from the device
// in one copy. Errors print "ERROR: <message>" on stdout and abort(), like the
// reference (eval/idash.h:12-14).
#ifndef IDASH_2019_IDASH_H
#define IDASH_2019_IDASH_H

#include <array>
#include <cstdint>
#include <iostream>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "tlwe.h"

#define REQUIRE_DRAMATICALLY(cond, message) \
  if (!(cond)) {                            \
    std::cout << "ERROR: " << message << std::endl; \
    abort();                                \
  }

#define DIE_DRAMATICALLY(message)                   \
  {                                                 \
    std::cout << "ERROR: " << message << std::endl; \
    abort();                                        \
  }

// default file names (eval/idash.h:16-34)
#define CHALLENGE_FILE "../../data/tag_testing.txt"
#define MODEL_FILE "../../ml/model/final"
#define TARGET_FILE_POS_ONLY true
#if TARGET_FILE_POS_ONLY == true
#define TARGET_FILE "../../data/target_geno_model_coordinates.txt"
#else
#define TARGET_FILE "../../data/target_testing.txt"
#endif

#define PARAMS_FILE "params.bin"
#define KEYS_FILE "keys.bin"
#define ENCRYPTED_DATA_FILE "encrypted_data.bin"
#define ENCRYPTED_PREDICTION_FILE "encrypted_prediction.bin"
#define RESULT_FILE "result.csv"
#define RESULT_BYPOS_FILE "result_bypos.csv"

// host threads of the file and packing loops (eval/idash.h:36-38)
#ifndef NB_THREADS
#define NB_THREADS 4
#endif

typedef uint32_t FeatBigIndex;
typedef uint32_t FeatIndex;
typedef uint32_t FeatRegion;

struct IdashParams {
  static const uint32_t NUM_SNP_PER_POSITIONS = 3;

  // TRLWE parameters (eval/idash.h:49-52, eval/idash.cpp:20-23)
  static const uint32_t N = 1024;
  static const uint32_t k = 1;
  static const double alpha;
  static const TLweParams *tlweParams;

  // one-hot encoding scale and values (eval/idash.cpp:29-45)
  static const double IN_SCALING_FACTOR;
  static const double ONE_IN_D;
  static const double NAN_0_IN_D;
  static const double NAN_1_IN_D;
  static const double NAN_2_IN_D;
  static const Torus32 ONE_IN_T32;
  static const Torus32 NAN_0_IN_T32;
  static const Torus32 NAN_1_IN_T32;
  static const Torus32 NAN_2_IN_T32;

  uint32_t NUM_SAMPLES;
  uint32_t NUM_INPUT_POSITIONS;
  uint32_t NUM_OUTPUT_POSITIONS;
  uint32_t NUM_INPUT_FEATURES;  // including the constant
  uint32_t NUM_OUTPUT_FEATURES;
  uint32_t NUM_REGIONS;
  uint32_t REGION_SIZE;  // N / NUM_REGIONS >= NUM_SAMPLES

  std::unordered_map<uint64_t, std::array<FeatBigIndex, 3>> in_features_index;
  std::unordered_map<uint64_t, std::array<FeatBigIndex, 3>> out_features_index;
  std::vector<std::pair<uint64_t, std::string>> out_position_names;  // target file order

  FeatIndex feature_indexOf(uint32_t big_index) const { return big_index / NUM_REGIONS; }
  FeatRegion feature_regionOf(uint32_t big_index) const { return big_index % NUM_REGIONS; }
  FeatBigIndex feature_bigIndexOf(uint32_t index, uint32_t region) const {
    return index * NUM_REGIONS + region;
  }
  FeatBigIndex constant_bigIndex() const { return 0xFFFFFFFFu; }
  FeatBigIndex inBigIdx(const uint64_t &pos, uint64_t snp) const {
    return in_features_index.at(pos).at(snp);
  }
  FeatBigIndex outBigIdx(const uint64_t &pos, uint64_t snp) const {
    return out_features_index.at(pos).at(snp);
  }
  void registerInBigIdx(const uint64_t &pos, uint64_t snp, FeatBigIndex bidx) {
    in_features_index[pos][snp] = bidx;
  }
  void registerOutBigIdx(const uint64_t &pos, uint64_t snp, FeatBigIndex bidx) {
    out_features_index[pos][snp] = bidx;
  }
  void registerOutPositionName(const uint64_t &pos, const std::string &name) {
    out_position_names.push_back({pos, name});
  }
};

struct IdashKey {
  const IdashParams *idashParams;
  const TLweKey *tlweKey;

  IdashKey(const IdashParams *idashParams, const TLweKey *tlweKey);
  IdashKey() {}
};

struct PlaintextData {
  // per tag position: genotype per sample (0, 1, 2) or -1 when missing
  std::unordered_map<uint64_t, std::vector<int8_t>> data;
};

struct Model {
  // out big index -> (in big index -> integer coefficient); the constant
  // feature is constant_bigIndex()
  std::unordered_map<FeatBigIndex, std::unordered_map<FeatBigIndex, int32_t>> model;
};

// Encrypts zero into one sample (tLweSymEncryptZero at alpha_min).
void encrypt_zero_sample(TLweSample *s, const IdashKey &key);

struct EncryptedData {
  std::unordered_map<FeatIndex, TLweSample *> enc_data;
  // Dense storage: when set, sample FeatIndex f lives in slot f of `block`.
  std::shared_ptr<SampleBlock> block;
  std::vector<std::shared_ptr<SampleBlock>> extra;  // samples added one at a time

  void ensure_exists(const FeatBigIndex bidx, TLweSample *&nextSample, const IdashParams &params) {
    const FeatIndex f = params.feature_indexOf(bidx);
    if (enc_data.count(f) == 0) {
      enc_data.emplace(f, nextSample);
      ++nextSample;
    }
  }

  void ensure_exists(const FeatBigIndex bidx, const IdashKey &key) {
    const FeatIndex f = key.idashParams->feature_indexOf(bidx);
    if (enc_data.count(f) == 0) {
      extra.push_back(std::make_shared<SampleBlock>(1));
      TLweSample *s = extra.back()->sample(0);
      encrypt_zero_sample(s, key);
      enc_data.emplace(f, s);
    }
  }

  void setScore(FeatBigIndex bidx, uint64_t sampleId, Torus32 score, const IdashParams &params) {
    const FeatIndex idx = params.feature_indexOf(bidx);
    const FeatRegion rgn = params.feature_regionOf(bidx);
    Torus32 &c = enc_data.at(idx)->b[sampleId + rgn * params.REGION_SIZE];
    c = (Torus32)((uint32_t)c + (uint32_t)score);
  }

  const TLweSample *getTLWE(FeatBigIndex inBidx, const IdashParams &params) const {
    const FeatIndex idx = params.feature_indexOf(inBidx);
    REQUIRE_DRAMATICALLY(enc_data.count(idx) != 0, "missing input ciphertext");
    return enc_data.at(idx);
  }
};

struct EncryptedPredictions {
  // per output big index: one sample packing the NUM_SAMPLES scores
  std::unordered_map<FeatBigIndex, TLweSample *> score;
  // Dense storage: when set, output big index o lives in slot o of `block`.
  std::shared_ptr<SampleBlock> block;
  std::vector<std::shared_ptr<SampleBlock>> extra;

  // creates an (uninitialized) sample at bidx; a second create of the same
  // bidx is an error (eval/idash.h:174-183)
  TLweSample *createAndGet(FeatBigIndex bidx, const TLweParams *) {
    if (score.count(bidx) != 0) DIE_DRAMATICALLY("output ciphertext created twice");
    extra.push_back(std::make_shared<SampleBlock>(1));
    TLweSample *s = extra.back()->sample(0);
    score.emplace(bidx, s);
    return s;
  }

  TLweSample *get(FeatBigIndex bidx, const TLweParams *) { return score.at(bidx); }
};

struct DecryptedPredictions {
  // per target position: the three class scores of every sample
  std::unordered_map<uint64_t, std::array<std::vector<float>, 3>> score;

  // prints the entries where the two predictions differ (eval/idash.h:194-218)
  static bool testEquals(const DecryptedPredictions &val1, const DecryptedPredictions &val2,
                         const IdashParams &params);
};

void write_params(const IdashParams &params, const std::string &filename);
void read_params(IdashParams &params, const std::string &filename);
void write_key(const IdashKey &key, const std::string &filename);
void read_key(IdashKey &key, const std::string &filename);
void read_model(Model &model, const IdashParams &params, const std::string &path);
void read_plaintext_data(PlaintextData &plaintext_data, const IdashParams &params,
                         const std::string &filename);
void write_encrypted_data(const EncryptedData &encrypted_data, const IdashParams &params,
                          const std::string &filename);
void read_encrypted_data(EncryptedData &encrypted_data, const IdashParams &params,
                         const std::string &filename);
void write_encrypted_predictions(const EncryptedPredictions &encrypted_preds,
                                 const IdashParams &params, const std::string &filename);
void read_encrypted_predictions(EncryptedPredictions &encrypted_preds, const IdashParams &params,
                                const std::string &filename);
void write_decrypted_predictions(const DecryptedPredictions &predictions, const IdashParams &params,
                                 const std::string &filename, const bool PRINT_POS_NAME = true);

void encrypt_data_ph1(EncryptedData &enc_data, const PlaintextData &plain_data, const IdashKey &key);
void encrypt_data_ph2(EncryptedData &enc_data, const PlaintextData &plain_data, const IdashKey &key);

void cloud_compute_score(EncryptedPredictions &enc_preds, const EncryptedData &enc_data,
                         const Model &model, const IdashParams &params);

void decrypt_predictions(DecryptedPredictions &predictions, const EncryptedPredictions &enc_preds,
                         const IdashKey &key);

typedef std::map<FeatBigIndex, std::vector<double>> PlaintextOnehot;
PlaintextOnehot compute_plaintext_onehot(const PlaintextData &X, const IdashParams &params);
void compute_score(DecryptedPredictions &predictions, const PlaintextData &X, const Model &M,
                   const IdashParams &params);

IdashParams *keygen_ph1(const std::string &targetFile, const std::string &challengeFile,
                        const bool targetFilePosOnly);
IdashKey *keygen_ph2(IdashParams *idashParams);

// Times a scope from construction (eval/idash.h:267-284).
class Profiler {
 public:
  Profiler();
  static double universalWallTime();
  static double universalClockTime();
  double walltime() const;
  double clocktime() const;
  long int maxrss() const;  // bytes

 private:
  const double tw0;
  const double tc0;
};

#endif  // IDASH_2019_IDASH_H
