// decrypt [any argument]: reads keys.bin and encrypted_prediction.bin, writes
// result.csv (target names), or result_bypos.csv (positions) when given an
// argument (eval/decrypt.cpp:13-49).
#include "idash.h"

int main(int argc, char **argv) {
  (void)argv;
  Profiler profiler;
  IdashKey key;
  EncryptedPredictions enc;
  DecryptedPredictions dec;
  read_key(key, KEYS_FILE);
  read_encrypted_predictions(enc, *key.idashParams, ENCRYPTED_PREDICTION_FILE);
  const double t0 = profiler.walltime();
  decrypt_predictions(dec, enc, key);
  const double t1 = profiler.walltime();
  if (argc == 1)
    write_decrypted_predictions(dec, *key.idashParams, RESULT_FILE, true);
  else
    write_decrypted_predictions(dec, *key.idashParams, RESULT_BYPOS_FILE, false);
  const double t2 = profiler.walltime();

  std::cout << "----------------- BENCHMARK ----------------- " << std::endl;
  std::cout << "decrypt wall time (seconds)......: " << t1 - t0 << std::endl;
  std::cout << "serialization wall time (seconds): " << t2 - t1 + t0 << std::endl;
  std::cout << "total wall time (seconds)........: " << t2 << std::endl;
  std::cout << "RAM usage (MB)...................: " << profiler.maxrss() / 1e6 << std::endl;
  return 0;
}
