// encrypt [tag file]: reads keys.bin, writes encrypted_data.bin
// (eval/encrypt.cpp:13-38).
#include "idash.h"

int main(int argc, char **argv) {
  Profiler profiler;
  const std::string tags = argc >= 2 ? argv[1] : CHALLENGE_FILE;
  std::cout << "using tag file (challenge): " << tags << std::endl;

  IdashKey key;
  PlaintextData plain;
  EncryptedData enc;
  read_key(key, KEYS_FILE);
  read_plaintext_data(plain, *key.idashParams, tags);
  const double t0 = profiler.walltime();
  encrypt_data_ph1(enc, plain, key);
  encrypt_data_ph2(enc, plain, key);
  const double t1 = profiler.walltime();
  write_encrypted_data(enc, *key.idashParams, ENCRYPTED_DATA_FILE);
  const double t2 = profiler.walltime();

  std::cout << "----------------- BENCHMARK ----------------- " << std::endl;
  std::cout << "encrypt wall time (seconds)......: " << t1 - t0 << std::endl;
  std::cout << "serialization wall time (seconds): " << t2 - t1 + t0 << std::endl;
  std::cout << "total wall time (seconds)........: " << t2 << std::endl;
  std::cout << "RAM usage (MB)...................: " << profiler.maxrss() / 1e6 << std::endl;
  return 0;
}
