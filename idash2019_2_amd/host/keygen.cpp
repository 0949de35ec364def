// keygen [target file] [tag file] [target file has positions only (0/1)]
// Writes params.bin and keys.bin (eval/keygen.cpp:3-32).
#include "idash.h"

int main(int argc, char **argv) {
  Profiler profiler;
  const std::string target = argc >= 2 ? argv[1] : TARGET_FILE;
  const std::string tags = argc >= 3 ? argv[2] : CHALLENGE_FILE;
  const bool pos_only = argc >= 4 ? (bool)atoi(argv[3]) : TARGET_FILE_POS_ONLY;
  std::cout << "using target file (headers): " << target << (pos_only ? " (only positions)" : "")
            << std::endl;
  std::cout << "using tag file (challenge): " << tags << std::endl;

  IdashParams *params = keygen_ph1(target, tags, pos_only);
  const double t0 = profiler.walltime();
  IdashKey *key = keygen_ph2(params);
  const double t1 = profiler.walltime();
  write_params(*key->idashParams, PARAMS_FILE);
  write_key(*key, KEYS_FILE);

  std::cout << "----------------- BENCHMARK ----------------- " << std::endl;
  std::cout << "Number of threads ...............: " << NB_THREADS << std::endl;
  std::cout << "Keygen time (seconds)............: " << t1 - t0 << std::endl;
  std::cout << "Total wall time (seconds)........: " << profiler.walltime() << std::endl;
  std::cout << "RAM usage (MB)...................: " << profiler.maxrss() / 1e6 << std::endl;
  return 0;
}
