// cloud [model dir]: reads params.bin and encrypted_data.bin, writes
// encrypted_prediction.bin (eval/cloud.cpp:3-29).
#include "idash.h"

int main(int argc, char **argv) {
  Profiler profiler;
  const std::string model_dir = argc >= 2 ? argv[1] : MODEL_FILE;
  std::cout << "using model dir: " << model_dir << std::endl;

  IdashParams params;
  Model model;
  EncryptedData enc;
  EncryptedPredictions preds;
  read_params(params, PARAMS_FILE);
  read_model(model, params, model_dir);
  read_encrypted_data(enc, params, ENCRYPTED_DATA_FILE);
  const double t0 = profiler.walltime();
  cloud_compute_score(preds, enc, model, params);
  const double t1 = profiler.walltime();
  write_encrypted_predictions(preds, params, ENCRYPTED_PREDICTION_FILE);
  const double t2 = profiler.walltime();

  std::cout << "----------------- BENCHMARK ----------------- " << std::endl;
  std::cout << "fhe wall time (seconds)..........: " << t1 - t0 << std::endl;
  std::cout << "serialization wall time (seconds): " << t2 - t1 + t0 << std::endl;
  std::cout << "total wall time (seconds)........: " << t2 << std::endl;
  std::cout << "RAM usage (MB)...................: " << profiler.maxrss() / 1e6 << std::endl;
  return 0;
}
