// Internal declarations shared by the HIP sources of libidash_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "idash_hip.h"

namespace ihe {

constexpr uint32_t kN = 1024;         // TRLWE dimension (eval/idash.h:49)
constexpr uint32_t kCtWords = 2 * kN; // one TLWE sample: a then b
constexpr uint32_t kOne = 1u << 18;   // ONE_IN_T32 = dtot32(1/16384) (eval/idash.cpp:39)
constexpr uint32_t kNan[3] = {131072u, 87381u, 43690u};  // NAN_{0,1,2}_IN_T32 (:40-45)

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define IHE_HIP(call)                                   \
  do {                                                  \
    hipError_t _e = (call);                             \
    if (_e != hipSuccess) return ::ihe::hip_fail(_e, #call); \
  } while (0)

struct TimerSlot {
  std::vector<hipEvent_t> starts, stops;  // pool; [0, used) recorded since reset
  uint32_t used = 0;
};

// Binary key as a bit mask; passed to kernels by value.
struct KeyBits {
  uint32_t w[kN / 32];
};

}  // namespace ihe

struct ihe_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  bool timing = false;
  ihe::TimerSlot timers[IHE_TIMER_COUNT];
  uint64_t *d_noise_table = nullptr;  // encrypt noise CDF, uploaded at context creation
  uint32_t *d_noise_lut = nullptr;    // its bucket index (same allocation, after the table)
  uint32_t *d_ring_queue = nullptr;   // ring kernels' work counters (same allocation, after the index)
};

struct ihe_key {
  ihe_ctx *ctx = nullptr;
  ihe::KeyBits bits;
};

namespace ihe {

// Brackets one launch with events when timing is enabled.
struct ScopedTimer {
  ihe_ctx *ctx;
  int slot;
  hipEvent_t stop = nullptr;
  ScopedTimer(ihe_ctx *c, int s);
  ~ScopedTimer();
};

int check_ctx(ihe_ctx *ctx);
int make_current(ihe_ctx *ctx);
void build_noise_table(uint64_t *table, uint32_t len);
// Bucket index of the CDF table: for m in [b 2^51, (b + 1) 2^51), the count of
// table entries <= m (among the first len - 1) lies in [lo, lo + span],
// lut[b] = lo | span << 16.
constexpr uint32_t kNoiseLutBits = 12;
void build_noise_lut(const uint64_t *table, uint32_t len, uint32_t *lut);
// Work counters of the persistent ring kernels: one per XCD, then a count of
// finished workgroups (the last one clears all of them for the next launch).
constexpr uint32_t kRingQueueWords = 16;
// Looks up (once per device, that device current) the device address of the
// work counters ihe_decrypt_noalloc uses; ihe_ctx_create calls it.
int noalloc_setup(int device);

}  // namespace ihe
