// Context, device memory, errors and timing of libidash_hip.so.
#include <algorithm>
#include <cmath>
#include <cstring>

#include <vector>

#include "ihe_internal.h"

namespace ihe {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return IHE_ERR_HIP;
}

int check_ctx(ihe_ctx *ctx) {
  if (!ctx) return fail(IHE_ERR_ARG, "null context");
  return IHE_OK;
}

int make_current(ihe_ctx *ctx) {
  IHE_HIP(hipSetDevice(ctx->device));
  return IHE_OK;
}

// P(|trunc(128 g)| <= k) = 1 - erfc((k+1) / (128 sqrt 2)), scaled by 2^63.
// gaussian32(0, alpha) in TFHE is dtot32(alpha * g); at alpha = 2^-25 that
// is trunc(128 g) torus units (eval/idash.cpp:20-21, :625).
void build_noise_table(uint64_t *table, uint32_t len) {
  const double two63 = 9223372036854775808.0;
  for (uint32_t k = 0; k < len; ++k) {
    const double tail = std::erfc((double)(k + 1) / (128.0 * std::sqrt(2.0))) * two63;
    uint64_t t = (uint64_t)1 << 63;
    if (tail >= 1.0) t -= (uint64_t)tail;
    table[k] = t;
  }
  table[len - 1] = (uint64_t)1 << 63;
}

void build_noise_lut(const uint64_t *table, uint32_t len, uint32_t *lut) {
  const uint32_t shift = 63 - kNoiseLutBits;
  auto count = [&](uint64_t m) {  // entries <= m among table[0 .. len - 2]
    return (uint32_t)(std::upper_bound(table, table + len - 1, m) - table);
  };
  for (uint32_t b = 0; b < (1u << kNoiseLutBits); ++b) {
    const uint32_t lo = count((uint64_t)b << shift);
    const uint32_t hi = count((((uint64_t)b + 1) << shift) - 1);
    lut[b] = lo | (hi - lo) << 16;
  }
}

ScopedTimer::ScopedTimer(ihe_ctx *c, int s) : ctx(c), slot(s) {
  if (!ctx->timing) return;
  TimerSlot &t = ctx->timers[slot];
  if (t.used == t.starts.size()) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return;
    if (hipEventCreate(&b) != hipSuccess) {
      (void)hipEventDestroy(a);
      return;
    }
    t.starts.push_back(a);
    t.stops.push_back(b);
  }
  (void)hipEventRecord(t.starts[t.used], ctx->stream);
  stop = t.stops[t.used];
  t.used++;
}

ScopedTimer::~ScopedTimer() {
  if (stop) (void)hipEventRecord(stop, ctx->stream);
}

}  // namespace ihe

using namespace ihe;

extern "C" {

const char *ihe_last_error(void) { return g_last_error.c_str(); }

int ihe_device_count(int *count) {
  if (!count) return fail(IHE_ERR_ARG, "null count");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return IHE_OK;
}

int ihe_ctx_create(int device, ihe_ctx **out) {
  if (!out) return fail(IHE_ERR_ARG, "null output");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    return fail(IHE_ERR_HIP, "no HIP device visible");
  if (device < 0 || device >= n) return fail(IHE_ERR_ARG, "device index out of range");
  IHE_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  IHE_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(IHE_ERR_UNSUPPORTED, std::string("kernels are built for gfx950, device is ") +
                                         prop.gcnArchName);
  ihe_ctx *c = new ihe_ctx();
  c->device = device;
  c->num_cus = prop.multiProcessorCount;
  hipError_t e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "hipStreamCreateWithFlags");
  }
  c->stream = c->own_stream;
  // the CDF table, its bucket index and the ring kernels' work counters (zero)
  // in one allocation
  std::vector<uint64_t> table(IHE_NOISE_TABLE_LEN + (1u << kNoiseLutBits) / 2 + kRingQueueWords / 2, 0);
  build_noise_table(table.data(), IHE_NOISE_TABLE_LEN);
  build_noise_lut(table.data(), IHE_NOISE_TABLE_LEN,
                  reinterpret_cast<uint32_t *>(table.data() + IHE_NOISE_TABLE_LEN));
  e = hipMalloc(&c->d_noise_table, table.size() * sizeof(uint64_t));
  if (e == hipSuccess)
    e = hipMemcpy(c->d_noise_table, table.data(), table.size() * sizeof(uint64_t),
                  hipMemcpyHostToDevice);
  c->d_noise_lut = reinterpret_cast<uint32_t *>(c->d_noise_table + IHE_NOISE_TABLE_LEN);
  c->d_ring_queue = c->d_noise_lut + (1u << kNoiseLutBits);
  if (e != hipSuccess) {
    if (c->d_noise_table) (void)hipFree(c->d_noise_table);
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return hip_fail(e, "noise table upload");
  }
  // the device address of ihe_decrypt_noalloc's work counters, looked up here
  // so that its calls (also under graph capture) make none
  if (int rc = noalloc_setup(device)) {
    (void)hipFree(c->d_noise_table);
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return rc;
  }
  *out = c;
  return IHE_OK;
}

int ihe_ctx_destroy(ihe_ctx *ctx) {
  if (!ctx) return IHE_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &t : ctx->timers) {
    for (auto ev : t.starts) (void)hipEventDestroy(ev);
    for (auto ev : t.stops) (void)hipEventDestroy(ev);
  }
  if (ctx->d_noise_table) (void)hipFree(ctx->d_noise_table);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  delete ctx;
  return IHE_OK;
}

int ihe_ctx_set_stream(ihe_ctx *ctx, void *hip_stream) {
  if (int rc = check_ctx(ctx)) return rc;
  // the ring kernels' work counters serve one launch at a time: let the
  // old stream's work end before launches go to the new one
  IHE_HIP(hipSetDevice(ctx->device));
  IHE_HIP(hipStreamSynchronize(ctx->stream));
  ctx->stream = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;
  return IHE_OK;
}

int ihe_ctx_get_stream(ihe_ctx *ctx, void **hip_stream) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!hip_stream) return fail(IHE_ERR_ARG, "null output");
  *hip_stream = ctx->stream;
  return IHE_OK;
}

int ihe_sync(ihe_ctx *ctx) {
  if (int rc = check_ctx(ctx)) return rc;
  IHE_HIP(hipSetDevice(ctx->device));
  IHE_HIP(hipStreamSynchronize(ctx->stream));
  return IHE_OK;
}

int ihe_malloc(ihe_ctx *ctx, size_t bytes, void **dptr) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!dptr) return fail(IHE_ERR_ARG, "null output");
  *dptr = nullptr;
  if (bytes == 0) return IHE_OK;
  IHE_HIP(hipSetDevice(ctx->device));
  hipError_t e = hipMalloc(dptr, bytes);
  if (e != hipSuccess) {
    *dptr = nullptr;
    (void)hipGetLastError();
    return fail(IHE_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  return IHE_OK;
}

int ihe_free(ihe_ctx *ctx, void *dptr) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!dptr) return IHE_OK;
  IHE_HIP(hipSetDevice(ctx->device));
  IHE_HIP(hipFree(dptr));
  return IHE_OK;
}

int ihe_host_alloc(ihe_ctx *ctx, size_t bytes, void **hptr) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!hptr) return fail(IHE_ERR_ARG, "null output");
  *hptr = nullptr;
  if (bytes == 0) return IHE_OK;
  IHE_HIP(hipSetDevice(ctx->device));
  hipError_t e = hipHostMalloc(hptr, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    *hptr = nullptr;
    return fail(IHE_ERR_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
  }
  return IHE_OK;
}

int ihe_host_free(ihe_ctx *ctx, void *hptr) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!hptr) return IHE_OK;
  IHE_HIP(hipHostFree(hptr));
  return IHE_OK;
}

static int copy(ihe_ctx *ctx, void *dst, const void *src, size_t bytes, hipMemcpyKind kind,
                bool blocking) {
  if (int rc = check_ctx(ctx)) return rc;
  if (bytes == 0) return IHE_OK;
  if (!dst || !src) return fail(IHE_ERR_ARG, "null copy pointer");
  IHE_HIP(hipSetDevice(ctx->device));
  IHE_HIP(hipMemcpyAsync(dst, src, bytes, kind, ctx->stream));
  if (blocking) IHE_HIP(hipStreamSynchronize(ctx->stream));
  return IHE_OK;
}

int ihe_memcpy_h2d(ihe_ctx *ctx, void *dst, const void *src, size_t bytes) {
  return copy(ctx, dst, src, bytes, hipMemcpyHostToDevice, true);
}
int ihe_memcpy_d2h(ihe_ctx *ctx, void *dst, const void *src, size_t bytes) {
  return copy(ctx, dst, src, bytes, hipMemcpyDeviceToHost, true);
}
int ihe_memcpy_h2d_async(ihe_ctx *ctx, void *dst, const void *src, size_t bytes) {
  return copy(ctx, dst, src, bytes, hipMemcpyHostToDevice, false);
}
int ihe_memcpy_d2h_async(ihe_ctx *ctx, void *dst, const void *src, size_t bytes) {
  return copy(ctx, dst, src, bytes, hipMemcpyDeviceToHost, false);
}

int ihe_memset(ihe_ctx *ctx, void *dst, int value, size_t bytes) {
  if (int rc = check_ctx(ctx)) return rc;
  if (bytes == 0) return IHE_OK;
  if (!dst) return fail(IHE_ERR_ARG, "null pointer");
  IHE_HIP(hipSetDevice(ctx->device));
  IHE_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
  return IHE_OK;
}

}  // extern "C"

namespace ihe {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_copy16(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst,
                                               size_t n16) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
    __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// Checksum partials: word i of the buffer contributes mix((i << 32) ^ w_i), mix
// being the output mix of splitmix64, summed mod 2^64 per block; the host adds
// the block partials.
constexpr uint32_t kSumBlocks = 1024;

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ __launch_bounds__(256) void k_checksum(const uint32_t *__restrict__ p, size_t n,
                                                  uint64_t *__restrict__ partial) {
  __shared__ uint64_t red[256];
  uint64_t acc = 0;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    acc += mix64(((uint64_t)i << 32) ^ p[i]);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Each of num_waves waves writes its share of the output in 1 KiB stores
// (at least one), with its share of the input read in 1 KiB loads spread
// evenly between them. Cache policies: LP on the loads, SP on the stores
// (buffer-instruction bits: sc0 1, nt 2, sc1 16). ihe_stream_probe uses one
// wave per 64 KiB of output, default loads and nt | sc1 (18) stores, the
// fastest documented policy for the cloud step's mix (the cloud kernels' too).
template <int LP, int SP>
__global__ __launch_bounds__(256) void k_stream_probe(const int *in, uint32_t in_bytes, int *out,
                                                      uint32_t out_bytes, uint32_t num_waves) {
  typedef int int4v __attribute__((ext_vector_type(4)));
  const uint32_t wave = blockIdx.x * 4 + threadIdx.x / 64, lane = threadIdx.x % 64;
  if (wave >= num_waves) return;
  const uint32_t in_kib = in_bytes / 1024, out_kib = out_bytes / 1024;
  const uint32_t r0 = (uint32_t)((uint64_t)in_kib * wave / num_waves);
  const uint32_t r1 = (uint32_t)((uint64_t)in_kib * (wave + 1) / num_waves);
  const uint32_t o0 = (uint32_t)((uint64_t)out_kib * wave / num_waves);
  const uint32_t o1 = (uint32_t)((uint64_t)out_kib * (wave + 1) / num_waves);
  const __amdgpu_buffer_rsrc_t ri =
      __builtin_amdgcn_make_buffer_rsrc((void *)in, (short)0, (int)in_bytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t ro =
      __builtin_amdgcn_make_buffer_rsrc((void *)out, (short)0, (int)out_bytes, 0x00020000);
  int4v acc = {(int)lane, 0, 0, 0};
  uint32_t r = r0;
  for (uint32_t o = o0; o < o1; ++o) {
    const uint32_t until = r0 + (uint32_t)((uint64_t)(r1 - r0) * (o - o0 + 1) / (o1 - o0));
    for (; r < until; ++r) acc += __builtin_amdgcn_raw_buffer_load_b128(ri, r * 1024u + lane * 16u, 0, LP);
    __builtin_amdgcn_raw_buffer_store_b128(acc, ro, o * 1024u + lane * 16u, 0, SP);
  }
}

template <int LP, int SP>
void launch_stream(hipStream_t s, const void *in, uint32_t in_bytes, void *out, uint32_t out_bytes,
                   uint32_t waves) {
  hipLaunchKernelGGL((k_stream_probe<LP, SP>), dim3((waves + 3) / 4), dim3(256), 0, s,
                     (const int *)in, in_bytes, (int *)out, out_bytes, waves);
}

}  // namespace ihe

extern "C" {

int ihe_stream_probe_ex(ihe_ctx *ctx, const void *d_in, size_t in_bytes, void *d_out,
                        size_t out_bytes, uint32_t num_waves, uint32_t load_policy,
                        uint32_t store_policy) {
  if (int rc = check_ctx(ctx)) return rc;
  if ((in_bytes && !d_in) || !d_out || in_bytes % 1024 || out_bytes % 1024 ||
      in_bytes >= (1ull << 31) || out_bytes >= (1ull << 31) ||
      ((uintptr_t)d_in | (uintptr_t)d_out) % 16 || num_waves == 0 || out_bytes / 1024 < num_waves)
    return fail(IHE_ERR_ARG, "stream probe: sizes must be multiples of 1 KiB below 2^31, at "
                             "least 1 KiB written per wave, buffers 16-byte aligned");
  IHE_HIP(hipSetDevice(ctx->device));
  const hipStream_t s = ctx->stream;
  const uint32_t ib = (uint32_t)in_bytes, ob = (uint32_t)out_bytes, w = num_waves;
  ScopedTimer t(ctx, IHE_TIMER_STREAM);
  switch (load_policy * 32 + store_policy) {
    case 0 * 32 + 0: launch_stream<0, 0>(s, d_in, ib, d_out, ob, w); break;
    case 0 * 32 + 1: launch_stream<0, 1>(s, d_in, ib, d_out, ob, w); break;
    case 0 * 32 + 2: launch_stream<0, 2>(s, d_in, ib, d_out, ob, w); break;
    case 0 * 32 + 16: launch_stream<0, 16>(s, d_in, ib, d_out, ob, w); break;
    case 0 * 32 + 18: launch_stream<0, 18>(s, d_in, ib, d_out, ob, w); break;
    case 2 * 32 + 2: launch_stream<2, 2>(s, d_in, ib, d_out, ob, w); break;
    case 2 * 32 + 18: launch_stream<2, 18>(s, d_in, ib, d_out, ob, w); break;
    default:
      return fail(IHE_ERR_ARG, "stream probe: policies (loads, stores) must be one of (0, 0), "
                               "(0, 1), (0, 2), (0, 16), (0, 18), (2, 2), (2, 18)");
  }
  IHE_HIP(hipGetLastError());
  return IHE_OK;
}

int ihe_stream_probe(ihe_ctx *ctx, const void *d_in, size_t in_bytes, void *d_out,
                     size_t out_bytes) {
  if (out_bytes < 65536)
    return fail(IHE_ERR_ARG, "stream probe: sizes must be multiples of 1 KiB below 2^31 "
                             "(at least 64 KiB written), buffers 16-byte aligned");
  return ihe_stream_probe_ex(ctx, d_in, in_bytes, d_out, out_bytes, (uint32_t)(out_bytes / 65536),
                             0, 18);
}

int ihe_checksum(ihe_ctx *ctx, const void *d_ptr, size_t bytes, uint64_t *out) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!out || (bytes && !d_ptr) || bytes % 4) return fail(IHE_ERR_ARG, "bad checksum arguments");
  *out = 0;
  if (bytes == 0) return IHE_OK;
  IHE_HIP(hipSetDevice(ctx->device));
  uint64_t *d_part = nullptr;
  IHE_HIP(hipMalloc(&d_part, kSumBlocks * sizeof(uint64_t)));
  hipLaunchKernelGGL(k_checksum, dim3(kSumBlocks), dim3(256), 0, ctx->stream,
                     (const uint32_t *)d_ptr, bytes / 4, d_part);
  std::vector<uint64_t> h(kSumBlocks);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), d_part, kSumBlocks * sizeof(uint64_t), hipMemcpyDeviceToHost,
                       ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d_part);
  if (e != hipSuccess) return fail(IHE_ERR_HIP, std::string("checksum: ") + hipGetErrorString(e));
  for (uint64_t v : h) *out += v;
  return IHE_OK;
}

int ihe_memcpy_d2d(ihe_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (int rc = check_ctx(ctx)) return rc;
  if (bytes == 0) return IHE_OK;
  if (!dst || !src) return fail(IHE_ERR_ARG, "null copy pointer");
  IHE_HIP(hipSetDevice(ctx->device));
  if ((bytes | (uintptr_t)dst | (uintptr_t)src) % 16 != 0) {
    IHE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return IHE_OK;
  }
  const size_t n16 = bytes / 16;
  const size_t blocks = std::min<size_t>((n16 + 255) / 256, (size_t)ctx->num_cus * 32);
  hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                     (const u32x4 *)src, (u32x4 *)dst, n16);
  IHE_HIP(hipGetLastError());
  return IHE_OK;
}

int ihe_timing_enable(ihe_ctx *ctx, int on) {
  if (int rc = check_ctx(ctx)) return rc;
  ctx->timing = on != 0;
  return IHE_OK;
}

int ihe_timing_region(ihe_ctx *ctx, int begin) {
  if (int rc = check_ctx(ctx)) return rc;
  IHE_HIP(hipSetDevice(ctx->device));
  TimerSlot &t = ctx->timers[IHE_TIMER_REGION];
  if (begin) {
    if (t.used == t.starts.size()) {
      hipEvent_t a, b;
      IHE_HIP(hipEventCreate(&a));
      if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return fail(IHE_ERR_HIP, "event creation failed");
      }
      t.starts.push_back(a);
      t.stops.push_back(b);
    }
    IHE_HIP(hipEventRecord(t.starts[t.used], ctx->stream));
    t.used++;
    return IHE_OK;
  }
  if (t.used == 0) return fail(IHE_ERR_ARG, "region end without a begin");
  IHE_HIP(hipEventRecord(t.stops[t.used - 1], ctx->stream));
  return IHE_OK;
}

int ihe_timing_reset(ihe_ctx *ctx) {
  if (int rc = check_ctx(ctx)) return rc;
  IHE_HIP(hipSetDevice(ctx->device));
  IHE_HIP(hipStreamSynchronize(ctx->stream));
  for (auto &t : ctx->timers) t.used = 0;
  return IHE_OK;
}

int ihe_timing_read(ihe_ctx *ctx, int timer, double *total_ms, uint32_t *count) {
  if (int rc = check_ctx(ctx)) return rc;
  if (timer < 0 || timer >= IHE_TIMER_COUNT || !total_ms || !count)
    return fail(IHE_ERR_ARG, "bad timer query");
  IHE_HIP(hipSetDevice(ctx->device));
  TimerSlot &t = ctx->timers[timer];
  double sum = 0.0;
  for (uint32_t i = 0; i < t.used; ++i) {
    IHE_HIP(hipEventSynchronize(t.stops[i]));
    float ms = 0.f;
    IHE_HIP(hipEventElapsedTime(&ms, t.starts[i], t.stops[i]));
    sum += ms;
  }
  *total_ms = sum;
  *count = t.used;
  return IHE_OK;
}

int ihe_noise_table(uint64_t *table, uint32_t len) {
  if (!table || len != IHE_NOISE_TABLE_LEN) return fail(IHE_ERR_ARG, "bad noise table buffer");
  build_noise_table(table, len);
  return IHE_OK;
}

}  // extern "C"
