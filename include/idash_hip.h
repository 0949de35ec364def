/*
 * idash_hip.h: C ABI of the MI355X (gfx950) secure-imputation engine.
 *
 * The reference (ssmiler/idash2019_2) has no FFI: its hot path is the C++
 * free-function API of eval/idash.h:221-264 on top of the TFHE library. This
 * ABI is what the build's idash-compatible C++ host (idash2019_2_amd/host)
 * binds in place of those function bodies; each entry point below names the
 * reference function it replaces. Plain pointers and sizes only, int status
 * codes, no exceptions across the boundary.
 *
 * Ciphertext layout in device memory (HBM): one TLWE sample (k = 1, N = 1024)
 * is 2048 consecutive int32 words, a[0..N) then b[0..N); arrays of samples
 * are dense, indexed by FeatIndex (inputs) or by output big index (outputs).
 * Variances are separate double arrays with one entry per sample.
 *
 * Threading: a context owns one HIP stream (default priority). Calls on one
 * context must not overlap from several host threads. Device pointers passed
 * in must come from ihe_malloc (or any hipMalloc on the context's device).
 */
#ifndef IDASH_HIP_H
#define IDASH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  IHE_OK = 0,
  IHE_ERR_ARG = 1,         /* invalid argument (null pointer, bad size, bad layout) */
  IHE_ERR_HIP = 2,         /* a HIP runtime call failed */
  IHE_ERR_RANGE = 3,       /* model references a ciphertext that is not present */
  IHE_ERR_NOMEM = 4,       /* device or host allocation failed / over budget */
  IHE_ERR_UNSUPPORTED = 5, /* shape the kernels do not handle */
  IHE_ERR_BUSY = 6         /* serving queue full; band still held by a batch or server */
};

#define IHE_N 1024u                     /* TRLWE dimension, eval/idash.h:49 */
#define IHE_CONSTANT_BIDX 0xFFFFFFFFu   /* constant_bigIndex(), eval/idash.h:93 */

typedef struct ihe_ctx ihe_ctx;
typedef struct ihe_band ihe_band;
typedef struct ihe_key ihe_key;

/* Packing layout of one cohort (IdashParams fields, eval/idash.h:72-79;
 * derived at eval/idash.cpp:409-412: R = N / NS, RS = N / R). */
typedef struct {
  uint32_t num_samples; /* NS, 1..1024 */
  uint32_t num_regions; /* R */
  uint32_t region_size; /* RS */
} ihe_layout;

/* Model in CSR form over output slots: the reference Model
 * (eval/idash.h:129-134) flattened. Row o holds the (in big index, weight)
 * pairs of output slot o in the order the caller wants them visited; the
 * entry with in_bidx == IHE_CONSTANT_BIDX is the bias. No duplicate in_bidx
 * within a row. */
typedef struct {
  uint32_t num_outputs;
  const uint32_t *row_ptr; /* [num_outputs + 1] */
  const uint32_t *in_bidx; /* [nnz] */
  const int32_t *weight;   /* [nnz] */
} ihe_model_csr;

/* Cloud kernel variants. */
enum {
  IHE_VARIANT_AUTO = 0,  /* MFMA when the weights fit its byte split, else VALU or WIDE */
  IHE_VARIANT_VALU = 1,  /* int16-pair dot2 VALU kernel (weights within int16) */
  IHE_VARIANT_WIDE = 2,  /* int32 weights, 32-bit multiply VALU kernel */
  IHE_VARIANT_MFMA = 3   /* byte-split int8 MFMA kernel (weights within int16) */
};

typedef struct {
  uint32_t num_outputs, num_groups, num_tiles, num_cols, nnz;
  uint32_t variant;        /* the variant the band was compiled for */
  uint32_t max_window;     /* widest column list of a group */
  uint64_t device_bytes;   /* device memory the band holds */
  /* MFMA bands (0 otherwise): output rows per tile (32, 64 or 128), K-blocks of
     32 columns per tile, and 32-word column tiles per wave (rows <= 64) or per
     workgroup (128 rows) */
  uint32_t tile_rows, tile_kblocks, col_tiles;
} ihe_band_info;

/* ---------------- errors, devices, context ---------------- */
const char *ihe_last_error(void); /* thread-local message of the last failure */
int ihe_device_count(int *count);
int ihe_ctx_create(int device, ihe_ctx **out);
int ihe_ctx_destroy(ihe_ctx *ctx);
/* NULL: back to the context's own. Waits for the current stream's work first:
 * a context's launches run one at a time. */
int ihe_ctx_set_stream(ihe_ctx *ctx, void *hip_stream);
int ihe_ctx_get_stream(ihe_ctx *ctx, void **hip_stream);
int ihe_sync(ihe_ctx *ctx);

/* ---------------- device memory (plain HIP wrappers) ---------------- */
int ihe_malloc(ihe_ctx *ctx, size_t bytes, void **dptr);
int ihe_free(ihe_ctx *ctx, void *dptr);
int ihe_host_alloc(ihe_ctx *ctx, size_t bytes, void **hptr); /* pinned */
int ihe_host_free(ihe_ctx *ctx, void *hptr);
int ihe_memcpy_h2d(ihe_ctx *ctx, void *dst, const void *src, size_t bytes); /* stream-ordered, blocking */
int ihe_memcpy_d2h(ihe_ctx *ctx, void *dst, const void *src, size_t bytes);
int ihe_memcpy_h2d_async(ihe_ctx *ctx, void *dst, const void *src, size_t bytes);
int ihe_memcpy_d2h_async(ihe_ctx *ctx, void *dst, const void *src, size_t bytes);
int ihe_memset(ihe_ctx *ctx, void *dst, int value, size_t bytes);
/* Device-to-device copy by a streaming kernel (16-byte aligned buffers and
 * sizes; otherwise hipMemcpyAsync). Also the reference copy of known size
 * that HBM traffic counters are calibrated against. */
int ihe_memcpy_d2d(ihe_ctx *ctx, void *dst, const void *src, size_t bytes);

/* Order-sensitive 64-bit checksum of a device buffer (bytes % 4 == 0):
 * sum over words i of mix((i << 32) ^ w_i) mod 2^64, mix being splitmix64's
 * output mix. Compares large
 * outputs of two runs without copying them to the host. Synchronizes. */
int ihe_checksum(ihe_ctx *ctx, const void *d_ptr, size_t bytes, uint64_t *out);

/* HBM streaming probe: reads in_bytes from d_in and writes out_bytes to d_out
 * (both multiples of 1024 and below 2^31, d_in and d_out 16-byte aligned),
 * contiguous 16-byte accesses, reads and writes interleaved in every wave in
 * the ratio of the two sizes. Timed under IHE_TIMER_STREAM, it measures the
 * bandwidth the device sustains for a read:write mix, such as the cloud
 * step's (input ciphertexts read once, outputs written once). The written
 * bytes are unspecified. Asynchronous. */
int ihe_stream_probe(ihe_ctx *ctx, const void *d_in, size_t in_bytes, void *d_out,
                     size_t out_bytes);
/* The same probe with num_waves waves (each writing at least 1 KiB: out_bytes
 * >= 1024 * num_waves; in_bytes may be 0 for writes alone, and out_bytes as
 * small as that for reads alone) and the cache policy bits of its buffer
 * loads and stores (sc0 = 1, nt = 2, sc1 = 16): (loads, stores) one of (0, 0),
 * (0, 1), (0, 2), (0, 16), (0, 18), (2, 2), (2, 18). ihe_stream_probe is
 * num_waves = out_bytes / 65536 with (0, 18). */
int ihe_stream_probe_ex(ihe_ctx *ctx, const void *d_in, size_t in_bytes, void *d_out,
                        size_t out_bytes, uint32_t num_waves, uint32_t load_policy,
                        uint32_t store_policy);

/* ---------------- timing of the library's own kernels ---------------- */
enum {
  IHE_TIMER_CLOUD = 0,     /* main cloud kernel */
  IHE_TIMER_CLOUD_VAR = 1, /* cloud variance kernel */
  IHE_TIMER_ENCRYPT = 2,
  IHE_TIMER_DECRYPT = 3,
  IHE_TIMER_BATCH = 4,
  IHE_TIMER_STREAM = 5,    /* ihe_stream_probe */
  IHE_TIMER_REGION = 6,    /* caller-marked regions (ihe_timing_region) */
  IHE_TIMER_COUNT = 7
};
/* When enabled, each launch is bracketed by HIP events on the launch stream. */
int ihe_timing_enable(ihe_ctx *ctx, int on);
/* Records the start (begin != 0) or the end (begin == 0) of a region of the
 * context stream's work under IHE_TIMER_REGION, whether or not per-launch
 * timing is enabled: a timed loop bracketed by two events, with nothing
 * recorded between its launches. */
int ihe_timing_region(ihe_ctx *ctx, int begin);
int ihe_timing_reset(ihe_ctx *ctx);
/* Synchronizes, then returns the summed duration and launch count since reset. */
int ihe_timing_read(ihe_ctx *ctx, int timer, double *total_ms, uint32_t *count);

/* ---------------- cloud step ---------------- */
/* Compiles a model into the device band format and uploads it (the model half
 * of cloud_compute_score, eval/idash.cpp:772-777 and :800-819). num_in_cts is
 * the number of input ciphertexts the step will read (dense FeatIndex range).
 * Fails with IHE_ERR_RANGE when a coefficient's FeatIndex >= num_in_cts. */
int ihe_band_create(ihe_ctx *ctx, const ihe_layout *layout, uint32_t num_in_cts,
                    const ihe_model_csr *model, int variant, ihe_band **out);
/* Autotuner: compiles the model once per feasible MFMA tile shape (output rows
 * per tile, i.e. targets per workgroup; K-blocks of 32 columns, the window
 * length a tile covers; 32-word column tiles per wave, the coefficient chunk),
 * times `reps` cloud launches of each on the caller's cohort buffers (d_out is
 * overwritten) and returns the fastest band in *out. results[0 .. *num_results)
 * receive the timings (*num_results: capacity in, count out). Needs weights
 * within the MFMA split; IHE_ERR_UNSUPPORTED otherwise. */
typedef struct {
  uint32_t rows_per_tile, kblocks, col_tiles_per_wave;
  float ms;
} ihe_tune_entry;
int ihe_band_create_tuned(ihe_ctx *ctx, const ihe_layout *layout, uint32_t num_in_cts,
                          const ihe_model_csr *model, const int32_t *d_in, int32_t *d_out,
                          uint32_t reps, ihe_band **out, ihe_tune_entry *results,
                          uint32_t *num_results);
/* IHE_ERR_BUSY while a batch or server holds the band. */
int ihe_band_destroy(ihe_band *band);
int ihe_band_get_info(const ihe_band *band, ihe_band_info *info);
/* Runs the band compiler on the host only (no device, no upload) and reports
 * what ihe_band_create would build, or the error it would return. */
int ihe_band_plan(const ihe_layout *layout, uint32_t num_in_cts, const ihe_model_csr *model,
                  int variant, ihe_band_info *info);

/* Replaces cloud_compute_score (eval/idash.cpp:763-848; declared at
 * eval/idash.h:251-252). d_in: [num_in_cts][2][N] int32, d_in_var: [num_in_cts]
 * doubles; d_out: [num_outputs][2][N] int32, d_out_var: [num_outputs] doubles
 * (NULL skips the variance). Enqueued on the context stream. */
int ihe_cloud_score(ihe_ctx *ctx, const ihe_band *band, const int32_t *d_in,
                    const double *d_in_var, int32_t *d_out, double *d_out_var);

/* The same step for a cohort with another packing layout or ciphertext count
 * than the band was created with: MFMA bands hold their columns as big
 * indices, so one resident band serves cohorts of every sample count (the
 * reference re-reads the model per run, eval/cloud.cpp:14). VALU bands accept
 * only their own layout (IHE_ERR_UNSUPPORTED otherwise). Fails with
 * IHE_ERR_RANGE when a coefficient's FeatIndex (bidx / R) >= num_in_cts. */
int ihe_cloud_score_layout(ihe_ctx *ctx, const ihe_band *band, const ihe_layout *layout,
                           uint32_t num_in_cts, const int32_t *d_in, const double *d_in_var,
                           int32_t *d_out, double *d_out_var);

/* ---------------- batched cloud step ---------------- */
/* Many cohorts in one step, in place of one cloud_compute_score call per
 * cohort (eval/cloud.cpp:17): each cohort has its own packing layout, input
 * ciphertexts and resident MFMA band (bands may be shared). The work is one
 * persistent kernel whose workgroups take every cohort's tiles and variance
 * blocks from device counters (bands tiled 32 or 64 rows x 1 K-block, or 128
 * rows x up to 6 K-blocks); a batch with other tilings, or built with
 * IHE_BATCH_PERSISTENT=0, runs one launch per kernel family instead.
 * ihe_batch_capture records the work into a hipGraph so that ihe_batch_run is
 * a single graph launch. Results equal ihe_cloud_score_layout per cohort.
 * Runs of one batch are stream-ordered; two batches may run concurrently. */
typedef struct ihe_batch ihe_batch;
typedef struct {
  const ihe_band *band;    /* MFMA band */
  ihe_layout layout;       /* this cohort's packing */
  uint32_t num_in_cts;
  const int32_t *d_in;     /* [num_in_cts][2][N] */
  const double *d_in_var;  /* [num_in_cts]; may be NULL when d_out_var is */
  int32_t *d_out;          /* [band outputs][2][N] */
  double *d_out_var;       /* NULL: no variances for this cohort */
} ihe_cohort;
typedef struct {
  uint32_t num_cohorts, num_launches, num_tiles, captured;
  uint32_t persistent; /* 1: one persistent kernel runs every family and the variances */
} ihe_batch_info;
/* The batch keeps device pointers to the cohorts' buffers and bands: both must
 * outlive it (ihe_band_destroy fails with IHE_ERR_BUSY while a batch holds the
 * band; buffers are the caller's to keep). */
int ihe_batch_create(ihe_ctx *ctx, const ihe_cohort *cohorts, uint32_t num_cohorts,
                     ihe_batch **out);
int ihe_batch_capture(ihe_batch *batch);
int ihe_batch_run(ihe_batch *batch); /* enqueued on the context stream */
int ihe_batch_get_info(const ihe_batch *batch, ihe_batch_info *info);
int ihe_batch_destroy(ihe_batch *batch);

/* ---------------- serving ---------------- */
/* A request loop over resident model bands: each request is one cohort's
 * encrypt -> cloud -> decrypt (encrypt_data_ph1/ph2, cloud_compute_score,
 * decrypt_predictions; eval/idash.cpp:615-761) run on the device, from host
 * genotypes to host scores. Requests above 1024 samples run as chunks of at
 * most 1024, and the request's host buffers are chunk-major: chunk k (samples
 * 1024k .. 1024k + ns_k - 1) is a contiguous [num_tags][ns_k] block of
 * genotypes and a contiguous [num_outputs][ns_k] block of scores (for
 * NS <= 1024 simply [num_tags][NS] and [num_outputs][NS]). A worker thread admits chunks in arrival order whenever one of
 * num_slots slots (each with its own stream) and space in a device arena of
 * hbm_budget bytes are free, so requests that arrive while others run start
 * at once and overlap on the device. Results equal the per-request path
 * (ihe_encrypt with seed + chunk index, ihe_cloud_score_layout, ihe_decrypt). */
typedef struct ihe_server ihe_server;
typedef struct {
  uint64_t hbm_budget;     /* bytes of device memory for requests in flight */
  uint32_t num_slots;      /* chunks on the device at once, 1..64 */
  uint32_t queue_capacity; /* waiting requests before submit returns IHE_ERR_BUSY (0: 1024) */
} ihe_server_config;
typedef struct {
  uint32_t model;       /* from ihe_server_add_model */
  uint32_t num_samples; /* NS >= 1 */
  const int8_t *geno;   /* host, chunk-major [num_tags][ns_k] blocks; outside {0,1,2}: missing */
  const int32_t *key;   /* host [N] binary key */
  uint64_t seed;        /* chunk k encrypts with seed + k */
  float *out;           /* host, chunk-major [num_outputs][ns_k] blocks of decrypted scores */
  uint64_t arrival_ns;  /* scheduled arrival, CLOCK_MONOTONIC nanoseconds; 0: the submit time */
} ihe_request;
typedef struct {
  uint64_t completed;
  uint32_t pending, running;
  uint64_t arena_bytes, arena_used, arena_peak;
} ihe_server_stats;
int ihe_server_create(ihe_ctx *ctx, const ihe_server_config *cfg, ihe_server **out);
/* band: an MFMA band over the big indices of num_tags tag SNPs. The band must
 * outlive the server: ihe_band_destroy fails with IHE_ERR_BUSY until the
 * server is destroyed. */
int ihe_server_add_model(ihe_server *server, const ihe_band *band, uint32_t num_tags,
                         uint32_t *model);
/* The request's host buffers must stay valid until it completes. */
int ihe_server_submit(ihe_server *server, const ihe_request *req, uint64_t *ticket);
/* Completed requests report their latency (arrival, or submit when the
 * request gives none, to scores on the host) and are forgotten: one waiter per
 * ticket (any other poll or wait on it then fails with IHE_ERR_ARG). */
int ihe_server_poll(ihe_server *server, uint64_t ticket, int *done, double *latency_ms);
int ihe_server_wait(ihe_server *server, uint64_t ticket, double *latency_ms);
int ihe_server_get_stats(ihe_server *server, ihe_server_stats *stats);
int ihe_server_destroy(ihe_server *server); /* drains the queue first */

/* ---------------- keys, encrypt, decrypt ---------------- */
/* Binary TLWE key s[0..N) (tLweKeyGen, eval/idash.cpp:429-430). */
int ihe_key_create(ihe_ctx *ctx, const int32_t *key, ihe_key **out);
int ihe_key_destroy(ihe_key *key);

/* Replaces encrypt_data_ph1 + encrypt_data_ph2 (eval/idash.cpp:615-679;
 * eval/idash.h:248-249). d_geno: [num_tags][NS] int8 genotypes in tag order
 * (tag q owns big indices 3q..3q+2); values outside {0,1,2} are missing.
 * d_geno == NULL encrypts zero in every slot: encrypt_data_ph1 alone
 * (eval/idash.cpp:615-640), whose scores the caller adds afterwards.
 * Writes ceil(3*num_tags / R) ciphertexts (FeatIndex order) to d_out and
 * alpha^2 = 2^-50 to d_out_var. seed selects the counter-based random streams. */
int ihe_encrypt(ihe_ctx *ctx, const ihe_key *key, const ihe_layout *layout, uint32_t num_tags,
                const int8_t *d_geno, uint64_t seed, int32_t *d_out, double *d_out_var);

/* Replaces the per-ciphertext body of decrypt_predictions
 * (eval/idash.cpp:711-719; eval/idash.h:254): out[c][j] =
 * float(t32tod(phase_c[j])), phase = b - a*s, for j < num_samples. */
int ihe_decrypt(ihe_ctx *ctx, const ihe_key *key, uint32_t num_samples, const int32_t *d_in,
                uint32_t num_cts, float *d_out);

/* The same decryption as ihe_decrypt, allocating nothing: no context, no
 * ihe_key, no device or host allocation. It has no counterpart in the
 * reference; it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). key: host
 * [N] binary key; d_in, d_out as for ihe_decrypt; enqueued on hip_stream
 * (NULL: the default stream) of the current device. The ring kernel's work
 * counters (16 words) live in a device array of the library, one per GPU:
 * calls that overlap on one GPU race on them and must be ordered (one stream,
 * or events between streams). At most 2^26 ciphertexts per call
 * (IHE_ERR_ARG beyond). The array's address is looked up when a context is
 * created on the device (ihe_ctx_create) or at the first call there: make
 * that first call, or create a context, before capturing a graph. */
int ihe_decrypt_noalloc(const int32_t *key, uint32_t num_samples, const int32_t *d_in,
                        uint32_t num_cts, float *d_out, void *hip_stream);

/* Noise table of the encrypt kernel: table[k] = P(|e| <= k) * 2^63 for
 * e = trunc(128 g), g ~ N(0,1) (TFHE gaussian32 at alpha = 2^-25). */
#define IHE_NOISE_TABLE_LEN 1280u
int ihe_noise_table(uint64_t *table, uint32_t len);

#ifdef __cplusplus
}
#endif

#endif /* IDASH_HIP_H */
