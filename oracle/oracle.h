/*
 * CPU oracle for the secure-imputation hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline). The product path never links or calls it.
 *
 * Every function restates the reference algorithm it cites (file:line in
 * ssmiler/idash2019_2). The TFHE library the reference calls is absent from
 * the reference tree (submodule tfhe/tfhe not checked out, version unpinned);
 * its used subset is restated from the library's published semantics:
 *   dtot32/t32tod, torusPolynomialMulByXai, tLweAddMulTo/tLweAddTo/tLweCopy/
 *   tLweClear, torusPolynomialAddMulR/SubMulR (negacyclic product), the
 *   Gaussian noise gaussian32 = dtot32(normal(0, alpha)).
 *
 * Pinning (each is checked by tests/test_oracle.py against something other
 * than this file's own formula):
 *   orc_dtot32/orc_t32tod        known answers 2^18, 131072, 87381, 43690
 *   orc_negacyclic_mul           numpy polynomial product folded mod X^N+1
 *   orc_mul_by_xai               product with the monomial X^a (numpy)
 *   orc_cloud_reference          closed form (orc_cloud_closed_form), the
 *                                numpy weighted sum at R=1, monomial special
 *                                cases at R>1, shuffled-order invariance
 *   orc_phase / orc_decrypt      schoolbook identity b - a*s, float rounding
 *                                against numpy's float32 cast
 *   orc_philox4x32_10            Random123 known-answer vectors
 *   orc_noise_table              erfc closed form and monotonicity; sampled
 *                                moments of the noise
 *   orc_encrypt                  round trip: phase == e + message exactly
 */
#ifndef IDASH_ORACLE_H
#define IDASH_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_CONSTANT_BIDX 0xFFFFFFFFu
#define ORC_NOISE_TABLE_LEN 1280

/* dtot32 / t32tod (TFHE numeric functions; used at eval/idash.cpp:39-45, 718). */
int32_t orc_dtot32(double d);
double orc_t32tod(int32_t x);

/* ONE_IN_T32, NAN_0_IN_T32, NAN_1_IN_T32, NAN_2_IN_T32 (eval/idash.cpp:29-45). */
void orc_encoding_constants(int32_t out[4]);

/* out = X^a * in mod (X^N + 1), a in [0, 2N): TFHE torusPolynomialMulByXai,
 * called at eval/idash.cpp:832. */
void orc_mul_by_xai(int32_t *out, int32_t a, const int32_t *in, int32_t N);

/* out = a * s mod (X^N + 1, 2^32), exact int64 schoolbook (TFHE
 * torusPolynomialAddMulR with a zero accumulator; eval/idash.cpp:625, 714). */
void orc_negacyclic_mul(int32_t *out, const int32_t *a, const int32_t *s, int32_t N);

/*
 * Cloud step, restated literally from cloud_compute_score
 * (eval/idash.cpp:763-848): per-output region temporaries cleared
 * (:795-797), tLweAddMulTo per model coefficient into region bidx%R (:812-817),
 * bias added to region 0 for samples < NS (:805-810), copy of region 0 (:825),
 * regions r>=1 rotated by X^(2N - r*RS) and added (:826-836), b[RS..N) zeroed
 * (:838-841). Variances follow tLweAddMulTo / tLweCopy / tLweAddTo.
 *
 * Model: CSR over output slots; entries (in_bidx, weight), in_bidx ==
 * ORC_CONSTANT_BIDX is the bias (eval/idash.h:93). Entries are visited in the
 * given order (the reference visits them in unordered_map order).
 * ct:  [num_in][2][N] int32 (a then b), dense by FeatIndex (bidx / R).
 * out: [num_out][2][N]; var_out: [num_out].
 * Returns 0, or -1 if a coefficient references a ciphertext >= num_in
 * (the reference aborts at eval/idash.h:164).
 */
int orc_cloud_reference(int32_t N, uint32_t NS, uint32_t R, uint32_t RS,
                        const int32_t *ct, const double *var_in, uint32_t num_in,
                        uint32_t num_out, const uint32_t *row_ptr,
                        const uint32_t *in_bidx, const int32_t *weight,
                        int32_t *out, double *var_out);

/* The same step from its closed form: out[i] = sum_c w_c * rho_{(c%R)*RS}(CT[c/R])[i],
 * rho_m(P)[i] = P[i+m] (i+m < N) else -P[i+m-N]; bias and tail as above. */
int orc_cloud_closed_form(int32_t N, uint32_t NS, uint32_t R, uint32_t RS,
                          const int32_t *ct, const double *var_in, uint32_t num_in,
                          uint32_t num_out, const uint32_t *row_ptr,
                          const uint32_t *in_bidx, const int32_t *weight,
                          int32_t *out, double *var_out);

/* phase = b - a*s (TFHE tLwePhase, eval/idash.cpp:714). ct: [2][N]. */
void orc_phase(int32_t N, const int32_t *ct, const int32_t *key, int32_t *phase);

/* decrypt_predictions per ciphertext (eval/idash.cpp:711-719):
 * out[c][j] = float(t32tod(phase_c[j])) for j < NS. */
void orc_decrypt(int32_t N, uint32_t NS, const int32_t *cts, uint32_t num_cts,
                 const int32_t *key, float *out);

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Cumulative table of |e| for e = trunc(128 * g), g ~ N(0,1): the distribution
 * of TFHE's gaussian32(0, 2^-25) on Torus32. table[k] = P(|e| <= k) * 2^63. */
void orc_noise_table(uint64_t table[ORC_NOISE_TABLE_LEN]);

/*
 * Encryption of the tag genotypes, the semantics of encrypt_data_ph1 +
 * encrypt_data_ph2 (eval/idash.cpp:615-679) with the counter-based random
 * streams of this build: for ciphertext c (FeatIndex) and word w,
 *   a[w]  = Philox(ctr = {w/4, c, 0, 0}, key = seed)[w % 4]
 *   e[j]  = noise from Philox(ctr = {j/2, c, 1, 0}, key = seed) 64-bit half j%2
 *   b     = a*s + e, then setScore (eval/idash.h:156-160) adds ONE_IN_T32 at
 *           b[j + r*RS] for column bidx = c*R + r whose one-hot value matches
 *           genotype geno[bidx/3][j], or NAN_{bidx%3}_IN_T32 when the genotype
 *           is not in {0,1,2} (eval/idash.cpp:656-676).
 * geno: [num_tags][NS] int8. Ciphertexts: ceil(3*num_tags / R).
 * var_out = alpha^2 = 2^-50 (tLweSymEncryptZero).
 */
void orc_encrypt(int32_t N, uint32_t NS, uint32_t R, uint32_t RS,
                 uint32_t num_tags, const int8_t *geno, const int32_t *key,
                 uint64_t seed, int32_t *out, double *var_out);
/* The same for the listed ciphertexts only: out[i] and var_out[i] receive
 * ciphertext cts[i] of orc_encrypt's output (cts NULL: ciphertext i). */
void orc_encrypt_cts(int32_t N, uint32_t NS, uint32_t R, uint32_t RS,
                     uint32_t num_tags, const int8_t *geno, const int32_t *key,
                     uint64_t seed, const uint32_t *cts, uint32_t count,
                     int32_t *out, double *var_out);

/* Plaintext scorer compute_score (eval/idash.cpp:888-927), one output slot:
 * score[j] = sum over entries of w * onehot(bidx)[j] (bias: w * ONE_IN_D). */
void orc_plaintext_score(uint32_t NS, uint32_t num_tags, const int8_t *geno,
                         uint32_t nnz, const uint32_t *in_bidx, const int32_t *weight,
                         double *score);

/* CPU baseline: cloud_compute_score with the reference's data structures
 * (hash-map model and ciphertext index, OpenMP over outputs, per-thread region
 * temporaries). Returns the wall time of the compute in seconds; outputs like
 * orc_cloud_reference. Implemented in cpu_baseline.cpp. */
double orc_cpu_baseline_cloud(int num_threads, uint32_t NS, uint32_t R, uint32_t RS,
                              const int32_t *ct, const double *var_in, uint32_t num_in,
                              uint32_t num_out, const uint32_t *row_ptr,
                              const uint32_t *in_bidx, const int32_t *weight,
                              int32_t *out, double *var_out);

#ifdef __cplusplus
}
#endif

#endif
