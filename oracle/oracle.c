/*
 * CPU oracle for the secure-imputation hot path. TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
 * never by the product. See oracle.h for what each function restates and how
 * it is pinned.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* TFHE numeric functions: dtot32(d) = int32(int64((d - int64(d)) * 2^32)),
 * t32tod(x) = x / 2^32 (used at eval/idash.cpp:39-45 and :718). */
int32_t orc_dtot32(double d) {
    double frac = d - (double)(int64_t)d;
    return (int32_t)(uint32_t)(uint64_t)(int64_t)(frac * 4294967296.0);
}

double orc_t32tod(int32_t x) { return (double)x / 4294967296.0; }

/* eval/idash.cpp:29-45: IN_SCALING_FACTOR = 1/16384, ONE = 1 * f, NAN_v = (3-v)/6 * f. */
void orc_encoding_constants(int32_t out[4]) {
    const double f = 1. / 16384;
    out[0] = orc_dtot32(f);
    out[1] = orc_dtot32(3. / 6. * f);
    out[2] = orc_dtot32(2. / 6. * f);
    out[3] = orc_dtot32(1. / 6. * f);
}

/* TFHE torusPolynomialMulByXai: out = X^a * in mod X^N+1, a in [0, 2N)
 * (eval/idash.cpp:832 calls it with a = 2N - r*RS). */
void orc_mul_by_xai(int32_t *out, int32_t a, const int32_t *in, int32_t N) {
    if (a < N) {
        for (int32_t i = 0; i < a; ++i) out[i] = (int32_t)(0u - (uint32_t)in[i - a + N]);
        for (int32_t i = a; i < N; ++i) out[i] = in[i - a];
    } else {
        const int32_t aa = a - N;
        for (int32_t i = 0; i < aa; ++i) out[i] = in[i - aa + N];
        for (int32_t i = aa; i < N; ++i) out[i] = (int32_t)(0u - (uint32_t)in[i - aa]);
    }
}

/* Exact negacyclic product: out[j] = sum_i a[i] s[j-i] with X^N = -1, mod 2^32. */
void orc_negacyclic_mul(int32_t *out, const int32_t *a, const int32_t *s, int32_t N) {
    for (int32_t j = 0; j < N; ++j) {
        int64_t acc = 0;
        for (int32_t i = 0; i <= j; ++i) acc += (int64_t)a[i] * (int64_t)s[j - i];
        for (int32_t i = j + 1; i < N; ++i) acc -= (int64_t)a[i] * (int64_t)s[j - i + N];
        out[j] = (int32_t)(uint32_t)(uint64_t)acc;
    }
}

static inline uint32_t wrap_mul(int32_t a, int32_t b) { return (uint32_t)a * (uint32_t)b; }

/* cloud_compute_score (eval/idash.cpp:763-848), literal structure. */
int orc_cloud_reference(int32_t N, uint32_t NS, uint32_t R, uint32_t RS,
                        const int32_t *ct, const double *var_in, uint32_t num_in,
                        uint32_t num_out, const uint32_t *row_ptr,
                        const uint32_t *in_bidx, const int32_t *weight,
                        int32_t *out, double *var_out) {
    int32_t ONE4[4];
    orc_encoding_constants(ONE4);
    const int32_t ONE = ONE4[0];
    const size_t ctw = 2 * (size_t)N;
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * ctw * R);
    double *tmp_var = (double *)malloc(sizeof(double) * R);
    int32_t *rot = (int32_t *)malloc(sizeof(int32_t) * N);
    int rc = 0;
    for (uint32_t o = 0; o < num_out && rc == 0; ++o) {
        /* tLweClear on every region temporary (:795-797) */
        memset(tmp, 0, sizeof(uint32_t) * ctw * R);
        for (uint32_t r = 0; r < R; ++r) tmp_var[r] = 0.0;
        for (uint32_t e = row_ptr[o]; e < row_ptr[o + 1]; ++e) {
            const uint32_t bidx = in_bidx[e];
            const int32_t coeff = weight[e];
            if (bidx == ORC_CONSTANT_BIDX) {
                /* :805-810, int32 product that may wrap */
                const uint32_t scaled = wrap_mul(coeff, ONE);
                for (uint32_t j = 0; j < NS; ++j) tmp[N + j] += scaled;
            } else {
                const uint32_t region = bidx % R, idx = bidx / R;
                if (idx >= num_in) { rc = -1; break; }
                const int32_t *src = ct + ctw * idx;
                uint32_t *dst = tmp + ctw * region;
                /* tLweAddMulTo (:817): a[i] += p*s.a[i] for both polynomials,
                 * variance += (p*p) * var with p*p an int32 product */
                for (size_t i = 0; i < ctw; ++i) dst[i] += wrap_mul(coeff, src[i]);
                tmp_var[region] += (double)(int32_t)wrap_mul(coeff, coeff) * var_in[idx];
            }
        }
        int32_t *dst = out + ctw * o;
        /* tLweCopy of region 0 (:825) */
        memcpy(dst, tmp, sizeof(int32_t) * ctw);
        double v = tmp_var[0];
        /* rotate regions >= 1 by X^(2N - r*RS) and tLweAddTo (:826-836); the
         * rotation temporary never receives a variance, so adds 0 */
        for (uint32_t r = 1; r < R; ++r) {
            const int32_t amount = 2 * N - (int32_t)(r * RS);
            for (int p = 0; p < 2; ++p) {
                orc_mul_by_xai(rot, amount, (const int32_t *)(tmp + ctw * r + (size_t)p * N), N);
                for (int32_t i = 0; i < N; ++i)
                    dst[(size_t)p * N + i] = (int32_t)((uint32_t)dst[(size_t)p * N + i] + (uint32_t)rot[i]);
            }
            v += 0.0;
        }
        /* hide b[RS..N) (:838-841) */
        for (int32_t j = (int32_t)RS; j < N; ++j) dst[N + j] = 0;
        var_out[o] = v;
    }
    free(tmp);
    free(tmp_var);
    free(rot);
    return rc;
}

/* Closed form of the same step. */
int orc_cloud_closed_form(int32_t N, uint32_t NS, uint32_t R, uint32_t RS,
                          const int32_t *ct, const double *var_in, uint32_t num_in,
                          uint32_t num_out, const uint32_t *row_ptr,
                          const uint32_t *in_bidx, const int32_t *weight,
                          int32_t *out, double *var_out) {
    int32_t ONE4[4];
    orc_encoding_constants(ONE4);
    const size_t ctw = 2 * (size_t)N;
    for (uint32_t o = 0; o < num_out; ++o) {
        int32_t *dst = out + ctw * o;
        double v = 0.0;
        for (int p = 0; p < 2; ++p) {
            for (int32_t i = 0; i < N; ++i) {
                uint32_t acc = 0;
                for (uint32_t e = row_ptr[o]; e < row_ptr[o + 1]; ++e) {
                    const uint32_t bidx = in_bidx[e];
                    if (bidx == ORC_CONSTANT_BIDX) {
                        if (p == 1 && (uint32_t)i < NS) acc += wrap_mul(weight[e], ONE4[0]);
                        continue;
                    }
                    const uint32_t idx = bidx / R;
                    if (idx >= num_in) return -1;
                    const int32_t m = (int32_t)((bidx % R) * RS);
                    const int32_t *src = ct + ctw * idx + (size_t)p * N;
                    const int32_t t = i + m;
                    const uint32_t x = t < N ? (uint32_t)src[t] : 0u - (uint32_t)src[t - N];
                    acc += (uint32_t)weight[e] * x;
                }
                if (p == 1 && (uint32_t)i >= RS) acc = 0;
                dst[(size_t)p * N + i] = (int32_t)acc;
            }
        }
        for (uint32_t e = row_ptr[o]; e < row_ptr[o + 1]; ++e) {
            const uint32_t bidx = in_bidx[e];
            if (bidx == ORC_CONSTANT_BIDX || bidx % R != 0) continue;
            v += (double)(int32_t)wrap_mul(weight[e], weight[e]) * var_in[bidx / R];
        }
        var_out[o] = v;
    }
    return 0;
}

/* TFHE tLwePhase (eval/idash.cpp:714): phase = b - a*s. */
void orc_phase(int32_t N, const int32_t *ct, const int32_t *key, int32_t *phase) {
    int32_t *as = (int32_t *)malloc(sizeof(int32_t) * N);
    orc_negacyclic_mul(as, ct, key, N);
    for (int32_t j = 0; j < N; ++j) phase[j] = (int32_t)((uint32_t)ct[N + j] - (uint32_t)as[j]);
    free(as);
}

/* decrypt_predictions inner loop (eval/idash.cpp:711-719). */
void orc_decrypt(int32_t N, uint32_t NS, const int32_t *cts, uint32_t num_cts,
                 const int32_t *key, float *out) {
    /* ciphertexts are independent (the reference's omp for, eval/idash.cpp:701-705) */
#pragma omp parallel
    {
        int32_t *phase = (int32_t *)malloc(sizeof(int32_t) * N);
#pragma omp for schedule(dynamic, 16)
        for (int64_t c = 0; c < (int64_t)num_cts; ++c) {
            orc_phase(N, cts + (size_t)c * 2 * N, key, phase);
            for (uint32_t j = 0; j < NS; ++j) out[(size_t)c * NS + j] = (float)orc_t32tod(phase[j]);
        }
        free(phase);
    }
}

/* Philox4x32-10: ten rounds of (hi,lo) = mulhilo(M, x) with Weyl key bumps. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* P(|trunc(128 g)| <= k) = P(|g| < (k+1)/128) = 1 - erfc((k+1)/(128 sqrt 2)). */
void orc_noise_table(uint64_t table[ORC_NOISE_TABLE_LEN]) {
    const double two63 = 9223372036854775808.0;
    for (int k = 0; k < ORC_NOISE_TABLE_LEN; ++k) {
        const double tail = erfc((double)(k + 1) / (128.0 * sqrt(2.0))) * two63;
        uint64_t t = (uint64_t)1 << 63;
        if (tail >= 1.0) t -= (uint64_t)tail;
        table[k] = t;
    }
    table[ORC_NOISE_TABLE_LEN - 1] = (uint64_t)1 << 63;
}

static int32_t noise_sample(uint64_t u, const uint64_t *table) {
    const uint64_t m = u >> 1;
    /* smallest k with m < table[k] */
    int lo = 0, hi = ORC_NOISE_TABLE_LEN - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (m < table[mid]) hi = mid; else lo = mid + 1;
    }
    return (u & 1) ? -lo : lo;
}

/* One ciphertext c of encrypt_data_ph1 + ph2: a from Philox stream 0, b =
 * a*s + e (e from Philox stream 1 through the noise table), then the one-hot
 * messages of the big indices c*R .. c*R + R - 1 at slots r*RS + j. */
static void encrypt_one(int32_t N, uint32_t NS, uint32_t R, uint32_t RS, uint32_t num_tags,
                        const int8_t *geno, const int32_t *key, const uint32_t skey[2],
                        const uint64_t *table, const int32_t C4[4], uint32_t c, int32_t *as,
                        int32_t *a, double *var) {
    const uint32_t ncols = 3 * num_tags;
    int32_t *b = a + N;
    for (int32_t w = 0; w < N; w += 4) {
        const uint32_t ctr[4] = {(uint32_t)(w / 4), c, 0, 0};
        uint32_t r4[4];
        orc_philox4x32_10(ctr, skey, r4);
        for (int q = 0; q < 4 && w + q < N; ++q) a[w + q] = (int32_t)r4[q];
    }
    orc_negacyclic_mul(as, a, key, N);
    for (int32_t j = 0; j < N; j += 2) {
        const uint32_t ctr[4] = {(uint32_t)(j / 2), c, 1, 0};
        uint32_t r4[4];
        orc_philox4x32_10(ctr, skey, r4);
        const uint64_t u0 = (uint64_t)r4[0] | ((uint64_t)r4[1] << 32);
        const uint64_t u1 = (uint64_t)r4[2] | ((uint64_t)r4[3] << 32);
        b[j] = (int32_t)((uint32_t)as[j] + (uint32_t)noise_sample(u0, table));
        if (j + 1 < N) b[j + 1] = (int32_t)((uint32_t)as[j + 1] + (uint32_t)noise_sample(u1, table));
    }
    for (uint32_t r = 0; r < R; ++r) {
        const uint32_t bidx = c * R + r;
        if (bidx >= ncols) break;
        const uint32_t tag = bidx / 3, v = bidx % 3;
        for (uint32_t j = 0; j < NS; ++j) {
            const int8_t g = geno[(size_t)tag * NS + j];
            uint32_t add = 0;
            if (g == 0 || g == 1 || g == 2) {
                if ((uint32_t)g == v) add = (uint32_t)C4[0];
            } else {
                add = (uint32_t)C4[1 + v];
            }
            b[r * RS + j] = (int32_t)((uint32_t)b[r * RS + j] + add);
        }
    }
    *var = ldexp(1.0, -50);
}

void orc_encrypt_cts(int32_t N, uint32_t NS, uint32_t R, uint32_t RS,
                     uint32_t num_tags, const int8_t *geno, const int32_t *key,
                     uint64_t seed, const uint32_t *cts, uint32_t count,
                     int32_t *out, double *var_out) {
    int32_t C4[4];
    orc_encoding_constants(C4);
    uint64_t table[ORC_NOISE_TABLE_LEN];
    orc_noise_table(table);
    const uint32_t skey[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    /* ciphertexts are independent (the reference's omp for, eval/idash.cpp:623) */
#pragma omp parallel
    {
        int32_t *as = (int32_t *)malloc(sizeof(int32_t) * N);
#pragma omp for schedule(dynamic, 4)
        for (int64_t i = 0; i < (int64_t)count; ++i)
            encrypt_one(N, NS, R, RS, num_tags, geno, key, skey, table, C4,
                        cts ? cts[i] : (uint32_t)i, as, out + (size_t)i * 2 * N, var_out + i);
        free(as);
    }
}

void orc_encrypt(int32_t N, uint32_t NS, uint32_t R, uint32_t RS,
                 uint32_t num_tags, const int8_t *geno, const int32_t *key,
                 uint64_t seed, int32_t *out, double *var_out) {
    const uint32_t num_cts = (3 * num_tags + R - 1) / R;
    orc_encrypt_cts(N, NS, R, RS, num_tags, geno, key, seed, NULL, num_cts, out, var_out);
}

/* compute_score (eval/idash.cpp:888-927) with the one-hot of
 * compute_plaintext_onehot (:850-886), for one output. */
void orc_plaintext_score(uint32_t NS, uint32_t num_tags, const int8_t *geno,
                         uint32_t nnz, const uint32_t *in_bidx, const int32_t *weight,
                         double *score) {
    const double ONE = 1. / 16384;
    const double NANV[3] = {3. / 6. * ONE, 2. / 6. * ONE, 1. / 6. * ONE};
    for (uint32_t j = 0; j < NS; ++j) score[j] = 0.0;
    for (uint32_t e = 0; e < nnz; ++e) {
        const uint32_t bidx = in_bidx[e];
        for (uint32_t j = 0; j < NS; ++j) {
            if (bidx == ORC_CONSTANT_BIDX) {
                score[j] += weight[e] * ONE;
            } else {
                const uint32_t tag = bidx / 3, v = bidx % 3;
                if (tag >= num_tags) continue;
                const int8_t g = geno[(size_t)tag * NS + j];
                double x;
                if (g == 0 || g == 1 || g == 2) x = ((uint32_t)g == v) ? ONE : 0.0;
                else x = NANV[v];
                score[j] += weight[e] * x;
            }
        }
    }
}
