/*
 * svs_oracle.c — CPU oracle (plain C restatement) of the reference's
 * encrypted-score matrix-vector path. TEST INFRASTRUCTURE ONLY: the checker
 * and the CPU baseline, never the product. See svs_oracle.h for the pinning
 * status of each function.
 *
 * Reference: sacs-epfl/secure-vector-search (Rust, edition 2024). All paths
 * below are relative to the reference repository root.
 */
#include "svs_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------
 * Field arithmetic, p = 2^61 - 1.
 * scorer-bntm/src/crypto.rs:47-67 (mersenne_reduce), :70-75 (add_mod),
 * :78-81 (sub_mod), :88-91 (mul_mod).
 * ---------------------------------------------------------------------- */
static uint64_t reduce128(u128 x) {
    /* Two Mersenne folds, then at most two conditional subtractions. */
    u128 s1 = (x & (u128)OR_P) + (x >> 61);
    uint64_t r = (uint64_t)(s1 & (u128)OR_P) + (uint64_t)(s1 >> 61);
    if (r >= OR_P) r -= OR_P;
    if (r >= OR_P) r -= OR_P;
    return r;
}

uint64_t or_mersenne_reduce(uint64_t lo, uint64_t hi) {
    return reduce128(((u128)hi << 64) | lo);
}

uint64_t or_add_mod(uint64_t a, uint64_t b) {
    uint64_t s = a + b;
    return s >= OR_P ? s - OR_P : s;
}

uint64_t or_sub_mod(uint64_t a, uint64_t b) { return a >= b ? a - b : a + OR_P - b; }

uint64_t or_mul_mod(uint64_t a, uint64_t b) { return reduce128((u128)a * (u128)b); }

/* ------------------------------------------------------------------------
 * ChaCha20 keystream as consumed by rand_chacha's ChaCha20Rng (0.10):
 * state = "expand 32-byte k" | key (seed as 8 LE u32) | 64-bit block
 * counter (words 12-13) | 64-bit stream id = 0 (words 14-15). 20 rounds.
 * next_u32 reads keystream words in order; next_u64 = lo word | hi word << 32
 * of the next two words (rand_core BlockRng). For counter < 2^32 this equals
 * RFC 7539 with an all-zero nonce, which its A.1 test vectors pin.
 * ---------------------------------------------------------------------- */
#define ROTL32(v, c) (((v) << (c)) | ((v) >> (32 - (c))))
#define QR(a, b, c, d)            \
    a += b; d ^= a; d = ROTL32(d, 16); \
    c += d; b ^= c; b = ROTL32(b, 12); \
    a += b; d ^= a; d = ROTL32(d, 8);  \
    c += d; b ^= c; b = ROTL32(b, 7);

void or_chacha20_block(const uint32_t key[8], uint64_t counter, uint32_t out[16]) {
    uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                       key[0], key[1], key[2], key[3], key[4], key[5], key[6], key[7],
                       (uint32_t)counter, (uint32_t)(counter >> 32), 0u, 0u};
    uint32_t x[16];
    memcpy(x, in, sizeof x);
    for (int i = 0; i < 10; ++i) {
        QR(x[0], x[4], x[8], x[12]);
        QR(x[1], x[5], x[9], x[13]);
        QR(x[2], x[6], x[10], x[14]);
        QR(x[3], x[7], x[11], x[15]);
        QR(x[0], x[5], x[10], x[15]);
        QR(x[1], x[6], x[11], x[12]);
        QR(x[2], x[7], x[8], x[13]);
        QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
}

void or_rng_seed(or_rng* r, const uint8_t seed[32]) {
    for (int i = 0; i < 8; ++i)
        r->key[i] = (uint32_t)seed[4 * i] | ((uint32_t)seed[4 * i + 1] << 8) |
                    ((uint32_t)seed[4 * i + 2] << 16) | ((uint32_t)seed[4 * i + 3] << 24);
    r->block = 0;
    r->idx = 16;
}

uint32_t or_next_u32(or_rng* r) {
    if (r->idx >= 16) {
        or_chacha20_block(r->key, r->block++, r->buf);
        r->idx = 0;
    }
    return r->buf[r->idx++];
}

uint64_t or_next_u64(or_rng* r) {
    uint64_t lo = or_next_u32(r);
    uint64_t hi = or_next_u32(r);
    return lo | (hi << 32);
}

/* sample_fp, scorer-bntm/src/crypto.rs:155-162: mask to 61 bits, reject P. */
uint64_t or_sample_fp(or_rng* r) {
    for (;;) {
        uint64_t x = or_next_u64(r) & OR_P;
        if (x < OR_P) return x;
    }
}

/* domain_separate, crypto.rs:142-150: XOR the tag into the seed cyclically,
 * then draw a [u8; 32] from ChaCha20 (each byte = next_u32 truncated). */
void or_domain_separate(const uint8_t seed[32], const uint8_t* tag, size_t tag_len,
                        uint8_t out[32]) {
    uint8_t s[32];
    memcpy(s, seed, 32);
    for (size_t i = 0; i < tag_len; ++i) s[i % 32] ^= tag[i];
    or_rng r;
    or_rng_seed(&r, s);
    for (int i = 0; i < 32; ++i) out[i] = (uint8_t)or_next_u32(&r);
}

/* Threshold of sample_sparse_pm1 (crypto.rs:211-212) at mu = n^(eps-1),
 * n = 1024, eps = 0.7 (params.rs:242-244). In f64, 0.7 - 1.0 is
 * -0.30000000000000004, so mu = 0.12499999999999996 and the threshold is
 * 2^29 - 1, not 2^29. */
uint32_t or_sparse_threshold(void) {
    double mu = pow(1024.0, 0.7 - 1.0);
    double t = mu * 4294967296.0;
    uint64_t th = (uint64_t)t;
    if (th > 0xFFFFFFFFull) th = 0xFFFFFFFFull;
    return (uint32_t)th;
}

/* sample_sparse_pm1, crypto.rs:207-223. */
size_t or_sample_sparse_pm1(or_rng* r, size_t n, uint32_t threshold, uint32_t* idx,
                            int8_t* sign) {
    size_t cnt = 0;
    for (size_t j = 0; j < n; ++j) {
        uint32_t w = or_next_u32(r);
        if (w < threshold) {
            idx[cnt] = (uint32_t)j;
            sign[cnt] = (or_next_u32(r) & 1u) == 0 ? -1 : 1;
            ++cnt;
        }
    }
    return cnt;
}

/* ------------------------------------------------------------------------
 * Encoding helpers, crypto.rs:105-122 (quantise_vector), :128-135
 * (fp_to_signed).
 * ---------------------------------------------------------------------- */
void or_quantise_vector(const float* v, size_t d, size_t n, uint64_t q, uint64_t* out) {
    double qs = (double)q;
    for (size_t i = 0; i < n; ++i) out[i] = 0;
    for (size_t i = 0; i < d; ++i) {
        double scaled = (double)v[i] * qs;
        double rd = round(scaled); /* f64::round: half away from 0 */
        /* `as i64` saturates and maps NaN to 0; the negation below wraps */
        int64_t rounded = rd != rd ? 0
                          : rd >= 9223372036854775808.0 ? INT64_MAX
                          : rd <= -9223372036854775808.0 ? INT64_MIN
                          : (int64_t)rd;
        if (rounded >= 0) {
            out[i] = (uint64_t)rounded % OR_P;
        } else {
            uint64_t k = 0ull - (uint64_t)rounded;
            uint64_t rr = k % OR_P;
            out[i] = rr == 0 ? 0 : OR_P - rr;
        }
    }
}

/* quantise_vector over `rows` row-major vectors of d floats each. */
void or_quantise_rows(const float* v, size_t rows, size_t d, size_t n, uint64_t q,
                      uint64_t* out) {
    for (size_t r = 0; r < rows; ++r) or_quantise_vector(v + r * d, d, n, q, out + r * n);
}

int64_t or_fp_to_signed(uint64_t s) {
    return s < OR_P / 2 ? (int64_t)s : (int64_t)s - (int64_t)OR_P;
}

/* ------------------------------------------------------------------------
 * Keygen / encryption.
 * ---------------------------------------------------------------------- */
/* generate_subspace, crypto.rs:235-238: L (n x n1) row-major, sample_fp. */
void or_generate_subspace(const uint8_t key_seed[32], size_t n, size_t n1, uint64_t* l_out) {
    or_rng r;
    or_rng_seed(&r, key_seed);
    for (size_t i = 0; i < n * n1; ++i) l_out[i] = or_sample_fp(&r);
}

/* regenerate_h_s, crypto.rs:258-270: H (m x n1) first, then S row by row. */
size_t or_regenerate_h_s(const uint8_t h_seed[32], size_t m, size_t n, size_t n1,
                         uint64_t* h_out, uint32_t* s_rowptr, uint32_t* s_idx,
                         int8_t* s_sign, size_t s_cap) {
    or_rng r;
    or_rng_seed(&r, h_seed);
    for (size_t i = 0; i < m * n1; ++i) h_out[i] = or_sample_fp(&r);
    uint32_t th = or_sparse_threshold();
    uint32_t* tmp_idx = (uint32_t*)malloc(n * sizeof(uint32_t));
    int8_t* tmp_sign = (int8_t*)malloc(n);
    size_t total = 0;
    s_rowptr[0] = 0;
    for (size_t i = 0; i < m; ++i) {
        size_t c = or_sample_sparse_pm1(&r, n, th, tmp_idx, tmp_sign);
        if (total + c > s_cap) {
            free(tmp_idx);
            free(tmp_sign);
            return (size_t)-1;
        }
        memcpy(s_idx + total, tmp_idx, c * sizeof(uint32_t));
        memcpy(s_sign + total, tmp_sign, c);
        total += c;
        s_rowptr[i + 1] = (uint32_t)total;
    }
    free(tmp_idx);
    free(tmp_sign);
    return total;
}

static int8_t sparse_lookup(const uint32_t* idx, const int8_t* sign, uint32_t lo, uint32_t hi,
                            uint32_t col) {
    /* binary search in a sorted row, crypto.rs:484-489 */
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (idx[mid] == col) return sign[mid];
        if (idx[mid] < col) lo = mid + 1; else hi = mid;
    }
    return 0;
}

/* encrypt_cluster, crypto.rs:405-478: M_enc = M + (H L^T + S), AL = M L. */
void or_encrypt_cluster(const uint64_t* m_plain, size_t m, size_t n, size_t n1,
                        const uint64_t* l, const uint64_t* h, const uint32_t* s_rowptr,
                        const uint32_t* s_idx, const int8_t* s_sign, uint64_t* m_enc_out,
                        uint64_t* a_l_out) {
    for (size_t i = 0; i < m; ++i) {
        const uint64_t* h_row = h + i * n1;
        for (size_t j = 0; j < n; ++j) {
            const uint64_t* l_row = l + j * n1;
            u128 acc = 0;
            for (size_t k = 0; k < n1; ++k) {
                acc += (u128)h_row[k] * l_row[k];
                if ((k & 31) == 31) acc = reduce128(acc);
            }
            uint64_t hl = reduce128(acc);
            int8_t st = sparse_lookup(s_idx, s_sign, s_rowptr[i], s_rowptr[i + 1], (uint32_t)j);
            uint64_t mp = st == 1 ? or_add_mod(hl, 1) : st == -1 ? or_sub_mod(hl, 1) : hl;
            m_enc_out[i * n + j] = or_add_mod(m_plain[i * n + j], mp);
        }
        for (size_t k = 0; k < n1; ++k) {
            u128 acc = 0;
            for (size_t j = 0; j < n; ++j) {
                acc += (u128)m_plain[i * n + j] * l[j * n1 + k];
                if ((j & 31) == 31) acc = reduce128(acc);
            }
            a_l_out[i * n1 + k] = reduce128(acc);
        }
    }
}

/* encode_query, crypto.rs:508-553: G (n1 sample_fp), T (sparse), then
 * v_enc = L G + T + v. */
size_t or_encode_query(const uint64_t* v, const uint64_t* l, size_t n, size_t n1,
                       const uint8_t q_seed[32], uint64_t* v_enc, uint64_t* g,
                       uint32_t* t_idx, int8_t* t_sign) {
    or_rng r;
    or_rng_seed(&r, q_seed);
    for (size_t k = 0; k < n1; ++k) g[k] = or_sample_fp(&r);
    size_t t_nnz = or_sample_sparse_pm1(&r, n, or_sparse_threshold(), t_idx, t_sign);
    for (size_t j = 0; j < n; ++j) {
        u128 acc = 0;
        for (size_t k = 0; k < n1; ++k) {
            acc += (u128)l[j * n1 + k] * g[k];
            if ((k & 31) == 31) acc = reduce128(acc);
        }
        v_enc[j] = reduce128(acc);
    }
    for (size_t e = 0; e < t_nnz; ++e) {
        uint32_t j = t_idx[e];
        v_enc[j] = t_sign[e] == 1 ? or_add_mod(v_enc[j], 1) : or_sub_mod(v_enc[j], 1);
    }
    for (size_t j = 0; j < n; ++j) v_enc[j] = or_add_mod(v_enc[j], v[j]);
    return t_nnz;
}

/* ------------------------------------------------------------------------
 * Server product and Protocol-2 verification.
 * ---------------------------------------------------------------------- */
/* mersenne_matvec_scalar, crypto.rs:606-626 (u128 acc, fold every 32). */
void or_matvec(const uint64_t* mat, size_t rows, size_t cols, const uint64_t* vec,
               uint64_t* out) {
    for (size_t i = 0; i < rows; ++i) {
        const uint64_t* row = mat + i * cols;
        u128 acc = 0;
        for (size_t j = 0; j < cols; ++j) {
            acc += (u128)row[j] * vec[j];
            if ((j & 31) == 31) acc = reduce128(acc);
        }
        out[i] = reduce128(acc);
    }
}

/* mersenne_matvec_transpose, crypto.rs:753-769. */
void or_matvec_transpose(const uint64_t* mat, size_t rows, size_t cols, const uint64_t* vec,
                         uint64_t* out) {
    for (size_t j = 0; j < cols; ++j) {
        u128 acc = 0;
        for (size_t i = 0; i < rows; ++i) {
            acc += (u128)vec[i] * mat[i * cols + j];
            if ((i & 31) == 31) acc = reduce128(acc);
        }
        out[j] = reduce128(acc);
    }
}

/* compute_products_batch, crypto.rs:782-828: each row read once for all
 * queries; output per query (nvec x rows). */
void or_matvec_batch(const uint64_t* mat, size_t rows, size_t cols, const uint64_t* vecs,
                     size_t nvec, uint64_t* out) {
    u128* accs = (u128*)malloc((nvec ? nvec : 1) * sizeof(u128));
    for (size_t i = 0; i < rows; ++i) {
        const uint64_t* row = mat + i * cols;
        for (size_t q = 0; q < nvec; ++q) accs[q] = 0;
        for (size_t j = 0; j < cols; ++j) {
            u128 rj = row[j];
            for (size_t q = 0; q < nvec; ++q) {
                accs[q] += rj * vecs[q * cols + j];
                if ((j & 31) == 31) accs[q] = reduce128(accs[q]);
            }
        }
        for (size_t q = 0; q < nvec; ++q) out[q * rows + i] = reduce128(accs[q]);
    }
    free(accs);
}

/* dot_product, crypto.rs:1049-1059. */
uint64_t or_dot_product(const uint64_t* a, const uint64_t* b, size_t len) {
    u128 acc = 0;
    for (size_t i = 0; i < len; ++i) {
        acc += (u128)a[i] * b[i];
        if ((i & 31) == 31) acc = reduce128(acc);
    }
    return reduce128(acc);
}

/* verify_response, crypto.rs:1013-1046: lambda' Freivalds trials; each draws
 * u (m sample_fp), checks u.r == (u^T M_enc).v_enc; returns the first
 * failing trial or -1. */
int or_verify_response_rng(const uint64_t* r, const uint64_t* m_enc, size_t m, size_t n,
                           const uint64_t* v_enc, int trials, or_rng* rng) {
    uint64_t* u = (uint64_t*)malloc((m ? m : 1) * sizeof(uint64_t));
    uint64_t* z = (uint64_t*)malloc(n * sizeof(uint64_t));
    int result = -1;
    for (int t = 0; t < trials; ++t) {
        for (size_t i = 0; i < m; ++i) u[i] = or_sample_fp(rng);
        uint64_t ur = or_dot_product(u, r, m);
        or_matvec_transpose(m_enc, m, n, u, z);
        uint64_t zv = or_dot_product(z, v_enc, n);
        if (ur != zv) {
            result = t;
            break;
        }
    }
    free(u);
    free(z);
    return result;
}

int or_verify_response(const uint64_t* r, const uint64_t* m_enc, size_t m, size_t n,
                       const uint64_t* v_enc, int trials, const uint8_t rng_seed[32]) {
    or_rng rng;
    or_rng_seed(&rng, rng_seed);
    return or_verify_response_rng(r, m_enc, m, n, v_enc, trials, &rng);
}

/* ------------------------------------------------------------------------
 * Client decode, crypto.rs:901-940:
 *   Mv = r - (AL G + M T) - (H (L^T v_enc) + S v_enc)   (mod p)
 * dense_times_sparse :852-873, sparse_times_dense :877-894.
 * ---------------------------------------------------------------------- */
void or_decode_scores(const uint64_t* r, size_t m, size_t n, size_t n1,
                      const uint64_t* a_l, const uint64_t* m_plain, const uint64_t* h,
                      const uint32_t* s_rowptr, const uint32_t* s_idx, const int8_t* s_sign,
                      const uint64_t* l, const uint64_t* v_enc, const uint64_t* g,
                      const uint32_t* t_idx, const int8_t* t_sign, size_t t_nnz,
                      uint64_t* mv_out) {
    uint64_t* al_g = (uint64_t*)malloc((m ? m : 1) * 8);
    uint64_t* h_lt = (uint64_t*)malloc((m ? m : 1) * 8);
    uint64_t* lt = (uint64_t*)malloc(n1 * 8);
    or_matvec(a_l, m, n1, g, al_g);
    or_matvec_transpose(l, n, n1, v_enc, lt);
    or_matvec(h, m, n1, lt, h_lt);
    for (size_t i = 0; i < m; ++i) {
        uint64_t mt = 0;
        for (size_t e = 0; e < t_nnz; ++e) {
            uint64_t x = m_plain[i * n + t_idx[e]];
            mt = t_sign[e] == 1 ? or_add_mod(mt, x) : or_sub_mod(mt, x);
        }
        uint64_t sv = 0;
        for (uint32_t e = s_rowptr[i]; e < s_rowptr[i + 1]; ++e) {
            uint64_t x = v_enc[s_idx[e]];
            sv = s_sign[e] == 1 ? or_add_mod(sv, x) : or_sub_mod(sv, x);
        }
        uint64_t mvp = or_add_mod(al_g[i], mt);
        uint64_t mpv = or_add_mod(h_lt[i], sv);
        mv_out[i] = or_sub_mod(or_sub_mod(r[i], mvp), mpv);
    }
    free(al_g);
    free(h_lt);
    free(lt);
}

/* ------------------------------------------------------------------------
 * Top-k. The reference sorts (id, i64 score) with sort_unstable_by_key on
 * Reverse(score) and truncates (lib.rs:452-469; ivf.rs:806-820); its tie
 * order is unspecified. The oracle fixes ties to input order (stable), the
 * rule the GPU path implements too.
 * ---------------------------------------------------------------------- */
size_t or_topk(const uint32_t* ids, const int64_t* scores, size_t count, size_t k,
               uint32_t* out_ids, int64_t* out_scores) {
    size_t kk = k < count ? k : count;
    size_t filled = 0;
    for (size_t e = 0; e < count; ++e) {
        int64_t s = scores[e];
        if (filled == kk && (kk == 0 || s <= out_scores[kk - 1])) continue;
        size_t pos = filled < kk ? filled : kk - 1;
        /* shift down while the new score is strictly greater */
        while (pos > 0 && out_scores[pos - 1] < s) {
            out_scores[pos] = out_scores[pos - 1];
            out_ids[pos] = out_ids[pos - 1];
            --pos;
        }
        out_scores[pos] = s;
        out_ids[pos] = ids[e];
        if (filled < kk) ++filled;
    }
    return filled;
}

/* ------------------------------------------------------------------------
 * IVF routing: l2_f32_scalar (ivf-index/src/distance.rs:30-37) = sqrt of a
 * left-to-right f32 sum of squared differences (no FMA contraction: this
 * file is built with -ffp-contract=off); probe_route (ivf-index/src/ivf.rs:
 * 150-164) sorts ascending by distance under f32::total_cmp and keeps the
 * first nprobe. The reference's sort is unstable, so ties in distance have
 * no fixed order there; here they keep centroid order.
 * ---------------------------------------------------------------------- */
float or_l2_f32(const float* a, const float* b, size_t d) {
    float acc = -0.0f;
    for (size_t i = 0; i < d; ++i) {
        float diff = a[i] - b[i];
        acc = acc + diff * diff;
    }
    return sqrtf(acc);
}

/* f32::total_cmp order as an unsigned key (-inf < ... < -0 < +0 < ... <
 * +inf). Every NaN is taken as the largest value: the sign of a NaN
 * distance follows the platform's NaN propagation (x86 keeps an operand's
 * NaN, and its default NaN is negative), which the reference does not fix. */
static uint32_t total_key(float f) {
    uint32_t b;
    memcpy(&b, &f, 4);
    if (f != f) return 0xFFFFFFFFu;
    return (b & 0x80000000u) ? ~b : b | 0x80000000u;
}

void or_probe_route(const float* centroids, size_t n_centroids, size_t d, const float* query,
                    size_t nprobe, uint32_t* out) {
    if (nprobe > n_centroids) nprobe = n_centroids;
    uint32_t* best_d = (uint32_t*)malloc((nprobe ? nprobe : 1) * sizeof(uint32_t));
    size_t filled = 0;
    for (size_t c = 0; c < n_centroids; ++c) {
        uint32_t dist = total_key(or_l2_f32(query, centroids + c * d, d));
        if (filled == nprobe && (nprobe == 0 || dist >= best_d[nprobe - 1])) continue;
        size_t pos = filled < nprobe ? filled : nprobe - 1;
        while (pos > 0 && best_d[pos - 1] > dist) {
            best_d[pos] = best_d[pos - 1];
            out[pos] = out[pos - 1];
            --pos;
        }
        best_d[pos] = dist;
        out[pos] = (uint32_t)c;
        if (filled < nprobe) ++filled;
    }
    free(best_d);
}

/* ------------------------------------------------------------------------
 * EMVP score GEMV (scorer-emvp): fp_reduce single fold (crypto.rs:19-24),
 * fp_add (:27-30), compute_products (:200-222): partials[si*m + i] =
 * fold over l < b of fp_add(acc, fp_mul(m_hat[i*n + si*b + l], q[si*b + l])).
 * SEC128: n = 1292, s = 76, b = 17 (scorer-emvp/src/params.rs).
 * ---------------------------------------------------------------------- */
#define EMVP_N 1292
#define EMVP_S 76
#define EMVP_B 17

uint64_t or_emvp_fp_reduce(uint64_t lo, uint64_t hi) {
    u128 x = ((u128)hi << 64) | lo;
    uint64_t l = (uint64_t)(x & (u128)OR_P);
    uint64_t hh = (uint64_t)(x >> 61);
    uint64_t v = l + hh;
    return v >= OR_P ? v - OR_P : v;
}

static uint64_t emvp_mul(uint64_t a, uint64_t b) {
    u128 x = (u128)a * b;
    return or_emvp_fp_reduce((uint64_t)x, (uint64_t)(x >> 64));
}

void or_emvp_compute_products(const uint64_t* m_hat, size_t m, const uint64_t* q_hat,
                              uint64_t* partials) {
    for (size_t si = 0; si < EMVP_S; ++si) {
        const uint64_t* qb = q_hat + si * EMVP_B;
        for (size_t i = 0; i < m; ++i) {
            const uint64_t* row = m_hat + i * EMVP_N + si * EMVP_B;
            uint64_t acc = 0;
            for (size_t l = 0; l < EMVP_B; ++l) acc = or_add_mod(acc, emvp_mul(row[l], qb[l]));
            partials[si * m + i] = acc;
        }
    }
}

/* compute_products_batch, scorer-emvp/src/crypto.rs:369-415. */
void or_emvp_compute_products_batch(const uint64_t* m_hat, size_t m, const uint64_t* q_hats,
                                    size_t nq, uint64_t* partials) {
    for (size_t q = 0; q < nq; ++q)
        or_emvp_compute_products(m_hat, m, q_hats + q * EMVP_N, partials + q * EMVP_S * m);
}

/* ------------------------------------------------------------------------
 * EMVP client side (scorer-emvp/src/crypto.rs). SEC128: ell0 = 1024,
 * k = 268 (n - ell0), Q = 2^20 (params.rs).
 * ---------------------------------------------------------------------- */
#define EMVP_ELL0 1024
#define EMVP_K 268
#define EMVP_Q 1048576.0

static uint64_t emvp_neg(uint64_t a) { return a == 0 ? 0 : OR_P - a; } /* :38-41 */

/* quantize, crypto.rs:47-52: round(v * Q) as i64 (f64 round half away from
 * zero, saturating `as`), then ((vi % p) + p) % p with Rust's truncating %. */
uint64_t or_emvp_quantize(float v) {
    double rd = round((double)v * EMVP_Q);
    int64_t vi = rd != rd ? 0
                 : rd >= 9223372036854775808.0 ? INT64_MAX
                 : rd <= -9223372036854775808.0 ? INT64_MIN
                 : (int64_t)rd;
    int64_t p = (int64_t)OR_P;
    return (uint64_t)(((vi % p) + p) % p);
}

/* decode_field, crypto.rs:54-62: signed lift (v > p/2 is negative), then
 * (signed as f64 / (Q as f64 * Q as f64)) as f32. */
float or_emvp_decode_field(uint64_t v) {
    int64_t s = v > OR_P / 2 ? (int64_t)v - (int64_t)OR_P : (int64_t)v;
    return (float)((double)s / (EMVP_Q * EMVP_Q));
}

/* next_u64 % P of a ChaCha20Rng (rng.random::<u64>() % P). */
static uint64_t emvp_draw(or_rng* r) { return or_next_u64(r) % OR_P; }

/* generate_d, crypto.rs:72-97: D = [I_ell0 | D'] (ell0 x n, row-major); D'
 * row by row from ChaCha20Rng(key_seed). */
void or_emvp_generate_d(const uint8_t key_seed[32], uint64_t* d) {
    memset(d, 0, sizeof(uint64_t) * EMVP_ELL0 * EMVP_N);
    for (size_t i = 0; i < EMVP_ELL0; ++i) d[i * EMVP_N + i] = 1;
    or_rng r;
    or_rng_seed(&r, key_seed);
    for (size_t i = 0; i < EMVP_ELL0; ++i)
        for (size_t j = 0; j < EMVP_K; ++j) d[i * EMVP_N + EMVP_ELL0 + j] = emvp_draw(&r);
}

/* row_rng, crypto.rs:100-107: r_seed with its first 8 bytes xor the row
 * index (u64 LE). */
static void emvp_row_rng(or_rng* r, const uint8_t r_seed[32], uint64_t row) {
    uint8_t seed[32];
    memcpy(seed, r_seed, 32);
    for (int i = 0; i < 8; ++i) seed[i] ^= (uint8_t)(row >> (8 * i));
    or_rng_seed(r, seed);
}

/* encrypt_cluster, crypto.rs:124-163: M_hat[i][j] = fp_add(fold_l fp_add(acc,
 * fp_mul(M0[i][l], D[l][j])), R[i][j]) with M0 = quantize(vectors[i]) zero-
 * padded to ell0 (d floats per row, the first min(d, ell0) used) and R[i]
 * drawn from row_rng(r_seed, i) in column order. */
void or_emvp_encrypt_cluster(const uint64_t* d_mat, const uint8_t r_seed[32], const float* vecs,
                             size_t m, size_t d, uint64_t* m_hat) {
    uint64_t* m0 = (uint64_t*)malloc(sizeof(uint64_t) * EMVP_ELL0);
    size_t take = d < EMVP_ELL0 ? d : EMVP_ELL0;
    for (size_t i = 0; i < m; ++i) {
        for (size_t l = 0; l < EMVP_ELL0; ++l) m0[l] = l < take ? or_emvp_quantize(vecs[i * d + l]) : 0;
        or_rng r;
        emvp_row_rng(&r, r_seed, (uint64_t)i);
        for (size_t j = 0; j < EMVP_N; ++j) {
            uint64_t dot = 0;
            for (size_t l = 0; l < EMVP_ELL0; ++l)
                dot = or_add_mod(dot, emvp_mul(m0[l], d_mat[l * EMVP_N + j]));
            m_hat[i * EMVP_N + j] = or_add_mod(dot, emvp_draw(&r));
        }
    }
    free(m0);
}

/* encode_query, crypto.rs:171-190: q_hat = (q - D' u | u) with
 * q[i] = quantize(query[i] as f32). */
void or_emvp_encode_query(const uint64_t* d_mat, const double* query, const uint64_t* u,
                          uint64_t* q_hat) {
    for (size_t i = 0; i < EMVP_ELL0; ++i) {
        uint64_t dp_u = 0;
        for (size_t j = 0; j < EMVP_K; ++j)
            dp_u = or_add_mod(dp_u, emvp_mul(d_mat[i * EMVP_N + EMVP_ELL0 + j], u[j]));
        q_hat[i] = or_add_mod(or_emvp_quantize((float)query[i]), emvp_neg(dp_u));
    }
    for (size_t j = 0; j < EMVP_K; ++j) q_hat[EMVP_ELL0 + j] = u[j];
}

/* decode_scores, crypto.rs:435-453: score[i] = decode_field(fp_add(
 * sum_si partials[si*m + i], fp_neg(R[i] . q_hat))), R[i] replayed from
 * row_rng(r_seed, i). */
void or_emvp_decode_scores(size_t m, const uint8_t r_seed[32], const uint64_t* q_hat,
                           const uint64_t* partials, float* scores) {
    for (size_t i = 0; i < m; ++i) {
        uint64_t block_sum = 0;
        for (size_t si = 0; si < EMVP_S; ++si) block_sum = or_add_mod(block_sum, partials[si * m + i]);
        or_rng r;
        emvp_row_rng(&r, r_seed, (uint64_t)i);
        uint64_t r_dot = 0;
        for (size_t j = 0; j < EMVP_N; ++j) r_dot = or_add_mod(r_dot, emvp_mul(emvp_draw(&r), q_hat[j]));
        scores[i] = or_emvp_decode_field(or_add_mod(block_sum, emvp_neg(r_dot)));
    }
}

/* ------------------------------------------------------------------------
 * Tiptoe SimplePIR server answer, scorer-tiptoe/src/pir/simplepir.rs:191-204:
 * ans[i] = wrapping sum_j m_db[i*c + j] * ct[j] (mod 2^64). The mod-2^32
 * variant is the same sum truncated (2^32 divides 2^64).
 * ---------------------------------------------------------------------- */
void or_tiptoe_server_answer_u64(const uint64_t* m_db, size_t r, size_t c, const uint64_t* ct,
                                 uint64_t* ans) {
    for (size_t i = 0; i < r; ++i) {
        uint64_t acc = 0;
        for (size_t j = 0; j < c; ++j) acc += m_db[i * c + j] * ct[j];
        ans[i] = acc;
    }
}

void or_tiptoe_server_answer_u32(const uint32_t* m_db, size_t r, size_t c, const uint32_t* ct,
                                 uint32_t* ans) {
    for (size_t i = 0; i < r; ++i) {
        uint32_t acc = 0;
        for (size_t j = 0; j < c; ++j) acc += m_db[i * c + j] * ct[j];
        ans[i] = acc;
    }
}

/* ------------------------------------------------------------------------
 * BnTmScorer::score (scorer-bntm/src/lib.rs:326-475) with q_seed injected:
 * encode -> compute_products -> [verify with domain_separate(q_seed,
 * "protocol2")] -> decode -> fp_to_signed -> top-k (k capped at m).
 * ---------------------------------------------------------------------- */
int or_bntm_flat_score(const uint64_t* m_enc, const uint64_t* a_l, const uint64_t* m_plain,
                       const uint64_t* h, const uint32_t* s_rowptr, const uint32_t* s_idx,
                       const int8_t* s_sign, const uint64_t* l, size_t m, const uint64_t* v_q,
                       const uint8_t q_seed[32], int verify, size_t k, uint32_t* out_ids,
                       int64_t* out_scores, size_t* out_count) {
    const size_t n = 1024, n1 = 128;
    uint64_t v_enc[1024], g[128];
    uint32_t t_idx[1024];
    int8_t t_sign[1024];
    size_t t_nnz = or_encode_query(v_q, l, n, n1, q_seed, v_enc, g, t_idx, t_sign);
    uint64_t* r = (uint64_t*)malloc((m ? m : 1) * 8);
    uint64_t* mv = (uint64_t*)malloc((m ? m : 1) * 8);
    or_matvec(m_enc, m, n, v_enc, r);
    if (verify) {
        uint8_t vs[32];
        or_domain_separate(q_seed, (const uint8_t*)"protocol2", 9, vs);
        int t = or_verify_response(r, m_enc, m, n, v_enc, 3, vs);
        if (t >= 0) {
            free(r);
            free(mv);
            *out_count = 0;
            return 1 + t;
        }
    }
    or_decode_scores(r, m, n, n1, a_l, m_plain, h, s_rowptr, s_idx, s_sign, l, v_enc, g, t_idx,
                     t_sign, t_nnz, mv);
    uint32_t* ids = (uint32_t*)malloc((m ? m : 1) * 4);
    int64_t* sc = (int64_t*)malloc((m ? m : 1) * 8);
    for (size_t i = 0; i < m; ++i) {
        ids[i] = (uint32_t)i;
        sc[i] = or_fp_to_signed(mv[i]);
    }
    *out_count = or_topk(ids, sc, m, k, out_ids, out_scores);
    free(ids);
    free(sc);
    free(r);
    free(mv);
    return 0;
}

/* ------------------------------------------------------------------------
 * Row-parallel CPU baseline of mersenne_matvec_scalar (rayon over rows,
 * crypto.rs:614-625) with pthreads.
 * ---------------------------------------------------------------------- */
typedef struct {
    const uint64_t* mat;
    const uint64_t* vec;
    uint64_t* out;
    size_t r0, r1, cols;
} mv_job;

static void* mv_worker(void* p) {
    mv_job* j = (mv_job*)p;
    or_matvec(j->mat + j->r0 * j->cols, j->r1 - j->r0, j->cols, j->vec, j->out + j->r0);
    return NULL;
}

void or_matvec_mt(const uint64_t* mat, size_t rows, size_t cols, const uint64_t* vec,
                  uint64_t* out, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    mv_job jobs[256];
    size_t chunk = (rows + (size_t)nthreads - 1) / (size_t)nthreads;
    int started = 0;
    for (int t = 0; t < nthreads; ++t) {
        size_t r0 = (size_t)t * chunk, r1 = r0 + chunk;
        if (r0 >= rows) break;
        if (r1 > rows) r1 = rows;
        jobs[t] = (mv_job){mat, vec, out, r0, r1, cols};
        pthread_create(&th[t], NULL, mv_worker, &jobs[t]);
        ++started;
    }
    for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
}
