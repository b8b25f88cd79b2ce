/*
 * svs_oracle.h — CPU oracle for the encrypted-score matrix-vector path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference
 * (sacs-epfl/secure-vector-search, Rust) algorithms on the hot path. Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
 * and only as the checker (or as the timed CPU baseline). The product library
 * never links or calls it.
 *
 * Every function cites the reference file:line it restates. Pinning status
 * (what checks each function against something other than itself) is listed
 * per function below. The ChaCha20Rng stream (rand_chacha 0.10, a
 * dependency the reference does not vendor) is pinned against the RFC 7539
 * A.1 keystreams the way rand_chacha's own tests use them: seed = key,
 * blocks in counter order, u32 words in order. next_u64 taking the low word
 * first (rand_core's BlockRng) and random::<u8>() = next_u32() truncated
 * (rand's StandardUniform) restate those crates' documented behaviour; no
 * reference file holds a ChaCha20Rng output to pin them further.
 */
#ifndef SVS_ORACLE_H
#define SVS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_P ((uint64_t)0x1FFFFFFFFFFFFFFFULL) /* 2^61 - 1, params.rs:10 */

/* ---- field arithmetic (crypto.rs:47-89) -------------------------------- */
/* Pinned: reference KATs crypto.rs:1079-1101 + big-int golden vectors. */
uint64_t or_mersenne_reduce(uint64_t lo, uint64_t hi);
uint64_t or_add_mod(uint64_t a, uint64_t b);
uint64_t or_sub_mod(uint64_t a, uint64_t b);
uint64_t or_mul_mod(uint64_t a, uint64_t b);

/* ---- ChaCha20 stream with rand_chacha conventions ---------------------- */
/* Block function pinned by RFC 7539 A.1 test vectors (key 0, nonce 0). */
typedef struct {
    uint32_t key[8];
    uint64_t block;     /* next block counter to generate */
    uint32_t buf[16];
    int idx;            /* next unread word in buf (16 = empty) */
} or_rng;

void or_chacha20_block(const uint32_t key[8], uint64_t counter, uint32_t out[16]);
void or_rng_seed(or_rng* r, const uint8_t seed[32]);
uint32_t or_next_u32(or_rng* r);
uint64_t or_next_u64(or_rng* r);

/* ---- sampling / seeds (crypto.rs:142-223) ------------------------------ */
uint64_t or_sample_fp(or_rng* r);                                    /* :155-162 */
void or_domain_separate(const uint8_t seed[32], const uint8_t* tag, size_t tag_len,
                        uint8_t out[32]);                            /* :142-150 */
uint32_t or_sparse_threshold(void);                                  /* :211-212 + params.rs:242-244 */
/* sample_sparse_pm1 (:207-223): writes up to n entries; returns count. */
size_t or_sample_sparse_pm1(or_rng* r, size_t n, uint32_t threshold,
                            uint32_t* idx, int8_t* sign);

/* ---- encoding helpers (crypto.rs:105-135) ------------------------------ */
void or_quantise_vector(const float* v, size_t d, size_t n, uint64_t q, uint64_t* out);
int64_t or_fp_to_signed(uint64_t s);

/* ---- keygen / encryption (crypto.rs:235-478) --------------------------- */
void or_generate_subspace(const uint8_t key_seed[32], size_t n, size_t n1, uint64_t* l_out);
/* regenerate_h_s (:258-270). s_rowptr has m+1 entries; s_idx/s_sign hold
 * up to s_cap entries. Returns total nnz or (size_t)-1 if s_cap too small. */
size_t or_regenerate_h_s(const uint8_t h_seed[32], size_t m, size_t n, size_t n1,
                         uint64_t* h_out, uint32_t* s_rowptr, uint32_t* s_idx,
                         int8_t* s_sign, size_t s_cap);
/* encrypt_cluster (:405-478) given H, S already regenerated. */
void or_encrypt_cluster(const uint64_t* m_plain, size_t m, size_t n, size_t n1,
                        const uint64_t* l, const uint64_t* h, const uint32_t* s_rowptr,
                        const uint32_t* s_idx, const int8_t* s_sign, uint64_t* m_enc_out,
                        uint64_t* a_l_out);

/* encode_query (:508-553). Writes v_enc (n), g (n1) and T (returns nnz). */
size_t or_encode_query(const uint64_t* v, const uint64_t* l, size_t n, size_t n1,
                       const uint8_t q_seed[32], uint64_t* v_enc, uint64_t* g,
                       uint32_t* t_idx, int8_t* t_sign);

/* ---- server product / verification (crypto.rs:606-1059) ---------------- */
void or_matvec(const uint64_t* mat, size_t rows, size_t cols, const uint64_t* vec,
               uint64_t* out);                                       /* :606-626 */
void or_matvec_transpose(const uint64_t* mat, size_t rows, size_t cols,
                         const uint64_t* vec, uint64_t* out);        /* :753-769 */
void or_matvec_batch(const uint64_t* mat, size_t rows, size_t cols, const uint64_t* vecs,
                     size_t nvec, uint64_t* out /* nvec x rows */);  /* :782-828 */
uint64_t or_dot_product(const uint64_t* a, const uint64_t* b, size_t len); /* :1049-1059 */
/* verify_response (:1013-1046): returns -1 on accept, failing trial index. */
int or_verify_response(const uint64_t* r, const uint64_t* m_enc, size_t m, size_t n,
                       const uint64_t* v_enc, int trials, const uint8_t rng_seed[32]);
/* Same, with the rng passed in (the IVF path re-seeds per cluster). */
int or_verify_response_rng(const uint64_t* r, const uint64_t* m_enc, size_t m, size_t n,
                           const uint64_t* v_enc, int trials, or_rng* rng);

/* decode_scores (:901-940). */
void or_decode_scores(const uint64_t* r, size_t m, size_t n, size_t n1,
                      const uint64_t* a_l, const uint64_t* m_plain, const uint64_t* h,
                      const uint32_t* s_rowptr, const uint32_t* s_idx, const int8_t* s_sign,
                      const uint64_t* l, const uint64_t* v_enc, const uint64_t* g,
                      const uint32_t* t_idx, const int8_t* t_sign, size_t t_nnz,
                      uint64_t* mv_out);

/* ---- top-k (lib.rs:452-469, ivf.rs:806-820) ---------------------------- */
/* Descending signed score; ties keep input order (documented tie rule). */
size_t or_topk(const uint32_t* ids, const int64_t* scores, size_t count, size_t k,
               uint32_t* out_ids, int64_t* out_scores);

/* ---- IVF routing (ivf-index/src/ivf.rs:150-164, distance.rs:30-37) ----- */
float or_l2_f32(const float* a, const float* b, size_t d);
void or_probe_route(const float* centroids, size_t n_centroids, size_t d, const float* query,
                    size_t nprobe, uint32_t* out);

/* ---- EMVP score GEMV (scorer-emvp/src/crypto.rs:19-40,200-222,369-415) -- */
uint64_t or_emvp_fp_reduce(uint64_t lo, uint64_t hi);
void or_emvp_compute_products(const uint64_t* m_hat, size_t m, const uint64_t* q_hat,
                              uint64_t* partials /* s x m */);
void or_emvp_compute_products_batch(const uint64_t* m_hat, size_t m, const uint64_t* q_hats,
                                    size_t nq, uint64_t* partials /* nq x s x m */);
/* ---- EMVP client side (scorer-emvp/src/crypto.rs:47-62,72-190,435-453) --
 * Pinned by the reference's encode_decode_single_vector (crypto.rs:509-542)
 * and by the exact round trip decode(compute(encrypt(M0), encode(q))) =
 * fl32(signed(M0 . q mod p) / 2^40). */
uint64_t or_emvp_quantize(float v);
float or_emvp_decode_field(uint64_t v);
void or_emvp_generate_d(const uint8_t key_seed[32], uint64_t* d);
void or_emvp_encrypt_cluster(const uint64_t* d_mat, const uint8_t r_seed[32], const float* vecs,
                             size_t m, size_t d, uint64_t* m_hat);
void or_emvp_encode_query(const uint64_t* d_mat, const double* query, const uint64_t* u,
                          uint64_t* q_hat);
void or_emvp_decode_scores(size_t m, const uint8_t r_seed[32], const uint64_t* q_hat,
                           const uint64_t* partials, float* scores);

/* ---- Tiptoe SimplePIR answer (scorer-tiptoe/src/pir/simplepir.rs:191-204) */
void or_tiptoe_server_answer_u64(const uint64_t* m_db, size_t r, size_t c, const uint64_t* ct,
                                 uint64_t* ans);
void or_tiptoe_server_answer_u32(const uint32_t* m_db, size_t r, size_t c, const uint32_t* ct,
                                 uint32_t* ans);

/* ---- whole-query orchestration --------------------------------------- */
/* BnTmScorer::score (lib.rs:326-475) on one flat cluster, with q_seed
 * injected. Returns 0 on success, 1 + trial on verification failure.
 * Writes min(k, m) hits. */
int or_bntm_flat_score(const uint64_t* m_enc, const uint64_t* a_l, const uint64_t* m_plain,
                       const uint64_t* h, const uint32_t* s_rowptr, const uint32_t* s_idx,
                       const int8_t* s_sign, const uint64_t* l, size_t m, const uint64_t* v_q,
                       const uint8_t q_seed[32], int verify, size_t k, uint32_t* out_ids,
                       int64_t* out_scores, size_t* out_count);

/* Multi-threaded CPU baseline of the server matvec (rows split across
 * nthreads, mirroring rayon's row parallelism in crypto.rs:614-625). */
void or_matvec_mt(const uint64_t* mat, size_t rows, size_t cols, const uint64_t* vec,
                  uint64_t* out, int nthreads);

/* ---- CPU baseline (svs_oracle_simd.c) ----------------------------------- */
typedef struct {
    const uint64_t* m_enc;    /* m x n */
    const uint64_t* a_l;      /* m x n1 */
    const uint64_t* m_plain;  /* m x n */
    const uint64_t* h;        /* m x n1 */
    const uint32_t* s_rowptr; /* m + 1 */
    const uint32_t* s_idx;
    const int8_t* s_sign;
    uint64_t m;
} or_cluster;

int or_has_avx512f(void);
/* mersenne_matvec_avx512, crypto.rs:643-739 (-1 without AVX-512F). */
int or_matvec_avx512(const uint64_t* mat, size_t rows, size_t cols, const uint64_t* vec,
                     uint64_t* out);
/* compute_products + decode_scores of one query over its probed lists,
 * rows split across nthreads (ivf.rs:746-822, crypto.rs:609-625). */
int or_ivf_rows_mt(const or_cluster* cl, size_t ncl, size_t n, size_t n1, const uint64_t* v_enc,
                   const uint64_t* g, const uint64_t* lt, const uint32_t* t_idx,
                   const int8_t* t_sign, size_t t_nnz, int nthreads, int simd,
                   int64_t* scores_out);

/* ---- Tiptoe BFV compression (svs_oracle_bfv.c; compress.rs) ------------ */
uint64_t or_bfv_modulus(void);
uint64_t or_bfv_psi(void);
void or_bfv_ntt(uint64_t* a);
void or_bfv_intt(uint64_t* a);
void or_bfv_negacyclic_naive(const uint64_t* a, const uint64_t* b, uint64_t* out);
void or_bfv_keygen(const uint8_t seed[32], uint64_t* s_ntt);
void or_bfv_encrypt(const uint64_t* s_ntt, const uint64_t* m, const uint8_t seed[32],
                    uint64_t* c0, uint64_t* c1);
void or_bfv_decrypt(const uint64_t* s_ntt, const uint64_t* c0, const uint64_t* c1, uint64_t* m);
uint64_t or_tiptoe_get_chunk(uint64_t v, size_t idx);
int64_t or_tiptoe_from_modulo_t(uint64_t t, uint64_t v);
void or_tiptoe_decompose_hint(const uint64_t* hint, size_t r, size_t n_lwe, uint64_t* pt);
int or_tiptoe_encrypt_secret(const uint8_t* s, size_t n_lwe, const uint64_t* sk_ntt,
                             const uint8_t seed[32], uint64_t* ct);
void or_tiptoe_apply_hint(const uint64_t* pt, size_t chunks, size_t n_lwe, const uint64_t* ct,
                          uint64_t* out);
void or_tiptoe_recover_h_s(const uint64_t* answers, size_t r, const uint64_t* sk_ntt,
                           uint64_t* h_s);

#ifdef __cplusplus
}
#endif
#endif
