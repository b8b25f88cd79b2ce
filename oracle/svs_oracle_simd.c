/*
 * svs_oracle_simd.c — the reference CPU path as it is timed: the AVX-512
 * lane-decomposed matvec, and the closed-loop one-query IVF score with
 * row-parallel threads.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ (as a checker) and by
 * bench.py's cpu_baseline leg (as the timed CPU baseline). The product
 * library never links or calls it.
 *
 * - or_matvec_avx512 restates mersenne_matvec_avx512 / avx512_row_dot /
 *   avx512_fold (scorer-bntm/src/crypto.rs:643-739): each u64 x u64 is split
 *   into four u32 x u32 -> u64 products with _mm512_mul_epu32 over 8 lanes;
 *   ll and lh+hl partials are folded (x & P) + (x >> 61) after every chunk,
 *   the hh partials every 8 chunks; the three lane sums are reduced and
 *   recombined as r_ll + (r_cross << 32) + 8 r_hh (2^64 = 8 mod p). The
 *   dispatcher (crypto.rs:584-598) uses it only when cols % 8 == 0; the
 *   same rule applies here, with the scalar row (crypto.rs:606-626)
 *   otherwise. Pinned by the reference's own gate
 *   (mersenne_matvec_avx512_matches_scalar, crypto.rs:1184-1198): equal to
 *   the scalar matvec, at its shapes and at all-(p-1) extremes.
 * - or_ivf_rows_mt is the per-query server + client work of
 *   BnTmIvfScorer::score (ivf.rs:746-822) over the probed lists, verification
 *   off: compute_products (crypto.rs:560-562) and decode_scores
 *   (crypto.rs:901-940) for every row, rows split across threads as rayon's
 *   par_iter over rows splits them (crypto.rs:609-625). Per row it computes
 *   r = M_enc . v_enc, A_L . G, H . (L^T v_enc), M_plain . T and S . v_enc
 *   and writes fp_to_signed(r - (A_L G + M T) - (H L^T v + S v)), the
 *   decode_scores formula; the row-wise fusion of decode's separate passes
 *   changes no value (every term is computed per row).
 */
#include <immintrin.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>

#include "svs_oracle.h"

typedef unsigned __int128 u128;

static uint64_t red128(u128 x) {
    u128 s1 = (x & (u128)OR_P) + (x >> 61);
    uint64_t r = (uint64_t)(s1 & (u128)OR_P) + (uint64_t)(s1 >> 61);
    if (r >= OR_P) r -= OR_P;
    if (r >= OR_P) r -= OR_P;
    return r;
}

int or_has_avx512f(void) {
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx512f") ? 1 : 0;
}

/* mersenne_matvec_scalar's row (crypto.rs:612-622): u128 accumulator,
 * folded every 32 products. */
static uint64_t row_dot_scalar(const uint64_t* row, const uint64_t* v, size_t cols) {
    u128 acc = 0;
    for (size_t j = 0; j < cols; ++j) {
        acc += (u128)row[j] * v[j];
        if ((j & 31) == 31) acc = red128(acc);
    }
    return red128(acc);
}

/* avx512_fold (crypto.rs:669-680): one Mersenne fold per lane. */
__attribute__((target("avx512f"))) static inline __m512i fold8(__m512i x, __m512i mask) {
    return _mm512_add_epi64(_mm512_and_si512(x, mask), _mm512_srli_epi64(x, 61));
}

/* avx512_row_dot (crypto.rs:682-739); cols % 8 == 0. */
__attribute__((target("avx512f"))) static uint64_t row_dot_avx512(const uint64_t* row,
                                                                  const uint64_t* v,
                                                                  size_t cols) {
    const __m512i mask = _mm512_set1_epi64((long long)OR_P);
    __m512i ll = _mm512_setzero_si512();
    __m512i cross = _mm512_setzero_si512();
    __m512i hh = _mm512_setzero_si512();
    for (size_t j = 0; j < cols;) {
        __m512i a = _mm512_loadu_si512((const void*)(row + j));
        __m512i b = _mm512_loadu_si512((const void*)(v + j));
        __m512i ah = _mm512_srli_epi64(a, 32);
        __m512i bh = _mm512_srli_epi64(b, 32);
        __m512i p_ll = _mm512_mul_epu32(a, b);
        __m512i p_x = _mm512_add_epi64(_mm512_mul_epu32(a, bh), _mm512_mul_epu32(ah, b));
        __m512i p_hh = _mm512_mul_epu32(ah, bh);
        ll = fold8(_mm512_add_epi64(ll, fold8(p_ll, mask)), mask);
        cross = fold8(_mm512_add_epi64(cross, fold8(p_x, mask)), mask);
        hh = _mm512_add_epi64(hh, p_hh);
        j += 8;
        if ((j / 8) % 8 == 0) hh = fold8(hh, mask);
    }
    uint64_t bl[8], bx[8], bh8[8];
    _mm512_storeu_si512((void*)bl, ll);
    _mm512_storeu_si512((void*)bx, cross);
    _mm512_storeu_si512((void*)bh8, hh);
    u128 sl = 0, sx = 0, sh = 0;
    for (int k = 0; k < 8; ++k) {
        sl += bl[k];
        sx += bx[k];
        sh += bh8[k];
    }
    u128 total = (u128)red128(sl) + ((u128)red128(sx) << 32) + (u128)red128(sh) * 8;
    return red128(total);
}

static uint64_t row_dot(const uint64_t* row, const uint64_t* v, size_t cols, int simd) {
    if (simd && cols > 0 && cols % 8 == 0) return row_dot_avx512(row, v, cols);
    return row_dot_scalar(row, v, cols);
}

/* mersenne_matvec with the AVX-512 kernel where the dispatcher would take
 * it. Returns -1 (nothing written) when the CPU lacks AVX-512F. */
int or_matvec_avx512(const uint64_t* mat, size_t rows, size_t cols, const uint64_t* vec,
                     uint64_t* out) {
    if (!or_has_avx512f()) return -1;
    for (size_t i = 0; i < rows; ++i) out[i] = row_dot(mat + i * cols, vec, cols, 1);
    return 0;
}

/* ------------------------------------------------------------------------
 * One query over its probed lists, rows split across threads.
 * ---------------------------------------------------------------------- */
typedef struct {
    const or_cluster* cl;
    size_t ncl, n, n1;
    const uint64_t *v_enc, *g, *lt;
    const uint32_t* t_idx;
    const int8_t* t_sign;
    size_t t_nnz;
    int simd, nthreads, tid;
    int64_t* out;
    const size_t* base;
} rows_job;

static void* rows_worker(void* arg) {
    rows_job* j = (rows_job*)arg;
    for (size_t c = 0; c < j->ncl; ++c) {
        const or_cluster* cl = &j->cl[c];
        size_t m = (size_t)cl->m;
        size_t chunk = (m + (size_t)j->nthreads - 1) / (size_t)j->nthreads;
        size_t r0 = (size_t)j->tid * chunk, r1 = r0 + chunk;
        if (r1 > m) r1 = m;
        int64_t* out = j->out + j->base[c];
        for (size_t i = r0; i < r1; ++i) {
            uint64_t r = row_dot(cl->m_enc + i * j->n, j->v_enc, j->n, j->simd);
            uint64_t al_g = row_dot(cl->a_l + i * j->n1, j->g, j->n1, j->simd);
            uint64_t h_lt = row_dot(cl->h + i * j->n1, j->lt, j->n1, j->simd);
            uint64_t mt = 0;
            const uint64_t* mp = cl->m_plain + i * j->n;
            for (size_t e = 0; e < j->t_nnz; ++e) {
                uint64_t x = mp[j->t_idx[e]];
                mt = j->t_sign[e] == 1 ? or_add_mod(mt, x) : or_sub_mod(mt, x);
            }
            uint64_t sv = 0;
            for (uint32_t e = cl->s_rowptr[i]; e < cl->s_rowptr[i + 1]; ++e) {
                uint64_t x = j->v_enc[cl->s_idx[e]];
                sv = cl->s_sign[e] == 1 ? or_add_mod(sv, x) : or_sub_mod(sv, x);
            }
            uint64_t mv = or_sub_mod(or_sub_mod(r, or_add_mod(al_g, mt)), or_add_mod(h_lt, sv));
            out[i] = or_fp_to_signed(mv);
        }
    }
    return NULL;
}

/* Writes the signed decoded score of every row of the ncl lists, list after
 * list (sum of m entries). simd = 1 takes the AVX-512 rows; returns -1 if
 * the CPU lacks AVX-512F, else 0. */
int or_ivf_rows_mt(const or_cluster* cl, size_t ncl, size_t n, size_t n1, const uint64_t* v_enc,
                   const uint64_t* g, const uint64_t* lt, const uint32_t* t_idx,
                   const int8_t* t_sign, size_t t_nnz, int nthreads, int simd,
                   int64_t* scores_out) {
    if (simd && !or_has_avx512f()) return -1;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    size_t* base = (size_t*)malloc((ncl ? ncl : 1) * sizeof(size_t));
    size_t acc = 0;
    for (size_t c = 0; c < ncl; ++c) {
        base[c] = acc;
        acc += (size_t)cl[c].m;
    }
    pthread_t th[256];
    rows_job jobs[256];
    for (int t = 0; t < nthreads; ++t) {
        jobs[t] = (rows_job){cl, ncl, n, n1, v_enc, g, lt, t_idx, t_sign, t_nnz, simd,
                             nthreads, t, scores_out, base};
        if (t > 0) pthread_create(&th[t], NULL, rows_worker, &jobs[t]);
    }
    rows_worker(&jobs[0]);
    for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
    free(base);
    return 0;
}
