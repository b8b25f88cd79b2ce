/*
 * svs_oracle_bfv.c — Tiptoe's BFV second-layer compression on the oracle
 * (scorer-tiptoe/src/bfv/compress.rs): decompose_hint (:146-190),
 * encrypt_secret_under_bfv (:201-227), apply_hint (:241-266), recover_h_s
 * (:277-318), with the BFV they run on.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker. The product
 * library never links or calls it.
 *
 * The reference takes BFV from the `fhe` crate 0.1.1 (tiptoe_text_params,
 * bfv/mod.rs:50-70: degree N = 2048, plaintext modulus t = 65537, one 38-bit
 * ciphertext modulus), which is not vendored. This file restates textbook
 * BFV at those parameters, with these choices of its own:
 *   q    = the largest prime below 2^38 with q = 1 (mod 2N), NTT-friendly;
 *   psi  = g^((q-1)/2N) for g the smallest primitive root mod q;
 *   NTT  = negacyclic Cooley-Tukey, natural order in, bit-reversed order
 *          out (twiddles psi^bitrev(i)); pointwise products in this domain
 *          are products in Z_q[X]/(X^N + 1) whatever the slot order;
 *   keys = ternary secret s, secret-key encryption c1 = a uniform,
 *          c0 = -a s + e + floor(q/t) m, e centred binomial (eta = 10);
 *   decrypt m = round(t (c0 + c1 s) / q) mod t.
 * Ciphertexts are held in the NTT domain, as the plaintext x ciphertext
 * products of apply_hint use them. Ciphertext bytes therefore are not
 * comparable with the fhe crate's; parity is pinned at the plaintext level
 * by the reference's own invariants (compress.rs:343-476): the limb
 * decomposition recombines to the top 32 bits, and decrypting apply_hint's
 * output recovers the top 32 bits of H s.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "svs_oracle.h"

typedef unsigned __int128 u128;

#define BN 2048
#define BLOG 11
#define BT 65537ull
#define LIMB_BITS 4
#define MAX_LIMBS 16 /* MAX_LIMBS_64, compress.rs:58 */
#define NUM_LIMBS 8  /* NUM_LIMBS_64, compress.rs:62 */

static uint64_t powmod(uint64_t a, uint64_t e, uint64_t m) {
    u128 r = 1, b = a % m;
    while (e) {
        if (e & 1) r = (r * b) % m;
        b = (b * b) % m;
        e >>= 1;
    }
    return (uint64_t)r;
}

/* Deterministic Miller-Rabin for 64-bit n (bases 2..37). */
static int is_prime(uint64_t n) {
    static const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return 0;
    for (int i = 0; i < 12; ++i) {
        if (n == bases[i]) return 1;
        if (n % bases[i] == 0) return 0;
    }
    uint64_t d = n - 1;
    int r = 0;
    while ((d & 1) == 0) {
        d >>= 1;
        ++r;
    }
    for (int i = 0; i < 12; ++i) {
        uint64_t x = powmod(bases[i], d, n);
        if (x == 1 || x == n - 1) continue;
        int comp = 1;
        for (int k = 1; k < r; ++k) {
            x = (uint64_t)(((u128)x * x) % n);
            if (x == n - 1) {
                comp = 0;
                break;
            }
        }
        if (comp) return 0;
    }
    return 1;
}

typedef struct {
    uint64_t q, psi, n_inv, delta;
    uint64_t psi_br[BN], psi_inv_br[BN];
    int ready;
} bfv_ctx;

static bfv_ctx G;

static uint32_t bitrev(uint32_t x, int bits) {
    uint32_t r = 0;
    for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}

static void bfv_init(void) {
    if (G.ready) return;
    const uint64_t m2 = 2 * BN;
    uint64_t q = (((1ull << 38) - 2) / m2) * m2 + 1;
    while (!is_prime(q)) q -= m2;
    /* smallest primitive root: g^((q-1)/f) != 1 for each prime f | q - 1 */
    uint64_t fac[64];
    int nf = 0;
    uint64_t rest = q - 1;
    for (uint64_t f = 2; f * f <= rest; ++f)
        if (rest % f == 0) {
            fac[nf++] = f;
            while (rest % f == 0) rest /= f;
        }
    if (rest > 1) fac[nf++] = rest;
    uint64_t g = 2;
    for (;; ++g) {
        int ok = 1;
        for (int i = 0; i < nf && ok; ++i) ok = powmod(g, (q - 1) / fac[i], q) != 1;
        if (ok) break;
    }
    G.q = q;
    G.psi = powmod(g, (q - 1) / m2, q);
    uint64_t psi_inv = powmod(G.psi, q - 2, q);
    for (uint32_t i = 0; i < BN; ++i) {
        uint32_t e = bitrev(i, BLOG);
        G.psi_br[i] = powmod(G.psi, e, q);
        G.psi_inv_br[i] = powmod(psi_inv, e, q);
    }
    G.n_inv = powmod(BN, q - 2, q);
    G.delta = q / BT;
    G.ready = 1;
}

uint64_t or_bfv_modulus(void) {
    bfv_init();
    return G.q;
}

uint64_t or_bfv_psi(void) {
    bfv_init();
    return G.psi;
}

static inline uint64_t mm(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * b) % G.q); }
static inline uint64_t am(uint64_t a, uint64_t b) {
    uint64_t s = a + b;
    return s >= G.q ? s - G.q : s;
}
static inline uint64_t sm(uint64_t a, uint64_t b) { return a >= b ? a - b : a + G.q - b; }

/* Forward negacyclic NTT in place: natural order in, bit-reversed out. */
void or_bfv_ntt(uint64_t* a) {
    bfv_init();
    uint32_t t = BN;
    for (uint32_t m = 1; m < BN; m <<= 1) {
        t >>= 1;
        for (uint32_t i = 0; i < m; ++i) {
            uint32_t j1 = 2 * i * t;
            uint64_t s = G.psi_br[m + i];
            for (uint32_t j = j1; j < j1 + t; ++j) {
                uint64_t u = a[j], v = mm(a[j + t], s);
                a[j] = am(u, v);
                a[j + t] = sm(u, v);
            }
        }
    }
}

/* Inverse: bit-reversed order in, natural out, scaled by N^-1. */
void or_bfv_intt(uint64_t* a) {
    bfv_init();
    uint32_t t = 1;
    for (uint32_t m = BN; m > 1; m >>= 1) {
        uint32_t h = m >> 1, j1 = 0;
        for (uint32_t i = 0; i < h; ++i) {
            uint64_t s = G.psi_inv_br[h + i];
            for (uint32_t j = j1; j < j1 + t; ++j) {
                uint64_t u = a[j], v = a[j + t];
                a[j] = am(u, v);
                a[j + t] = mm(sm(u, v), s);
            }
            j1 += 2 * t;
        }
        t <<= 1;
    }
    for (uint32_t j = 0; j < BN; ++j) a[j] = mm(a[j], G.n_inv);
}

/* a * b in Z_q[X]/(X^N + 1) by schoolbook (the NTT's reference). */
void or_bfv_negacyclic_naive(const uint64_t* a, const uint64_t* b, uint64_t* out) {
    bfv_init();
    for (uint32_t k = 0; k < BN; ++k) out[k] = 0;
    for (uint32_t i = 0; i < BN; ++i)
        for (uint32_t j = 0; j < BN; ++j) {
            uint64_t p = mm(a[i], b[j]);
            uint32_t k = i + j;
            if (k < BN) out[k] = am(out[k], p);
            else out[k - BN] = sm(out[k - BN], p);
        }
}

/* Ternary secret key from a 32-byte seed; writes s (coefficients mod q)
 * in the NTT domain. */
void or_bfv_keygen(const uint8_t seed[32], uint64_t* s_ntt) {
    bfv_init();
    or_rng r;
    or_rng_seed(&r, seed);
    for (uint32_t i = 0; i < BN; ++i) {
        uint32_t v = or_next_u32(&r) % 3u;
        s_ntt[i] = v == 0 ? 0 : v == 1 ? 1 : G.q - 1;
    }
    or_bfv_ntt(s_ntt);
}

/* Secret-key encryption of m (N coefficients in [0, t)); ciphertext
 * (c0, c1) in the NTT domain. */
void or_bfv_encrypt(const uint64_t* s_ntt, const uint64_t* m, const uint8_t seed[32],
                    uint64_t* c0, uint64_t* c1) {
    bfv_init();
    or_rng r;
    or_rng_seed(&r, seed);
    const uint64_t mask = (1ull << 38) - 1;
    for (uint32_t i = 0; i < BN; ++i) {
        uint64_t x;
        do x = or_next_u64(&r) & mask;
        while (x >= G.q);
        c1[i] = x;
    }
    for (uint32_t i = 0; i < BN; ++i) {
        uint32_t w = or_next_u32(&r);
        int e = __builtin_popcount(w & 0x3FFu) - __builtin_popcount((w >> 10) & 0x3FFu);
        uint64_t ev = e >= 0 ? (uint64_t)e : G.q - (uint64_t)(-e);
        c0[i] = am(ev, mm(G.delta, m[i] % BT));
    }
    or_bfv_ntt(c0);
    for (uint32_t i = 0; i < BN; ++i) c0[i] = sm(c0[i], mm(c1[i], s_ntt[i]));
}

/* m = round(t (c0 + c1 s) / q) mod t, coefficient by coefficient. */
void or_bfv_decrypt(const uint64_t* s_ntt, const uint64_t* c0, const uint64_t* c1, uint64_t* m) {
    bfv_init();
    uint64_t* x = (uint64_t*)malloc(sizeof(uint64_t) * BN);
    for (uint32_t i = 0; i < BN; ++i) x[i] = am(c0[i], mm(c1[i], s_ntt[i]));
    or_bfv_intt(x);
    for (uint32_t i = 0; i < BN; ++i) {
        u128 num = (u128)2 * BT * x[i] + G.q;
        m[i] = (uint64_t)((num / ((u128)2 * G.q)) % BT);
    }
    free(x);
}

/* get_chunk, compress.rs:137-141. */
uint64_t or_tiptoe_get_chunk(uint64_t v, size_t idx) {
    return (v >> (idx * LIMB_BITS)) & ((1ull << LIMB_BITS) - 1);
}

/* from_modulo_p, compress.rs:324-332. */
int64_t or_tiptoe_from_modulo_t(uint64_t t, uint64_t v) {
    return v > t / 2 ? (int64_t)v - (int64_t)t : (int64_t)v;
}

/* decompose_hint, compress.rs:146-190: for limb b (limb_idx = 15 - b),
 * chunk c and column j, the plaintext whose coefficient i is limb_idx of
 * H[c N + i][j] (0 past row r), here in the NTT domain:
 * pt[((b * chunks + c) * n_lwe + j) * N + slot]. */
void or_tiptoe_decompose_hint(const uint64_t* hint, size_t r, size_t n_lwe, uint64_t* pt) {
    bfv_init();
    size_t chunks = (r + BN - 1) / BN;
    for (size_t b = 0; b < NUM_LIMBS; ++b) {
        size_t limb = MAX_LIMBS - b - 1;
        for (size_t c = 0; c < chunks; ++c)
            for (size_t j = 0; j < n_lwe; ++j) {
                uint64_t* p = pt + ((b * chunks + c) * n_lwe + j) * BN;
                for (size_t i = 0; i < BN; ++i) {
                    size_t row = c * BN + i;
                    p[i] = row < r ? or_tiptoe_get_chunk(hint[row * n_lwe + j], limb) : 0;
                }
                or_bfv_ntt(p);
            }
    }
}

/* encrypt_secret_under_bfv, compress.rs:201-227: ciphertext j encrypts the
 * plaintext with coefficient 0 = s[j] (s in {0, 1, 2}, checked: -1-j for an
 * out-of-range s[j]), seeds derived per coordinate from `seed`.
 * ct[(j * 2 + k) * N + slot]. */
int or_tiptoe_encrypt_secret(const uint8_t* s, size_t n_lwe, const uint64_t* sk_ntt,
                             const uint8_t seed[32], uint64_t* ct) {
    bfv_init();
    uint64_t* m = (uint64_t*)calloc(BN, sizeof(uint64_t));
    for (size_t j = 0; j < n_lwe; ++j) {
        if (s[j] > 2) {
            free(m);
            return -1 - (int)j;
        }
        uint8_t sj[32];
        memcpy(sj, seed, 32);
        for (int i = 0; i < 8; ++i) sj[i] ^= (uint8_t)((uint64_t)j >> (8 * i));
        m[0] = s[j];
        or_bfv_encrypt(sk_ntt, m, sj, ct + (j * 2) * BN, ct + (j * 2 + 1) * BN);
    }
    free(m);
    return 0;
}

/* apply_hint, compress.rs:241-266: out[b][c] = sum_j ct[j] pt[b][c][j],
 * the plaintext x ciphertext inner product (dot_product_scalar), slot by
 * slot in the NTT domain. out[((b * chunks + c) * 2 + k) * N + slot]. */
void or_tiptoe_apply_hint(const uint64_t* pt, size_t chunks, size_t n_lwe, const uint64_t* ct,
                          uint64_t* out) {
    bfv_init();
    for (size_t b = 0; b < NUM_LIMBS; ++b)
        for (size_t c = 0; c < chunks; ++c)
            for (int k = 0; k < 2; ++k) {
                uint64_t* o = out + ((b * chunks + c) * 2 + k) * BN;
                for (size_t x = 0; x < BN; ++x) {
                    uint64_t acc = 0;
                    for (size_t j = 0; j < n_lwe; ++j)
                        acc = am(acc, mm(ct[(j * 2 + k) * BN + x],
                                         pt[((b * chunks + c) * n_lwe + j) * BN + x]));
                    o[x] = acc;
                }
            }
}

/* recover_h_s, compress.rs:277-318: decrypt every (limb, chunk) answer,
 * lift each coefficient to (-t/2, t/2], scale by 2^(4 limb_idx) and sum
 * with wrapping i64 arithmetic; H s mod 2^64 per row. */
void or_tiptoe_recover_h_s(const uint64_t* answers, size_t r, const uint64_t* sk_ntt,
                           uint64_t* h_s) {
    bfv_init();
    size_t chunks = (r + BN - 1) / BN;
    uint64_t* m = (uint64_t*)malloc(sizeof(uint64_t) * BN);
    for (size_t i = 0; i < r; ++i) h_s[i] = 0;
    for (size_t b = 0; b < NUM_LIMBS; ++b) {
        size_t limb = MAX_LIMBS - b - 1;
        uint64_t scale = 1ull << (LIMB_BITS * limb);
        for (size_t c = 0; c < chunks; ++c) {
            const uint64_t* ans = answers + ((b * chunks + c) * 2) * BN;
            or_bfv_decrypt(sk_ntt, ans, ans + BN, m);
            for (size_t i = 0; i < BN && c * BN + i < r; ++i) {
                int64_t sgn = or_tiptoe_from_modulo_t(BT, m[i]);
                h_s[c * BN + i] += (uint64_t)sgn * scale; /* wrapping i64 as u64 */
            }
        }
    }
    free(m);
}
