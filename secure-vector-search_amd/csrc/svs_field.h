// svs_field.h — F_p (p = 2^61 - 1) arithmetic and the ChaCha20 keystream,
// shared by host and gfx950 device code.
//
// Semantics follow the reference scorer:
//   field ops            scorer-bntm/src/crypto.rs:47-89
//   sample_fp            scorer-bntm/src/crypto.rs:155-162
//   ChaCha20Rng stream   rand_chacha 0.10 (key = seed, 64-bit block counter,
//                        zero stream id; next_u64 = two consecutive words,
//                        low word first)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SVS_HD __host__ __device__ __forceinline__

namespace svs {

constexpr uint64_t P = 0x1FFFFFFFFFFFFFFFull;  // 2^61 - 1
constexpr int N = 1024;                        // ring dimension (Sec128)
constexpr int N1 = 128;                        // subspace dimension (Sec128)
constexpr int LAMBDA = 3;                      // Freivalds trials (params.rs:270-277)
constexpr uint32_t SPARSE_THRESHOLD = 536870911u;  // (mu * 2^32) as u64, mu = 1024^(0.7-1.0)

// ---------------------------------------------------------------- field ops
SVS_HD uint64_t add_mod(uint64_t a, uint64_t b) {
    uint64_t s = a + b;
    return s >= P ? s - P : s;
}

SVS_HD uint64_t sub_mod(uint64_t a, uint64_t b) { return a >= b ? a - b : a + P - b; }

// Canonical reduction of any 128-bit value hi:lo. Uses 2^61 = 1 (mod p), so
// 2^64 = 8: x = lo + 8*hi; the high word is split at bit 58 so that no term
// overflows. Result in [0, p).
SVS_HD uint64_t reduce128(uint64_t lo, uint64_t hi) {
    uint64_t s = (lo & P) + (lo >> 61) + ((hi << 3) & P) + (hi >> 58);  // < 2^62 + 72
    uint64_t r = (s & P) + (s >> 61);                                   // <= p + 1
    return r >= P ? r - P : r;
}

// Lazy 128-bit accumulator: acc += a*b without reduction. Callers bound the
// number of accumulated products (each < 2^122) so the sum stays < 2^128.
struct Acc128 {
    uint64_t lo, hi;
    SVS_HD void zero() { lo = 0; hi = 0; }
    SVS_HD void mac(uint64_t a, uint64_t b) {
        unsigned __int128 p = (unsigned __int128)a * b;
        unsigned __int128 s = ((unsigned __int128)hi << 64 | lo) + p;
        lo = (uint64_t)s;
        hi = (uint64_t)(s >> 64);
    }
    SVS_HD uint64_t reduce() const { return reduce128(lo, hi); }
};

SVS_HD uint64_t mul_mod(uint64_t a, uint64_t b) {
    unsigned __int128 p = (unsigned __int128)a * b;
    return reduce128((uint64_t)p, (uint64_t)(p >> 64));
}

// x * 2^s mod p for x < p, 0 <= s < 61: a 61-bit rotation (2^61 = 1).
SVS_HD uint64_t mul_pow2(uint64_t x, int s) {
    if (s == 0) return x;
    uint64_t r = ((x << s) & P) | (x >> (61 - s));
    return r == P ? 0 : r;
}

// Byte-plane fold of the plane product kernels: with c[a] = C_ab, the
// contraction of signed byte plane a of a row with signed byte plane b of a
// query (|c[a]| <= 1024 * 128 * 128 = 2^24), returns
//   2^(8b) * sum_a 2^(8a) c[a]  (mod p)   as a value in [0, p]
// (p itself stands for 0), rot = 8b in 0..56. Exact: lo and hi (the a < 4
// and a >= 4 halves) stay below 2^49 in magnitude; hi 2^32 = h1 2^61 +
// h0 2^32 with hi = h1 2^29 + h0 is h1 + h0 2^32 (mod p), so z = lo +
// h0 2^32 + h1 lies in (-2^50, 2^62); z + p for negative z, one Mersenne
// fold and a 61-bit rotation finish it. Eight such values sum to < 2^64.
SVS_HD uint64_t plane_fold(int32_t c0, int32_t c1, int32_t c2, int32_t c3, int32_t c4,
                           int32_t c5, int32_t c6, int32_t c7, int rot, int32_t m8 = 256,
                           int32_t m16 = 65536, int32_t m24 = 16777216) {
    // m8, m16, m24 = 2^8, 2^16, 2^24 (callers may pass them opaque, so that
    // each term is one 32 x 32 + 64-bit multiply-add instead of a sign
    // extension, a 64-bit shift and an add)
    const int64_t lo = (int64_t)c0 + (int64_t)c1 * m8 + (int64_t)c2 * m16 + (int64_t)c3 * m24;
    const int64_t hi = (int64_t)c4 + (int64_t)c5 * m8 + (int64_t)c6 * m16 + (int64_t)c7 * m24;
    const int64_t h1 = hi >> 29;  // arithmetic: floor(hi / 2^29)
    const uint64_t h0 = (uint64_t)hi & ((1ull << 29) - 1);
    const int64_t z = lo + (int64_t)(h0 << 32) + h1;
    const uint64_t zu = z < 0 ? (uint64_t)(z + (int64_t)P) : (uint64_t)z;
    const uint64_t r = (zu & P) + (zu >> 61);
    // r < 2^61, so the rotation by 0 (r >> 61 = 0) needs no special case
    return ((r << rot) & P) | (r >> (61 - rot));
}

// A sum of up to eight plane_fold values plus two canonical terms, as a
// canonical field element.
SVS_HD uint64_t plane_finish(uint64_t folded_sum, uint64_t t1, uint64_t t2) {
    uint64_t k = (folded_sum & P) + (folded_sum >> 61);  // <= p + 7
    k += t1 + t2;                                         // < 2^63
    k = (k & P) + (k >> 61);                              // <= p + 3
    return k >= P ? k - P : k;
}

// quantise_vector's element map (crypto.rs:105-122): round(x * q) with f64
// round-half-away-from-zero, then into [0, p) with negatives as p - (-k mod p).
// The f64 -> i64 conversion follows Rust's `as` (NaN -> 0, out of range
// saturates) and the negation wraps, as in a release build.
SVS_HD uint64_t quantise_scalar(float x, double q_scale) {
    const double r = round((double)x * q_scale);
    int64_t k;
    if (r != r) k = 0;
    else if (r >= 9223372036854775808.0) k = INT64_MAX;
    else if (r <= -9223372036854775808.0) k = INT64_MIN;
    else k = (int64_t)r;
    if (k >= 0) return (uint64_t)k % P;
    const uint64_t rr = (0ull - (uint64_t)k) % P;
    return rr == 0 ? 0 : P - rr;
}

// ------------------------------------------------------------------ ChaCha20
SVS_HD uint32_t rotl32(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }

#define SVS_QR(a, b, c, d)                     \
    a += b; d ^= a; d = svs::rotl32(d, 16);    \
    c += d; b ^= c; b = svs::rotl32(b, 12);    \
    a += b; d ^= a; d = svs::rotl32(d, 8);     \
    c += d; b ^= c; b = svs::rotl32(b, 7);

// One 64-byte ChaCha20 block: key = 8 LE words of the 32-byte seed, block
// counter in words 12-13, stream id (words 14-15) = 0.
SVS_HD void chacha20_block(const uint32_t key[8], uint64_t counter, uint32_t out[16]) {
    uint32_t x0 = 0x61707865u, x1 = 0x3320646eu, x2 = 0x79622d32u, x3 = 0x6b206574u;
    uint32_t x4 = key[0], x5 = key[1], x6 = key[2], x7 = key[3];
    uint32_t x8 = key[4], x9 = key[5], x10 = key[6], x11 = key[7];
    uint32_t x12 = (uint32_t)counter, x13 = (uint32_t)(counter >> 32), x14 = 0, x15 = 0;
    for (int i = 0; i < 10; ++i) {
        SVS_QR(x0, x4, x8, x12) SVS_QR(x1, x5, x9, x13)
        SVS_QR(x2, x6, x10, x14) SVS_QR(x3, x7, x11, x15)
        SVS_QR(x0, x5, x10, x15) SVS_QR(x1, x6, x11, x12)
        SVS_QR(x2, x7, x8, x13) SVS_QR(x3, x4, x9, x14)
    }
    out[0] = x0 + 0x61707865u; out[1] = x1 + 0x3320646eu;
    out[2] = x2 + 0x79622d32u; out[3] = x3 + 0x6b206574u;
    out[4] = x4 + key[0]; out[5] = x5 + key[1]; out[6] = x6 + key[2]; out[7] = x7 + key[3];
    out[8] = x8 + key[4]; out[9] = x9 + key[5]; out[10] = x10 + key[6]; out[11] = x11 + key[7];
    out[12] = x12 + (uint32_t)counter; out[13] = x13 + (uint32_t)(counter >> 32);
    out[14] = x14; out[15] = x15;
}

SVS_HD void seed_to_key(const uint8_t seed[32], uint32_t key[8]) {
    for (int i = 0; i < 8; ++i)
        key[i] = (uint32_t)seed[4 * i] | ((uint32_t)seed[4 * i + 1] << 8) |
                 ((uint32_t)seed[4 * i + 2] << 16) | ((uint32_t)seed[4 * i + 3] << 24);
}

// Sequential word stream (ChaCha20Rng::next_u32 / next_u64).
struct Stream {
    uint32_t key[8];
    uint64_t block;
    uint32_t buf[16];
    int idx;
    SVS_HD void init(const uint32_t k[8]) {
        for (int i = 0; i < 8; ++i) key[i] = k[i];
        block = 0;
        idx = 16;
    }
    SVS_HD uint32_t u32() {
        if (idx >= 16) {
            chacha20_block(key, block++, buf);
            idx = 0;
        }
        return buf[idx++];
    }
    SVS_HD uint64_t u64() {
        uint64_t lo = u32();
        return lo | ((uint64_t)u32() << 32);
    }
    // sample_fp: 61-bit mask, reject the single value p.
    SVS_HD uint64_t fp() {
        for (;;) {
            uint64_t x = u64() & P;
            if (x < P) return x;
        }
    }
};

}  // namespace svs
