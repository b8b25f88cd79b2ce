// svs_host.cpp — host-side client functions behind the C ABI: keygen,
// quantisation and IVF routing. Trapdoor regeneration for the device index
// build: H and S from ChaCha20Rng(h_seed) exactly as regenerate_h_s
// (scorer-bntm/src/crypto.rs:258-270; sample_fp :155-162,
// sample_sparse_pm1 :207-223). Keystream blocks are produced in parallel
// chunks (ChaCha20 is seekable by block counter); the words are consumed in
// stream order, so rejections and the data-dependent sign draws of S keep
// their exact positions.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "svs_internal.h"

namespace svs {
namespace {

unsigned worker_count() {
    unsigned hc = std::thread::hardware_concurrency();
    return std::max(1u, std::min(hc ? hc : 1u, 16u));
}

class WordSource {
public:
    WordSource(const uint8_t seed[32], size_t expected_words)
        : chunk_words_((std::max<size_t>(4096, std::min<size_t>(expected_words + 64, 1u << 24)) +
                        15) & ~(size_t)15) {
        seed_to_key(seed, key_);
        buf_.resize(chunk_words_);
        refill();
    }
    uint32_t next() {
        if (pos_ == buf_.size()) refill();
        return buf_[pos_++];
    }

private:
    void refill() {
        const size_t nblocks = chunk_words_ / 16;
        const unsigned nt = nblocks >= 4096 ? worker_count() : 1u;
        auto work = [&](unsigned t) {
            const size_t b0 = nblocks * t / nt, b1 = nblocks * (t + 1) / nt;
            for (size_t b = b0; b < b1; ++b) chacha20_block(key_, next_block_ + b, &buf_[b * 16]);
        };
        if (nt == 1) {
            work(0);
        } else {
            std::vector<std::thread> th;
            for (unsigned t = 0; t < nt; ++t) th.emplace_back(work, t);
            for (auto& x : th) x.join();
        }
        next_block_ += nblocks;
        pos_ = 0;
    }
    uint32_t key_[8];
    size_t chunk_words_;
    uint64_t next_block_ = 0;
    std::vector<uint32_t> buf_;
    size_t pos_ = 0;
};

}  // namespace

void regenerate_h_s_host(const uint8_t h_seed[32], uint32_t m, std::vector<uint64_t>& h,
                         std::vector<uint32_t>& rowptr, std::vector<uint32_t>& sidx,
                         std::vector<int8_t>& ssign) {
    const size_t expected = (size_t)m * (2 * N1 + N + N / 8);
    WordSource ws(h_seed, expected);
    h.resize((size_t)m * N1);
    for (size_t i = 0; i < h.size(); ++i) {
        for (;;) {
            uint64_t lo = ws.next();
            uint64_t x = (lo | ((uint64_t)ws.next() << 32)) & P;
            if (x < P) {
                h[i] = x;
                break;
            }
        }
    }
    rowptr.assign((size_t)m + 1, 0);
    sidx.clear();
    ssign.clear();
    sidx.reserve((size_t)m * N / 7);
    ssign.reserve((size_t)m * N / 7);
    for (uint32_t i = 0; i < m; ++i) {
        for (uint32_t j = 0; j < (uint32_t)N; ++j) {
            if (ws.next() < SPARSE_THRESHOLD) {
                sidx.push_back(j);
                ssign.push_back((ws.next() & 1u) == 0 ? -1 : 1);
            }
        }
        rowptr[i + 1] = (uint32_t)sidx.size();
    }
}

}  // namespace svs

// ============================================================ C ABI (host)
using namespace svs;

// generate_subspace (crypto.rs:235-238): L = n x n1 sample_fp draws.
extern "C" int svs_keygen_subspace(const uint8_t key_seed[32], uint64_t* l_out) {
    if (!key_seed || !l_out) return fail(SVS_ERR_ARG, "null argument");
    uint32_t key[8];
    seed_to_key(key_seed, key);
    Stream s;
    s.init(key);
    for (size_t i = 0; i < (size_t)N * N1; ++i) l_out[i] = s.fp();
    return SVS_OK;
}

// derive_h_seed (ivf.rs:258-264): seed[0..8] = upload_seed LE,
// seed[8..16] = cluster index LE, then a [u8; 32] draw (one byte per word).
extern "C" int svs_keygen_derive_h_seed(uint64_t upload_seed, uint64_t cluster, uint8_t out[32]) {
    if (!out) return fail(SVS_ERR_ARG, "null argument");
    uint8_t seed[32] = {0};
    for (int i = 0; i < 8; ++i) {
        seed[i] = (uint8_t)(upload_seed >> (8 * i));
        seed[8 + i] = (uint8_t)(cluster >> (8 * i));
    }
    uint32_t key[8];
    seed_to_key(seed, key);
    Stream s;
    s.init(key);
    for (int i = 0; i < 32; ++i) out[i] = (uint8_t)s.u32();
    return SVS_OK;
}

// domain_separate (crypto.rs:142-150).
extern "C" int svs_keygen_domain_separate(const uint8_t seed[32], const uint8_t* tag,
                                          size_t tag_len, uint8_t out[32]) {
    if (!seed || !out || (tag_len && !tag)) return fail(SVS_ERR_ARG, "null argument");
    uint8_t s[32];
    memcpy(s, seed, 32);
    for (size_t i = 0; i < tag_len; ++i) s[i % 32] ^= tag[i];
    uint32_t key[8];
    seed_to_key(s, key);
    Stream st;
    st.init(key);
    for (int i = 0; i < 32; ++i) out[i] = (uint8_t)st.u32();
    return SVS_OK;
}

// regenerate_h_s (crypto.rs:258-270) into caller buffers: h (m x n1),
// s_rowptr (m + 1), s_idx / s_sign (capacity s_cap). Returns nnz through
// *nnz_out; SVS_ERR_ARG if s_cap is too small.
extern "C" int svs_keygen_regenerate_h_s(const uint8_t h_seed[32], uint32_t m, uint64_t* h,
                                         uint32_t* s_rowptr, uint32_t* s_idx, int8_t* s_sign,
                                         size_t s_cap, size_t* nnz_out) {
    if (!h_seed || !s_rowptr || !nnz_out || (m && !h)) return fail(SVS_ERR_ARG, "null argument");
    std::vector<uint64_t> hv;
    std::vector<uint32_t> rp, si;
    std::vector<int8_t> ss;
    regenerate_h_s_host(h_seed, m, hv, rp, si, ss);
    *nnz_out = si.size();
    if (si.size() > s_cap) return fail(SVS_ERR_ARG, "s_cap %zu < nnz %zu", s_cap, si.size());
    if (m) memcpy(h, hv.data(), hv.size() * sizeof(uint64_t));
    memcpy(s_rowptr, rp.data(), rp.size() * sizeof(uint32_t));
    if (!si.empty()) {
        memcpy(s_idx, si.data(), si.size() * sizeof(uint32_t));
        memcpy(s_sign, ss.data(), ss.size());
    }
    return SVS_OK;
}

// quantise_vector (crypto.rs:105-122) for `rows` vectors of dimension d,
// zero-padded to n: round(x * q) half away from zero, negatives wrapped.
extern "C" int svs_quantise(const float* v, uint32_t rows, uint32_t d, uint64_t q, uint64_t* out) {
    if ((rows && (!v || !out)) || d > (uint32_t)N) return fail(SVS_ERR_ARG, "bad argument");
    const double qs = (double)q;
    auto work = [&](uint32_t r0, uint32_t r1) {
        for (uint32_t r = r0; r < r1; ++r) {
            uint64_t* o = out + (size_t)r * N;
            for (uint32_t j = 0; j < (uint32_t)N; ++j) o[j] = 0;
            for (uint32_t j = 0; j < d; ++j) o[j] = quantise_scalar(v[(size_t)r * d + j], qs);
        }
    };
    const unsigned nt = rows >= 4096 ? worker_count() : 1u;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back(work, (uint32_t)((uint64_t)rows * t / nt),
                        (uint32_t)((uint64_t)rows * (t + 1) / nt));
    for (auto& x : th) x.join();
    return SVS_OK;
}

// probe_route (ivf-index/src/ivf.rs:150-164) with l2_f32_scalar
// (distance.rs:30-37): f32 squared differences summed left to right (this
// file is built with -ffp-contract=off), correctly rounded sqrt, ascending
// sort under f32::total_cmp, first nprobe. Equal distances keep centroid
// order.
// (every NaN is taken as the largest value: a NaN distance's sign depends
// on the platform's NaN propagation)
static uint32_t total_key(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    if (f != f) return 0xFFFFFFFFu;
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

extern "C" int svs_probe_route(const float* centroids, uint32_t n_centroids, uint32_t d,
                               const float* queries, uint32_t n_queries, uint32_t nprobe,
                               uint32_t* out) {
    if (!centroids || !queries || !out || nprobe == 0 || nprobe > n_centroids)
        return fail(SVS_ERR_ARG, "bad argument");
    auto work = [&](uint32_t q0, uint32_t q1) {
        std::vector<std::pair<uint32_t, uint32_t>> dist(n_centroids);
        for (uint32_t q = q0; q < q1; ++q) {
            const float* x = queries + (size_t)q * d;
            for (uint32_t c = 0; c < n_centroids; ++c) {
                const float* y = centroids + (size_t)c * d;
                float acc = -0.0f;
                for (uint32_t j = 0; j < d; ++j) {
                    const float diff = x[j] - y[j];
                    acc = acc + diff * diff;
                }
                dist[c] = {total_key(std::sqrt(acc)), c};
            }
            std::partial_sort(dist.begin(), dist.begin() + nprobe, dist.end());
            for (uint32_t p = 0; p < nprobe; ++p) out[(size_t)q * nprobe + p] = dist[p].second;
        }
    };
    const unsigned nt = n_queries >= 64 ? worker_count() : 1u;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back(work, (uint32_t)((uint64_t)n_queries * t / nt),
                        (uint32_t)((uint64_t)n_queries * (t + 1) / nt));
    for (auto& x : th) x.join();
    return SVS_OK;
}
