// svs_cache.cpp — loads the reference's BN-IVF cache file straight into HBM.
//
// Layout (scorer-bntm/src/ivf.rs:1683-1902; header scorer-core/src/cache.rs:
// 19-73; IVF payload ivf-index/src/ivf.rs:68-131), all little-endian:
//   "\x89SVScach" | u32 version = 1 | u64 fingerprint64(key parts)
//   u64 dim | u64 n_centroids | n_centroids x dim f32
//   u64 n_clusters | per cluster: u64 size | size x (u32 id | dim f32)
//   zero pad to an 8-byte offset
//   u64 n_clusters | per cluster: u64 m | 32 B h_seed | m x n u64 M_enc |
//       m x n1 u64 AL | m x n u64 M_plain | m x n1 u64 H
//   n x n1 u64 L
// The file is memory-mapped and each list goes from the mapping to the
// device; S is regenerated from h_seed (the reference does not persist it)
// and the regenerated H must equal the stored one.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstring>
#include <vector>

#include "svs_internal.h"

using namespace svs;

namespace {

constexpr uint8_t kMagic[8] = {0x89, 'S', 'V', 'S', 'c', 'a', 'c', 'h'};
constexpr uint32_t kVersion = 1;

struct Mapping {
    const uint8_t* p = nullptr;
    size_t n = 0;
    int fd = -1;
    ~Mapping() {
        if (p) munmap(const_cast<uint8_t*>(p), n);
        if (fd >= 0) close(fd);
    }
};

// Bounds-checked little-endian reader over the mapping.
struct Cursor {
    const uint8_t* p;
    size_t n, off = 0;
    bool ok = true;
    bool need(size_t bytes) {
        if (!ok || n - off < bytes || off > n) ok = false;
        return ok;
    }
    uint64_t u64() {
        uint64_t v = 0;
        if (need(8)) memcpy(&v, p + off, 8), off += 8;
        return v;
    }
    const uint8_t* take(size_t bytes) {
        if (!need(bytes)) return nullptr;
        const uint8_t* q = p + off;
        off += bytes;
        return q;
    }
};

}  // namespace

// FNV-1a 64 over the length-prefixed concatenation of the parts
// (cache.rs:22-37).
extern "C" uint64_t svs_cache_fingerprint64(const uint8_t* const* parts, const uint64_t* lens,
                                            uint32_t n_parts) {
    uint64_t h = 0xcbf29ce484222325ull;
    auto eat = [&](uint8_t b) {
        h ^= b;
        h *= 0x100000001b3ull;
    };
    for (uint32_t i = 0; i < n_parts; ++i) {
        for (int k = 0; k < 8; ++k) eat((uint8_t)(lens[i] >> (8 * k)));
        for (uint64_t j = 0; j < lens[i]; ++j) eat(parts[i][j]);
    }
    return h;
}

extern "C" int svs_index_load_ivf_cache(svs_ctx* ctx, const char* path, uint64_t fingerprint,
                                        int check_fingerprint, svs_index** out, uint32_t* n_lists,
                                        uint32_t* dim) {
    if (!ctx || !path || !out) return fail(SVS_ERR_ARG, "null argument");
    *out = nullptr;
    Mapping mp;
    mp.fd = open(path, O_RDONLY);
    if (mp.fd < 0) return fail(SVS_ERR_ARG, "cannot open %s", path);
    struct stat st;
    if (fstat(mp.fd, &st) != 0) return fail(SVS_ERR_ARG, "cannot stat %s", path);
    mp.n = (size_t)st.st_size;
    if (mp.n < 20) return SVS_NONE;
    void* p = mmap(nullptr, mp.n, PROT_READ, MAP_PRIVATE, mp.fd, 0);
    if (p == MAP_FAILED) return fail(SVS_ERR_ARG, "cannot map %s", path);
    mp.p = (const uint8_t*)p;
    // header: every "unusable" case is Ok(None) in the reference
    uint32_t version = 0;
    uint64_t fp = 0;
    memcpy(&version, mp.p + 8, 4);
    memcpy(&fp, mp.p + 12, 8);
    if (memcmp(mp.p, kMagic, 8) != 0 || version != kVersion ||
        (check_fingerprint && fp != fingerprint))
        return SVS_NONE;
    Cursor c{mp.p, mp.n, 20};
    // IVF payload: centroids and the member ids of each list
    const uint64_t d = c.u64(), nc = c.u64();
    if (!c.ok || d == 0 || d > (uint64_t)N || nc == 0 || nc > (1u << 24))
        return fail(SVS_ERR_ARG, "cache: bad IVF payload header");
    const uint8_t* cent = c.take(nc * d * 4);
    const uint64_t ncl = c.u64();
    if (!c.ok || ncl != nc) return fail(SVS_ERR_ARG, "cache: truncated or inconsistent centroids");
    std::vector<std::vector<uint32_t>> ids(ncl);
    for (uint64_t ci = 0; ci < ncl; ++ci) {
        const uint64_t size = c.u64();
        if (!c.ok || size > c.n) return fail(SVS_ERR_ARG, "cache: truncated IVF payload");
        ids[ci].resize(size);
        for (uint64_t i = 0; i < size; ++i) {
            const uint8_t* e = c.take(4 + 4 * d);
            if (!e) return fail(SVS_ERR_ARG, "cache: truncated IVF payload");
            memcpy(&ids[ci][i], e, 4);
        }
    }
    c.off += (8 - c.off % 8) % 8;
    const uint64_t ncl2 = c.u64();
    if (!c.ok || ncl2 != ncl)
        return fail(SVS_ERR_ARG, "cache cluster count %llu != IVF cluster count %llu",
                    (unsigned long long)ncl2, (unsigned long long)ncl);
    struct ListRef {
        uint64_t m;
        const uint8_t *seed, *menc, *al, *mp, *h;
    };
    std::vector<ListRef> refs(ncl);
    for (uint64_t ci = 0; ci < ncl; ++ci) {
        ListRef& r = refs[ci];
        r.m = c.u64();
        r.seed = c.take(32);
        if (!c.ok || r.m > c.n / ((size_t)N * 8) || r.m != ids[ci].size())
            return fail(SVS_ERR_ARG, "cache m != IVF cluster %llu membership size",
                        (unsigned long long)ci);
        r.menc = c.take(r.m * N * 8);
        r.al = c.take(r.m * N1 * 8);
        r.mp = c.take(r.m * N * 8);
        r.h = c.take(r.m * N1 * 8);
        if (!c.ok) return fail(SVS_ERR_ARG, "cache: truncated list %llu", (unsigned long long)ci);
    }
    const uint8_t* l = c.take((size_t)N * N1 * 8);
    if (!l) return fail(SVS_ERR_ARG, "cache: truncated subspace L");
    // the index; S from h_seed, H checked against it
    svs_index* idx = nullptr;
    SVS_TRY(svs_index_create(ctx, (uint32_t)ncl, (const uint64_t*)l, &idx));
    int rc = SVS_OK;
    std::vector<uint64_t> h;
    std::vector<uint32_t> rowptr, sidx;
    std::vector<int8_t> ssign;
    for (uint64_t ci = 0; ci < ncl && rc == SVS_OK; ++ci) {
        const ListRef& r = refs[ci];
        const uint32_t m = (uint32_t)r.m;
        if (m == 0) {
            rc = svs_index_upload_list(idx, (uint32_t)ci, 0, nullptr, nullptr, nullptr, nullptr,
                                       nullptr, nullptr, nullptr, nullptr);
            continue;
        }
        size_t cap = (size_t)m * N / 4 + 1024, nnz = 0;
        h.resize((size_t)m * N1);
        rowptr.resize(m + 1);
        sidx.resize(cap);
        ssign.resize(cap);
        rc = svs_keygen_regenerate_h_s(r.seed, m, h.data(), rowptr.data(), sidx.data(),
                                       ssign.data(), cap, &nnz);
        if (rc != SVS_OK && nnz > cap) {
            cap = nnz;
            sidx.resize(cap);
            ssign.resize(cap);
            rc = svs_keygen_regenerate_h_s(r.seed, m, h.data(), rowptr.data(), sidx.data(),
                                           ssign.data(), cap, &nnz);
        }
        if (rc != SVS_OK) break;
        if (memcmp(h.data(), r.h, h.size() * 8) != 0) {
            rc = fail(SVS_ERR_ARG, "cache list %llu: H does not match its h_seed",
                      (unsigned long long)ci);
            break;
        }
        rc = svs_index_upload_list(idx, (uint32_t)ci, m, (const uint64_t*)r.menc,
                                   (const uint64_t*)r.al, (const uint64_t*)r.h,
                                   (const uint64_t*)r.mp, rowptr.data(), sidx.data(),
                                   ssign.data(), ids[ci].data());
    }
    if (rc == SVS_OK && ncl <= (uint64_t)RT_SELECT_MAX)
        rc = svs_index_set_centroids(idx, (const float*)cent, (uint32_t)d);
    if (rc != SVS_OK) {
        char msg[512];
        svs_last_error(msg, sizeof(msg));
        svs_index_destroy(idx);
        return fail(rc, "%s", msg);
    }
    *out = idx;
    if (n_lists) *n_lists = (uint32_t)ncl;
    if (dim) *dim = (uint32_t)d;
    return SVS_OK;
}

extern "C" int svs_index_get_centroids(svs_index* idx, float* out) {
    if (!idx || !out) return fail(SVS_ERR_ARG, "null argument");
    if (!idx->cent_dev) return fail(SVS_ERR_STATE, "the index has no centroids");
    std::lock_guard<std::mutex> g(idx->ctx->mu);
    HIP_TRY(hipSetDevice(idx->ctx->device));
    HIP_TRY(hipMemcpy(out, idx->cent_dev, sizeof(float) * idx->lists.size() * idx->cent_d,
                      hipMemcpyDeviceToHost));
    return SVS_OK;
}

extern "C" int svs_index_get_subspace(svs_index* idx, uint64_t* l_out) {
    if (!idx || !l_out) return fail(SVS_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> g(idx->ctx->mu);
    HIP_TRY(hipSetDevice(idx->ctx->device));
    HIP_TRY(hipMemcpy(l_out, idx->l_dev, sizeof(uint64_t) * N * N1, hipMemcpyDeviceToHost));
    return SVS_OK;
}
