// svs_api.hip — C ABI of the gfx950 BN scoring backend (include/svs_gpu.h):
// context, resident index, device index build, per-list drop-ins and the
// gathered whole-query pipeline.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/svs_gpu.h"
#include "svs_internal.h"
#include "svs_kernels.h"
#include "svs_plan.h"
#include "svs_route.h"

using namespace svs;

// ===================================================================== errors
static thread_local std::string g_last_error;

int svs::fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

extern "C" int svs_last_error(char* buf, size_t len) {
    if (buf && len) {
        size_t n = std::min(len - 1, g_last_error.size());
        memcpy(buf, g_last_error.data(), n);
        buf[n] = 0;
    }
    return (int)g_last_error.size();
}

// ==================================================================== context
int DevBuf::ensure(svs_ctx* ctx, size_t bytes) {
    if (bytes <= cap) return SVS_OK;
    if (ptr) {
        hipFree(ptr);
        ctx->scratch_bytes -= cap;
        ptr = nullptr;
        cap = 0;
    }
    // geometric growth: a buffer that keeps growing is reallocated O(log) times
    size_t want = std::max({bytes, (size_t)4096, cap + cap / 2});
    want = (want + 4095) & ~(size_t)4095;
    HIP_TRY(hipMalloc(&ptr, want));
    cap = want;
    ctx->scratch_bytes += cap;
    ctx->note_peak();
    return SVS_OK;
}

void DevBuf::release(svs_ctx* ctx) {
    if (ptr) {
        hipFree(ptr);
        if (ctx) ctx->scratch_bytes -= cap;
    }
    ptr = nullptr;
    cap = 0;
}

int HostBuf::ensure(size_t bytes) {
    if (bytes <= cap) return SVS_OK;
    const size_t want = std::max(bytes, cap + cap / 2);
    if (ptr) hipHostFree(ptr);
    ptr = nullptr;
    cap = 0;
    HIP_TRY(hipHostMalloc(&ptr, want, hipHostMallocDefault));
    cap = want;
    return SVS_OK;
}

void HostBuf::release() {
    if (ptr) hipHostFree(ptr);
    ptr = nullptr;
    cap = 0;
}

void svs_ctx::note_peak() {
    uint64_t now = resident_bytes + scratch_bytes;
    if (now > peak_bytes) peak_bytes = now;
}

int svs::dev_alloc(svs_ctx* ctx, void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return SVS_OK;
    if (ctx->resident_bytes + bytes > ctx->budget_bytes)
        return fail(SVS_ERR_BUDGET,
                    "list payload %zu bytes exceeds budget %llu bytes (resident %llu)", bytes,
                    (unsigned long long)ctx->budget_bytes,
                    (unsigned long long)ctx->resident_bytes);
    HIP_TRY(hipMalloc(p, bytes));
    ctx->resident_bytes += bytes;
    ctx->note_peak();
    return SVS_OK;
}

void svs::dev_free(svs_ctx* ctx, void* p, size_t bytes) {
    if (!p) return;
    hipFree(p);
    ctx->resident_bytes -= bytes;
}

// Caller-held device buffers (query inputs and outputs kept in HBM for
// svs_score_batch_dev), counted against the context budget.
extern "C" int svs_dev_alloc(svs_ctx* ctx, uint64_t bytes, void** out) {
    if (!ctx || !out || bytes == 0) return fail(SVS_ERR_ARG, "null argument or zero size");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    SVS_TRY(dev_alloc(ctx, out, bytes));
    ctx->user_allocs[*out] = bytes;
    return SVS_OK;
}

extern "C" int svs_dev_free(svs_ctx* ctx, void* p) {
    if (!ctx) return fail(SVS_ERR_ARG, "null context");
    if (!p) return SVS_OK;
    std::lock_guard<std::mutex> g(ctx->mu);
    auto it = ctx->user_allocs.find(p);
    if (it == ctx->user_allocs.end()) return fail(SVS_ERR_ARG, "pointer not from svs_dev_alloc");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    dev_free(ctx, p, it->second);
    ctx->user_allocs.erase(it);
    return SVS_OK;
}

extern "C" int svs_memcpy(svs_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return fail(SVS_ERR_ARG, "null argument");
    if (bytes == 0) return SVS_OK;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SVS_OK;
}

extern "C" int svs_ctx_create(int device, uint64_t budget_bytes, svs_ctx** out) {
    if (!out) return fail(SVS_ERR_ARG, "out is null");
    *out = nullptr;
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(SVS_ERR_ARG, "device %d of %d", device, n);
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(SVS_ERR_STATE, "device %d is %s; this library is built for gfx950", device,
                    prop.gcnArchName);
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    auto* ctx = new svs_ctx();
    ctx->device = device;
    ctx->n_cu = prop.multiProcessorCount;
    if (const char* g = getenv("SVS_TILE_GRID")) ctx->tile_grid = std::max(1, std::min(8, atoi(g)));
    if (const char* g = getenv("SVS_PROD_SPLIT")) ctx->prod_split = atoi(g) ? 1 : 0;
    if (const char* g = getenv("SVS_PROD_SPLIT_PF")) ctx->prod_split_pf = atoi(g);
    if (const char* g = getenv("SVS_DENSE_PLANES")) ctx->dense_planes = atoi(g) ? 1 : 0;
    if (const char* g = getenv("SVS_PROD_LDSB")) ctx->prod_ldsb = atoi(g) ? 1 : 0;
    // graphs off for the whole process (profilers that trace kernels have
    // crashed on the host inside graph capture; the kernels are the same)
    if (const char* g = getenv("SVS_GRAPHS")) ctx->graphs = atoi(g) ? 1 : 0;
    ctx->budget_bytes = budget_bytes ? budget_bytes : (uint64_t)(free_b * 0.9);
    ctx->hbm_total = total_b;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return fail(SVS_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    // the fixed scratch of svs_verify_noalloc on this device
    const int rs = noalloc_scratch_init(device);
    if (rs != SVS_OK) {
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return rs;
    }
    *out = ctx;
    return SVS_OK;
}

// A context outlives its indexes: destroying it while indexes exist only
// marks it, and the last svs_index_destroy frees it (callers with garbage-
// collected handles cannot order the two calls).
static std::mutex g_ctx_refs;

static void ctx_free(svs_ctx* ctx) {
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    release_graphs(*ctx);
    ctx->h_stage.release();
    for (DevBuf* b : ctx->all_scratch()) b->release(ctx);
    for (auto& kv : ctx->user_allocs) dev_free(ctx, kv.first, kv.second);
    ctx->user_allocs.clear();
    if (ctx->prof.created)
        for (auto& e : ctx->prof.ev) hipEventDestroy(e);
    hipStreamDestroy(ctx->stream);
    delete ctx;
}

void svs::ctx_retain(svs_ctx* ctx) {
    std::lock_guard<std::mutex> g(g_ctx_refs);
    ctx->index_refs++;
}

void svs::ctx_release(svs_ctx* ctx) {
    bool free_ctx;
    {
        std::lock_guard<std::mutex> g(g_ctx_refs);
        free_ctx = --ctx->index_refs == 0 && ctx->destroy_pending;
    }
    if (free_ctx) ctx_free(ctx);
}

extern "C" int svs_ctx_destroy(svs_ctx* ctx) {
    if (!ctx) return SVS_OK;
    {
        std::lock_guard<std::mutex> g(g_ctx_refs);
        if (ctx->index_refs > 0) {
            ctx->destroy_pending = true;
            return SVS_OK;
        }
    }
    ctx_free(ctx);
    return SVS_OK;
}

extern "C" int svs_ctx_stats(svs_ctx* ctx, svs_stats* out) {
    if (!ctx || !out) return fail(SVS_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    out->budget_bytes = ctx->budget_bytes;
    out->resident_bytes = ctx->resident_bytes;
    out->peak_resident_bytes = ctx->peak_bytes;
    out->hbm_total_bytes = total_b;
    out->hbm_free_bytes = free_b;
    out->uploads = ctx->uploads;
    out->launches = ctx->launches;
    out->graph_launches = ctx->graph_launches;
    out->graph_captures = ctx->graph_captures;
    return SVS_OK;
}

extern "C" int svs_ctx_set_graphs(svs_ctx* ctx, int on) {
    if (!ctx) return fail(SVS_ERR_ARG, "null context");
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->graphs = on ? 1 : 0;
    return SVS_OK;
}

// ====================================================================== index
// Generations key the captured batch graphs: a graph holds the device
// addresses of its index (L, L^T, the list table, the centroids) and the
// list count, so a new index, or new centroids, must never replay a graph
// captured before, even at a reused host address.
static std::atomic<uint64_t> g_index_gen{1};
uint64_t svs::next_index_gen() { return g_index_gen.fetch_add(1); }

extern "C" int svs_index_create(svs_ctx* ctx, uint32_t n_lists, const uint64_t* l_host,
                                svs_index** out) {
    if (!ctx || !out || !l_host) return fail(SVS_ERR_ARG, "null argument");
    if (n_lists == 0) return fail(SVS_ERR_ARG, "n_lists must be > 0");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    for (size_t i = 0; i < (size_t)N * N1; ++i)
        if (l_host[i] >= P) return fail(SVS_ERR_ARG, "L[%zu] is not canonical", i);
    auto* idx = new svs_index();
    idx->ctx = ctx;
    idx->gen = next_index_gen();
    idx->lists.resize(n_lists);
    int rc = dev_alloc(ctx, (void**)&idx->l_dev, sizeof(uint64_t) * N * N1);
    if (rc == SVS_OK) rc = dev_alloc(ctx, (void**)&idx->lt_dev, sizeof(uint64_t) * N * N1);
    if (rc == SVS_OK) rc = dev_alloc(ctx, (void**)&idx->lists_dev, sizeof(ListDev) * n_lists);
    if (rc != SVS_OK) {
        dev_free(ctx, idx->l_dev, sizeof(uint64_t) * N * N1);
        dev_free(ctx, idx->lt_dev, sizeof(uint64_t) * N * N1);
        delete idx;
        return rc;
    }
    std::vector<uint64_t> lt((size_t)N * N1);
    for (int j = 0; j < N; ++j)
        for (int k = 0; k < N1; ++k) lt[(size_t)k * N + j] = l_host[(size_t)j * N1 + k];
    HIP_TRY(hipMemcpy(idx->l_dev, l_host, sizeof(uint64_t) * N * N1, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(idx->lt_dev, lt.data(), sizeof(uint64_t) * N * N1, hipMemcpyHostToDevice));
    std::vector<ListDev> zero(n_lists);
    memset(zero.data(), 0, sizeof(ListDev) * n_lists);
    HIP_TRY(hipMemcpy(idx->lists_dev, zero.data(), sizeof(ListDev) * n_lists,
                      hipMemcpyHostToDevice));
    ctx_retain(ctx);
    *out = idx;
    return SVS_OK;
}

static void free_list(svs_ctx* ctx, ListHost& L) {
    for (auto& a : L.allocs) dev_free(ctx, a.first, a.second);
    L = ListHost();
}

extern "C" int svs_index_destroy(svs_index* idx) {
    if (!idx) return SVS_OK;
    svs_ctx* ctx = idx->ctx;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        hipSetDevice(ctx->device);
        hipStreamSynchronize(ctx->stream);
        for (auto& L : idx->lists) free_list(ctx, L);
        dev_free(ctx, idx->l_dev, sizeof(uint64_t) * N * N1);
        dev_free(ctx, idx->lt_dev, sizeof(uint64_t) * N * N1);
        dev_free(ctx, idx->lists_dev, sizeof(ListDev) * idx->lists.size());
        if (idx->cent_dev)
            dev_free(ctx, idx->cent_dev, sizeof(float) * idx->lists.size() * idx->cent_d);
        if (ctx->cached_plan_index == idx) ctx->cached_plan_index = nullptr;
        delete idx;
    }
    ctx_release(ctx);
    return SVS_OK;
}

extern "C" uint32_t svs_index_list_rows(const svs_index* idx, uint32_t list) {
    if (!idx || list >= idx->lists.size()) return 0;
    return idx->lists[list].m;
}

// Sliced-ELL S from CSR: per 64-row slice, width = max row nnz; entry
// (e, lane) at slice_off + e * 64 + lane holds idx | 0x8000 for sign -1.
static void build_s_ell(uint32_t m, const uint32_t* rowptr, const uint32_t* sidx,
                        const int8_t* ssign, std::vector<uint32_t>& slice_off,
                        std::vector<uint16_t>& ell) {
    const uint32_t nslices = (m + SLICE - 1) / SLICE;
    slice_off.assign(nslices + 1, 0);
    for (uint32_t s = 0; s < nslices; ++s) {
        uint32_t w = 0;
        for (uint32_t r = s * SLICE; r < std::min(m, (s + 1) * SLICE); ++r)
            w = std::max(w, rowptr[r + 1] - rowptr[r]);
        slice_off[s + 1] = slice_off[s] + w * SLICE;
    }
    ell.assign(slice_off[nslices], S_PAD);
    for (uint32_t r = 0; r < m; ++r) {
        const uint32_t s = r / SLICE, lane = r % SLICE;
        for (uint32_t e = rowptr[r]; e < rowptr[r + 1]; ++e)
            ell[slice_off[s] + (size_t)(e - rowptr[r]) * SLICE + lane] =
                (uint16_t)((sidx[e] & 0x3FF) | (ssign[e] < 0 ? 0x8000 : 0));
    }
}

// Allocates and fills the list's device state except menc / alh content.
static int list_alloc(svs_index* idx, uint32_t list, uint32_t m, bool mp32, bool want_mp,
                      bool planes) {
    svs_ctx* ctx = idx->ctx;
    ListHost& L = idx->lists[list];
    free_list(ctx, L);
    L.m = m;
    if (m == 0) return SVS_OK;
    auto take = [&](void** p, size_t bytes) -> int {
        int rc = dev_alloc(ctx, p, bytes);
        if (rc == SVS_OK && *p) L.allocs.push_back({*p, bytes});
        return rc;
    };
    int rc = take((void**)&L.d.menc, sizeof(uint64_t) * (size_t)m * N);
    if (rc == SVS_OK && want_mp) {
        if (mp32) rc = take((void**)&L.d.mp_t32, sizeof(int32_t) * (size_t)m * N);
        else rc = take((void**)&L.d.mp_t64, sizeof(uint64_t) * (size_t)m * N);
    }
    if (rc == SVS_OK) rc = take((void**)&L.d.ids, sizeof(uint32_t) * m);
    if (rc == SVS_OK) rc = take((void**)&L.d.rowcorr, sizeof(uint64_t) * m);
    if (rc == SVS_OK) rc = take((void**)&L.d.alh, sizeof(uint64_t) * (size_t)m * 2 * N1);
    if (rc == SVS_OK) rc = take((void**)&L.d.alh_rowcorr, sizeof(uint64_t) * m);
    // the decode tiles' planes: rows padded to 16 (their tiled layout)
    const size_t m16 = ((size_t)m + 15) & ~(size_t)15;
    if (rc == SVS_OK && planes) rc = take((void**)&L.d.mp8, m16 * 3 * N);
    if (rc == SVS_OK) rc = take((void**)&L.d.s8, m16 * N);
    if (rc == SVS_OK) rc = take((void**)&L.d.s_rowterm, sizeof(uint64_t) * m);
    if (rc != SVS_OK) free_list(ctx, L);
    L.d.m = m;
    // no M_plain: no column of it may be read (the decode kernels bound
    // their M_plain gathers by mp_cols)
    L.d.mp_cols = want_mp ? N : 0;
    return rc;
}

static int list_upload_s_ids(svs_index* idx, uint32_t list, const uint32_t* rowptr,
                             const uint32_t* sidx, const int8_t* ssign, const uint32_t* ids) {
    svs_ctx* ctx = idx->ctx;
    ListHost& L = idx->lists[list];
    const uint32_t m = L.m;
    std::vector<uint32_t> soff;
    std::vector<uint16_t> ell;
    build_s_ell(m, rowptr, sidx, ssign, soff, ell);
    void* p = nullptr;
    int rc = dev_alloc(ctx, &p, soff.size() * sizeof(uint32_t));
    if (rc != SVS_OK) return rc;
    L.allocs.push_back({p, soff.size() * sizeof(uint32_t)});
    L.d.s_slice = (const uint32_t*)p;
    HIP_TRY(hipMemcpy(p, soff.data(), soff.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    if (!ell.empty()) {
        rc = dev_alloc(ctx, &p, ell.size() * sizeof(uint16_t));
        if (rc != SVS_OK) return rc;
        L.allocs.push_back({p, ell.size() * sizeof(uint16_t)});
        L.d.s_ell = (const uint16_t*)p;
        HIP_TRY(hipMemcpy(p, ell.data(), ell.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    }
    std::vector<uint32_t> idv(m);
    for (uint32_t i = 0; i < m; ++i) idv[i] = ids ? ids[i] : i;
    HIP_TRY(hipMemcpy((void*)L.d.ids, idv.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice));
    // dense int8 S and its row term for the MFMA decode
    const uint64_t words = (((uint64_t)m + 15) & ~15ull) * N / 4;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)((words + 255) / 256)), dim3(256), 0,
                       ctx->stream, (uint32_t*)L.d.s8, (uint32_t)words, 0u);
    hipLaunchKernelGGL(k_make_s8, dim3((m + 255) / 256), dim3(256), 0, ctx->stream, L.d.s_slice,
                       L.d.s_ell, m, (int8_t*)L.d.s8, (uint64_t*)L.d.s_rowterm);
    HIP_TRY(hipGetLastError());
    ctx->launches += 2;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SVS_OK;
}

// Signed lifts within (-2^23, 2^23): the rows get the MFMA decode planes.
static bool fits_i24(const uint64_t* mp, size_t count) {
    for (size_t i = 0; i < count; ++i) {
        const uint64_t x = mp[i];
        if (!(x < (uint64_t)MP8_LIMIT || x > P - (uint64_t)MP8_LIMIT)) return false;
    }
    return true;
}

// The smallest multiple of 64 beyond which every row of M_plain is zero:
// quantise_vector zero-pads d < n dimensions (crypto.rs:105-122), and the
// decode skips M T over those columns.
static uint32_t plain_cols(const uint64_t* mp, uint32_t m) {
    uint32_t w = 0;
    for (uint32_t r = 0; r < m && w < (uint32_t)N; ++r) {
        const uint64_t* row = mp + (size_t)r * N;
        uint32_t c = N;
        while (c > w && row[c - 1] == 0) --c;
        w = c;
    }
    return (w + 63) / 64 * 64;
}

static bool fits_i32(const uint64_t* mp, size_t count) {
    for (size_t i = 0; i < count; ++i) {
        const uint64_t x = mp[i];
        if (!(x <= 0x7FFFFFFFull || x >= P - 0x80000000ull)) return false;
    }
    return true;
}

// Derived device state shared by both upload paths, once menc and alh are
// in place: row-major [AL | H] for the MFMA decode, and the row terms of the
// MFMA identity for M_enc and for [AL | H].
static int finish_list(svs_ctx* ctx, ListHost& L) {
    if (!L.m) return SVS_OK;
    const uint32_t m = L.m;
    const uint32_t blocks = std::min<uint32_t>((m + 3) / 4, (uint32_t)ctx->n_cu * 32);
    hipLaunchKernelGGL(k_row_corr<N>, dim3(blocks), dim3(256), 0, ctx->stream,
                       (const uint64_t*)L.d.menc, m, (uint64_t*)L.d.rowcorr);
    hipLaunchKernelGGL(k_row_corr<2 * N1>, dim3(blocks), dim3(256), 0, ctx->stream,
                       (const uint64_t*)L.d.alh, m, (uint64_t*)L.d.alh_rowcorr);
    HIP_TRY(hipGetLastError());
    ctx->launches += 2;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SVS_OK;
}

static int publish_list(svs_index* idx, uint32_t list) {
    ListDev d = idx->lists[list].d;
    d.m = idx->lists[list].m;
    HIP_TRY(hipMemcpy(idx->lists_dev + list, &d, sizeof(ListDev), hipMemcpyHostToDevice));
    idx->ctx->uploads++;
    idx->ctx->cached_plan_list = -1;
    idx->top_rows_dirty = true;
    return SVS_OK;
}

// Transposes a host row-major rows x cols u64 matrix (via a temp device
// buffer) into dst (cols x dst_ld).
static int upload_transposed(svs_ctx* ctx, const uint64_t* host, uint32_t rows, uint32_t cols,
                             uint64_t* dst, uint32_t dst_ld) {
    DevBuf& tmp = ctx->tmp_a;
    SVS_TRY(tmp.ensure(ctx, sizeof(uint64_t) * (size_t)rows * cols));
    HIP_TRY(hipMemcpy(tmp.ptr, host, sizeof(uint64_t) * (size_t)rows * cols,
                      hipMemcpyHostToDevice));
    dim3 grid((cols + 31) / 32, (rows + 31) / 32);
    if (grid.y > 65535) return fail(SVS_ERR_ARG, "list of %u rows too large to transpose", rows);
    hipLaunchKernelGGL(k_transpose_u64, grid, dim3(32, 8), 0, ctx->stream,
                       (const uint64_t*)tmp.ptr, rows, cols, dst, dst_ld);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SVS_OK;
}

static int upload_mplain(svs_ctx* ctx, ListHost& L, const uint64_t* mp) {
    const uint32_t m = L.m;
    if (L.d.mp_t32) {
        DevBuf& tmp = ctx->tmp_a;
        SVS_TRY(tmp.ensure(ctx, sizeof(uint64_t) * (size_t)m * N));
        HIP_TRY(hipMemcpy(tmp.ptr, mp, sizeof(uint64_t) * (size_t)m * N, hipMemcpyHostToDevice));
        dim3 grid(N / 32, (m + 31) / 32);
        if (grid.y > 65535) return fail(SVS_ERR_ARG, "list of %u rows too large", m);
        hipLaunchKernelGGL(k_transpose_mp32, grid, dim3(32, 8), 0, ctx->stream,
                           (const uint64_t*)tmp.ptr, m, (int32_t*)L.d.mp_t32);
        if (L.d.mp8)
            hipLaunchKernelGGL(k_make_mp8, dim3(std::min<uint32_t>(m * 4, 65536)),
                               dim3(256), 0, ctx->stream, (const uint64_t*)tmp.ptr, m,
                               (int8_t*)L.d.mp8);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        return SVS_OK;
    }
    if (L.d.mp_t64) return upload_transposed(ctx, mp, m, N, (uint64_t*)L.d.mp_t64, m);
    return SVS_OK;
}

static int check_canonical(const uint64_t* a, size_t count, const char* what) {
    for (size_t i = 0; i < count; ++i)
        if (a[i] >= P) return fail(SVS_ERR_ARG, "%s[%zu] is not canonical", what, i);
    return SVS_OK;
}

extern "C" int svs_index_upload_list(svs_index* idx, uint32_t list, uint32_t m,
                                     const uint64_t* m_enc, const uint64_t* a_l,
                                     const uint64_t* h, const uint64_t* m_plain,
                                     const uint32_t* s_rowptr, const uint32_t* s_idx,
                                     const int8_t* s_sign, const uint32_t* ids) {
    if (!idx) return fail(SVS_ERR_ARG, "null index");
    if (list >= idx->lists.size()) return fail(SVS_ERR_ARG, "list %u out of range", list);
    if (m > 0 && (!m_enc || !a_l || !h || !s_rowptr))
        return fail(SVS_ERR_ARG, "m_enc, a_l, h and s_rowptr are required");
    if (m > 0 && s_rowptr[m] > 0 && (!s_idx || !s_sign))
        return fail(SVS_ERR_ARG, "s_idx and s_sign are required");
    svs_ctx* ctx = idx->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    SVS_TRY(check_canonical(m_enc, (size_t)m * N, "m_enc"));
    SVS_TRY(check_canonical(a_l, (size_t)m * N1, "a_l"));
    SVS_TRY(check_canonical(h, (size_t)m * N1, "h"));
    if (m_plain) SVS_TRY(check_canonical(m_plain, (size_t)m * N, "m_plain"));
    for (uint32_t i = 0; i < m; ++i)
        for (uint32_t e = s_rowptr[i]; e < s_rowptr[i + 1]; ++e)
            if (s_idx[e] >= (uint32_t)N || (s_sign[e] != 1 && s_sign[e] != -1) ||
                (e > s_rowptr[i] && s_idx[e] <= s_idx[e - 1]))
                return fail(SVS_ERR_ARG, "S row %u is not a sorted +-1 row", i);
    const bool mp32 = m_plain && fits_i32(m_plain, (size_t)m * N);
    const bool planes = mp32 && fits_i24(m_plain, (size_t)m * N);
    SVS_TRY(list_alloc(idx, list, m, mp32, m_plain != nullptr, planes));
    if (m_plain) idx->lists[list].d.mp_cols = plain_cols(m_plain, m);
    ListHost& L = idx->lists[list];
    if (m > 0) {
        HIP_TRY(hipMemcpy((void*)L.d.menc, m_enc, sizeof(uint64_t) * (size_t)m * N,
                          hipMemcpyHostToDevice));
        // [AL | H] row-major: AL into columns 0..127, H into 128..255
        HIP_TRY(hipMemcpy2D((void*)L.d.alh, sizeof(uint64_t) * 2 * N1, a_l,
                            sizeof(uint64_t) * N1, sizeof(uint64_t) * N1, m,
                            hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy2D((void*)(L.d.alh + N1), sizeof(uint64_t) * 2 * N1, h,
                            sizeof(uint64_t) * N1, sizeof(uint64_t) * N1, m,
                            hipMemcpyHostToDevice));
        SVS_TRY(finish_list(ctx, L));
        if (m_plain) SVS_TRY(upload_mplain(ctx, L, m_plain));
        SVS_TRY(list_upload_s_ids(idx, list, s_rowptr, s_idx, s_sign, ids));
    }
    return publish_list(idx, list);
}

extern "C" int svs_index_build_list(svs_index* idx, uint32_t list, uint32_t m,
                                    const uint64_t* m_plain, const uint8_t h_seed[32],
                                    const uint32_t* ids) {
    if (!idx || !h_seed) return fail(SVS_ERR_ARG, "null argument");
    if (list >= idx->lists.size()) return fail(SVS_ERR_ARG, "list %u out of range", list);
    if (m > 0 && !m_plain) return fail(SVS_ERR_ARG, "m_plain is required");
    svs_ctx* ctx = idx->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    SVS_TRY(check_canonical(m_plain, (size_t)m * N, "m_plain"));
    // H and S from h_seed on the host (regenerate_h_s, crypto.rs:258-270).
    std::vector<uint64_t> h;
    std::vector<uint32_t> rowptr, sidx;
    std::vector<int8_t> ssign;
    regenerate_h_s_host(h_seed, m, h, rowptr, sidx, ssign);
    const bool mp32 = fits_i32(m_plain, (size_t)m * N);
    const bool planes = mp32 && fits_i24(m_plain, (size_t)m * N);
    SVS_TRY(list_alloc(idx, list, m, mp32, true, planes));
    ListHost& L = idx->lists[list];
    L.d.mp_cols = plain_cols(m_plain, m);
    if (m > 0) {
        DevBuf& tm = ctx->tmp_a;
        DevBuf& th = ctx->tmp_b;
        DevBuf& ts = ctx->tmp_c;
        const size_t nnz = sidx.size();
        SVS_TRY(tm.ensure(ctx, sizeof(uint64_t) * (size_t)m * N));
        SVS_TRY(th.ensure(ctx, sizeof(uint64_t) * (size_t)m * N1));
        SVS_TRY(ts.ensure(ctx, sizeof(uint32_t) * (m + 1 + nnz) + nnz + 16));
        HIP_TRY(hipMemcpy(tm.ptr, m_plain, sizeof(uint64_t) * (size_t)m * N,
                          hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(th.ptr, h.data(), sizeof(uint64_t) * (size_t)m * N1,
                          hipMemcpyHostToDevice));
        uint32_t* d_rowptr = (uint32_t*)ts.ptr;
        uint32_t* d_sidx = d_rowptr + m + 1;
        int8_t* d_ssign = (int8_t*)(d_sidx + nnz);
        HIP_TRY(hipMemcpy(d_rowptr, rowptr.data(), sizeof(uint32_t) * (m + 1),
                          hipMemcpyHostToDevice));
        if (nnz) {
            HIP_TRY(hipMemcpy(d_sidx, sidx.data(), sizeof(uint32_t) * nnz, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(d_ssign, ssign.data(), nnz, hipMemcpyHostToDevice));
        }
        const uint64_t* d_m = (const uint64_t*)tm.ptr;
        hipLaunchKernelGGL(k_encrypt_menc, dim3(m * (N / 256)), dim3(256), 0, ctx->stream, d_m,
                           (const uint64_t*)th.ptr, (const uint64_t*)idx->lt_dev, d_rowptr,
                           (const uint32_t*)d_sidx, (const int8_t*)d_ssign, m,
                           (uint64_t*)L.d.menc);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_encrypt_al, dim3((m + 7) / 8), dim3(128), 0, ctx->stream, d_m,
                           (const uint64_t*)idx->l_dev, m, (uint64_t*)L.d.alh);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpy2DAsync((void*)(L.d.alh + N1), sizeof(uint64_t) * 2 * N1, th.ptr,
                                 sizeof(uint64_t) * N1, sizeof(uint64_t) * N1, m,
                                 hipMemcpyDeviceToDevice, ctx->stream));
        if (L.d.mp_t32) {
            dim3 mg(N / 32, (m + 31) / 32);
            hipLaunchKernelGGL(k_transpose_mp32, mg, dim3(32, 8), 0, ctx->stream, d_m, m,
                               (int32_t*)L.d.mp_t32);
            if (L.d.mp8)
                hipLaunchKernelGGL(k_make_mp8, dim3(std::min<uint32_t>(m * 4, 65536)), dim3(256),
                                   0, ctx->stream, d_m, m, (int8_t*)L.d.mp8);
        } else {
            dim3 mg(N / 32, (m + 31) / 32);
            hipLaunchKernelGGL(k_transpose_u64, mg, dim3(32, 8), 0, ctx->stream, d_m, m,
                               (uint32_t)N, (uint64_t*)L.d.mp_t64, m);
        }
        HIP_TRY(hipGetLastError());
        ctx->launches += 4;
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        SVS_TRY(finish_list(ctx, L));
        SVS_TRY(list_upload_s_ids(idx, list, rowptr.data(), sidx.data(), ssign.data(), ids));
    }
    return publish_list(idx, list);
}

extern "C" int svs_index_download_list(svs_index* idx, uint32_t list, uint64_t* m_enc,
                                       uint64_t* a_l, uint64_t* h) {
    if (!idx) return fail(SVS_ERR_ARG, "null argument");
    if (list >= idx->lists.size()) return fail(SVS_ERR_ARG, "list %u out of range", list);
    std::lock_guard<std::mutex> g(idx->ctx->mu);
    HIP_TRY(hipSetDevice(idx->ctx->device));
    const ListHost& L = idx->lists[list];
    const size_t m = L.m;
    if (!m) return SVS_OK;
    if (m_enc)
        HIP_TRY(hipMemcpy(m_enc, L.d.menc, sizeof(uint64_t) * m * N, hipMemcpyDeviceToHost));
    if (a_l)
        HIP_TRY(hipMemcpy2D(a_l, sizeof(uint64_t) * N1, L.d.alh, sizeof(uint64_t) * 2 * N1,
                            sizeof(uint64_t) * N1, m, hipMemcpyDeviceToHost));
    if (h)
        HIP_TRY(hipMemcpy2D(h, sizeof(uint64_t) * N1, L.d.alh + N1, sizeof(uint64_t) * 2 * N1,
                            sizeof(uint64_t) * N1, m, hipMemcpyDeviceToHost));
    return SVS_OK;
}

// ================================================================== pipeline
// Device-side work description of a batch: items, prefixes, flags.

int svs::make_plan(const svs_index* idx, uint32_t nq, const uint32_t* probe_ptr,
              const uint32_t* probe_lists, const uint8_t* verify, BatchPlan& pl,
              uint32_t tile_rows, uint32_t tile_qg) {
    pl = BatchPlan();
    pl.n_queries = nq;
    if (probe_ptr[0] != 0) return fail(SVS_ERR_ARG, "probe_ptr[0] must be 0");
    pl.item_ptr.resize(nq + 1);
    for (uint32_t q = 0; q < nq; ++q) {
        if (probe_ptr[q + 1] < probe_ptr[q]) return fail(SVS_ERR_ARG, "probe_ptr not monotone");
        pl.item_ptr[q] = (uint32_t)pl.items.size();
        for (uint32_t p = probe_ptr[q]; p < probe_ptr[q + 1]; ++p) {
            const uint32_t c = probe_lists[p];
            const uint32_t m = c < idx->lists.size() ? idx->lists[c].m : 0;
            pl.items.push_back(Item{q, c < idx->lists.size() ? c : 0, p - probe_ptr[q], m});
            const bool v = verify && verify[q];
            pl.item_verify.push_back(v ? 1 : 0);
            pl.any_verify |= v && m > 0;
            if (m && !idx->lists[c].d.mp8) pl.planes_ok = false;
        }
    }
    pl.item_ptr[nq] = (uint32_t)pl.items.size();
    pl.n_items = (uint32_t)pl.items.size();
    pl.chunk_ptr.resize(pl.n_items + 1);
    pl.cand_ptr.resize(pl.n_items + 1);
    uint64_t ch = 0, cd = 0;
    for (uint32_t i = 0; i < pl.n_items; ++i) {
        pl.chunk_ptr[i] = (uint32_t)ch;
        pl.cand_ptr[i] = (uint32_t)cd;
        ch += (pl.items[i].m + CHUNK_ROWS - 1) / CHUNK_ROWS;
        cd += pl.items[i].m;
    }
    if (cd >= 0xFFFFFFFFull) return fail(SVS_ERR_ARG, "batch has too many candidate rows");
    pl.chunk_ptr[pl.n_items] = (uint32_t)ch;
    pl.cand_ptr[pl.n_items] = (uint32_t)cd;
    pl.n_chunks = (uint32_t)ch;
    pl.n_cands = (uint32_t)cd;
    pl.qchunk_ptr.resize(nq + 1);
    uint32_t tch = 0;
    for (uint32_t q = 0; q < nq; ++q) {
        pl.qchunk_ptr[q] = tch;
        const uint32_t nc = pl.cand_ptr[pl.item_ptr[q + 1]] - pl.cand_ptr[pl.item_ptr[q]];
        tch += (nc + TOPK_WCHUNK - 1) / TOPK_WCHUNK;
    }
    pl.qchunk_ptr[nq] = tch;
    pl.n_tchunks = tch;
    std::vector<uint32_t> seen;
    seen.reserve(pl.n_items);
    for (const Item& it : pl.items) seen.push_back(it.list);
    std::sort(seen.begin(), seen.end());
    seen.erase(std::unique(seen.begin(), seen.end()), seen.end());
    for (uint32_t c : seen)
        if (c < idx->lists.size()) pl.distinct_bytes += (uint64_t)idx->lists[c].m * N * 8;
    // MFMA tiles: list-major, row block, then query group innermost so that
    // consecutive tiles re-read the same rows while they are in cache.
    // (items of empty or out-of-range lists have no rows and no tiles)
    for (uint32_t i = 0; i < pl.n_items; ++i)
        if (pl.items[i].m) pl.gitems.push_back(i);
    std::stable_sort(pl.gitems.begin(), pl.gitems.end(), [&](uint32_t x, uint32_t y) {
        return pl.items[x].list < pl.items[y].list;
    });
    auto make_tiles = [&](const std::vector<uint32_t>& gi, std::vector<MTile>& tiles) {
        const uint32_t ng = (uint32_t)gi.size();
        for (uint32_t g = 0; g < ng;) {
            uint32_t e = g;
            while (e < ng && pl.items[gi[e]].list == pl.items[gi[g]].list) ++e;
            const uint32_t m = pl.items[gi[g]].m;
            for (uint32_t r0 = 0; r0 < m; r0 += tile_rows)
                for (uint32_t q0 = g; q0 < e; q0 += tile_qg)
                    tiles.push_back(MTile{pl.items[gi[g]].list, r0, q0,
                                          std::min<uint32_t>(tile_qg, e - q0)});
            g = e;
        }
    };
    make_tiles(pl.gitems, pl.tiles);
    if (pl.any_verify) {
        for (uint32_t i : pl.gitems)
            if (pl.item_verify[i]) pl.vgitems.push_back(i);
        make_tiles(pl.vgitems, pl.vtiles);
    }
    pl.hdr[PH_NT] = (uint32_t)pl.tiles.size();
    pl.hdr[PH_NVT] = (uint32_t)pl.vtiles.size();
    pl.hdr[PH_NCANDS] = pl.n_cands;
    return SVS_OK;
}

// Per query: status = 0, or 1 + trial + 4 * list of the first failing item.
__global__ void k_query_status(const Item* items, const uint32_t* item_ptr, uint32_t nq,
                               const int32_t* item_verdict, int32_t* q_status, int any_verify) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    int32_t st = 0;
    if (any_verify)
        for (uint32_t it = item_ptr[q]; it < item_ptr[q + 1]; ++it)
            if (item_verdict[it] >= 0) {
                st = 1 + item_verdict[it] + 4 * (int32_t)items[it].list;
                break;
            }
    q_status[q] = st;
}

static uint32_t row_grid(svs_ctx* ctx, uint64_t waves_needed) {
    const uint64_t blocks = (waves_needed + 3) / 4;
    const uint64_t cap = (uint64_t)ctx->n_cu * 16;
    return (uint32_t)std::max<uint64_t>(1, std::min(blocks, cap));
}

// ------------------------------------------------------------ stage timing
// HIP events recorded on the context stream around each pipeline stage when
// profiling is on; elapsed times are accumulated after the batch syncs.
static int prof_mark(svs_slot& sl, int stage, bool begin) {
    if (!sl.prof.on) return SVS_OK;
    HIP_TRY(hipEventRecord(sl.prof.ev[2 * stage + (begin ? 0 : 1)], sl.stream));
    if (!begin) sl.prof.pending |= 1u << stage;
    return SVS_OK;
}

static int prof_collect(svs_slot& sl) {
    if (!sl.prof.on) return SVS_OK;
    for (int s = 0; s < SVS_PROF_STAGES; ++s) {
        if (!(sl.prof.pending & (1u << s))) continue;
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, sl.prof.ev[2 * s], sl.prof.ev[2 * s + 1]));
        sl.prof.ms[s] += ms;
        sl.prof.calls[s] += 1;
    }
    sl.prof.pending = 0;
    return SVS_OK;
}

extern "C" int svs_ctx_profile(svs_ctx* ctx, int enable, int reset, svs_profile* out) {
    if (!ctx) return fail(SVS_ERR_ARG, "null context");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    if (enable && !ctx->prof.created) {
        for (auto& e : ctx->prof.ev) HIP_TRY(hipEventCreate(&e));
        ctx->prof.created = true;
    }
    if (out) {
        for (int s = 0; s < SVS_PROF_STAGES; ++s) {
            out->ms[s] = ctx->prof.ms[s];
            out->calls[s] = ctx->prof.calls[s];
        }
        out->gemv_bytes = ctx->prof.gemv_bytes;
        out->gemv_rows = ctx->prof.gemv_rows;
    }
    if (reset) {
        for (int s = 0; s < SVS_PROF_STAGES; ++s) ctx->prof.ms[s] = 0, ctx->prof.calls[s] = 0;
        ctx->prof.gemv_bytes = ctx->prof.gemv_rows = 0;
    }
    ctx->prof.on = enable != 0;
    return SVS_OK;
}

// Runs encode (+ product, verify, decode, top-k) for one sub-batch whose
// query inputs are already on the device (ctx->q_*). Asynchronous.
// Product kernels: VALU row streaming (one query per row pass) or the int8
// MFMA limb-split tiles (many queries per row pass).
enum { KERNEL_VALU = 0, KERNEL_MFMA = 1, KERNEL_MFMA_WAVE = 2 };

// Block-tile variants of the MFMA path (product, dense decode and decode
// tiles share the plan's tiles): query group, prefetch ring depth, waves per
// SIMD. All give identical results; svs_autotune times them.
struct TileVariant {
    uint32_t qg, pf, occ;
    uint32_t regs;  // 1: B windows built in registers (k_fp_gemm_vb), 0: in LDS
    uint32_t vpf;   // query-word ring depth (LDS kernel)
};
// (variant 8 also builds two steps of B windows per block barrier)
static constexpr TileVariant kTileVariants[SVS_TILE_VARIANTS] = {
    {8, 8, 1, 0, 8}, {8, 2, 2, 0, 2}, {16, 4, 1, 0, 4}, {4, 4, 2, 0, 4}, {16, 2, 1, 0, 2},
    {8, 4, 1, 1, 4}, {8, 2, 1, 1, 2}, {16, 2, 1, 1, 2}, {8, 4, 1, 0, 8}, {16, 4, 1, 0, 8}};

// batch-size class of a tile-variant choice
static int nq_class(uint32_t nq) { return nq <= 16 ? 0 : nq <= 64 ? 1 : nq <= 256 ? 2 : 3; }
static int ctx_variant_for(const svs_ctx* ctx, uint32_t nq) { return ctx->tile_variant[nq_class(nq)]; }

extern "C" int svs_ctx_set_tile_variant(svs_ctx* ctx, uint32_t n_queries, int variant) {
    if (!ctx) return fail(SVS_ERR_ARG, "null context");
    if (variant < 0 || variant >= SVS_TILE_VARIANTS)
        return fail(SVS_ERR_ARG, "tile variant %d out of range", variant);
    std::lock_guard<std::mutex> g(ctx->mu);
    if (n_queries == 0)
        for (int& v : ctx->tile_variant) v = variant;
    else
        ctx->tile_variant[nq_class(n_queries)] = variant;
    return SVS_OK;
}

extern "C" int svs_ctx_get_tile_variant(svs_ctx* ctx, uint32_t n_queries, int* variant,
                                        uint32_t* query_group, uint32_t* ring_depth,
                                        uint32_t* waves_per_simd) {
    if (!ctx || !variant) return fail(SVS_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> g(ctx->mu);
    const int v = ctx_variant_for(ctx, std::max<uint32_t>(n_queries, 1));
    *variant = v;
    if (query_group) *query_group = kTileVariants[v].qg;
    if (ring_depth) *ring_depth = kTileVariants[v].pf;
    if (waves_per_simd) *waves_per_simd = kTileVariants[v].occ;
    return SVS_OK;
}

struct TileArgs {
    const ListDev* lists;
    const MTile* tiles;
    const uint32_t* nt;  // tile count, in device memory
    const uint32_t* gitems;
    const Item* items;
    const uint32_t* cand_ptr;
};

template <int KE, int SRC, int QG, int PF, int OCC, int VPF = PF, int SB = 1>
static void launch_gemm_v(hipStream_t st, uint32_t blocks, const TileArgs& a,
                          const uint64_t* vecs, const uint64_t* qc, uint64_t* out) {
    hipLaunchKernelGGL((k_fp_gemm_tiles<KE, SRC, QG, PF, OCC, VPF, SB>), dim3(blocks), dim3(256), 0, st,
                       a.lists, a.tiles, a.nt, a.gitems, a.items, a.cand_ptr, vecs, qc, out);
}

template <int KE, int SRC, int QG, int PF>
static void launch_gemm_vb(hipStream_t st, uint32_t blocks, const TileArgs& a,
                           const uint64_t* vecs, const uint64_t* qc, uint64_t* out) {
    hipLaunchKernelGGL((k_fp_gemm_vb<KE, SRC, QG, PF>), dim3(blocks), dim3(256), 0, st, a.lists,
                       a.tiles, a.nt, a.gitems, a.items, a.cand_ptr, vecs, qc, out);
}

// The batched server product on the byte-plane kernel for the variant's
// query group (a plan tile holds up to QG queries of one list). The grid is
// the resident blocks (one per CU and wave slot); units are dealt as
// described at k_prod_planes.
template <int KE, int SRC, int QW, int PF, int OCC>
static void launch_planes_v(svs_ctx* ctx, hipStream_t st, uint64_t cap_tiles, const TileArgs& a,
                            const uint32_t* tsel, const int8_t* qplanes, const uint64_t* qc,
                            uint64_t* out) {
    const uint64_t units = 64 * ((std::max<uint64_t>(cap_tiles, 1) + 3) / 4);
    const uint32_t blocks = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>((units + 3) / 4, (uint64_t)ctx->n_cu * OCC * ctx->tile_grid));
    hipLaunchKernelGGL((k_prod_planes<KE, SRC, QW, PF, OCC>), dim3(blocks), dim3(256), 0, st,
                       a.lists, a.tiles, a.nt, tsel, a.gitems, a.items, a.cand_ptr, qplanes, qc,
                       out);
}

// The plane product over the plan's tiles: SRC 0 the server product
// (KE = n over M_enc), SRC 1 the dense decode term (KE = 2 n1 over [AL | H]).
template <int KE, int SRC>
static void launch_planes(svs_ctx* ctx, int v, hipStream_t st, uint64_t cap_tiles,
                          const TileArgs& a, const int8_t* qplanes, const uint64_t* qc,
                          uint64_t* out) {
    constexpr int PF16 = KE == 256 ? 4 : 2;
    switch (kTileVariants[v].qg) {
        case 16:
            if (ctx->prod_ldsb) {
                // whole tiles per block, query planes staged in LDS; the
                // dense decode term (KE = 2 n1) with half-tile waves, two
                // blocks per CU
                constexpr uint32_t per_cu = KE == 2 * N1 ? 2 : 1;
                const uint32_t blocks = (uint32_t)std::max<uint64_t>(
                    1, std::min<uint64_t>(std::max<uint64_t>(cap_tiles, 1),
                                          (uint64_t)ctx->n_cu * per_cu * ctx->tile_grid));
                if constexpr (KE == 2 * N1)
                    hipLaunchKernelGGL((k_prod_planes_lh<KE, SRC, 1>), dim3(blocks), dim3(256), 0,
                                       st, a.lists, a.tiles, a.nt, a.gitems, a.items, a.cand_ptr,
                                       qplanes, qc, out);
                else
                    hipLaunchKernelGGL((k_prod_planes_lb<KE, SRC, PF16>), dim3(blocks), dim3(256), 0,
                                       st, a.lists, a.tiles, a.nt, a.gitems, a.items, a.cand_ptr,
                                       qplanes, qc, out);
            } else {
                launch_planes_v<KE, SRC, 16, PF16, 1>(ctx, st, cap_tiles, a, nullptr, qplanes, qc,
                                                      out);
            }
            break;
        case 4:
            launch_planes_v<KE, SRC, 4, 4, 1>(ctx, st, cap_tiles, a, nullptr, qplanes, qc, out);
            break;
        default:
            launch_planes_v<KE, SRC, 8, 4, 1>(ctx, st, cap_tiles, a, nullptr, qplanes, qc, out);
            break;
    }
}

// The product (SRC 0, KE = n) or dense decode (SRC 1, KE = 2 n1) tiles.
template <int KE, int SRC>
static void launch_gemm(int v, hipStream_t st, uint32_t blocks, const TileArgs& a,
                        const uint64_t* vecs, const uint64_t* qc, uint64_t* out) {
    switch (v) {
        case 1: launch_gemm_v<KE, SRC, 8, 2, 2>(st, blocks, a, vecs, qc, out); break;
        case 2: launch_gemm_v<KE, SRC, 16, 4, 1>(st, blocks, a, vecs, qc, out); break;
        case 4: launch_gemm_v<KE, SRC, 16, 2, 1>(st, blocks, a, vecs, qc, out); break;
        case 5: launch_gemm_vb<KE, SRC, 8, 4>(st, blocks, a, vecs, qc, out); break;
        case 6: launch_gemm_vb<KE, SRC, 8, 2>(st, blocks, a, vecs, qc, out); break;
        case 7: launch_gemm_vb<KE, SRC, 16, 2>(st, blocks, a, vecs, qc, out); break;
        case 8: launch_gemm_v<KE, SRC, 8, 4, 1, 8, 2>(st, blocks, a, vecs, qc, out); break;
        case 9: launch_gemm_v<KE, SRC, 16, 4, 1, 8>(st, blocks, a, vecs, qc, out); break;
        case 3: launch_gemm_v<KE, SRC, 4, 4, 2>(st, blocks, a, vecs, qc, out); break;
        default: launch_gemm_v<KE, SRC, 8, 8, 1>(st, blocks, a, vecs, qc, out); break;
    }
}

template <int QG, int NS = 2>
static void launch_decode_v(hipStream_t st, uint32_t blocks, const TileArgs& a,
                            const int8_t* t_dense, const int8_t* vplanes, const uint64_t* mt_q,
                            const uint64_t* r, const uint64_t* dense, int64_t* scores) {
    hipLaunchKernelGGL((k_decode_tiles<QG, NS>), dim3(blocks), dim3(256), 0, st, a.lists,
                       a.tiles, a.nt, a.gitems, a.items, a.cand_ptr, t_dense, vplanes, mt_q, r,
                       dense, scores);
}

// Decode tiles for the variant's query group: all four row tiles of a wave
// in one pass (one wave per SIMD measured faster than two passes of two).
static void launch_decode_tiles(int v, hipStream_t st, uint32_t blocks, const TileArgs& a,
                                const int8_t* t_dense, const int8_t* vplanes,
                                const uint64_t* mt_q, const uint64_t* r, const uint64_t* dense,
                                int64_t* scores) {
    switch (kTileVariants[v].qg) {
        case 16: launch_decode_v<16>(st, blocks, a, t_dense, vplanes, mt_q, r, dense, scores); break;
        case 4: launch_decode_v<4>(st, blocks, a, t_dense, vplanes, mt_q, r, dense, scores); break;
        default: launch_decode_v<8>(st, blocks, a, t_dense, vplanes, mt_q, r, dense, scores); break;
    }
}

// Tile shape a product kernel's plan is built with.
static void tile_shape(int kernel, int variant, uint32_t& rows, uint32_t& qg) {
    if (kernel == KERNEL_MFMA_WAVE) rows = MT_ROWS, qg = MT_QG;
    else rows = PB_ROWS, qg = kTileVariants[variant].qg;
}

// Plan layout in a slot. prefixes: chunk_ptr [n_items + 1], cand_ptr
// [n_items + 1], item_ptr [nq + 1], qchunk_ptr [nq + 1], hdr [PH_WORDS],
// planner scratch [2 (n_items + 1)]. mtiles: tiles [cap_tiles], gitems
// [n_items], then 16-byte aligned vtiles [cap_tiles], vgitems [n_items].
int svs::plan_layout(svs_ctx* ctx, svs_slot& sl, uint32_t nq, uint32_t n_items,
                     uint64_t cap_tiles, PlanView& v) {
    const uint64_t ni = n_items;
    SVS_TRY(sl.items.ensure(ctx, sizeof(Item) * std::max<uint64_t>(ni, 1)));
    SVS_TRY(sl.prefixes.ensure(ctx, sizeof(uint32_t) * (4 * (ni + 1) + 2 * (nq + 1) + PH_WORDS)));
    SVS_TRY(sl.flags.ensure(ctx, ni + 64));
    const size_t half = (sizeof(MTile) * cap_tiles + sizeof(uint32_t) * ni + 15) & ~(size_t)15;
    SVS_TRY(sl.mtiles.ensure(ctx, 2 * half + 16));
    uint32_t* pre = (uint32_t*)sl.prefixes.ptr;
    v.nq = nq;
    v.n_items = n_items;
    v.items = (const Item*)sl.items.ptr;
    v.chunk_ptr = pre;
    v.cand_ptr = pre + ni + 1;
    v.item_ptr = pre + 2 * (ni + 1);
    v.qchunk_ptr = v.item_ptr + nq + 1;
    v.hdr = v.qchunk_ptr + nq + 1;
    v.flags = (const uint8_t*)sl.flags.ptr;
    v.tiles = (const MTile*)sl.mtiles.ptr;
    v.gitems = (const uint32_t*)(v.tiles + cap_tiles);
    v.vtiles = (const MTile*)((char*)sl.mtiles.ptr + half);
    v.vgitems = (const uint32_t*)(v.vtiles + cap_tiles);
    v.cap_tiles = v.cap_vtiles = cap_tiles;
    return SVS_OK;
}

// Uploads a host plan into the slot's plan layout.
static int upload_plan(svs_ctx* ctx, svs_slot& sl, const BatchPlan& pl, PlanView& v) {
    hipStream_t st = sl.stream;
    const uint64_t cap_tiles = std::max<size_t>(pl.tiles.size(), pl.vtiles.size());
    SVS_TRY(plan_layout(ctx, sl, pl.n_queries, pl.n_items, cap_tiles, v));
    v.cap_cands = pl.n_cands;
    v.cap_chunks = pl.n_chunks;
    v.cap_tiles = pl.tiles.size();
    v.cap_vtiles = pl.vtiles.size();
    v.cap_tchunks = pl.n_tchunks;
    v.any_verify = pl.any_verify;
    v.planes_ok = pl.planes_ok;
    v.distinct_bytes = pl.distinct_bytes;
    v.n_cands = pl.n_cands;
    const uint32_t nq = pl.n_queries, ni = pl.n_items;
    auto up = [&](const void* dst, const void* src, size_t bytes) -> int {
        if (bytes)
            HIP_TRY(hipMemcpyAsync(const_cast<void*>(dst), src, bytes, hipMemcpyHostToDevice, st));
        return SVS_OK;
    };
    SVS_TRY(up(v.items, pl.items.data(), sizeof(Item) * ni));
    SVS_TRY(up(v.chunk_ptr, pl.chunk_ptr.data(), sizeof(uint32_t) * (ni + 1)));
    SVS_TRY(up(v.cand_ptr, pl.cand_ptr.data(), sizeof(uint32_t) * (ni + 1)));
    SVS_TRY(up(v.item_ptr, pl.item_ptr.data(), sizeof(uint32_t) * (nq + 1)));
    SVS_TRY(up(v.qchunk_ptr, pl.qchunk_ptr.data(), sizeof(uint32_t) * (nq + 1)));
    SVS_TRY(up(v.hdr, pl.hdr, sizeof(pl.hdr)));
    SVS_TRY(up(v.flags, pl.item_verify.data(), ni));
    SVS_TRY(up(v.tiles, pl.tiles.data(), sizeof(MTile) * pl.tiles.size()));
    SVS_TRY(up(v.gitems, pl.gitems.data(), sizeof(uint32_t) * pl.gitems.size()));
    SVS_TRY(up(v.vtiles, pl.vtiles.data(), sizeof(MTile) * pl.vtiles.size()));
    SVS_TRY(up(v.vgitems, pl.vgitems.data(), sizeof(uint32_t) * pl.vgitems.size()));
    return SVS_OK;
}

bool svs::device_plan_ok(const svs_index* idx, uint32_t n_items) {
    return n_items <= PLAN_MAX_ITEMS && idx->lists.size() < PLAN_MAX_LISTS;
}

uint64_t svs::top_rows(svs_index* idx, uint32_t n) {
    if (idx->top_rows_dirty) {
        std::vector<uint32_t> m(idx->lists.size());
        for (size_t c = 0; c < m.size(); ++c) m[c] = idx->lists[c].m;
        std::sort(m.begin(), m.end(), std::greater<uint32_t>());
        idx->top_rows_prefix.assign(m.size() + 1, 0);
        for (size_t c = 0; c < m.size(); ++c)
            idx->top_rows_prefix[c + 1] = idx->top_rows_prefix[c] + m[c];
        idx->top_rows_dirty = false;
    }
    return idx->top_rows_prefix[std::min<size_t>(n, idx->lists.size())];
}

int svs::plan_setup(svs_index* idx, svs_slot& sl, uint32_t nq, uint32_t n_items,
                    const uint8_t* verify, uint64_t cands_bound, int variant, PlanView& v) {
    svs_ctx* ctx = idx->ctx;
    if (!device_plan_ok(idx, n_items))
        return fail(SVS_ERR_ARG, "device planning takes up to %u items and %u lists",
                    PLAN_MAX_ITEMS, PLAN_MAX_LISTS - 1);
    const uint64_t cap_cands = std::min<uint64_t>(cands_bound, 0xFFFFFFFEull);
    uint32_t trows, tqg;
    tile_shape(KERNEL_MFMA, variant, trows, tqg);
    // tiles = sum over list groups of ceil(m / trows) ceil(n_g / qg)
    //      <= sum of n_g (m / trows + 1) = rows / trows + items
    const uint64_t cap_tiles = cap_cands / trows + n_items;
    SVS_TRY(plan_layout(ctx, sl, nq, n_items, cap_tiles, v));
    v.cap_cands = cap_cands;
    v.cap_chunks = cap_cands / CHUNK_ROWS + n_items;
    v.cap_tchunks = cap_cands / TOPK_WCHUNK + nq;
    v.any_verify = false;
    if (verify)
        for (uint32_t q = 0; q < nq; ++q) v.any_verify |= verify[q] != 0;
    v.planes_ok = true;
    for (const ListHost& L : idx->lists)
        if (L.m && !L.d.mp8) v.planes_ok = false;
    SVS_TRY(sl.q_probes.ensure(ctx, sizeof(uint32_t) * ((size_t)n_items + nq + 1) + nq + 64));
    v.probes = (const uint32_t*)sl.q_probes.ptr;
    return SVS_OK;
}

int svs::plan_stage(svs_slot& sl, const PlanView& v, const uint32_t* probe_ptr,
                    const uint32_t* probe_lists, const uint8_t* verify) {
    const uint32_t n_items = v.n_items, nq = v.nq;
    uint32_t* d_probes = (uint32_t*)sl.q_probes.ptr;
    uint32_t* d_pptr = d_probes + n_items;
    uint8_t* d_ver = (uint8_t*)(d_pptr + nq + 1);
    hipStream_t st = sl.stream;
    if (probe_lists && n_items)
        HIP_TRY(hipMemcpyAsync(d_probes, probe_lists, sizeof(uint32_t) * n_items,
                               hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d_pptr, probe_ptr, sizeof(uint32_t) * (nq + 1), hipMemcpyHostToDevice,
                           st));
    if (verify)
        HIP_TRY(hipMemcpyAsync(d_ver, verify, nq, hipMemcpyHostToDevice, st));
    else
        hipLaunchKernelGGL(k_fill_u32, dim3(((nq + 3) / 4 + 255) / 256), dim3(256), 0, st,
                           (uint32_t*)d_ver, (nq + 3) / 4, 0u);
    return SVS_OK;
}

int svs::plan_launch(svs_index* idx, svs_slot& sl, const PlanView& v, int variant) {
    svs_ctx* ctx = idx->ctx;
    uint32_t trows, tqg;
    tile_shape(KERNEL_MFMA, variant, trows, tqg);
    const uint32_t n_items = v.n_items, nq = v.nq;
    const uint32_t* d_probes = (const uint32_t*)sl.q_probes.ptr;
    const uint32_t* d_pptr = d_probes + n_items;
    const uint8_t* d_ver = (const uint8_t*)(d_pptr + nq + 1);
    uint32_t* tmp = const_cast<uint32_t*>(v.hdr) + PH_WORDS;
    hipLaunchKernelGGL(k_plan, dim3(1), dim3(PLAN_THREADS), 0, sl.stream, idx->lists_dev,
                       (uint32_t)idx->lists.size(), d_probes, d_pptr, d_ver, nq, n_items, trows,
                       tqg, (uint32_t)v.cap_cands, (uint32_t)v.cap_tchunks,
                       (uint32_t)v.cap_tiles, const_cast<Item*>(v.items),
                       const_cast<uint8_t*>(v.flags), const_cast<uint32_t*>(v.chunk_ptr),
                       const_cast<uint32_t*>(v.cand_ptr), const_cast<uint32_t*>(v.item_ptr),
                       const_cast<uint32_t*>(v.qchunk_ptr), tmp, const_cast<uint32_t*>(v.gitems),
                       const_cast<MTile*>(v.tiles), const_cast<uint32_t*>(v.vgitems),
                       const_cast<MTile*>(v.vtiles), const_cast<uint32_t*>(v.hdr));
    ctx->launches++;
    HIP_TRY(hipGetLastError());
    return SVS_OK;
}

int svs::route_setup(svs_index* idx, svs_slot& sl, uint32_t nq) {
    svs_ctx* ctx = idx->ctx;
    SVS_TRY(sl.q_vq.ensure(ctx, sizeof(uint64_t) * (size_t)nq * N));
    SVS_TRY(sl.q_raw.ensure(ctx, sizeof(float) * (size_t)nq * idx->cent_d));
    // (a cross-tenant table's list count bounds every tenant's)
    SVS_TRY(sl.route_dist.ensure(ctx, sizeof(float) * (size_t)nq * idx->lists.size()));
    return SVS_OK;
}

int svs::route_launch(svs_index* idx, svs_slot& sl, uint32_t nq, uint32_t n_items,
                      double q_scale) {
    svs_ctx* ctx = idx->ctx;
    const uint32_t d = idx->cent_d;
    hipStream_t st = sl.stream;
    const float* x = (const float*)sl.q_raw.ptr;
    hipLaunchKernelGGL(k_quantise, dim3((uint32_t)std::min<uint64_t>(((uint64_t)nq * N + 255) / 256,
                                                                      (uint64_t)ctx->n_cu * 8)),
                       dim3(256), 0, st, x, nq, d, q_scale, (uint64_t*)sl.q_vq.ptr);
    ctx->launches++;
    // per tenant run of queries: its centroids, its list ids shifted to
    // where its lists sit in the batch's list table
    uint32_t* d_probes = (uint32_t*)sl.q_probes.ptr;
    const TenantSeg whole{0, nq, 0, idx};
    const size_t nseg = sl.segs.empty() ? 1 : sl.segs.size();
    size_t dist_off = 0;
    for (size_t si = 0; si < nseg; ++si) {
        const TenantSeg& sg = sl.segs.empty() ? whole : sl.segs[si];
        if (sg.nq == 0) continue;
        const uint32_t nl = (uint32_t)sg.idx->lists.size();
        if (!sg.idx->cent_dev || sg.idx->cent_d != d)
            return fail(SVS_ERR_STATE, "routing needs centroids of dimension %u", d);
        float* dist = (float*)sl.route_dist.ptr + dist_off;
        hipLaunchKernelGGL(k_route_dist, dim3((sg.nq + RT_Q - 1) / RT_Q, (nl + RT_C - 1) / RT_C),
                           dim3(256), 0, st, x + (size_t)sg.q0 * d, sg.nq,
                           (const float*)sg.idx->cent_dev, nl, d, dist);
        uint32_t p2 = 1;
        while (p2 < nl) p2 <<= 1;
        hipLaunchKernelGGL(k_route_select, dim3(sg.nq), dim3(1024), sizeof(uint64_t) * p2, st,
                           (const float*)dist, nl,
                           (const uint32_t*)(d_probes + n_items + sg.q0), sg.list_base, d_probes);
        dist_off += (size_t)sg.nq * nl;
        ctx->launches += 2;
    }
    HIP_TRY(hipGetLastError());
    return SVS_OK;
}

// Blocks of a persistent tile kernel: `resident` blocks fit on a CU at
// once; the grid is that many per CU times ctx->tile_grid (1: every block
// resident from the start, so the tiles t = w, w + G, ... that the XCDs take
// in step are processed together), capped at one block per tile.
static uint32_t tile_blocks(svs_ctx* ctx, uint64_t cap, uint32_t resident) {
    return (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>(cap, (uint64_t)ctx->n_cu * resident * ctx->tile_grid));
}

int svs::run_plan(svs_index* idx, svs_slot& sl, const BatchPlan& pl, uint32_t k,
                  bool product_only, int kernel, bool encode, int variant) {
    PlanView v;
    SVS_TRY(upload_plan(idx->ctx, sl, pl, v));
    return run_view(idx, sl, v, k, product_only, kernel, encode, variant);
}

int svs::run_view(svs_index* idx, svs_slot& sl, const PlanView& v, uint32_t k,
                  bool product_only, int kernel, bool encode, int variant) {
    svs_ctx* ctx = idx->ctx;
    hipStream_t st = sl.stream;
    const uint32_t nq = v.nq;
    const uint32_t ni1 = std::max<uint32_t>(v.n_items, 1);
    const uint64_t nc1 = std::max<uint64_t>(v.cap_cands, 1);
    const Item* d_items = v.items;
    const uint32_t* d_chunk_ptr = v.chunk_ptr;
    const uint32_t* d_cand_ptr = v.cand_ptr;
    const uint32_t* d_item_ptr = v.item_ptr;
    const uint32_t* d_qchunk_ptr = v.qchunk_ptr;
    const uint32_t* d_nt = v.hdr + PH_NT;
    const uint32_t* d_nvt = v.hdr + PH_NVT;
    const ListDev* d_lists = idx->lists_dev;
    // per-query encode
    uint64_t* d_venc = (uint64_t*)sl.q_venc.ptr;
    uint64_t* d_gl = (uint64_t*)sl.q_gl.ptr;
    int8_t* d_tdense = (int8_t*)sl.q_tdense.ptr;
    uint16_t* d_tlist = (uint16_t*)sl.q_tlist.ptr;
    uint32_t* d_tcnt = (uint32_t*)sl.q_tcnt.ptr;
    uint32_t* d_vseed = (uint32_t*)sl.q_vseed.ptr;
    // verdicts, per-query status and the RNG flag (set by encode or verify)
    SVS_TRY(sl.verdict.ensure(ctx, sizeof(int32_t) * (ni1 + nq) + 64));
    int32_t* d_item_verdict = (int32_t*)sl.verdict.ptr;
    int32_t* d_qstatus = d_item_verdict + ni1;
    uint32_t* d_rng_flag = (uint32_t*)(d_qstatus + nq);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, d_rng_flag, 1u, 0u);
    SVS_TRY(prof_mark(sl, 0, true));
    if (encode) {
        hipLaunchKernelGGL(k_encode_keystream, dim3(nq), dim3(256), 0, st,
                           (const uint8_t*)sl.q_seeds.ptr, d_gl, d_tdense, d_tlist, d_tcnt,
                           d_vseed, d_rng_flag);
        // v_enc = L G + T + v and L^T v_enc with each query's own keys: one
        // segment for a batch of one index, one per tenant otherwise
        const TenantSeg whole{0, nq, 0, idx};
        const size_t nseg = sl.segs.empty() ? 1 : sl.segs.size();
        for (size_t si = 0; si < nseg; ++si) {
            const TenantSeg& sg = sl.segs.empty() ? whole : sl.segs[si];
            if (sg.nq == 0) continue;
            if (!sg.idx->lt_dev || !sg.idx->l_dev)
                return fail(SVS_ERR_STATE, "encoding needs an index with keys");
            hipLaunchKernelGGL(k_encode_venc, dim3(N / 256, sg.nq), dim3(256), 0, st,
                               (const uint64_t*)sg.idx->lt_dev,
                               (const uint64_t*)(d_gl + (size_t)sg.q0 * 2 * N1),
                               (const int8_t*)(d_tdense + (size_t)sg.q0 * N),
                               (const uint64_t*)sl.q_vq.ptr + (size_t)sg.q0 * N,
                               d_venc + (size_t)sg.q0 * N);
            hipLaunchKernelGGL(k_lt, dim3(sg.nq), dim3(N1 * LT_PARTS), 0, st,
                               (const uint64_t*)sg.idx->l_dev,
                               (const uint64_t*)(d_venc + (size_t)sg.q0 * N),
                               d_gl + (size_t)sg.q0 * 2 * N1);
            ctx->launches += 2;
        }
        ctx->launches += 1;
        HIP_TRY(hipGetLastError());
    }
    SVS_TRY(prof_mark(sl, 0, false));
    // server product
    SVS_TRY(prof_mark(sl, 1, true));
    SVS_TRY(sl.r.ensure(ctx, sizeof(uint64_t) * nc1));
    if (v.cap_chunks && kernel == KERNEL_VALU) {
        hipLaunchKernelGGL((k_gemv_chunks<N, 0>), dim3(row_grid(ctx, v.cap_chunks)), dim3(ROW_THREADS), 0,
                           st, d_lists, d_items, v.n_items, d_chunk_ptr, d_cand_ptr,
                           (const uint64_t*)d_venc, (uint64_t*)sl.r.ptr);
        ctx->launches++;
        HIP_TRY(hipGetLastError());
    } else if (v.cap_chunks) {
        SVS_TRY(sl.q_corr.ensure(ctx, sizeof(uint64_t) * (3 + MT_BLOCKS) * nq));
        hipLaunchKernelGGL(k_query_corr<N>, dim3(nq), dim3(WAVE), 0, st, (const uint64_t*)d_venc,
                           (uint64_t*)sl.q_corr.ptr);
        if (kernel == KERNEL_MFMA_WAVE) {
            const uint32_t blocks = tile_blocks(ctx, (v.cap_tiles + 3) / 4, 8);
            hipLaunchKernelGGL(k_gemv_mfma, dim3(blocks), dim3(256), 0, st, d_lists, v.tiles, d_nt,
                               v.gitems, d_items, d_cand_ptr, (const uint64_t*)d_venc,
                               (const uint64_t*)sl.q_corr.ptr, (uint64_t*)sl.r.ptr);
        } else {
            // byte planes of the queries, then the plane-by-plane product
            SVS_TRY(sl.q_pplanes.ensure(ctx, (size_t)nq * 8 * N));
            hipLaunchKernelGGL(k_query_planes<N>, dim3(nq), dim3(256), 0, st,
                               (const uint64_t*)d_venc, (int8_t*)sl.q_pplanes.ptr);
            const TileArgs ta{d_lists, v.tiles, d_nt, v.gitems, d_items, d_cand_ptr};
            if (ctx->prod_split && kTileVariants[variant].qg == 16) {
                // the staged kernel over work units of 16..64 rows x 1..4
                // query groups (k_plan_units)
                const uint64_t cap = std::max<uint64_t>(v.cap_tiles, 1);
                SVS_TRY(sl.split.ensure(ctx, sizeof(uint32_t) * (17 * cap + 2)));
                uint32_t* units = (uint32_t*)sl.split.ptr;
                uint32_t* hdr = const_cast<uint32_t*>(v.hdr);
                hipLaunchKernelGGL(k_plan_units, dim3(1), dim3(PLAN_THREADS), 0, st, d_lists,
                                   v.tiles, hdr, units, units + 16 * cap);
                const uint32_t blocks = (uint32_t)std::max<uint64_t>(
                    1, std::min<uint64_t>(16 * cap, (uint64_t)ctx->n_cu * ctx->tile_grid));
                auto staged = [&](auto kern) {
                    hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, st, d_lists, v.tiles,
                                       (const uint32_t*)units, (const uint32_t*)(hdr + PH_NU),
                                       v.gitems, d_items, d_cand_ptr,
                                       (const int8_t*)sl.q_pplanes.ptr,
                                       (const uint64_t*)sl.q_corr.ptr, (uint64_t*)sl.r.ptr);
                };
                // ring depth: steps of 40 KB of LDS per block in flight
                staged(k_prod_staged<2>);
                ctx->launches += 1;
            } else {
                launch_planes<N, 0>(ctx, variant, st, v.cap_tiles, ta,
                                    (const int8_t*)sl.q_pplanes.ptr,
                                    (const uint64_t*)sl.q_corr.ptr, (uint64_t*)sl.r.ptr);
            }
            ctx->launches++;
        }
        ctx->launches += 2;
        HIP_TRY(hipGetLastError());
    }
    SVS_TRY(prof_mark(sl, 1, false));
    sl.prof.gemv_bytes += v.distinct_bytes;
    sl.prof.gemv_rows += v.n_cands;
    if (ctx->fault_list >= 0 && v.n_items) {
        hipLaunchKernelGGL(k_inject_fault, dim3((v.n_items + 255) / 256), dim3(256), 0, st,
                           d_items, v.n_items, d_cand_ptr, (uint32_t)ctx->fault_list,
                           ctx->fault_row, ctx->fault_delta, (uint64_t*)sl.r.ptr);
        HIP_TRY(hipGetLastError());
    }
    if (product_only) return SVS_OK;
    // verification
    SVS_TRY(prof_mark(sl, 2, true));
    if (v.any_verify) {
        SVS_TRY(sl.partials.ensure(ctx, sizeof(uint64_t) * 2 * LAMBDA *
                                            std::max<uint64_t>(v.cap_chunks, 1)));
        // The check's right-hand side sum_i u_t[i] (M_enc[i] . v_enc) needs
        // the row products again, computed independently of the r the server
        // returned: on the MFMA path by the window-formulation tiles
        // (k_fp_gemm_tiles: other operand layouts, other limb products, other
        // epilogue than the byte-plane kernel that produced r) over the
        // verified items only; otherwise k_verify_chunks recomputes them row
        // by row on the VALU.
        const uint64_t* r_check = nullptr;
        if (kernel == KERNEL_MFMA && v.cap_vtiles) {
            SVS_TRY(sl.rv.ensure(ctx, sizeof(uint64_t) * nc1));
            const TileArgs va{d_lists, v.vtiles, d_nvt, v.vgitems, d_items, d_cand_ptr};
            const uint32_t blocks = tile_blocks(ctx, v.cap_vtiles, kTileVariants[variant].occ);
            launch_gemm<N, 0>(variant, st, blocks, va, (const uint64_t*)d_venc,
                              (const uint64_t*)sl.q_corr.ptr, (uint64_t*)sl.rv.ptr);
            ctx->launches++;
            r_check = (const uint64_t*)sl.rv.ptr;
        }
        if (r_check)
            hipLaunchKernelGGL(k_verify_rows,
                               dim3(row_grid(ctx, std::max<uint64_t>(
                                                      (v.cap_chunks + VR_GROUP - 1) / VR_GROUP, 1))),
                               dim3(ROW_THREADS), 0, st, d_items, v.n_items, d_chunk_ptr,
                               d_cand_ptr, (const uint64_t*)sl.r.ptr, r_check,
                               (const uint32_t*)d_vseed, v.flags, (uint64_t*)sl.partials.ptr,
                               d_rng_flag);
        else
            hipLaunchKernelGGL(k_verify_chunks,
                               dim3(row_grid(ctx, std::max<uint64_t>(v.cap_chunks, 1))),
                               dim3(ROW_THREADS), 0, st, d_lists, d_items, v.n_items, d_chunk_ptr,
                               d_cand_ptr, (const uint64_t*)d_venc, (const uint64_t*)sl.r.ptr,
                               r_check, (const uint32_t*)d_vseed, v.flags,
                               (uint64_t*)sl.partials.ptr, d_rng_flag);
        hipLaunchKernelGGL(k_verify_finalize,
                           dim3((uint32_t)std::min<uint64_t>(ni1, (uint64_t)ctx->n_cu * 8)),
                           dim3(VF_THREADS), 0, st, d_items,
                           v.n_items, d_chunk_ptr, v.flags, (const uint64_t*)sl.partials.ptr,
                           d_item_verdict);
        ctx->launches += 2;
        HIP_TRY(hipGetLastError());
    }
    SVS_TRY(prof_mark(sl, 2, false));
    hipLaunchKernelGGL(k_query_status, dim3((nq + 255) / 256), dim3(256), 0, st, d_items,
                       d_item_ptr, nq, (const int32_t*)d_item_verdict, d_qstatus,
                       v.any_verify ? 1 : 0);
    // decode
    SVS_TRY(sl.scores.ensure(ctx, sizeof(int64_t) * nc1));
    SVS_TRY(prof_mark(sl, 3, true));
    if (v.cap_cands) {
        // The dense decode terms AL G + H L^T v_enc: block tiles over the
        // row-major [AL | H] on the product's tiles when the batch runs on
        // block tiles, else a row-streaming pass before the per-candidate
        // decode kernel.
        const uint64_t* dense = nullptr;
        const bool tiles = kernel == KERNEL_MFMA && v.cap_chunks;
        const TileArgs ta{d_lists, v.tiles, d_nt, v.gitems, d_items, d_cand_ptr};
        if (!tiles && v.cap_chunks) {
            // one query per list: the dense terms stream the row-major
            // [AL | H] like the server product streams M_enc
            SVS_TRY(sl.dense.ensure(ctx, sizeof(uint64_t) * nc1));
            hipLaunchKernelGGL((k_gemv_chunks<2 * N1, 1>), dim3(row_grid(ctx, v.cap_chunks)),
                               dim3(ROW_THREADS), 0, st, d_lists, d_items, v.n_items,
                               d_chunk_ptr, d_cand_ptr, (const uint64_t*)d_gl,
                               (uint64_t*)sl.dense.ptr);
            ctx->launches++;
            dense = (const uint64_t*)sl.dense.ptr;
        }
        if (tiles) {
            // the dense terms AL G + H L^T v_enc over [AL | H] (K = 256) on
            // the plane kernel (with the plane_fold epilogue 0.17 ms faster
            // per 256-query batch than the window-formulation tiles, which
            // remain as ctx->dense_planes = 0)
            SVS_TRY(sl.dense.ensure(ctx, sizeof(uint64_t) * nc1));
            uint64_t* qc_gl = (uint64_t*)sl.q_corr.ptr + nq;
            hipLaunchKernelGGL(k_query_corr<2 * N1>, dim3(nq), dim3(WAVE), 0, st,
                               (const uint64_t*)d_gl, qc_gl);
            if (ctx->dense_planes) {
                // the plane kernel over [AL | H]; the product's query planes
                // are spent, so their buffer takes [G ; L^T v_enc]'s
                hipLaunchKernelGGL(k_query_planes<2 * N1>, dim3(nq), dim3(256), 0, st,
                                   (const uint64_t*)d_gl, (int8_t*)sl.q_pplanes.ptr);
                launch_planes<2 * N1, 1>(ctx, variant, st, v.cap_tiles, ta,
                                         (const int8_t*)sl.q_pplanes.ptr,
                                         (const uint64_t*)qc_gl, (uint64_t*)sl.dense.ptr);
                ctx->launches++;
            } else {
                launch_gemm<2 * N1, 1>(variant, st,
                                       tile_blocks(ctx, v.cap_tiles, kTileVariants[variant].occ),
                                       ta, (const uint64_t*)d_gl, (const uint64_t*)qc_gl,
                                       (uint64_t*)sl.dense.ptr);
            }
            ctx->launches += 2;
            dense = (const uint64_t*)sl.dense.ptr;
        }
        if (tiles && v.planes_ok) {
            // M T and S v_enc on the same tiles; the decode ends there.
            SVS_TRY(sl.q_vplanes.ensure(ctx, (size_t)nq * 8 * N));
            uint64_t* mt_q = (uint64_t*)sl.q_corr.ptr + 2 * nq;
            hipLaunchKernelGGL(k_decode_prep, dim3(nq), dim3(256), 0, st, (const uint64_t*)d_venc,
                               (const int8_t*)d_tdense, (int8_t*)sl.q_vplanes.ptr, mt_q);
            launch_decode_tiles(variant, st, tile_blocks(ctx, v.cap_tiles, 1), ta,
                                (const int8_t*)d_tdense, (const int8_t*)sl.q_vplanes.ptr,
                                (const uint64_t*)mt_q, (const uint64_t*)sl.r.ptr, dense,
                                (int64_t*)sl.scores.ptr);
            ctx->launches += 2;
        } else {
            const uint32_t blocks = (uint32_t)std::min<uint64_t>(
                (v.cap_cands + DC_ROWS - 1) / DC_ROWS, (uint64_t)ctx->n_cu * 32);
            if (!dense) return fail(SVS_ERR_STATE, "decode without its dense terms");
            hipLaunchKernelGGL(k_decode_cands, dim3(blocks), dim3(DC_THREADS), 0, st, d_lists,
                               d_items, v.n_items, d_cand_ptr, (const uint64_t*)sl.r.ptr,
                               (const uint64_t*)d_venc, (const uint16_t*)d_tlist,
                               (const uint32_t*)d_tcnt, dense, (int64_t*)sl.scores.ptr,
                               (uint64_t*)nullptr);
            ctx->launches++;
        }
        HIP_TRY(hipGetLastError());
    }
    SVS_TRY(prof_mark(sl, 3, false));
    // top-k
    SVS_TRY(prof_mark(sl, 4, true));
    SVS_TRY(sl.out.ensure(ctx, (sizeof(uint32_t) + sizeof(int64_t)) * (size_t)nq * k +
                                     sizeof(uint32_t) * nq + 64));
    uint32_t* d_oids = (uint32_t*)sl.out.ptr;
    int64_t* d_oscores = (int64_t*)(d_oids + (size_t)nq * k + ((nq * k) & 1));
    uint32_t* d_ocount = (uint32_t*)(d_oscores + (size_t)nq * k);
    if (k <= (uint32_t)TOPK_LANE) {
        SVS_TRY(sl.topk_part.ensure(ctx, sizeof(Cand) * std::max<uint64_t>(v.cap_tchunks, 1) * k));
        if (v.cap_tchunks)
            hipLaunchKernelGGL(k_topk_chunks,
                               dim3((uint32_t)std::min<uint64_t>((v.cap_tchunks + 3) / 4,
                                                                 (uint64_t)ctx->n_cu * 16)),
                               dim3(256), 0, st, d_qchunk_ptr, nq, d_item_ptr, d_cand_ptr,
                               (const int64_t*)sl.scores.ptr, k, (Cand*)sl.topk_part.ptr);
        hipLaunchKernelGGL((k_topk_final<256>), dim3(nq), dim3(256), 0, st, d_items, d_item_ptr,
                           d_cand_ptr, d_lists, d_qchunk_ptr, (const Cand*)sl.topk_part.ptr,
                           (const int32_t*)d_qstatus, k, d_oids, d_oscores, d_ocount);
        ctx->launches++;
    } else if (k <= 128)
        hipLaunchKernelGGL((k_topk<128, 32>), dim3(nq), dim3(32), 0, st, d_items, d_item_ptr,
                           d_cand_ptr, d_lists, (const int64_t*)sl.scores.ptr,
                           (const int32_t*)d_qstatus, k, d_oids, d_oscores, d_ocount);
    else
        hipLaunchKernelGGL(k_topk_large, dim3(nq), dim3(TOPK_LARGE_THREADS), 0, st, d_items,
                           d_item_ptr, d_cand_ptr, d_lists, (const int64_t*)sl.scores.ptr,
                           (const int32_t*)d_qstatus, k, d_oids, d_oscores, d_ocount);
    ctx->launches += 2;
    HIP_TRY(hipGetLastError());
    SVS_TRY(prof_mark(sl, 4, false));
    return SVS_OK;
}

int svs::view_reserve(svs_ctx* ctx, svs_slot& sl, const PlanView& v, uint32_t k) {
    const uint32_t nq = v.nq;
    const uint64_t ni1 = std::max<uint32_t>(v.n_items, 1), nc1 = std::max<uint64_t>(v.cap_cands, 1);
    SVS_TRY(sl.verdict.ensure(ctx, sizeof(int32_t) * (ni1 + nq) + 64));
    SVS_TRY(sl.r.ensure(ctx, sizeof(uint64_t) * nc1));
    SVS_TRY(sl.rv.ensure(ctx, sizeof(uint64_t) * nc1));
    SVS_TRY(sl.scores.ensure(ctx, sizeof(int64_t) * nc1));
    SVS_TRY(sl.dense.ensure(ctx, sizeof(uint64_t) * nc1));
    SVS_TRY(sl.q_corr.ensure(ctx, sizeof(uint64_t) * (3 + MT_BLOCKS) * nq));
    SVS_TRY(sl.partials.ensure(ctx, sizeof(uint64_t) * 2 * LAMBDA *
                                        std::max<uint64_t>(v.cap_chunks, 1)));
    SVS_TRY(sl.q_vplanes.ensure(ctx, (size_t)nq * 8 * N));
    SVS_TRY(sl.q_pplanes.ensure(ctx, (size_t)nq * 8 * N));
    SVS_TRY(sl.out.ensure(ctx, (sizeof(uint32_t) + sizeof(int64_t)) * (size_t)nq * k +
                                   sizeof(uint32_t) * nq + 64));
    if (k <= (uint32_t)TOPK_LANE)
        SVS_TRY(sl.topk_part.ensure(ctx, sizeof(Cand) * std::max<uint64_t>(v.cap_tchunks, 1) * k));
    return SVS_OK;
}

uint64_t svs::cap_bucket(uint64_t rows) {
    uint64_t c = 1ull << 16;
    while (c < rows) c <<= 1;
    return c;
}

void svs::release_graphs(svs_slot& sl) {
    for (auto& g : sl.graphs)
        if (g.exec) hipGraphExecDestroy(g.exec);
    sl.graphs.clear();
    sl.graph_seen.clear();
}

int svs::run_graphed(svs_ctx* ctx, svs_slot& sl, std::vector<uint64_t> key,
                     const std::function<int()>& enqueue) {
    if (!ctx->graphs || sl.prof.on || ctx->fault_list >= 0) return enqueue();
    key.push_back((uint64_t)ctx->prod_split);
    key.push_back((uint64_t)ctx->prod_split_pf);
    key.push_back((uint64_t)ctx->dense_planes);
    key.push_back((uint64_t)ctx->prod_ldsb);
    for (DevBuf* b : sl.slot_scratch()) {
        key.push_back((uint64_t)(uintptr_t)b->ptr);
        key.push_back(b->cap);
    }
    const uint64_t now = ++sl.graph_clock;
    for (auto& g : sl.graphs)
        if (g.key == key) {
            g.last_use = now;
            HIP_TRY(hipGraphLaunch(g.exec, sl.stream));
            ++sl.graph_launches;
            return SVS_OK;
        }
    auto it = std::find(sl.graph_seen.begin(), sl.graph_seen.end(), key);
    if (it == sl.graph_seen.end()) {
        if (sl.graph_seen.size() >= 64) sl.graph_seen.erase(sl.graph_seen.begin());
        sl.graph_seen.push_back(std::move(key));
        return enqueue();
    }
    sl.graph_seen.erase(it);
    HIP_TRY(hipStreamBeginCapture(sl.stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue();
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(sl.stream, &graph);
    hipGraphExec_t exec = nullptr;
    hipError_t ei = hipErrorUnknown;
    if (rc == SVS_OK && e == hipSuccess && graph)
        ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (graph) hipGraphDestroy(graph);
    if (rc != SVS_OK || e != hipSuccess || ei != hipSuccess) {
        // not capturable after all: run it as launches
        (void)hipGetLastError();
        if (exec) hipGraphExecDestroy(exec);
        return enqueue();
    }
    if (sl.graphs.size() >= 32) {
        size_t lru = 0;
        for (size_t i = 1; i < sl.graphs.size(); ++i)
            if (sl.graphs[i].last_use < sl.graphs[lru].last_use) lru = i;
        HIP_TRY(hipStreamSynchronize(sl.stream));
        hipGraphExecDestroy(sl.graphs[lru].exec);
        sl.graphs.erase(sl.graphs.begin() + lru);
    }
    sl.graphs.push_back(svs_slot::Graph{std::move(key), exec, now});
    ++sl.graph_captures;
    HIP_TRY(hipGraphLaunch(exec, sl.stream));
    ++sl.graph_launches;
    return SVS_OK;
}

int svs::stage_queries(svs_ctx* ctx, svs_slot& sl, uint32_t nq, const uint64_t* v_q,
                  const uint8_t* seeds) {
    SVS_TRY(sl.q_vq.ensure(ctx, sizeof(uint64_t) * (size_t)nq * N));
    SVS_TRY(sl.q_seeds.ensure(ctx, (size_t)nq * 32));
    SVS_TRY(sl.q_venc.ensure(ctx, sizeof(uint64_t) * (size_t)nq * N));
    SVS_TRY(sl.q_gl.ensure(ctx, sizeof(uint64_t) * (size_t)nq * 2 * N1));
    SVS_TRY(sl.q_tdense.ensure(ctx, (size_t)nq * N));
    SVS_TRY(sl.q_tlist.ensure(ctx, sizeof(uint16_t) * (size_t)nq * N));
    SVS_TRY(sl.q_tcnt.ensure(ctx, sizeof(uint32_t) * nq));
    SVS_TRY(sl.q_vseed.ensure(ctx, sizeof(uint32_t) * 8 * nq));
    if (v_q)
        HIP_TRY(hipMemcpyAsync(sl.q_vq.ptr, v_q, sizeof(uint64_t) * (size_t)nq * N,
                               hipMemcpyDefault, sl.stream));
    if (seeds)
        HIP_TRY(hipMemcpyAsync(sl.q_seeds.ptr, seeds, (size_t)nq * 32, hipMemcpyDefault,
                               sl.stream));
    return SVS_OK;
}

// Grows a slot's batch scratch to hold a batch of nq queries, n_items
// (query, list) items and n_cands candidate rows with top-k of k, so that
// serving does not reallocate (a device free synchronises every stream).
int svs::reserve_slot(svs_ctx* ctx, svs_slot& sl, uint32_t nq, uint32_t n_items,
                      uint64_t n_cands, uint32_t k) {
    const uint64_t ni = std::max<uint32_t>(n_items, 1), nc = std::max<uint64_t>(n_cands, 1);
    const uint64_t chunks = nc / CHUNK_ROWS + ni;
    const uint64_t tiles = nc / 64 + ni;  // generous for any tile shape
    PlanView v;
    SVS_TRY(plan_layout(ctx, sl, nq, (uint32_t)ni, tiles, v));
    SVS_TRY(sl.q_probes.ensure(ctx, sizeof(uint32_t) * ni + nq + 64));
    SVS_TRY(sl.q_vq.ensure(ctx, sizeof(uint64_t) * (size_t)nq * N));
    SVS_TRY(sl.q_seeds.ensure(ctx, (size_t)nq * 32));
    SVS_TRY(sl.q_venc.ensure(ctx, sizeof(uint64_t) * (size_t)nq * N));
    SVS_TRY(sl.q_gl.ensure(ctx, sizeof(uint64_t) * (size_t)nq * 2 * N1));
    SVS_TRY(sl.q_tdense.ensure(ctx, (size_t)nq * N));
    SVS_TRY(sl.q_tlist.ensure(ctx, sizeof(uint16_t) * (size_t)nq * N));
    SVS_TRY(sl.q_tcnt.ensure(ctx, sizeof(uint32_t) * nq));
    SVS_TRY(sl.q_vseed.ensure(ctx, sizeof(uint32_t) * 8 * nq));
    SVS_TRY(sl.q_vplanes.ensure(ctx, (size_t)nq * 8 * N));
    SVS_TRY(sl.q_pplanes.ensure(ctx, (size_t)nq * 8 * N));
    SVS_TRY(sl.q_corr.ensure(ctx, sizeof(uint64_t) * (3 + MT_BLOCKS) * nq));
    SVS_TRY(sl.r.ensure(ctx, sizeof(uint64_t) * nc));
    SVS_TRY(sl.scores.ensure(ctx, sizeof(int64_t) * nc));
    SVS_TRY(sl.dense.ensure(ctx, sizeof(uint64_t) * nc));
    SVS_TRY(sl.rv.ensure(ctx, sizeof(uint64_t) * nc));
    SVS_TRY(sl.partials.ensure(ctx, sizeof(uint64_t) * 2 * LAMBDA * chunks));
    SVS_TRY(sl.verdict.ensure(ctx, sizeof(int32_t) * (ni + nq) + 64));
    SVS_TRY(sl.out.ensure(ctx, (sizeof(uint32_t) + sizeof(int64_t)) * (size_t)nq * k +
                                   sizeof(uint32_t) * nq + 64));
    SVS_TRY(sl.topk_part.ensure(ctx, sizeof(Cand) * (nc / TOPK_WCHUNK + nq) * k));
    return SVS_OK;
}

// Enqueues the batch's result copies (ids, scores, counts, per-query status
// and the RNG flag) on the slot's stream.
int svs::collect_async(svs_slot& sl, uint32_t nq, uint32_t k, uint32_t n_items, uint32_t* ids,
                  int64_t* scores, uint32_t* count, int32_t* status, uint32_t* rng_flag) {
    hipStream_t st = sl.stream;
    uint32_t* d_oids = (uint32_t*)sl.out.ptr;
    int64_t* d_oscores = (int64_t*)(d_oids + (size_t)nq * k + ((nq * k) & 1));
    uint32_t* d_ocount = (uint32_t*)(d_oscores + (size_t)nq * k);
    int32_t* d_qstatus = (int32_t*)sl.verdict.ptr + std::max<uint32_t>(n_items, 1);
    uint32_t* d_rng_flag = (uint32_t*)(d_qstatus + nq);
    HIP_TRY(hipMemcpyAsync(ids, d_oids, sizeof(uint32_t) * (size_t)nq * k, hipMemcpyDefault, st));
    HIP_TRY(hipMemcpyAsync(scores, d_oscores, sizeof(int64_t) * (size_t)nq * k, hipMemcpyDefault, st));
    HIP_TRY(hipMemcpyAsync(count, d_ocount, sizeof(uint32_t) * nq, hipMemcpyDefault, st));
    HIP_TRY(hipMemcpyAsync(status, d_qstatus, sizeof(int32_t) * nq, hipMemcpyDefault, st));
    HIP_TRY(hipMemcpyAsync(rng_flag, d_rng_flag, sizeof(uint32_t), hipMemcpyDefault, st));
    return SVS_OK;
}

// The batch's RNG flag: a Freivalds u-draw equal to p, or an encode
// keystream that ran out (both have probability < 2^-60).
int svs::check_rng_flag(uint32_t rng_flag) {
    if (rng_flag & 1u)
        return fail(SVS_ERR_RNG, "a Freivalds u-draw equalled p (probability 2^-61 per draw)");
    if (rng_flag & 2u)
        return fail(SVS_ERR_RNG, "a query's encode keystream ran past %d words", ENC_WORDS);
    return SVS_OK;
}

// plan_hdr: a device plan's header, whose overflow word is checked too
// (its bounds are exact or upper bounds, so a set flag is an internal error).
static int collect(svs_ctx* ctx, svs_slot& sl, uint32_t nq, uint32_t k, uint32_t n_items,
                   uint32_t* ids, int64_t* scores, uint32_t* count, int32_t* status,
                   const uint32_t* plan_hdr = nullptr) {
    (void)ctx;
    uint32_t rng_flag = 0, over = 0;
    SVS_TRY(collect_async(sl, nq, k, n_items, ids, scores, count, status, &rng_flag));
    if (plan_hdr)
        HIP_TRY(hipMemcpyAsync(&over, plan_hdr + PH_OVERFLOW, sizeof(uint32_t),
                               hipMemcpyDeviceToHost, sl.stream));
    HIP_TRY(hipStreamSynchronize(sl.stream));
    SVS_TRY(prof_collect(sl));
    if (over) return fail(SVS_ERR_STATE, "device plan exceeded its bounds");
    return check_rng_flag(rng_flag);
}

// Kernel, tile variant and plan for a batch run in `mode`.
int svs::plan_for_mode(svs_index* idx, int mode, uint32_t nq, const uint32_t* probe_ptr,
                  const uint32_t* probe_lists, const uint8_t* verify, BatchPlan& pl,
                  int& kernel, int& variant) {
    kernel = mode == SVS_MODE_BATCHED ? KERNEL_MFMA : KERNEL_VALU;
    variant = ctx_variant_for(idx->ctx, nq);
    uint32_t trows, tqg;
    tile_shape(kernel, variant, trows, tqg);
    SVS_TRY(make_plan(idx, nq, probe_ptr, probe_lists, verify, pl, trows, tqg));
    for (const Item& it : pl.items)
        if (it.m && !idx->lists[it.list].d.mp_t32 && !idx->lists[it.list].d.mp_t64)
            return fail(SVS_ERR_STATE, "list %u was uploaded without M_plain", it.list);
    return SVS_OK;
}

// Rows and validity of a batch's probe lists (host inputs): the exact
// candidate-row count, and the checks make_plan / plan_for_mode apply.
static int probe_rows(const svs_index* idx, uint32_t nq, const uint32_t* probe_ptr,
                      const uint32_t* probe_lists, uint64_t& rows) {
    if (probe_ptr[0] != 0) return fail(SVS_ERR_ARG, "probe_ptr[0] must be 0");
    rows = 0;
    for (uint32_t q = 0; q < nq; ++q) {
        if (probe_ptr[q + 1] < probe_ptr[q]) return fail(SVS_ERR_ARG, "probe_ptr not monotone");
        for (uint32_t p = probe_ptr[q]; p < probe_ptr[q + 1]; ++p) {
            const uint32_t c = probe_lists[p];
            if (c >= idx->lists.size() || !idx->lists[c].m) continue;
            const ListHost& L = idx->lists[c];
            if (!L.d.mp_t32 && !L.d.mp_t64)
                return fail(SVS_ERR_STATE, "list %u was uploaded without M_plain", c);
            rows += L.m;
        }
    }
    if (rows >= 0xFFFFFFFFull) return fail(SVS_ERR_ARG, "batch has too many candidate rows");
    return SVS_OK;
}


// One batched run (mode BATCHED or BATCHED_VALU) planned on the device: the
// host only validates the probe lists and counts their rows.
// Pinned staging of a batch's host inputs in ctx->h_stage: probe_ptr,
// probe lists, verify flags, then (optionally) raw f32 queries, v_q and seeds.
struct Staged {
    uint32_t* pptr = nullptr;
    uint32_t* plist = nullptr;
    uint8_t* ver = nullptr;
    float* x = nullptr;
    uint64_t* vq = nullptr;
    uint8_t* seeds = nullptr;
};
static int stage_host(svs_ctx* ctx, uint32_t nq, const uint32_t* probe_ptr,
                      const uint32_t* probe_lists, const uint8_t* verify, uint32_t d,
                      const float* x, const uint64_t* vq, const uint8_t* seeds, Staged& o) {
    const uint32_t ni = probe_ptr[nq];
    auto al = [](size_t b) { return (b + 63) & ~(size_t)63; };
    const size_t o_plist = al(sizeof(uint32_t) * (nq + 1));
    const size_t o_ver = o_plist + al(sizeof(uint32_t) * ni);
    const size_t o_x = o_ver + al(nq);
    const size_t o_vq = o_x + al(x ? sizeof(float) * (size_t)nq * d : 0);
    const size_t o_seeds = o_vq + al(vq ? sizeof(uint64_t) * (size_t)nq * N : 0);
    const size_t total = o_seeds + al(seeds ? (size_t)nq * 32 : 0);
    SVS_TRY(ctx->h_stage.ensure(total));
    char* base = (char*)ctx->h_stage.ptr;
    o.pptr = (uint32_t*)base;
    for (uint32_t q = 0; q <= nq; ++q) o.pptr[q] = probe_ptr[q] - probe_ptr[0];
    o.plist = (uint32_t*)(base + o_plist);
    if (probe_lists) std::memcpy(o.plist, probe_lists + probe_ptr[0], sizeof(uint32_t) * ni);
    o.ver = (uint8_t*)(base + o_ver);
    if (verify) std::memcpy(o.ver, verify, nq);
    else std::memset(o.ver, 0, nq);
    if (x) std::memcpy(o.x = (float*)(base + o_x), x, sizeof(float) * (size_t)nq * d);
    if (vq) std::memcpy(o.vq = (uint64_t*)(base + o_vq), vq, sizeof(uint64_t) * (size_t)nq * N);
    if (seeds) std::memcpy(o.seeds = (uint8_t*)(base + o_seeds), seeds, (size_t)nq * 32);
    return SVS_OK;
}

// One batched run (mode BATCHED or BATCHED_VALU) planned on the device: the
// host only validates the probe lists and counts their rows. dev: v_q,
// seeds and outputs are device pointers.
static int batch_device_planned(svs_index* idx, const svs_query_batch* b, int mode, bool dev) {
    svs_ctx* ctx = idx->ctx;
    svs_slot& sl = *ctx;
    const uint32_t nq = b->n_queries;
    uint64_t rows = 0;
    SVS_TRY(probe_rows(idx, nq, b->probe_ptr, b->probe_lists, rows));
    const int kernel = mode == SVS_MODE_BATCHED ? KERNEL_MFMA : KERNEL_VALU;
    const int variant = ctx_variant_for(ctx, nq);
    const uint32_t n_items = b->probe_ptr[nq];
    const uint64_t cap = ctx->graphs ? cap_bucket(rows) : rows;
    Staged h;
    SVS_TRY(stage_host(ctx, nq, b->probe_ptr, b->probe_lists, b->verify, 0, nullptr,
                       dev ? nullptr : b->v_q, dev ? nullptr : b->q_seeds, h));
    PlanView v;
    SVS_TRY(plan_setup(idx, sl, nq, n_items, h.ver, cap, variant, v));
    SVS_TRY(stage_queries(ctx, sl, nq, nullptr, nullptr));
    SVS_TRY(view_reserve(ctx, sl, v, b->k));
    if (sl.prof.on) {
        // profile counters: the product's algorithmic bytes (distinct lists)
        std::vector<uint8_t> seen(idx->lists.size(), 0);
        for (uint32_t i = 0; i < n_items; ++i) {
            const uint32_t c = h.plist[i];
            if (c < seen.size() && !seen[c]) {
                seen[c] = 1;
                v.distinct_bytes += (uint64_t)idx->lists[c].m * N * 8;
            }
        }
        v.n_cands = rows;
    }
    if (dev) {
        HIP_TRY(hipMemcpyAsync(sl.q_vq.ptr, b->v_q, sizeof(uint64_t) * (size_t)nq * N,
                               hipMemcpyDeviceToDevice, sl.stream));
        HIP_TRY(hipMemcpyAsync(sl.q_seeds.ptr, b->q_seeds, (size_t)nq * 32,
                               hipMemcpyDeviceToDevice, sl.stream));
    }
    const std::vector<uint64_t> key{1, idx->gen, nq, n_items, b->k,
                                    (uint64_t)kernel, (uint64_t)variant, v.any_verify,
                                    v.planes_ok, cap, dev, (uint64_t)(uintptr_t)h.pptr};
    SVS_TRY(run_graphed(ctx, sl, key, [&]() -> int {
        SVS_TRY(plan_stage(sl, v, h.pptr, h.plist, h.ver));
        if (!dev) {
            HIP_TRY(hipMemcpyAsync(sl.q_vq.ptr, h.vq, sizeof(uint64_t) * (size_t)nq * N,
                                   hipMemcpyHostToDevice, sl.stream));
            HIP_TRY(hipMemcpyAsync(sl.q_seeds.ptr, h.seeds, (size_t)nq * 32,
                                   hipMemcpyHostToDevice, sl.stream));
        }
        SVS_TRY(plan_launch(idx, sl, v, variant));
        return run_view(idx, sl, v, b->k, false, kernel, true, variant);
    }));
    return collect(ctx, sl, nq, b->k, v.n_items, b->out_ids, b->out_scores, b->out_count,
                   b->out_status, v.hdr);
}

// dev: v_q, q_seeds and the outputs are device pointers (checked for
// canonical values on the device); otherwise host pointers.
static int score_batch_impl(svs_index* idx, const svs_query_batch* b, int mode, bool dev) {
    if (!idx || !b) return fail(SVS_ERR_ARG, "null argument");
    if (b->n_queries == 0) return SVS_OK;
    if (!b->v_q || !b->q_seeds || !b->probe_ptr || !b->probe_lists || !b->out_ids ||
        !b->out_scores || !b->out_count || !b->out_status)
        return fail(SVS_ERR_ARG, "missing batch array");
    if (b->k == 0 || b->k > (uint32_t)TOPK_LARGE_MAX)
        return fail(SVS_ERR_ARG, "k = %u must be in 1..%d", b->k, TOPK_LARGE_MAX);
    if (mode != SVS_MODE_PLAIN && mode != SVS_MODE_BATCHED && mode != SVS_MODE_BATCHED_VALU)
        return fail(SVS_ERR_ARG, "unknown mode %d", mode);
    svs_ctx* ctx = idx->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    uint32_t* d_badv = nullptr;
    if (dev) {
        SVS_TRY(ctx->io_b.ensure(ctx, 64));
        d_badv = (uint32_t*)ctx->io_b.ptr;
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, d_badv, 1u, 0u);
        const uint64_t n = (uint64_t)b->n_queries * N;
        hipLaunchKernelGGL(k_check_canonical, dim3((uint32_t)std::min<uint64_t>((n + 255) / 256, 4096)),
                           dim3(256), 0, ctx->stream, b->v_q, n, d_badv);
        HIP_TRY(hipGetLastError());
    } else {
        SVS_TRY(check_canonical(b->v_q, (size_t)b->n_queries * N, "v_q"));
    }
    const uint32_t nq_total = b->n_queries;
    if (mode != SVS_MODE_PLAIN && ctx->planner &&
        device_plan_ok(idx, b->probe_ptr[nq_total] - b->probe_ptr[0])) {
        SVS_TRY(batch_device_planned(idx, b, mode, dev));
        if (dev) {
            uint32_t bad = 0;
            HIP_TRY(hipMemcpy(&bad, d_badv, sizeof(uint32_t), hipMemcpyDeviceToHost));
            if (bad) return fail(SVS_ERR_ARG, "v_q holds a non-canonical value (>= p)");
        }
        return SVS_OK;
    }
    const uint32_t step = mode == SVS_MODE_PLAIN ? 1 : nq_total;
    for (uint32_t q0 = 0; q0 < nq_total; q0 += step) {
        const uint32_t nq = std::min(step, nq_total - q0);
        std::vector<uint32_t> pptr(nq + 1);
        for (uint32_t q = 0; q <= nq; ++q) pptr[q] = b->probe_ptr[q0 + q] - b->probe_ptr[q0];
        BatchPlan pl;
        int kernel, variant;
        SVS_TRY(plan_for_mode(idx, mode, nq, pptr.data(), b->probe_lists + b->probe_ptr[q0],
                              b->verify ? b->verify + q0 : nullptr, pl, kernel, variant));
        SVS_TRY(stage_queries(ctx, *ctx, nq, b->v_q + (size_t)q0 * N,
                              b->q_seeds + (size_t)q0 * 32));
        SVS_TRY(run_plan(idx, *ctx, pl, b->k, false, kernel, true, variant));
        SVS_TRY(collect(ctx, *ctx, nq, b->k, pl.n_items, b->out_ids + (size_t)q0 * b->k,
                        b->out_scores + (size_t)q0 * b->k, b->out_count + q0,
                        b->out_status + q0));
    }
    if (dev) {
        uint32_t bad = 0;
        HIP_TRY(hipMemcpy(&bad, d_badv, sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (bad) return fail(SVS_ERR_ARG, "v_q holds a non-canonical value (>= p)");
    }
    return SVS_OK;
}

extern "C" int svs_score_batch(svs_index* idx, const svs_query_batch* b, int mode) {
    return score_batch_impl(idx, b, mode, false);
}

extern "C" int svs_score_batch_dev(svs_index* idx, const svs_query_batch* b, int mode) {
    return score_batch_impl(idx, b, mode, true);
}

extern "C" int svs_ctx_set_planner(svs_ctx* ctx, int on_device) {
    if (!ctx) return fail(SVS_ERR_ARG, "null context");
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->planner = on_device ? 1 : 0;
    return SVS_OK;
}

extern "C" int svs_ctx_set_product_split(svs_ctx* ctx, int on) {
    if (!ctx) return fail(SVS_ERR_ARG, "null context");
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->prod_split = on ? 1 : 0;
    return SVS_OK;
}

// ======================================================= routing on device
extern "C" int svs_index_set_centroids(svs_index* idx, const float* centroids, uint32_t d) {
    if (!idx || !centroids) return fail(SVS_ERR_ARG, "null argument");
    if (d == 0 || d > (uint32_t)N) return fail(SVS_ERR_ARG, "d = %u must be in 1..%d", d, N);
    svs_ctx* ctx = idx->ctx;
    const size_t nl = idx->lists.size();
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (idx->cent_dev) dev_free(ctx, idx->cent_dev, sizeof(float) * nl * idx->cent_d);
    idx->cent_dev = nullptr;
    idx->cent_d = 0;
    float* p = nullptr;
    SVS_TRY(dev_alloc(ctx, (void**)&p, sizeof(float) * nl * d));
    HIP_TRY(hipMemcpy(p, centroids, sizeof(float) * nl * d, hipMemcpyHostToDevice));
    idx->cent_dev = p;
    idx->cent_d = d;
    // graphs captured for this index hold the old centroid buffer: a new
    // generation keys fresh captures
    idx->gen = next_index_gen();
    return SVS_OK;
}

static int route_ok(const svs_index* idx) {
    if (!idx->cent_dev) return fail(SVS_ERR_STATE, "the index has no centroids (svs_index_set_centroids)");
    if (idx->lists.size() > (size_t)RT_SELECT_MAX)
        return fail(SVS_ERR_ARG, "device routing takes up to %d lists", RT_SELECT_MAX);
    return SVS_OK;
}

// Stages nq raw queries and a route-only probe layout in the slot:
// q_raw <- queries, q_probes = [probes][probe_ptr][verify], q_seeds.
static int stage_raw(svs_index* idx, svs_slot& sl, uint32_t nq, uint32_t d, const float* x,
                     const uint32_t* probe_ptr, const uint8_t* seeds) {
    svs_ctx* ctx = idx->ctx;
    const uint32_t n_items = probe_ptr[nq];
    hipStream_t st = sl.stream;
    SVS_TRY(route_setup(idx, sl, nq));
    HIP_TRY(hipMemcpyAsync(sl.q_raw.ptr, x, sizeof(float) * (size_t)nq * d, hipMemcpyHostToDevice,
                           st));
    SVS_TRY(sl.q_probes.ensure(ctx, sizeof(uint32_t) * ((size_t)n_items + nq + 1) + nq + 64));
    HIP_TRY(hipMemcpyAsync((uint32_t*)sl.q_probes.ptr + n_items, probe_ptr,
                           sizeof(uint32_t) * (nq + 1), hipMemcpyHostToDevice, st));
    if (seeds) {
        SVS_TRY(sl.q_seeds.ensure(ctx, (size_t)nq * 32));
        HIP_TRY(hipMemcpyAsync(sl.q_seeds.ptr, seeds, (size_t)nq * 32, hipMemcpyHostToDevice, st));
    }
    return SVS_OK;
}

extern "C" int svs_route_batch(svs_index* idx, const float* queries, uint32_t n_queries,
                               uint32_t nprobe, uint32_t* out) {
    if (!idx || (n_queries && (!queries || !out))) return fail(SVS_ERR_ARG, "null argument");
    SVS_TRY(route_ok(idx));
    if (n_queries == 0) return SVS_OK;
    svs_ctx* ctx = idx->ctx;
    const uint32_t np = std::min<uint32_t>(nprobe, (uint32_t)idx->lists.size());
    std::vector<uint32_t> pptr(n_queries + 1);
    for (uint32_t q = 0; q <= n_queries; ++q) pptr[q] = q * np;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    SVS_TRY(stage_raw(idx, *ctx, n_queries, idx->cent_d, queries, pptr.data(), nullptr));
    SVS_TRY(route_launch(idx, *ctx, n_queries, pptr[n_queries], (double)(1u << 20)));
    HIP_TRY(hipMemcpyAsync(out, ctx->q_probes.ptr, sizeof(uint32_t) * pptr[n_queries],
                           hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SVS_OK;
}

extern "C" int svs_score_raw_batch(svs_index* idx, const svs_raw_batch* b, int mode) {
    if (!idx || !b) return fail(SVS_ERR_ARG, "null argument");
    if (b->n_queries == 0) return SVS_OK;
    if (!b->queries || !b->q_seeds || !b->out_ids || !b->out_scores || !b->out_count ||
        !b->out_status)
        return fail(SVS_ERR_ARG, "missing batch array");
    if (b->k == 0 || b->k > (uint32_t)TOPK_LARGE_MAX)
        return fail(SVS_ERR_ARG, "k = %u must be in 1..%d", b->k, TOPK_LARGE_MAX);
    if (mode != SVS_MODE_PLAIN && mode != SVS_MODE_BATCHED && mode != SVS_MODE_BATCHED_VALU)
        return fail(SVS_ERR_ARG, "unknown mode %d", mode);
    SVS_TRY(route_ok(idx));
    svs_ctx* ctx = idx->ctx;
    const uint32_t nq_total = b->n_queries, nl = (uint32_t)idx->lists.size(), d = idx->cent_d;
    for (const ListHost& L : idx->lists)
        if (L.m && !L.d.mp_t32 && !L.d.mp_t64)
            return fail(SVS_ERR_STATE, "a list was uploaded without M_plain");
    std::vector<uint32_t> np(nq_total);
    for (uint32_t q = 0; q < nq_total; ++q)
        np[q] = std::min<uint32_t>(b->nprobe ? b->nprobe[q] : b->nprobe_all, nl);
    const double q_scale = (double)b->quantisation_q;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const uint32_t step = mode == SVS_MODE_PLAIN ? 1 : nq_total;
    uint32_t probe_base = 0;
    for (uint32_t q0 = 0; q0 < nq_total; q0 += step) {
        const uint32_t nq = std::min(step, nq_total - q0);
        std::vector<uint32_t> pptr(nq + 1, 0);
        uint64_t bound = 0;
        for (uint32_t q = 0; q < nq; ++q) {
            pptr[q + 1] = pptr[q] + np[q0 + q];
            bound += top_rows(idx, np[q0 + q]);
        }
        const uint32_t n_items = pptr[nq];
        const uint8_t* ver = b->verify ? b->verify + q0 : nullptr;
        const int kernel = mode == SVS_MODE_BATCHED ? KERNEL_MFMA : KERNEL_VALU;
        const int variant = ctx_variant_for(ctx, nq);
        const float* x = b->queries + (size_t)q0 * d;
        const uint8_t* seeds = b->q_seeds + (size_t)q0 * 32;
        uint32_t* o_ids = b->out_ids + (size_t)q0 * b->k;
        int64_t* o_scores = b->out_scores + (size_t)q0 * b->k;
        if (mode != SVS_MODE_PLAIN && ctx->planner && device_plan_ok(idx, n_items) &&
            bound < 0xFFFFFFFFull) {
            // route -> plan -> pipeline, one stream, no host round trip
            svs_slot& sl = *ctx;
            Staged h;
            SVS_TRY(stage_host(ctx, nq, pptr.data(), nullptr, ver, d, x, nullptr, seeds, h));
            PlanView v;
            SVS_TRY(plan_setup(idx, sl, nq, n_items, h.ver, bound, variant, v));
            SVS_TRY(route_setup(idx, sl, nq));
            SVS_TRY(stage_queries(ctx, sl, nq, nullptr, nullptr));
            SVS_TRY(view_reserve(ctx, sl, v, b->k));
            uint64_t qbits;
            std::memcpy(&qbits, &q_scale, 8);
            const std::vector<uint64_t> key{2, idx->gen, nq, n_items, b->k,
                                            (uint64_t)kernel, (uint64_t)variant, v.any_verify,
                                            v.planes_ok, bound, d, qbits,
                                            (uint64_t)(uintptr_t)h.pptr};
            SVS_TRY(run_graphed(ctx, sl, key, [&]() -> int {
                SVS_TRY(plan_stage(sl, v, h.pptr, nullptr, h.ver));
                HIP_TRY(hipMemcpyAsync(sl.q_raw.ptr, h.x, sizeof(float) * (size_t)nq * d,
                                       hipMemcpyHostToDevice, sl.stream));
                HIP_TRY(hipMemcpyAsync(sl.q_seeds.ptr, h.seeds, (size_t)nq * 32,
                                       hipMemcpyHostToDevice, sl.stream));
                SVS_TRY(route_launch(idx, sl, nq, n_items, q_scale));
                SVS_TRY(plan_launch(idx, sl, v, variant));
                return run_view(idx, sl, v, b->k, false, kernel, true, variant);
            }));
            if (b->out_probes && n_items)
                HIP_TRY(hipMemcpyAsync(b->out_probes + probe_base, v.probes,
                                       sizeof(uint32_t) * n_items, hipMemcpyDeviceToHost,
                                       sl.stream));
            SVS_TRY(collect(ctx, sl, nq, b->k, n_items, o_ids, o_scores, b->out_count + q0,
                            b->out_status + q0, v.hdr));
        } else {
            // route on the device, then plan on the host
            SVS_TRY(stage_raw(idx, *ctx, nq, d, x, pptr.data(), seeds));
            SVS_TRY(stage_queries(ctx, *ctx, nq, nullptr, nullptr));
            SVS_TRY(route_launch(idx, *ctx, nq, n_items, q_scale));
            std::vector<uint32_t> probes(std::max<uint32_t>(n_items, 1));
            HIP_TRY(hipMemcpyAsync(probes.data(), ctx->q_probes.ptr, sizeof(uint32_t) * n_items,
                                   hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            if (b->out_probes)
                std::copy(probes.begin(), probes.begin() + n_items, b->out_probes + probe_base);
            BatchPlan pl;
            int kern, var;
            SVS_TRY(plan_for_mode(idx, mode, nq, pptr.data(), probes.data(), ver, pl, kern, var));
            SVS_TRY(run_plan(idx, *ctx, pl, b->k, false, kern, true, var));
            SVS_TRY(collect(ctx, *ctx, nq, b->k, pl.n_items, o_ids, o_scores, b->out_count + q0,
                            b->out_status + q0));
        }
        probe_base += n_items;
    }
    return SVS_OK;
}

// Times every block-tile variant on a real batch (MODE_BATCHED, `reps` timed
// runs after one warm-up each), checks that all variants return identical
// results, and keeps the fastest for the batch's size class.
extern "C" int svs_autotune(svs_index* idx, const svs_query_batch* b, uint32_t reps, int* best,
                            double* ms_per_variant) {
    if (!idx || !b || reps == 0) return fail(SVS_ERR_ARG, "null argument or reps == 0");
    if (b->n_queries == 0) return fail(SVS_ERR_ARG, "empty batch");
    svs_ctx* ctx = idx->ctx;
    const size_t nk = (size_t)b->n_queries * b->k;
    std::vector<uint32_t> ids0(nk), cnt0(b->n_queries);
    std::vector<int64_t> sc0(nk);
    std::vector<int32_t> st0(b->n_queries);
    int prev;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        prev = ctx_variant_for(ctx, b->n_queries);
    }
    int win = -1;
    double win_ms = 0;
    for (int v = 0; v < SVS_TILE_VARIANTS; ++v) {
        SVS_TRY(svs_ctx_set_tile_variant(ctx, b->n_queries, v));
        double tot = 0;
        for (uint32_t r = 0; r <= reps; ++r) {
            const auto t0 = std::chrono::steady_clock::now();
            const int rc = svs_score_batch(idx, b, SVS_MODE_BATCHED);
            const auto t1 = std::chrono::steady_clock::now();
            if (rc != SVS_OK) {
                svs_ctx_set_tile_variant(ctx, b->n_queries, prev);
                return rc;
            }
            if (r) tot += std::chrono::duration<double, std::milli>(t1 - t0).count();
        }
        const double ms = tot / reps;
        if (ms_per_variant) ms_per_variant[v] = ms;
        if (v == 0) {
            std::copy(b->out_ids, b->out_ids + nk, ids0.begin());
            std::copy(b->out_scores, b->out_scores + nk, sc0.begin());
            std::copy(b->out_count, b->out_count + b->n_queries, cnt0.begin());
            std::copy(b->out_status, b->out_status + b->n_queries, st0.begin());
        } else {
            for (uint32_t q = 0; q < b->n_queries; ++q) {
                bool same = b->out_count[q] == cnt0[q] && b->out_status[q] == st0[q];
                for (uint32_t i = 0; same && i < b->out_count[q]; ++i)
                    same = b->out_ids[(size_t)q * b->k + i] == ids0[(size_t)q * b->k + i] &&
                           b->out_scores[(size_t)q * b->k + i] == sc0[(size_t)q * b->k + i];
                if (!same) {
                    svs_ctx_set_tile_variant(ctx, b->n_queries, prev);
                    return fail(SVS_ERR_STATE, "tile variant %d disagrees with variant 0 on query %u",
                                v, q);
                }
            }
        }
        if (win < 0 || ms < win_ms) win = v, win_ms = ms;
    }
    if (best) *best = win;
    return svs_ctx_set_tile_variant(ctx, b->n_queries, win);
}

// ============================================================ per-list drop-ins
static int one_item_plan(svs_index* idx, uint32_t list, BatchPlan& pl, bool verify) {
    uint32_t pptr[2] = {0, 1};
    uint8_t v = verify ? 1 : 0;
    return make_plan(idx, 1, pptr, &list, &v, pl, PB_ROWS, PB_QG);
}

static int list_ok(svs_index* idx, uint32_t list) {
    if (!idx) return fail(SVS_ERR_ARG, "null index");
    if (list >= idx->lists.size() || idx->lists[list].m == 0) return SVS_NONE;
    return SVS_OK;
}

extern "C" int svs_compute_products_dev(svs_index* idx, uint32_t list, const uint64_t* v_enc_dev,
                                        uint64_t* r_dev) {
    int rc = list_ok(idx, list);
    if (rc != SVS_OK) return rc;
    svs_ctx* ctx = idx->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    // Single-item plan, cached per list so repeated calls launch without
    // host-side planning or plan uploads.
    if (ctx->cached_plan_list != (int64_t)list || ctx->cached_plan_index != idx) {
        BatchPlan pl;
        SVS_TRY(one_item_plan(idx, list, pl, false));
        SVS_TRY(ctx->items1.ensure(ctx, sizeof(Item)));
        SVS_TRY(ctx->prefixes1.ensure(ctx, sizeof(uint32_t) * 4));
        uint32_t pre[4] = {pl.chunk_ptr[0], pl.chunk_ptr[1], pl.cand_ptr[0], pl.cand_ptr[1]};
        HIP_TRY(hipMemcpy(ctx->items1.ptr, pl.items.data(), sizeof(Item), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(ctx->prefixes1.ptr, pre, sizeof(pre), hipMemcpyHostToDevice));
        ctx->cached_plan_list = list;
        ctx->cached_plan_index = idx;
        ctx->cached_plan_chunks = pl.n_chunks;
    }
    const uint32_t* pre = (const uint32_t*)ctx->prefixes1.ptr;
    hipLaunchKernelGGL((k_gemv_chunks<N, 0>), dim3(row_grid(ctx, ctx->cached_plan_chunks)),
                       dim3(ROW_THREADS), 0, ctx->stream, (const ListDev*)idx->lists_dev,
                       (const Item*)ctx->items1.ptr, 1u, pre, pre + 2, v_enc_dev, r_dev);
    ctx->launches++;
    HIP_TRY(hipGetLastError());
    return SVS_OK;
}

extern "C" int svs_compute_products(svs_index* idx, uint32_t list, const uint64_t* v_enc,
                                    uint64_t* r) {
    int rc = list_ok(idx, list);
    if (rc != SVS_OK) return rc;
    if (!v_enc || !r) return fail(SVS_ERR_ARG, "null argument");
    SVS_TRY(check_canonical(v_enc, N, "v_enc"));
    svs_ctx* ctx = idx->ctx;
    const uint32_t m = idx->lists[list].m;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        HIP_TRY(hipSetDevice(ctx->device));
        SVS_TRY(ctx->io_a.ensure(ctx, sizeof(uint64_t) * N));
        SVS_TRY(ctx->io_b.ensure(ctx, sizeof(uint64_t) * m));
        HIP_TRY(hipMemcpyAsync(ctx->io_a.ptr, v_enc, sizeof(uint64_t) * N, hipMemcpyHostToDevice,
                               ctx->stream));
    }
    SVS_TRY(svs_compute_products_dev(idx, list, (const uint64_t*)ctx->io_a.ptr,
                                     (uint64_t*)ctx->io_b.ptr));
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipMemcpyAsync(r, ctx->io_b.ptr, sizeof(uint64_t) * m, hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SVS_OK;
}

// Stages a one-item plan with an explicit v_enc and claimed r on the device.
static int stage_single(svs_index* idx, uint32_t list, const uint64_t* v_enc, bool verify,
                        BatchPlan& pl) {
    svs_ctx* ctx = idx->ctx;
    SVS_TRY(one_item_plan(idx, list, pl, verify));
    SVS_TRY(ctx->items.ensure(ctx, sizeof(Item)));
    SVS_TRY(ctx->prefixes.ensure(ctx, sizeof(uint32_t) * 8));
    SVS_TRY(ctx->flags.ensure(ctx, 64));
    SVS_TRY(ctx->q_venc.ensure(ctx, sizeof(uint64_t) * N));
    uint32_t pre[6] = {pl.chunk_ptr[0], pl.chunk_ptr[1], pl.cand_ptr[0], pl.cand_ptr[1], 0, 1};
    HIP_TRY(hipMemcpyAsync(ctx->items.ptr, pl.items.data(), sizeof(Item), hipMemcpyHostToDevice,
                           ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->prefixes.ptr, pre, sizeof(pre), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->flags.ptr, pl.item_verify.data(), 1, hipMemcpyHostToDevice,
                           ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->q_venc.ptr, v_enc, sizeof(uint64_t) * N, hipMemcpyHostToDevice,
                           ctx->stream));
    return SVS_OK;
}

extern "C" int svs_verify(svs_index* idx, uint32_t list, const uint64_t* r,
                          const uint64_t* v_enc, const uint8_t rng_seed[32], uint32_t trials,
                          int32_t* failed_trial) {
    int rc = list_ok(idx, list);
    if (rc != SVS_OK) return rc;
    if (!r || !v_enc || !rng_seed || !failed_trial) return fail(SVS_ERR_ARG, "null argument");
    if (trials != (uint32_t)LAMBDA)
        return fail(SVS_ERR_ARG, "trials = %u; Sec128 uses lambda' = %d", trials, LAMBDA);
    SVS_TRY(check_canonical(v_enc, N, "v_enc"));
    svs_ctx* ctx = idx->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const uint32_t m = idx->lists[list].m;
    BatchPlan pl;
    SVS_TRY(stage_single(idx, list, v_enc, true, pl));
    SVS_TRY(ctx->r.ensure(ctx, sizeof(uint64_t) * m));
    SVS_TRY(ctx->q_vseed.ensure(ctx, 32));
    SVS_TRY(ctx->partials.ensure(ctx, sizeof(uint64_t) * 2 * LAMBDA * pl.n_chunks));
    SVS_TRY(ctx->verdict.ensure(ctx, 64));
    uint32_t key[8];
    seed_to_key(rng_seed, key);
    HIP_TRY(hipMemcpyAsync(ctx->q_vseed.ptr, key, 32, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->r.ptr, r, sizeof(uint64_t) * m, hipMemcpyHostToDevice, ctx->stream));
    int32_t* d_verdict = (int32_t*)ctx->verdict.ptr;
    uint32_t* d_flag = (uint32_t*)(d_verdict + 2);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, d_flag, 1u, 0u);
    const uint32_t* pre = (const uint32_t*)ctx->prefixes.ptr;
    hipLaunchKernelGGL(k_verify_chunks, dim3(row_grid(ctx, pl.n_chunks)), dim3(ROW_THREADS), 0,
                       ctx->stream, (const ListDev*)idx->lists_dev, (const Item*)ctx->items.ptr,
                       1u, pre, pre + 2, (const uint64_t*)ctx->q_venc.ptr,
                       (const uint64_t*)ctx->r.ptr, (const uint64_t*)nullptr,
                       (const uint32_t*)ctx->q_vseed.ptr, (const uint8_t*)ctx->flags.ptr,
                       (uint64_t*)ctx->partials.ptr, d_flag);
    hipLaunchKernelGGL(k_verify_finalize, dim3(1), dim3(VF_THREADS), 0, ctx->stream,
                       (const Item*)ctx->items.ptr, 1u, pre, (const uint8_t*)ctx->flags.ptr,
                       (const uint64_t*)ctx->partials.ptr, d_verdict);
    ctx->launches += 2;
    HIP_TRY(hipGetLastError());
    int32_t host[3];
    HIP_TRY(hipMemcpyAsync(host, d_verdict, sizeof(host), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (host[2]) return fail(SVS_ERR_RNG, "a Freivalds u-draw equalled p");
    *failed_trial = host[0];
    return SVS_OK;
}

// ---------------------------------------------- verification without allocation
// svs_verify_noalloc: the Protocol-2 check of svs_verify for callers that
// cannot allocate device memory. Its scratch (the one-item plan, the seed
// words, the per-chunk partial sums, the verdict and error flags) is this
// file-scope device array; each GPU holds its own copy at its own address,
// looked up once per device with hipGetSymbolAddress (noalloc_scratch_init,
// called by svs_ctx_create) and kept in g_scratch_ptr. Host code passes that
// pointer, never the array's name, to kernels.
constexpr uint32_t NOALLOC_MAX_ROWS = 1u << 20;
constexpr size_t NA_ITEM = 0;        // Item
constexpr size_t NA_PREFIX = 32;     // chunk_ptr[2], cand_ptr[2]
constexpr size_t NA_FLAG = 64;       // verify flag of the item
constexpr size_t NA_SEED = 96;       // 8 seed words
constexpr size_t NA_VERDICT = 128;   // verdict, unused, rng flag, non-canonical v_enc flag
constexpr size_t NA_PARTIALS = 256;  // 2 LAMBDA u64 per 16-row chunk
constexpr size_t NOALLOC_SCRATCH_BYTES =
    NA_PARTIALS + sizeof(uint64_t) * 2 * LAMBDA * (NOALLOC_MAX_ROWS / CHUNK_ROWS);
static __device__ __align__(256) uint64_t scratch[NOALLOC_SCRATCH_BYTES / sizeof(uint64_t)];

static std::mutex g_scratch_mu;
static std::vector<uint8_t*> g_scratch_ptr;  // per device; null until looked up

int svs::noalloc_scratch_init(int device) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize(device + 1, nullptr);
    if (g_scratch_ptr[device]) return SVS_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(SVS_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                    hipGetErrorString(e));
    g_scratch_ptr[device] = (uint8_t*)p;
    return SVS_OK;
}

static uint8_t* noalloc_scratch(int device) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    return (size_t)device < g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

struct SeedWords {
    uint32_t w[8];
};

// Writes the one-item plan, the seed words and cleared flags into the
// scratch from kernel arguments (no host-to-device copy).
__global__ void k_verify_noalloc_setup(uint8_t* __restrict__ base, uint32_t list, uint32_t m,
                                       uint32_t chunks, SeedWords seed) {
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        Item* it = (Item*)(base + NA_ITEM);
        it->q = 0;
        it->list = list;
        it->probe_pos = 0;
        it->m = m;
        uint32_t* pre = (uint32_t*)(base + NA_PREFIX);
        pre[0] = 0;
        pre[1] = chunks;
        pre[2] = 0;
        pre[3] = m;
        base[NA_FLAG] = 1;
    }
    if (t < 8) ((uint32_t*)(base + NA_SEED))[t] = seed.w[t];
    if (t < 4) ((int32_t*)(base + NA_VERDICT))[t] = 0;
}

extern "C" int svs_verify_noalloc(svs_index* idx, uint32_t list, const uint64_t* r_dev,
                                  const uint64_t* v_enc_dev, const uint8_t rng_seed[32],
                                  uint32_t trials, int32_t* failed_trial) {
    int rc = list_ok(idx, list);
    if (rc != SVS_OK) return rc;
    if (!r_dev || !v_enc_dev || !rng_seed || !failed_trial)
        return fail(SVS_ERR_ARG, "null argument");
    if (trials != (uint32_t)LAMBDA)
        return fail(SVS_ERR_ARG, "trials = %u; Sec128 uses lambda' = %d", trials, LAMBDA);
    const uint32_t m = idx->lists[list].m;
    if (m > NOALLOC_MAX_ROWS)
        return fail(SVS_ERR_ARG,
                    "list %u has %u rows; svs_verify_noalloc takes lists of up to %u rows "
                    "(its fixed scratch); use svs_verify",
                    list, m, NOALLOC_MAX_ROWS);
    svs_ctx* ctx = idx->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    uint8_t* base = noalloc_scratch(ctx->device);
    if (!base) return fail(SVS_ERR_STATE, "no scratch address for device %d", ctx->device);
    const uint32_t chunks = (m + CHUNK_ROWS - 1) / CHUNK_ROWS;
    SeedWords sw;
    seed_to_key(rng_seed, sw.w);
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(k_verify_noalloc_setup, dim3(1), dim3(64), 0, st, base, list, m, chunks,
                       sw);
    int32_t* d_verdict = (int32_t*)(base + NA_VERDICT);
    hipLaunchKernelGGL(k_check_canonical, dim3(4), dim3(256), 0, st, v_enc_dev, (uint64_t)N,
                       (uint32_t*)(d_verdict + 3));
    const Item* d_item = (const Item*)(base + NA_ITEM);
    const uint32_t* pre = (const uint32_t*)(base + NA_PREFIX);
    const uint8_t* d_flag = base + NA_FLAG;
    uint64_t* d_partials = (uint64_t*)(base + NA_PARTIALS);
    hipLaunchKernelGGL(k_verify_chunks, dim3(row_grid(ctx, chunks)), dim3(ROW_THREADS), 0, st,
                       (const ListDev*)idx->lists_dev, d_item, 1u, pre, pre + 2, v_enc_dev, r_dev,
                       (const uint64_t*)nullptr, (const uint32_t*)(base + NA_SEED), d_flag,
                       d_partials, (uint32_t*)(d_verdict + 2));
    hipLaunchKernelGGL(k_verify_finalize, dim3(1), dim3(VF_THREADS), 0, st, d_item, 1u, pre,
                       d_flag, (const uint64_t*)d_partials, d_verdict);
    ctx->launches += 4;
    HIP_TRY(hipGetLastError());
    int32_t host[4];
    HIP_TRY(hipMemcpyAsync(host, d_verdict, sizeof(host), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (host[3]) return fail(SVS_ERR_ARG, "v_enc holds a non-canonical value (>= p)");
    if (host[2]) return fail(SVS_ERR_RNG, "a Freivalds u-draw equalled p");
    *failed_trial = host[0];
    return SVS_OK;
}

extern "C" int svs_encode_query(svs_index* idx, uint32_t n_queries, const uint64_t* v_q,
                                const uint8_t* q_seeds, uint64_t* v_enc, uint64_t* g, int8_t* t) {
    if (!idx || (n_queries && (!v_q || !q_seeds || !v_enc || !g || !t)))
        return fail(SVS_ERR_ARG, "null argument");
    if (n_queries == 0) return SVS_OK;
    SVS_TRY(check_canonical(v_q, (size_t)n_queries * N, "v_q"));
    svs_ctx* ctx = idx->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t nq = n_queries;
    SVS_TRY(ctx->q_seeds.ensure(ctx, 32 * nq));
    SVS_TRY(ctx->q_vq.ensure(ctx, sizeof(uint64_t) * N * nq));
    SVS_TRY(ctx->q_gl.ensure(ctx, sizeof(uint64_t) * 2 * N1 * nq));
    SVS_TRY(ctx->q_tdense.ensure(ctx, (size_t)N * nq));
    SVS_TRY(ctx->q_tlist.ensure(ctx, sizeof(uint16_t) * N * nq));
    SVS_TRY(ctx->q_tcnt.ensure(ctx, sizeof(uint32_t) * nq));
    SVS_TRY(ctx->q_vseed.ensure(ctx, sizeof(uint32_t) * 8 * nq));
    SVS_TRY(ctx->q_venc.ensure(ctx, sizeof(uint64_t) * N * nq));
    SVS_TRY(ctx->tmp_c.ensure(ctx, 64));
    hipStream_t st = ctx->stream;
    uint32_t* d_flag = (uint32_t*)ctx->tmp_c.ptr;
    HIP_TRY(hipMemcpyAsync(ctx->q_seeds.ptr, q_seeds, 32 * nq, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(ctx->q_vq.ptr, v_q, sizeof(uint64_t) * N * nq, hipMemcpyHostToDevice,
                           st));
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, d_flag, 1u, 0u);
    hipLaunchKernelGGL(k_encode_keystream, dim3(n_queries), dim3(256), 0, st,
                       (const uint8_t*)ctx->q_seeds.ptr, (uint64_t*)ctx->q_gl.ptr,
                       (int8_t*)ctx->q_tdense.ptr, (uint16_t*)ctx->q_tlist.ptr,
                       (uint32_t*)ctx->q_tcnt.ptr, (uint32_t*)ctx->q_vseed.ptr, d_flag);
    hipLaunchKernelGGL(k_encode_venc, dim3(N / 256, n_queries), dim3(256), 0, st,
                       (const uint64_t*)idx->lt_dev, (const uint64_t*)ctx->q_gl.ptr,
                       (const int8_t*)ctx->q_tdense.ptr, (const uint64_t*)ctx->q_vq.ptr,
                       (uint64_t*)ctx->q_venc.ptr);
    ctx->launches += 3;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(v_enc, ctx->q_venc.ptr, sizeof(uint64_t) * N * nq,
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpy2DAsync(g, sizeof(uint64_t) * N1, ctx->q_gl.ptr, sizeof(uint64_t) * 2 * N1,
                             sizeof(uint64_t) * N1, nq, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(t, ctx->q_tdense.ptr, (size_t)N * nq, hipMemcpyDeviceToHost, st));
    uint32_t flag = 0;
    HIP_TRY(hipMemcpyAsync(&flag, d_flag, sizeof(flag), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return check_rng_flag(flag);
}

extern "C" int svs_decode(svs_index* idx, uint32_t list, const uint64_t* r, const uint64_t* v_enc,
                          const uint64_t* g, const uint32_t* t_idx, const int8_t* t_sign,
                          uint32_t t_nnz, uint64_t* mv) {
    int rc = list_ok(idx, list);
    if (rc != SVS_OK) return rc;
    if (!r || !v_enc || !g || !mv || (t_nnz && (!t_idx || !t_sign)))
        return fail(SVS_ERR_ARG, "null argument");
    if (t_nnz > (uint32_t)N) return fail(SVS_ERR_ARG, "t_nnz > n");
    SVS_TRY(check_canonical(v_enc, N, "v_enc"));
    SVS_TRY(check_canonical(g, N1, "g"));
    svs_ctx* ctx = idx->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const uint32_t m = idx->lists[list].m;
    if (!idx->lists[list].d.mp_t32 && !idx->lists[list].d.mp_t64)
        return fail(SVS_ERR_STATE, "list %u was uploaded without M_plain", list);
    BatchPlan pl;
    SVS_TRY(stage_single(idx, list, v_enc, false, pl));
    SVS_TRY(ctx->r.ensure(ctx, sizeof(uint64_t) * m));
    SVS_TRY(ctx->scores.ensure(ctx, sizeof(uint64_t) * m));
    SVS_TRY(ctx->q_gl.ensure(ctx, sizeof(uint64_t) * 2 * N1));
    SVS_TRY(ctx->q_tlist.ensure(ctx, sizeof(uint16_t) * N));
    SVS_TRY(ctx->q_tcnt.ensure(ctx, sizeof(uint32_t)));
    std::vector<uint16_t> tl(t_nnz);
    for (uint32_t e = 0; e < t_nnz; ++e) {
        if (t_idx[e] >= (uint32_t)N || (t_sign[e] != 1 && t_sign[e] != -1))
            return fail(SVS_ERR_ARG, "T entry %u invalid", e);
        tl[e] = (uint16_t)(t_idx[e] | (t_sign[e] < 0 ? 0x8000 : 0));
    }
    HIP_TRY(hipMemcpyAsync(ctx->r.ptr, r, sizeof(uint64_t) * m, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->q_gl.ptr, g, sizeof(uint64_t) * N1, hipMemcpyHostToDevice,
                           ctx->stream));
    if (t_nnz)
        HIP_TRY(hipMemcpyAsync(ctx->q_tlist.ptr, tl.data(), sizeof(uint16_t) * t_nnz,
                               hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->q_tcnt.ptr, &t_nnz, sizeof(uint32_t), hipMemcpyHostToDevice,
                           ctx->stream));
    hipLaunchKernelGGL(k_lt, dim3(1), dim3(N1 * LT_PARTS), 0, ctx->stream, (const uint64_t*)idx->l_dev,
                       (const uint64_t*)ctx->q_venc.ptr, (uint64_t*)ctx->q_gl.ptr);
    SVS_TRY(ctx->dense.ensure(ctx, sizeof(uint64_t) * m));
    hipLaunchKernelGGL(k_dense_rows, dim3(std::max<uint32_t>(1, std::min<uint32_t>((m + 3) / 4,
                                                                        ctx->n_cu * 16))),
                       dim3(256), 0, ctx->stream, idx->lists[list].d.alh, m,
                       (const uint64_t*)ctx->q_gl.ptr, (uint64_t*)ctx->dense.ptr);
    const uint32_t* pre = (const uint32_t*)ctx->prefixes.ptr;
    hipLaunchKernelGGL(k_decode_cands, dim3((m + DC_ROWS - 1) / DC_ROWS), dim3(DC_THREADS), 0,
                       ctx->stream,
                       (const ListDev*)idx->lists_dev, (const Item*)ctx->items.ptr, 1u, pre + 2,
                       (const uint64_t*)ctx->r.ptr, (const uint64_t*)ctx->q_venc.ptr,
                       (const uint16_t*)ctx->q_tlist.ptr, (const uint32_t*)ctx->q_tcnt.ptr,
                       (const uint64_t*)ctx->dense.ptr, (int64_t*)nullptr,
                       (uint64_t*)ctx->scores.ptr);
    ctx->launches += 3;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(mv, ctx->scores.ptr, sizeof(uint64_t) * m, hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SVS_OK;
}

// decode_dense_terms: AL g + H lt for one list (the reference's GPU decode
// split, gpu/mod.rs:390-408). Runs the decode kernel with T = 0, S = 0
// contributions removed by subtraction: out = r - decode(r) with r = 0.
extern "C" int svs_decode_dense_terms(svs_index* idx, uint32_t list, const uint64_t* g,
                                      const uint64_t* lt_venc, uint64_t* out) {
    int rc = list_ok(idx, list);
    if (rc != SVS_OK) return rc;
    if (!g || !lt_venc || !out) return fail(SVS_ERR_ARG, "null argument");
    SVS_TRY(check_canonical(g, N1, "g"));
    SVS_TRY(check_canonical(lt_venc, N1, "lt_venc"));
    svs_ctx* ctx = idx->ctx;
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const uint32_t m = idx->lists[list].m;
    SVS_TRY(ctx->q_gl.ensure(ctx, sizeof(uint64_t) * 2 * N1));
    SVS_TRY(ctx->r.ensure(ctx, sizeof(uint64_t) * m));
    HIP_TRY(hipMemcpyAsync(ctx->q_gl.ptr, g, sizeof(uint64_t) * N1, hipMemcpyHostToDevice,
                           ctx->stream));
    HIP_TRY(hipMemcpyAsync((uint64_t*)ctx->q_gl.ptr + N1, lt_venc, sizeof(uint64_t) * N1,
                           hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_dense_rows, dim3(std::max<uint32_t>(1, std::min<uint32_t>((m + 3) / 4,
                                                                        ctx->n_cu * 16))),
                       dim3(256), 0, ctx->stream, idx->lists[list].d.alh, m,
                       (const uint64_t*)ctx->q_gl.ptr, (uint64_t*)ctx->r.ptr);
    ctx->launches++;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, ctx->r.ptr, sizeof(uint64_t) * m, hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SVS_OK;
}

// Standalone top-k (fp_to_signed + sort by Reverse(score) + truncate,
// lib.rs:452-469; IVF concatenates the lists' candidates in probe order,
// ivf.rs:806-820) over caller segments, host buffers.
extern "C" int svs_topk(svs_ctx* ctx, uint32_t n_queries, const uint32_t* seg_ptr,
                        const uint64_t* mv, const uint32_t* ids, uint32_t k, uint32_t* out_ids,
                        int64_t* out_scores, uint32_t* out_count) {
    if (!ctx || (n_queries && (!seg_ptr || !out_ids || !out_scores || !out_count)))
        return fail(SVS_ERR_ARG, "null argument");
    if (k == 0 || k > 128) return fail(SVS_ERR_ARG, "k = %u must be in 1..128", k);
    if (n_queries == 0) return SVS_OK;
    if (seg_ptr[0] != 0) return fail(SVS_ERR_ARG, "seg_ptr[0] must be 0");
    for (uint32_t q = 0; q < n_queries; ++q)
        if (seg_ptr[q + 1] < seg_ptr[q]) return fail(SVS_ERR_ARG, "seg_ptr not monotone");
    const uint64_t nc = seg_ptr[n_queries];
    if (nc && (!mv || !ids)) return fail(SVS_ERR_ARG, "null argument");
    SVS_TRY(check_canonical(mv, nc, "mv"));
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t out_b = (sizeof(uint32_t) + sizeof(int64_t)) * (size_t)n_queries * k +
                         sizeof(uint32_t) * n_queries + 64;
    SVS_TRY(ctx->tmp_a.ensure(ctx, sizeof(uint64_t) * std::max<uint64_t>(nc, 1)));
    SVS_TRY(ctx->tmp_b.ensure(ctx, sizeof(uint32_t) * (std::max<uint64_t>(nc, 1) + n_queries + 1)));
    SVS_TRY(ctx->tmp_c.ensure(ctx, out_b));
    hipStream_t st = ctx->stream;
    uint64_t* d_mv = (uint64_t*)ctx->tmp_a.ptr;
    uint32_t* d_ids = (uint32_t*)ctx->tmp_b.ptr;
    uint32_t* d_seg = d_ids + std::max<uint64_t>(nc, 1);
    uint32_t* d_oids = (uint32_t*)ctx->tmp_c.ptr;
    int64_t* d_osc = (int64_t*)(d_oids + (size_t)n_queries * k + ((n_queries * k) & 1));
    uint32_t* d_ocnt = (uint32_t*)(d_osc + (size_t)n_queries * k);
    if (nc) {
        HIP_TRY(hipMemcpyAsync(d_mv, mv, sizeof(uint64_t) * nc, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(d_ids, ids, sizeof(uint32_t) * nc, hipMemcpyHostToDevice, st));
    }
    HIP_TRY(hipMemcpyAsync(d_seg, seg_ptr, sizeof(uint32_t) * (n_queries + 1),
                           hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL((k_topk_seg<128, 32>), dim3(n_queries), dim3(32), 0, st,
                       (const uint32_t*)d_seg, (const uint64_t*)d_mv, (const uint32_t*)d_ids, k,
                       d_oids, d_osc, d_ocnt);
    ctx->launches++;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out_ids, d_oids, sizeof(uint32_t) * (size_t)n_queries * k,
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(out_scores, d_osc, sizeof(int64_t) * (size_t)n_queries * k,
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(out_count, d_ocnt, sizeof(uint32_t) * n_queries,
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return SVS_OK;
}

extern "C" int svs_debug_inject_fault(svs_ctx* ctx, int32_t list, uint32_t row, uint64_t delta) {
    if (!ctx) return fail(SVS_ERR_ARG, "null context");
    if (delta >= P) return fail(SVS_ERR_ARG, "delta must be canonical");
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->fault_list = list;
    ctx->fault_row = row;
    ctx->fault_delta = delta;
    return SVS_OK;
}

extern "C" int svs_compute_products_batch(svs_index* idx, uint32_t n_queries,
                                          const uint64_t* v_enc, const uint32_t* probe_ptr,
                                          const uint32_t* probe_lists, int kernel, uint64_t* r) {
    if (!idx || (n_queries && (!v_enc || !probe_ptr || !probe_lists || !r)))
        return fail(SVS_ERR_ARG, "null argument");
    if (kernel != SVS_KERNEL_VALU && kernel != SVS_KERNEL_MFMA && kernel != SVS_KERNEL_MFMA_WAVE)
        return fail(SVS_ERR_ARG, "unknown kernel %d", kernel);
    const int kern = kernel == SVS_KERNEL_MFMA ? KERNEL_MFMA
                     : kernel == SVS_KERNEL_MFMA_WAVE ? KERNEL_MFMA_WAVE : KERNEL_VALU;
    uint32_t trows, tqg;
    const int variant = ctx_variant_for(idx->ctx, n_queries);
    tile_shape(kern, variant, trows, tqg);
    if (n_queries == 0) return SVS_OK;
    SVS_TRY(check_canonical(v_enc, (size_t)n_queries * N, "v_enc"));
    svs_ctx* ctx = idx->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    BatchPlan pl;
    SVS_TRY(make_plan(idx, n_queries, probe_ptr, probe_lists, nullptr, pl, trows, tqg));
    SVS_TRY(ctx->q_venc.ensure(ctx, sizeof(uint64_t) * (size_t)n_queries * N));
    HIP_TRY(hipMemcpyAsync(ctx->q_venc.ptr, v_enc, sizeof(uint64_t) * (size_t)n_queries * N,
                           hipMemcpyHostToDevice, ctx->stream));
    SVS_TRY(run_plan(idx, *ctx, pl, 1, true, kern, false, variant));
    if (pl.n_cands)
        HIP_TRY(hipMemcpyAsync(r, ctx->r.ptr, sizeof(uint64_t) * pl.n_cands,
                               hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    SVS_TRY(prof_collect(*ctx));
    return SVS_OK;
}

extern "C" int svs_ctx_read_probe(svs_ctx* ctx, uint64_t bytes, double* ms) {
    if (!ctx || bytes == 0 || bytes % 8192)
        return fail(SVS_ERR_ARG, "bytes must be a multiple of 8192");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    SVS_TRY(ctx->tmp_a.ensure(ctx, bytes));
    const uint32_t blocks = ctx->n_cu * 8;
    SVS_TRY(ctx->tmp_b.ensure(ctx, sizeof(uint32_t) * blocks * 256));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, ctx->stream));
    hipLaunchKernelGGL(k_read_probe, dim3(blocks), dim3(256), 0, ctx->stream,
                       (const uint64_t*)ctx->tmp_a.ptr, bytes / 8192, (uint32_t*)ctx->tmp_b.ptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(e1, ctx->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return SVS_OK;
}

extern "C" int svs_ctx_copy_probe(svs_ctx* ctx, uint64_t bytes, double* ms) {
    if (!ctx || bytes == 0 || bytes % 16) return fail(SVS_ERR_ARG, "bytes must be a multiple of 16");
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    SVS_TRY(ctx->tmp_a.ensure(ctx, bytes));
    SVS_TRY(ctx->tmp_b.ensure(ctx, bytes));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, ctx->stream));
    hipLaunchKernelGGL(k_copy_probe, dim3(ctx->n_cu * 16), dim3(256), 0, ctx->stream,
                       (const uint4*)ctx->tmp_a.ptr, (uint4*)ctx->tmp_b.ptr, bytes / 16);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(e1, ctx->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return SVS_OK;
}
