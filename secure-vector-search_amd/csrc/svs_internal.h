// svs_internal.h — host-side state behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <mutex>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/svs_gpu.h"
#include "svs_types.h"

struct svs_ctx;

namespace svs {

int fail(int code, const char* fmt, ...);
// A process-unique index generation (see svs_index::gen).
uint64_t next_index_gen();
// Looks up (once per device, with that device current) the device address
// of svs_verify_noalloc's fixed scratch array.
int noalloc_scratch_init(int device);

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ::svs::fail(SVS_ERR_HIP, "%s failed: %s (%s:%d)", #expr,            \
                               hipGetErrorString(e_), __FILE__, __LINE__);             \
    } while (0)

#define SVS_TRY(expr)              \
    do {                           \
        int rc_ = (expr);          \
        if (rc_ != SVS_OK) return rc_; \
    } while (0)

// Growable device scratch buffer owned by a context.
struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    int ensure(svs_ctx* ctx, size_t bytes);
    void release(svs_ctx* ctx);
};

// Pinned host staging that grows (hipHostMalloc): the host side of copies
// that a captured graph replays.
struct HostBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes);
    void release();
};

// Objects that live on a context (indexes, EMVP clusters, Tiptoe databases)
// retain it; svs_ctx_destroy is deferred until the last one is released.
void ctx_retain(svs_ctx* ctx);
void ctx_release(svs_ctx* ctx);

// Resident allocation against the context budget.
int dev_alloc(svs_ctx* ctx, void** p, size_t bytes);
void dev_free(svs_ctx* ctx, void* p, size_t bytes);

// regenerate_h_s (crypto.rs:258-270) on the host: H (m x n1 row-major) then
// S as CSR, from ChaCha20Rng(h_seed). Keystream blocks are generated in
// parallel; consumption is sequential and exact (rejections included).
void regenerate_h_s_host(const uint8_t h_seed[32], uint32_t m, std::vector<uint64_t>& h,
                         std::vector<uint32_t>& rowptr, std::vector<uint32_t>& sidx,
                         std::vector<int8_t>& ssign);

struct ListHost {
    uint32_t m = 0;
    ListDev d{};
    std::vector<std::pair<void*, size_t>> allocs;
};

}  // namespace svs

constexpr int SVS_PROF_STAGES = 5;

struct svs_prof {
    bool on = false, created = false;
    hipEvent_t ev[2 * SVS_PROF_STAGES] = {};
    double ms[SVS_PROF_STAGES] = {};
    uint64_t calls[SVS_PROF_STAGES] = {};
    uint64_t gemv_bytes = 0, gemv_rows = 0;
    uint32_t pending = 0;
};

// Per-batch device scratch and the stream a batch runs on. The context is
// itself the default slot (used by every synchronous entry point); a server
// owns further slots so that batches overlap on the device.
// One tenant's run of queries in a cross-tenant batch (queries q0 ..
// q0 + nq - 1 of the batch, sorted by tenant): encoding uses the tenant's
// keys, routing its centroids, and its lists sit at list_base.. in the
// batch's combined list table.
struct TenantSeg {
    uint32_t q0, nq, list_base;
    const svs_index* idx;
};

struct svs_slot {
    hipStream_t stream = nullptr;
    // set for a cross-tenant batch (a server over several tenants); empty for
    // a batch of one index
    std::vector<TenantSeg> segs;
    svs::DevBuf items, prefixes, flags;
    svs::DevBuf q_vq, q_seeds, q_venc, q_gl, q_tdense, q_tlist, q_tcnt, q_vseed, q_vplanes,
        q_pplanes, split;
    svs::DevBuf r, scores, partials, verdict, out, mtiles, q_corr, dense, topk_part, rv;
    // inputs of device-planned batches: probes + per-query flags, raw f32
    // queries, routing distances
    svs::DevBuf q_probes, q_raw, route_dist;
    svs_prof prof;
    // captured batch graphs (LRU) keyed by the batch's shape and the slot's
    // buffers, and the keys seen once (a key is captured on its second use)
    struct Graph {
        std::vector<uint64_t> key;
        hipGraphExec_t exec = nullptr;
        uint64_t last_use = 0;
    };
    std::vector<Graph> graphs;
    std::vector<std::vector<uint64_t>> graph_seen;
    uint64_t graph_clock = 0, graph_launches = 0, graph_captures = 0;
    std::vector<svs::DevBuf*> slot_scratch() {
        return {&items, &prefixes, &flags, &q_vq, &q_seeds, &q_venc, &q_gl, &q_tdense,
                &q_tlist, &q_tcnt, &q_vseed, &q_vplanes, &q_pplanes, &split, &r,
                &scores, &partials, &verdict,
                &out, &mtiles, &q_corr, &dense, &topk_part, &rv, &q_probes, &q_raw,
                &route_dist};
    }
};

struct svs_ctx : svs_slot {
    int device = 0;
    int n_cu = 256;
    uint64_t budget_bytes = 0, hbm_total = 0;
    uint64_t resident_bytes = 0, scratch_bytes = 0, peak_bytes = 0;
    uint64_t uploads = 0, launches = 0;
    std::mutex mu;
    // scratch of the per-list entry points and index builds
    svs::DevBuf tmp_a, tmp_b, tmp_c;
    svs::DevBuf items1, prefixes1, io_a, io_b;
    // cached single-item plan of svs_compute_products_dev
    int64_t cached_plan_list = -1;
    const void* cached_plan_index = nullptr;
    uint32_t cached_plan_chunks = 0;
    // test hook: server-side fault injected after the product (list < 0: off)
    int32_t fault_list = -1;
    // live indexes on this context; destroy is deferred while any remain
    int index_refs = 0;
    bool destroy_pending = false;
    uint32_t fault_row = 0;
    uint64_t fault_delta = 0;
    // chosen block-tile variant of the MFMA path per batch-size class
    // (n <= 16, <= 64, <= 256, larger)
    // tile variant per batch-size class (<= 16, <= 64, <= 256, more
    // queries) until svs_autotune or svs_ctx_set_tile_variant sets one:
    // 4-query tiles for the smallest batches, 16-query tiles (the plane
    // product with the query planes in LDS) above, the fastest measured
    // on the 1M-row k-means workload
    int tile_variant[4] = {3, 4, 4, 4};
    // batched runs planned on the device (1) or on the host (0)
    int planner = 1;
    // persistent tile grids: resident blocks per CU times this
    int tile_grid = 1;
    // the batched product on 16-query tile variants: the staged kernel over
    // work units that share each staged row and query group (1), or every
    // tile on k_prod_planes (0, the default: measured faster so far)
    int prod_split = 0;
    int prod_split_pf = 2;  // k_prod_staged's ring depth in steps: 2 or 4
    int prod_ldsb = 1;      // 16-query plane tiles: query planes staged in LDS per tile
    // the batched dense decode terms on the plane kernel (1, the default:
    // 0.17 ms faster per 256-query batch) or the window tiles (0)
    int dense_planes = 1;
    // device-planned batches replay captured HIP graphs (1) or launch (0)
    int graphs = 1;
    // pinned staging of caller inputs for graphed batches
    svs::HostBuf h_stage;
    // svs_dev_alloc buffers and their sizes
    std::unordered_map<void*, size_t> user_allocs;
    void note_peak();
    std::vector<svs::DevBuf*> all_scratch() {
        std::vector<svs::DevBuf*> v = slot_scratch();
        for (svs::DevBuf* b : {&tmp_a, &tmp_b, &tmp_c, &items1, &prefixes1, &io_a, &io_b})
            v.push_back(b);
        return v;
    }
};

struct svs_index {
    svs_ctx* ctx = nullptr;
    uint64_t* l_dev = nullptr;   // L, n x n1 row-major
    uint64_t* lt_dev = nullptr;  // L^T, n1 x n
    svs::ListDev* lists_dev = nullptr;
    std::vector<svs::ListHost> lists;
    // IVF centroids for device routing (n_lists x cent_d f32), or null
    float* cent_dev = nullptr;
    uint32_t cent_d = 0;
    // prefix sums of the list sizes sorted descending (rebuilt when a list
    // changes; top_rows)
    std::vector<uint64_t> top_rows_prefix;
    bool top_rows_dirty = true;
    // generation: unique per index and per centroid set (graph cache key)
    uint64_t gen = 0;
    // a combined list table over other indexes' lists (a cross-tenant
    // server's): its lists are not owned, it has no keys of its own
    bool view = false;
};

// ---------------------------------------------------------------- batches
namespace svs {

struct BatchPlan {
    uint32_t n_queries = 0, n_items = 0, n_chunks = 0, n_cands = 0;
    uint64_t distinct_bytes = 0;  // sum of m_c * n * 8 over distinct lists
    bool any_verify = false;
    bool planes_ok = true;  // every probed list has the MFMA decode planes
    std::vector<Item> items;
    std::vector<uint32_t> item_ptr, chunk_ptr, cand_ptr;
    std::vector<uint32_t> qchunk_ptr;  // per query: its top-k wave chunks
    uint32_t n_tchunks = 0;
    std::vector<uint8_t> item_verify;
    // MFMA product: items grouped by list (probe order kept within a list)
    // and work tiles of MT_ROWS rows x up to MT_QG of those items.
    std::vector<uint32_t> gitems;
    std::vector<MTile> tiles;
    // the same over the items to verify (the Freivalds recompute)
    std::vector<uint32_t> vgitems;
    std::vector<MTile> vtiles;
    // device header words: tile and verify-tile counts, overflow, rows
    uint32_t hdr[PH_WORDS] = {};
};

// What the pipeline kernels read about a batch: device arrays in the slot's
// buffers, plus host-side bounds that size grids and scratch. A plan built
// on the device knows its row and tile counts only there (hdr words and the
// prefix arrays' last entries); the caps bound them. For an uploaded host
// plan the caps are the exact counts.
struct PlanView {
    uint32_t nq = 0, n_items = 0;
    uint64_t cap_cands = 0, cap_chunks = 0, cap_tiles = 0, cap_vtiles = 0, cap_tchunks = 0;
    bool any_verify = false;  // some item may need verification
    const uint32_t* probes = nullptr;  // device-planned: the probe lists
    bool planes_ok = true;    // every list the batch can touch has the decode planes
    uint64_t distinct_bytes = 0, n_cands = 0;  // profile counters (host plans only)
    const Item* items = nullptr;
    const uint32_t *chunk_ptr = nullptr, *cand_ptr = nullptr, *item_ptr = nullptr,
                   *qchunk_ptr = nullptr, *hdr = nullptr;
    const uint8_t* flags = nullptr;
    const MTile *tiles = nullptr, *vtiles = nullptr;
    const uint32_t *gitems = nullptr, *vgitems = nullptr;
};

int make_plan(const svs_index* idx, uint32_t nq, const uint32_t* probe_ptr,
              const uint32_t* probe_lists, const uint8_t* verify, BatchPlan& pl,
              uint32_t tile_rows, uint32_t tile_qg);
int plan_for_mode(svs_index* idx, int mode, uint32_t nq, const uint32_t* probe_ptr,
                  const uint32_t* probe_lists, const uint8_t* verify, BatchPlan& pl,
                  int& kernel, int& variant);
// Encode + product + verify + decode + top-k of one planned batch whose
// query inputs are staged on the slot; asynchronous on the slot's stream.
int run_plan(svs_index* idx, svs_slot& sl, const BatchPlan& pl, uint32_t k, bool product_only,
             int kernel, bool encode, int variant);
// The same from a plan view (uploaded or built on the device).
int run_view(svs_index* idx, svs_slot& sl, const PlanView& v, uint32_t k, bool product_only,
             int kernel, bool encode, int variant);
// Device planning of a batched run. plan_setup sizes the slot for nq queries
// and n_items items (host work and allocation only): the plan layout, the
// caps (cands_bound bounds the candidate rows) and sl.q_probes = [probes:
// n_items u32][probe_ptr: nq + 1 u32][verify: nq bytes]. plan_stage copies
// probe_ptr, verify and (unless null: a routing kernel writes them) the
// probe lists from the host. plan_launch enqueues the planner; a plan beyond
// the bounds comes back flagged in hdr[PH_OVERFLOW] and empty.
int plan_setup(svs_index* idx, svs_slot& sl, uint32_t nq, uint32_t n_items, const uint8_t* verify,
               uint64_t cands_bound, int variant, PlanView& v);
int plan_stage(svs_slot& sl, const PlanView& v, const uint32_t* probe_ptr,
               const uint32_t* probe_lists, const uint8_t* verify);
int plan_launch(svs_index* idx, svs_slot& sl, const PlanView& v, int variant);
// Quantises and routes nq raw queries already in sl.q_raw (f32, d each)
// into sl.q_vq and the probe lists of sl.q_probes (layout above);
// route_setup sizes the buffers.
int route_setup(svs_index* idx, svs_slot& sl, uint32_t nq);
int route_launch(svs_index* idx, svs_slot& sl, uint32_t nq, uint32_t n_items, double q_scale);
// Sizes every buffer run_view uses for this view and k.
int view_reserve(svs_ctx* ctx, svs_slot& sl, const PlanView& v, uint32_t k);
// Enqueues a batch on the slot's stream by replaying the HIP graph captured
// for `key` (plus the slot's buffer addresses); a key is captured on its
// second use, and runs eagerly otherwise or when graphs are off, profiling
// is on or a fault is injected. enqueue() must only enqueue device work and
// copies from pinned host memory, and must not allocate.
int run_graphed(svs_ctx* ctx, svs_slot& sl, std::vector<uint64_t> key,
                const std::function<int()>& enqueue);
void release_graphs(svs_slot& sl);
// Candidate-row capacity bucket: rows rounded up to a power of two (at
// least 2^16), so batches of similar size share a capture.
uint64_t cap_bucket(uint64_t rows);
// Plan layout in the slot for nq queries, n_items items and cap_tiles tiles:
// ensures the buffers and sets v's array pointers.
int plan_layout(svs_ctx* ctx, svs_slot& sl, uint32_t nq, uint32_t n_items, uint64_t cap_tiles,
                PlanView& v);
// Whether device planning can take a batch of this index and size.
bool device_plan_ok(const svs_index* idx, uint32_t n_items);
// Sum of the n largest list sizes of the index, n = 0 .. n_lists.
uint64_t top_rows(svs_index* idx, uint32_t n);
int stage_queries(svs_ctx* ctx, svs_slot& sl, uint32_t nq, const uint64_t* v_q,
                  const uint8_t* seeds);
int collect_async(svs_slot& sl, uint32_t nq, uint32_t k, uint32_t n_items, uint32_t* ids,
                  int64_t* scores, uint32_t* count, int32_t* status, uint32_t* rng_flag);
int check_rng_flag(uint32_t rng_flag);
int reserve_slot(svs_ctx* ctx, svs_slot& sl, uint32_t nq, uint32_t n_items, uint64_t n_cands,
                 uint32_t k);

}  // namespace svs
