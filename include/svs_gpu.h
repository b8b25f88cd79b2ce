/*
 * svs_gpu.h — C ABI of the MI355X (gfx950) backend for the encrypted-score
 * matrix-vector path of sacs-epfl/secure-vector-search.
 *
 * This library replaces the CUDA path behind the reference's `Scorer` trait:
 * scorer-gpu-common (ClusterStore, budget, sampler) and the per-scheme GPU
 * states scorer-bntm/src/gpu/mod.rs (FlatGpuState, IvfGpuState) and
 * scorer-emvp/src/gpu/mod.rs, plus Tiptoe's server answer
 * (scorer-tiptoe/src/pir/simplepir.rs:191-204), which has no GPU path in the
 * reference. A Rust FFI crate binds these symbols as plain extern "C"
 * declarations; every call returns an int status (svs_last_error for text).
 *
 * Conventions
 *  - Every field element in and out is canonical: a u64 in [0, p), p = 2^61-1.
 *  - Ring dimension n = 1024, subspace n1 = 128, lambda' = 3 (Sec128,
 *    scorer-bntm/src/params.rs:102-119,198-205).
 *  - Seeds are explicit inputs (the reference draws q_seed / h_seed from OS
 *    entropy, lib.rs:882-885, ivf.rs:770); the caller injects them.
 *  - Pointers named `*_host` are host memory; `*_dev` are device memory on the
 *    context's GPU. No allocation crosses the ABI: outputs are caller-owned.
 *  - Return value: SVS_OK (0), a non-negative status, or a negative error;
 *    svs_last_error() gives the thread-local message of the last error.
 *  - Thread safety: calls on one context are serialised internally.
 */
#ifndef SVS_GPU_H
#define SVS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------
 * SVS_NONE mirrors `Ok(None)` for an empty / out-of-range list
 * (scorer-bntm/src/gpu/mod.rs:506-536). SVS_ERR_BUDGET mirrors
 * GpuError::BudgetTooSmall (gpu/mod.rs:97-144, store.rs:183-211). */
#define SVS_OK 0
#define SVS_NONE 1
#define SVS_ERR_ARG (-1)
#define SVS_ERR_HIP (-2)
#define SVS_ERR_BUDGET (-3)
#define SVS_ERR_STATE (-4)
#define SVS_ERR_RNG (-5) /* a device u-draw hit the rejected value p (prob 2^-61) */

typedef struct svs_ctx svs_ctx;
typedef struct svs_index svs_index;

typedef struct {
    uint64_t budget_bytes;        /* resolved HBM budget for resident state */
    uint64_t resident_bytes;      /* index bytes resident now */
    uint64_t peak_resident_bytes; /* max of resident + scratch since create */
    uint64_t hbm_total_bytes;     /* hipMemGetInfo total */
    uint64_t hbm_free_bytes;      /* hipMemGetInfo free, sampled at this call */
    uint64_t uploads;             /* lists uploaded */
    uint64_t launches;            /* kernels launched (a graph's once, at capture) */
    uint64_t graph_launches;      /* batches replayed from captured HIP graphs */
    uint64_t graph_captures;      /* HIP graphs captured */
} svs_stats;

/* Thread-local message of the last failing call; returns its length. */
int svs_last_error(char* buf, size_t len);

/* Create a context on `device`. budget_bytes = 0 resolves to 90 % of the
 * free HBM. Replaces the process-wide `kernel()` OnceLock and
 * `resolve_budget` (scorer-bntm/src/gpu/mod.rs:169-200,
 * scorer-gpu-common/src/budget.rs:132-150). */
int svs_ctx_create(int device, uint64_t budget_bytes, svs_ctx** out);
int svs_ctx_destroy(svs_ctx* ctx);
/* peak_vram_bytes / vram_budget_bytes / global_stats
 * (gpu/mod.rs:336-349, scorer-gpu-common/src/store.rs:74-81). */
int svs_ctx_stats(svs_ctx* ctx, svs_stats* out);

/* Stage timing of svs_score_batch: HIP events recorded on the context's own
 * stream around each stage (0 encode, 1 server product, 2 verification,
 * 3 decode, 4 top-k). Accumulated ms and call counts, plus the algorithmic
 * bytes of the server product (sum over each batch's distinct probed lists
 * of m_c * n * 8) and its candidate rows. The reference's per-substep
 * breakdown (BnTmIvfTiming, ivf.rs:1380-1677) is the host-clock analogue. */
typedef struct {
    double ms[8];
    uint64_t calls[8];
    uint64_t gemv_bytes;
    uint64_t gemv_rows;
} svs_profile;
int svs_ctx_profile(svs_ctx* ctx, int enable, int reset, svs_profile* out);

/* Copies `bytes` (a multiple of 16) between two device buffers with a plain
 * 16-byte-vector kernel (k_copy_probe) and returns its HIP-event time: the
 * reference dispatch for calibrating HBM byte counters (FETCH_SIZE /
 * WRITE_SIZE) and the measured copy bandwidth. */
int svs_ctx_copy_probe(svs_ctx* ctx, uint64_t bytes, double* ms);
/* Read-only probe of `bytes` (a multiple of 8192) in the batched product's
 * A-operand access pattern (64-byte segments of 16 rows of 8 KiB per load
 * instruction): the HBM-counter calibration for that kernel. */
int svs_ctx_read_probe(svs_ctx* ctx, uint64_t bytes, double* ms);

/* ---- resident index -----------------------------------------------------
 * An index holds n_lists encrypted lists (a flat scorer is one list) and the
 * shared subspace L (n x n1 row-major, generate_subspace crypto.rs:235-238).
 * Everything stays resident in HBM; no LRU streaming. Replaces
 * IvfGpuState::build / FlatGpuState::build (gpu/mod.rs:300-332,443-481) and
 * ClusterStore::get_or_upload (scorer-gpu-common/src/store.rs:327-380). */
int svs_index_create(svs_ctx* ctx, uint32_t n_lists, const uint64_t* l_host, svs_index** out);
int svs_index_destroy(svs_index* idx);

/* Upload one list's server state (M_enc) and client state (AL, H, M_plain,
 * S as CSR with sorted column indices, signs +-1) as produced by
 * encrypt_cluster (crypto.rs:405-478), plus the list's global vector ids
 * (index.clusters[ci][local].0, ivf.rs:815-817). m = 0 uploads an empty
 * list. Row-major host arrays; m_plain may be NULL when decode is unused. */
int svs_index_upload_list(svs_index* idx, uint32_t list, uint32_t m, const uint64_t* m_enc_host,
                          const uint64_t* a_l_host, const uint64_t* h_host,
                          const uint64_t* m_plain_host, const uint32_t* s_rowptr_host,
                          const uint32_t* s_idx_host, const int8_t* s_sign_host,
                          const uint32_t* ids_host);

/* Build one list on the device from its quantised plaintext rows: H and S
 * are regenerated from h_seed (regenerate_h_s, crypto.rs:258-270) on the
 * host, then M_enc = M + H L^T + S and AL = M L are computed on the GPU
 * (encrypt_cluster, crypto.rs:405-478). Bit-identical to uploading the
 * host-encrypted list. */
int svs_index_build_list(svs_index* idx, uint32_t list, uint32_t m, const uint64_t* m_plain_host,
                         const uint8_t h_seed[32], const uint32_t* ids_host);

/* Copy a list's M_enc, AL and H back to the host (row-major; NULL skips an
 * array): lets tests compare the device build with the host encryption and
 * lets a CPU baseline reuse the device-built state. */
int svs_index_download_list(svs_index* idx, uint32_t list, uint64_t* m_enc_host,
                            uint64_t* a_l_host, uint64_t* h_host);

/* Load a BN-IVF index from the reference's cache file
 * (.bntm-ivf-cache-<fingerprint>.bin: layout scorer-bntm/src/ivf.rs:1683-1902,
 * header scorer-core/src/cache.rs:19-73, IVF payload ivf-index/src/ivf.rs:
 * 68-131) straight into HBM: the file is memory-mapped and each list's
 * M_enc, AL, H and M_plain go from the mapping to the device; S is
 * regenerated from the list's h_seed and the regenerated H must equal the
 * stored one; the centroids are kept for device routing. A header that does
 * not match (magic, version, or the fingerprint when check_fingerprint) is
 * the reference's Ok(None): SVS_NONE and *out = NULL. Replaces
 * load_bntm_ivf_cache + IvfGpuState::build. */
int svs_index_load_ivf_cache(svs_ctx* ctx, const char* path, uint64_t fingerprint,
                             int check_fingerprint, svs_index** out, uint32_t* n_lists,
                             uint32_t* dim);
/* fingerprint64 (cache.rs:22-37): FNV-1a 64 over the length-prefixed parts. */
uint64_t svs_cache_fingerprint64(const uint8_t* const* parts, const uint64_t* lens,
                                 uint32_t n_parts);
/* Copies of an index's routing centroids (n_lists x d f32) and subspace L
 * (n x n1 u64) back to the host. */
int svs_index_get_centroids(svs_index* idx, float* out_host);
int svs_index_get_subspace(svs_index* idx, uint64_t* l_out_host);

/* Rows of a list (0 for empty/out-of-range). */
uint32_t svs_index_list_rows(const svs_index* idx, uint32_t list);

/* ---- plain per-list drop-ins (reference GPU state API) ------------------ */
/* r = M_enc . v_enc (mod p) for one list; mersenne_matvec semantics
 * (crypto.rs:606-626). Replaces IvfGpuState::compute_products(ci, v_enc)
 * (gpu/mod.rs:506-536) / FlatGpuState::compute_products (:362-379) and the
 * bntm_compute_products kernel (kernels.cu:100-178). Returns SVS_NONE for an
 * empty or out-of-range list. */
int svs_compute_products(svs_index* idx, uint32_t list, const uint64_t* v_enc_host,
                         uint64_t* r_host);

/* Same product with device pointers, asynchronous on the context stream:
 * the bench's resident-input path. */
int svs_compute_products_dev(svs_index* idx, uint32_t list, const uint64_t* v_enc_dev,
                             uint64_t* r_dev);

/* Gathered server products for many (query, list) pairs in one launch:
 * compute_products_batch (crypto.rs:782-828) generalised to uneven IVF probe
 * sets. Query q's v_enc (host, n u64) is multiplied with every list
 * probe_lists[probe_ptr[q] .. probe_ptr[q+1]); r receives, item after item
 * in that order, m_c values per item (nothing for empty lists). `kernel`
 * picks the VALU row-streaming kernel, the exact int8 MFMA limb tiles with
 * 256-row x 8-query block tiles (SVS_KERNEL_MFMA) or the same arithmetic on
 * 64-row x 4-query single-wave tiles (SVS_KERNEL_MFMA_WAVE); all are
 * bit-identical to mersenne_matvec. */
#define SVS_KERNEL_VALU 0
#define SVS_KERNEL_MFMA 1
#define SVS_KERNEL_MFMA_WAVE 2
int svs_compute_products_batch(svs_index* idx, uint32_t n_queries, const uint64_t* v_enc_host,
                               const uint32_t* probe_ptr_host, const uint32_t* probe_lists_host,
                               int kernel, uint64_t* r_host);

/* AL . g + H . lt_venc (mod p): decode_dense_terms (gpu/mod.rs:390-408,
 * 541-577) with the reference's stale N1 = 512 kernel replaced by n1 = 128. */
int svs_decode_dense_terms(svs_index* idx, uint32_t list, const uint64_t* g_host,
                           const uint64_t* lt_venc_host, uint64_t* out_host);

/* Protocol-2 iterated Freivalds check of a claimed r (verify_response,
 * crypto.rs:1013-1046). `rng_seed` seeds ChaCha20Rng for the u draws; the
 * scorers pass domain_separate(q_seed, "protocol2") (lib.rs:420-431).
 * *failed_trial = -1 on accept, else the first failing trial. */
int svs_verify(svs_index* idx, uint32_t list, const uint64_t* r_host, const uint64_t* v_enc_host,
               const uint8_t rng_seed[32], uint32_t trials, int32_t* failed_trial);

/* The same check with r and v_enc in device memory (r_dev: the list's m
 * rows; v_enc_dev: n canonical u64), calling no hipMalloc, hipFree or other
 * allocating routine: its scratch (plan, seed words, per-chunk partial sums,
 * verdict) is a fixed device array of the library, 3 MiB per GPU, held until
 * the process ends. No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation). Lists of up to 1,048,576 rows (larger:
 * SVS_ERR_ARG; use svs_verify). Calls on one GPU share the scratch: calls
 * through one context are serialised, but calls that overlap from different
 * contexts on the same GPU race. Non-canonical v_enc: SVS_ERR_ARG. */
int svs_verify_noalloc(svs_index* idx, uint32_t list, const uint64_t* r_dev,
                       const uint64_t* v_enc_dev, const uint8_t rng_seed[32], uint32_t trials,
                       int32_t* failed_trial);

/* Client query encoding on the device (encode_query, crypto.rs:508-553):
 * for each query, G = 128 sample_fp draws and T = sample_sparse_pm1 from
 * ChaCha20Rng(q_seed), v_enc = v_q + L G + T. v_q: n_queries x n canonical
 * (quantise_vector's output, zero-padded); q_seeds: n_queries x 32 bytes.
 * Outputs v_enc (n_queries x n), g (n_queries x n1) and the dense T
 * (n_queries x n, entries -1 / 0 / +1). The batch pipeline runs the same
 * kernels; this entry exposes them. */
int svs_encode_query(svs_index* idx, uint32_t n_queries, const uint64_t* v_q_host,
                     const uint8_t* q_seeds_host, uint64_t* v_enc_host, uint64_t* g_host,
                     int8_t* t_host);

/* Full client decode (decode_scores, crypto.rs:901-940): Mv from r, the
 * per-query mask (g, sparse T) and v_enc. */
int svs_decode(svs_index* idx, uint32_t list, const uint64_t* r_host, const uint64_t* v_enc_host,
               const uint64_t* g_host, const uint32_t* t_idx_host, const int8_t* t_sign_host,
               uint32_t t_nnz, uint64_t* mv_host);

/* Top-k of decoded scores (fp_to_signed, crypto.rs:128-135, then the
 * sort by Reverse(score) and truncate of lib.rs:452-469): query q's
 * candidates are mv[seg_ptr[q] .. seg_ptr[q+1]) (canonical F_p, e.g. from
 * svs_decode) with their global ids, concatenated in probe order as the IVF
 * scorer does (ivf.rs:806-820). Outputs up to k (id, signed score) per query
 * (n_queries x k), descending, ties by position; out_count[q] = min(k,
 * candidates). k in 1..128. */
int svs_topk(svs_ctx* ctx, uint32_t n_queries, const uint32_t* seg_ptr_host,
             const uint64_t* mv_host, const uint32_t* ids_host, uint32_t k,
             uint32_t* out_ids_host, int64_t* out_scores_host, uint32_t* out_count_host);

/* ---- whole-query pipeline (encode -> product -> verify -> decode -> top-k)
 * For each query q: v_q (quantised, n u64), q_seed (32 B), its probe lists
 * (probe_ptr[q]..probe_ptr[q+1] into probe_lists, in probe order), a verify
 * flag and k. Encoding uses q_seed exactly as encode_query
 * (crypto.rs:508-553); verification re-seeds ChaCha20Rng with
 * domain_separate(q_seed, "protocol2") per probed list (ivf.rs:790-804).
 * Outputs per query: up to k (id, signed score) hits, descending, ties in
 * probe-order position (the reference leaves ties unspecified, lib.rs:458);
 * status[q] = 0, or 1 + trial + 4 * list for the first probed list (in
 * probe order) whose verification failed (then no hits are written).
 *
 * mode SVS_MODE_PLAIN launches one gathered product per query (the
 * one-query-per-launch drop-in); SVS_MODE_BATCHED runs the whole batch
 * through one gathered launch per stage. Results are identical. */
#define SVS_MODE_PLAIN 0
#define SVS_MODE_BATCHED 1      /* server product on the int8 MFMA limb tiles */
#define SVS_MODE_BATCHED_VALU 2 /* batched, server product on the VALU kernel */

typedef struct {
    uint32_t n_queries;
    const uint64_t* v_q;          /* n_queries x 1024 */
    const uint8_t* q_seeds;       /* n_queries x 32 */
    const uint32_t* probe_ptr;    /* n_queries + 1 */
    const uint32_t* probe_lists;  /* probe_ptr[n_queries] list ids */
    const uint8_t* verify;        /* n_queries flags (NULL = none) */
    uint32_t k;                   /* hits per query (1..4096) */
    uint32_t* out_ids;            /* n_queries x k */
    int64_t* out_scores;          /* n_queries x k */
    uint32_t* out_count;          /* n_queries */
    int32_t* out_status;          /* n_queries */
} svs_query_batch;

int svs_score_batch(svs_index* idx, const svs_query_batch* batch_host, int mode);
/* Device buffers for callers that keep query inputs and outputs in HBM (no
 * reference counterpart; plumbing for svs_score_batch_dev). Counted against
 * the context budget; svs_memcpy copies in any direction (synchronous). */
int svs_dev_alloc(svs_ctx* ctx, uint64_t bytes, void** out);
int svs_dev_free(svs_ctx* ctx, void* p);
int svs_memcpy(svs_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* Same with v_q, q_seeds and the four output arrays in device memory
 * (probe_ptr, probe_lists and verify stay on the host: they plan the
 * batch). Non-canonical v_q is reported as SVS_ERR_ARG after the batch. */
int svs_score_batch_dev(svs_index* idx, const svs_query_batch* batch_dev, int mode);

/* Where batched runs are planned (items, prefix arrays, MFMA tiles): 1 = on
 * the device by one planning kernel (default), 0 = on the host and uploaded.
 * Both give the same plan; no reference counterpart (the reference plans
 * per list on the host, ivf.rs:1289-1296). */
int svs_ctx_set_planner(svs_ctx* ctx, int on_device);

/* Device-planned batches (svs_score_batch[_dev] and svs_score_raw_batch in
 * the batched modes, and every server batch) run as one HIP graph per batch
 * shape: the shape's second batch is captured (staging copies from pinned
 * memory, routing, planning, the whole pipeline) and later batches of that
 * shape replay it; 0 launches every kernel instead (default 1). Results are
 * identical either way. No reference counterpart. */
int svs_ctx_set_graphs(svs_ctx* ctx, int on);

/* How the batched server product runs on 16-query tile variants: 1 runs
 * the staged kernel, whose work units (16 rows x up to four query groups of
 * a 256-row block, 32 rows x two, or 64 rows x one) stage each step's rows
 * and query planes once in LDS for the block's four waves, so a row is
 * read from HBM once for up to 64 queries; 0 (default, measured faster)
 * runs every tile on the per-wave plane kernel. Results are identical
 * either way. No reference counterpart (the reference launches one GEMV
 * per probed list, gpu/mod.rs:584-614). */
int svs_ctx_set_product_split(svs_ctx* ctx, int on);

/* ---- client routing and quantisation on the device -------------------------
 * The reference client routes and quantises each query on the host before
 * encryption (BnTmIvfScorer::score steps 1-2, scorer-bntm/src/ivf.rs:
 * 762-766): probe_route (ivf-index/src/ivf.rs:150-164) picks the nprobe
 * centroids nearest by l2_f32_scalar (ivf-index/src/distance.rs:30-37), and
 * quantise_vector (crypto.rs:105-122) scales by Q. Here both run on the
 * device with the same f32 / f64 results: each distance is a left-to-right
 * f32 sum of rounded squared differences and a correctly rounded sqrt;
 * probes ascend by f32::total_cmp with ties in centroid order (the
 * reference's unstable sort leaves ties unspecified). */
/* The IVF centroids (IvfIndex::centroids): n_lists rows of d <= 1024 f32,
 * row-major. Replaces any earlier set. Device routing takes up to 8192. */
int svs_index_set_centroids(svs_index* idx, const float* centroids_host, uint32_t d);
/* probe_route for n_queries queries (n_queries x d f32, host): out
 * (n_queries x nprobe, host) receives each query's probe list. nprobe is
 * clamped to n_lists as in the reference. */
int svs_route_batch(svs_index* idx, const float* queries_host, uint32_t n_queries,
                    uint32_t nprobe, uint32_t* out_host);

/* A batch of raw queries: the whole BnTmIvfScorer::score (scorer-bntm/src/
 * ivf.rs:746-822) per query on the device: quantise, route, encode, product,
 * verify, decode, top-k. SVS_MODE_BATCHED runs the batch as one stream of
 * kernels with no host round trip (routing feeds the device planner);
 * SVS_MODE_PLAIN routes and runs one query at a time. Outputs as
 * svs_query_batch; out_probes (optional) receives the routed lists, query
 * after query (sum of the clamped nprobe words). */
typedef struct {
    uint32_t n_queries;
    const float* queries;      /* n_queries x d f32, d = the centroids' d */
    const uint32_t* nprobe;    /* per query (NULL: nprobe_all for every query) */
    uint32_t nprobe_all;
    uint64_t quantisation_q;   /* Q of quantise_vector (params.rs:19: 2^20) */
    const uint8_t* q_seeds;    /* n_queries x 32 */
    const uint8_t* verify;     /* n_queries flags (NULL = none) */
    uint32_t k;                /* 1..4096 */
    uint32_t* out_ids;         /* n_queries x k */
    int64_t* out_scores;       /* n_queries x k */
    uint32_t* out_count;       /* n_queries */
    int32_t* out_status;       /* n_queries */
    uint32_t* out_probes;      /* optional */
} svs_raw_batch;
int svs_score_raw_batch(svs_index* idx, const svs_raw_batch* batch_host, int mode);

/* ---- block-tile variants and the autotuner ---------------------------------
 * No reference counterpart (the reference has one CUDA kernel shape): the
 * MFMA batch path (SVS_MODE_BATCHED, SVS_KERNEL_MFMA) runs on 256-row block
 * tiles whose query group, prefetch depth and waves per SIMD come from a
 * small table of variants, chosen per batch-size class (n_queries <= 16,
 * <= 64, <= 256, larger). Every variant returns identical results. */
#define SVS_TILE_VARIANTS 10
/* n_queries == 0 sets every class. */
int svs_ctx_set_tile_variant(svs_ctx* ctx, uint32_t n_queries, int variant);
int svs_ctx_get_tile_variant(svs_ctx* ctx, uint32_t n_queries, int* variant,
                             uint32_t* query_group, uint32_t* ring_depth,
                             uint32_t* waves_per_simd);
/* Runs the batch once per variant (one warm-up plus `reps` timed runs,
 * SVS_MODE_BATCHED), fails with SVS_ERR_STATE if any variant's results
 * differ from variant 0's, and keeps the fastest for this batch size.
 * ms_per_variant (SVS_TILE_VARIANTS entries, may be NULL) receives the mean
 * wall time per run; the batch's outputs hold the last variant's results. */
int svs_autotune(svs_index* idx, const svs_query_batch* batch_host, uint32_t reps, int* best,
                 double* ms_per_variant);

/* ---- serving: continuous batching -----------------------------------------
 * No single reference counterpart: the reference scores one query per call
 * (BnTmIvfScorer::score, scorer-bntm/src/ivf.rs:746-822) and batches only a
 * chunk handed to it (eval.rs:2428-2433). A server keeps one or more
 * resident indexes (tenants: each its own keys and lists, all on one
 * context) and a queue of requests. A dispatcher thread packs the pending
 * requests (of every tenant, or of one tenant: cross_tenant) into a batch
 * whenever one of `slots` device
 * scratch/stream slots is free, admitting requests while the batch's device
 * scratch estimate stays within slot_bytes (at least one request), and runs
 * the batch asynchronously; batches on different slots overlap. Results are
 * those of svs_score_batch for the same request alone (bit-exact). Indexes
 * must outlive the server and must not be modified while it runs. */
typedef struct svs_server svs_server;
typedef struct {
    uint32_t max_batch;   /* queries per batch, 1..4096 */
    uint32_t slots;       /* batches in flight, 1..8 */
    uint64_t slot_bytes;  /* device scratch budget per batch (0 = 1/16 of the
                             context budget, at least 1 GiB); a request routed
                             on the device counts its nprobe largest lists */
    int mode;             /* SVS_MODE_BATCHED or SVS_MODE_BATCHED_VALU */
    uint32_t cross_tenant;  /* with several tenants: 1 = a batch takes the
                               pending requests of every tenant (each query
                               encoded with its tenant's keys, routed on its
                               tenant's centroids, scored on its tenant's
                               lists, through one table of all tenants'
                               lists); 0 (the better default when tenants'
                               queries share their lists) = a batch holds one
                               tenant */
} svs_server_config;
typedef struct {
    uint32_t tenant;         /* position in the tenants array */
    const uint64_t* v_q;     /* 1024 canonical u64 (quantised query) */
    const uint8_t* q_seed;   /* 32 bytes */
    const uint32_t* probes;  /* probed list ids, in probe order */
    uint32_t n_probes;
    uint8_t verify;          /* Freivalds on every probed list */
    uint32_t k;              /* 1..4096 */
    double arrival_s;        /* 0 (= now), or a scheduled arrival on the
                                library's steady clock (load generator) */
    /* A raw query instead (as svs_score_raw_batch): when `query` is set,
     * v_q and probes are ignored and the batch quantises with
     * quantisation_q and routes to `nprobe` lists on the device. */
    const float* query;      /* d f32, d = the tenant's centroids' d */
    uint32_t nprobe;
    uint64_t quantisation_q;
} svs_request;
typedef struct {
    uint32_t* ids;        /* caller buffer of k (may be NULL) */
    int64_t* scores;      /* caller buffer of k (may be NULL) */
    uint32_t count;
    int32_t status;       /* as svs_query_batch.out_status */
    int32_t rc;           /* SVS_OK or the batch's error code */
    uint32_t batch_size;  /* queries in the batch the request ran in */
    double latency_ms;    /* submit (or scheduled arrival) to completion */
} svs_result;
typedef struct {
    uint64_t submitted, completed, batches, batched_queries;
    uint32_t max_batch, max_in_flight, pending;
    uint64_t graph_launches, graph_captures;  /* over the server's slots */
} svs_server_counters;

int svs_server_create(svs_index* const* tenants, uint32_t n_tenants, const svs_server_config* cfg,
                      svs_server** out);
/* Sizes every slot's device scratch and pinned staging for batches of up to
 * n_queries queries, n_items (query, list) items, n_candidates candidate
 * rows and top-k of k, so that serving never reallocates (a device free
 * synchronises all streams). Optional, and only before the first request;
 * buffers otherwise grow on demand. */
int svs_server_reserve(svs_server* s, uint32_t n_queries, uint32_t n_items,
                       uint64_t n_candidates, uint32_t k);
/* Finishes every pending request, then stops. */
int svs_server_destroy(svs_server* s);
int svs_server_submit(svs_server* s, const svs_request* req, uint64_t* ticket);
/* Blocks until the request is done; each ticket is waited for once. */
int svs_server_wait(svs_server* s, uint64_t ticket, svs_result* out);
int svs_server_stats(svs_server* s, svs_server_counters* out);
/* Open-loop Poisson load generator: request i is submitted at its scheduled
 * arrival (exponential gaps at `qps`, seeded) regardless of earlier ones;
 * results[i] as svs_server_wait, latency from the scheduled arrival. */
int svs_server_run_poisson(svs_server* s, const svs_request* reqs, uint32_t n, double qps,
                           uint64_t seed, svs_result* results, double* wall_s);

/* ---- EMVP score GEMV (same kernel family) ---------------------------------
 * Resident EMVP cluster M_hat (m x 1292 canonical u64, SEC128 n = 1292,
 * s = 76, b = 17; scorer-emvp/src/params.rs). Replaces EMVP's
 * FlatGpuState/IvfGpuState::build + compute_products
 * (scorer-emvp/src/gpu/mod.rs:217-380) and emvp_compute_products
 * (scorer-emvp/src/gpu/kernels.cu:90-110). */
typedef struct svs_emvp svs_emvp;
int svs_emvp_create(svs_ctx* ctx, uint32_t m, const uint64_t* m_hat_host, svs_emvp** out);
int svs_emvp_destroy(svs_emvp* e);
/* compute_products / compute_products_batch (scorer-emvp/src/crypto.rs:200-222,
 * 369-415): for each query q, partials[q][si * m + i] (s x m, block-major)
 * = sum_{l<17} M_hat[i][17 si + l] q_hat[q][17 si + l] mod p. SVS_NONE for
 * an empty cluster. */
int svs_emvp_compute_products(svs_emvp* e, uint32_t n_queries, const uint64_t* q_hat_host,
                              uint64_t* partials_host);
/* Same with device pointers (q_hat_dev canonical, n_queries x 1292;
 * partials_dev n_queries x 76 x m), asynchronous on the context stream:
 * the resident-input path, without the partials' trip over PCIe. */
int svs_emvp_compute_products_dev(svs_emvp* e, uint32_t n_queries, const uint64_t* q_hat_dev,
                                  uint64_t* partials_dev);
/* decode_scores (scorer-emvp/src/crypto.rs:435-453; the reference runs it
 * on the client CPU, scorer-emvp/src/lib.rs:549,567): for each query q and
 * row i, scores[q][i] = decode_field(sum_si partials[q][si * m + i] -
 * R[i] . q_hat[q]) as f32, R[i] replayed on the device from row_rng(r_seed,
 * i) (crypto.rs:100-107). q_hat and partials must be canonical (checked:
 * SVS_ERR_ARG). Host buffers; SVS_NONE for an empty cluster. */
int svs_emvp_decode_scores(svs_emvp* e, uint32_t n_queries, const uint8_t* r_seed,
                           const uint64_t* q_hat_host, const uint64_t* partials_host,
                           float* scores_host);
/* Same with device pointers (not checked), asynchronous on the context
 * stream. */
int svs_emvp_decode_scores_dev(svs_emvp* e, uint32_t n_queries, const uint8_t* r_seed,
                               const uint64_t* q_hat_dev, const uint64_t* partials_dev,
                               float* scores_dev);
/* compute_products then decode_scores for n_queries queries of this cluster
 * (EmvpScorer's per-cluster work, scorer-emvp/src/lib.rs:536-567): q_hat in,
 * f32 scores (n_queries x m) out; the partials never leave the device. */
int svs_emvp_score(svs_emvp* e, uint32_t n_queries, const uint8_t* r_seed,
                   const uint64_t* q_hat_host, float* scores_host);

/* ---- Tiptoe BFV hint compression -------------------------------------------
 * decompose_hint + apply_hint (scorer-tiptoe/src/bfv/compress.rs:146-190,
 * 241-266; the reference runs both on the CPU through the fhe crate) with
 * the limb plaintexts resident in HBM. BFV at tiptoe_text_params' sizes
 * (scorer-tiptoe/src/bfv/mod.rs:50-70): N = 2048, t = 65537, one 38-bit
 * modulus q = the largest prime below 2^38 that is 1 mod 4096
 * (274877820929). Polynomials are in the negacyclic NTT domain of
 * psi = g^((q-1)/4096), g the smallest primitive root mod q, natural order
 * in and bit-reversed order out (slot i of the transform of a is
 * sum_k a_k psi^((2 bitrev(i) + 1) k)); query tokens must be encrypted in
 * that domain. */
typedef struct svs_tiptoe_hint svs_tiptoe_hint;
/* The hint H (r x n_lwe u64, row-major) split into its top 8 four-bit limbs,
 * 2048-row chunks and columns, each limb polynomial transformed on the
 * device: 8 x ceil(r / 2048) x n_lwe x 2048 u64 resident. */
int svs_tiptoe_hint_create(svs_ctx* ctx, uint32_t r, uint32_t n_lwe, const uint64_t* hint_host,
                           svs_tiptoe_hint** out);
int svs_tiptoe_hint_destroy(svs_tiptoe_hint* h);
/* The modulus q, the chunk count ceil(r / 2048) and the resident bytes. */
int svs_tiptoe_hint_info(const svs_tiptoe_hint* h, uint64_t* q, uint32_t* chunks,
                         uint64_t* stored_bytes);
/* apply_hint for n_queries query tokens. ct: n_queries x n_lwe x 2 x 2048
 * (ciphertext j of a token, its two polynomials, NTT domain, every value
 * below q: SVS_ERR_ARG otherwise); out: n_queries x 8 x chunks x 2 x 2048,
 * out[b][c] = sum_j ct[j] (x) pt[b][c][j] slot by slot mod q, limb b being
 * bits 60 - 4b .. 63 - 4b. Host buffers. */
int svs_tiptoe_apply_hint(svs_tiptoe_hint* h, uint32_t n_queries, const uint64_t* ct_host,
                          uint64_t* out_host);
/* Same with device pointers (not checked), asynchronous on the context
 * stream. */
int svs_tiptoe_apply_hint_dev(svs_tiptoe_hint* h, uint32_t n_queries, const uint64_t* ct_dev,
                              uint64_t* out_dev);

/* ---- Tiptoe SimplePIR answer (same kernel family) --------------------------
 * Resident database M_db (rows x cols, row-major) of u32 (elem_bytes 4) or
 * u64 (elem_bytes 8) entries. A u32 database whose entries are all < 128 is
 * stored one byte per entry (the int8 MFMA operand). server_answer
 * (scorer-tiptoe/src/pir/simplepir.rs:191-204) has no GPU path in the
 * reference. */
typedef struct svs_tiptoe svs_tiptoe;
int svs_tiptoe_create(svs_ctx* ctx, uint32_t rows, uint32_t cols, int elem_bytes,
                      const void* m_db_host, svs_tiptoe** out);
int svs_tiptoe_destroy(svs_tiptoe* t);
/* Bytes stored on the device and the stored entry width (1, 4 or 8). */
int svs_tiptoe_stored_bytes(const svs_tiptoe* t, uint64_t* bytes, int* elem_bytes);
/* ans[q][i] = sum_j M_db[i][j] ct[q][j] mod 2^32 (a u32 database): the
 * reference's wrapping u64 answer truncated to 32 bits, since 2^32 divides
 * 2^64. kernel: SVS_KERNEL_VALU, SVS_KERNEL_MFMA (byte-stored database,
 * cols % 64 == 0) or -1 (MFMA for batches when possible). */
int svs_tiptoe_answer_u32(svs_tiptoe* t, uint32_t n_queries, const uint32_t* ct_host,
                          uint32_t* ans_host, int kernel);
/* Same with device pointers, asynchronous on the context stream. */
int svs_tiptoe_answer_u32_dev(svs_tiptoe* t, uint32_t n_queries, const uint32_t* ct_dev,
                              uint32_t* ans_dev, int kernel);
/* ans[q][i] = sum_j M_db[i][j] ct[q][j] mod 2^64 (a u64 database), exactly
 * server_answer. */
int svs_tiptoe_answer_u64(svs_tiptoe* t, uint32_t n_queries, const uint64_t* ct_host,
                          uint64_t* ans_host);

/* ---- test hook -------------------------------------------------------------
 * After every server product of svs_score_batch, add `delta` to r[row] of
 * each item probing `list` (list < 0 turns it off): a malicious or faulty
 * server. No counterpart in the reference's API; it models the reference's
 * CPU-forced fault (crypto.rs:1446-1474), which the tests use to show that
 * verification rejects and reports (list, trial). */
int svs_debug_inject_fault(svs_ctx* ctx, int32_t list, uint32_t row, uint64_t delta);

/* ---- host-side client functions (no GPU) ---------------------------------
 * The client half of the protocol that the Rust scorers run on the host,
 * kept native so callers need nothing else. */

/* generate_subspace (crypto.rs:235-238): L, n x n1 row-major. */
int svs_keygen_subspace(const uint8_t key_seed[32], uint64_t* l_out);
/* derive_h_seed (ivf.rs:258-264). */
int svs_keygen_derive_h_seed(uint64_t upload_seed, uint64_t cluster, uint8_t out[32]);
/* domain_separate (crypto.rs:142-150). */
int svs_keygen_domain_separate(const uint8_t seed[32], const uint8_t* tag, size_t tag_len,
                               uint8_t out[32]);
/* regenerate_h_s (crypto.rs:258-270): H (m x n1) and S as CSR. */
int svs_keygen_regenerate_h_s(const uint8_t h_seed[32], uint32_t m, uint64_t* h_out,
                              uint32_t* s_rowptr_out, uint32_t* s_idx_out, int8_t* s_sign_out,
                              size_t s_cap, size_t* nnz_out);
/* quantise_vector (crypto.rs:105-122) for rows x d floats, padded to n. */
int svs_quantise(const float* v, uint32_t rows, uint32_t d, uint64_t q, uint64_t* out);
/* probe_route (ivf-index/src/ivf.rs:150-164): nprobe nearest centroids by
 * f32 L2 (distance.rs:30-37), per query, in ascending distance order. */
int svs_probe_route(const float* centroids, uint32_t n_centroids, uint32_t d,
                    const float* queries, uint32_t n_queries, uint32_t nprobe, uint32_t* out);

#ifdef __cplusplus
}
#endif
#endif /* SVS_GPU_H */
