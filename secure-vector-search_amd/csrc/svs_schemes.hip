// svs_schemes.hip — the two other schemes on the BN kernel family:
//   * EMVP's score GEMV (block partials over F_p), scorer-emvp/src/crypto.rs
//     :200-222 / :369-415, replacing scorer-emvp/src/gpu/mod.rs and its
//     emvp_compute_products kernel (scorer-emvp/src/gpu/kernels.cu:90-110),
//     and EMVP's client decode (crypto.rs:435-453), which has no GPU path in
//     the reference;
//   * Tiptoe's SimplePIR server answer, scorer-tiptoe/src/pir/simplepir.rs
//     :191-204, in the mod-2^32 regime (a u32 database) and mod 2^64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/svs_gpu.h"
#include "svs_internal.h"

using namespace svs;

namespace {

constexpr int EMVP_N = 1292, EMVP_S = 76, EMVP_B = 17;  // SEC128, scorer-emvp/src/params.rs
constexpr int TWAVE = 64;

// ===================================================================== EMVP
// partials[q][si * m + i] = sum_{l < 17} m_hat[i][17 si + l] * q_hat[q][17 si + l]
// (mod p). A work unit is 64 rows x EM_G column blocks (EM_G * 17 u64 =
// 544 contiguous bytes of each row): the block stages it in LDS with 16-byte
// loads that run along the rows, then wave w takes block si = EM_G g + w with
// one lane per row, holds its 17 values in registers across the batch's
// queries (a batch reads m_hat once) and writes 64 consecutive partials per
// query. 17 products of canonical values sum below 2^127: one reduction per
// partial. Units are ordered row tile major, column group minor, and the
// XCD remap gives each XCD a contiguous range of them, so the cache lines
// two neighbouring column groups share stay in one L2.
constexpr int EM_ROWS = 64, EM_G = 4, EM_W = EM_G * EMVP_B, EM_GROUPS = EMVP_S / EM_G;
static_assert(EMVP_S % EM_G == 0 && EM_W % 2 == 0, "EMVP tile shape");

__device__ __forceinline__ uint32_t em_xcd_unit() {
    const uint32_t b = blockIdx.x, g = gridDim.x;
    return (g & 7) ? b : (b & 7) * (g >> 3) + (b >> 3);
}

__global__ void __launch_bounds__(256)
k_emvp(const uint64_t* __restrict__ m_hat, uint32_t m, const uint64_t* __restrict__ q_hat,
       uint32_t nq, uint64_t* __restrict__ partials) {
    __shared__ uint64_t tile[EM_ROWS][EM_W + 1];  // odd row stride: conflict-light reads
    const uint32_t lane = threadIdx.x % TWAVE;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / TWAVE);
    const uint32_t n_units = (m + EM_ROWS - 1) / EM_ROWS * EM_GROUPS;
    for (uint32_t u = em_xcd_unit(); u < n_units; u += gridDim.x) {
        const uint32_t i0 = u / EM_GROUPS * EM_ROWS, g = u % EM_GROUPS;
        __syncthreads();  // the previous unit's LDS reads are done
        for (uint32_t c = threadIdx.x; c < EM_ROWS * (EM_W / 2); c += blockDim.x) {
            const uint32_t r = c / (EM_W / 2), k = c % (EM_W / 2);
            if (i0 + r < m) {
                const ulonglong2 v = *reinterpret_cast<const ulonglong2*>(
                    m_hat + (size_t)(i0 + r) * EMVP_N + (size_t)g * EM_W + 2 * k);
                tile[r][2 * k] = v.x;
                tile[r][2 * k + 1] = v.y;
            }
        }
        __syncthreads();
        const uint32_t i = i0 + lane, si = g * EM_G + wave;
        if (i >= m) continue;
        uint64_t a[EMVP_B];
#pragma unroll
        for (int l = 0; l < EMVP_B; ++l) a[l] = tile[lane][wave * EMVP_B + l];
        for (uint32_t q = 0; q < nq; ++q) {
            const uint64_t* qb = q_hat + (size_t)q * EMVP_N + (size_t)si * EMVP_B;
            Acc128 acc;
            acc.zero();
#pragma unroll
            for (int l = 0; l < EMVP_B; ++l) acc.mac(a[l], qb[l]);
            partials[((size_t)q * EMVP_S + si) * m + i] = acc.reduce();
        }
    }
}

// ============================================================ EMVP decode
// decode_scores (scorer-emvp/src/crypto.rs:435-453) on the device: for row i
// the client replays R[i] from row_rng(r_seed, i) (r_seed with its first 8
// bytes xor i, crypto.rs:100-107): R[i][j] = next_u64 % p, the u64 draw j
// being keystream words 2j (low) and 2j + 1 (high), i.e. block j / 8 of the
// row's ChaCha20 stream. score = decode_field(sum_si partials[si m + i] -
// R[i] . q_hat). One wave per row: lane l generates blocks l, l + 64 and
// l + 128 (162 blocks cover the 1292 draws), keeps its 24 draws in registers
// for all of the batch's queries, multiplies them into a 128-bit lazy sum
// (24 products of canonical values stay below 2^127) and adds its one or two
// block partials; an add_mod butterfly over the wave gives the row's value.
// The stream generation is VALU work (about 1,000 integer ops a block), so
// the kernel is bound by the ALUs, not HBM: a batch of queries shares it.
constexpr int EMVP_BLOCKS = (2 * EMVP_N + 15) / 16;  // 162
constexpr int ED_T = (EMVP_BLOCKS + TWAVE - 1) / TWAVE;  // 3 blocks per lane
static_assert(EMVP_S <= 2 * TWAVE, "two block partials per lane");

__device__ __forceinline__ uint64_t mod_p_u64(uint64_t x) {
    const uint64_t r = (x & P) + (x >> 61);  // <= 2^61 + 6
    return r >= P ? r - P : r;
}

// decode_field (crypto.rs:54-62): signed lift, / Q^2 in f64, then f32.
__device__ __forceinline__ float emvp_decode_field(uint64_t v) {
    const int64_t s = v > P / 2 ? (int64_t)v - (int64_t)P : (int64_t)v;
    return (float)((double)s / (1048576.0 * 1048576.0));
}

__global__ void __launch_bounds__(256)
k_emvp_decode(uint32_t m, uint4 seed_lo, uint4 seed_hi, const uint64_t* __restrict__ q_hat,
              uint32_t nq, const uint64_t* __restrict__ partials, float* __restrict__ scores) {
    const uint32_t lane = threadIdx.x % TWAVE;
    const uint32_t wpb = blockDim.x / TWAVE;
    for (uint32_t i = blockIdx.x * wpb + threadIdx.x / TWAVE; i < m; i += gridDim.x * wpb) {
        // row_rng: the row index (u64 LE) xor the seed's first 8 bytes
        const uint32_t key[8] = {seed_lo.x ^ i, seed_lo.y, seed_lo.z, seed_lo.w,
                                 seed_hi.x,     seed_hi.y, seed_hi.z, seed_hi.w};
        uint64_t r[ED_T][8];
#pragma unroll
        for (int t = 0; t < ED_T; ++t) {
            const uint32_t b = lane + TWAVE * t;
            uint32_t w[16];
            chacha20_block(key, b, w);
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const bool live = b < EMVP_BLOCKS && 8 * b + e < (uint32_t)EMVP_N;
                r[t][e] = live ? mod_p_u64((uint64_t)w[2 * e] | ((uint64_t)w[2 * e + 1] << 32)) : 0;
            }
        }
        for (uint32_t q = 0; q < nq; ++q) {
            const uint64_t* qh = q_hat + (size_t)q * EMVP_N;
            Acc128 acc;
            acc.zero();
#pragma unroll
            for (int t = 0; t < ED_T; ++t) {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const uint32_t j = 8 * (lane + TWAVE * t) + e;
                    if (j < (uint32_t)EMVP_N) acc.mac(r[t][e], qh[j]);
                }
            }
            const uint64_t* pq = partials + (size_t)q * EMVP_S * m + i;
            uint64_t bs = lane < (uint32_t)EMVP_S ? pq[(size_t)lane * m] : 0;
            if (lane + TWAVE < (uint32_t)EMVP_S) bs = add_mod(bs, pq[(size_t)(lane + TWAVE) * m]);
            uint64_t v = sub_mod(bs, acc.reduce());
#pragma unroll
            for (int off = TWAVE / 2; off >= 1; off >>= 1) {
                const uint64_t o = ((uint64_t)__shfl_xor((uint32_t)(v >> 32), off, TWAVE) << 32) |
                                   (uint32_t)__shfl_xor((uint32_t)v, off, TWAVE);
                v = add_mod(v, o);
            }
            if (lane == 0) scores[(size_t)q * m + i] = emvp_decode_field(v);
        }
    }
}

// =================================================================== Tiptoe
// VALU answer, one wave per row: ans[q][i] = sum_j M[i][j] ct[q][j] wrapping
// at 2^32 (or 2^64). For 8-bit database entries the 32-bit product is split
// into two 24-bit multiply-adds: M * ct = M * lo16 + (M * hi16) << 16.
template <typename T>
__global__ void __launch_bounds__(256)
k_tiptoe_u32(const T* __restrict__ db, uint32_t rows, uint32_t cols,
             const uint32_t* __restrict__ ct, uint32_t nq, uint32_t* __restrict__ ans) {
    const uint32_t lane = threadIdx.x % TWAVE;
    const uint32_t wpb = blockDim.x / TWAVE;
    for (uint32_t i = blockIdx.x * wpb + threadIdx.x / TWAVE; i < rows; i += gridDim.x * wpb) {
        const T* row = db + (size_t)i * cols;
        for (uint32_t q = 0; q < nq; ++q) {
            const uint32_t* c = ct + (size_t)q * cols;
            uint32_t lo = 0, hi = 0;
            for (uint32_t j = lane; j < cols; j += TWAVE) {
                const uint32_t mv = (uint32_t)row[j];
                const uint32_t cv = c[j];
                if (sizeof(T) == 1) {
                    lo += __umul24(mv, cv & 0xFFFFu);
                    hi += __umul24(mv, cv >> 16);
                } else {
                    lo += mv * cv;
                }
            }
            uint32_t s = lo + (hi << 16);
#pragma unroll
            for (int off = TWAVE / 2; off >= 1; off >>= 1) s += __shfl_xor(s, off, TWAVE);
            if (lane == 0) ans[(size_t)q * rows + i] = s;
        }
    }
}

__global__ void __launch_bounds__(256)
k_tiptoe_u64(const uint64_t* __restrict__ db, uint32_t rows, uint32_t cols,
             const uint64_t* __restrict__ ct, uint32_t nq, uint64_t* __restrict__ ans) {
    const uint32_t lane = threadIdx.x % TWAVE;
    const uint32_t wpb = blockDim.x / TWAVE;
    for (uint32_t i = blockIdx.x * wpb + threadIdx.x / TWAVE; i < rows; i += gridDim.x * wpb) {
        const uint64_t* row = db + (size_t)i * cols;
        for (uint32_t q = 0; q < nq; ++q) {
            const uint64_t* c = ct + (size_t)q * cols;
            uint64_t s = 0;
            for (uint32_t j = lane; j < cols; j += TWAVE) s += row[j] * c[j];
#pragma unroll
            for (int off = TWAVE / 2; off >= 1; off >>= 1) {
                uint32_t lo = __shfl_xor((uint32_t)s, off, TWAVE);
                uint32_t hi = __shfl_xor((uint32_t)(s >> 32), off, TWAVE);
                s += ((uint64_t)hi << 32) | lo;
            }
            if (lane == 0) ans[(size_t)q * rows + i] = s;
        }
    }
}

// MFMA answer for a batch (database entries < 128, cols % 64 == 0).
// ct[q][j] = sum_b 256^b c_b with bytes c_b; with t_b = c_b - 128,
//   sum_j M_ij c_b[j] = D_b + 128 rowsum_i,  D_b = sum_j M_ij t_b[j],
// |D_b| < 2^31, and ans = sum_b 256^b (D_b + 128 rowsum_i) mod 2^32.
// A = database rows (u8, K = j: plain 16-byte loads), B column (query, b) =
// the signed byte plane t_b of that query, prepared once per batch.
typedef int v4i __attribute__((ext_vector_type(4)));
constexpr int TT_RT = 2;  // 16-row tiles per wave
constexpr int TT_QT = 4;  // 16-column tiles (4 queries each) per wave

// planes[q][b][j] = (int8)(byte b of ct[q][j]) - 128
__global__ void k_tiptoe_planes(const uint32_t* __restrict__ ct, uint32_t cols, uint32_t nq,
                                int8_t* __restrict__ planes) {
    const uint64_t total = (uint64_t)nq * cols;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t q = (uint32_t)(t / cols), j = (uint32_t)(t % cols);
        const uint32_t v = ct[t];
#pragma unroll
        for (int b = 0; b < 4; ++b)
            planes[((size_t)q * 4 + b) * cols + j] = (int8_t)((int)((v >> (8 * b)) & 0xFF) - 128);
    }
}

__global__ void __launch_bounds__(256)
k_tiptoe_mfma(const uint8_t* __restrict__ db, const uint32_t* __restrict__ rowsum, uint32_t rows,
              uint32_t cols, const int8_t* __restrict__ planes, uint32_t nq,
              uint32_t* __restrict__ ans) {
    const int lane = threadIdx.x % TWAVE;
    const int grp = lane / 16, col = lane % 16;
    const uint32_t wpb = blockDim.x / TWAVE;
    const uint32_t row_tiles = (rows + 16 * TT_RT - 1) / (16 * TT_RT);
    const uint32_t q_tiles = (nq + 4 * TT_QT - 1) / (4 * TT_QT);
    const uint64_t n_work = (uint64_t)row_tiles * q_tiles;
    for (uint64_t w = (uint64_t)blockIdx.x * wpb + threadIdx.x / TWAVE; w < n_work;
         w += (uint64_t)gridDim.x * wpb) {
        const uint32_t rtile = (uint32_t)(w % row_tiles), qtile = (uint32_t)(w / row_tiles);
        const uint32_t r0 = rtile * 16 * TT_RT, q0 = qtile * 4 * TT_QT;
        const uint8_t* arow[TT_RT];
#pragma unroll
        for (int rt = 0; rt < TT_RT; ++rt)
            arow[rt] = db + (size_t)min(r0 + 16 * rt + col, rows - 1) * cols + 16 * grp;
        const int8_t* bcol[TT_QT];
#pragma unroll
        for (int qt = 0; qt < TT_QT; ++qt) {
            const uint32_t q = min(q0 + 4 * qt + col / 4, nq - 1);
            bcol[qt] = planes + ((size_t)q * 4 + col % 4) * cols + 16 * grp;
        }
        v4i acc[TT_RT][TT_QT];
#pragma unroll
        for (int rt = 0; rt < TT_RT; ++rt)
#pragma unroll
            for (int qt = 0; qt < TT_QT; ++qt) acc[rt][qt] = v4i{0, 0, 0, 0};
        for (uint32_t s = 0; s < cols / 64; ++s) {
            v4i a[TT_RT], b[TT_QT];
#pragma unroll
            for (int rt = 0; rt < TT_RT; ++rt) a[rt] = *reinterpret_cast<const v4i*>(arow[rt] + 64 * s);
#pragma unroll
            for (int qt = 0; qt < TT_QT; ++qt) b[qt] = *reinterpret_cast<const v4i*>(bcol[qt] + 64 * s);
#pragma unroll
            for (int rt = 0; rt < TT_RT; ++rt)
#pragma unroll
                for (int qt = 0; qt < TT_QT; ++qt)
                    acc[rt][qt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[rt], b[qt], acc[rt][qt],
                                                                         0, 0, 0);
        }
        // lane (col = 4 qq + b) holds D_b of rows 4 grp + i: weight by 256^b,
        // sum the four limb lanes, add the offset term.
#pragma unroll
        for (int rt = 0; rt < TT_RT; ++rt)
#pragma unroll
            for (int qt = 0; qt < TT_QT; ++qt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    uint32_t v = (uint32_t)acc[rt][qt][i] << (8 * (col % 4));
                    v += __shfl_xor(v, 1, TWAVE);
                    v += __shfl_xor(v, 2, TWAVE);
                    const uint32_t row = r0 + 16 * rt + 4 * grp + i;
                    const uint32_t q = q0 + 4 * qt + col / 4;
                    if (col % 4 == 0 && row < rows && q < nq)
                        ans[(size_t)q * rows + row] = v + 128u * rowsum[row] * 0x01010101u;
                }
    }
}

__global__ void k_tiptoe_rowsum(const uint8_t* __restrict__ db, uint32_t rows, uint32_t cols,
                                uint32_t* __restrict__ rowsum) {
    const uint32_t lane = threadIdx.x % TWAVE;
    for (uint32_t i = blockIdx.x * (blockDim.x / TWAVE) + threadIdx.x / TWAVE; i < rows;
         i += gridDim.x * (blockDim.x / TWAVE)) {
        uint32_t s = 0;
        for (uint32_t j = lane; j < cols; j += TWAVE) s += db[(size_t)i * cols + j];
#pragma unroll
        for (int off = TWAVE / 2; off >= 1; off >>= 1) s += __shfl_xor(s, off, TWAVE);
        if (lane == 0) rowsum[i] = s;
    }
}

}  // namespace

// ------------------------------------------------------------------- state
struct svs_emvp {
    svs_ctx* ctx = nullptr;
    uint32_t m = 0;
    uint64_t* m_hat = nullptr;
    DevBuf q, out, sc;
};

struct svs_tiptoe {
    svs_ctx* ctx = nullptr;
    uint32_t rows = 0, cols = 0;
    int elem = 0;             // 1: packed u8, 4: u32, 8: u64
    void* db = nullptr;
    size_t db_bytes = 0;
    uint32_t* rowsum = nullptr;
    DevBuf ct, planes, ans;
};

static uint32_t grid_for(svs_ctx* ctx, uint64_t work, uint32_t per_block) {
    return (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>((work + per_block - 1) / per_block, (uint64_t)ctx->n_cu * 16));
}

extern "C" int svs_emvp_create(svs_ctx* ctx, uint32_t m, const uint64_t* m_hat, svs_emvp** out) {
    if (!ctx || !out || (m && !m_hat)) return fail(SVS_ERR_ARG, "null argument");
    for (size_t i = 0; i < (size_t)m * EMVP_N; ++i)
        if (m_hat[i] >= P) return fail(SVS_ERR_ARG, "m_hat[%zu] is not canonical", i);
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    auto* e = new svs_emvp();
    e->ctx = ctx;
    e->m = m;
    int rc = dev_alloc(ctx, (void**)&e->m_hat, sizeof(uint64_t) * (size_t)m * EMVP_N);
    if (rc != SVS_OK) {
        delete e;
        return rc;
    }
    if (m) HIP_TRY(hipMemcpy(e->m_hat, m_hat, sizeof(uint64_t) * (size_t)m * EMVP_N,
                             hipMemcpyHostToDevice));
    ctx_retain(ctx);
    *out = e;
    return SVS_OK;
}

extern "C" int svs_emvp_destroy(svs_emvp* e) {
    if (!e) return SVS_OK;
    svs_ctx* ctx = e->ctx;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        hipSetDevice(ctx->device);
        hipStreamSynchronize(ctx->stream);
        dev_free(ctx, e->m_hat, sizeof(uint64_t) * (size_t)e->m * EMVP_N);
        e->q.release(ctx);
        e->out.release(ctx);
        e->sc.release(ctx);
        delete e;
    }
    ctx_release(ctx);
    return SVS_OK;
}

static int emvp_launch(svs_emvp* e, uint32_t n_queries, const uint64_t* q_dev, uint64_t* out_dev) {
    svs_ctx* ctx = e->ctx;
    const uint64_t units = (uint64_t)(e->m + EM_ROWS - 1) / EM_ROWS * EM_GROUPS;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(units, (uint64_t)ctx->n_cu * 4);
    hipLaunchKernelGGL(k_emvp, dim3(blocks), dim3(256), 0, ctx->stream, (const uint64_t*)e->m_hat,
                       e->m, q_dev, n_queries, out_dev);
    HIP_TRY(hipGetLastError());
    ctx->launches++;
    return SVS_OK;
}

extern "C" int svs_emvp_compute_products_dev(svs_emvp* e, uint32_t n_queries,
                                             const uint64_t* q_hat_dev, uint64_t* partials_dev) {
    if (!e || (n_queries && (!q_hat_dev || !partials_dev))) return fail(SVS_ERR_ARG, "null argument");
    if (e->m == 0 || n_queries == 0) return e->m == 0 ? SVS_NONE : SVS_OK;
    svs_ctx* ctx = e->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    return emvp_launch(e, n_queries, q_hat_dev, partials_dev);
}

extern "C" int svs_emvp_compute_products(svs_emvp* e, uint32_t n_queries, const uint64_t* q_hat,
                                         uint64_t* partials) {
    if (!e || (n_queries && (!q_hat || !partials))) return fail(SVS_ERR_ARG, "null argument");
    if (e->m == 0 || n_queries == 0) return e->m == 0 ? SVS_NONE : SVS_OK;
    for (size_t i = 0; i < (size_t)n_queries * EMVP_N; ++i)
        if (q_hat[i] >= P) return fail(SVS_ERR_ARG, "q_hat[%zu] is not canonical", i);
    svs_ctx* ctx = e->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t out_n = (size_t)n_queries * EMVP_S * e->m;
    SVS_TRY(e->q.ensure(ctx, sizeof(uint64_t) * (size_t)n_queries * EMVP_N));
    SVS_TRY(e->out.ensure(ctx, sizeof(uint64_t) * out_n));
    HIP_TRY(hipMemcpyAsync(e->q.ptr, q_hat, sizeof(uint64_t) * (size_t)n_queries * EMVP_N,
                           hipMemcpyHostToDevice, ctx->stream));
    SVS_TRY(emvp_launch(e, n_queries, (const uint64_t*)e->q.ptr, (uint64_t*)e->out.ptr));
    HIP_TRY(hipMemcpyAsync(partials, e->out.ptr, sizeof(uint64_t) * out_n, hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SVS_OK;
}

// ------------------------------------------------------------------ Tiptoe
static int emvp_decode_launch(svs_emvp* e, uint32_t n_queries, const uint8_t* r_seed,
                              const uint64_t* q_dev, const uint64_t* partials_dev,
                              float* scores_dev) {
    svs_ctx* ctx = e->ctx;
    uint32_t k[8];
    seed_to_key(r_seed, k);
    const uint4 lo = make_uint4(k[0], k[1], k[2], k[3]), hi = make_uint4(k[4], k[5], k[6], k[7]);
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(((uint64_t)e->m + 3) / 4,
                                                         (uint64_t)ctx->n_cu * 16);
    hipLaunchKernelGGL(k_emvp_decode, dim3(blocks), dim3(256), 0, ctx->stream, e->m, lo, hi,
                       q_dev, n_queries, partials_dev, scores_dev);
    HIP_TRY(hipGetLastError());
    ctx->launches++;
    return SVS_OK;
}

extern "C" int svs_emvp_decode_scores_dev(svs_emvp* e, uint32_t n_queries, const uint8_t* r_seed,
                                          const uint64_t* q_hat_dev, const uint64_t* partials_dev,
                                          float* scores_dev) {
    if (!e || !r_seed || (n_queries && (!q_hat_dev || !partials_dev || !scores_dev)))
        return fail(SVS_ERR_ARG, "null argument");
    if (e->m == 0 || n_queries == 0) return e->m == 0 ? SVS_NONE : SVS_OK;
    svs_ctx* ctx = e->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    return emvp_decode_launch(e, n_queries, r_seed, q_hat_dev, partials_dev, scores_dev);
}

static int check_canonical(const uint64_t* a, size_t n, const char* what) {
    for (size_t i = 0; i < n; ++i)
        if (a[i] >= P) return fail(SVS_ERR_ARG, "%s[%zu] is not canonical", what, i);
    return SVS_OK;
}

extern "C" int svs_emvp_decode_scores(svs_emvp* e, uint32_t n_queries, const uint8_t* r_seed,
                                      const uint64_t* q_hat, const uint64_t* partials,
                                      float* scores) {
    if (!e || !r_seed || (n_queries && (!q_hat || !partials || !scores)))
        return fail(SVS_ERR_ARG, "null argument");
    if (e->m == 0 || n_queries == 0) return e->m == 0 ? SVS_NONE : SVS_OK;
    const size_t nq_n = (size_t)n_queries * EMVP_N, np_ = (size_t)n_queries * EMVP_S * e->m;
    SVS_TRY(check_canonical(q_hat, nq_n, "q_hat"));
    SVS_TRY(check_canonical(partials, np_, "partials"));
    svs_ctx* ctx = e->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    SVS_TRY(e->q.ensure(ctx, sizeof(uint64_t) * nq_n));
    SVS_TRY(e->out.ensure(ctx, sizeof(uint64_t) * np_));
    SVS_TRY(e->sc.ensure(ctx, sizeof(float) * (size_t)n_queries * e->m));
    HIP_TRY(hipMemcpyAsync(e->q.ptr, q_hat, sizeof(uint64_t) * nq_n, hipMemcpyHostToDevice,
                           ctx->stream));
    HIP_TRY(hipMemcpyAsync(e->out.ptr, partials, sizeof(uint64_t) * np_, hipMemcpyHostToDevice,
                           ctx->stream));
    SVS_TRY(emvp_decode_launch(e, n_queries, r_seed, (const uint64_t*)e->q.ptr,
                               (const uint64_t*)e->out.ptr, (float*)e->sc.ptr));
    HIP_TRY(hipMemcpyAsync(scores, e->sc.ptr, sizeof(float) * (size_t)n_queries * e->m,
                           hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SVS_OK;
}

extern "C" int svs_emvp_score(svs_emvp* e, uint32_t n_queries, const uint8_t* r_seed,
                              const uint64_t* q_hat, float* scores) {
    if (!e || !r_seed || (n_queries && (!q_hat || !scores))) return fail(SVS_ERR_ARG, "null argument");
    if (e->m == 0 || n_queries == 0) return e->m == 0 ? SVS_NONE : SVS_OK;
    const size_t nq_n = (size_t)n_queries * EMVP_N, np_ = (size_t)n_queries * EMVP_S * e->m;
    SVS_TRY(check_canonical(q_hat, nq_n, "q_hat"));
    svs_ctx* ctx = e->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    SVS_TRY(e->q.ensure(ctx, sizeof(uint64_t) * nq_n));
    SVS_TRY(e->out.ensure(ctx, sizeof(uint64_t) * np_));
    SVS_TRY(e->sc.ensure(ctx, sizeof(float) * (size_t)n_queries * e->m));
    HIP_TRY(hipMemcpyAsync(e->q.ptr, q_hat, sizeof(uint64_t) * nq_n, hipMemcpyHostToDevice,
                           ctx->stream));
    SVS_TRY(emvp_launch(e, n_queries, (const uint64_t*)e->q.ptr, (uint64_t*)e->out.ptr));
    SVS_TRY(emvp_decode_launch(e, n_queries, r_seed, (const uint64_t*)e->q.ptr,
                               (const uint64_t*)e->out.ptr, (float*)e->sc.ptr));
    HIP_TRY(hipMemcpyAsync(scores, e->sc.ptr, sizeof(float) * (size_t)n_queries * e->m,
                           hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SVS_OK;
}

extern "C" int svs_tiptoe_create(svs_ctx* ctx, uint32_t rows, uint32_t cols, int elem_bytes,
                                 const void* m_db, svs_tiptoe** out) {
    if (!ctx || !out || !m_db || rows == 0 || cols == 0) return fail(SVS_ERR_ARG, "bad argument");
    if (elem_bytes != 4 && elem_bytes != 8) return fail(SVS_ERR_ARG, "elem_bytes must be 4 or 8");
    const size_t n = (size_t)rows * cols;
    // A u32 database whose entries all fit in 7 bits is stored packed (one
    // byte per entry) — the layout the MFMA path reads directly.
    bool pack = elem_bytes == 4;
    if (pack) {
        const uint32_t* d = (const uint32_t*)m_db;
        for (size_t i = 0; i < n && pack; ++i) pack = d[i] < 128;
    }
    std::vector<uint8_t> packed;
    if (pack) {
        packed.resize(n);
        const uint32_t* d = (const uint32_t*)m_db;
        for (size_t i = 0; i < n; ++i) packed[i] = (uint8_t)d[i];
    }
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    auto* t = new svs_tiptoe();
    t->ctx = ctx;
    t->rows = rows;
    t->cols = cols;
    t->elem = pack ? 1 : elem_bytes;
    t->db_bytes = n * (size_t)t->elem;
    int rc = dev_alloc(ctx, &t->db, t->db_bytes);
    if (rc == SVS_OK && pack) rc = dev_alloc(ctx, (void**)&t->rowsum, sizeof(uint32_t) * rows);
    if (rc != SVS_OK) {
        dev_free(ctx, t->db, t->db_bytes);
        delete t;
        return rc;
    }
    HIP_TRY(hipMemcpy(t->db, pack ? (const void*)packed.data() : m_db, t->db_bytes,
                      hipMemcpyHostToDevice));
    if (pack) {
        hipLaunchKernelGGL(k_tiptoe_rowsum, dim3(grid_for(ctx, rows, 4)), dim3(256), 0, ctx->stream,
                           (const uint8_t*)t->db, rows, cols, t->rowsum);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    ctx_retain(ctx);
    *out = t;
    return SVS_OK;
}

extern "C" int svs_tiptoe_destroy(svs_tiptoe* t) {
    if (!t) return SVS_OK;
    svs_ctx* ctx = t->ctx;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        hipSetDevice(ctx->device);
        hipStreamSynchronize(ctx->stream);
        dev_free(ctx, t->db, t->db_bytes);
        if (t->rowsum) dev_free(ctx, t->rowsum, sizeof(uint32_t) * t->rows);
        t->ct.release(ctx);
        t->planes.release(ctx);
        t->ans.release(ctx);
        delete t;
    }
    ctx_release(ctx);
    return SVS_OK;
}

extern "C" int svs_tiptoe_stored_bytes(const svs_tiptoe* t, uint64_t* bytes, int* elem_bytes) {
    if (!t) return fail(SVS_ERR_ARG, "null argument");
    if (bytes) *bytes = t->db_bytes;
    if (elem_bytes) *elem_bytes = t->elem;
    return SVS_OK;
}

// Device-pointer variant (asynchronous on the context stream).
extern "C" int svs_tiptoe_answer_u32_dev(svs_tiptoe* t, uint32_t n_queries, const uint32_t* ct_dev,
                                         uint32_t* ans_dev, int kernel) {
    if (!t || !ct_dev || !ans_dev) return fail(SVS_ERR_ARG, "null argument");
    if (t->elem == 8) return fail(SVS_ERR_STATE, "database was uploaded as u64");
    if (kernel != -1 && kernel != SVS_KERNEL_VALU && kernel != SVS_KERNEL_MFMA)
        return fail(SVS_ERR_ARG, "unknown kernel %d", kernel);
    if (n_queries == 0) return SVS_OK;
    svs_ctx* ctx = t->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const bool mfma_ok = t->elem == 1 && t->cols % 64 == 0;
    const bool use_mfma = kernel == SVS_KERNEL_MFMA || (kernel < 0 && mfma_ok && n_queries > 1);
    if (use_mfma) {
        if (!mfma_ok)
            return fail(SVS_ERR_STATE, "MFMA answer needs 7-bit entries and cols %% 64 == 0");
        SVS_TRY(t->planes.ensure(ctx, (size_t)n_queries * 4 * t->cols));
        hipLaunchKernelGGL(k_tiptoe_planes, dim3(grid_for(ctx, (uint64_t)n_queries * t->cols, 256)),
                           dim3(256), 0, ctx->stream, ct_dev, t->cols, n_queries,
                           (int8_t*)t->planes.ptr);
        const uint64_t work = (uint64_t)((t->rows + 16 * TT_RT - 1) / (16 * TT_RT)) *
                              ((n_queries + 4 * TT_QT - 1) / (4 * TT_QT));
        hipLaunchKernelGGL(k_tiptoe_mfma, dim3(grid_for(ctx, work, 4)), dim3(256), 0, ctx->stream,
                           (const uint8_t*)t->db, (const uint32_t*)t->rowsum, t->rows, t->cols,
                           (const int8_t*)t->planes.ptr, n_queries, ans_dev);
        ctx->launches += 2;
    } else if (t->elem == 1) {
        hipLaunchKernelGGL(k_tiptoe_u32<uint8_t>, dim3(grid_for(ctx, t->rows, 4)), dim3(256), 0,
                           ctx->stream, (const uint8_t*)t->db, t->rows, t->cols, ct_dev,
                           n_queries, ans_dev);
        ctx->launches++;
    } else {
        hipLaunchKernelGGL(k_tiptoe_u32<uint32_t>, dim3(grid_for(ctx, t->rows, 4)), dim3(256), 0,
                           ctx->stream, (const uint32_t*)t->db, t->rows, t->cols, ct_dev,
                           n_queries, ans_dev);
        ctx->launches++;
    }
    HIP_TRY(hipGetLastError());
    return SVS_OK;
}

extern "C" int svs_tiptoe_answer_u32(svs_tiptoe* t, uint32_t n_queries, const uint32_t* ct,
                                     uint32_t* ans, int kernel) {
    if (!t || (n_queries && (!ct || !ans))) return fail(SVS_ERR_ARG, "null argument");
    if (n_queries == 0) return SVS_OK;
    svs_ctx* ctx = t->ctx;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        HIP_TRY(hipSetDevice(ctx->device));
        SVS_TRY(t->ct.ensure(ctx, sizeof(uint32_t) * (size_t)n_queries * t->cols));
        SVS_TRY(t->ans.ensure(ctx, sizeof(uint32_t) * (size_t)n_queries * t->rows));
        HIP_TRY(hipMemcpyAsync(t->ct.ptr, ct, sizeof(uint32_t) * (size_t)n_queries * t->cols,
                               hipMemcpyHostToDevice, ctx->stream));
    }
    SVS_TRY(svs_tiptoe_answer_u32_dev(t, n_queries, (const uint32_t*)t->ct.ptr,
                                      (uint32_t*)t->ans.ptr, kernel));
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipMemcpyAsync(ans, t->ans.ptr, sizeof(uint32_t) * (size_t)n_queries * t->rows,
                           hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SVS_OK;
}

extern "C" int svs_tiptoe_answer_u64(svs_tiptoe* t, uint32_t n_queries, const uint64_t* ct,
                                     uint64_t* ans) {
    if (!t || (n_queries && (!ct || !ans))) return fail(SVS_ERR_ARG, "null argument");
    if (t->elem != 8) return fail(SVS_ERR_STATE, "database was not uploaded as u64");
    if (n_queries == 0) return SVS_OK;
    svs_ctx* ctx = t->ctx;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    SVS_TRY(t->ct.ensure(ctx, sizeof(uint64_t) * (size_t)n_queries * t->cols));
    SVS_TRY(t->ans.ensure(ctx, sizeof(uint64_t) * (size_t)n_queries * t->rows));
    HIP_TRY(hipMemcpyAsync(t->ct.ptr, ct, sizeof(uint64_t) * (size_t)n_queries * t->cols,
                           hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_tiptoe_u64, dim3(grid_for(ctx, t->rows, 4)), dim3(256), 0, ctx->stream,
                       (const uint64_t*)t->db, t->rows, t->cols, (const uint64_t*)t->ct.ptr,
                       n_queries, (uint64_t*)t->ans.ptr);
    HIP_TRY(hipGetLastError());
    ctx->launches++;
    HIP_TRY(hipMemcpyAsync(ans, t->ans.ptr, sizeof(uint64_t) * (size_t)n_queries * t->rows,
                           hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SVS_OK;
}
