// svs_bfv.hip — Tiptoe's BFV hint compression on the device
// (scorer-tiptoe/src/bfv/compress.rs, which runs it on the CPU through the
// `fhe` crate):
//   * decompose_hint (:146-190): each 64-bit entry of the SimplePIR hint
//     H (r x n_lwe) split into its top 8 four-bit limbs; for limb b, row
//     chunk c (N = 2048 rows) and column j, the polynomial whose coefficient
//     i is limb b of H[c N + i][j], held in the NTT domain, resident in HBM;
//   * apply_hint (:241-266): out[b][c] = sum_j ct[j] pt[b][c][j] for the
//     query token's n_lwe ciphertexts, a pointwise multiply-accumulate in
//     the NTT domain (the plaintext x ciphertext dot_product_scalar).
//
// BFV at tiptoe_text_params' sizes (bfv/mod.rs:50-70): N = 2048, t = 65537,
// one 38-bit ciphertext modulus q: the largest prime below 2^38 that is
// 1 mod 2N. The transform is the negacyclic NTT with psi = g^((q-1)/2N), g
// the smallest primitive root, natural order in and bit-reversed order out
// (twiddle i = psi^bitrev(i)); the client must encrypt its token in the
// same domain. The reference's fhe crate is not vendored, so these are this
// library's choices, stated here and in the header.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/svs_gpu.h"
#include "svs_internal.h"

using namespace svs;

namespace {

constexpr int TN = 2048, TLOG = 11;
constexpr int TLIMBS = 8, TMAXLIMBS = 16;  // NUM_LIMBS_64, MAX_LIMBS_64 (compress.rs:58-62)
constexpr int BW = 64;                      // wave size

typedef unsigned __int128 u128;

// ---------------------------------------------------------------- host math
uint64_t h_powmod(uint64_t a, uint64_t e, uint64_t m) {
    u128 r = 1, b = a % m;
    while (e) {
        if (e & 1) r = r * b % m;
        b = b * b % m;
        e >>= 1;
    }
    return (uint64_t)r;
}

bool h_is_prime(uint64_t n) {
    static const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2) return false;
    for (uint64_t b : bases) {
        if (n == b) return true;
        if (n % b == 0) return false;
    }
    uint64_t d = n - 1;
    int r = 0;
    while ((d & 1) == 0) d >>= 1, ++r;
    for (uint64_t b : bases) {
        uint64_t x = h_powmod(b, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int k = 1; k < r && comp; ++k) {
            x = (uint64_t)((u128)x * x % n);
            comp = x != n - 1;
        }
        if (comp) return false;
    }
    return true;
}

struct BfvTables {
    uint64_t q = 0, psi = 0;
    std::vector<uint64_t> w, ws;  // psi^bitrev(i) and its Shoup constant floor(w 2^64 / q)
};

const BfvTables& tables() {
    static BfvTables T;
    static std::once_flag once;
    std::call_once(once, [] {
        const uint64_t m2 = 2 * TN;
        uint64_t q = (((1ull << 38) - 2) / m2) * m2 + 1;
        while (!h_is_prime(q)) q -= m2;
        std::vector<uint64_t> fac;
        uint64_t rest = q - 1;
        for (uint64_t f = 2; f * f <= rest; ++f)
            if (rest % f == 0) {
                fac.push_back(f);
                while (rest % f == 0) rest /= f;
            }
        if (rest > 1) fac.push_back(rest);
        uint64_t g = 2;
        for (;; ++g) {
            bool ok = true;
            for (uint64_t f : fac) ok = ok && h_powmod(g, (q - 1) / f, q) != 1;
            if (ok) break;
        }
        T.q = q;
        T.psi = h_powmod(g, (q - 1) / m2, q);
        T.w.resize(TN);
        T.ws.resize(TN);
        for (uint32_t i = 0; i < TN; ++i) {
            uint32_t e = 0;
            for (int b = 0; b < TLOG; ++b) e |= ((i >> b) & 1u) << (TLOG - 1 - b);
            T.w[i] = h_powmod(T.psi, e, q);
            T.ws[i] = (uint64_t)(((u128)T.w[i] << 64) / q);
        }
    });
    return T;
}

// -------------------------------------------------------------- device math
// a * w mod q for a < q, w < q < 2^63, ws = floor(w 2^64 / q) (Shoup).
__device__ __forceinline__ uint64_t mul_shoup(uint64_t a, uint64_t w, uint64_t ws, uint64_t q) {
    const uint64_t r = a * w - __umul64hi(a, ws) * q;  // in [0, 2q)
    return r >= q ? r - q : r;
}

__device__ __forceinline__ uint64_t addq(uint64_t a, uint64_t b, uint64_t q) {
    const uint64_t s = a + b;
    return s >= q ? s - q : s;
}

__device__ __forceinline__ uint64_t subq(uint64_t a, uint64_t b, uint64_t q) {
    return a >= b ? a - b : a + q - b;
}

// decompose_hint: one workgroup per (chunk, column); the column's 2048 rows
// of the chunk are read once (8 per thread), then for each of the 8 limbs
// the limb polynomial is transformed in LDS (11 Cooley-Tukey stages of 1024
// butterflies, 4 per thread) and written out, 16 KB contiguous.
// pt[((b * chunks + c) * n_lwe + j) * N + slot].
__global__ void __launch_bounds__(256)
k_hint_ntt(const uint64_t* __restrict__ hint, uint32_t r, uint32_t n_lwe, uint32_t chunks,
           const uint64_t* __restrict__ w, const uint64_t* __restrict__ ws, uint64_t q,
           uint64_t* __restrict__ pt) {
    __shared__ uint64_t poly[TN];
    const uint32_t c = blockIdx.x / n_lwe, j = blockIdx.x % n_lwe;
    const uint32_t t = threadIdx.x;
    uint64_t h[TN / 256];
#pragma unroll
    for (int e = 0; e < TN / 256; ++e) {
        const uint32_t row = c * TN + t + 256 * e;
        h[e] = row < r ? hint[(size_t)row * n_lwe + j] : 0;
    }
    for (int b = 0; b < TLIMBS; ++b) {
        const int shift = 4 * (TMAXLIMBS - b - 1);
#pragma unroll
        for (int e = 0; e < TN / 256; ++e) poly[t + 256 * e] = (h[e] >> shift) & 0xFu;
        __syncthreads();
        uint32_t half = TN;
        for (uint32_t m = 1; m < TN; m <<= 1) {
            half >>= 1;  // butterfly span
#pragma unroll
            for (int e = 0; e < TN / 512; ++e) {
                const uint32_t k = t + 256 * e;        // butterfly 0..1023
                const uint32_t grp = k / half, off = k % half;
                const uint32_t i0 = 2 * grp * half + off, i1 = i0 + half;
                const uint64_t u = poly[i0];
                const uint64_t v = mul_shoup(poly[i1], w[m + grp], ws[m + grp], q);
                poly[i0] = addq(u, v, q);
                poly[i1] = subq(u, v, q);
            }
            __syncthreads();
        }
        uint64_t* out = pt + (((size_t)b * chunks + c) * n_lwe + j) * TN;
#pragma unroll
        for (int e = 0; e < TN / 256; ++e) out[t + 256 * e] = poly[t + 256 * e];
        __syncthreads();
    }
}

// apply_hint for QB query tokens: slot x = 64 xt + lane; wave v of the
// block takes the (limb, chunk) pairs 16 g + 4 v .. + 3; split s takes the
// columns [s len, (s + 1) len). Each lane streams its 4 plaintext slots and
// the tokens' two ciphertext slots along j with 128-bit lazy sums (at most
// 2^87 for 2048 products below 2^76), reduced once at the end into
// part[s][qi][bc][k][x]. Every plaintext slot is read once per launch.
template <int QB>
__global__ void __launch_bounds__(256)
k_apply_hint(const uint64_t* __restrict__ pt, uint32_t n_bc, uint32_t n_lwe, uint32_t len,
             const uint64_t* __restrict__ ct, uint32_t nq, uint64_t q,
             uint64_t* __restrict__ part) {
    const uint32_t lane = threadIdx.x % BW, wv = threadIdx.x / BW;
    const uint32_t xt = blockIdx.x, g = blockIdx.y, s = blockIdx.z;
    const uint32_t x = xt * BW + lane;
    const uint32_t bc0 = 16 * g + 4 * wv;
    const uint32_t j0 = s * len, j1 = min(n_lwe, j0 + len);
    u128 acc[4][QB][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int qi = 0; qi < QB; ++qi) acc[a][qi][0] = acc[a][qi][1] = 0;
    const bool live[4] = {bc0 < n_bc, bc0 + 1 < n_bc, bc0 + 2 < n_bc, bc0 + 3 < n_bc};
    for (uint32_t j = j0; j < j1; ++j) {
        uint64_t p[4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
            p[a] = live[a] ? pt[((size_t)(bc0 + a) * n_lwe + j) * TN + x] : 0;
#pragma unroll
        for (int qi = 0; qi < QB; ++qi) {
            if ((uint32_t)qi >= nq) break;
            const uint64_t* cq = ct + ((size_t)qi * n_lwe + j) * 2 * TN;
            const uint64_t c0 = cq[x], c1 = cq[TN + x];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                acc[a][qi][0] += (u128)c0 * p[a];
                acc[a][qi][1] += (u128)c1 * p[a];
            }
        }
    }
#pragma unroll
    for (int qi = 0; qi < QB; ++qi) {
        if ((uint32_t)qi >= nq) break;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            if (!live[a]) continue;
#pragma unroll
            for (int k = 0; k < 2; ++k)
                part[((((size_t)s * nq + qi) * n_bc + bc0 + a) * 2 + k) * TN + x] =
                    (uint64_t)(acc[a][qi][k] % q);
        }
    }
}

// out[qi][bc][k][x] = sum over the splits, mod q.
__global__ void __launch_bounds__(256)
k_apply_sum(const uint64_t* __restrict__ part, uint32_t splits, size_t per_split, uint64_t q,
            uint64_t* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < per_split;
         i += (size_t)gridDim.x * blockDim.x) {
        uint64_t v = 0;
        for (uint32_t s = 0; s < splits; ++s) v = addq(v, part[s * per_split + i], q);
        out[i] = v;
    }
}

}  // namespace

struct svs_tiptoe_hint {
    svs_ctx* ctx = nullptr;
    uint32_t r = 0, n_lwe = 0, chunks = 0;
    uint64_t q = 0;
    uint64_t* pt = nullptr;
    size_t pt_bytes = 0;
    uint64_t *w = nullptr, *ws = nullptr;
    DevBuf ct, part, out;
};

extern "C" int svs_tiptoe_hint_create(svs_ctx* ctx, uint32_t r, uint32_t n_lwe,
                                      const uint64_t* hint, svs_tiptoe_hint** out) {
    if (!ctx || !out || (r && n_lwe && !hint)) return fail(SVS_ERR_ARG, "null argument");
    if (r == 0 || n_lwe == 0) return fail(SVS_ERR_ARG, "empty hint (r = %u, n_lwe = %u)", r, n_lwe);
    const BfvTables& T = tables();
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    auto* h = new svs_tiptoe_hint();
    h->ctx = ctx;
    h->r = r;
    h->n_lwe = n_lwe;
    h->chunks = (r + TN - 1) / TN;
    h->q = T.q;
    h->pt_bytes = sizeof(uint64_t) * (size_t)TLIMBS * h->chunks * n_lwe * TN;
    auto undo = [&](int rc) {
        if (h->pt) dev_free(ctx, h->pt, h->pt_bytes);
        if (h->w) dev_free(ctx, h->w, sizeof(uint64_t) * TN);
        if (h->ws) dev_free(ctx, h->ws, sizeof(uint64_t) * TN);
        delete h;
        return rc;
    };
    int rc = dev_alloc(ctx, (void**)&h->pt, h->pt_bytes);
    if (rc == SVS_OK) rc = dev_alloc(ctx, (void**)&h->w, sizeof(uint64_t) * TN);
    if (rc == SVS_OK) rc = dev_alloc(ctx, (void**)&h->ws, sizeof(uint64_t) * TN);
    if (rc != SVS_OK) return undo(rc);
    // the hint goes through a scratch buffer, released after the transform
    const size_t hint_bytes = sizeof(uint64_t) * (size_t)r * n_lwe;
    void* dh = nullptr;
    if ((rc = dev_alloc(ctx, &dh, hint_bytes)) != SVS_OK) return undo(rc);
    hipError_t e = hipMemcpy(dh, hint, hint_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->w, T.w.data(), sizeof(uint64_t) * TN, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->ws, T.ws.data(), sizeof(uint64_t) * TN, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_hint_ntt, dim3(h->chunks * n_lwe), dim3(256), 0, ctx->stream,
                           (const uint64_t*)dh, r, n_lwe, h->chunks, (const uint64_t*)h->w,
                           (const uint64_t*)h->ws, h->q, h->pt);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        ctx->launches++;
    }
    dev_free(ctx, dh, hint_bytes);
    if (e != hipSuccess) {
        undo(SVS_ERR_HIP);
        return fail(SVS_ERR_HIP, "tiptoe hint transform: %s", hipGetErrorString(e));
    }
    ctx_retain(ctx);
    *out = h;
    return SVS_OK;
}

extern "C" int svs_tiptoe_hint_destroy(svs_tiptoe_hint* h) {
    if (!h) return SVS_OK;
    svs_ctx* ctx = h->ctx;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        hipSetDevice(ctx->device);
        hipStreamSynchronize(ctx->stream);
        dev_free(ctx, h->pt, h->pt_bytes);
        dev_free(ctx, h->w, sizeof(uint64_t) * TN);
        dev_free(ctx, h->ws, sizeof(uint64_t) * TN);
        h->ct.release(ctx);
        h->part.release(ctx);
        h->out.release(ctx);
        delete h;
    }
    ctx_release(ctx);
    return SVS_OK;
}

extern "C" int svs_tiptoe_hint_info(const svs_tiptoe_hint* h, uint64_t* q, uint32_t* chunks,
                                    uint64_t* stored_bytes) {
    if (!h) return fail(SVS_ERR_ARG, "null argument");
    if (q) *q = h->q;
    if (chunks) *chunks = h->chunks;
    if (stored_bytes) *stored_bytes = h->pt_bytes;
    return SVS_OK;
}

static int apply_launch(svs_tiptoe_hint* h, uint32_t n_queries, const uint64_t* ct_dev,
                        uint64_t* out_dev) {
    svs_ctx* ctx = h->ctx;
    const uint32_t n_bc = TLIMBS * h->chunks;
    constexpr uint32_t SPLITS = 4;
    const uint32_t len = (h->n_lwe + SPLITS - 1) / SPLITS;
    const size_t per_q = (size_t)n_bc * 2 * TN;
    constexpr uint32_t QB = 4;
    SVS_TRY(h->part.ensure(ctx, sizeof(uint64_t) * SPLITS * std::min<uint32_t>(n_queries, QB) * per_q));
    const dim3 grid(TN / BW, (n_bc + 15) / 16, SPLITS);
    for (uint32_t q0 = 0; q0 < n_queries; q0 += QB) {
        const uint32_t nq = std::min<uint32_t>(QB, n_queries - q0);
        const uint64_t* ct = ct_dev + (size_t)q0 * h->n_lwe * 2 * TN;
        if (nq == 1)
            hipLaunchKernelGGL(k_apply_hint<1>, grid, dim3(256), 0, ctx->stream, h->pt, n_bc,
                               h->n_lwe, len, ct, nq, h->q, (uint64_t*)h->part.ptr);
        else
            hipLaunchKernelGGL(k_apply_hint<QB>, grid, dim3(256), 0, ctx->stream, h->pt, n_bc,
                               h->n_lwe, len, ct, nq, h->q, (uint64_t*)h->part.ptr);
        HIP_TRY(hipGetLastError());
        const size_t per_split = (size_t)nq * per_q;
        hipLaunchKernelGGL(k_apply_sum, dim3((uint32_t)std::min<size_t>((per_split + 255) / 256,
                                                                        (size_t)ctx->n_cu * 8)),
                           dim3(256), 0, ctx->stream, (const uint64_t*)h->part.ptr, SPLITS,
                           per_split, h->q, out_dev + (size_t)q0 * per_q);
        HIP_TRY(hipGetLastError());
        ctx->launches += 2;
    }
    return SVS_OK;
}

extern "C" int svs_tiptoe_apply_hint_dev(svs_tiptoe_hint* h, uint32_t n_queries,
                                         const uint64_t* ct_dev, uint64_t* out_dev) {
    if (!h || (n_queries && (!ct_dev || !out_dev))) return fail(SVS_ERR_ARG, "null argument");
    if (n_queries == 0) return SVS_OK;
    std::lock_guard<std::mutex> g(h->ctx->mu);
    HIP_TRY(hipSetDevice(h->ctx->device));
    return apply_launch(h, n_queries, ct_dev, out_dev);
}

extern "C" int svs_tiptoe_apply_hint(svs_tiptoe_hint* h, uint32_t n_queries, const uint64_t* ct,
                                     uint64_t* out) {
    if (!h || (n_queries && (!ct || !out))) return fail(SVS_ERR_ARG, "null argument");
    if (n_queries == 0) return SVS_OK;
    const size_t ct_n = (size_t)n_queries * h->n_lwe * 2 * TN;
    for (size_t i = 0; i < ct_n; ++i)
        if (ct[i] >= h->q) return fail(SVS_ERR_ARG, "ct[%zu] is not below q", i);
    svs_ctx* ctx = h->ctx;
    const size_t out_n = (size_t)n_queries * TLIMBS * h->chunks * 2 * TN;
    std::lock_guard<std::mutex> g(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    SVS_TRY(h->ct.ensure(ctx, sizeof(uint64_t) * ct_n));
    SVS_TRY(h->out.ensure(ctx, sizeof(uint64_t) * out_n));
    HIP_TRY(hipMemcpyAsync(h->ct.ptr, ct, sizeof(uint64_t) * ct_n, hipMemcpyHostToDevice,
                           ctx->stream));
    SVS_TRY(apply_launch(h, n_queries, (const uint64_t*)h->ct.ptr, (uint64_t*)h->out.ptr));
    HIP_TRY(hipMemcpyAsync(out, h->out.ptr, sizeof(uint64_t) * out_n, hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SVS_OK;
}
