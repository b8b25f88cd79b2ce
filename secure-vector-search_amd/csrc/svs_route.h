// svs_route.h — the client's per-query steps before encryption, on the
// device: quantise_vector (scorer-bntm/src/crypto.rs:105-122) and IVF routing
// probe_route (ivf-index/src/ivf.rs:150-164) over l2_f32_scalar
// (ivf-index/src/distance.rs:30-37).
//
// Routing must reproduce the reference's f32 results bit for bit: each
// distance is a left-to-right f32 sum of rounded squared differences (no
// fused multiply-add, no reassociation), then a correctly rounded sqrt. So
// one thread owns a whole (query, centroid) sum; the parallelism is across
// pairs. The order of the probe set is ascending f32::total_cmp with ties
// in centroid order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "svs_types.h"

namespace svs {

constexpr int RT_Q = 32;   // queries per distance tile
constexpr int RT_C = 32;   // centroids per distance tile
constexpr int RT_J = 32;   // dimensions staged per round

// v_q rows (n = 1024 words, zero padded) from f32 query rows of d <= 1024.
__global__ void __launch_bounds__(256)
k_quantise(const float* __restrict__ x, uint32_t nq, uint32_t d, double q_scale,
           uint64_t* __restrict__ v_q) {
    const uint64_t total = (uint64_t)nq * N;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t q = (uint32_t)(e / N), j = (uint32_t)(e % N);
        v_q[e] = j < d ? quantise_scalar(x[(size_t)q * d + j], q_scale) : 0;
    }
}

// dist[q][c] = l2_f32_scalar(query q, centroid c). Block tile: 32 queries x
// 32 centroids; thread t owns query t % 32 and centroids 4 (t / 32) + 0..3.
// Dimensions beyond d are staged as zeros: (0 - 0)^2 = +0 leaves a partial
// sum unchanged (every partial sum is >= +0 or NaN). The next round's 32
// dimensions are loaded into registers (clamped addresses, unconditional
// loads) while the current round is summed.
__global__ void __launch_bounds__(256)
k_route_dist(const float* __restrict__ x, uint32_t nq, const float* __restrict__ cent,
             uint32_t nc, uint32_t d, float* __restrict__ dist) {
    __shared__ float xs[RT_J][RT_Q + 1];
    __shared__ __align__(16) float cs[RT_J][RT_C + 4];
    constexpr int PER = RT_Q * RT_J / 256;  // staged elements per thread and array
    const uint32_t q0 = blockIdx.x * RT_Q, c0 = blockIdx.y * RT_C;
    const int t = threadIdx.x;
    const int tq = t % RT_Q, tc = 4 * (t / RT_Q);
    float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f, acc3 = 0.f;
    float xn[PER], cn[PER];
    auto fetch = [&](uint32_t j0) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int e = t + 256 * k, r = e / RT_J, jj = e % RT_J;
            const uint32_t j = min(j0 + jj, d - 1);
            xn[k] = x[(size_t)min(q0 + r, nq - 1) * d + j];
            cn[k] = cent[(size_t)min(c0 + r, nc - 1) * d + j];
        }
    };
    fetch(0);
    for (uint32_t j0 = 0; j0 < d; j0 += RT_J) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int e = t + 256 * k, r = e / RT_J, jj = e % RT_J;
            const bool jok = j0 + jj < d;
            xs[jj][r] = (q0 + r < nq && jok) ? xn[k] : 0.f;
            cs[jj][r] = (c0 + r < nc && jok) ? cn[k] : 0.f;
        }
        __syncthreads();
        if (j0 + RT_J < d) fetch(j0 + RT_J);
#pragma unroll 8
        for (int jj = 0; jj < RT_J; ++jj) {
            const float xv = xs[jj][tq];
            const float4 cv = *reinterpret_cast<const float4*>(&cs[jj][tc]);
            float df;
            df = __fsub_rn(xv, cv.x);
            acc0 = __fadd_rn(acc0, __fmul_rn(df, df));
            df = __fsub_rn(xv, cv.y);
            acc1 = __fadd_rn(acc1, __fmul_rn(df, df));
            df = __fsub_rn(xv, cv.z);
            acc2 = __fadd_rn(acc2, __fmul_rn(df, df));
            df = __fsub_rn(xv, cv.w);
            acc3 = __fadd_rn(acc3, __fmul_rn(df, df));
        }
        __syncthreads();
    }
    const uint32_t q = q0 + tq;
    if (q >= nq) return;
    float* out = dist + (size_t)q * nc;
    const uint32_t c = c0 + tc;
    if (c + 0 < nc) out[c + 0] = __fsqrt_rn(acc0);
    if (c + 1 < nc) out[c + 1] = __fsqrt_rn(acc1);
    if (c + 2 < nc) out[c + 2] = __fsqrt_rn(acc2);
    if (c + 3 < nc) out[c + 3] = __fsqrt_rn(acc3);
}

// f32::total_cmp as an unsigned order, with every NaN taken as the largest
// value (the sign of a NaN distance depends on how the host's floating point
// propagates NaNs, so a NaN centroid or query sorts last everywhere here).
__device__ __forceinline__ uint32_t f32_total_key(float f) {
    const uint32_t b = __float_as_uint(f);
    if (f != f) return 0xFFFFFFFFu;
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// Per query (one block of 1024 threads): sorts (total_key(dist), centroid)
// for all nc <= RT_SELECT_MAX centroids in LDS (bitonic) and writes the
// first probe_ptr[q+1] - probe_ptr[q] centroids to probes[probe_ptr[q]..],
// each plus list_base (the tenant's first list in a cross-tenant table).
// Unused slots of the padded sort hold all-ones keys, which sort last.
__global__ void __launch_bounds__(1024)
k_route_select(const float* __restrict__ dist, uint32_t nc, const uint32_t* __restrict__ probe_ptr,
               uint32_t list_base, uint32_t* __restrict__ probes) {
    extern __shared__ uint64_t keys[];
    const uint32_t q = blockIdx.x;
    const uint32_t np = probe_ptr[q + 1] - probe_ptr[q];
    if (np == 0) return;  // uniform per block
    uint32_t p2 = 1;
    while (p2 < nc) p2 <<= 1;
    const float* dq = dist + (size_t)q * nc;
    for (uint32_t i = threadIdx.x; i < p2; i += blockDim.x)
        keys[i] = i < nc ? ((uint64_t)f32_total_key(dq[i]) << 32) | i : ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= p2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < p2; i += blockDim.x) {
                const uint32_t o = i ^ j;
                if (o > i) {
                    const uint64_t a = keys[i], b = keys[o];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        keys[i] = b;
                        keys[o] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    uint32_t* out = probes + probe_ptr[q];
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) out[i] = list_base + (uint32_t)keys[i];
}

}  // namespace svs
