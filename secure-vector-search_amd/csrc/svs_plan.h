// svs_plan.h — the batch plan built on the device.
//
// A batch of (query, list) items is described to the pipeline kernels by
// prefix arrays (rows and wave-chunks per item, top-k chunks per query) and
// by MFMA work tiles over the items grouped by list. The host builds the same
// arrays in svs::make_plan; k_plan builds them on the device from the probe
// lists, so that routing, planning and the whole pipeline run back to back on
// one stream with no host round trip. The arrays match the host plan's:
// items in probe order, gitems = the items with rows grouped by list (the
// host orders each list's items by item; the device's counting sort leaves
// them in no set order), tiles list-major, then row block, then query group.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "svs_types.h"

namespace svs {

constexpr int PLAN_THREADS = 1024;
constexpr uint32_t PLAN_ITEM_MASK = PLAN_MAX_ITEMS - 1;

// Exclusive scan of a[0..n) in place (u32 values, u64 running sums) by the
// whole block; a[n] = total. Returns the total to every thread. Each thread
// sums a contiguous segment; wave prefix sums by shuffles, then one wave
// scans the wave totals (red: at least 2 x the block's wave count words).
__device__ uint64_t plan_scan(uint32_t* a, uint32_t n, uint64_t* red) {
    const uint32_t t = threadIdx.x, T = blockDim.x;
    const uint32_t lane = t % WAVE, w = t / WAVE, nw = T / WAVE;
    const uint32_t seg = (n + T - 1) / T;
    const uint32_t b = min(n, t * seg), e = min(n, b + seg);
    uint64_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += a[i];
    uint64_t x = s;  // inclusive prefix within the wave
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const uint64_t y = __shfl_up(x, off, WAVE);
        if ((int)lane >= off) x += y;
    }
    if (lane == WAVE - 1) red[w] = x;
    __syncthreads();
    if (w == 0) {
        uint64_t v = lane < nw ? red[lane] : 0;
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) {
            const uint64_t y = __shfl_up(v, off, WAVE);
            if ((int)lane >= off) v += y;
        }
        if (lane < nw) red[nw + lane] = v;  // inclusive over the waves
    }
    __syncthreads();
    const uint64_t total = red[2 * nw - 1];
    uint64_t run = (w ? red[nw + w - 1] : 0) + x - s;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t v = a[i];
        a[i] = (uint32_t)run;
        run += v;
    }
    if (t == 0) a[n] = (uint32_t)total;
    __syncthreads();
    return total;
}

// Ascending bitonic sort of keys[0..p2) in LDS, p2 a power of two, by a
// block whose thread t handles positions t, t + blockDim, ... . A stage
// whose partner distance j is below WAVE pairs positions of one aligned
// 64-chunk, and every chunk belongs to one wave, so a stage that follows
// such a stage needs only the wave's own LDS ordering (a wave-scope fence);
// the block barrier is kept before and after stages that pair across waves.
__device__ void plan_sort(uint32_t* keys, uint32_t p2) {
    for (uint32_t k = 2; k <= p2; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < p2; i += blockDim.x) {
                const uint32_t o = i ^ j;
                if (o > i) {
                    const uint32_t a = keys[i], b = keys[o];
                    if ((a > b) == ((i & k) == 0)) {
                        keys[i] = b;
                        keys[o] = a;
                    }
                }
            }
            // the next stage: j / 2, or (k * 2) / 2 = k when j == 1
            const uint32_t next_j = j > 1 ? j >> 1 : k;
            if (j < WAVE && next_j < WAVE) {
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
            } else {
                __syncthreads();
            }
        }
    __syncthreads();  // every wave reads the sorted keys next
}

// End of the list group that starts at g: the first position after g whose
// key has another list (keys sorted, positions < ng).
__device__ __forceinline__ uint32_t plan_group_end(const uint32_t* keys, uint32_t g, uint32_t ng) {
    const uint32_t list = keys[g] >> PLAN_ITEM_BITS;
    uint32_t lo = g + 1, hi = ng;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((keys[mid] >> PLAN_ITEM_BITS) == list) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Tiles of tile_rows rows x up to tile_qg items over the sorted keys
// keys[0..ng) (list << PLAN_ITEM_BITS | item): per list group, row blocks
// outer, query groups inner. gitems[g] = item of keys[g]. tmp (LDS, ng + 1
// words) takes the tile offsets, so that each tile's binary search for its
// group stays in LDS. Returns the tile count; writes no tile past cap.
__device__ uint64_t plan_tiles(const uint32_t* keys, uint32_t ng, const Item* items,
                               uint32_t tile_rows, uint32_t tile_qg, uint32_t* tmp,
                               uint32_t* gitems, MTile* tiles, uint32_t cap, uint64_t* red) {
    for (uint32_t g = threadIdx.x; g < ng; g += blockDim.x) {
        gitems[g] = keys[g] & PLAN_ITEM_MASK;
        const bool start = g == 0 || (keys[g - 1] >> PLAN_ITEM_BITS) != (keys[g] >> PLAN_ITEM_BITS);
        uint32_t cnt = 0;
        if (start) {
            const uint32_t e = plan_group_end(keys, g, ng);
            const uint32_t m = items[keys[g] & PLAN_ITEM_MASK].m;
            cnt = ((m + tile_rows - 1) / tile_rows) * ((e - g + tile_qg - 1) / tile_qg);
        }
        tmp[g] = cnt;
    }
    __syncthreads();
    const uint64_t nt = plan_scan(tmp, ng, red);
    if (nt > cap) return nt;
    for (uint32_t x = threadIdx.x; x < (uint32_t)nt; x += blockDim.x) {
        // the group start s whose tile range holds x (tmp[ng] = nt > x)
        uint32_t lo = 0, hi = ng;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (tmp[mid] <= x) lo = mid;
            else hi = mid;
        }
        const uint32_t s = lo, e = plan_group_end(keys, s, ng);
        const uint32_t nqg = (e - s + tile_qg - 1) / tile_qg;
        const uint32_t local = x - tmp[s];
        const uint32_t rb = local / nqg, qg = local % nqg;
        tiles[x] = MTile{keys[s] >> PLAN_ITEM_BITS, rb * tile_rows, s + qg * tile_qg,
                         min(tile_qg, e - s - qg * tile_qg)};
    }
    __syncthreads();
    return nt;
}

// One block of PLAN_THREADS. Inputs: probes[n_items] (list ids, items in
// probe order), item_ptr[nq + 1] (each query's items), qverify[nq].
// Outputs: items, item_verify, chunk_ptr / cand_ptr [n_items + 1], a copy
// of item_ptr, qchunk_ptr [nq + 1], gitems + tiles, vgitems + vtiles (items
// to verify), and hdr (tile counts, candidate count, overflow). tmp:
// n_items words of global scratch. A plan over the capacities (candidate
// rows, top-k chunks, tiles) is emptied (no rows, no tiles) and flagged in
// hdr[PH_OVERFLOW].
__global__ void __launch_bounds__(PLAN_THREADS)
k_plan(const ListDev* __restrict__ lists, uint32_t n_lists, const uint32_t* __restrict__ probes,
       const uint32_t* __restrict__ item_ptr, const uint8_t* __restrict__ qverify, uint32_t nq,
       uint32_t n_items, uint32_t tile_rows, uint32_t tile_qg, uint32_t cap_cands,
       uint32_t cap_tchunks, uint32_t cap_tiles, Item* __restrict__ items,
       uint8_t* __restrict__ item_verify, uint32_t* __restrict__ chunk_ptr,
       uint32_t* __restrict__ cand_ptr, uint32_t* __restrict__ item_ptr_out,
       uint32_t* __restrict__ qchunk_ptr, uint32_t* __restrict__ tmp,
       uint32_t* __restrict__ gitems, MTile* __restrict__ tiles, uint32_t* __restrict__ vgitems,
       MTile* __restrict__ vtiles, uint32_t* __restrict__ hdr) {
    __shared__ uint32_t keys[PLAN_MAX_ITEMS];
    // per item its query, then the tile offsets and verify flags
    __shared__ uint32_t scratch[PLAN_MAX_ITEMS + 1];
    __shared__ uint64_t red[PLAN_THREADS];
    const uint32_t T = blockDim.x;
    for (uint32_t q = threadIdx.x; q <= nq; q += T) item_ptr_out[q] = item_ptr[q];
    for (uint32_t q = threadIdx.x; q < nq; q += T)
        for (uint32_t i = item_ptr[q]; i < item_ptr[q + 1]; ++i) scratch[i] = q;
    __syncthreads();
    // items in probe order; rows and chunks per item
    for (uint32_t i = threadIdx.x; i < n_items; i += T) {
        const uint32_t q = scratch[i], c = probes[i];
        const bool ok = c < n_lists;
        const uint32_t m = ok ? lists[c].m : 0;
        items[i] = Item{q, ok ? c : 0, i - item_ptr[q], m};
        item_verify[i] = qverify[q];
        cand_ptr[i] = m;
        chunk_ptr[i] = (m + CHUNK_ROWS - 1) / CHUNK_ROWS;
        keys[i] = m ? (c << PLAN_ITEM_BITS) | i : 0xFFFFFFFFu;
    }
    __syncthreads();
    const uint64_t n_cands = plan_scan(cand_ptr, n_items, red);
    plan_scan(chunk_ptr, n_items, red);
    bool over = n_cands > cap_cands;
    if (!over) {
        for (uint32_t q = threadIdx.x; q < nq; q += T) {
            const uint32_t nc = cand_ptr[item_ptr[q + 1]] - cand_ptr[item_ptr[q]];
            qchunk_ptr[q] = (nc + TOPK_WCHUNK - 1) / TOPK_WCHUNK;
        }
        __syncthreads();
        over = plan_scan(qchunk_ptr, nq, red) > cap_tchunks;
    }
    uint64_t nt = 0, nvt = 0;
    if (!over) {
        uint32_t ng;
        if (n_lists <= PLAN_MAX_ITEMS) {
            // Counting sort by list: per-list counts (LDS atomics in
            // scratch, free until the tiles), their scan, then a scatter
            // with per-list cursors. Items of one list come out in no set
            // order; a list's tiles then group its queries differently
            // from run to run, which changes no result (every (query, row)
            // product is exact and written at the item's own place).
            for (uint32_t l = threadIdx.x; l <= n_lists; l += T) scratch[l] = 0;
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < n_items; i += T) {
                const uint32_t k = keys[i];
                tmp[i] = k;
                if (k != 0xFFFFFFFFu) atomicAdd(&scratch[k >> PLAN_ITEM_BITS], 1u);
            }
            __syncthreads();
            ng = (uint32_t)plan_scan(scratch, n_lists, red);
            for (uint32_t i = threadIdx.x; i < n_items; i += T) {
                const uint32_t k = tmp[i];
                if (k != 0xFFFFFFFFu) keys[atomicAdd(&scratch[k >> PLAN_ITEM_BITS], 1u)] = k;
            }
            __syncthreads();
        } else {
            uint32_t p2 = 1;
            while (p2 < n_items) p2 <<= 1;
            for (uint32_t i = n_items + threadIdx.x; i < p2; i += T) keys[i] = 0xFFFFFFFFu;
            __syncthreads();
            plan_sort(keys, p2);
            // items with rows: the keys below the all-ones padding
            uint32_t lo = 0, hi = n_items;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (keys[mid] != 0xFFFFFFFFu) lo = mid + 1;
                else hi = mid;
            }
            ng = lo;
        }
        nt = plan_tiles(keys, ng, items, tile_rows, tile_qg, scratch, gitems, tiles, cap_tiles,
                        red);
        over = nt > cap_tiles;
        if (!over) {
            // the verified items' keys, in order (a stable compaction)
            uint32_t* vflag = scratch;
            for (uint32_t g = threadIdx.x; g < ng; g += T)
                vflag[g] = item_verify[keys[g] & PLAN_ITEM_MASK] ? 1u : 0u;
            __syncthreads();
            const uint32_t nvg = (uint32_t)plan_scan(vflag, ng, red);
            for (uint32_t g = threadIdx.x; g < ng; g += T)
                if (vflag[g + 1] != vflag[g]) tmp[vflag[g]] = keys[g];
            __syncthreads();
            for (uint32_t g = threadIdx.x; g < nvg; g += T) keys[g] = tmp[g];
            __syncthreads();
            nvt = plan_tiles(keys, nvg, items, tile_rows, tile_qg, scratch, vgitems, vtiles,
                             cap_tiles, red);
            over = nvt > cap_tiles;
        }
    }
    if (over) {
        // an empty plan: every kernel after this one finds no work
        for (uint32_t i = threadIdx.x; i <= n_items; i += T) {
            if (i < n_items) items[i].m = 0;
            cand_ptr[i] = 0;
            chunk_ptr[i] = 0;
        }
        for (uint32_t q = threadIdx.x; q <= nq; q += T) qchunk_ptr[q] = 0;
        nt = nvt = 0;
    }
    if (threadIdx.x == 0) {
        hdr[PH_NT] = (uint32_t)nt;
        hdr[PH_NVT] = (uint32_t)nvt;
        hdr[PH_OVERFLOW] = over ? 1u : 0u;
        hdr[PH_NCANDS] = over ? 0u : (uint32_t)n_cands;
    }
}

// Work units of the staged product kernel over a plan's tiles (tiles
// list-major, then row block, then 16-query group). A row block's groups
// go out in entries of up to four consecutive groups (full quads, then the
// rest); an entry of c groups covers the row block in units of
//   c >= 3: 16 rows x the c groups (one wave per group; c = 3 leaves one
//           wave idle),
//   c == 2: 32 rows x the 2 groups (waves: 2 row blocks of 16 x 2 groups),
//   c == 1: 64 rows x the group (one wave per 16 rows),
// so that a block's four waves always share what one staging of a step
// serves (each unit stages 40 KB per 64-element step). units[] =
// (first tile << 8 | c << 4 | k), k the unit's index within its entry;
// hdr[PH_NU] = the unit count. One block of PLAN_THREADS; tmp: nt + 1
// words of scratch.
__global__ void __launch_bounds__(PLAN_THREADS)
k_plan_units(const ListDev* __restrict__ lists, const MTile* __restrict__ tiles,
             uint32_t* __restrict__ hdr, uint32_t* __restrict__ units,
             uint32_t* __restrict__ tmp) {
    __shared__ uint64_t red[PLAN_THREADS];
    __shared__ uint32_t wmax[PLAN_THREADS / WAVE];
    const uint32_t nt = hdr[PH_NT];
    const uint32_t t = threadIdx.x, T = blockDim.x;
    auto same = [&](uint32_t i, uint32_t j) {
        return tiles[i].list == tiles[j].list && tiles[i].row0 == tiles[j].row0;
    };
    // each thread walks a contiguous segment; the row block in progress at
    // its start is the last row-block start of the segments before it
    const uint32_t seg = (nt + T - 1) / T;
    const uint32_t b = min(nt, t * seg), e = min(nt, b + seg);
    uint32_t last = 0;  // tile 0 starts a row block
    for (uint32_t i = b; i < e; ++i)
        if (i == 0 || !same(i - 1, i)) last = i;
    // exclusive max over the earlier threads' values (later segments hold
    // later starts)
    const uint32_t lane = t % WAVE, w = t / WAVE;
    uint32_t x = last;  // inclusive max within the wave
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, WAVE);
        if ((int)lane >= off) x = max(x, y);
    }
    if (lane == WAVE - 1) wmax[w] = x;
    __syncthreads();
    const uint32_t below = __shfl_up(x, 1, WAVE);
    uint32_t rbs = lane ? below : 0u;
    for (uint32_t k = 0; k < w; ++k) rbs = max(rbs, wmax[k]);
    // per tile: the units of the entry it starts (0 if it starts none)
    uint32_t* cnt = tmp;
    for (uint32_t i = b; i < e; ++i) {
        if (i == 0 || !same(i - 1, i)) rbs = i;
        uint32_t n = 0;
        if ((i - rbs) % 4 == 0) {
            uint32_t c = 1;
            while (c < 4 && i + c < nt && same(i, i + c)) ++c;
            const MTile tl = tiles[i];
            const uint32_t rows = min(256u, lists[tl.list].m - tl.row0);
            const uint32_t nsb = (rows + 15) / 16;
            n = c >= 3 ? nsb : c == 2 ? (nsb + 1) / 2 : (nsb + 3) / 4;
        }
        cnt[i] = n;
    }
    __syncthreads();
    const uint32_t nu = (uint32_t)plan_scan(cnt, nt, red);
    for (uint32_t i = t; i < nt; i += T) {
        const uint32_t n = cnt[i + 1] - cnt[i];
        if (!n) continue;
        uint32_t c = 1;
        while (c < 4 && i + c < nt && same(i, i + c)) ++c;
        for (uint32_t k = 0; k < n; ++k) units[cnt[i] + k] = i << 8 | c << 4 | k;
    }
    if (t == 0) hdr[PH_NU] = nu;
}

}  // namespace svs
