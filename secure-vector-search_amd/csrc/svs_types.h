// svs_types.h — device-visible layout of the resident index and of batch
// work items, shared by kernels and host code.
#pragma once

#include "svs_field.h"

namespace svs {

constexpr int WAVE = 64;
constexpr int CHUNK_ROWS = 16;          // rows per wave-chunk in row-major kernels
constexpr int ROW_THREADS = 256;        // threads per block in row kernels
constexpr int SLICE = 64;               // rows per S sliced-ELL slice
constexpr uint16_t S_PAD = 0xFFFF;      // padding entry in S sliced-ELL

struct ListDev {
    const uint64_t* menc;    // m x 1024, row-major (server state)
    const int32_t* mp_t32;   // 1024 x m, signed lift of M_plain (when it fits)
    const uint64_t* mp_t64;  // 1024 x m, canonical M_plain (otherwise)
    const uint32_t* s_slice; // ceil(m/64)+1 offsets (in entries) into s_ell
    const uint16_t* s_ell;   // per slice: width x 64 entries, idx | 0x8000 if sign < 0
    const uint32_t* ids;     // m global vector ids
    const uint64_t* rowcorr; // m: c * sum_k M_enc[row][k] mod p, c = 128 * 0x0101010101010101
    const uint64_t* alh;     // m x 256, row-major [AL | H] (the dense decode terms)
    const uint64_t* alh_rowcorr;  // m: c * sum_k alh[row][k] mod p
    // MFMA decode (null when M_plain does not fit the 24-bit planes):
    const int8_t* mp8;       // m x 3 x 1024: byte planes s_a of M_plain + 2^23, each ^ 0x80
    const int8_t* s8;        // m x 1024: S as dense int8 (+-1 / 0)
    const uint64_t* s_rowterm;  // m: 128 * 0x0101010101010101 * sum_k S[row][k] mod p
    uint32_t m;
    uint32_t mp_cols;        // M_plain is zero in columns >= mp_cols (a multiple of 64)
};

// M_plain = sum_a 256^a s_a + MP8_OFFSET for the planes of mp8 (|M| < 2^23):
// 128 * (1 + 256 + 65536) - 2^23.
constexpr int64_t MP8_OFFSET = 32896;
// Per query, MP8_OFFSET * sum_{k < 64 b} T[k] (mod p) for b = 0..MT_BLOCKS:
// the offset term of M T over a list's nonzero plaintext columns.
constexpr int MT_BLOCKS = 1024 / 64;
constexpr int64_t MP8_LIMIT = 1 << 23;

// One MFMA work tile: MT_ROWS rows of a list times up to MT_QG items
// (queries) that probe it; the items are gitems[g0 .. g0 + gn).
constexpr int MT_RT = 4;                 // 16-row MFMA tiles per wave
constexpr int MT_ROWS = 16 * MT_RT;      // rows per work tile
constexpr int MT_QG = 4;                 // queries per work tile
struct MTile {
    uint32_t list;
    uint32_t row0;
    uint32_t g0;
    uint32_t gn;
};

// Block-cooperative MFMA tiles: PB_WAVES waves x PB_RT 16-row tiles each
// (PB_ROWS rows) times up to PB_QG queries of one list.
constexpr int PB_WAVES = 4;
constexpr int PB_RT = 4;
constexpr int PB_ROWS = 16 * PB_RT * PB_WAVES;  // 256
constexpr int PB_QG = 8;
constexpr int PB_PF = 8;   // K'-steps of A rows and query words loaded ahead

// c = 128 * sum_{a<8} 2^(8a) reduced mod p: the offset between a u64 and its
// signed-byte-limb value (each byte b read as b - 128).
constexpr uint64_t LIMB_OFFSET_C = 0x8080808080808080ull % 0x1FFFFFFFFFFFFFFFull;

// Device batch planning (k_plan): a sort key packs (list, item) into 32
// bits, so a batch has at most PLAN_MAX_ITEMS items and an index fewer than
// PLAN_MAX_LISTS lists.
constexpr uint32_t PLAN_ITEM_BITS = 14;
constexpr uint32_t PLAN_MAX_ITEMS = 1u << PLAN_ITEM_BITS;
constexpr uint32_t PLAN_MAX_LISTS = 1u << (32 - PLAN_ITEM_BITS);
// Header words of a plan in device memory.
// PH_NU: the staged product's work units (k_plan_units).
enum : uint32_t { PH_NT = 0, PH_NVT = 1, PH_OVERFLOW = 2, PH_NCANDS = 3, PH_NU = 4, PH_WORDS = 8 };
// Device routing sorts a query's centroid distances in LDS (64 KiB).
constexpr int RT_SELECT_MAX = 8192;

struct Item {
    uint32_t q;          // query slot in the batch
    uint32_t list;       // list id
    uint32_t probe_pos;  // position in the query's probe order
    uint32_t m;          // rows (0 for empty / out of range)
};

}  // namespace svs
