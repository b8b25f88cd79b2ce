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
    const uint64_t* alh_t;   // 256 x m: rows 0..127 = AL^T, rows 128..255 = H^T
    const int32_t* mp_t32;   // 1024 x m, signed lift of M_plain (when it fits)
    const uint64_t* mp_t64;  // 1024 x m, canonical M_plain (otherwise)
    const uint32_t* s_slice; // ceil(m/64)+1 offsets (in entries) into s_ell
    const uint16_t* s_ell;   // per slice: width x 64 entries, idx | 0x8000 if sign < 0
    const uint32_t* ids;     // m global vector ids
    uint32_t m;
    uint32_t pad;
};

struct Item {
    uint32_t q;          // query slot in the batch
    uint32_t list;       // list id
    uint32_t probe_pos;  // position in the query's probe order
    uint32_t m;          // rows (0 for empty / out of range)
};

}  // namespace svs
