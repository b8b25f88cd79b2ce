"""ctypes binding of libsvs_gpu.so (include/svs_gpu.h).

The library must have been built in-tree (secure-vector-search_amd/build.py);
there is no fallback: a missing library raises at import of the first call.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsvs_gpu.so")

SVS_OK, SVS_NONE = 0, 1
SVS_ERR_ARG, SVS_ERR_HIP, SVS_ERR_BUDGET, SVS_ERR_STATE, SVS_ERR_RNG = -1, -2, -3, -4, -5
MODE_PLAIN, MODE_BATCHED, MODE_BATCHED_VALU = 0, 1, 2
KERNEL_VALU, KERNEL_MFMA, KERNEL_MFMA_WAVE = 0, 1, 2
N, N1, LAMBDA = 1024, 128, 3
P = (1 << 61) - 1

EXPORTS = [
    "svs_last_error", "svs_ctx_create", "svs_ctx_destroy", "svs_ctx_stats", "svs_index_create",
    "svs_index_destroy", "svs_index_upload_list", "svs_index_build_list",
    "svs_index_download_list", "svs_index_list_rows", "svs_compute_products",
    "svs_compute_products_dev", "svs_decode_dense_terms", "svs_verify", "svs_decode",
    "svs_score_batch", "svs_keygen_subspace", "svs_keygen_derive_h_seed",
    "svs_keygen_domain_separate", "svs_keygen_regenerate_h_s", "svs_quantise", "svs_probe_route",
    "svs_debug_inject_fault", "svs_ctx_profile", "svs_compute_products_batch",
    "svs_emvp_create", "svs_emvp_destroy", "svs_emvp_compute_products",
    "svs_emvp_compute_products_dev", "svs_emvp_decode_scores", "svs_emvp_decode_scores_dev",
    "svs_emvp_score", "svs_encode_query", "svs_tiptoe_hint_create", "svs_tiptoe_hint_destroy",
    "svs_tiptoe_hint_info", "svs_tiptoe_apply_hint", "svs_tiptoe_apply_hint_dev",
    "svs_tiptoe_create",
    "svs_tiptoe_destroy", "svs_tiptoe_stored_bytes", "svs_tiptoe_answer_u32",
    "svs_tiptoe_answer_u32_dev", "svs_tiptoe_answer_u64", "svs_ctx_copy_probe",
    "svs_ctx_set_tile_variant", "svs_ctx_get_tile_variant", "svs_autotune",
    "svs_server_create", "svs_server_destroy", "svs_server_submit", "svs_server_wait",
    "svs_server_stats", "svs_server_run_poisson", "svs_score_batch_dev", "svs_dev_alloc",
    "svs_dev_free", "svs_memcpy", "svs_ctx_read_probe", "svs_server_reserve",
    "svs_ctx_set_planner", "svs_ctx_set_product_split", "svs_index_set_centroids",
    "svs_route_batch", "svs_score_raw_batch",
    "svs_ctx_set_graphs", "svs_topk", "svs_index_load_ivf_cache", "svs_cache_fingerprint64",
    "svs_index_get_centroids", "svs_index_get_subspace", "svs_verify_noalloc",
]
TILE_VARIANTS = 10
ROUTE_MAX_LISTS = 8192  # device routing sorts a query's distances in LDS
PLAN_MAX_ITEMS = 16384  # items of a batch planned on the device


class GpuError(RuntimeError):
    """Device / argument error from the backend (reference: GpuError,
    scorer-bntm/src/gpu/mod.rs:97-144)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class BudgetTooSmall(GpuError):
    pass


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("budget_bytes", "resident_bytes", "peak_resident_bytes",
                                          "hbm_total_bytes", "hbm_free_bytes", "uploads",
                                          "launches", "graph_launches", "graph_captures")]


class Profile(C.Structure):
    _fields_ = [("ms", C.c_double * 8), ("calls", C.c_uint64 * 8), ("gemv_bytes", C.c_uint64),
                ("gemv_rows", C.c_uint64)]


STAGES = ("encode", "product", "verify", "decode", "topk")


class QueryBatch(C.Structure):
    _fields_ = [
        ("n_queries", C.c_uint32),
        ("v_q", C.c_void_p), ("q_seeds", C.c_void_p), ("probe_ptr", C.c_void_p),
        ("probe_lists", C.c_void_p), ("verify", C.c_void_p), ("k", C.c_uint32),
        ("out_ids", C.c_void_p), ("out_scores", C.c_void_p), ("out_count", C.c_void_p),
        ("out_status", C.c_void_p),
    ]


class RawBatch(C.Structure):
    _fields_ = [
        ("n_queries", C.c_uint32), ("queries", C.c_void_p), ("nprobe", C.c_void_p),
        ("nprobe_all", C.c_uint32), ("quantisation_q", C.c_uint64), ("q_seeds", C.c_void_p),
        ("verify", C.c_void_p), ("k", C.c_uint32), ("out_ids", C.c_void_p),
        ("out_scores", C.c_void_p), ("out_count", C.c_void_p), ("out_status", C.c_void_p),
        ("out_probes", C.c_void_p),
    ]


class ServerConfig(C.Structure):
    _fields_ = [("max_batch", C.c_uint32), ("slots", C.c_uint32), ("slot_bytes", C.c_uint64),
                ("mode", C.c_int), ("cross_tenant", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("tenant", C.c_uint32), ("v_q", C.c_void_p), ("q_seed", C.c_void_p),
                ("probes", C.c_void_p), ("n_probes", C.c_uint32), ("verify", C.c_uint8),
                ("k", C.c_uint32), ("arrival_s", C.c_double), ("query", C.c_void_p),
                ("nprobe", C.c_uint32), ("quantisation_q", C.c_uint64)]


class Result(C.Structure):
    _fields_ = [("ids", C.c_void_p), ("scores", C.c_void_p), ("count", C.c_uint32),
                ("status", C.c_int32), ("rc", C.c_int32), ("batch_size", C.c_uint32),
                ("latency_ms", C.c_double)]


class ServerCounters(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("batches", C.c_uint64),
                ("batched_queries", C.c_uint64), ("max_batch", C.c_uint32),
                ("max_in_flight", C.c_uint32), ("pending", C.c_uint32),
                ("graph_launches", C.c_uint64), ("graph_captures", C.c_uint64)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with secure-vector-search_amd/build.py")
        L = C.CDLL(LIB_PATH)
        vp, u32, u64, sz, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_size_t, C.c_int
        sig = {
            "svs_last_error": ([C.c_char_p, sz], i32),
            "svs_ctx_create": ([i32, u64, C.POINTER(vp)], i32),
            "svs_ctx_destroy": ([vp], i32),
            "svs_ctx_stats": ([vp, C.POINTER(Stats)], i32),
            "svs_index_create": ([vp, u32, vp, C.POINTER(vp)], i32),
            "svs_index_destroy": ([vp], i32),
            "svs_index_upload_list": ([vp, u32, u32, vp, vp, vp, vp, vp, vp, vp, vp], i32),
            "svs_index_build_list": ([vp, u32, u32, vp, vp, vp], i32),
            "svs_index_download_list": ([vp, u32, vp, vp, vp], i32),
            "svs_index_list_rows": ([vp, u32], u32),
            "svs_compute_products": ([vp, u32, vp, vp], i32),
            "svs_compute_products_dev": ([vp, u32, vp, vp], i32),
            "svs_decode_dense_terms": ([vp, u32, vp, vp, vp], i32),
            "svs_verify": ([vp, u32, vp, vp, vp, u32, C.POINTER(C.c_int32)], i32),
            "svs_verify_noalloc": ([vp, u32, vp, vp, vp, u32, C.POINTER(C.c_int32)], i32),
            "svs_decode": ([vp, u32, vp, vp, vp, vp, vp, u32, vp], i32),
            "svs_score_batch": ([vp, C.POINTER(QueryBatch), i32], i32),
            "svs_score_batch_dev": ([vp, C.POINTER(QueryBatch), i32], i32),
            "svs_dev_alloc": ([vp, u64, C.POINTER(vp)], i32),
            "svs_dev_free": ([vp, vp], i32),
            "svs_memcpy": ([vp, vp, vp, u64], i32),
            "svs_keygen_subspace": ([vp, vp], i32),
            "svs_keygen_derive_h_seed": ([u64, u64, vp], i32),
            "svs_keygen_domain_separate": ([vp, vp, sz, vp], i32),
            "svs_keygen_regenerate_h_s": ([vp, u32, vp, vp, vp, vp, sz, C.POINTER(sz)], i32),
            "svs_quantise": ([vp, u32, u32, u64, vp], i32),
            "svs_probe_route": ([vp, u32, u32, vp, u32, u32, vp], i32),
            "svs_debug_inject_fault": ([vp, C.c_int32, u32, u64], i32),
            "svs_ctx_profile": ([vp, i32, i32, C.POINTER(Profile)], i32),
            "svs_compute_products_batch": ([vp, u32, vp, vp, vp, i32, vp], i32),
            "svs_emvp_create": ([vp, u32, vp, C.POINTER(vp)], i32),
            "svs_emvp_destroy": ([vp], i32),
            "svs_emvp_compute_products": ([vp, u32, vp, vp], i32),
            "svs_emvp_compute_products_dev": ([vp, u32, vp, vp], i32),
            "svs_emvp_decode_scores": ([vp, u32, vp, vp, vp, vp], i32),
            "svs_emvp_decode_scores_dev": ([vp, u32, vp, vp, vp, vp], i32),
            "svs_emvp_score": ([vp, u32, vp, vp, vp], i32),
            "svs_tiptoe_hint_create": ([vp, u32, u32, vp, C.POINTER(vp)], i32),
            "svs_tiptoe_hint_destroy": ([vp], i32),
            "svs_tiptoe_hint_info": ([vp, C.POINTER(u64), C.POINTER(u32), C.POINTER(u64)], i32),
            "svs_tiptoe_apply_hint": ([vp, u32, vp, vp], i32),
            "svs_tiptoe_apply_hint_dev": ([vp, u32, vp, vp], i32),
            "svs_encode_query": ([vp, u32, vp, vp, vp, vp, vp], i32),
            "svs_tiptoe_create": ([vp, u32, u32, i32, vp, C.POINTER(vp)], i32),
            "svs_tiptoe_destroy": ([vp], i32),
            "svs_tiptoe_stored_bytes": ([vp, C.POINTER(u64), C.POINTER(i32)], i32),
            "svs_tiptoe_answer_u32": ([vp, u32, vp, vp, i32], i32),
            "svs_tiptoe_answer_u32_dev": ([vp, u32, vp, vp, i32], i32),
            "svs_tiptoe_answer_u64": ([vp, u32, vp, vp], i32),
            "svs_ctx_copy_probe": ([vp, u64, C.POINTER(C.c_double)], i32),
            "svs_ctx_read_probe": ([vp, u64, C.POINTER(C.c_double)], i32),
            "svs_ctx_set_tile_variant": ([vp, u32, i32], i32),
            "svs_ctx_get_tile_variant": ([vp, u32, C.POINTER(i32), C.POINTER(u32),
                                          C.POINTER(u32), C.POINTER(u32)], i32),
            "svs_autotune": ([vp, C.POINTER(QueryBatch), u32, C.POINTER(i32),
                              C.POINTER(C.c_double)], i32),
            "svs_server_create": ([vp, u32, C.POINTER(ServerConfig), C.POINTER(vp)], i32),
            "svs_server_destroy": ([vp], i32),
            "svs_server_reserve": ([vp, u32, u32, u64, u32], i32),
            "svs_server_submit": ([vp, C.POINTER(Request), C.POINTER(u64)], i32),
            "svs_server_wait": ([vp, u64, C.POINTER(Result)], i32),
            "svs_server_stats": ([vp, C.POINTER(ServerCounters)], i32),
            "svs_server_run_poisson": ([vp, vp, u32, C.c_double, u64, vp,
                                        C.POINTER(C.c_double)], i32),
            "svs_ctx_set_planner": ([vp, i32], i32),
            "svs_ctx_set_product_split": ([vp, i32], i32),
            "svs_ctx_set_graphs": ([vp, i32], i32),
            "svs_topk": ([vp, u32, vp, vp, vp, u32, vp, vp, vp], i32),
            "svs_index_load_ivf_cache": ([vp, C.c_char_p, u64, i32, C.POINTER(vp),
                                          C.POINTER(u32), C.POINTER(u32)], i32),
            "svs_cache_fingerprint64": ([C.POINTER(C.c_char_p), C.POINTER(u64), u32], u64),
            "svs_index_get_centroids": ([vp, vp], i32),
            "svs_index_get_subspace": ([vp, vp], i32),
            "svs_index_set_centroids": ([vp, vp, u32], i32),
            "svs_route_batch": ([vp, vp, u32, u32, vp], i32),
            "svs_score_raw_batch": ([vp, C.POINTER(RawBatch), i32], i32),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes, f.restype = args, res
        _lib = L
    return _lib


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().svs_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int) -> int:
    if rc < 0:
        msg = last_error()
        if rc == SVS_ERR_BUDGET:
            raise BudgetTooSmall(rc, msg)
        raise GpuError(rc, msg)
    return rc


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def u64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint64)


def seed_array(seed) -> np.ndarray:
    a = np.frombuffer(bytes(seed), dtype=np.uint8).copy() if not isinstance(seed, np.ndarray) \
        else np.ascontiguousarray(seed, dtype=np.uint8)
    if a.size != 32:
        raise ValueError("seeds are 32 bytes")
    return a
