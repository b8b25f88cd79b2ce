"""Braverman-Newman scorers on the gfx950 backend.

Mirrors the reference's plugin surface (the async `Scorer` trait,
scorer-core/src/lib.rs:105-187) for the BN flat scorer
(scorer-bntm/src/lib.rs:203-625) and the BN-IVF scorer
(scorer-bntm/src/ivf.rs:209-1366): `upload_cluster`, `score`,
`score_batch`, `communication_cost`, `score_with_realised_cost`, with the
same argument meaning and error behaviour. Every server product,
verification, decode and top-k runs on the GPU through libsvs_gpu.so; the
host keeps what the reference keeps on the client (seed handling,
quantisation, routing), also in native code.

Seeds that the reference draws from OS entropy (h_seed, q_seed) are
injectable here so results can be checked; they default to os.urandom.
"""
from __future__ import annotations

import atexit
import ctypes as C
import math
import os
import time
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from ._lib import N, N1, P, GpuError, check, ptr, seed_array, u64

FIELD_BYTES = 8
Q_DEFAULT = 1 << 20  # params.rs:19


# ---------------------------------------------------------------- core types
@dataclass
class Hit:
    """scorer-core/src/lib.rs:18-22."""
    id: int
    score: float


@dataclass
class CommunicationCost:
    """scorer-core/src/lib.rs:25-55."""
    query_bytes: int
    response_bytes: int
    cluster_response_bytes: int
    setup_bytes: int
    pre_query_offline_up_bytes: int = 0
    pre_query_offline_down_bytes: int = 0
    effective_bytes_per_query: int = 0


@dataclass
class BuildOutcome:
    """scorer-core/src/lib.rs:83-87."""
    cache_hit: bool
    build_duration: float


class BnTmError(Exception):
    """scorer-bntm/src/lib.rs:176-199 and ivf.rs:165-195."""


class DimensionTooLarge(BnTmError):
    def __init__(self, actual, n):
        super().__init__(f"dimension too large: vectors have dim {actual}, n = {n}")
        self.actual, self.n = actual, n


class DimensionMismatch(BnTmError):
    def __init__(self, query_dim, index_dim):
        super().__init__(f"query dimension {query_dim} does not match index dimension {index_dim}")
        self.query_dim, self.index_dim = query_dim, index_dim


class EmptyCorpus(BnTmError):
    pass


class TooManyCentroids(BnTmError):
    pass


class InvalidNprobe(BnTmError):
    def __init__(self, nprobe, n_centroids):
        super().__init__(f"invalid nprobe {nprobe}: must be in 1..={n_centroids}")
        self.nprobe, self.n_centroids = nprobe, n_centroids


class VerificationFailed(BnTmError):
    def __init__(self, trial, cluster=None):
        where = "" if cluster is None else f" at cluster {cluster},"
        super().__init__(f"Protocol 2 verification failed{where} trial {trial}")
        self.trial, self.cluster = trial, cluster


# ------------------------------------------------------------ client helpers
def quantise(vectors, q: int = Q_DEFAULT) -> np.ndarray:
    """quantise_vector (crypto.rs:105-122) row-wise, zero-padded to n."""
    v = np.ascontiguousarray(vectors, dtype=np.float32)
    if v.ndim == 1:
        v = v[None, :]
    rows, d = v.shape
    if d > N:
        raise DimensionTooLarge(d, N)
    out = np.zeros((rows, N), np.uint64)
    if rows:
        check(L.lib().svs_quantise(ptr(v), rows, d, q, ptr(out)))
    return out


def generate_subspace(key_seed) -> np.ndarray:
    out = np.zeros(N * N1, np.uint64)
    check(L.lib().svs_keygen_subspace(ptr(seed_array(key_seed)), ptr(out)))
    return out


def derive_h_seed(upload_seed: int, cluster: int) -> np.ndarray:
    out = np.zeros(32, np.uint8)
    check(L.lib().svs_keygen_derive_h_seed(upload_seed, cluster, ptr(out)))
    return out


def domain_separate(seed, tag: bytes) -> np.ndarray:
    out = np.zeros(32, np.uint8)
    t = np.frombuffer(tag, np.uint8).copy()
    check(L.lib().svs_keygen_domain_separate(ptr(seed_array(seed)), ptr(t), t.size, ptr(out)))
    return out


def regenerate_h_s(h_seed, m: int):
    h = np.zeros(max(m * N1, 1), np.uint64)
    rowptr = np.zeros(m + 1, np.uint32)
    cap = m * N + 1
    idx = np.zeros(cap, np.uint32)
    sign = np.zeros(cap, np.int8)
    nnz = C.c_size_t(0)
    check(L.lib().svs_keygen_regenerate_h_s(ptr(seed_array(h_seed)), m, ptr(h), ptr(rowptr),
                                            ptr(idx), ptr(sign), cap, C.byref(nnz)))
    return h[: m * N1], rowptr, idx[: nnz.value].copy(), sign[: nnz.value].copy()


def probe_route(centroids, queries, nprobe: int) -> np.ndarray:
    cents = np.ascontiguousarray(centroids, np.float32)
    qs = np.ascontiguousarray(queries, np.float32)
    if qs.ndim == 1:
        qs = qs[None, :]
    out = np.zeros((qs.shape[0], nprobe), np.uint32)
    check(L.lib().svs_probe_route(ptr(cents), cents.shape[0], cents.shape[1], ptr(qs),
                                  qs.shape[0], nprobe, ptr(out)))
    return out


def fp_to_signed(mv) -> np.ndarray:
    mv = u64(mv)
    s = mv.astype(np.int64)
    s[mv >= np.uint64(P // 2)] -= np.int64(P)
    return s


# ------------------------------------------------------------- GPU context
class Context:
    """One GPU: replaces the reference's process-wide CUDA context and the
    ClusterStore budget (gpu/mod.rs:169-200, store.rs:133-243)."""

    def __init__(self, device: int = 0, budget_bytes: int | None = None):
        h = C.c_void_p()
        check(L.lib().svs_ctx_create(device, budget_bytes or 0, C.byref(h)))
        self.h = h

    def stats(self) -> dict:
        s = L.Stats()
        check(L.lib().svs_ctx_stats(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in L.Stats._fields_}

    def profile(self, enable: bool = True, reset: bool = False) -> dict:
        """Stage times (ms, HIP events on the library's stream) of
        score_batch since the last reset; enables/disables recording."""
        pr = L.Profile()
        check(L.lib().svs_ctx_profile(self.h, int(enable), int(reset), C.byref(pr)))
        out = {s: (pr.ms[i], pr.calls[i]) for i, s in enumerate(L.STAGES)}
        out["gemv_bytes"] = pr.gemv_bytes
        out["gemv_rows"] = pr.gemv_rows
        return out

    def copy_probe(self, nbytes: int) -> float:
        """ms for a plain device copy of nbytes (HBM-counter calibration)."""
        ms = C.c_double(0)
        check(L.lib().svs_ctx_copy_probe(self.h, nbytes, C.byref(ms)))
        return ms.value

    def set_tile_variant(self, variant: int, n_queries: int = 0):
        """Block-tile variant of the MFMA batch path for batches of
        n_queries (0: every batch size)."""
        check(L.lib().svs_ctx_set_tile_variant(self.h, n_queries, variant))

    def tile_variant(self, n_queries: int) -> dict:
        v, qg, pf, occ = C.c_int(0), C.c_uint32(0), C.c_uint32(0), C.c_uint32(0)
        check(L.lib().svs_ctx_get_tile_variant(self.h, n_queries, C.byref(v), C.byref(qg),
                                               C.byref(pf), C.byref(occ)))
        return dict(variant=v.value, query_group=qg.value, ring_depth=pf.value,
                    waves_per_simd=occ.value)

    def read_probe(self, nbytes: int) -> float:
        """ms for a read of nbytes in the batched product's access pattern."""
        ms = C.c_double(0)
        check(L.lib().svs_ctx_read_probe(self.h, nbytes, C.byref(ms)))
        return ms.value

    def set_planner(self, on_device: bool = True):
        """Plan batched runs on the device (default) or on the host."""
        check(L.lib().svs_ctx_set_planner(self.h, 1 if on_device else 0))

    def set_product_split(self, on: bool = True):
        """Run the batched product of 16-query tile variants on the staged
        kernel (each step's rows and query planes staged once in LDS per
        work unit of up to 64 queries); off (the default): the per-wave
        plane kernel for every tile. Same results either way."""
        check(L.lib().svs_ctx_set_product_split(self.h, 1 if on else 0))

    def set_graphs(self, on: bool = True):
        """Replay device-planned batches from captured HIP graphs (default)
        or launch every kernel."""
        check(L.lib().svs_ctx_set_graphs(self.h, 1 if on else 0))

    def topk(self, seg_ptr, mv, ids, k: int):
        """svs_topk: per segment the k best fp_to_signed(mv) with their ids
        (score desc, position asc); returns (ids, scores, count)."""
        sp = np.ascontiguousarray(seg_ptr, np.uint32)
        nq = sp.size - 1
        m = u64(mv).reshape(-1)
        i = np.ascontiguousarray(ids, np.uint32).reshape(-1)
        oi = np.zeros((nq, k), np.uint32)
        os_ = np.zeros((nq, k), np.int64)
        oc = np.zeros(nq, np.uint32)
        check(L.lib().svs_topk(self.h, nq, ptr(sp), ptr(m) if m.size else None,
                               ptr(i) if i.size else None, k, ptr(oi), ptr(os_), ptr(oc)))
        return oi, os_, oc

    def inject_fault(self, list_id: int, row: int = 0, delta: int = 1):
        """Test hook: a server that adds `delta` to r[row] of `list_id`
        (list_id < 0 disables)."""
        check(L.lib().svs_debug_inject_fault(self.h, list_id, row, delta))

    def close(self):
        if self.h:
            L.lib().svs_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Context | None = None
_live_indexes: "weakref.WeakSet[GpuIndex]" = weakref.WeakSet()
_live_servers: weakref.WeakSet = weakref.WeakSet()


@atexit.register
def _close_all():
    """Release servers, then indexes, then contexts at interpreter exit
    (finalisers run in no particular order then)."""
    for srv in list(_live_servers):
        srv.close()
    for idx in list(_live_indexes):
        idx.close()
    if _default_ctx is not None:
        _default_ctx.close()


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class GpuIndex:
    """Resident encrypted lists on one GPU (svs_index_*)."""

    def __init__(self, ctx: Context, n_lists: int, l_subspace):
        self.ctx = ctx
        self.l = u64(l_subspace)
        h = C.c_void_p()
        check(L.lib().svs_index_create(ctx.h, n_lists, ptr(self.l), C.byref(h)))
        self.h = h
        self.n_lists = n_lists
        self.cent_d = 0
        _live_indexes.add(self)

    def set_centroids(self, centroids):
        """The IVF centroids (n_lists x d f32) for routing on the device."""
        c = np.ascontiguousarray(centroids, np.float32)
        if c.ndim != 2 or c.shape[0] != self.n_lists:
            raise ValueError(f"centroids must be {self.n_lists} x d")
        check(L.lib().svs_index_set_centroids(self.h, ptr(c), c.shape[1]))
        self.cent_d = c.shape[1]

    def route(self, queries, nprobe: int) -> np.ndarray:
        """probe_route on the device: (n_queries, min(nprobe, n_lists))."""
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, max(self.cent_d, 1))
        np_ = min(int(nprobe), self.n_lists)
        out = np.zeros((q.shape[0], max(np_, 1)), np.uint32)
        check(L.lib().svs_route_batch(self.h, ptr(q), q.shape[0], np_, ptr(out)))
        return out[:, :np_]

    def score_raw_batch(self, queries, q_seeds, nprobe, verify, k, quantisation_q=Q_DEFAULT,
                        mode=L.MODE_BATCHED):
        """svs_score_raw_batch: f32 queries quantised and routed on the
        device, then scored. nprobe: one value or one per query. Returns
        (ids, scores, count, status, probes) with probes concatenated."""
        q = np.ascontiguousarray(queries, np.float32).reshape(-1, max(self.cent_d, 1))
        nq = q.shape[0]
        seeds = np.ascontiguousarray(q_seeds, np.uint8).reshape(nq, 32)
        per = None
        if np.ndim(nprobe):
            per = np.minimum(np.ascontiguousarray(nprobe, np.uint32), self.n_lists)
            total = int(per.sum())
        else:
            total = nq * min(int(nprobe), self.n_lists)
        ver = None if verify is None else np.ascontiguousarray(verify, np.uint8)
        ids = np.zeros((nq, k), np.uint32)
        scores = np.zeros((nq, k), np.int64)
        count = np.zeros(nq, np.uint32)
        status = np.zeros(nq, np.int32)
        probes = np.zeros(max(total, 1), np.uint32)
        b = L.RawBatch(nq, ptr(q), ptr(per), 0 if per is not None else int(nprobe),
                       int(quantisation_q), ptr(seeds), ptr(ver), k, ptr(ids), ptr(scores),
                       ptr(count), ptr(status), ptr(probes))
        check(L.lib().svs_score_raw_batch(self.h, C.byref(b), mode))
        return ids, scores, count, status, probes[:total]

    def upload_list(self, list_id, cl: dict, ids=None):
        m = int(cl["m"])
        ids = None if ids is None else np.ascontiguousarray(ids, np.uint32)
        z1 = np.zeros(1, np.uint64)
        check(L.lib().svs_index_upload_list(
            self.h, list_id, m, ptr(u64(cl["m_enc"]) if m else z1),
            ptr(u64(cl["a_l"]) if m else z1), ptr(u64(cl["h"]) if m else z1),
            ptr(u64(cl["m_plain"])) if cl.get("m_plain") is not None and m else None,
            ptr(np.ascontiguousarray(cl["s_rowptr"], np.uint32)),
            ptr(np.ascontiguousarray(cl["s_idx"], np.uint32)) if len(cl["s_idx"]) else None,
            ptr(np.ascontiguousarray(cl["s_sign"], np.int8)) if len(cl["s_sign"]) else None,
            ptr(ids)))

    def build_list(self, list_id, m_plain, h_seed, ids=None):
        mp = u64(m_plain).reshape(-1)
        m = mp.size // N
        ids = None if ids is None else np.ascontiguousarray(ids, np.uint32)
        check(L.lib().svs_index_build_list(self.h, list_id, m, ptr(mp) if m else None,
                                           ptr(seed_array(h_seed)), ptr(ids)))

    def rows(self, list_id) -> int:
        return L.lib().svs_index_list_rows(self.h, list_id)

    def download_list(self, list_id):
        """(M_enc, AL, H) of a list, row-major, copied from the device."""
        m = self.rows(list_id)
        me = np.zeros(max(m * N, 1), np.uint64)
        al = np.zeros(max(m * N1, 1), np.uint64)
        hh = np.zeros(max(m * N1, 1), np.uint64)
        check(L.lib().svs_index_download_list(self.h, list_id, ptr(me), ptr(al), ptr(hh)))
        return me[: m * N], al[: m * N1], hh[: m * N1]

    def download_menc(self, list_id) -> np.ndarray:
        return self.download_list(list_id)[0]

    def compute_products(self, list_id, v_enc):
        """IvfGpuState::compute_products: None for an empty/out-of-range list."""
        m = self.rows(list_id)
        r = np.zeros(max(m, 1), np.uint64)
        rc = check(L.lib().svs_compute_products(self.h, list_id, ptr(u64(v_enc)), ptr(r)))
        return None if rc == L.SVS_NONE else r[:m]

    def compute_products_batch(self, v_enc, probe_ptr, probe_lists, kernel=L.KERNEL_MFMA):
        """Gathered products: list of r vectors, one per (query, probed list)
        item in order (None for empty lists)."""
        v_enc = u64(v_enc).reshape(-1, N)
        pp = np.ascontiguousarray(probe_ptr, np.uint32)
        pl = np.ascontiguousarray(probe_lists, np.uint32)
        sizes = [self.rows(int(c)) for c in pl]
        r = np.zeros(max(sum(sizes), 1), np.uint64)
        check(L.lib().svs_compute_products_batch(self.h, v_enc.shape[0], ptr(v_enc), ptr(pp),
                                                 ptr(pl if pl.size else np.zeros(1, np.uint32)),
                                                 kernel, ptr(r)))
        out, off = [], 0
        for s in sizes:
            out.append(r[off:off + s] if s else None)
            off += s
        return out

    def decode_dense_terms(self, list_id, g, lt_venc):
        m = self.rows(list_id)
        out = np.zeros(max(m, 1), np.uint64)
        rc = check(L.lib().svs_decode_dense_terms(self.h, list_id, ptr(u64(g)), ptr(u64(lt_venc)),
                                                  ptr(out)))
        return None if rc == L.SVS_NONE else out[:m]

    def verify(self, list_id, r, v_enc, rng_seed, trials=L.LAMBDA) -> int:
        """-1 on accept, else the first failing trial."""
        ft = C.c_int32(0)
        rc = check(L.lib().svs_verify(self.h, list_id, ptr(u64(r)), ptr(u64(v_enc)),
                                      ptr(seed_array(rng_seed)), trials, C.byref(ft)))
        if rc == L.SVS_NONE:
            return -1
        return ft.value

    def verify_noalloc(self, list_id, r_dev_ptr: int, v_enc_dev_ptr: int, rng_seed,
                       trials=L.LAMBDA) -> int:
        """svs_verify_noalloc: r and v_enc already in device memory (e.g.
        DeviceArray.ptr()); no device allocation. -1 on accept, else the
        first failing trial."""
        ft = C.c_int32(0)
        rc = check(L.lib().svs_verify_noalloc(self.h, list_id, C.c_void_p(r_dev_ptr),
                                              C.c_void_p(v_enc_dev_ptr),
                                              ptr(seed_array(rng_seed)), trials, C.byref(ft)))
        if rc == L.SVS_NONE:
            return -1
        return ft.value

    def decode(self, list_id, r, v_enc, g, t_idx, t_sign):
        m = self.rows(list_id)
        out = np.zeros(max(m, 1), np.uint64)
        t_idx = np.ascontiguousarray(t_idx, np.uint32)
        t_sign = np.ascontiguousarray(t_sign, np.int8)
        rc = check(L.lib().svs_decode(self.h, list_id, ptr(u64(r)), ptr(u64(v_enc)), ptr(u64(g)),
                                      ptr(t_idx) if t_idx.size else None,
                                      ptr(t_sign) if t_sign.size else None, t_idx.size,
                                      ptr(out)))
        return None if rc == L.SVS_NONE else out[:m]

    def encode_query(self, v_q, q_seeds):
        """Device encode_query (crypto.rs:508-553) with this index's L:
        (v_enc, g, t) for each query, t dense in {-1, 0, 1}."""
        v = u64(v_q).reshape(-1, N)
        nq = v.shape[0]
        seeds = np.ascontiguousarray(q_seeds, np.uint8).reshape(nq, 32)
        v_enc = np.zeros((nq, N), np.uint64)
        g = np.zeros((nq, N1), np.uint64)
        t = np.zeros((nq, N), np.int8)
        check(L.lib().svs_encode_query(self.h, nq, ptr(v), ptr(seeds), ptr(v_enc), ptr(g), ptr(t)))
        return v_enc, g, t

    def _batch(self, v_q, q_seeds, probe_ptr, probe_lists, verify, k):
        v_q = u64(v_q).reshape(-1, N)
        nq = v_q.shape[0]
        seeds = np.ascontiguousarray(q_seeds, np.uint8).reshape(nq, 32)
        pp = np.ascontiguousarray(probe_ptr, np.uint32)
        pl = np.ascontiguousarray(probe_lists, np.uint32)
        if pl.size == 0:
            pl = np.zeros(1, np.uint32)
        ver = None if verify is None else np.ascontiguousarray(verify, np.uint8)
        ids = np.zeros((nq, k), np.uint32)
        scores = np.zeros((nq, k), np.int64)
        count = np.zeros(nq, np.uint32)
        status = np.zeros(nq, np.int32)
        b = L.QueryBatch(nq, ptr(v_q), ptr(seeds), ptr(pp), ptr(pl), ptr(ver), k, ptr(ids),
                         ptr(scores), ptr(count), ptr(status))
        # the arrays stay referenced by the returned tuple while b is used
        return b, (v_q, seeds, pp, pl, ver), (ids, scores, count, status)

    def score_batch(self, v_q, q_seeds, probe_ptr, probe_lists, verify, k,
                    mode=L.MODE_BATCHED):
        b, _keep, out = self._batch(v_q, q_seeds, probe_ptr, probe_lists, verify, k)
        check(L.lib().svs_score_batch(self.h, C.byref(b), mode))
        return out

    def score_batch_dev(self, n_queries, v_q_dev, q_seeds_dev, probe_ptr, probe_lists, verify,
                        k, ids_dev, scores_dev, count_dev, status_dev, mode=L.MODE_BATCHED):
        """svs_score_batch_dev: query inputs and outputs already in device
        memory (device addresses as ints, e.g. torch's data_ptr()); the probe
        plan and verification flags stay on the host."""
        pp = np.ascontiguousarray(probe_ptr, np.uint32)
        pl = np.ascontiguousarray(probe_lists, np.uint32)
        if pl.size == 0:
            pl = np.zeros(1, np.uint32)
        ver = None if verify is None else np.ascontiguousarray(verify, np.uint8)
        b = L.QueryBatch(n_queries, v_q_dev, q_seeds_dev, ptr(pp), ptr(pl), ptr(ver), k,
                         ids_dev, scores_dev, count_dev, status_dev)
        check(L.lib().svs_score_batch_dev(self.h, C.byref(b), mode))

    def autotune(self, v_q, q_seeds, probe_ptr, probe_lists, verify, k, reps: int = 2):
        """Times each block-tile variant on this batch and keeps the fastest
        for its size class; returns (best, ms per variant)."""
        b, _keep, _out = self._batch(v_q, q_seeds, probe_ptr, probe_lists, verify, k)
        best = C.c_int(0)
        ms = (C.c_double * L.TILE_VARIANTS)()
        check(L.lib().svs_autotune(self.h, C.byref(b), reps, C.byref(best), ms))
        return best.value, list(ms)

    def close(self):
        if getattr(self, "h", None):
            L.lib().svs_index_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _hits(ids, scores, count) -> list[Hit]:
    return [Hit(int(ids[i]), float(np.float32(scores[i]))) for i in range(int(count))]


def _seeds(n, q_seeds):
    if q_seeds is None:
        return np.frombuffer(os.urandom(32 * n), np.uint8).reshape(n, 32).copy()
    if isinstance(q_seeds, (list, tuple)):
        return np.stack([seed_array(s) for s in q_seeds]).reshape(n, 32)
    return np.ascontiguousarray(q_seeds, np.uint8).reshape(n, 32)


# ------------------------------------------------------------------ BN flat
@dataclass
class BnTmConfig:
    """scorer-bntm/src/lib.rs:45-101 (Sec128 only: n = 1024, n1 = 128)."""
    key_seed: bytes = bytes(32)
    verification_enabled: bool = False
    quantisation_q: int = Q_DEFAULT
    vram_budget_bytes: int | None = None


@dataclass
class BnTmHandle:
    index: GpuIndex
    l_subspace: np.ndarray
    m: int
    dim: int
    verification_enabled: bool
    quantisation_q: int


class BnTmScorer:
    """BnTmScorer (lib.rs:203) with the GPU path for every per-query stage."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx

    def upload_cluster(self, config: BnTmConfig, vectors, h_seed=None):
        t0 = time.perf_counter()
        v = np.ascontiguousarray(vectors, np.float32)
        if v.ndim == 1:
            v = v[None, :]
        m, dim = v.shape if v.size else (0, 0)
        if dim > N:
            raise DimensionTooLarge(dim, N)
        ctx = self.ctx or default_context()
        l = generate_subspace(config.key_seed)
        idx = GpuIndex(ctx, 1, l)
        h_seed = os.urandom(32) if h_seed is None else h_seed
        idx.build_list(0, quantise(v, config.quantisation_q), h_seed)
        handle = BnTmHandle(idx, l, m, dim, config.verification_enabled, config.quantisation_q)
        return handle, BuildOutcome(False, time.perf_counter() - t0)

    def score(self, handle: BnTmHandle, query, k: int, q_seed=None, mode=L.MODE_PLAIN):
        return self.score_batch(handle, np.asarray(query, np.float32)[None, :], k,
                                None if q_seed is None else [q_seed], mode=mode)[0]

    def score_batch(self, handle: BnTmHandle, queries, k: int, q_seeds=None,
                    mode=L.MODE_BATCHED):
        qs = np.ascontiguousarray(queries, np.float32)
        if qs.ndim == 1:
            qs = qs[None, :]
        nq = qs.shape[0]
        if qs.shape[1] > N:
            raise DimensionTooLarge(qs.shape[1], N)
        if nq == 0:
            return []
        k_capped = min(k, handle.m)
        if k_capped == 0:
            return [[] for _ in range(nq)]
        v_q = quantise(qs, handle.quantisation_q)
        seeds = _seeds(nq, q_seeds)
        pp = np.arange(nq + 1, dtype=np.uint32)
        pl = np.zeros(nq, np.uint32)
        ver = np.full(nq, 1 if handle.verification_enabled else 0, np.uint8)
        ids, scores, count, status = handle.index.score_batch(v_q, seeds, pp, pl, ver, k_capped,
                                                             mode)
        out = []
        for q in range(nq):
            if status[q]:
                raise VerificationFailed((int(status[q]) - 1) % 4)
            out.append(_hits(ids[q], scores[q], count[q]))
        return out

    def communication_cost(self, handle: BnTmHandle, k: int) -> CommunicationCost:
        """lib.rs:477-511."""
        m = handle.m
        return CommunicationCost(N * FIELD_BYTES, m * FIELD_BYTES, m * FIELD_BYTES,
                                 m * N * FIELD_BYTES, 0, 0, m * N * FIELD_BYTES)


# ------------------------------------------------------------------- BN-IVF
@dataclass
class IvfIndex:
    """Plaintext-side IVF (ivf-index/src/ivf.rs:10-14): centroids and the
    member ids of each list in ascending order. k-means training is an
    offline step; the index is an input here."""
    centroids: np.ndarray
    lists: list = field(default_factory=list)


@dataclass
class BnTmIvfConfig:
    """scorer-bntm/src/ivf.rs:52-92."""
    index: IvfIndex
    nprobe: int
    key_seed: bytes = bytes(32)
    upload_seed: int = 7
    verification_enabled: bool = False
    quantisation_q: int = Q_DEFAULT
    vram_budget_bytes: int | None = None


@dataclass
class BnTmIvfHandle:
    index: GpuIndex
    ivf: IvfIndex
    l_subspace: np.ndarray
    sizes: np.ndarray
    nprobe: int
    dim: int
    verification_enabled: bool
    quantisation_q: int


def cache_fingerprint64(parts) -> int:
    """fingerprint64 of the cache key parts (scorer-core/src/cache.rs:22-37)."""
    bs = [bytes(p) for p in parts]
    arr = (C.c_char_p * max(len(bs), 1))(*bs)
    lens = (C.c_uint64 * max(len(bs), 1))(*[len(b) for b in bs])
    return int(L.lib().svs_cache_fingerprint64(arr, lens, len(bs)))


def load_bntm_ivf_cache(path, ctx: "Context | None" = None, fingerprint: int | None = None,
                        nprobe: int = 1, verification_enabled: bool = False,
                        quantisation_q: int = Q_DEFAULT):
    """The reference's BN-IVF cache file loaded straight into HBM
    (load_bntm_ivf_cache, scorer-bntm/src/ivf.rs:1761-1902): a handle for
    BnTmIvfScorer, or None when the header does not match (magic, version,
    and the fingerprint when given)."""
    ctx = ctx or default_context()
    h = C.c_void_p()
    nl, d = C.c_uint32(0), C.c_uint32(0)
    rc = check(L.lib().svs_index_load_ivf_cache(ctx.h, os.fsencode(path), fingerprint or 0,
                                                1 if fingerprint is not None else 0, C.byref(h),
                                                C.byref(nl), C.byref(d)))
    if rc == L.SVS_NONE:
        return None
    idx = GpuIndex.__new__(GpuIndex)
    idx.ctx, idx.h, idx.n_lists, idx.cent_d = ctx, h, nl.value, d.value
    idx.l = np.zeros(N * N1, np.uint64)
    check(L.lib().svs_index_get_subspace(h, ptr(idx.l)))
    _live_indexes.add(idx)
    cent = np.zeros((nl.value, d.value), np.float32)
    check(L.lib().svs_index_get_centroids(h, ptr(cent)))
    sizes = np.array([idx.rows(c) for c in range(nl.value)], np.int64)
    if nprobe == 0 or nprobe > nl.value:
        raise InvalidNprobe(nprobe, nl.value)
    return BnTmIvfHandle(idx, IvfIndex(cent, []), idx.l, sizes, nprobe, d.value,
                         verification_enabled, quantisation_q)


class BnTmIvfScorer:
    """BnTmIvfScorer (ivf.rs:209) on the gfx950 backend."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx

    def upload_cluster(self, config: BnTmIvfConfig, vectors):
        t0 = time.perf_counter()
        v = np.ascontiguousarray(vectors, np.float32)
        if v.shape[0] == 0:
            raise EmptyCorpus("corpus must contain at least one vector")
        nc = len(config.index.lists)
        if nc > v.shape[0]:
            raise TooManyCentroids(nc)
        if config.nprobe == 0 or config.nprobe > nc:
            raise InvalidNprobe(config.nprobe, nc)
        dim = v.shape[1]
        if dim > N:
            raise DimensionTooLarge(dim, N)
        ctx = self.ctx or default_context()
        l = generate_subspace(config.key_seed)
        idx = GpuIndex(ctx, nc, l)
        sizes = np.zeros(nc, np.int64)
        for ci, members in enumerate(config.index.lists):
            members = np.ascontiguousarray(members, np.uint32)
            sizes[ci] = members.size
            mq = quantise(v[members], config.quantisation_q) if members.size else \
                np.zeros((0, N), np.uint64)
            idx.build_list(ci, mq, derive_h_seed(config.upload_seed, ci), members)
        cent = np.asarray(config.index.centroids, np.float32)
        if nc <= L.ROUTE_MAX_LISTS and cent.ndim == 2 and cent.shape[1] == dim:
            idx.set_centroids(cent)
        h = BnTmIvfHandle(idx, config.index, l, sizes, config.nprobe, dim,
                          config.verification_enabled, config.quantisation_q)
        return h, BuildOutcome(False, time.perf_counter() - t0)

    def _run(self, handle, qs, k, q_seeds, mode):
        qs = np.ascontiguousarray(qs, np.float32)
        if qs.ndim == 1:
            qs = qs[None, :]
        if qs.shape[1] != handle.dim:
            raise DimensionMismatch(qs.shape[1], handle.dim)
        nq = qs.shape[0]
        seeds = _seeds(nq, q_seeds)
        ver = np.full(nq, 1 if handle.verification_enabled else 0, np.uint8)
        if handle.index.cent_d:
            # steps 1-2 of score (routing, quantisation) on the device too
            *res, probes = handle.index.score_raw_batch(qs, seeds, handle.nprobe, ver, max(k, 1),
                                                        handle.quantisation_q, mode)
            return probes.reshape(nq, -1), tuple(res)
        probes = probe_route(handle.ivf.centroids, qs, handle.nprobe)
        v_q = quantise(qs, handle.quantisation_q)
        pp = np.arange(0, nq * handle.nprobe + 1, handle.nprobe, dtype=np.uint32)
        res = handle.index.score_batch(v_q, seeds, pp, probes.reshape(-1), ver, max(k, 1), mode)
        return probes, res

    def score(self, handle: BnTmIvfHandle, query, k: int, q_seed=None, mode=L.MODE_PLAIN):
        return self.score_batch(handle, np.asarray(query, np.float32)[None, :], k,
                                None if q_seed is None else [q_seed], mode=mode)[0]

    def score_batch(self, handle: BnTmIvfHandle, queries, k: int, q_seeds=None,
                    mode=L.MODE_BATCHED):
        qs = np.asarray(queries, np.float32)
        if qs.ndim == 2 and qs.shape[0] == 0:
            return []
        _, (ids, scores, count, status) = self._run(handle, qs, k, q_seeds, mode)
        out = []
        for q in range(ids.shape[0]):
            if status[q]:
                st = int(status[q]) - 1
                raise VerificationFailed(st % 4, st // 4)
            out.append(_hits(ids[q], scores[q], min(int(count[q]), k)))
        return out

    def communication_cost(self, handle: BnTmIvfHandle, k: int) -> CommunicationCost:
        """ivf.rs:928-970."""
        mean_m = float(handle.sizes.mean()) if handle.sizes.size else 0.0
        # f64::round: half away from zero (Python's round is half to even)
        avg = int(math.floor(mean_m * FIELD_BYTES + 0.5))
        return CommunicationCost(N * FIELD_BYTES, handle.nprobe * avg, avg,
                                 int(handle.sizes.sum()) * N * FIELD_BYTES, 0, 0,
                                 handle.nprobe * int(mean_m) * N * FIELD_BYTES)

    def score_with_realised_cost(self, handle: BnTmIvfHandle, query, k: int, q_seed=None):
        """ivf.rs:972-1140: realised response bytes from the probed sizes."""
        probes, (ids, scores, count, status) = self._run(
            handle, np.asarray(query, np.float32)[None, :], k,
            None if q_seed is None else [q_seed], L.MODE_PLAIN)
        if status[0]:
            st = int(status[0]) - 1
            raise VerificationFailed(st % 4, st // 4)
        sizes = handle.sizes[probes[0]]
        cost = self.communication_cost(handle, k)
        per = [int(s) * FIELD_BYTES for s in sizes]
        cost.response_bytes = sum(per)
        cost.cluster_response_bytes = sum(per) // len(per) if per else 0
        cost.effective_bytes_per_query = int(sizes.sum()) * N * FIELD_BYTES
        return _hits(ids[0], scores[0], min(int(count[0]), k)), cost


class DeviceArray:
    """A numpy-shaped buffer in HBM (svs_dev_alloc), for the device-resident
    batch entry point; `ptr(offset)` is a device address."""

    def __init__(self, ctx: Context, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        h = C.c_void_p()
        check(L.lib().svs_dev_alloc(ctx.h, max(self.nbytes, 1), C.byref(h)))
        self.h = h

    @classmethod
    def from_host(cls, ctx: Context, a):
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.shape, a.dtype)
        check(L.lib().svs_memcpy(ctx.h, d.h, a.ctypes.data_as(C.c_void_p), a.nbytes))
        return d

    def ptr(self, offset_bytes: int = 0) -> int:
        return self.h.value + offset_bytes

    def copy_from(self, a, offset_bytes: int = 0):
        a = np.ascontiguousarray(a)
        check(L.lib().svs_memcpy(self.ctx.h, C.c_void_p(self.ptr(offset_bytes)),
                                 a.ctypes.data_as(C.c_void_p), a.nbytes))

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(L.lib().svs_memcpy(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.h,
                                 self.nbytes))
        return out

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            L.lib().svs_dev_free(self.ctx.h, self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
