"""Serving: continuous batching of BN-IVF queries over resident indexes.

Thin mirror of the svs_server_* C ABI (include/svs_gpu.h). A Server keeps
one or more GpuIndex tenants (each with its own keys and lists, all on one
Context) and a request queue; its native dispatcher thread packs pending
requests into batches whenever a device slot frees up. The reference
serves one query per BnTmIvfScorer::score call (scorer-bntm/src/ivf.rs:
746-822); a request here carries the same inputs as one such call after
routing and quantisation (v_q, q_seed, probe order, verification flag, k)
and returns the same hits.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import check, ptr, seed_array, u64
from . import bntm as _bntm
from ._lib import N
from .bntm import GpuIndex


@dataclass
class ServedResult:
    ids: np.ndarray
    scores: np.ndarray
    status: int
    batch_size: int
    latency_ms: float


def _request(tenant, v_q, q_seed, probes, verify, k, keep):
    v = u64(v_q).reshape(-1)
    if v.size != N:
        raise ValueError(f"v_q must have {N} entries")
    s = seed_array(q_seed)
    p = np.ascontiguousarray(probes, np.uint32).reshape(-1)
    keep.extend((v, s, p))
    return L.Request(tenant, ptr(v), ptr(s), ptr(p) if p.size else None, p.size,
                     1 if verify else 0, k, 0.0, None, 0, 0)


def _raw_request(tenant, query, q_seed, nprobe, verify, k, quantisation_q, keep):
    x = np.ascontiguousarray(query, np.float32).reshape(-1)
    s = seed_array(q_seed)
    keep.extend((x, s))
    return L.Request(tenant, None, ptr(s), None, 0, 1 if verify else 0, k, 0.0, ptr(x),
                     int(nprobe), int(quantisation_q))


def _any_request(rq, keep):
    if rq.get("query") is not None:
        return _raw_request(rq.get("tenant", 0), rq["query"], rq["q_seed"], rq["nprobe"],
                            rq.get("verify", False), rq.get("k", 10),
                            rq.get("quantisation_q", _bntm.Q_DEFAULT), keep)
    return _request(rq.get("tenant", 0), rq["v_q"], rq["q_seed"], rq["probes"],
                    rq.get("verify", False), rq.get("k", 10), keep)


class Server:
    def __init__(self, tenants, max_batch: int = 256, slots: int = 2, slot_bytes: int = 0,
                 mode: int = L.MODE_BATCHED, cross_tenant: bool = False):
        tenants = [tenants] if isinstance(tenants, GpuIndex) else list(tenants)
        self._tenants = tenants  # the indexes must outlive the server
        arr = (C.c_void_p * len(tenants))(*[t.h for t in tenants])
        cfg = L.ServerConfig(max_batch, slots, slot_bytes, mode, 1 if cross_tenant else 0)
        h = C.c_void_p()
        check(L.lib().svs_server_create(arr, len(tenants), C.byref(cfg), C.byref(h)))
        self.h = h
        self._keep = {}
        _bntm._live_servers.add(self)

    def reserve(self, n_queries: int, n_items: int, n_candidates: int, k: int = 10):
        """Pre-size every slot for batches up to these sizes (no reallocation
        while serving)."""
        check(L.lib().svs_server_reserve(self.h, n_queries, n_items, n_candidates, k))

    def submit(self, v_q, q_seed, probes, verify: bool = False, k: int = 10,
               tenant: int = 0) -> int:
        keep = []
        r = _request(tenant, v_q, q_seed, probes, verify, k, keep)
        t = C.c_uint64(0)
        check(L.lib().svs_server_submit(self.h, C.byref(r), C.byref(t)))
        self._keep[t.value] = k
        return t.value

    def submit_raw(self, query, q_seed, nprobe: int, verify: bool = False, k: int = 10,
                   tenant: int = 0, quantisation_q: int = _bntm.Q_DEFAULT) -> int:
        """A raw f32 query: quantised and routed to `nprobe` lists on the
        device (the tenant's index needs its centroids)."""
        keep = []
        r = _raw_request(tenant, query, q_seed, nprobe, verify, k, quantisation_q, keep)
        t = C.c_uint64(0)
        check(L.lib().svs_server_submit(self.h, C.byref(r), C.byref(t)))
        self._keep[t.value] = k
        return t.value

    def wait(self, ticket: int) -> ServedResult:
        k = self._keep.pop(ticket)
        ids = np.zeros(k, np.uint32)
        scores = np.zeros(k, np.int64)
        res = L.Result(ptr(ids), ptr(scores), 0, 0, 0, 0, 0.0)
        check(L.lib().svs_server_wait(self.h, ticket, C.byref(res)))
        return ServedResult(ids[:res.count].copy(), scores[:res.count].copy(), res.status,
                            res.batch_size, res.latency_ms)

    def run_poisson(self, requests, qps: float, seed: int = 1):
        """Open-loop Poisson arrivals at `qps` over `requests` (dicts with
        v_q, q_seed, probes, verify, k, tenant, or a raw `query` with
        `nprobe` instead of v_q and probes); returns (results, wall_s)."""
        keep = []
        n = len(requests)
        reqs = (L.Request * n)()
        outs = []
        res = (L.Result * n)()
        for i, rq in enumerate(requests):
            reqs[i] = _any_request(rq, keep)
            ids = np.zeros(rq.get("k", 10), np.uint32)
            sc = np.zeros(rq.get("k", 10), np.int64)
            outs.append((ids, sc))
            res[i] = L.Result(ptr(ids), ptr(sc), 0, 0, 0, 0, 0.0)
        wall = C.c_double(0)
        check(L.lib().svs_server_run_poisson(self.h, reqs, n, qps, seed, res, C.byref(wall)))
        out = [ServedResult(outs[i][0][:res[i].count].copy(), outs[i][1][:res[i].count].copy(),
                            res[i].status, res[i].batch_size, res[i].latency_ms)
               for i in range(n)]
        return out, wall.value

    def stats(self) -> dict:
        c = L.ServerCounters()
        check(L.lib().svs_server_stats(self.h, C.byref(c)))
        return {f: getattr(c, f) for f, _ in L.ServerCounters._fields_}

    def close(self):
        if getattr(self, "h", None):
            L.lib().svs_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
