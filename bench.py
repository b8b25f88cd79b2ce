#!/usr/bin/env python3
"""BN-IVF on one MI355X: served (continuous batching, open-loop arrivals)
against the plain per-query GPU path and the reference CPU path.

Workload (defaults): N = 1,000,000 synthetic d = 768 unit vectors from a
seeded 1024-component Gaussian mixture with Zipf(1.1) component weights;
nlist = 1024 IVF lists from a seeded k-means of the corpus (`--lists
mixture`: the mixture components); nprobe = 32; k = 10;
Sec128 BN encryption (n = 1024, n1 = 128, p = 2^61 - 1).

`value` is the served throughput at the latency SLO: raw f32 queries arrive
open-loop (Poisson) at each offered rate, are quantised and routed on the
device, batched as they arrive and scored (encode -> server product ->
[Freivalds] -> decode -> top-k); the highest achieved rate whose p99
(scheduled arrival -> result) stays under 10 ms is reported with its p50 and
p99. Beside it: the plain one-query-per-launch path over 1,000 queries, the
closed-loop batched rate (a step = one 256-query batch with device-resident
quantised queries and probe lists, replayed from a HIP graph), a
half-verified batched leg, the product kernel's roofline (HIP events on the
library's stream), the plain BN flat drop-in at N = 1M (verification off and
on), the 16-tenant served configuration (N = 8.8M), and the CPU baseline:
the C restatement of the reference's CPU path in its closed-loop B=1
protocol with row-parallel threads, scalar and AVX-512 builds.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "secure-vector-search_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def pct(xs, q):
    return float(np.percentile(np.asarray(xs, np.float64), q)) if len(xs) else None


def run_batches(index, v_q, seeds, probes, verify, k, batch, nbatches, mode, offset=0):
    """Runs nbatches batches; returns per-batch wall seconds and results."""
    nprobe = probes.shape[1]
    lat, res = [], []
    for b in range(nbatches):
        s = slice(offset + b * batch, offset + (b + 1) * batch)
        nq = v_q[s].shape[0]
        pp = np.arange(0, nq * nprobe + 1, nprobe, dtype=np.uint32)
        t0 = time.perf_counter()
        out = index.score_batch(v_q[s], seeds[s], pp, probes[s].reshape(-1), verify[s], k, mode)
        lat.append(time.perf_counter() - t0)
        res.append(out)
    return lat, res


def run_raw(index, qs, seeds, nprobe, verify, k, batch, nbatches, mode, offset=0):
    """Raw f32 queries (host buffers): quantise + route on the device, then
    the pipeline; per-batch wall seconds and results."""
    lat, res = [], []
    for b in range(nbatches):
        s = slice(offset + b * batch, offset + (b + 1) * batch)
        t0 = time.perf_counter()
        out = index.score_raw_batch(qs[s], seeds[s], nprobe, verify[s], k, mode=mode)
        lat.append(time.perf_counter() - t0)
        res.append(out)
    return lat, res


class DeviceQueries:
    """The workload's query vectors and seeds resident in HBM, plus device
    outputs for one batch (svs_dev_alloc buffers)."""

    def __init__(self, ctx, v_q, seeds, batch, k):
        import svs
        self.n = v_q.shape[0]
        self.v = svs.DeviceArray.from_host(ctx, v_q)
        self.s = svs.DeviceArray.from_host(ctx, seeds)
        self.ids = svs.DeviceArray(ctx, (batch, k), np.uint32)
        self.sc = svs.DeviceArray(ctx, (batch, k), np.int64)
        self.cnt = svs.DeviceArray(ctx, (batch,), np.uint32)
        self.st = svs.DeviceArray(ctx, (batch,), np.int32)

    def results(self, nq):
        return (self.ids.to_host()[:nq], self.sc.to_host()[:nq], self.cnt.to_host()[:nq],
                self.st.to_host()[:nq])


def run_batches_dev(index, dq, probes, verify, k, batch, nbatches, offset=0):
    """Device-resident batches: per-batch wall seconds; results of the last."""
    nprobe = probes.shape[1]
    lat = []
    for b in range(nbatches):
        q0 = offset + b * batch
        nq = min(batch, dq.n - q0)
        pp = np.arange(0, nq * nprobe + 1, nprobe, dtype=np.uint32)
        t0 = time.perf_counter()
        index.score_batch_dev(nq, dq.v.ptr(q0 * 1024 * 8), dq.s.ptr(q0 * 32), pp,
                              probes[q0:q0 + nq].reshape(-1), verify[q0:q0 + nq], k,
                              dq.ids.ptr(), dq.sc.ptr(), dq.cnt.ptr(), dq.st.ptr())
        lat.append(time.perf_counter() - t0)
    return lat


def schemes(ctx, seconds=1.0):
    """The other schemes on the kernel family, device-resident inputs and
    outputs: Tiptoe's SimplePIR answer (32768 x 8192 u32 database, entries
    in [0, 32), mod 2^32) at batch 1 and 64, and EMVP's partials GEMV
    (20,000 rows x 1292, batch 1 and 16; also with host buffers)."""
    import ctypes as C
    import svs
    from svs import _lib as L
    from svs import schemes as S
    rng = np.random.default_rng(11)
    out = {}
    rows, cols = 32768, 8192
    db = rng.integers(0, 32, size=(rows, cols), dtype=np.uint32)
    tt = S.TiptoeDatabase(db, ctx)
    stored, elem = tt.stored()
    del db
    res = {}
    for b in (1, 64):
        ct = svs.DeviceArray.from_host(ctx, rng.integers(0, 2**32, size=(b, cols), dtype=np.uint32))
        ans = svs.DeviceArray(ctx, (b, rows), np.uint32)
        one = np.zeros(1, np.uint32)

        def run(n):
            for _ in range(n):
                L.check(L.lib().svs_tiptoe_answer_u32_dev(tt.h, b, C.c_void_p(ct.ptr()),
                                                          C.c_void_p(ans.ptr()), -1))
            L.check(L.lib().svs_memcpy(ctx.h, one.ctypes.data_as(C.c_void_p),
                                       C.c_void_p(ans.ptr()), 4))
        run(3)
        t0 = time.perf_counter()
        n = 0
        while time.perf_counter() - t0 < seconds / 2:
            run(10)
            n += 10
        dt = (time.perf_counter() - t0) / n
        res[str(b)] = dict(ms_per_batch=round(dt * 1e3, 4), queries_per_s=round(b / dt, 1),
                           u32_layout_gbs=round(rows * cols * 4 / dt / 1e9, 1),
                           stored_gbs=round(stored / dt / 1e9, 1))
        ct.close()
        ans.close()
    tt.close()
    out["tiptoe_simplepir_32768x8192_mod2^32"] = dict(
        stored_bytes=int(stored), stored_entry_bytes=int(elem), batches=res,
        note="algorithmic bytes per batch 1.07 GB in the u32 layout; the database is held "
             "one byte per entry (all entries < 128)")
    m = 20000
    em = S.EmvpCluster(rng.integers(0, (1 << 61) - 1, size=(m, 1292), dtype=np.uint64),
                                 ctx)
    eres = {}
    for b in (1, 16):
        q = rng.integers(0, (1 << 61) - 1, size=(b, 1292), dtype=np.uint64)
        qd = svs.DeviceArray.from_host(ctx, q)
        pd = svs.DeviceArray(ctx, (b, 76 * m), np.uint64)
        one = np.zeros(1, np.uint64)

        def run_dev(n):
            for _ in range(n):
                em.compute_products_dev(b, qd.ptr(), pd.ptr())
            L.check(L.lib().svs_memcpy(ctx.h, one.ctypes.data_as(C.c_void_p),
                                       C.c_void_p(pd.ptr()), 8))
        run_dev(3)
        t0 = time.perf_counter()
        n = 0
        while time.perf_counter() - t0 < seconds / 2:
            run_dev(20)
            n += 20
        dt = (time.perf_counter() - t0) / n
        # the client decode on the device (R replayed per row) from the
        # resident partials, and the fused product + decode with host buffers
        sd = svs.DeviceArray(ctx, (b, m), np.float32)
        r_seed = bytes(range(32))

        def run_decode(n):
            for _ in range(n):
                em.decode_scores_dev(b, r_seed, qd.ptr(), pd.ptr(), sd.ptr())
            L.check(L.lib().svs_memcpy(ctx.h, one.ctypes.data_as(C.c_void_p),
                                       C.c_void_p(sd.ptr()), 4))
        run_decode(2)
        t0 = time.perf_counter()
        nd = 0
        while time.perf_counter() - t0 < seconds / 4:
            run_decode(5)
            nd += 5
        dtd = (time.perf_counter() - t0) / nd
        sd.close()
        em.score_batch(r_seed, q)
        t0 = time.perf_counter()
        nf = 0
        while time.perf_counter() - t0 < seconds / 4:
            em.score_batch(r_seed, q)
            nf += 1
        dtf = (time.perf_counter() - t0) / nf
        em.compute_products_batch(q)
        t0 = time.perf_counter()
        nh = 0
        while time.perf_counter() - t0 < seconds / 2:
            em.compute_products_batch(q)
            nh += 1
        dth = (time.perf_counter() - t0) / nh
        eres[str(b)] = dict(ms_per_batch=round(dt * 1e3, 4), queries_per_s=round(b / dt, 1),
                            gbs=round(m * 1292 * 8 / dt / 1e9, 1),
                            host_buffers_ms_per_batch=round(dth * 1e3, 3),
                            decode_ms_per_batch=round(dtd * 1e3, 4),
                            score_host_buffers_ms_per_batch=round(dtf * 1e3, 3))
        qd.close()
        pd.close()
    em.close()
    # Tiptoe's BFV hint compression: the hint of the 32768-row database
    # (n_lwe = 2048, 16 chunks, 4 GiB of NTT-domain limb plaintexts), one
    # apply_hint per query token, device-resident tokens and answers
    r, n_lwe = 32768, 2048
    hint = rng.integers(0, 2 ** 64, size=(r, n_lwe), dtype=np.uint64)
    t0 = time.perf_counter()
    th = S.TiptoeHint(hint, r, n_lwe, ctx)
    build_s = time.perf_counter() - t0
    del hint
    hres = {}
    for b in (1, 4):
        ct = svs.DeviceArray.from_host(ctx, rng.integers(0, th.q, size=(b, n_lwe, 2, 2048),
                                                          dtype=np.uint64))
        ans = svs.DeviceArray(ctx, (b, 8, th.chunks, 2, 2048), np.uint64)
        one = np.zeros(1, np.uint64)

        def run_hint(n):
            for _ in range(n):
                th.apply_hint_dev(b, ct.ptr(), ans.ptr())
            L.check(L.lib().svs_memcpy(ctx.h, one.ctypes.data_as(C.c_void_p),
                                       C.c_void_p(ans.ptr()), 8))
        run_hint(2)
        t0 = time.perf_counter()
        n = 0
        while time.perf_counter() - t0 < seconds / 2:
            run_hint(2)
            n += 2
        dt = (time.perf_counter() - t0) / n
        moved = th.stored_bytes + b * n_lwe * 2 * 2048 * 8 * 8  # tokens re-read per group
        hres[str(b)] = dict(ms_per_batch=round(dt * 1e3, 3), tokens_per_s=round(b / dt, 1),
                            plaintext_gbs=round(th.stored_bytes / dt / 1e9, 1),
                            est_traffic_gbs=round(moved / dt / 1e9, 1))
        ct.close()
        ans.close()
    out["tiptoe_apply_hint_32768x2048"] = dict(
        stored_bytes=int(th.stored_bytes), build_s=round(build_s, 2), batches=hres,
        note="decompose_hint on the device at build; per token the 4 GiB of limb plaintexts "
             "are read once per batch of tokens and each token's 64 MB of ciphertexts once "
             "per 16 (limb, chunk) pairs")
    th.close()
    out["emvp_partials_20000x1292"] = dict(
        batches=eres, note="device-resident queries and partials; algorithmic bytes per batch "
        "m x 1292 x 8 = 207 MB (m_hat read once); host_buffers: each query's 76 x m partials "
        "(12 MB) come back over PCIe; decode: the client's decode_scores on the device (R "
        "replayed from ChaCha20 per row, 162 blocks a row, shared by the batch); "
        "score_host_buffers: product + decode fused, q_hat in and f32 scores out")
    return out


def traffic_profile(args):
    """The committed counter summary of this workload, if there is one."""
    if not (args.n == 1_000_000 and args.nprobe == 32 and args.batch == 256 and
            args.nlist == 1024):
        return None
    tag = "bn_ivf_kmeans_1M_nprobe32_b256" if args.lists == "kmeans" else \
        "bn_ivf_1M_nprobe32_b256"
    return f"profiles/hbm_traffic_{tag}.json"


def measured_traffic(kernel, args):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE / WRITE_SIZE scaled by the copy-probe factors), when
    it was taken on this workload; else None."""
    rel = traffic_profile(args)
    if rel is None or not os.path.exists(os.path.join(ROOT, rel)):
        return None, None
    with open(os.path.join(ROOT, rel)) as f:
        k = json.load(f)["kernels"].get(kernel)
    if not k:
        return None, None
    return k["traffic_bytes"], rel


def served(index, qs, v_q, probes, verify, k, levels, seconds, max_batch, slots, pool, raw,
           seed=5):
    """Open-loop Poisson arrivals through the continuous-batching server at
    each offered rate; per level: achieved q/s and latency percentiles
    (scheduled arrival -> result). raw: requests carry the f32 query and
    nprobe (quantised and routed on the device inside each batch); else the
    quantised query and its probe list."""
    import svs
    rng = np.random.default_rng(seed)
    out = []
    nprobe = probes.shape[1]
    sizes = np.array([index.rows(c) for c in range(index.n_lists)], np.int64)
    # raw batches are sized for their bound (the nprobe largest lists)
    per_q = np.sort(sizes)[::-1][:nprobe].sum() if raw else None
    rows = np.sort(sizes[probes[:pool]].sum(axis=1))[::-1]
    cand = int(per_q) * max_batch if raw else int(rows[:max_batch].sum())

    def req(i):
        base = dict(q_seed=rng.integers(0, 256, 32, np.uint8), verify=bool(verify[i]), k=k)
        if raw:
            base.update(query=qs[i], nprobe=nprobe)
        else:
            base.update(v_q=v_q[i], probes=probes[i])
        return base
    with svs.Server(index, max_batch=max_batch, slots=slots) as srv:
        # size the slots for the largest possible batch of this pool, then warm up
        srv.reserve(max_batch, max_batch * nprobe, cand, k)
        warm = [req(i) for i in range(min(pool, 512))]
        srv.run_poisson(warm, qps=levels[0], seed=seed)
        for qps in levels:
            n = max(64, int(qps * seconds))
            reqs = [req(j % pool) for j in range(n)]
            s0 = srv.stats()
            res, wall = srv.run_poisson(reqs, qps=qps, seed=seed + int(qps))
            s1 = srv.stats()
            lat = [r.latency_ms for r in res]
            nb = s1["batches"] - s0["batches"]
            lvl = dict(offered_qps=qps, achieved_qps=round(n / wall, 1), requests=n,
                       p50_ms=round(pct(lat, 50), 3), p99_ms=round(pct(lat, 99), 3),
                       p999_ms=round(pct(lat, 99.9), 3),
                       mean_batch=round((s1["batched_queries"] - s0["batched_queries"]) /
                                        max(nb, 1), 1), max_in_flight=s1["max_in_flight"])
            log(f"served: {lvl}")
            out.append(lvl)
    return out


def served_at_slo(levels, slo_ms, raw_plain_qps, plain_qps):
    """The highest achieved served rate whose p99 stays under the SLO, and
    its ratio to the plain one-query paths (raw queries, as served, and
    pre-routed)."""
    ok = [x for x in (levels or []) if x["p99_ms"] < slo_ms]
    if not ok:
        return None
    best = max(ok, key=lambda x: x["achieved_qps"])
    return dict(slo_ms=slo_ms, achieved_qps=best["achieved_qps"],
                offered_qps=best["offered_qps"], p50_ms=best["p50_ms"], p99_ms=best["p99_ms"],
                p999_ms=best.get("p999_ms"), slots=best.get("slots"),
                max_in_flight=best.get("max_in_flight"), mean_batch=best.get("mean_batch"),
                vs_raw_plain=round(best["achieved_qps"] / raw_plain_qps, 2),
                vs_plain=round(best["achieved_qps"] / plain_qps, 2))


def served_mix(index, qs, k, qps, seconds, max_batch, slots, seed=9):
    """One open-loop level with the serving request mix of the reference's
    served configuration: raw queries, nprobe uniform in {8, 16, 32, 64,
    128}, half of them verified."""
    import svs
    rng = np.random.default_rng(seed)
    n = max(64, int(qps * seconds))
    choices = np.array([8, 16, 32, 64, 128])

    def req(i):
        return dict(query=qs[i % qs.shape[0]], nprobe=int(rng.choice(choices)),
                    q_seed=rng.integers(0, 256, 32, np.uint8), verify=bool(rng.integers(0, 2)),
                    k=k)
    with svs.Server(index, max_batch=max_batch, slots=slots) as srv:
        srv.run_poisson([req(i) for i in range(512)], qps=qps, seed=seed)  # sizes the slot
        reqs = [req(i) for i in range(n)]
        s0 = srv.stats()
        res, wall = srv.run_poisson(reqs, qps=qps, seed=seed + 1)
        s1 = srv.stats()
    lat = [r.latency_ms for r in res]
    nb = s1["batches"] - s0["batches"]
    out = dict(offered_qps=qps, achieved_qps=round(n / wall, 1), requests=n,
               p50_ms=round(pct(lat, 50), 3), p99_ms=round(pct(lat, 99), 3),
               mean_batch=round((s1["batched_queries"] - s0["batched_queries"]) / max(nb, 1), 1),
               failed=int(sum(1 for r in res if r.status != 0)),
               mix="raw queries, nprobe uniform in {8,16,32,64,128}, 50% verified")
    log(f"served mix: {out}")
    return out


def served_tenants(args, ctx=None):
    """The served configuration at full scale: `tenants` BN-IVF sub-indices
    of `tenant_rows` rows each (d = 768, `tenant_nlist` Zipf(1.1)-weighted
    mixture lists, own key seed per tenant) resident in HBM together, raw
    f32 queries arriving open-loop (Poisson) from uniformly drawn tenants,
    nprobe uniform in {8, 16, 32, 64, 128}, half of them verified. Returns
    (one dict), per offered rate, the achieved q/s, p50/p99/p99.9 latency
    (scheduled arrival -> result) and mean batch, beside the plain
    one-query-per-launch rate of the same request mix."""
    import svs
    from svs import data
    t_all = time.perf_counter()
    ctx = ctx or svs.Context(0)
    scorer = svs.BnTmIvfScorer(ctx)
    handles, pools = [], []
    for t in range(args.served_tenants):
        corpus = data.make_corpus(args.tenant_rows, 768, components=args.tenant_nlist,
                                  zipf_s=1.1, seed=1000 + t)
        lists = data.ivf_lists(corpus)
        cfg = svs.BnTmIvfConfig(index=svs.IvfIndex(corpus.centres, lists), nprobe=8,
                                key_seed=bytes([t + 1] * 32), upload_seed=100 + t)
        h, _ = scorer.upload_cluster(cfg, corpus.vectors)
        handles.append(h)
        pools.append(data.make_queries(corpus, 1024, seed=2000 + t))
        del corpus
        log(f"tenant {t}: {args.tenant_rows} rows in {time.perf_counter() - t_all:.0f} s; "
            f"resident {ctx.stats()['resident_bytes'] / 1e9:.1f} GB")
    st = ctx.stats()
    rng = np.random.default_rng(17)
    choices = np.array([8, 16, 32, 64, 128])

    def req(i):
        t = int(rng.integers(0, len(handles)))
        return dict(tenant=t, query=pools[t][int(rng.integers(0, 1024))],
                    nprobe=int(rng.choice(choices)), q_seed=rng.integers(0, 256, 32, np.uint8),
                    verify=bool(rng.integers(0, 2)), k=args.k)
    # plain: the same kind of requests, one query per launch sequence
    plain = [req(i) for i in range(args.tenant_plain_queries)]
    for r in plain[:4]:
        handles[r["tenant"]].index.score_raw_batch(r["query"][None, :], r["q_seed"][None, :],
                                                   r["nprobe"], np.array([r["verify"]], np.uint8),
                                                   args.k, mode=0)
    t0 = time.perf_counter()
    for r in plain:
        handles[r["tenant"]].index.score_raw_batch(r["query"][None, :], r["q_seed"][None, :],
                                                   r["nprobe"], np.array([r["verify"]], np.uint8),
                                                   args.k, mode=0)
    plain_qps = len(plain) / (time.perf_counter() - t0)
    log(f"plain (one query per launch sequence): {plain_qps:.0f} q/s")
    # one query alone, per nprobe: the plain sequence and a batch of one
    alone = {}
    for npb in choices.tolist():
        r0 = [req(i) for i in range(24)]
        row = {}
        for mode, name in ((0, "plain_ms"), (1, "batch_of_one_ms")):
            ts = []
            for r in r0:
                t1 = time.perf_counter()
                handles[r["tenant"]].index.score_raw_batch(
                    r["query"][None, :], r["q_seed"][None, :], npb,
                    np.array([r["verify"]], np.uint8), args.k, mode=mode)
                ts.append(time.perf_counter() - t1)
            row[name] = round(float(np.median(ts[4:])) * 1e3, 3)
        alone[str(npb)] = row
    log(f"one query alone by nprobe: {alone}")
    levels = []
    configs = [(int(x), int(c)) for c in args.tenant_cross.split(",") if c
               for x in args.tenant_slots.split(",") if x]
    for slots, cross in configs:
        with svs.Server([h.index for h in handles], max_batch=args.tenant_max_batch,
                        slots=slots, cross_tenant=bool(cross)) as srv:
            srv.run_poisson([req(i) for i in range(2048)], qps=2000, seed=3)  # sizes the slots
            for qps in [float(x) for x in args.tenant_qps.split(",") if x]:
                n = max(64, int(qps * args.serve_seconds))
                reqs = [req(i) for i in range(n)]
                s0 = srv.stats()
                res, wall = srv.run_poisson(reqs, qps=qps, seed=int(qps))
                s1 = srv.stats()
                lat = [r.latency_ms for r in res]
                nb = s1["batches"] - s0["batches"]
                lvl = dict(slots=slots, cross_tenant=bool(cross), offered_qps=qps,
                           achieved_qps=round(n / wall, 1),
                           requests=n, p50_ms=round(pct(lat, 50), 3),
                           p99_ms=round(pct(lat, 99), 3), p999_ms=round(pct(lat, 99.9), 3),
                           mean_batch=round((s1["batched_queries"] - s0["batched_queries"]) /
                                            max(nb, 1), 1),
                           failed=int(sum(1 for r in res if r.status != 0)))
                log(f"served tenants: {lvl}")
                levels.append(lvl)
    rows = args.served_tenants * args.tenant_rows
    line = dict(metric="BN-IVF served queries/s + p50/p99 ms, tenant sub-indices resident in HBM",
                unit="queries/s", n_gpus=1,
                value=max((x["achieved_qps"] for x in levels
                           if x["p99_ms"] < args.slo_ms and x["failed"] == 0), default=None),
                higher_is_better=True, dtype="u64 (F_p, p = 2^61-1)",
                data="synthetic: per tenant a seeded Zipf(1.1) Gaussian mixture, d=768, unit-norm",
                config=dict(workload=f"served bn-ivf {args.served_tenants} tenants x "
                                     f"{args.tenant_rows} rows (N={rows}) d=768 nlist="
                                     f"{args.tenant_nlist} per tenant, raw queries, nprobe "
                                     "uniform in {8,16,32,64,128}, 50% verified, k="
                                     f"{args.k}", max_batch=args.tenant_max_batch),
                resident_bytes=st["resident_bytes"], plain_qps=round(plain_qps, 1),
                alone_by_nprobe=alone,
                levels=levels, slo_ms=args.slo_ms,
                setup_s=round(time.perf_counter() - t_all, 1))
    best = [x for x in levels if x["p99_ms"] < args.slo_ms and x["failed"] == 0]
    line["served_at_slo"] = None if not best else dict(
        max(best, key=lambda x: x["achieved_qps"]),
        vs_plain=round(max(x["achieved_qps"] for x in best) / plain_qps, 2))
    for h in handles:
        h.index.close()
    return line


def flat_plain(args, ctx=None):
    """The plain GPU drop-in at full size: BN flat over N = `flat_n` rows
    (d = 768, a seeded 1024-component mixture with equal weights), one query
    per launch sequence, verification off and on. Reports the server
    product alone (compute_products: M_enc . v_enc for one v_enc, device
    pointers, the 8.19 GB M_enc read once per call) against HBM, the whole
    query (encode, product, [verify], decode, top-k) as q/s and p50/p99, the
    stage split, and the zero-noise check of the returned scores (one dict)."""
    import ctypes as C
    import svs
    from svs import _lib as L
    from svs import data
    t_all = time.perf_counter()
    corpus = data.make_corpus(args.flat_n, 768, components=1024, zipf_s=None, seed=42)
    ctx = ctx or svs.Context(0)
    handle, _ = svs.BnTmScorer(ctx).upload_cluster(
        svs.BnTmConfig(key_seed=bytes([42] * 32)), corpus.vectors, h_seed=bytes([7] * 32))
    idx = handle.index
    m = idx.rows(0)
    log(f"flat index of {m} rows built in {time.perf_counter() - t_all:.1f} s; "
        f"resident {ctx.stats()['resident_bytes'] / 1e9:.1f} GB")
    rng = np.random.default_rng(21)
    v = svs.DeviceArray.from_host(ctx, rng.integers(0, (1 << 61) - 1, 1024, dtype=np.uint64))
    r = svs.DeviceArray(ctx, (m,), np.uint64)
    one = np.zeros(1, np.uint64)

    def products(n):
        for _ in range(n):
            L.check(L.lib().svs_compute_products_dev(idx.h, 0, C.c_void_p(v.ptr()),
                                                     C.c_void_p(r.ptr())))
        L.check(L.lib().svs_memcpy(ctx.h, one.ctypes.data_as(C.c_void_p), C.c_void_p(r.ptr()), 8))
    products(3)
    t0 = time.perf_counter()
    products(args.flat_product_calls)
    prod_ms = (time.perf_counter() - t0) / args.flat_product_calls * 1e3
    prod_bytes = m * 1024 * 8
    log(f"product alone: {prod_ms:.3f} ms per call, {prod_bytes / prod_ms / 1e6:.0f} GB/s")
    v.close()
    r.close()
    qs = data.make_queries(corpus, args.flat_queries, seed=4242)
    v_q = svs.quantise(qs)
    seeds = rng.integers(0, 256, (qs.shape[0], 32), dtype=np.uint8)
    pp = np.array([0, 1], np.uint32)
    pl = np.zeros(1, np.uint32)
    legs = {}
    for ver in (0, 1):
        vv = np.array([ver], np.uint8)
        for q in range(3):
            idx.score_batch(v_q[q:q + 1], seeds[q:q + 1], pp, pl, vv, args.k, 0)
        ctx.profile(True, True)
        lat, res = [], []
        for q in range(qs.shape[0]):
            t0 = time.perf_counter()
            out = idx.score_batch(v_q[q:q + 1], seeds[q:q + 1], pp, pl, vv, args.k, 0)
            lat.append(time.perf_counter() - t0)
            res.append(out)
        prof = ctx.profile(False, False)
        failed = sum(int(o[3][0] != 0) for o in res)
        stages = {k: round(prof[k][0] / max(prof[k][1], 1), 4) for k in L.STAGES}
        legs["verify_on" if ver else "verify_off"] = dict(
            queries=qs.shape[0], queries_per_s=round(len(lat) / sum(lat), 1),
            p50_ms=round(pct(lat, 50) * 1e3, 3), p99_ms=round(pct(lat, 99) * 1e3, 3),
            stage_ms=stages, failed=failed)
        log(f"flat plain verify={ver}: {legs['verify_on' if ver else 'verify_off']}")
    # zero noise: every returned score equals fp_to_signed(M_q . v_q)
    p = (1 << 61) - 1
    checked = wrong = 0
    for q in range(min(16, qs.shape[0])):
        ids, sc = res[q][0][0], res[q][1][0]
        for j in range(int(res[q][2][0])):
            mq = svs.quantise(corpus.vectors[int(ids[j]):int(ids[j]) + 1])[0]
            t = sum(a * b for a, b in zip(mq[:768].tolist(), v_q[q][:768].tolist())) % p
            wrong += int((t if t < p // 2 else t - p) != int(sc[j]))
            checked += 1
    line = dict(metric="BN flat plain GPU drop-in queries/s + p50/p99 ms, one query per launch",
                unit="queries/s", n_gpus=1, value=legs["verify_on"]["queries_per_s"],
                higher_is_better=True, dtype="u64 (F_p, p = 2^61-1)",
                data="synthetic: seeded 1024-component Gaussian mixture, d=768, unit-norm",
                config=dict(workload=f"bn flat N={m} d=768 k={args.k}, one query per launch "
                                     "sequence, verification off and on"),
                product_alone=dict(ms=round(prod_ms, 4), bytes=prod_bytes,
                                   gbs=round(prod_bytes / prod_ms / 1e6, 1),
                                   frac=round(prod_bytes / prod_ms / 1e6 / HBM_PEAK_GBS, 4),
                                   kernel="k_gemv_chunks"),
                hbm_ceiling_qps=round(HBM_PEAK_GBS * 1e9 / prod_bytes, 1),
                legs=legs, zero_noise_scores_checked=checked, zero_noise_mismatches=wrong,
                setup_s=round(time.perf_counter() - t_all, 1))
    idx.close()
    return line


def host_cpu():
    """The host the CPU baseline runs on: model, logical CPUs, the CPUs this
    process may use, and AVX-512F."""
    from oracle import oracle as O
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return dict(model=model, nproc=os.cpu_count(), usable_cpus=usable, avx512f=O.has_avx512f())


def cpu_baseline(index, corpus, centroids, lists, l_sub, qs, nprobe, k, budget_s, max_threads,
                 upload_seed=7):
    """The reference CPU path on this host, timed as the reference times it:
    the closed-loop B=1 protocol (eval.rs:2338-2344: one query, then the
    next) of BnTmIvfScorer::score (ivf.rs:746-822), verification off, with
    compute_products and decode_scores' matvecs row-parallel over the
    threads (rayon, crypto.rs:609-625). Two builds, as the reference has
    two: the scalar u128 matvec (crypto.rs:606-626) and the AVX-512
    lane-decomposed one (crypto.rs:643-739). Also the query-parallel rate
    (one query per thread). Encrypted state comes from the device build
    (downloaded), so the CPU scores the same ciphertexts."""
    from oracle import oracle as O
    import svs

    host = host_cpu()
    threads = max(1, min(max_threads, host["usable_cpus"]))
    pool = min(qs.shape[0], 256)
    probes = svs.probe_route(centroids, qs[:pool], nprobe)
    state = {}
    lock = threading.Lock()

    def load(ci):
        members = lists[ci]
        m = members.size
        me, al, hh = index.download_list(ci)
        mq = O.quantise_rows(corpus.vectors[members]).reshape(-1) if m else np.zeros(0, np.uint64)
        _, rowptr, sidx, ssign = O.regenerate_h_s(O.derive_h_seed(upload_seed, ci), m)
        with lock:
            state[ci] = dict(m=m, m_enc=me, a_l=al, m_plain=mq, h=hh, s_rowptr=rowptr,
                             s_idx=sidx, s_sign=ssign)

    todo = sorted(set(probes.reshape(-1).tolist()))
    for i in range(0, len(todo), threads):
        ts = [threading.Thread(target=load, args=(c,)) for c in todo[i:i + threads]]
        [t.start() for t in ts]
        [t.join() for t in ts]
    cpu = O.CpuIvf(state, lists, centroids, l_sub)
    seeds = [bytes([qi % 256]) * 32 for qi in range(pool)]

    def closed_loop(simd, seconds):
        lat, hits, qi = [], [], 0
        t_end = time.perf_counter() + seconds
        while qi < pool and (qi < 2 or time.perf_counter() < t_end):
            t0 = time.perf_counter()
            hits.append(cpu.score(qs[qi], nprobe, seeds[qi], k, threads, simd))
            lat.append(time.perf_counter() - t0)
            qi += 1
        return lat, hits

    legs = dict(scalar=0.4, avx512=0.4) if host["avx512f"] else dict(scalar=0.8)
    out = {}
    for name, frac in legs.items():
        lat, hits = closed_loop(name == "avx512", budget_s * frac)
        out[name] = dict(queries_per_s=round(len(lat) / sum(lat), 3), queries=len(lat),
                         p50_ms=round(pct(lat, 50) * 1e3, 2), p99_ms=round(pct(lat, 99) * 1e3, 2),
                         seconds=round(sum(lat), 2), hits=hits)
        log(f"cpu baseline ({name}, closed-loop B=1, {threads} threads): "
            f"{out[name]['queries_per_s']} q/s")
    same = True
    if "avx512" in out:
        n = min(len(out["scalar"]["hits"]), len(out["avx512"]["hits"]))
        same = all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
                   for a, b in zip(out["scalar"]["hits"][:n], out["avx512"]["hits"][:n]))
    # query-parallel: one query per thread, each on one thread
    done = []

    def worker(tid, n):
        for qi in range(tid, n, threads):
            cpu.score(qs[qi], nprobe, seeds[qi], k, 1, False)
            done.append(qi)
    nq = min(pool, threads * 4)
    t0 = time.perf_counter()
    ts = [threading.Thread(target=worker, args=(i, nq)) for i in range(threads)]
    [t.start() for t in ts]
    [t.join() for t in ts]
    qp = len(done) / (time.perf_counter() - t0)
    first = out["scalar"]
    rows = int(sum(sum(lists[c].size for c in probes[q]) for q in range(first["queries"])))
    for v in out.values():
        v.pop("hits")
    return dict(value=first["queries_per_s"], unit="queries/s", cores=threads, kind="port",
                sample=f"the first {first['queries']} queries of the timed workload (nprobe "
                       f"{nprobe}, {rows} probed rows), closed-loop B=1, whole query (route, "
                       "quantise, encode, product, decode, top-k), verification off, scalar "
                       "u128 matvec; C restatement of the Rust CPU path",
                p50_ms=first["p50_ms"], p99_ms=first["p99_ms"],
                avx512=out.get("avx512"), avx512_equals_scalar=bool(same),
                query_parallel=dict(queries_per_s=round(qp, 3), queries=nq,
                                    note="one query per thread, scalar"),
                host=dict(host, threads_used=threads,
                          note="threads = min(--cpu-threads, usable CPUs); 16 is one GPU's "
                               "share of the host's cores"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--nlist", type=int, default=1024)
    ap.add_argument("--nprobe", type=int, default=32)
    ap.add_argument("--lists", choices=("kmeans", "mixture"), default="kmeans",
                    help="IVF lists: a seeded k-means of the corpus, or its mixture components")
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--plain-queries", type=int, default=1000)
    ap.add_argument("--verify-frac", type=float, default=0.0)
    ap.add_argument("--cpu-seconds", type=float, default=16.0)
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="threads of the CPU baseline (capped at the usable CPUs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tile-variant", type=int, default=-1,
                    help="force a block-tile variant (skips the autotuner)")
    ap.add_argument("--serve-qps", default="20000,28000,32000,36000,40000,44000,48000,52000,54000,56000,58000,60000,62000,64000",
                    help="offered rates of the served (open-loop Poisson) levels; '' = skip")
    ap.add_argument("--serve-seconds", type=float, default=1.5)
    ap.add_argument("--serve-max-batch", type=int, default=512)
    ap.add_argument("--serve-slots", default="1,2",
                    help="batches in flight, one set of served levels per value")
    ap.add_argument("--serve-mix-qps", default="4000,8000,12000",
                    help="offered rates of the mixed-nprobe, half-verified served levels")
    ap.add_argument("--serve-routed", action="store_true",
                    help="served requests carry quantised queries and probe lists "
                         "(default: raw f32 queries routed on the device)")
    ap.add_argument("--slo-ms", type=float, default=10.0)
    ap.add_argument("--no-schemes", action="store_true",
                    help="skip the Tiptoe / EMVP measurements")
    ap.add_argument("--copy-probe-mib", type=int, default=1024,
                    help="size of the plain-copy calibration dispatch (0 = skip)")
    ap.add_argument("--served-tenants", type=int, default=16,
                    help="tenant sub-indices of the multi-tenant served configuration")
    ap.add_argument("--tenants-only", action="store_true",
                    help="run only the multi-tenant served configuration")
    ap.add_argument("--no-tenants", action="store_true",
                    help="skip the multi-tenant served configuration")
    ap.add_argument("--tenant-rows", type=int, default=550_000)
    ap.add_argument("--tenant-nlist", type=int, default=512)
    ap.add_argument("--tenant-qps", default="2000,4000,8000,12000")
    ap.add_argument("--tenant-plain-queries", type=int, default=200)
    ap.add_argument("--tenant-slots", default="1,2",
                    help="slot counts the multi-tenant served levels are run with")
    ap.add_argument("--tenant-cross", default="0",
                    help="batching of the multi-tenant levels: 0 = a batch holds one tenant, "
                         "1 = batches across tenants")
    ap.add_argument("--tenant-max-batch", type=int, default=512)
    ap.add_argument("--flat-n", type=int, default=1_000_000,
                    help="rows of the plain flat drop-in configuration")
    ap.add_argument("--flat-only", action="store_true",
                    help="run only the plain flat drop-in configuration")
    ap.add_argument("--no-flat", action="store_true",
                    help="skip the plain flat drop-in configuration")
    ap.add_argument("--flat-queries", type=int, default=1000)
    ap.add_argument("--flat-product-calls", type=int, default=50)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(lists would be bin-packed across GPUs with one top-k exchange per batch; "
              "only the one-GPU share is built and measured)", file=sys.stderr)
        sys.exit(2)

    if args.tenants_only:
        print(json.dumps(served_tenants(args)), flush=True)
        return
    if args.flat_only:
        print(json.dumps(flat_plain(args)), flush=True)
        return

    import svs
    from svs import data

    t_all = time.perf_counter()
    log(f"corpus N={args.n} nlist={args.nlist}")
    corpus = data.make_corpus(args.n, 768, components=args.nlist, zipf_s=1.1, seed=42)
    t0 = time.perf_counter()
    if args.lists == "kmeans":
        centroids, lists = data.kmeans_lists(corpus, args.nlist, seed=42, max_iter=25)
    else:
        centroids, lists = corpus.centres, data.ivf_lists(corpus)
    kmeans_s = time.perf_counter() - t0
    sizes = np.array([x.size for x in lists])
    log(f"{args.lists} lists in {kmeans_s:.1f} s: max {sizes.max()} min {sizes.min()} "
        f"mean {sizes.mean():.1f} empty {(sizes == 0).sum()} ({time.perf_counter() - t_all:.1f} s)")

    ctx = svs.Context(0)
    scorer = svs.BnTmIvfScorer(ctx)
    cfg = svs.BnTmIvfConfig(index=svs.IvfIndex(centroids, lists), nprobe=args.nprobe,
                            key_seed=bytes([42] * 32), upload_seed=7)
    t0 = time.perf_counter()
    handle, _ = scorer.upload_cluster(cfg, corpus.vectors)
    log(f"device index built in {time.perf_counter() - t0:.1f} s; {ctx.stats()}")

    total_q = (args.warmup + args.steps) * args.batch
    qs = data.make_queries(corpus, max(total_q, args.warmup * args.batch + args.plain_queries),
                           seed=4242)
    probes = svs.probe_route(centroids, qs, args.nprobe)
    v_q = svs.quantise(qs)
    rng = np.random.default_rng(7)
    seeds = rng.integers(0, 256, size=(qs.shape[0], 32), dtype=np.uint8)
    verify = (rng.random(qs.shape[0]) < args.verify_frac).astype(np.uint8)
    index = handle.index
    probed_rows = np.array([sizes[p].sum() for p in probes])
    probed_by_nprobe = {}
    for npb in (8, 32, 128):
        pr = sizes[svs.probe_route(centroids, qs[:2048], npb)].sum(axis=1)
        probed_by_nprobe[str(npb)] = dict(mean=round(float(pr.mean()), 1),
                                          p50=float(np.median(pr)), max=int(pr.max()))
    log(f"probed rows per query: {probed_by_nprobe}")

    # block-tile variant for this batch size: timed on the first warm-up batch
    tune = None
    if args.tile_variant >= 0:
        ctx.set_tile_variant(args.tile_variant)
    elif not args.no_autotune:
        # the timed batch size, plus the classes the server uses (batches
        # past 256 queries form when arrivals outrun the device)
        tune = {}
        for nb in sorted({args.batch, 16, 64, min(512, v_q.shape[0])}):
            s0 = slice(0, nb)
            pp0 = np.arange(0, v_q[s0].shape[0] * args.nprobe + 1, args.nprobe, dtype=np.uint32)
            best, tms = index.autotune(v_q[s0], seeds[s0], pp0, probes[s0].reshape(-1),
                                       verify[s0], args.k, reps=2)
            tune[str(nb)] = dict(best=best, ms_per_variant=[round(x, 3) for x in tms])
        log(f"autotune: {tune}")
    variant = ctx.tile_variant(args.batch)

    # batched path, query inputs and outputs resident in HBM
    dq = DeviceQueries(ctx, v_q, seeds, args.batch, args.k)
    run_batches_dev(index, dq, probes, verify, args.k, args.batch, args.warmup)
    off = args.warmup * args.batch
    # the timed region: each batch replays its captured HIP graph
    g0 = ctx.stats()
    lat = run_batches_dev(index, dq, probes, verify, args.k, args.batch, args.steps, offset=off)
    g1 = ctx.stats()
    wall = sum(lat)
    qps = args.steps * args.batch / wall
    graphs = dict(replayed=g1["graph_launches"] - g0["graph_launches"],
                  captured=g1["graph_captures"] - g0["graph_captures"])
    # the same batches again with HIP events around each stage (events are
    # recorded between launched kernels, so this pass runs without graphs):
    # the stage split and the product kernel's time for the roofline
    ctx.profile(True, True)
    lat_ev = run_batches_dev(index, dq, probes, verify, args.k, args.batch, args.steps, offset=off)
    prof = ctx.profile(False, False)
    qps_launch = args.steps * args.batch / sum(lat_ev)
    log(f"batched: {qps:.0f} q/s (graphs {graphs}), batch p50 {pct(lat, 50) * 1e3:.2f} ms; "
        f"launched with stage events {qps_launch:.0f} q/s; stages {prof}")
    # batch 0 of the timed region again, for the bit-exactness check below
    run_batches_dev(index, dq, probes, verify, args.k, args.batch, 1, offset=off)
    res_b = [dq.results(min(args.batch, v_q.shape[0] - off))]
    # the product at the served batch sizes (same queries, batches of 64):
    # fewer queries per list, so the tiles stream rows with less MFMA work
    by_batch = {}
    for nb in sorted({64, args.batch}):
        if nb > args.batch:
            continue
        nbat = max(1, min(args.steps * args.batch // nb, (v_q.shape[0] - off) // nb))
        run_batches_dev(index, dq, probes, verify, args.k, nb, 1, offset=off)
        ctx.profile(True, True)
        lat_b = run_batches_dev(index, dq, probes, verify, args.k, nb, nbat, offset=off)
        pb = ctx.profile(False, False)
        ms_b, calls_b = pb["product"]
        gbs = pb["gemv_bytes"] / (ms_b * 1e-3) / 1e9 if ms_b else 0.0
        by_batch[str(nb)] = dict(queries_per_s=round(nb * nbat / sum(lat_b), 1),
                                 product_ms=round(ms_b / max(calls_b, 1), 4),
                                 algorithmic_bytes_per_launch=int(pb["gemv_bytes"] /
                                                                  max(calls_b, 1)),
                                 achieved_gbs=round(gbs, 1),
                                 frac=round(gbs / HBM_PEAK_GBS, 4))
    log(f"product by batch size: {by_batch}")
    # the same batches through host buffers (PCIe-inclusive rate)
    lat_h, _ = run_batches(index, v_q, seeds, probes, verify, args.k, args.batch, args.steps, 1,
                           offset=off)
    qps_host = args.steps * args.batch / sum(lat_h)
    log(f"batched, host buffers: {qps_host:.0f} q/s")

    # plain path: one query per launch sequence, same queries as batch 0 of
    # the timed region -> also a bit-exactness check between the two modes
    npl = args.plain_queries
    ncmp = min(npl, args.batch)  # the plain queries that batch 0 also holds
    run_batches(index, v_q, seeds, probes, verify, args.k, 1, 2, 0, offset=off)
    lat_p, res_p = run_batches(index, v_q, seeds, probes, verify, args.k, 1, npl, 0, offset=off)
    qps_plain = npl / sum(lat_p)
    ids_b, sc_b = res_b[0][0][:ncmp], res_b[0][1][:ncmp]
    exact = all(np.array_equal(res_p[i][0][0], ids_b[i]) and np.array_equal(res_p[i][1][0], sc_b[i])
                for i in range(ncmp))
    log(f"plain: {qps_plain:.0f} q/s, p50 {pct(lat_p, 50) * 1e3:.2f} ms; batched==plain {exact}")

    # raw f32 queries: quantisation and routing on the device as well (the
    # reference's score does both on the client CPU), host buffers
    lat_r, res_r = run_raw(index, qs, seeds, args.nprobe, verify, args.k, args.batch,
                           args.steps, 1, offset=off)
    qps_raw = args.steps * args.batch / sum(lat_r)
    run_raw(index, qs, seeds, args.nprobe, verify, args.k, 1, 2, 0, offset=off)
    lat_rp, res_rp = run_raw(index, qs, seeds, args.nprobe, verify, args.k, 1, npl, 0,
                             offset=off)
    qps_raw_plain = npl / sum(lat_rp)
    exact_raw = np.array_equal(res_r[0][0][:ncmp], ids_b) and \
        np.array_equal(res_r[0][1][:ncmp], sc_b) and \
        all(np.array_equal(res_rp[i][0][0], ids_b[i]) for i in range(ncmp))
    log(f"raw batched: {qps_raw:.0f} q/s; raw plain: {qps_raw_plain:.0f} q/s; "
        f"raw == routed {exact_raw}")

    levels = [float(x) for x in args.serve_qps.split(",") if x.strip()]
    slot_counts = [int(x) for x in args.serve_slots.split(",") if x.strip()]
    serve = []
    for slots in (slot_counts if levels else []):
        for lvl in served(index, qs, v_q, probes, verify, args.k, levels, args.serve_seconds,
                          args.serve_max_batch, slots, qs.shape[0], not args.serve_routed):
            serve.append(dict(lvl, slots=slots))

    mix = []
    for qps_mix in [float(x) for x in args.serve_mix_qps.split(",") if x.strip()]:
        for slots in slot_counts:
            mix.append(dict(served_mix(index, qs, args.k, qps_mix, args.serve_seconds,
                                       args.serve_max_batch, slots), slots=slots))

    gemv_ms, gemv_calls = prof["product"]
    achieved = prof["gemv_bytes"] / (gemv_ms * 1e-3) / 1e9 if gemv_ms else 0.0
    rows = prof["gemv_rows"]
    mulmods = rows * 1024
    qw = variant["query_group"] if variant["query_group"] in (4, 16) else 8
    kname = "k_prod_planes_lb<1024,0,2>" if qw == 16 else f"k_prod_planes<{qw},4,1>"
    traffic, traffic_src = measured_traffic(kname, args)
    # useful int8 work of the byte-plane product: 1024 elements x 8 x 8 limb
    # byte products per (row, query) pair, 2 ops per MAC, against 5.0 POP/s
    ops_pair = 1024 * 64 * 2
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                    traffic_source=traffic_src,
                    kernel=kname,
                    algorithmic_bytes_per_launch=int(prof["gemv_bytes"] / max(gemv_calls, 1)),
                    avg_launch_ms=round(gemv_ms / max(gemv_calls, 1), 4),
                    mulmods_per_s=round(mulmods / (gemv_ms * 1e-3), 1) if gemv_ms else 0,
                    mfma_tops=round(rows * ops_pair / (gemv_ms * 1e-3) / 1e12, 1)
                    if gemv_ms else 0,
                    mfma_frac=round(rows * ops_pair / (gemv_ms * 1e-3) / 5.0e15, 4)
                    if gemv_ms else 0,
                    measured="HIP events on the library's stream around the product stage "
                             "(k_query_corr, k_query_planes and the product kernel), over a "
                             "second pass of the timed batches")

    copy = None
    if args.copy_probe_mib:
        nb = args.copy_probe_mib << 20
        ctx.copy_probe(nb)
        ms = min(ctx.copy_probe(nb) for _ in range(3))
        copy = dict(bytes=nb, ms=round(ms, 4), gbs=round(2 * nb / (ms * 1e-3) / 1e9, 1),
                    note="k_copy_probe reads and writes `bytes`; gbs counts both")
        ctx.read_probe(nb)
        rms = min(ctx.read_probe(nb) for _ in range(3))
        copy["read_probe"] = dict(bytes=nb, ms=round(rms, 4),
                                  gbs=round(nb / (rms * 1e-3) / 1e9, 1),
                                  note="k_read_probe: reads only, in the product's A pattern")
        log(f"copy probe: {copy}")

    other = None if args.no_schemes else schemes(ctx)
    if other:
        log(f"schemes: {other}")

    # the verified-query cost: the same batches with half of the queries
    # verified (Protocol 2, three Freivalds trials per probed list)
    ver_half = (np.random.default_rng(8).random(qs.shape[0]) < 0.5).astype(np.uint8)
    run_batches_dev(index, dq, probes, ver_half, args.k, args.batch, 2, offset=off)
    lat_v = run_batches_dev(index, dq, probes, ver_half, args.k, args.batch, args.steps, offset=off)
    ctx.profile(True, True)
    run_batches_dev(index, dq, probes, ver_half, args.k, args.batch, args.steps, offset=off)
    prof_v = ctx.profile(False, False)
    verified = dict(verify_frac=0.5, queries_per_s=round(args.steps * args.batch / sum(lat_v), 1),
                    stages_ms_per_batch={s_: round(prof_v[s_][0] / max(prof_v[s_][1], 1), 4)
                                         for s_ in ("encode", "product", "verify", "decode",
                                                    "topk")})
    log(f"batched, half verified: {verified}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(index, corpus, centroids, lists, handle.l_subspace, qs[off:],
                           args.nprobe, args.k,
                           args.cpu_seconds, args.cpu_threads)
    dq_close = [dq.v, dq.s, dq.ids, dq.sc, dq.cnt, dq.st]
    for a in dq_close:
        a.close()
    index.close()
    del corpus

    flat = None if args.no_flat else flat_plain(args, ctx)
    tenants = None if args.no_tenants else served_tenants(args, ctx)

    slo = served_at_slo(serve, args.slo_ms, qps_raw_plain, qps_plain)
    out = {
        "metric": "BN-IVF queries/s + p50/p99 ms, served-batched vs plain GEMV, 1 MI355X, bit-exact",
        "value": slo["achieved_qps"] if slo else 0.0,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (F_p, p = 2^61-1)",
        "data": "synthetic: seeded 1024-component Gaussian mixture (Zipf 1.1 weights), d=768, "
                "unit-norm; " + ("lists = seeded k-means (k-means++ init, Lloyd)"
                                 if args.lists == "kmeans" else "lists = mixture components"),
        "config": {"workload": f"bn-ivf N={args.n} d=768 nlist={args.nlist} nprobe={args.nprobe} "
                               f"k={args.k}, served: open-loop Poisson arrivals of raw f32 "
                               f"queries, continuous batching up to {args.serve_max_batch}, "
                               f"p99 < {args.slo_ms:g} ms",
                   "nprobe": args.nprobe, "mean_probed_rows": round(float(probed_rows.mean()), 1),
                   "lists": args.lists, "list_sizes": dict(max=int(sizes.max()),
                                                            min=int(sizes.min()),
                                                            p50=float(np.median(sizes)),
                                                            empty=int((sizes == 0).sum())),
                   "probed_rows_by_nprobe": probed_by_nprobe, "kmeans_s": round(kmeans_s, 1),
                   "closed_loop_batch": args.batch, "verify_frac": args.verify_frac},
        "value_is": "served_at_slo.achieved_qps: the highest served rate (over the offered "
                    "levels and slot counts) whose p99 latency, scheduled arrival to result, "
                    "stays under the SLO; steps/ms_per_step are the closed-loop batched steps",
        "p50_ms": slo["p50_ms"] if slo else None,
        "p99_ms": slo["p99_ms"] if slo else None,
        "served_at_slo": slo,
        "served": serve,
        "served_mix": mix,
        "served_requests": "routed (v_q + probes)" if args.serve_routed else
                           "raw f32 queries, quantised and routed on the device",
        "plain": {"value": round(qps_plain, 2), "unit": "queries/s",
                  "p50_ms": round(pct(lat_p, 50) * 1e3, 3),
                  "p99_ms": round(pct(lat_p, 99) * 1e3, 3), "queries": npl,
                  "note": "one query per launch sequence, quantised query and probe list given"},
        "batched_closed_loop": {"queries_per_s": round(qps, 2), "batch": args.batch,
                                "p50_batch_ms": round(pct(lat, 50) * 1e3, 3),
                                "p99_batch_ms": round(pct(lat, 99) * 1e3, 3),
                                "graphs": graphs,
                                "launched_with_stage_events_qps": round(qps_launch, 2),
                                "host_buffers_qps": round(qps_host, 2),
                                "speedup_vs_plain": round(qps / qps_plain, 2),
                                "note": "device-resident quantised queries and probe lists, one "
                                        "batch after the other"},
        "batched_half_verified": verified,
        "raw_queries": {"batched_host_buffers_qps": round(qps_raw, 2),
                        "plain_qps": round(qps_raw_plain, 2),
                        "plain_p50_ms": round(pct(lat_rp, 50) * 1e3, 3),
                        "plain_p99_ms": round(pct(lat_rp, 99) * 1e3, 3),
                        "bit_exact_vs_routed": bool(exact_raw),
                        "note": "f32 queries in host buffers; quantise_vector and probe_route "
                                "run on the device inside the batch"},
        "bit_exact_batched_vs_plain": bool(exact),
        "bit_exact_compared_queries": ncmp,
        "stages_ms_per_batch": {s_: round(prof[s_][0] / max(prof[s_][1], 1), 4)
                                for s_ in ("encode", "product", "verify", "decode", "topk")},
        "tile_variant": variant,
        "autotune": tune,
        "roofline": roofline,
        "roofline_by_batch": by_batch,
        "flat_plain": flat,
        "served_tenants": tenants,
        "schemes": other,
        "copy_probe": copy,
        "cpu_baseline": cpu,
        "wall_s": round(time.perf_counter() - t_all, 1),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
