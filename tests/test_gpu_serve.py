"""Serving (continuous batching): every request's result is the same as the
same request run alone, bit-exact, whatever it was batched with; also
against the oracle (BnTmIvfScorer::score, scorer-bntm/src/ivf.rs:746-822).
"""
import sys
import threading

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import svs
    return svs.Context(0)


def _tenant(ctx, seed, n=2500, comps=16, key=42, upload_seed=7):
    import svs
    from svs import data
    corpus = data.make_corpus(n, components=comps, zipf_s=1.1, seed=seed)
    lists = data.ivf_lists(corpus)
    ivf = svs.IvfIndex(corpus.centres, lists)
    cfg = svs.BnTmIvfConfig(index=ivf, nprobe=4, key_seed=bytes([key] * 32),
                            upload_seed=upload_seed, verification_enabled=True)
    handle, _ = svs.BnTmIvfScorer(ctx).upload_cluster(cfg, corpus.vectors)
    qs = data.make_queries(corpus, 48, seed=seed + 1)
    return dict(corpus=corpus, lists=lists, handle=handle, qs=qs,
                l=O.generate_subspace(bytes([key] * 32)), upload_seed=upload_seed)


@pytest.fixture(scope="module")
def tenant_a(ctx):
    return _tenant(ctx, 31)


@pytest.fixture(scope="module")
def tenant_b(ctx):
    return _tenant(ctx, 77, n=1800, comps=12, key=5, upload_seed=3)


def _requests(t, rng, n, tenant=0):
    import svs
    v_q = svs.quantise(t["qs"])
    out = []
    for i in range(n):
        qi = int(rng.integers(0, t["qs"].shape[0]))
        nprobe = int(rng.choice([1, 2, 4, 8]))
        probes = svs.probe_route(t["corpus"].centres, t["qs"][qi:qi + 1], nprobe)[0]
        out.append(dict(tenant=tenant, v_q=v_q[qi], q_seed=rng.integers(0, 256, 32, np.uint8),
                        probes=probes, verify=bool(rng.integers(0, 2)),
                        k=int(rng.choice([1, 5, 10, 16, 40])), qi=qi))
    return out


def _alone(t, rq):
    idx = t["handle"].index
    pp = np.array([0, len(rq["probes"])], np.uint32)
    ids, scores, count, status = idx.score_batch(rq["v_q"][None, :], rq["q_seed"][None, :], pp,
                                                 rq["probes"], np.array([rq["verify"]], np.uint8),
                                                 rq["k"])
    return ids[0, :count[0]], scores[0, :count[0]], int(status[0])


def _check(res, t, rq, what):
    ids, scores, status = _alone(t, rq)
    assert res.status == status, what
    assert res.ids.tolist() == ids.tolist(), what
    assert res.scores.tolist() == scores.tolist(), what


def test_served_results_equal_alone(ctx, tenant_a):
    import svs
    rng = np.random.default_rng(1)
    reqs = _requests(tenant_a, rng, 60)
    tickets = [None] * len(reqs)
    with svs.Server(tenant_a["handle"].index, max_batch=8, slots=2) as srv:
        def submit(part):
            for i in part:
                r = reqs[i]
                tickets[i] = srv.submit(r["v_q"], r["q_seed"], r["probes"], r["verify"], r["k"])
        ths = [threading.Thread(target=submit, args=(range(j, len(reqs), 3),)) for j in range(3)]
        [t.start() for t in ths]
        [t.join() for t in ths]
        results = [srv.wait(tk) for tk in tickets]
        st = srv.stats()
    say("server stats", st)
    assert st["completed"] == len(reqs) and st["pending"] == 0
    assert st["batches"] < len(reqs) and st["max_batch"] > 1  # requests were batched
    for i, (res, rq) in enumerate(zip(results, reqs)):
        _check(res, tenant_a, rq, f"request {i}")
    # and the reference's own answer for the unverified ones
    t = tenant_a
    clusters = []
    for ci, members in enumerate(t["lists"]):
        mq = np.stack([O.quantise_vector(t["corpus"].vectors[j]) for j in members])
        clusters.append(O.build_cluster(mq.reshape(-1), members.size, t["l"],
                                        O.derive_h_seed(t["upload_seed"], ci)))
    for i in range(6):
        rq, res = reqs[i], results[i]
        q = t["qs"][rq["qi"]]
        ids, scores = O.ivf_score(clusters, t["lists"], t["corpus"].centres, t["l"], q,
                                  len(rq["probes"]), bytes(rq["q_seed"]), rq["verify"], rq["k"])
        assert res.ids.tolist() == ids.tolist() and res.scores.tolist() == scores.tolist()


def test_served_two_tenants(ctx, tenant_a, tenant_b):
    import svs
    rng = np.random.default_rng(2)
    ra = _requests(tenant_a, rng, 20, tenant=0)
    rb = _requests(tenant_b, rng, 20, tenant=1)
    mixed = [x for pair in zip(ra, rb) for x in pair]
    with svs.Server([tenant_a["handle"].index, tenant_b["handle"].index], max_batch=16,
                    slots=3) as srv:
        tickets = [srv.submit(r["v_q"], r["q_seed"], r["probes"], r["verify"], r["k"],
                              tenant=r["tenant"]) for r in mixed]
        results = [srv.wait(tk) for tk in tickets]
    for i, (res, rq) in enumerate(zip(results, mixed)):
        _check(res, tenant_a if rq["tenant"] == 0 else tenant_b, rq, f"request {i}")


def test_served_poisson_open_loop(ctx, tenant_a):
    import svs
    rng = np.random.default_rng(3)
    reqs = _requests(tenant_a, rng, 80)
    with svs.Server(tenant_a["handle"].index, max_batch=32, slots=2) as srv:
        results, wall = srv.run_poisson(reqs, qps=4000.0, seed=9)
        st = srv.stats()
    say("poisson wall", wall, "stats", st)
    assert st["completed"] == len(reqs)
    for i, (res, rq) in enumerate(zip(results, reqs)):
        assert res.latency_ms > 0
        _check(res, tenant_a, rq, f"request {i}")


def test_served_admission_by_memory(ctx, tenant_a):
    """A small per-batch scratch budget caps how many requests a batch
    admits (always at least one); results are unchanged."""
    import svs
    rng = np.random.default_rng(4)
    reqs = _requests(tenant_a, rng, 30)
    with svs.Server(tenant_a["handle"].index, max_batch=64, slots=1,
                    slot_bytes=100 * 1024) as srv:
        tickets = [srv.submit(r["v_q"], r["q_seed"], r["probes"], r["verify"], r["k"])
                   for r in reqs]
        results = [srv.wait(tk) for tk in tickets]
        st = srv.stats()
    say("admission stats", st)
    # a batch needs >= 2^16 candidate rows of capacity (36 B each) plus ~30 KiB
    # per query: beyond 100 KiB, so every batch holds just its first request
    assert st["max_batch"] == 1
    for i, (res, rq) in enumerate(zip(results, reqs)):
        _check(res, tenant_a, rq, f"request {i}")


def test_served_verification_failure(ctx, tenant_a):
    """A tampered list fails the verified requests that probe it (status
    1 + trial + 4 * list), in the mix exactly as alone."""
    import svs
    rng = np.random.default_rng(5)
    reqs = _requests(tenant_a, rng, 24)
    for r in reqs:
        r["verify"] = True
    target = int(reqs[0]["probes"][0])
    ctx.inject_fault(target, 0, 1)
    try:
        with svs.Server(tenant_a["handle"].index, max_batch=8, slots=2) as srv:
            tickets = [srv.submit(r["v_q"], r["q_seed"], r["probes"], r["verify"], r["k"])
                       for r in reqs]
            results = [srv.wait(tk) for tk in tickets]
        for i, (res, rq) in enumerate(zip(results, reqs)):
            if target in rq["probes"].tolist():
                assert res.status == 1 + 4 * target and res.ids.size == 0
            _check(res, tenant_a, rq, f"request {i}")
    finally:
        ctx.inject_fault(-1)


def test_served_after_reserve(ctx, tenant_a):
    """Pre-sized slots serve the same results; reserving after the first
    request is refused."""
    import svs
    rng = np.random.default_rng(6)
    reqs = _requests(tenant_a, rng, 20)
    with svs.Server(tenant_a["handle"].index, max_batch=16, slots=2) as srv:
        srv.reserve(16, 16 * 8, 16 * 2500, 40)
        tickets = [srv.submit(r["v_q"], r["q_seed"], r["probes"], r["verify"], r["k"])
                   for r in reqs]
        results = [srv.wait(tk) for tk in tickets]
        with pytest.raises(svs.GpuError):
            srv.reserve(16, 128, 1000, 10)
    for i, (res, rq) in enumerate(zip(results, reqs)):
        _check(res, tenant_a, rq, f"request {i}")


def _alone_raw(t, rq):
    ids, scores, count, status, _ = t["handle"].index.score_raw_batch(
        rq["query"][None, :], rq["q_seed"][None, :], rq["nprobe"],
        np.array([rq["verify"]], np.uint8), rq["k"])
    return ids[0, :count[0]], scores[0, :count[0]], int(status[0])


def test_served_raw_queries_equal_alone(ctx, tenant_a, tenant_b):
    """Raw f32 requests (routed and quantised on the device inside the
    batch) mixed with pre-routed ones and across tenants: each result equals
    the same request alone."""
    import svs
    rng = np.random.default_rng(8)
    reqs = []
    for i in range(60):
        t = tenant_a if i % 3 else tenant_b
        qi = int(rng.integers(0, t["qs"].shape[0]))
        rq = dict(tenant=0 if t is tenant_a else 1, q_seed=rng.integers(0, 256, 32, np.uint8),
                  verify=bool(rng.integers(0, 2)), k=int(rng.choice([1, 10, 16, 33])))
        if i % 4:
            rq.update(query=np.asarray(t["qs"][qi], np.float32),
                      nprobe=int(rng.choice([1, 3, 8, 40])))
        else:
            rq.update(_requests(t, rng, 1, tenant=rq["tenant"])[0])
        reqs.append(rq)
    with svs.Server([tenant_a["handle"].index, tenant_b["handle"].index], max_batch=16,
                    slots=2) as srv:
        results, _ = srv.run_poisson(reqs, qps=5000.0, seed=4)
        st = srv.stats()
    say("raw served stats", st)
    assert st["completed"] == len(reqs)
    for i, (res, rq) in enumerate(zip(results, reqs)):
        t = tenant_a if rq["tenant"] == 0 else tenant_b
        if rq.get("query") is not None:
            ids, scores, status = _alone_raw(t, rq)
            assert res.status == status, f"request {i}"
            assert res.ids.tolist() == ids.tolist(), f"request {i}"
            assert res.scores.tolist() == scores.tolist(), f"request {i}"
        else:
            _check(res, t, rq, f"request {i}")


@pytest.mark.parametrize("cross", [True, False])
def test_served_cross_tenant_batches(ctx, tenant_a, tenant_b, cross):
    """Requests of two tenants queued together: with cross-tenant batching a
    batch mixes them (each query encoded with its tenant's keys, routed on
    its tenant's centroids and scored on its tenant's lists through the
    combined list table), without it a batch holds one tenant; either way
    every result equals the same request alone. Probe ids past a tenant's
    own lists (which would be the other tenant's lists in the combined
    table) score nothing, as alone."""
    import svs
    rng = np.random.default_rng(21 if cross else 22)
    reqs = []
    for i in range(64):
        t, ti = (tenant_a, 0) if i % 2 else (tenant_b, 1)
        if i % 3 == 0:
            qi = int(rng.integers(0, t["qs"].shape[0]))
            rq = dict(tenant=ti, query=np.asarray(t["qs"][qi], np.float32),
                      nprobe=int(rng.choice([1, 3, 8])), q_seed=rng.integers(0, 256, 32, np.uint8),
                      verify=bool(rng.integers(0, 2)), k=int(rng.choice([1, 10, 16])))
        else:
            rq = _requests(t, rng, 1, tenant=ti)[0]
            if i % 5 == 0:  # one probe past the tenant's own lists
                n_own = len(t["lists"])
                rq["probes"] = np.append(rq["probes"], np.uint32(n_own + 2))
        reqs.append(rq)
    with svs.Server([tenant_a["handle"].index, tenant_b["handle"].index], max_batch=32,
                    slots=1, cross_tenant=cross) as srv:
        tickets = []
        for rq in reqs:
            if rq.get("query") is not None:
                tickets.append(srv.submit_raw(rq["query"], rq["q_seed"], rq["nprobe"],
                                              rq["verify"], rq["k"], tenant=rq["tenant"]))
            else:
                tickets.append(srv.submit(rq["v_q"], rq["q_seed"], rq["probes"], rq["verify"],
                                          rq["k"], tenant=rq["tenant"]))
        results = [srv.wait(tk) for tk in tickets]
        st = srv.stats()
    say("cross-tenant" if cross else "one tenant per batch", st)
    assert st["completed"] == len(reqs)
    for i, (res, rq) in enumerate(zip(results, reqs)):
        t = tenant_a if rq["tenant"] == 0 else tenant_b
        if rq.get("query") is not None:
            ids, scores, status = _alone_raw(t, rq)
            assert res.status == status, f"request {i}"
            assert res.ids.tolist() == ids.tolist(), f"request {i}"
            assert res.scores.tolist() == scores.tolist(), f"request {i}"
        else:
            _check(res, t, rq, f"request {i}")
