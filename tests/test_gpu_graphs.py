"""Batches replayed from captured HIP graphs give the same results as the
same batches launched kernel by kernel, on every replay, with other work and
synchronisations between the replays.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import svs
    return svs.Context(0)


@pytest.fixture(scope="module")
def setup(ctx):
    import svs
    from svs import data
    corpus = data.make_corpus(2400, components=20, zipf_s=1.1, seed=31)
    lists = data.ivf_lists(corpus)
    ivf = svs.IvfIndex(corpus.centres, lists)
    cfg = svs.BnTmIvfConfig(index=ivf, nprobe=4, key_seed=bytes([3] * 32), upload_seed=2,
                            verification_enabled=True)
    handle, _ = svs.BnTmIvfScorer(ctx).upload_cluster(cfg, corpus.vectors)
    qs = data.make_queries(corpus, 64, seed=32)
    return dict(idx=handle.index, corpus=corpus, qs=qs, nl=len(lists))


def _routed_batch(setup, rng, nq, nprobe):
    import svs
    qs = setup["qs"][rng.integers(0, setup["qs"].shape[0], nq)]
    probes = svs.probe_route(setup["corpus"].centres, qs, nprobe)
    pp = np.arange(0, nq * nprobe + 1, nprobe, dtype=np.uint32)
    return dict(v_q=svs.quantise(qs), seeds=rng.integers(0, 256, (nq, 32), dtype=np.uint8),
                pp=pp, plist=probes.reshape(-1).astype(np.uint32),
                verify=rng.integers(0, 2, nq).astype(np.uint8), qs=qs)


def _run(idx, b, k=10, mode=1):
    return idx.score_batch(b["v_q"], b["seeds"], b["pp"], b["plist"], b["verify"], k, mode)


def test_graph_replays_equal_launches(ctx, setup):
    idx = setup["idx"]
    rng = np.random.default_rng(1)
    a = _routed_batch(setup, rng, 24, 4)
    b = _routed_batch(setup, rng, 9, 6)
    ctx.set_graphs(False)
    want_a, want_b = _run(idx, a), _run(idx, b)
    ctx.set_graphs(True)
    s0 = ctx.stats()
    try:
        for rep in range(6):
            say("replay", rep)
            got_a = _run(idx, a)
            for x, y in zip(got_a, want_a):
                assert np.array_equal(x, y), f"replay {rep}"
            got_b = _run(idx, b)  # another shape between replays
            for x, y in zip(got_b, want_b):
                assert np.array_equal(x, y), f"replay {rep}"
            ctx.stats()  # a synchronising call between replays
    finally:
        ctx.set_graphs(True)
    s1 = ctx.stats()
    say("graph counters", s1["graph_captures"] - s0["graph_captures"],
        s1["graph_launches"] - s0["graph_launches"])
    assert s1["graph_captures"] - s0["graph_captures"] >= 2
    assert s1["graph_launches"] - s0["graph_launches"] >= 8


def test_graph_replays_new_inputs(ctx, setup):
    """The same shape with new queries, seeds and verification flags: a
    replay copies the new inputs from the staging memory."""
    idx = setup["idx"]
    rng = np.random.default_rng(2)
    batches = [_routed_batch(setup, rng, 16, 4) for _ in range(5)]
    ctx.set_graphs(False)
    want = [_run(idx, x) for x in batches]
    ctx.set_graphs(True)
    for rep in range(2):
        for i, x in enumerate(batches):
            got = _run(idx, x)
            for g, w in zip(got, want[i]):
                assert np.array_equal(g, w), f"batch {i} pass {rep}"


def test_graph_raw_batches(ctx, setup):
    """Raw queries (routing and planning inside the graph)."""
    idx = setup["idx"]
    rng = np.random.default_rng(3)
    qsets = [setup["qs"][rng.integers(0, 64, 20)] for _ in range(4)]
    seeds = [rng.integers(0, 256, (20, 32), dtype=np.uint8) for _ in range(4)]
    ver = [rng.integers(0, 2, 20).astype(np.uint8) for _ in range(4)]
    ctx.set_graphs(False)
    want = [idx.score_raw_batch(q, s, 5, v, 10) for q, s, v in zip(qsets, seeds, ver)]
    ctx.set_graphs(True)
    for rep in range(3):
        for i in range(4):
            got = idx.score_raw_batch(qsets[i], seeds[i], 5, ver[i], 10)
            for g, w in zip(got, want[i]):
                assert np.array_equal(g, w), f"raw batch {i} pass {rep}"


def test_graph_dev_batches(ctx, setup):
    """Device-resident inputs and outputs (the inputs are copied into the
    slot before each replay)."""
    import svs
    idx = setup["idx"]
    rng = np.random.default_rng(4)
    batches = [_routed_batch(setup, rng, 12, 4) for _ in range(3)]
    ctx.set_graphs(False)
    want = [_run(idx, x) for x in batches]
    ctx.set_graphs(True)
    outs = [svs.DeviceArray(ctx, (12, 10), np.uint32), svs.DeviceArray(ctx, (12, 10), np.int64),
            svs.DeviceArray(ctx, (12,), np.uint32), svs.DeviceArray(ctx, (12,), np.int32)]
    for rep in range(3):
        for i, x in enumerate(batches):
            vq = svs.DeviceArray.from_host(ctx, x["v_q"])
            sd = svs.DeviceArray.from_host(ctx, x["seeds"])
            idx.score_batch_dev(12, vq.ptr(), sd.ptr(), x["pp"], x["plist"], x["verify"], 10,
                                *[o.ptr() for o in outs])
            got = [o.to_host() for o in outs]
            for g, w in zip(got, want[i]):
                assert np.array_equal(g, w), f"dev batch {i} pass {rep}"
            vq.close()
            sd.close()
    for o in outs:
        o.close()


def test_served_fixed_shape_replays(ctx, setup):
    """Server batches of one shape replay their graph; results equal alone."""
    import svs
    idx = setup["idx"]
    rng = np.random.default_rng(5)
    reqs = [dict(query=setup["qs"][int(rng.integers(0, 64))], nprobe=3,
                 q_seed=rng.integers(0, 256, 32, np.uint8), verify=bool(rng.integers(0, 2)), k=10)
            for _ in range(64)]
    with svs.Server(idx, max_batch=4, slots=1) as srv:
        tickets = [srv.submit_raw(r["query"], r["q_seed"], r["nprobe"], r["verify"], r["k"])
                   for r in reqs]
        results = [srv.wait(t) for t in tickets]
        st = srv.stats()
    say("server graph counters", st)
    for i, (res, r) in enumerate(zip(results, reqs)):
        ids, scores, count, status, _ = idx.score_raw_batch(
            r["query"][None, :], r["q_seed"][None, :], r["nprobe"],
            np.array([r["verify"]], np.uint8), r["k"])
        assert res.status == int(status[0]), f"request {i}"
        assert res.ids.tolist() == ids[0, :count[0]].tolist(), f"request {i}"
        assert res.scores.tolist() == scores[0, :count[0]].tolist(), f"request {i}"
    assert st["graph_launches"] > 0


def test_graph_cache_eviction(ctx, setup):
    """More shapes than a slot keeps graphs for: least recently used graphs
    are evicted and recaptured; every result still equals the launched one."""
    idx = setup["idx"]
    rng = np.random.default_rng(6)
    batches = [_routed_batch(setup, rng, nq, 3) for nq in range(1, 41)]
    ctx.set_graphs(False)
    want = [_run(idx, x, k=5) for x in batches]
    ctx.set_graphs(True)
    s0 = ctx.stats()
    for rep in range(4):
        for i, x in enumerate(batches):
            got = _run(idx, x, k=5)
            for g, w in zip(got, want[i]):
                assert np.array_equal(g, w), f"shape {i} pass {rep}"
    s1 = ctx.stats()
    say("eviction counters", s1["graph_captures"] - s0["graph_captures"],
        s1["graph_launches"] - s0["graph_launches"])
    assert s1["graph_captures"] - s0["graph_captures"] > 32
