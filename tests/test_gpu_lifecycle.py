"""Index lifetimes against the batch-graph cache, and the server's admission
of tenants whose lists cannot be decoded.

A captured batch graph holds the device addresses of its index (L, L^T,
the list table, the centroids) and its list count. A new index, or new
centroids on an index, must never replay a graph captured before: results
after destroy + rebuild (likely at the same host address) and after a
second svs_index_set_centroids equal the oracle (BnTmIvfScorer::score,
scorer-bntm/src/ivf.rs:746-822).
"""
import sys

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


def _tenant(ctx, n, comps, seed, key=42, upload_seed=7):
    import svs
    from svs import data
    corpus = data.make_corpus(n, components=comps, zipf_s=1.1, seed=seed)
    lists = data.ivf_lists(corpus)
    cfg = svs.BnTmIvfConfig(index=svs.IvfIndex(corpus.centres, lists), nprobe=3,
                            key_seed=bytes([key] * 32), upload_seed=upload_seed,
                            verification_enabled=True)
    handle, _ = svs.BnTmIvfScorer(ctx).upload_cluster(cfg, corpus.vectors)
    l = O.generate_subspace(bytes([key] * 32))
    clusters = []
    for ci, members in enumerate(lists):
        mq = np.stack([O.quantise_vector(corpus.vectors[i]) for i in members]) if members.size \
            else np.zeros((0, O.N), np.uint64)
        clusters.append(O.build_cluster(mq.reshape(-1), members.size, l,
                                        O.derive_h_seed(upload_seed, ci)))
    return dict(corpus=corpus, lists=lists, handle=handle, l=l, clusters=clusters)


def _check_raw(t, qs, seeds, nprobe, verify, k, got):
    ids, scores, count, status, _ = got
    for i in range(qs.shape[0]):
        want_ids, want_scores = O.ivf_score(t["clusters"], t["lists"], t["corpus"].centres, t["l"],
                                            qs[i], nprobe, seeds[i], bool(verify[i]), k)
        assert status[i] == 0, f"query {i}"
        assert ids[i, :count[i]].tolist() == want_ids.tolist(), f"query {i}"
        assert scores[i, :count[i]].tolist() == want_scores.tolist(), f"query {i}"


def test_graphs_after_index_rebuild(ctx):
    """The same raw-batch shape on an index, then (after it is destroyed) on
    a new index with a different list count: every result equals the
    oracle, so no graph of the first index is replayed on the second."""
    from svs import data
    rng = np.random.default_rng(11)
    nq, nprobe, k = 12, 3, 10
    ctx.set_graphs(True)
    a = _tenant(ctx, 1500, 14, seed=101)
    qs = data.make_queries(a["corpus"], nq, seed=102)
    seeds = rng.integers(0, 256, (nq, 32), dtype=np.uint8)
    ver = rng.integers(0, 2, nq).astype(np.uint8)
    s0 = ctx.stats()
    for rep in range(3):  # launched, captured, replayed
        say("first index, pass", rep)
        _check_raw(a, qs, seeds, nprobe, ver, k,
                   a["handle"].index.score_raw_batch(qs, seeds, nprobe, ver, k))
    assert ctx.stats()["graph_launches"] > s0["graph_launches"]
    addr = a["handle"].index.h.value
    a["handle"].index.close()
    b = _tenant(ctx, 1300, 9, seed=201, key=5, upload_seed=3)
    say("second index at the same host address:", b["handle"].index.h.value == addr)
    qs_b = data.make_queries(b["corpus"], nq, seed=202)
    for rep in range(3):
        say("second index, pass", rep)
        _check_raw(b, qs_b, seeds, nprobe, ver, k,
                   b["handle"].index.score_raw_batch(qs_b, seeds, nprobe, ver, k))
    b["handle"].index.close()


def test_graphs_after_new_centroids(ctx):
    """A second svs_index_set_centroids (a new device buffer) after a shape
    was captured: routing uses the new centroids, as the host router does."""
    import svs
    from svs import data
    rng = np.random.default_rng(12)
    nq, nprobe, k = 10, 2, 5
    ctx.set_graphs(True)
    t = _tenant(ctx, 1200, 10, seed=301)
    idx = t["handle"].index
    qs = data.make_queries(t["corpus"], nq, seed=302)
    seeds = rng.integers(0, 256, (nq, 32), dtype=np.uint8)
    ver = np.zeros(nq, np.uint8)
    for rep in range(3):
        idx.score_raw_batch(qs, seeds, nprobe, ver, k)
    # the centroids in reverse order: routing now picks other list ids
    flipped = np.ascontiguousarray(t["corpus"].centres[::-1], np.float32)
    idx.set_centroids(flipped)
    for rep in range(3):
        say("new centroids, pass", rep)
        got = idx.score_raw_batch(qs, seeds, nprobe, ver, k)
        probes = got[4].reshape(nq, nprobe)
        want = svs.probe_route(flipped, qs, nprobe)
        assert np.array_equal(probes, want), f"pass {rep}"
    idx.close()


def test_server_rejects_list_without_mplain(ctx):
    """A tenant list uploaded without M_plain (legal for product-only use)
    cannot be decoded: the server refuses the tenant on the host, as
    svs_score_batch refuses such a list."""
    import svs
    from svs import _lib as L
    rng = np.random.default_rng(13)
    l = O.generate_subspace(bytes([9] * 32))
    idx = svs.GpuIndex(ctx, 2, l)
    m = 40
    mp = np.array([O.quantise_vector(v) for v in
                   rng.standard_normal((m, 768)).astype(np.float32) / 30.0])
    cl = O.build_cluster(mp.reshape(-1), m, l, O.derive_h_seed(1, 0))
    cl = dict(cl)
    idx.upload_list(0, cl)
    cl_no_mp = dict(cl)
    cl_no_mp["m_plain"] = None
    idx.upload_list(1, cl_no_mp)
    with pytest.raises(L.GpuError) as e:
        svs.Server(idx, max_batch=4, slots=1)
    assert e.value.code == L.SVS_ERR_STATE
    say("refused:", e.value)
    # the same list scored directly is refused as before
    v_q = svs.quantise(rng.standard_normal((1, 768)).astype(np.float32) / 30.0)
    with pytest.raises(L.GpuError) as e2:
        idx.score_batch(v_q, rng.integers(0, 256, (1, 32), dtype=np.uint8),
                        np.array([0, 1], np.uint32), np.array([1], np.uint32),
                        np.zeros(1, np.uint8), 5, L.MODE_BATCHED)
    assert e2.value.code == L.SVS_ERR_STATE
    # its server product still works (M_enc only)
    v_enc = rng.integers(0, O.P, O.N, dtype=np.uint64)
    r = idx.compute_products(1, v_enc)
    assert np.array_equal(r, O.matvec(cl["m_enc"], m, O.N, v_enc))
    idx.close()
