"""Client routing and quantisation on the device, and batches planned on the
device, bit-exact against the oracle and the host paths.

Routing: probe_route (ivf-index/src/ivf.rs:150-164) over l2_f32_scalar
(ivf-index/src/distance.rs:30-37); quantisation: quantise_vector
(scorer-bntm/src/crypto.rs:105-122); the whole raw-query pipeline:
BnTmIvfScorer::score (scorer-bntm/src/ivf.rs:746-822).
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


@pytest.fixture(scope="module")
def setup(ctx):
    import svs
    from svs import data
    corpus = data.make_corpus(3000, components=24, zipf_s=1.1, seed=21)
    lists = data.ivf_lists(corpus)
    lists[3] = np.zeros(0, np.uint32)  # an empty list
    ivf = svs.IvfIndex(corpus.centres, lists)
    cfg = svs.BnTmIvfConfig(index=ivf, nprobe=4, key_seed=bytes([9] * 32), upload_seed=5,
                            verification_enabled=True)
    scorer = svs.BnTmIvfScorer(ctx)
    handle, _ = scorer.upload_cluster(cfg, corpus.vectors)
    qs = data.make_queries(corpus, 40, seed=22)
    return dict(scorer=scorer, handle=handle, idx=handle.index, lists=lists, corpus=corpus,
                centres=np.ascontiguousarray(corpus.centres, np.float32), qs=qs,
                l=O.generate_subspace(bytes([9] * 32)))


def test_route_matches_oracle_and_host(ctx, setup):
    import svs
    idx, c, qs = setup["idx"], setup["centres"], setup["qs"]
    for nprobe in (1, 4, 7, c.shape[0], c.shape[0] + 5):
        say("route nprobe", nprobe)
        got = idx.route(qs, nprobe)
        np_ = min(nprobe, c.shape[0])
        assert got.shape == (qs.shape[0], np_)
        host = svs.probe_route(c, qs, np_)
        assert np.array_equal(got, host)
        for i in range(0, qs.shape[0], 7):
            assert got[i].tolist() == O.probe_route(c, qs[i], np_).tolist(), f"query {i}"


def test_route_ties_and_special_values(ctx):
    """Equal distances keep centroid order; a query on a centroid is at
    distance 0; duplicated centroids, odd d and more centroids than a
    distance tile; a NaN distance sorts last."""
    import svs
    rng = np.random.default_rng(3)
    nl, d = 70, 37
    cent = rng.standard_normal((nl, d)).astype(np.float32)
    cent[10] = cent[4]
    cent[11] = cent[4]
    cent[50] = cent[20]
    cent[60, 5] = np.nan
    idx = svs.GpuIndex(ctx, nl, O.generate_subspace(bytes(32)))
    idx.set_centroids(cent)
    qs = np.stack([cent[4], cent[20], -cent[4], rng.standard_normal(d).astype(np.float32),
                   np.zeros(d, np.float32)])
    got = idx.route(qs, nl)
    for i, q in enumerate(qs):
        want = O.probe_route(cent, q, nl)
        assert got[i].tolist() == want.tolist(), f"query {i}"
    assert got[0][:3].tolist() == [4, 10, 11]
    assert got[1][:2].tolist() == [20, 50]
    assert got[0][-1] == 60 and got[3][-1] == 60  # NaN distance last
    idx.close()


def test_quantise_edges_match_oracle():
    """The host restatement and the oracle agree on round-half-away and on
    Rust's saturating `as i64` (NaN -> 0, +-inf and huge values clamp)."""
    import svs
    v = np.array([0.5 / 2**20, -0.5 / 2**20, 1.5 / 2**20, -2.5 / 2**20, np.nan, np.inf, -np.inf,
                  3e30, -3e30, 1.0, -1.0, 0.0, -0.0], np.float32)
    got = svs.quantise(v[None, :])[0]
    want = O.quantise_vector(v)
    assert got.tolist() == want.tolist()
    assert want[0] == 1 and want[1] == O.P - 1 and want[2] == 2 and want[3] == O.P - 3
    assert want[4] == 0 and want[5] == 3 and want[6] == O.P - 4


def _routed(setup, qs, seeds, nprobe, verify, k, mode):
    import svs
    idx = setup["idx"]
    probes = [svs.probe_route(setup["centres"], qs[i:i + 1], int(nprobe[i]))[0]
              if nprobe[i] else np.zeros(0, np.uint32) for i in range(qs.shape[0])]
    pp = np.zeros(qs.shape[0] + 1, np.uint32)
    pp[1:] = np.cumsum([p.size for p in probes])
    return idx.score_batch(svs.quantise(qs), seeds, pp, np.concatenate(probes).astype(np.uint32),
                           verify, k, mode), probes


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_raw_batch_equals_routed_batch(ctx, setup, mode):
    """Quantise + route on the device give the same results as the host
    client steps followed by svs_score_batch."""
    rng = np.random.default_rng(40 + mode)
    qs = setup["qs"].copy()
    qs[3] *= 0  # a zero query
    nq = qs.shape[0]
    seeds = rng.integers(0, 256, (nq, 32), dtype=np.uint8)
    nprobe = rng.choice([1, 2, 4, 8, 24], nq).astype(np.uint32)
    nprobe[5] = 0  # no probes: no hits
    verify = rng.integers(0, 2, nq).astype(np.uint8)
    for k in (1, 10, 40):
        say("raw batch mode", mode, "k", k)
        ids, scores, count, status, probes = setup["idx"].score_raw_batch(
            qs, seeds, nprobe, verify, k, mode=mode)
        (wids, wscores, wcount, wstatus), wprobes = _routed(setup, qs, seeds, nprobe, verify, k,
                                                           mode)
        assert np.array_equal(probes, np.concatenate(wprobes))
        assert np.array_equal(count, wcount) and np.array_equal(status, wstatus)
        for q in range(nq):
            c = int(count[q])
            assert ids[q, :c].tolist() == wids[q, :c].tolist(), f"query {q}"
            assert scores[q, :c].tolist() == wscores[q, :c].tolist(), f"query {q}"
        assert count[5] == 0


def test_raw_batch_matches_oracle(ctx, setup):
    """The whole score (ivf.rs:746-822) from raw f32 queries, on the oracle."""
    s = setup
    clusters = []
    for ci, members in enumerate(s["lists"]):
        mq = np.stack([O.quantise_vector(s["corpus"].vectors[i]) for i in members]) \
            if members.size else np.zeros((0, 1024), np.uint64)
        clusters.append(O.build_cluster(mq.reshape(-1), members.size, s["l"],
                                        O.derive_h_seed(5, ci)))
    rng = np.random.default_rng(77)
    qs = s["qs"][:6]
    seeds = rng.integers(0, 256, (6, 32), dtype=np.uint8)
    ids, scores, count, status, _ = s["idx"].score_raw_batch(qs, seeds, 4,
                                                             np.ones(6, np.uint8), 10)
    for i in range(6):
        wids, wscores = O.ivf_score(clusters, s["lists"], s["centres"], s["l"], qs[i], 4,
                                    bytes(seeds[i]), True, 10)
        assert status[i] == 0
        assert ids[i, :count[i]].tolist() == wids.tolist(), f"query {i}"
        assert scores[i, :count[i]].tolist() == wscores.tolist(), f"query {i}"


def test_device_plan_equals_host_plan(ctx, setup):
    """Batches planned by the planning kernel and on the host give identical
    results: out-of-range and empty lists, a list probed twice by one query,
    queries without probes, mixed verification, a tampered list, and every
    tile variant (each has its own tile shape)."""
    import svs
    idx = setup["idx"]
    nl = len(setup["lists"])
    rng = np.random.default_rng(5)
    nq = 50
    v_q = svs.quantise(setup["qs"][rng.integers(0, setup["qs"].shape[0], nq)])
    seeds = rng.integers(0, 256, (nq, 32), dtype=np.uint8)
    probes, pp = [], [0]
    for q in range(nq):
        n = int(rng.integers(0, 9))
        p = rng.integers(0, nl, n).tolist()
        if q == 1:
            p = [3, nl + 4, 7, 7]  # empty, out of range, twice
        if q == 2:
            p = []
        probes += p
        pp.append(len(probes))
    pp = np.array(pp, np.uint32)
    plist = np.array(probes, np.uint32)
    verify = rng.integers(0, 2, nq).astype(np.uint8)
    target = int(plist[0])
    for fault in (False, True):
        if fault:
            ctx.inject_fault(target, 0, 1)
        try:
            for variant in range(svs._lib.TILE_VARIANTS):
                ctx.set_tile_variant(variant)
                for mode in (1, 2):
                    say("plan check fault", fault, "variant", variant, "mode", mode)
                    ctx.set_planner(False)
                    want = idx.score_batch(v_q, seeds, pp, plist, verify, 12, mode)
                    ctx.set_planner(True)
                    got = idx.score_batch(v_q, seeds, pp, plist, verify, 12, mode)
                    for a, b in zip(got, want):
                        assert np.array_equal(a, b)
                    if fault:
                        assert (want[3] != 0).any()
        finally:
            ctx.set_planner(True)
            ctx.set_tile_variant(0)
            if fault:
                ctx.inject_fault(-1)


def test_device_plan_large_batch(ctx, setup):
    """A batch past the planner's item limit falls back to host planning
    and still agrees with per-query runs."""
    import svs
    idx = setup["idx"]
    nl = len(setup["lists"])
    rng = np.random.default_rng(6)
    nq = 700
    v_q = svs.quantise(setup["qs"][rng.integers(0, setup["qs"].shape[0], nq)])
    seeds = rng.integers(0, 256, (nq, 32), dtype=np.uint8)
    pp = np.arange(0, nq * nl + 1, nl, dtype=np.uint32)  # every list: > 16384 items
    plist = np.tile(np.arange(nl, dtype=np.uint32), nq)
    verify = np.zeros(nq, np.uint8)
    got = idx.score_batch(v_q, seeds, pp, plist, verify, 5, 1)
    for q in (0, 399, 699):
        one = idx.score_batch(v_q[q:q + 1], seeds[q:q + 1], np.array([0, nl], np.uint32),
                              np.arange(nl, dtype=np.uint32), verify[:1], 5, 1)
        for a, b in zip(got, one):
            assert np.array_equal(a[q:q + 1], b)


def test_scorer_routes_on_device(ctx, setup):
    """BnTmIvfScorer.score_batch (raw queries) equals the oracle's score."""
    s = setup
    seeds = np.arange(8 * 32, dtype=np.uint64).astype(np.uint8)
    got = s["scorer"].score_batch(s["handle"], s["qs"][:8], 10, q_seeds=seeds)
    one = [s["scorer"].score(s["handle"], s["qs"][i], 10, q_seed=seeds[i * 32:(i + 1) * 32])
           for i in range(8)]
    assert got == one


def test_device_plan_near_item_limit(ctx, setup):
    """16000 items (the planner's sort at its largest power of two): the
    device plan equals the host plan."""
    import svs
    idx = setup["idx"]
    nl = len(setup["lists"])
    rng = np.random.default_rng(8)
    per = nl + 1
    nq = 16000 // per
    v_q = svs.quantise(setup["qs"][rng.integers(0, setup["qs"].shape[0], nq)])
    seeds = rng.integers(0, 256, (nq, 32), dtype=np.uint8)
    plist = np.concatenate([rng.permutation(np.append(np.arange(nl), nl + 3)) for _ in range(nq)])
    plist = plist.astype(np.uint32)
    pp = np.arange(0, nq * per + 1, per, dtype=np.uint32)
    verify = rng.integers(0, 2, nq).astype(np.uint8)
    try:
        ctx.set_planner(False)
        want = idx.score_batch(v_q, seeds, pp, plist, verify, 7, 1)
        ctx.set_planner(True)
        got = idx.score_batch(v_q, seeds, pp, plist, verify, 7, 1)
    finally:
        ctx.set_planner(True)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
