"""The BASELINE configurations on the GPU, through the C ABI.

* BN flat, N = 10,000, d = 768, a 64-component Gaussian mixture (sigma 0.3),
  100 queries, verification on, top-10: every result equals the oracle's
  BnTmScorer::score (scorer-bntm/src/lib.rs:326-475), one query per launch
  sequence and batched.
* The plain drop-in at full size: BN flat at N = 1,000,000 (1024-component
  mixture), 1,000 queries one per launch sequence, verification off and on.
  Parity at this size rests on size-independent properties: every returned
  score equals fp_to_signed(M_q . v_q) of its row (the zero-noise invariant
  of decode, crypto.rs:1232-1266); for sampled queries the whole top-10
  equals the plaintext top-10 (all 1M plaintext products in float64, exact
  here: |product| < 2^40, |sum| < 2^50); sampled rows of the server product
  equal the oracle's matvec; honest queries verify; a server fault is
  rejected at trial 0 (crypto.rs:1446-1474).
* The served configuration: 16 tenant sub-indices, open-loop Poisson
  arrivals, nprobe uniform in {8, 16, 32, 64, 128}, half of the requests
  verified: every served result equals the same request alone and the
  oracle's BnTmIvfScorer::score (scorer-bntm/src/ivf.rs:746-822).
"""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
P = O.P


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import svs
    return svs.Context(0)


def _signed(x):
    x = int(x) % P
    return x if x < P // 2 else x - P


def _topk_plain(scores, k):
    """Descending by score, ties by position (the library's tie rule)."""
    order = np.lexsort((np.arange(scores.size), -scores))
    return order[:k], scores[order[:k]]


# ------------------------------------------------------------ configuration 1
def test_config1_flat_10k_100_queries_verified(ctx):
    import svs
    from svs import data
    n, k = 10_000, 10
    corpus = data.make_corpus(n, 768, components=64, zipf_s=None, seed=42)
    key = bytes([42] * 32)
    h_seed = bytes([9] * 32)
    handle, _ = svs.BnTmScorer(ctx).upload_cluster(
        svs.BnTmConfig(key_seed=key, verification_enabled=True), corpus.vectors, h_seed=h_seed)
    idx = handle.index
    qs = data.make_queries(corpus, 100, seed=43)
    v_q = svs.quantise(qs)
    rng = np.random.default_rng(44)
    seeds = rng.integers(0, 256, (100, 32), dtype=np.uint8)
    pp = np.arange(101, dtype=np.uint32)
    pl = np.zeros(100, np.uint32)
    ver = np.ones(100, np.uint8)
    say("config 1: batched and plain runs")
    got_b = idx.score_batch(v_q, seeds, pp, pl, ver, k, 1)
    got_p = idx.score_batch(v_q, seeds, pp, pl, ver, k, 0)
    for a, b in zip(got_b, got_p):
        assert np.array_equal(a, b)
    ids, scores, count, status = got_b
    assert (status == 0).all() and (count == k).all()
    say("config 1: oracle cluster")
    l = O.generate_subspace(key)
    cl = O.build_cluster(O.quantise_rows(corpus.vectors).reshape(-1), n, l,
                         np.frombuffer(h_seed, np.uint8))

    def one(q):
        return O.bntm_flat_score(cl, l, v_q[q], seeds[q], True, k)
    with ThreadPoolExecutor(8) as ex:
        want = list(ex.map(one, range(100)))
    for q, (st, w_ids, w_scores) in enumerate(want):
        assert st == 0, (q, st)
        assert ids[q].tolist() == w_ids.tolist(), q
        assert scores[q].tolist() == w_scores.tolist(), q
    idx.close()


# ------------------------------------------------------------ configuration 2
@pytest.fixture(scope="module")
def flat1m(ctx):
    import svs
    from svs import data
    say("config 2: building the 1M flat corpus and index")
    corpus = data.make_corpus(1_000_000, 768, components=1024, zipf_s=None, seed=42)
    handle, _ = svs.BnTmScorer(ctx).upload_cluster(
        svs.BnTmConfig(key_seed=bytes([42] * 32)), corpus.vectors, h_seed=bytes([7] * 32))
    say("config 2: index built", ctx.stats()["resident_bytes"])
    yield dict(corpus=corpus, idx=handle.index)
    handle.index.close()


def _run_plain(idx, v_q, seeds, verify, k):
    pp = np.array([0, 1], np.uint32)
    pl = np.zeros(1, np.uint32)
    out = [idx.score_batch(v_q[q:q + 1], seeds[q:q + 1], pp, pl,
                           np.array([verify], np.uint8), k, 0) for q in range(v_q.shape[0])]
    return (np.concatenate([o[0] for o in out]), np.concatenate([o[1] for o in out]),
            np.concatenate([o[2] for o in out]), np.concatenate([o[3] for o in out]))


def test_config2_flat_1m_plain_1000_queries(ctx, flat1m):
    import svs
    from svs import data
    corpus, idx = flat1m["corpus"], flat1m["idx"]
    nq, k = 1000, 10
    qs = data.make_queries(corpus, nq, seed=4242)
    v_q = svs.quantise(qs)
    rng = np.random.default_rng(45)
    seeds = rng.integers(0, 256, (nq, 32), dtype=np.uint8)
    runs = {}
    for ver in (0, 1):
        say("config 2: 1000 plain queries, verification", ver)
        runs[ver] = _run_plain(idx, v_q, seeds, ver, k)
        ids, scores, count, status = runs[ver]
        assert (status == 0).all(), f"honest queries rejected (verify={ver})"
        assert (count == k).all()
    # the same hits with verification on and off
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    ids, scores = runs[1][0], runs[1][1]
    # zero noise: every returned score is its row's plaintext product
    say("config 2: zero-noise check of 10,000 hits")
    mq = O.quantise_rows(corpus.vectors[ids.reshape(-1)]).astype(object)[:, :768]
    vq = np.repeat(v_q.astype(object)[:, :768], k, axis=0)
    plain = [_signed(x) for x in (mq * vq).sum(axis=1)]
    assert plain == scores.reshape(-1).tolist()
    # the whole top-10 of sampled queries against all 1M plaintext products
    say("config 2: plaintext top-10 of 16 sampled queries")
    sample = rng.choice(nq, 16, replace=False)
    def signed_f64(a):
        a = a.astype(np.int64)  # canonical values < 2^61 fit; lift [p/2, p) to negatives
        a[a > P // 2] -= P
        return a.astype(np.float64)  # |x| <= 2^20: exact
    m_signed = signed_f64(O.quantise_rows(corpus.vectors)[:, :768])
    q_signed = signed_f64(v_q[sample, :768])
    prods = m_signed @ q_signed.T  # exact: every partial sum is an integer below 2^53
    del m_signed
    for j, q in enumerate(sample):
        col = prods[:, j].astype(np.int64)
        w_ids, w_scores = _topk_plain(col, k)
        assert scores[q].tolist() == w_scores.tolist(), q
        assert ids[q].tolist() == w_ids.tolist(), q


def test_config2_products_and_fault(ctx, flat1m):
    """Sampled rows of svs_compute_products (the reference's
    compute_products drop-in) against the oracle's matvec of the same rows;
    then a faulty server product is rejected at trial 0."""
    import svs
    from svs import data
    corpus, idx = flat1m["corpus"], flat1m["idx"]
    m = idx.rows(0)
    rng = np.random.default_rng(46)
    v = rng.integers(0, P, 1024, dtype=np.uint64)
    say("config 2: server product of the 1M list")
    r = idx.compute_products(0, v)
    m_enc = idx.download_menc(0).reshape(m, 1024)
    rows = np.sort(rng.choice(m, 128, replace=False))
    want = O.matvec(np.ascontiguousarray(m_enc[rows]).reshape(-1), rows.size, 1024, v)
    assert r[rows].tolist() == want.tolist()
    del m_enc
    qs = data.make_queries(corpus, 4, seed=47)
    v_q = svs.quantise(qs)
    seeds = rng.integers(0, 256, (4, 32), dtype=np.uint8)
    say("config 2: injected server fault")
    try:
        ctx.inject_fault(0, row=int(rng.integers(0, m)), delta=1)
        _, _, _, status = _run_plain(idx, v_q, seeds, 1, 10)
    finally:
        ctx.inject_fault(-1)
    # status = 1 + trial + 4 * list: trial 0 of list 0
    assert status.tolist() == [1, 1, 1, 1]
    # unverified queries do not check, so the fault is not reported
    _, _, _, st_off = _run_plain(idx, v_q, seeds, 0, 10)
    assert (st_off == 0).all()


# ------------------------------------------------------------ configuration 4
def _tenant(ctx, seed, n, comps, key, upload_seed):
    import svs
    from svs import data
    corpus = data.make_corpus(n, components=comps, zipf_s=1.1, seed=seed)
    lists = data.ivf_lists(corpus)
    cfg = svs.BnTmIvfConfig(index=svs.IvfIndex(corpus.centres, lists), nprobe=8,
                            key_seed=bytes([key] * 32), upload_seed=upload_seed)
    handle, _ = svs.BnTmIvfScorer(ctx).upload_cluster(cfg, corpus.vectors)
    l = O.generate_subspace(bytes([key] * 32))
    clusters = []
    for ci, members in enumerate(lists):
        mq = np.stack([O.quantise_vector(corpus.vectors[i]) for i in members]) if members.size \
            else np.zeros((0, O.N), np.uint64)
        clusters.append(O.build_cluster(mq.reshape(-1), members.size, l,
                                        O.derive_h_seed(upload_seed, ci)))
    return dict(corpus=corpus, lists=lists, handle=handle, l=l, clusters=clusters,
                qs=data.make_queries(corpus, 64, seed=seed + 1))


def test_config4_sixteen_tenants_poisson_mix(ctx):
    import svs
    say("config 4: building 16 tenants")
    tenants = [_tenant(ctx, 500 + 7 * t, 1500 + 100 * t, 160, key=t + 1, upload_seed=100 + t)
               for t in range(16)]
    rng = np.random.default_rng(48)
    choices = [8, 16, 32, 64, 128]
    reqs = []
    for i in range(320):
        t = int(rng.integers(0, 16))
        reqs.append(dict(tenant=t, query=tenants[t]["qs"][int(rng.integers(0, 64))],
                         nprobe=int(rng.choice(choices)),
                         q_seed=rng.integers(0, 256, 32, np.uint8),
                         verify=bool(rng.integers(0, 2)), k=10))
    with svs.Server([t["handle"].index for t in tenants], max_batch=64, slots=2) as srv:
        say("config 4: 320 Poisson requests at 2000 q/s")
        results, wall = srv.run_poisson(reqs, qps=2000.0, seed=49)
        st = srv.stats()
    say("config 4: served", st, "wall", round(wall, 3))
    assert st["completed"] >= 320

    def check(i):
        r, res = reqs[i], results[i]
        t = tenants[r["tenant"]]
        # alone: the same request as a one-query raw batch on its tenant
        ids, scores, count, status, _ = t["handle"].index.score_raw_batch(
            r["query"][None, :], r["q_seed"][None, :], r["nprobe"],
            np.array([r["verify"]], np.uint8), r["k"])
        assert res.status == int(status[0]) == 0, i
        assert res.ids.tolist() == ids[0, :count[0]].tolist(), i
        assert res.scores.tolist() == scores[0, :count[0]].tolist(), i
        w_ids, w_scores = O.ivf_score(t["clusters"], t["lists"], t["corpus"].centres, t["l"],
                                      r["query"], r["nprobe"], r["q_seed"], r["verify"], r["k"])
        assert res.ids.tolist() == w_ids.tolist(), i
        assert res.scores.tolist() == w_scores.tolist(), i
        return True
    with ThreadPoolExecutor(4) as ex:
        assert all(ex.map(check, range(len(reqs))))
    for t in tenants:
        t["handle"].index.close()
