"""Full-size workloads through size-independent properties.

* BN-IVF at N = 1,000,000, d = 768, 1024 Zipf(1.1) lists, nprobe 32 (and 8,
  128, and a 512-query batch of mixed nprobe): the
  zero-noise invariant of decode (crypto.rs:1232-1266) — every returned
  score equals fp_to_signed(M_q . v_q) recomputed from the plaintext
  quantised vectors — plus the batch == per-query invariant
  (crypto.rs:1533-1559), raw queries == routed, and the server product r of
  sampled rows against the oracle's matvec.
* Tiptoe SimplePIR at 32768 x 8192 with entries in [0, 32), mod 2^32:
  linearity of the answer in the query and sampled rows against a direct
  mod-2^32 dot product.
"""
import sys

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


@pytest.fixture(scope="module")
def big(ctx):
    import svs
    from svs import data
    say("building the 1M corpus")
    corpus = data.make_corpus(1_000_000, 768, components=1024, zipf_s=1.1, seed=42)
    lists = data.ivf_lists(corpus)
    cfg = svs.BnTmIvfConfig(index=svs.IvfIndex(corpus.centres, lists), nprobe=32,
                            key_seed=bytes([42] * 32), upload_seed=7)
    handle, _ = svs.BnTmIvfScorer(ctx).upload_cluster(cfg, corpus.vectors)
    say("index built", ctx.stats()["resident_bytes"])
    yield dict(corpus=corpus, lists=lists, handle=handle, idx=handle.index)
    handle.index.close()


def _plain_score(corpus, qv, gid):
    """fp_to_signed(M_q . v_q mod p) for one corpus row."""
    import svs
    mq = svs.quantise(corpus.vectors[gid:gid + 1])[0]
    s = 0
    for a, b in zip(mq[:768].tolist(), qv[:768].tolist()):
        s += a * b
    s %= P
    return s if s < P // 2 else s - P


def test_fullsize_ivf_zero_noise_and_batch_invariants(ctx, big):
    import svs
    from svs import data
    idx, corpus = big["idx"], big["corpus"]
    qs = data.make_queries(corpus, 256, seed=77)
    rng = np.random.default_rng(78)
    seeds = rng.integers(0, 256, (256, 32), dtype=np.uint8)
    verify = (rng.random(256) < 0.5).astype(np.uint8)
    say("raw batch of 256")
    ids, scores, count, status, probes = idx.score_raw_batch(qs, seeds, 32, verify, 10)
    assert (status == 0).all() and (count == 10).all()
    # routed with the host C++ router: same probe lists, same results
    hp = svs.probe_route(corpus.centres, qs, 32)
    assert np.array_equal(probes, hp.reshape(-1))
    v_q = svs.quantise(qs)
    pp = np.arange(0, 256 * 32 + 1, 32, dtype=np.uint32)
    say("routed batch of 256")
    r_ids, r_scores, r_count, r_status = idx.score_batch(v_q, seeds, pp, hp.reshape(-1), verify,
                                                         10, 1)
    assert np.array_equal(r_ids, ids) and np.array_equal(r_scores, scores)
    # batch == per query (plain, one query per launch sequence)
    for q in (0, 17, 255):
        say("plain query", q)
        p_out = idx.score_batch(v_q[q:q + 1], seeds[q:q + 1], np.array([0, 32], np.uint32),
                                hp[q], verify[q:q + 1], 10, 0)
        assert p_out[0][0].tolist() == ids[q].tolist()
        assert p_out[1][0].tolist() == scores[q].tolist()
    # zero noise: every returned score is the plaintext product
    for q in (3, 128, 200):
        for i in range(10):
            assert scores[q, i] == _plain_score(corpus, v_q[q], int(ids[q, i])), (q, i)


def _zero_noise(corpus, v_q, ids, scores, count, queries):
    for q in queries:
        for i in range(int(count[q])):
            assert scores[q, i] == _plain_score(corpus, v_q[q], int(ids[q, i])), (q, i)


@pytest.mark.parametrize("nprobe", [8, 128])
def test_fullsize_ivf_nprobe_8_and_128(ctx, big, nprobe):
    """The other nprobe of the batched configuration at N = 1M: a raw batch
    of 256 queries (half verified) against the host router's probe lists,
    the plain one-query path and the zero-noise invariant."""
    import svs
    from svs import data
    idx, corpus = big["idx"], big["corpus"]
    qs = data.make_queries(corpus, 256, seed=90 + nprobe)
    rng = np.random.default_rng(91 + nprobe)
    seeds = rng.integers(0, 256, (256, 32), dtype=np.uint8)
    verify = (rng.random(256) < 0.5).astype(np.uint8)
    say("raw batch of 256, nprobe", nprobe)
    ids, scores, count, status, probes = idx.score_raw_batch(qs, seeds, nprobe, verify, 10)
    assert (status == 0).all() and (count == 10).all()
    hp = svs.probe_route(corpus.centres, qs, nprobe)
    assert np.array_equal(probes, hp.reshape(-1))
    v_q = svs.quantise(qs)
    for q in (0, 101, 255):
        say("plain query", q)
        p_out = idx.score_batch(v_q[q:q + 1], seeds[q:q + 1],
                                np.array([0, nprobe], np.uint32), hp[q], verify[q:q + 1], 10, 0)
        assert p_out[0][0].tolist() == ids[q].tolist()
        assert p_out[1][0].tolist() == scores[q].tolist()
    _zero_noise(corpus, v_q, ids, scores, count, (5, 77, 200))


def test_fullsize_ivf_mixed_nprobe_batch_512(ctx, big):
    """One 512-query batch with per-query nprobe in {8, 32, 128} (the
    batched configuration's mixed-nprobe launch): every query equals itself
    scored alone, and the zero-noise invariant holds."""
    import svs
    from svs import data
    idx, corpus = big["idx"], big["corpus"]
    rng = np.random.default_rng(92)
    qs = data.make_queries(corpus, 512, seed=93)
    npr = rng.choice([8, 32, 128], 512).astype(np.uint32)
    seeds = rng.integers(0, 256, (512, 32), dtype=np.uint8)
    verify = (rng.random(512) < 0.5).astype(np.uint8)
    say("mixed-nprobe raw batch of 512")
    ids, scores, count, status, probes = idx.score_raw_batch(qs, seeds, npr, verify, 10)
    assert (status == 0).all() and (count == 10).all()
    assert probes.size == int(npr.sum())
    for q in (0, 17, 300, 511):
        one = idx.score_raw_batch(qs[q:q + 1], seeds[q:q + 1], int(npr[q]), verify[q:q + 1],
                                  10, mode=0)
        assert one[0][0].tolist() == ids[q].tolist(), q
        assert one[1][0].tolist() == scores[q].tolist(), q
    _zero_noise(corpus, svs.quantise(qs), ids, scores, count, (3, 250, 499))


def test_fullsize_products_of_sampled_rows(ctx, big):
    """The server product r of the largest list (about 180k rows) for one
    v_enc, on sampled rows, against the oracle's matvec of those rows."""
    idx = big["idx"]
    sizes = np.array([x.size for x in big["lists"]])
    c = int(sizes.argmax())
    m = int(sizes[c])
    rng = np.random.default_rng(79)
    v = rng.integers(0, P, 1024, dtype=np.uint64)
    r = idx.compute_products(c, v)
    m_enc = idx.download_menc(c).reshape(m, 1024)
    rows = rng.choice(m, 64, replace=False)
    want = O.matvec(m_enc[rows].reshape(-1), rows.size, 1024, v)
    assert r[rows].tolist() == want.tolist()


def test_fullsize_tiptoe_linearity_and_rows(ctx):
    from svs.schemes import TiptoeDatabase
    rows, cols = 32768, 8192
    rng = np.random.default_rng(80)
    db = rng.integers(0, 32, size=(rows, cols), dtype=np.uint32)
    tt = TiptoeDatabase(db, ctx)
    q = rng.integers(0, 2**32, size=(4, cols), dtype=np.uint32)
    q[3] = q[0] + q[1]  # wraps mod 2^32
    for kernel in (0, 1):
        say("tiptoe kernel", kernel)
        ans = tt.server_answer(q, kernel=kernel)
        assert ans.shape == (4, rows)
        assert np.array_equal(ans[3], ans[0] + ans[1])  # linear mod 2^32
        sample = rng.choice(rows, 32, replace=False)
        for i in sample[:8]:
            for j in range(3):
                want = int((db[i].astype(np.uint64) * q[j].astype(np.uint64)).sum() % (1 << 32))
                assert int(ans[j, i]) == want, (kernel, i, j)
    tt.close()
