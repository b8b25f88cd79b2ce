"""GPU parity of the BN path against the CPU oracle (bit-exact).

Mirrors the reference's GPU tests (scorer-bntm/src/lib.rs:1191-1570,
ivf.rs:2345-2371) and the protocol tests of crypto.rs:1232-1559, called
through the C ABI.
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


def _field(rng, size):
    return rng.integers(0, P, size=size, dtype=np.uint64)


def _random_cluster(rng, m):
    """A list with random field state (shape-only checks of the product)."""
    rowptr = np.zeros(m + 1, np.uint32)
    return dict(m=m, m_enc=_field(rng, m * 1024), a_l=_field(rng, m * 128), h=_field(rng, m * 128),
                m_plain=None, s_rowptr=rowptr, s_idx=np.zeros(0, np.uint32),
                s_sign=np.zeros(0, np.int8))


@pytest.fixture(scope="module")
def l_sub():
    return O.generate_subspace(bytes([1] * 32))


def test_compute_products_matches_oracle(ctx, l_sub):
    import svs
    sizes = [1, 7, 32, 63, 64, 65, 255, 256, 257, 300, 511, 1024, 1500]
    idx = svs.GpuIndex(ctx, len(sizes), l_sub)
    rng = np.random.default_rng(0xC0DE)
    cls = []
    for i, m in enumerate(sizes):
        cl = _random_cluster(rng, m)
        idx.upload_list(i, cl)
        cls.append(cl)
    for i, m in enumerate(sizes):
        say("compute_products m =", m)
        v = _field(rng, 1024)
        got = idx.compute_products(i, v)
        want = O.matvec(cls[i]["m_enc"], m, 1024, v)
        assert np.array_equal(got, want), f"m={m}"


def test_batched_products_mfma_and_valu_match_oracle(ctx, l_sub):
    """compute_products_batch (crypto.rs:782-828) generalised to uneven probe
    sets; both product kernels against the oracle, bit-exact."""
    import svs
    sizes = [1, 15, 16, 17, 63, 64, 65, 130, 300, 0]
    idx = svs.GpuIndex(ctx, len(sizes), l_sub)
    rng = np.random.default_rng(0xBA7C)
    cls = []
    for i, m in enumerate(sizes):
        cl = _random_cluster(rng, m)
        if i == 2:  # extremes: all p-1 rows
            cl["m_enc"] = np.full(m * 1024, P - 1, np.uint64)
        if i == 3:  # zero rows
            cl["m_enc"] = np.zeros(m * 1024, np.uint64)
        idx.upload_list(i, cl)
        cls.append(cl)
    nq = 11
    v = _field(rng, (nq, 1024))
    v[0] = P - 1
    v[1] = 0
    v[2, ::2] = P - 1
    probes = [rng.choice(len(sizes), size=int(rng.integers(1, 8)), replace=False) for _ in range(nq)]
    probes[3] = np.arange(len(sizes))  # one query probes every list
    pp = np.concatenate([[0], np.cumsum([len(x) for x in probes])]).astype(np.uint32)
    pl = np.concatenate(probes).astype(np.uint32)
    for kernel in (svs._lib.KERNEL_MFMA, svs._lib.KERNEL_MFMA_WAVE, svs._lib.KERNEL_VALU):
        say("batched products kernel", kernel)
        got = idx.compute_products_batch(v, pp, pl, kernel)
        item = 0
        for q in range(nq):
            for c in probes[q]:
                m = sizes[c]
                if m == 0:
                    assert got[item] is None
                else:
                    want = O.matvec(cls[c]["m_enc"], m, 1024, v[q])
                    assert np.array_equal(got[item], want), f"kernel {kernel} q {q} list {c}"
                item += 1


def test_mfma_many_queries_per_list(ctx, l_sub):
    """A hot list probed by many queries (several MFMA query groups)."""
    import svs
    rng = np.random.default_rng(91)
    m = 300  # two 256-row block tiles, the second ragged
    cl = _random_cluster(rng, m)
    idx = svs.GpuIndex(ctx, 1, l_sub)
    idx.upload_list(0, cl)
    nq = 37  # 5 groups of 8 (block tiles) or 10 of 4 (wave tiles), last ragged
    v = _field(rng, (nq, 1024))
    pp = np.arange(nq + 1, dtype=np.uint32)
    pl = np.zeros(nq, np.uint32)
    want = O.matvec_batch(cl["m_enc"], m, 1024, v)
    for kernel in (svs._lib.KERNEL_MFMA, svs._lib.KERNEL_MFMA_WAVE):
        say("mfma hot list, queries", nq, "kernel", kernel)
        got = idx.compute_products_batch(v, pp, pl, kernel)
        for q in range(nq):
            assert np.array_equal(got[q], want[q]), f"kernel {kernel} query {q}"


def test_compute_products_extremes(ctx, l_sub):
    import svs
    idx = svs.GpuIndex(ctx, 2, l_sub)
    m = 70
    allmax = dict(_random_cluster(np.random.default_rng(1), m))
    allmax["m_enc"] = np.full(m * 1024, P - 1, np.uint64)
    idx.upload_list(0, allmax)
    say("compute_products all p-1")
    r = idx.compute_products(0, np.full(1024, P - 1, np.uint64))
    assert np.all(r == np.uint64((1024 * (P - 1) ** 2) % P))
    rnd = _random_cluster(np.random.default_rng(2), m)
    idx.upload_list(1, rnd)
    for col in (0, 1, 63, 64, 1023):
        say("one-hot column", col)
        v = np.zeros(1024, np.uint64)
        v[col] = 1
        assert np.array_equal(idx.compute_products(1, v), rnd["m_enc"].reshape(m, 1024)[:, col])


def test_empty_and_out_of_range_lists(ctx, l_sub):
    import svs
    idx = svs.GpuIndex(ctx, 2, l_sub)
    idx.upload_list(0, _random_cluster(np.random.default_rng(3), 0))
    say("empty list")
    assert idx.compute_products(0, np.zeros(1024, np.uint64)) is None
    assert idx.compute_products(1, np.zeros(1024, np.uint64)) is None  # never uploaded
    assert idx.compute_products(7, np.zeros(1024, np.uint64)) is None  # out of range


def test_non_canonical_input_rejected(ctx, l_sub):
    import svs
    idx = svs.GpuIndex(ctx, 1, l_sub)
    idx.upload_list(0, _random_cluster(np.random.default_rng(4), 8))
    v = np.zeros(1024, np.uint64)
    v[3] = P
    with pytest.raises(svs.GpuError):
        idx.compute_products(0, v)


def _real_cluster(m, seed, l):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((m, 768)).astype(np.float32)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    mq = np.stack([O.quantise_vector(r) for r in x]) if m else np.zeros((0, 1024), np.uint64)
    h_seed = bytes([seed % 256] * 32)
    return O.build_cluster(mq.reshape(-1), m, l, h_seed), x, h_seed


def test_device_build_matches_host_encryption(ctx, l_sub):
    import svs
    idx = svs.GpuIndex(ctx, 3, l_sub)
    for i, m in enumerate([1, 37, 130]):
        say("build_list m =", m)
        cl, _, h_seed = _real_cluster(m, 10 + i, l_sub)
        idx.build_list(i, cl["m_plain"], h_seed)
        assert np.array_equal(idx.download_menc(i), cl["m_enc"])
    # random field rows (M_plain not representable in 32 bits)
    rng = np.random.default_rng(5)
    mp = _field(rng, 20 * 1024)
    cl = O.build_cluster(mp, 20, l_sub, bytes([9] * 32))
    idx.build_list(0, mp, bytes([9] * 32))
    assert np.array_equal(idx.download_menc(0), cl["m_enc"])


def test_verify_honest_and_tampered(ctx, l_sub):
    import svs
    cl, _, _ = _real_cluster(16, 21, l_sub)
    idx = svs.GpuIndex(ctx, 1, l_sub)
    idx.upload_list(0, cl)
    rng = np.random.default_rng(21)
    for b in range(6):
        say("verify honest", b)
        v_enc = _field(rng, 1024)
        r = O.matvec(cl["m_enc"], 16, 1024, v_enc)
        seed = bytes([b ^ 0xAA] * 32)
        assert idx.verify(0, r, v_enc, seed) == -1 == O.verify_response(r, cl["m_enc"], 16, v_enc,
                                                                         seed)
    # one coordinate + 1: rejected at trial 0 (crypto.rs:1446-1474)
    v_enc = _field(rng, 1024)
    r = O.matvec(cl["m_enc"], 16, 1024, v_enc)
    r[8] = (int(r[8]) + 1) % P
    seed = bytes([35] * 32)
    assert idx.verify(0, r, v_enc, seed) == 0 == O.verify_response(r, cl["m_enc"], 16, v_enc, seed)
    # an error invisible to trial 0 (u_0 . e = 0) must be caught by trial 1
    r = O.matvec(cl["m_enc"], 16, 1024, v_enc)
    u0 = O.u_draws(seed, 16)
    ea = 5
    eb = (-ea * u0[3] * pow(u0[11], P - 2, P)) % P
    r[3] = (int(r[3]) + ea) % P
    r[11] = (int(r[11]) + eb) % P
    want = O.verify_response(r, cl["m_enc"], 16, v_enc, seed)
    say("verify trial-1 tamper, oracle says", want)
    assert want == 1
    assert idx.verify(0, r, v_enc, seed) == 1


def test_verify_chunk_boundaries(ctx, l_sub):
    """u draws straddle keystream blocks when t*m is not a multiple of 8."""
    import svs
    rng = np.random.default_rng(31)
    for m in (1, 5, 17, 33, 100):
        say("verify m =", m)
        cl = _random_cluster(rng, m)
        idx = svs.GpuIndex(ctx, 1, l_sub)
        idx.upload_list(0, cl)
        v = _field(rng, 1024)
        r = O.matvec(cl["m_enc"], m, 1024, v)
        seed = bytes([m] * 32)
        assert idx.verify(0, r, v, seed) == -1
        r[m - 1] = (int(r[m - 1]) + 3) % P
        assert idx.verify(0, r, v, seed) == O.verify_response(r, cl["m_enc"], m, v, seed) == 0


def test_decode_matches_oracle(ctx, l_sub):
    import svs
    for m, seed in ((1, 40), (50, 41), (129, 42)):
        say("decode m =", m)
        cl, _, _ = _real_cluster(m, seed, l_sub)
        idx = svs.GpuIndex(ctx, 1, l_sub)
        idx.upload_list(0, cl)
        v = O.quantise_vector(np.random.default_rng(seed).standard_normal(768).astype(np.float32)
                              / 30)
        v_enc, g, t_idx, t_sign = O.encode_query(v, l_sub, bytes([seed] * 32))
        r = O.matvec(cl["m_enc"], m, 1024, v_enc)
        want = O.decode_scores(r, m, cl["a_l"], cl["m_plain"], cl["h"], cl["s_rowptr"],
                               cl["s_idx"], cl["s_sign"], l_sub, v_enc, g, t_idx, t_sign)
        got = idx.decode(0, r, v_enc, g, t_idx, t_sign)
        assert np.array_equal(got, want)
        # zero-noise: decode == M_plain . v exactly
        assert np.array_equal(got, O.matvec(cl["m_plain"], m, 1024, v))
        lt = O.matvec_transpose(l_sub, 1024, 128, v_enc)
        dense = idx.decode_dense_terms(0, g, lt)
        want_dense = [(int(a) + int(b)) % P for a, b in
                      zip(O.matvec(cl["a_l"], m, 128, g), O.matvec(cl["h"], m, 128, lt))]
        assert [int(x) for x in dense] == want_dense


def test_device_encode_matches_oracle(ctx, l_sub):
    """The device encode (the batch pipeline's keystream, G draws, sparse T
    walk and v_enc = v + L G + T) equals encode_query (crypto.rs:508-553)
    for many seeds. The scores cannot show this: decode removes G and T
    exactly whatever they are (zero noise)."""
    import svs
    rng = np.random.default_rng(2027)
    idx = svs.GpuIndex(ctx, 1, l_sub)
    nq = 96
    v_q = _field(rng, (nq, 1024))
    v_q[:, 768:] = 0
    v_q[0] = 0
    seeds = rng.integers(0, 256, (nq, 32), dtype=np.uint8)
    seeds[1] = 0
    seeds[2] = 255
    v_enc, g, t = idx.encode_query(v_q, seeds)
    for q in range(nq):
        w_enc, w_g, w_idx, w_sign = O.encode_query(v_q[q], l_sub, bytes(seeds[q]))
        dense = np.zeros(1024, np.int8)
        dense[w_idx] = w_sign
        assert np.array_equal(g[q], w_g), f"g of query {q}"
        assert np.array_equal(t[q], dense), f"T of query {q}"
        assert np.array_equal(v_enc[q], w_enc), f"v_enc of query {q}"
    idx.close()


def test_decode_with_wide_plaintext_rows(ctx, l_sub):
    """M_plain outside the 32-bit signed range uses the 64-bit column store."""
    import svs
    rng = np.random.default_rng(77)
    mp = _field(rng, 9 * 1024)
    cl = O.build_cluster(mp, 9, l_sub, bytes([77] * 32))
    idx = svs.GpuIndex(ctx, 1, l_sub)
    idx.upload_list(0, cl)
    v = _field(rng, 1024)
    v_enc, g, t_idx, t_sign = O.encode_query(v, l_sub, bytes([78] * 32))
    r = O.matvec(cl["m_enc"], 9, 1024, v_enc)
    say("decode wide rows")
    assert np.array_equal(idx.decode(0, r, v_enc, g, t_idx, t_sign), O.matvec(mp, 9, 1024, v))


def test_flat_scorer_matches_oracle(ctx):
    import svs
    from svs import data
    corpus = data.make_corpus(300, components=8, seed=3)
    qs = data.make_queries(corpus, 6, seed=4)
    cfg = svs.BnTmConfig(key_seed=bytes([42] * 32), verification_enabled=True)
    scorer = svs.BnTmScorer(ctx)
    h_seed = bytes([5] * 32)
    handle, _ = scorer.upload_cluster(cfg, corpus.vectors, h_seed=h_seed)
    l = O.generate_subspace(bytes([42] * 32))
    mq = np.stack([O.quantise_vector(v) for v in corpus.vectors])
    cl = O.build_cluster(mq.reshape(-1), 300, l, h_seed)
    seeds = [bytes([100 + i] * 32) for i in range(6)]
    for mode in (0, 1, 2):
        say("flat scorer mode", mode)
        got = scorer.score_batch(handle, qs, 10, q_seeds=np.frombuffer(b"".join(seeds), np.uint8),
                                 mode=mode)
        for i, q in enumerate(qs):
            st, ids, scores = O.bntm_flat_score(cl, l, O.quantise_vector(q), seeds[i], True, 10)
            assert st == 0
            assert [h.id for h in got[i]] == ids.tolist()
            assert [h.score for h in got[i]] == [float(np.float32(s)) for s in scores]
    # k > m is capped (lib.rs:1069-1083); k > 16 uses the wide top-k kernel
    assert len(scorer.score(handle, qs[0], 1000, q_seed=seeds[0])) == 300
    for k in (40, 257):
        say("flat scorer k =", k)
        wide = scorer.score(handle, qs[0], k, q_seed=seeds[0])
        st, ids, _ = O.bntm_flat_score(cl, l, O.quantise_vector(qs[0]), seeds[0], False, k)
        assert [h.id for h in wide] == ids.tolist()


@pytest.fixture(scope="module")
def ivf_setup(ctx):
    import svs
    from svs import data
    corpus = data.make_corpus(2500, components=16, zipf_s=1.1, seed=11)
    lists = data.ivf_lists(corpus)
    lists[5] = np.zeros(0, np.uint32)  # an empty list (ivf.rs skips m == 0)
    ivf = svs.IvfIndex(corpus.centres, lists)
    cfg = svs.BnTmIvfConfig(index=ivf, nprobe=4, key_seed=bytes([42] * 32), upload_seed=7,
                            verification_enabled=True)
    scorer = svs.BnTmIvfScorer(ctx)
    handle, _ = scorer.upload_cluster(cfg, corpus.vectors)
    l = O.generate_subspace(bytes([42] * 32))
    clusters = []
    for ci, members in enumerate(lists):
        mq = np.stack([O.quantise_vector(corpus.vectors[i]) for i in members]) if members.size \
            else np.zeros((0, 1024), np.uint64)
        clusters.append(O.build_cluster(mq.reshape(-1), members.size, l, O.derive_h_seed(7, ci)))
    qs = data.make_queries(corpus, 12, seed=12)
    return dict(scorer=scorer, handle=handle, clusters=clusters, lists=lists, l=l,
                centres=corpus.centres, qs=qs)


def test_ivf_matches_oracle(ivf_setup):
    s = ivf_setup
    seeds = np.frombuffer(bytes(range(12 * 32 % 256)) * 2, np.uint8)[: 12 * 32]
    seeds = np.arange(12 * 32, dtype=np.uint64).astype(np.uint8)
    for mode in (0, 1, 2):
        say("ivf mode", mode)
        got = s["scorer"].score_batch(s["handle"], s["qs"], 10, q_seeds=seeds, mode=mode)
        for i, q in enumerate(s["qs"]):
            ids, scores = O.ivf_score(s["clusters"], s["lists"], s["centres"], s["l"], q, 4,
                                      seeds[i * 32:(i + 1) * 32], True, 10)
            assert [h.id for h in got[i]] == ids.tolist(), f"query {i}"
            assert [h.score for h in got[i]] == [float(np.float32(x)) for x in scores]


def test_ivf_batch_equals_sequential(ivf_setup):
    s = ivf_setup
    seeds = np.arange(12 * 32, dtype=np.uint64).astype(np.uint8) ^ np.uint8(0x5A)
    batch = s["scorer"].score_batch(s["handle"], s["qs"], 10, q_seeds=seeds)
    for i, q in enumerate(s["qs"]):
        one = s["scorer"].score(s["handle"], q, 10, q_seed=seeds[i * 32:(i + 1) * 32])
        assert one == batch[i]


def test_verification_overhead_in_expected_band(ctx, ivf_setup):
    """Verification does real work and no more than 20x the server product
    (crypto.rs:1482-1523), from the stage events on the library's stream,
    in both kernel families; with verification off its stage is a no-op."""
    import svs
    s = ivf_setup
    idx = s["handle"].index
    qs = s["qs"]
    nq = qs.shape[0]
    probes = svs.probe_route(s["centres"], qs, 4)
    pp = np.arange(0, nq * 4 + 1, 4, dtype=np.uint32)
    v_q = svs.quantise(qs)
    seeds = np.arange(nq * 32, dtype=np.uint64).astype(np.uint8).reshape(nq, 32)
    for mode in (0, 1):
        for ver in (1, 0):
            flags = np.full(nq, ver, np.uint8)
            idx.score_batch(v_q, seeds, pp, probes.reshape(-1), flags, 10, mode)
            ctx.profile(True, True)
            for _ in range(3):
                idx.score_batch(v_q, seeds, pp, probes.reshape(-1), flags, 10, mode)
            prof = ctx.profile(False, False)
            product_ms, verify_ms = prof["product"][0], prof["verify"][0]
            say("mode", mode, "verify", ver, "product ms", product_ms, "verify ms", verify_ms)
            assert product_ms > 0
            if ver:
                assert 0 < verify_ms <= 20 * product_ms
            else:
                assert verify_ms <= product_ms


def test_ivf_tampered_server_rejected_with_cluster(ctx, ivf_setup):
    import svs
    s = ivf_setup
    h = s["handle"]
    probes = svs.probe_route(s["centres"], s["qs"], 4)
    target = int(probes[0][1]) if len(s["lists"][int(probes[0][1])]) else int(probes[0][0])
    v_q = svs.quantise(s["qs"])
    seeds = np.zeros((12, 32), np.uint8)
    pp = np.arange(0, 12 * 4 + 1, 4, dtype=np.uint32)
    ctx.inject_fault(target, 0, 1)
    try:
        say("tamper list", target)
        _, _, count, status = h.index.score_batch(v_q, seeds, pp, probes.reshape(-1),
                                                  np.ones(12, np.uint8), 10)
    finally:
        ctx.inject_fault(-1)
    for q in range(12):
        if target in probes[q].tolist():
            assert status[q] == 1 + 0 + 4 * target and count[q] == 0
        else:
            assert status[q] == 0 and count[q] == 10
    with pytest.raises(svs.VerificationFailed):
        ctx.inject_fault(target, 0, 1)
        try:
            s["scorer"].score(h, s["qs"][0], 10, q_seed=bytes(32))
        finally:
            ctx.inject_fault(-1)


def test_single_list_ivf_equals_flat(ctx):
    """ivf.rs:2045-2078: one list probed with nprobe 1 == flat scorer."""
    import svs
    from svs import data
    corpus = data.make_corpus(200, components=4, seed=21)
    qs = data.make_queries(corpus, 3, seed=22)
    ivf = svs.IvfIndex(corpus.centres[:1], [np.arange(200, dtype=np.uint32)])
    cfg = svs.BnTmIvfConfig(index=ivf, nprobe=1, key_seed=bytes([3] * 32), upload_seed=0)
    hi, _ = svs.BnTmIvfScorer(ctx).upload_cluster(cfg, corpus.vectors)
    hf, _ = svs.BnTmScorer(ctx).upload_cluster(svs.BnTmConfig(key_seed=bytes([3] * 32)),
                                               corpus.vectors, h_seed=svs.derive_h_seed(0, 0))
    seeds = np.full(3 * 32, 9, np.uint8)
    a = svs.BnTmIvfScorer(ctx).score_batch(hi, qs, 10, q_seeds=seeds)
    b = svs.BnTmScorer(ctx).score_batch(hf, qs, 10, q_seeds=seeds)
    assert a == b


def test_realised_cost_counts_probed_rows(ivf_setup):
    import svs
    s = ivf_setup
    hits, cost = s["scorer"].score_with_realised_cost(s["handle"], s["qs"][0], 10, q_seed=bytes(32))
    probes = svs.probe_route(s["centres"], s["qs"][:1], 4)[0]
    rows = sum(len(s["lists"][c]) for c in probes)
    assert cost.response_bytes == rows * 8
    assert cost.effective_bytes_per_query == rows * 1024 * 8
    assert len(hits) == 10


def _oracle_probe_scores(clusters, probes, v_q, seed, l, k):
    """ivf_score's per-list steps (ivf.rs:746-822) for an explicit probe set."""
    v_enc, g, t_idx, t_sign = O.encode_query(v_q, l, seed)
    ids, scores = [], []
    for c in probes:
        cl = clusters[c]
        r = O.matvec(cl["m_enc"], cl["m"], 1024, v_enc)
        mv = O.decode_scores(r, cl["m"], cl["a_l"], cl["m_plain"], cl["h"], cl["s_rowptr"],
                             cl["s_idx"], cl["s_sign"], l, v_enc, g, t_idx, t_sign)
        ids.extend((1000 * c + np.arange(cl["m"])).tolist())
        scores.extend(O.fp_to_signed_vec(mv).tolist())
    return O.topk(np.array(ids, np.uint32), np.array(scores, np.int64), k)


def test_score_batch_decode_planes_and_fallback(ctx, l_sub):
    """Lists whose plaintext fits the 24-bit decode planes use the MFMA
    decode; a batch touching any other list (|M| >= 2^23 in 32 bits, or a
    full-width field element) decodes per candidate. Both bit-exact."""
    import svs
    rng = np.random.default_rng(2024)
    lim = 1 << 23
    m = 70

    def lift(v):
        v = np.asarray(v, np.int64)
        return np.array([int(x) % P for x in v], np.uint64)

    small = rng.integers(-lim + 1, lim, size=m * 1024)
    small[:4] = [lim - 1, -(lim - 1), 0, -1]
    edge = rng.integers(-1000, 1000, size=m * 1024)
    edge[7] = lim  # one entry just outside the planes
    wide32 = rng.integers(-(1 << 30), 1 << 30, size=m * 1024)
    mps = [lift(small), lift(edge), lift(wide32), _field(rng, m * 1024)]
    clusters = [O.build_cluster(mp, m, l_sub, bytes([60 + i] * 32)) for i, mp in enumerate(mps)]
    idx = svs.GpuIndex(ctx, len(clusters), l_sub)
    for i, cl in enumerate(clusters):
        idx.upload_list(i, cl, ids=1000 * i + np.arange(m, dtype=np.uint32))
    nq, k = 5, 12
    v_q = lift(rng.integers(-(1 << 20), 1 << 20, size=(nq, 1024)).reshape(-1)).reshape(nq, 1024)
    seeds = rng.integers(0, 256, size=(nq, 32), dtype=np.uint8)
    for probes in ([[0]] * nq, [[0, 1], [2], [3, 0], [1, 2, 3], [0]]):
        pp = np.concatenate([[0], np.cumsum([len(p) for p in probes])]).astype(np.uint32)
        pl = np.concatenate(probes).astype(np.uint32)
        want = [_oracle_probe_scores(clusters, probes[q], v_q[q], bytes(seeds[q]), l_sub, k)
                for q in range(nq)]
        for mode in (0, 1, 2):
            say("plane fallback probes", probes, "mode", mode)
            ids, scores, count, status = idx.score_batch(v_q, seeds, pp, pl, None, k, mode=mode)
            for q in range(nq):
                assert status[q] == 0 and count[q] == k
                assert ids[q].tolist() == want[q][0].tolist(), f"mode {mode} query {q}"
                assert scores[q].tolist() == want[q][1].tolist(), f"mode {mode} query {q}"


def test_decode_skips_zero_plaintext_columns(ctx, l_sub):
    """Lists whose plaintext is zero beyond some column (d < n: quantised
    vectors are zero-padded) decode M T over their nonzero columns only;
    lists of widths 64, 100 (-> 128), 768 and 1024, alone and in one batch,
    in every mode, equal the oracle. One row of the 100-wide list carries
    its last entry at column 99 and nothing else."""
    import svs
    rng = np.random.default_rng(2031)
    m = 90

    def lift(v):
        v = np.asarray(v, np.int64)
        return np.array([int(x) % P for x in v], np.uint64)

    mps = []
    for w in (64, 100, 768, 1024):
        x = np.zeros((m, 1024), np.int64)
        x[:, :w] = rng.integers(-(1 << 20), 1 << 20, size=(m, w))
        if w == 100:
            x[3, :] = 0
            x[3, 99] = -5
        mps.append(lift(x.reshape(-1)))
    clusters = [O.build_cluster(mp, m, l_sub, bytes([70 + i] * 32)) for i, mp in enumerate(mps)]
    idx = svs.GpuIndex(ctx, len(clusters), l_sub)
    for i, cl in enumerate(clusters):
        idx.upload_list(i, cl, ids=1000 * i + np.arange(m, dtype=np.uint32))
    nq, k = 6, 10
    v_q = lift(rng.integers(-(1 << 20), 1 << 20, size=(nq, 1024)).reshape(-1)).reshape(nq, 1024)
    seeds = rng.integers(0, 256, size=(nq, 32), dtype=np.uint8)
    for probes in ([[0]] * nq, [[1]] * nq, [[0, 1, 2, 3], [1], [3, 1], [2, 0], [1, 2], [3]]):
        pp = np.concatenate([[0], np.cumsum([len(p) for p in probes])]).astype(np.uint32)
        pl = np.concatenate(probes).astype(np.uint32)
        want = [_oracle_probe_scores(clusters, probes[q], v_q[q], bytes(seeds[q]), l_sub, k)
                for q in range(nq)]
        for mode in (0, 1, 2):
            say("plaintext widths probes", probes, "mode", mode)
            ids, scores, count, status = idx.score_batch(v_q, seeds, pp, pl, None, k, mode=mode)
            for q in range(nq):
                assert status[q] == 0 and count[q] == k
                assert ids[q].tolist() == want[q][0].tolist(), f"mode {mode} query {q}"
                assert scores[q].tolist() == want[q][1].tolist(), f"mode {mode} query {q}"
    # the standalone one-list decode
    v_enc, g, t_idx, t_sign = O.encode_query(v_q[0], l_sub, bytes(seeds[0]))
    for i, cl in enumerate(clusters):
        r = O.matvec(cl["m_enc"], m, 1024, v_enc)
        want = O.decode_scores(r, m, cl["a_l"], cl["m_plain"], cl["h"], cl["s_rowptr"],
                               cl["s_idx"], cl["s_sign"], l_sub, v_enc, g, t_idx, t_sign)
        assert np.array_equal(idx.decode(i, r, v_enc, g, t_idx, t_sign), want), f"list {i}"


def test_topk_ties_across_chunks(ctx):
    """Duplicate rows give equal scores; the top-k keeps the earliest
    positions (score desc, position asc) across 1024-candidate chunks, for
    k = 1, 10, 16 (two-phase kernel) and 17, 40 (block kernel)."""
    import svs
    from svs import data
    base = data.make_corpus(7, components=2, seed=5).vectors
    rng = np.random.default_rng(6)
    rows = base[rng.integers(0, 7, size=3000)]
    qs = data.make_queries(data.make_corpus(50, components=2, seed=5), 3, seed=8)
    cfg = svs.BnTmConfig(key_seed=bytes([42] * 32), verification_enabled=False)
    scorer = svs.BnTmScorer(ctx)
    h_seed = bytes([9] * 32)
    handle, _ = scorer.upload_cluster(cfg, rows, h_seed=h_seed)
    l = O.generate_subspace(bytes([42] * 32))
    mq = np.stack([O.quantise_vector(v) for v in rows])
    cl = O.build_cluster(mq.reshape(-1), rows.shape[0], l, h_seed)
    seeds = np.full(3 * 32, 3, np.uint8)
    for k in (1, 10, 16, 17, 40):
        say("ties k =", k)
        got = scorer.score_batch(handle, qs, k, q_seeds=seeds)
        for i, q in enumerate(qs):
            st, ids, scores = O.bntm_flat_score(cl, l, O.quantise_vector(q), bytes(seeds[:32]),
                                                False, k)
            assert [h.id for h in got[i]] == ids.tolist(), f"k {k} query {i}"
            assert len(set(scores.tolist())) < k or k == 1


@pytest.mark.parametrize("variant", range(10))
def test_tile_variants_match_oracle(ctx, ivf_setup, l_sub, variant):
    """Every block-tile variant (query group 4/8/16, ring depth, waves per
    SIMD) of the MFMA batch path is bit-exact: whole IVF queries and the
    hot-list product."""
    import svs
    s = ivf_setup
    ctx.set_tile_variant(variant)
    try:
        assert ctx.tile_variant(1)["variant"] == variant
        seeds = np.arange(12 * 32, dtype=np.uint64).astype(np.uint8) ^ np.uint8(0x33)
        say("tile variant", variant, ctx.tile_variant(12))
        got = s["scorer"].score_batch(s["handle"], s["qs"], 10, q_seeds=seeds,
                                      mode=svs._lib.MODE_BATCHED)
        for i, q in enumerate(s["qs"]):
            ids, scores = O.ivf_score(s["clusters"], s["lists"], s["centres"], s["l"], q, 4,
                                      seeds[i * 32:(i + 1) * 32], True, 10)
            assert [h.id for h in got[i]] == ids.tolist(), f"variant {variant} query {i}"
        rng = np.random.default_rng(500 + variant)
        m = 300
        cl = _random_cluster(rng, m)
        idx = svs.GpuIndex(ctx, 1, l_sub)
        idx.upload_list(0, cl)
        nq = 37
        v = _field(rng, (nq, 1024))
        got = idx.compute_products_batch(v, np.arange(nq + 1, dtype=np.uint32),
                                         np.zeros(nq, np.uint32), svs._lib.KERNEL_MFMA)
        want = O.matvec_batch(cl["m_enc"], m, 1024, v)
        for q in range(nq):
            assert np.array_equal(got[q], want[q]), f"variant {variant} query {q}"
    finally:
        ctx.set_tile_variant(0)


def test_autotune_picks_a_variant(ctx, ivf_setup):
    import svs
    s = ivf_setup
    h = s["handle"]
    probes = svs.probe_route(s["centres"], s["qs"], 4)
    v_q = svs.quantise(s["qs"])
    seeds = np.zeros((12, 32), np.uint8)
    pp = np.arange(0, 12 * 4 + 1, 4, dtype=np.uint32)
    say("autotune")
    try:
        best, ms = h.index.autotune(v_q, seeds, pp, probes.reshape(-1), None, 10, reps=1)
        assert 0 <= best < svs._lib.TILE_VARIANTS and len(ms) == svs._lib.TILE_VARIANTS
        assert ctx.tile_variant(12)["variant"] == best
        assert all(x > 0 for x in ms)
    finally:
        ctx.set_tile_variant(0)


def test_score_batch_dev_matches_host(ctx, ivf_setup):
    """svs_score_batch_dev (device-resident queries and outputs) returns the
    host entry point's results; a non-canonical v_q is rejected."""
    import svs
    s = ivf_setup
    h = s["handle"]
    probes = svs.probe_route(s["centres"], s["qs"], 4)
    v_q = svs.quantise(s["qs"])
    seeds = np.arange(12 * 32, dtype=np.uint64).astype(np.uint8).reshape(12, 32)
    pp = np.arange(0, 12 * 4 + 1, 4, dtype=np.uint32)
    ver = (np.arange(12) % 2).astype(np.uint8)
    want = h.index.score_batch(v_q, seeds, pp, probes.reshape(-1), ver, 10)
    dv = svs.DeviceArray.from_host(ctx, v_q)
    ds = svs.DeviceArray.from_host(ctx, seeds)
    ids = svs.DeviceArray(ctx, (12, 10), np.uint32)
    sc = svs.DeviceArray(ctx, (12, 10), np.int64)
    cnt = svs.DeviceArray(ctx, (12,), np.uint32)
    st = svs.DeviceArray(ctx, (12,), np.int32)
    say("score_batch_dev")
    h.index.score_batch_dev(12, dv.ptr(), ds.ptr(), pp, probes.reshape(-1), ver, 10,
                            ids.ptr(), sc.ptr(), cnt.ptr(), st.ptr())
    assert np.array_equal(ids.to_host(), want[0])
    assert np.array_equal(sc.to_host(), want[1])
    assert np.array_equal(cnt.to_host(), want[2])
    assert np.array_equal(st.to_host(), want[3])
    bad = v_q.copy()
    bad[3, 7] = np.uint64(O.P)
    dv.copy_from(bad)
    with pytest.raises(svs.GpuError):
        h.index.score_batch_dev(12, dv.ptr(), ds.ptr(), pp, probes.reshape(-1), ver, 10,
                                ids.ptr(), sc.ptr(), cnt.ptr(), st.ptr())
    for d in (dv, ds, ids, sc, cnt, st):
        d.close()


def test_topk_standalone_matches_oracle(ctx):
    """svs_topk (fp_to_signed + Reverse(score) sort + truncate, lib.rs:452-469)
    against the oracle: empty and short segments, ties (many equal scores),
    the signed boundary at p/2, k up to 128."""
    rng = np.random.default_rng(12)
    sizes = [0, 1, 5, 17, 128, 129, 1000, 3000, 0, 64]
    seg = np.zeros(len(sizes) + 1, np.uint32)
    seg[1:] = np.cumsum(sizes)
    n = int(seg[-1])
    mv = rng.integers(0, P, n, dtype=np.uint64)
    mv[::7] = np.uint64(12345)                      # ties
    mv[1::11] = np.uint64(P // 2)                   # the most negative score
    mv[2::13] = np.uint64(P // 2 - 1)               # the largest score
    mv[3::17] = np.uint64(0)
    ids = rng.permutation(n).astype(np.uint32)
    for k in (1, 10, 128):
        say("standalone top-k k =", k)
        oi, os_, oc = ctx.topk(seg, mv, ids, k)
        for q in range(len(sizes)):
            a, b = int(seg[q]), int(seg[q + 1])
            wi, ws = O.topk(ids[a:b], O.fp_to_signed_vec(mv[a:b]), k)
            assert oc[q] == wi.size, f"segment {q}"
            assert oi[q, :oc[q]].tolist() == wi.tolist(), f"segment {q}"
            assert os_[q, :oc[q]].tolist() == ws.tolist(), f"segment {q}"
