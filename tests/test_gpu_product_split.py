"""The batched server product on the staged kernel (work units of 16 rows x
up to four 16-query groups, 32 rows x two groups or 64 rows x one group,
each staging a step's rows and query planes once in LDS for its four
waves) against the per-wave plane kernel: compute_products_batch
(scorer-bntm/src/crypto.rs:782-828) generalised to uneven probe sets, every
(query, list) product against the oracle's matvec, with the staged product
on and off and on both planners; then whole IVF queries against the
oracle's BnTmIvfScorer::score (scorer-bntm/src/ivf.rs:746-822), launched
and through replays of the captured batch graph."""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
P = O.P

# list sizes (partial 16-row blocks, one and several 256-row blocks) and the
# number of queries probing each: one group (1, 5, 16: 64-row units), two
# (17: 32-row units), three and four (33, 64: 16-row units, one wave idle
# at three) and row blocks of several entries (65: four + one, 120: seven
# full groups and a partial one)
SIZES = [300, 17, 256, 1500, 257, 700, 1, 520]
PROBED = [1, 5, 16, 17, 33, 64, 65, 120]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def split_setup():
    import svs
    ctx = svs.Context(0)
    l_sub = O.generate_subspace(bytes([3] * 32))
    rng = np.random.default_rng(0x5EED)
    idx = svs.GpuIndex(ctx, len(SIZES), l_sub)
    cls = []
    for i, m in enumerate(SIZES):
        cl = dict(m=m, m_enc=rng.integers(0, P, m * 1024, dtype=np.uint64),
                  a_l=rng.integers(0, P, m * 128, dtype=np.uint64),
                  h=rng.integers(0, P, m * 128, dtype=np.uint64), m_plain=None,
                  s_rowptr=np.zeros(m + 1, np.uint32), s_idx=np.zeros(0, np.uint32),
                  s_sign=np.zeros(0, np.int8))
        idx.upload_list(i, cl)
        cls.append(cl)
    nq = max(PROBED)
    # query i probes the lists that more than i queries probe, in a rotated
    # order so that items of one list sit at different probe positions
    probe_ptr, probe_lists = [0], []
    for i in range(nq):
        mine = [c for c in range(len(SIZES)) if PROBED[c] > i]
        k = i % len(mine)
        probe_lists += mine[k:] + mine[:k]
        probe_ptr.append(len(probe_lists))
    v = rng.integers(0, P, (nq, 1024), dtype=np.uint64)
    want = {}
    for c, m in enumerate(SIZES):
        qs = [i for i in range(nq) if PROBED[c] > i]
        say("oracle products of list", c, "rows", m, "queries", len(qs))
        w = O.matvec_batch(cls[c]["m_enc"], m, 1024, v[qs])
        for j, q in enumerate(qs):
            want[(q, c)] = w[j]
    yield dict(ctx=ctx, idx=idx, v=v, pp=np.array(probe_ptr, np.uint32),
               pl=np.array(probe_lists, np.uint32), want=want)
    idx.close()


def _check(s, got, what):
    pp, pl = s["pp"], s["pl"]
    for q in range(len(pp) - 1):
        for it in range(pp[q], pp[q + 1]):
            c = int(pl[it])
            assert np.array_equal(got[it], s["want"][(q, c)]), f"{what}: query {q} list {c}"


@pytest.mark.parametrize("variant", [2, 4, 9])
@pytest.mark.parametrize("planner", [True, False])
def test_split_product_matches_oracle(split_setup, variant, planner):
    import svs
    s = split_setup
    ctx = s["ctx"]
    ctx.set_tile_variant(variant)
    ctx.set_planner(planner)
    try:
        assert ctx.tile_variant(120)["query_group"] == 16
        for split in (True, False, True):
            ctx.set_product_split(split)
            say("variant", variant, "device planner", planner, "split", split)
            got = s["idx"].compute_products_batch(s["v"], s["pp"], s["pl"], svs._lib.KERNEL_MFMA)
            _check(s, got, f"variant {variant} split {split}")
    finally:
        ctx.set_product_split(False)
        ctx.set_planner(True)
        ctx.set_tile_variant(0)


def test_split_ivf_queries_match_oracle_and_replays():
    """Whole IVF queries on a 16-query tile variant with the staged product:
    many queries on few lists (row blocks of several quads), through the
    launched batch and two graph replays, each against the oracle."""
    import svs
    from svs import data
    ctx = svs.Context(0)
    corpus = data.make_corpus(5000, components=12, zipf_s=1.1, seed=77)
    lists = data.ivf_lists(corpus)
    cfg = svs.BnTmIvfConfig(index=svs.IvfIndex(corpus.centres, lists), nprobe=5,
                            key_seed=bytes([21] * 32), upload_seed=3, verification_enabled=True)
    scorer = svs.BnTmIvfScorer(ctx)
    handle, _ = scorer.upload_cluster(cfg, corpus.vectors)
    l_sub = O.generate_subspace(bytes([21] * 32))
    clusters = []
    for ci, mem in enumerate(lists):
        mq = np.stack([O.quantise_vector(corpus.vectors[i]) for i in mem])
        clusters.append(O.build_cluster(mq.reshape(-1), mem.size, l_sub, O.derive_h_seed(3, ci)))
    qs = data.make_queries(corpus, 90, seed=78)
    ctx.set_tile_variant(4)
    ctx.set_product_split(True)
    try:
        for rep in range(3):
            seeds = np.random.default_rng(rep).integers(0, 256, 90 * 32, dtype=np.uint8)
            say("split IVF batch", rep)
            got = scorer.score_batch(handle, qs, 10, q_seeds=seeds, mode=svs._lib.MODE_BATCHED)
            for i in range(0, 90, 7 if rep else 1):
                ids, scores = O.ivf_score(clusters, lists, corpus.centres, l_sub, qs[i], 5,
                                          seeds[i * 32:(i + 1) * 32], True, 10)
                assert [h.id for h in got[i]] == ids.tolist(), f"batch {rep} query {i}"
                assert [h.score for h in got[i]] == [float(np.float32(x)) for x in scores], \
                    f"batch {rep} query {i}"
    finally:
        ctx.set_product_split(False)
        ctx.set_tile_variant(0)
        handle.index.close()
