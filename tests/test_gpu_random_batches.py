"""Seeded random raw batches against the oracle's BnTmIvfScorer::score
(scorer-bntm/src/ivf.rs:746-822): per-query nprobe, verification flags,
k across the three top-k kernels (<= 16, <= 128, larger), duplicate
queries, both kernel families and graphs on and off — every result bit
for bit."""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("seed, n, comps, iters", [(11, 4000, 24, 30), (23, 12000, 96, 12)])
def test_random_raw_batches_match_oracle(seed, n, comps, iters):
    import svs
    from svs import data
    ctx = svs.Context(0)
    rng = np.random.default_rng(seed)
    corpus = data.make_corpus(n, components=comps, zipf_s=1.1, seed=seed + 1)
    lists = data.ivf_lists(corpus)
    lists[3] = np.zeros(0, np.uint32)  # an empty list
    cfg = svs.BnTmIvfConfig(index=svs.IvfIndex(corpus.centres, lists), nprobe=4,
                            key_seed=bytes([9] * 32), upload_seed=5)
    handle, _ = svs.BnTmIvfScorer(ctx).upload_cluster(cfg, corpus.vectors)
    idx = handle.index
    l_sub = O.generate_subspace(bytes([9] * 32))
    clusters = []
    for ci, mem in enumerate(lists):
        mq = np.stack([O.quantise_vector(corpus.vectors[i]) for i in mem]) if mem.size \
            else np.zeros((0, 1024), np.uint64)
        clusters.append(O.build_cluster(mq.reshape(-1), mem.size, l_sub, O.derive_h_seed(5, ci)))
    pool = data.make_queries(corpus, 200, seed=seed + 2)
    max_np = min(comps, 64)
    try:
        for it in range(iters):
            nq = int(rng.integers(1, 40))
            qs = pool[rng.integers(0, pool.shape[0], nq)].copy()
            if nq > 1 and rng.random() < 0.3:
                qs[0] = qs[1]  # a duplicate query in the batch
            nprobe = rng.integers(1, max_np + 1, nq).astype(np.uint32)
            verify = (rng.random(nq) < 0.5).astype(np.uint8)
            k = int(rng.choice([1, 10, 16, 17, 40, 130]))
            seeds = rng.integers(0, 256, (nq, 32), dtype=np.uint8)
            mode = int(rng.integers(0, 2))
            ctx.set_graphs(bool(rng.integers(0, 2)))
            say("batch", it, "nq", nq, "k", k, "mode", mode)
            ids, scores, count, status, _ = idx.score_raw_batch(qs, seeds, nprobe, verify, k,
                                                                mode=mode)
            for q in range(nq):
                w_ids, w_sc = O.ivf_score(clusters, lists, corpus.centres, l_sub, qs[q],
                                          int(nprobe[q]), bytes(seeds[q]), bool(verify[q]), k)
                assert status[q] == 0 and count[q] == len(w_ids), (it, q)
                assert ids[q, :count[q]].tolist() == w_ids.tolist(), (it, q)
                assert scores[q, :count[q]].tolist() == w_sc.tolist(), (it, q)
    finally:
        ctx.set_graphs(True)
        idx.close()
