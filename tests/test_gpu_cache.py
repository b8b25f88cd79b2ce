"""Loading the reference's BN-IVF cache file into HBM
(scorer-bntm/src/ivf.rs:1683-1902, scorer-core/src/cache.rs:19-73).

The file is written here, from the oracle's encryption, in the layout the
reference's loader reads; the loaded index must score exactly as the oracle
and as the same lists uploaded directly.
"""
import struct
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

MAGIC = b"\x89SVScach"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fnv_parts(parts):
    """fingerprint64: FNV-1a 64 over the length-prefixed parts."""
    h = 0xcbf29ce484222325
    for part in parts:
        for b in struct.pack("<Q", len(part)) + bytes(part):
            h ^= b
            h = (h * 0x100000001b3) & ((1 << 64) - 1)
    return h


def write_cache(path, parts, centroids, members, vectors, clusters, seeds, l):
    """The reference's cache layout, field by field."""
    with open(path, "wb") as f:
        f.write(MAGIC + struct.pack("<IQ", 1, fnv_parts(parts)))
        n = 20
        d = centroids.shape[1]
        f.write(struct.pack("<QQ", d, centroids.shape[0]))
        f.write(np.ascontiguousarray(centroids, "<f4").tobytes())
        f.write(struct.pack("<Q", len(members)))
        n += 16 + centroids.size * 4 + 8
        for ids in members:
            f.write(struct.pack("<Q", ids.size))
            n += 8
            for i in ids.tolist():
                f.write(struct.pack("<I", i) + np.ascontiguousarray(vectors[i], "<f4").tobytes())
                n += 4 + 4 * d
        f.write(b"\0" * ((8 - n % 8) % 8))
        f.write(struct.pack("<Q", len(clusters)))
        for cl, seed in zip(clusters, seeds):
            f.write(struct.pack("<Q", cl["m"]) + bytes(seed))
            for key in ("m_enc", "a_l", "m_plain", "h"):
                f.write(np.ascontiguousarray(cl[key], "<u8").tobytes())
        f.write(np.ascontiguousarray(l, "<u8").tobytes())


@pytest.fixture(scope="module")
def ctx():
    import svs
    return svs.Context(0)


@pytest.fixture(scope="module")
def cached(ctx, tmp_path_factory):
    import svs
    from svs import data
    corpus = data.make_corpus(1500, components=10, zipf_s=1.1, seed=51)
    lists = data.ivf_lists(corpus)
    lists[2] = np.zeros(0, np.uint32)  # an empty list
    key = bytes([7] * 32)
    l = O.generate_subspace(key)
    clusters, seeds = [], []
    for ci, members in enumerate(lists):
        mq = np.stack([O.quantise_vector(corpus.vectors[i]) for i in members]) if members.size \
            else np.zeros((0, 1024), np.uint64)
        seed = O.derive_h_seed(9, ci)
        clusters.append(O.build_cluster(mq.reshape(-1), members.size, l, seed))
        seeds.append(seed)
    parts = [b"bntm-ivf", b"corpus-51", struct.pack("<Q", 1024)]
    path = tmp_path_factory.mktemp("cache") / "bntm-ivf-cache-test.bin"
    write_cache(path, parts, corpus.centres, lists, corpus.vectors, clusters, seeds, l)
    return dict(path=str(path), parts=parts, corpus=corpus, lists=lists, clusters=clusters, l=l)


def test_fingerprint_matches_fnv(ctx, cached):
    import svs
    assert svs.cache_fingerprint64(cached["parts"]) == fnv_parts(cached["parts"])
    assert svs.cache_fingerprint64([]) == 0xcbf29ce484222325


def test_load_cache_scores_like_oracle(ctx, cached):
    import svs
    c = cached
    handle = svs.load_bntm_ivf_cache(c["path"], ctx, fingerprint=svs.cache_fingerprint64(c["parts"]),
                                     nprobe=4, verification_enabled=True)
    assert handle is not None
    assert handle.sizes.tolist() == [x.size for x in c["lists"]]
    assert np.array_equal(handle.ivf.centroids, np.asarray(c["corpus"].centres, np.float32))
    assert np.array_equal(handle.l_subspace, c["l"])
    qs = svs.data.make_queries(c["corpus"], 8, seed=52)
    seeds = np.arange(8 * 32, dtype=np.uint64).astype(np.uint8)
    got = svs.BnTmIvfScorer(ctx).score_batch(handle, qs, 10, q_seeds=seeds)
    for i, q in enumerate(qs):
        ids, scores = O.ivf_score(c["clusters"], c["lists"], c["corpus"].centres, c["l"], q, 4,
                                  seeds[i * 32:(i + 1) * 32], True, 10)
        assert [h.id for h in got[i]] == ids.tolist(), f"query {i}"
    handle.index.close()


def test_load_cache_header_mismatch_is_none(ctx, cached, tmp_path):
    import svs
    c = cached
    assert svs.load_bntm_ivf_cache(c["path"], ctx, fingerprint=123, nprobe=4) is None
    bad = tmp_path / "bad.bin"
    raw = bytearray(open(c["path"], "rb").read())
    raw[8] = 2  # another version
    bad.write_bytes(bytes(raw))
    assert svs.load_bntm_ivf_cache(str(bad), ctx, nprobe=4) is None


def test_load_cache_truncated_or_tampered_fails(ctx, cached, tmp_path):
    import svs
    c = cached
    raw = open(c["path"], "rb").read()
    short = tmp_path / "short.bin"
    short.write_bytes(raw[: len(raw) - 1000])
    with pytest.raises(svs.GpuError):
        svs.load_bntm_ivf_cache(str(short), ctx, nprobe=4)
    # H is stored; one flipped bit no longer matches the list's h_seed
    tampered = bytearray(raw)
    tampered[len(raw) - 1024 * 128 * 8 - 8] ^= 1  # last word of the last list's H
    t = tmp_path / "tampered.bin"
    t.write_bytes(bytes(tampered))
    with pytest.raises(svs.GpuError):
        svs.load_bntm_ivf_cache(str(t), ctx, nprobe=4)
