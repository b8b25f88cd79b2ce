"""CPU-only checks of the product library: it loads, exports every symbol the
C ABI header declares, and its host-side client functions (keygen,
quantisation, routing) agree with the oracle. No GPU compute is called."""
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    syms = set()
    for name in os.listdir(os.path.join(ROOT, "include")):
        if name.endswith(".h"):
            text = open(os.path.join(ROOT, "include", name)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            syms |= set(re.findall(r"\b(svs_[a-z0-9_]+)\s*\(", text))
    return syms


def test_library_exports_every_declared_symbol():
    import svs
    lib = svs.lib()
    declared = _declared_symbols()
    assert declared, "no declarations found"
    for s in declared:
        assert hasattr(lib, s), f"{s} declared in include/ but not exported"
    assert declared == set(svs.EXPORTS)


def test_library_is_gfx950_code_object():
    import svs
    data = open(svs.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_keygen_matches_oracle():
    import svs
    for seed in (bytes([42] * 32), bytes(range(32))):
        assert np.array_equal(svs.generate_subspace(seed), O.generate_subspace(seed))
        assert np.array_equal(svs.domain_separate(seed, b"protocol2"),
                              O.domain_separate(seed, b"protocol2"))
    for m in (0, 1, 65, 300):
        got = svs.regenerate_h_s(bytes([m % 256] * 32), m)
        want = O.regenerate_h_s(bytes([m % 256] * 32), m)
        for a, b in zip(got, want):
            assert np.array_equal(a, b)


def test_derive_h_seed_is_domain_separated_and_deterministic():
    import svs
    a = svs.derive_h_seed(7, 0)
    assert np.array_equal(a, svs.derive_h_seed(7, 0))
    assert not np.array_equal(a, svs.derive_h_seed(7, 1))
    assert not np.array_equal(a, svs.derive_h_seed(8, 0))
    # ivf.rs:258-264: the derived seed is the first 32 stream bytes (one per
    # u32 word) of ChaCha20Rng(upload_seed LE || cluster LE || 0...).
    seed = (7).to_bytes(8, "little") + (3).to_bytes(8, "little") + bytes(16)
    key = np.frombuffer(seed, "<u4").astype(np.uint32)
    words = np.concatenate([O.chacha20_block(key, 0), O.chacha20_block(key, 1)])
    assert np.array_equal(svs.derive_h_seed(7, 3), (words & 0xFF).astype(np.uint8))


def test_quantise_matches_oracle(golden):
    import svs
    q = golden["quantise"]
    v = np.array(q["v"], np.float32)
    assert [str(int(x)) for x in svs.quantise(v)[0][: q["n"]]] == q["out"]
    rng = np.random.default_rng(3)
    x = rng.standard_normal((50, 768)).astype(np.float32)
    got = svs.quantise(x)
    for i in range(50):
        assert np.array_equal(got[i], O.quantise_vector(x[i]))


def test_quantise_rejects_oversized_dimension():
    import svs
    with pytest.raises(svs.DimensionTooLarge):
        svs.quantise(np.zeros((1, 1025), np.float32))


def test_probe_route_matches_oracle(golden):
    import svs
    g = golden["probe_route"]
    rng = np.random.default_rng(g["seed"])
    cents = rng.standard_normal((20, 24)).astype(np.float32)
    qs = rng.standard_normal((4, 24)).astype(np.float32)
    assert svs.probe_route(cents, qs, g["nprobe"]).tolist() == g["out"]
    from svs import data
    c = data.make_corpus(3000, components=64, zipf_s=1.1, seed=5)
    qs = data.make_queries(c, 40)
    got = svs.probe_route(c.centres, qs, 16)
    for i in range(40):
        assert got[i].tolist() == O.probe_route(c.centres, qs[i], 16).tolist()


def test_synthetic_lists_partition_the_corpus():
    from svs import data
    c = data.make_corpus(5000, components=32, zipf_s=1.1, seed=9)
    lists = data.ivf_lists(c)
    allids = np.concatenate(lists)
    assert np.array_equal(np.sort(allids), np.arange(5000))
    sizes = np.array([len(x) for x in lists])
    assert sizes[0] > sizes[-1]  # Zipf-skewed
    for ci, members in enumerate(lists):
        assert np.all(np.diff(members.astype(np.int64)) > 0)
        assert np.all(c.component[members] == ci)


def test_kmeans_lists_mirror_the_reference_tests():
    """The seeded k-means fixture against kmeans.rs's own tests (140-200):
    deterministic under the seed; every point assigned; three well-separated
    groups get one centroid each. Then its lists: a partition of the ids,
    ascending within a list, every vector in the list of its nearest
    centroid (brute force, float64), and a centroid equal to its list's mean
    after Lloyd converged."""
    from svs import data
    rng = np.random.default_rng(11)
    x = rng.random((400, 8), dtype=np.float32)
    c1 = data.kmeans(x, 12, seed=99, max_iter=20)
    c2 = data.kmeans(x, 12, seed=99, max_iter=20)
    assert np.array_equal(c1, c2)
    a = data._nearest(x, c1)
    assert a.shape == (400,) and (a < 12).all()
    pts = np.array([[base + i * 0.01, i * 0.01] for base in (0.0, 100.0, 200.0)
                    for i in range(30)], np.float32)
    c3 = data.kmeans(pts, 3, seed=1, max_iter=25)
    assert sorted(int(round(c[0] / 100.0)) for c in c3) == [0, 1, 2]
    corpus = data.make_corpus(6000, components=48, zipf_s=1.1, seed=5)
    cents, lists = data.kmeans_lists(corpus, 40, seed=3, max_iter=50, train_sample=6000)
    allids = np.concatenate(lists)
    assert np.array_equal(np.sort(allids), np.arange(6000))
    d = ((corpus.vectors.astype(np.float64)[:, None, :] -
          cents.astype(np.float64)[None, :, :]) ** 2).sum(axis=2)
    best = d.min(axis=1)
    for ci, members in enumerate(lists):
        assert np.all(np.diff(members.astype(np.int64)) > 0)
        # nearest up to float32 rounding of the distance expansion
        assert np.all(d[members, ci] <= best[members] + 1e-4)
        if members.size and not np.array_equal(cents[ci], 0):
            mean = corpus.vectors[members].astype(np.float64).mean(axis=0)
            assert np.allclose(cents[ci], mean, atol=1e-5)


def test_cache_fingerprint_is_fnv1a_over_length_prefixed_parts():
    """fingerprint64 (scorer-core/src/cache.rs:22-37) restated from the
    FNV-1a definition: offset basis 0xcbf29ce484222325, prime 0x100000001b3,
    each part preceded by its u64 little-endian length."""
    import struct
    import svs

    def fnv(parts):
        h = 0xcbf29ce484222325
        for p in parts:
            for b in struct.pack("<Q", len(p)) + p:
                h = ((h ^ b) * 0x100000001b3) & ((1 << 64) - 1)
        return h
    assert svs.cache_fingerprint64([]) == 0xcbf29ce484222325
    for parts in ([b""], [b"a"], [b"bntm", b"\x00\x01\x02", b"x" * 300]):
        assert svs.cache_fingerprint64(parts) == fnv(parts)
    # the length prefix separates ("ab", "") from ("a", "b")
    assert svs.cache_fingerprint64([b"ab", b""]) != svs.cache_fingerprint64([b"a", b"b"])


def test_communication_cost_formulas():
    """communication_cost of the flat and IVF scorers (lib.rs:477-511,
    ivf.rs:928-970; the reference's test lib.rs:1007-1029): query n*8,
    response and setup from the list sizes, the effective bytes the
    product streams per query; the IVF mean rounds half away from zero."""
    import svs
    from svs import bntm as B
    flat = B.BnTmHandle(index=None, l_subspace=None, m=10, dim=32, verification_enabled=False,
                        quantisation_q=B.Q_DEFAULT)
    c = svs.BnTmScorer().communication_cost(flat, 5)
    assert (c.query_bytes, c.response_bytes, c.cluster_response_bytes, c.setup_bytes) == \
        (1024 * 8, 10 * 8, 10 * 8, 10 * 1024 * 8)
    assert (c.pre_query_offline_up_bytes, c.pre_query_offline_down_bytes) == (0, 0)
    assert c.effective_bytes_per_query == 10 * 1024 * 8
    # 16 lists holding 1 row in total: mean 1/16 row, mean bytes 0.5 -> 1
    sizes = np.zeros(16, np.int64)
    sizes[3] = 1
    ivf = B.BnTmIvfHandle(index=None, ivf=None, l_subspace=None, sizes=sizes, nprobe=4, dim=32,
                          verification_enabled=False, quantisation_q=B.Q_DEFAULT)
    c = svs.BnTmIvfScorer().communication_cost(ivf, 10)
    assert c.cluster_response_bytes == 1 and c.response_bytes == 4
    assert c.setup_bytes == 1 * 1024 * 8
    assert c.effective_bytes_per_query == 0  # (mean_m as u64) truncates 0.0625
    sizes = np.array([100, 60, 20], np.int64)
    ivf.sizes, ivf.nprobe = sizes, 2
    c = svs.BnTmIvfScorer().communication_cost(ivf, 10)
    assert c.cluster_response_bytes == 480 and c.response_bytes == 960
    assert c.setup_bytes == 180 * 1024 * 8
    assert c.effective_bytes_per_query == 2 * 60 * 1024 * 8
