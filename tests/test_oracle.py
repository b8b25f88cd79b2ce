"""Pins the CPU oracle against published vectors, the reference's own
known-answer constants, exact big-int closed forms and protocol invariants.

Mirrors the reference's unit tests in scorer-bntm/src/crypto.rs:1070-1560,
scorer-bntm/src/params.rs:197-285, scorer-emvp/src/crypto.rs:455-543 and
scorer-tiptoe/src/pir/simplepir.rs:280-501.
"""
import hashlib
import struct

import numpy as np
import pytest

from oracle import oracle as O

P = O.P


def _sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def _field(rng, size):
    return rng.integers(0, P, size=size, dtype=np.uint64)


def _exact_matvec(mat, vec):
    v = [int(x) for x in vec]
    return [sum(int(a) * b for a, b in zip(row.tolist(), v)) % P for row in mat]


# ------------------------------------------------------------------ ChaCha20
def test_chacha20_block_rfc7539(golden):
    for tv in golden["chacha20_rfc7539_a1"]:
        key = np.frombuffer(bytes.fromhex(tv["key"]), dtype="<u4").astype(np.uint32)
        words = O.chacha20_block(key, tv["counter"])
        assert struct.pack("<16I", *[int(w) for w in words]).hex() == tv["keystream"]


def test_stream_conventions_cross_implementation(golden):
    # generate_subspace draws sample_fp from ChaCha20Rng(seed): its first
    # entries are the stream's first field samples (crypto.rs:235-238).
    for st in golden["streams"]:
        seed = bytes.fromhex(st["seed"])
        l = O.generate_subspace(seed, 8, 1)
        assert [str(int(x)) for x in l] == st["first_fp"]
        assert O.domain_separate(seed, b"protocol2").tobytes().hex() == st["protocol2"]


def test_chacha20rng_stream_against_rfc7539(golden):
    """ChaCha20Rng::from_seed (rand_chacha 0.10, the reference's workspace
    dependency) takes the 32-byte seed as the RFC 7539 key and emits block 0,
    1, 2, ... of the zero-nonce stream, 16 u32 words per block; next_u64 is
    (word 2i+1 << 32) | word 2i (rand_core's BlockRng), and random::<u8>()
    is next_u32() truncated (rand's StandardUniform). The same construction
    as rand_chacha's own test_chacha_true_values_{a,b,c}: the RFC A.1 keys
    used as seeds, earlier blocks skipped. Here it goes through the oracle's
    sample_fp (crypto.rs:155-162) and derive_h_seed (ivf.rs:258-264), so
    seed byte order, block sequencing and word order are all checked
    against the published keystreams."""
    for tv in golden["chacha20_rfc7539_a1"]:
        seed = bytes.fromhex(tv["key"])
        words = struct.unpack("<16I", bytes.fromhex(tv["keystream"]))
        c = tv["counter"]
        draws = O.u_draws(seed, 8 * (c + 1))[8 * c:]
        want = [((words[2 * i + 1] << 32) | words[2 * i]) & P for i in range(8)]
        assert all(w < P for w in want)  # no rejection inside these blocks
        assert draws == want, tv["key"]
    # derive_h_seed(0, 0) seeds ChaCha20Rng with 32 zero bytes: its 32 bytes
    # are the low bytes of the zero key's blocks 0 and 1
    blocks = {tv["counter"]: struct.unpack("<16I", bytes.fromhex(tv["keystream"]))
              for tv in golden["chacha20_rfc7539_a1"] if int(tv["key"], 16) == 0}
    want = bytes(w & 0xFF for w in blocks[0] + blocks[1])
    assert O.derive_h_seed(0, 0).tobytes() == want


# --------------------------------------------------------------------- field
def test_field_kat(golden):
    for row in golden["field_kat"]:
        op = row[0]
        vals = [int(v) for v in row[1:]]
        if op == "reduce":
            assert O.mersenne_reduce(vals[0]) == vals[1], row
        else:
            got = {"add": O.add_mod, "sub": O.sub_mod, "mul": O.mul_mod}[op](vals[0], vals[1])
            assert got == vals[2], row


def test_params_and_threshold(golden):
    p = golden["params"]
    assert O.sparse_threshold() == p["sparse_threshold"] == (1 << 29) - 1
    assert p["n1_by_k"]["3"] == O.N1 and p["lambda_prime"] == O.LAMBDA


def test_quantise_half_away_from_zero(golden):
    q = golden["quantise"]
    out = O.quantise_vector(np.array(q["v"], np.float32), q["n"], q["q"])
    assert [str(int(x)) for x in out] == q["out"]


def test_fp_to_signed(golden):
    for s, want in golden["fp_to_signed"]:
        assert O.fp_to_signed(int(s)) == int(want)


# ------------------------------------------------------------ server matvec
def test_matvec_golden(golden):
    for case in golden["matvec"]:
        rng = np.random.default_rng(case["seed"])
        mat = _field(rng, (case["rows"], case["cols"]))
        vec = _field(rng, case["cols"])
        assert _sha(mat, vec) == case["sha"], "input generator drifted"
        got = O.matvec(mat, case["rows"], case["cols"], vec)
        assert [str(int(x)) for x in got] == case["out"]


def test_matvec_max_accumulator(golden):
    c = golden["matvec_all_pm1"]
    mat = np.full((c["rows"], c["cols"]), P - 1, np.uint64)
    vec = np.full(c["cols"], P - 1, np.uint64)
    assert all(str(int(x)) == c["out"] for x in O.matvec(mat, c["rows"], c["cols"], vec))


def test_matvec_one_hot_extracts_column():
    rng = np.random.default_rng(3)
    mat = _field(rng, (9, 1024))
    for col in (0, 1, 511, 1023):
        v = np.zeros(1024, np.uint64)
        v[col] = 1
        assert np.array_equal(O.matvec(mat, 9, 1024, v), mat[:, col])


def test_matvec_transpose_golden(golden):
    c = golden["matvec_transpose"]
    rng = np.random.default_rng(c["seed"])
    mat = _field(rng, (c["rows"], c["cols"]))
    vec = _field(rng, c["rows"])
    assert _sha(mat, vec) == c["sha"]
    got = O.matvec_transpose(mat, c["rows"], c["cols"], vec)
    assert [str(int(x)) for x in got] == c["out"]


@pytest.mark.parametrize("big_b", [1, 8, 64])
def test_batch_matches_per_query(big_b):
    # crypto.rs:1533-1559
    rng = np.random.default_rng(57 + big_b)
    m = 16
    m_enc = _field(rng, (m, 1024))
    vs = _field(rng, (big_b, 1024))
    batch = O.matvec_batch(m_enc, m, 1024, vs)
    for q in range(big_b):
        assert np.array_equal(batch[q], O.matvec(m_enc, m, 1024, vs[q]))


def test_matvec_mt_matches_scalar():
    rng = np.random.default_rng(5)
    mat = _field(rng, (257, 1024))
    vec = _field(rng, 1024)
    assert np.array_equal(O.matvec_mt(mat, 257, 1024, vec, 4), O.matvec(mat, 257, 1024, vec))


@pytest.mark.skipif(not O.has_avx512f(), reason="this CPU lacks AVX-512F")
def test_matvec_avx512_matches_scalar(golden):
    """mersenne_matvec_avx512 (crypto.rs:643-739) equals the scalar matvec
    at the reference gate's shapes (crypto.rs:1184-1198), at the all-(p-1)
    extreme (every lane partial and fold at its maximum) and against the
    golden big-int sums; cols % 8 != 0 takes the scalar row."""
    rng = np.random.default_rng(0xC0DE)
    for rows, cols in ((100, 512), (100, 1024), (3, 8), (5, 64), (7, 12)):
        mat = _field(rng, (rows, cols))
        vec = _field(rng, cols)
        assert np.array_equal(O.matvec_avx512(mat, rows, cols, vec),
                              O.matvec(mat, rows, cols, vec)), (rows, cols)
    c = golden["matvec_all_pm1"]
    mat = np.full((c["rows"], c["cols"]), P - 1, np.uint64)
    vec = np.full(c["cols"], P - 1, np.uint64)
    assert all(str(int(x)) == c["out"] for x in O.matvec_avx512(mat, c["rows"], c["cols"], vec))
    for case in golden["matvec"]:
        r = np.random.default_rng(case["seed"])
        mat = _field(r, (case["rows"], case["cols"]))
        vec = _field(r, case["cols"])
        got = O.matvec_avx512(mat, case["rows"], case["cols"], vec)
        assert [str(int(x)) for x in got] == case["out"]
    # one-hot vectors pick columns: a wrong lane recombination would not
    mat = _field(rng, (4, 1024))
    for col in (0, 7, 8, 1023):
        v = np.zeros(1024, np.uint64)
        v[col] = 1
        assert np.array_equal(O.matvec_avx512(mat, 4, 1024, v), mat[:, col])


@pytest.mark.parametrize("simd", [False, True])
def test_cpu_ivf_rows_mt_equals_ivf_score(simd):
    """The timed CPU path (closed-loop B=1, rows across threads) returns
    exactly BnTmIvfScorer::score's hits, including empty and one-row lists,
    lists smaller than the thread count and ties in probe order."""
    if simd and not O.has_avx512f():
        pytest.skip("this CPU lacks AVX-512F")
    rng = np.random.default_rng(77)
    l = O.generate_subspace(bytes([3] * 32))
    sizes = [0, 1, 5, 40, 129, 7]
    lists, clusters, nxt = [], [], 0
    for ci, m in enumerate(sizes):
        vecs = rng.standard_normal((m, 768)).astype(np.float32) / 28.0
        mq = O.quantise_rows(vecs) if m else np.zeros((0, O.N), np.uint64)
        clusters.append(O.build_cluster(mq.reshape(-1), m, l, O.derive_h_seed(5, ci)))
        lists.append(np.arange(nxt, nxt + m, dtype=np.uint32))
        nxt += m
    centroids = rng.standard_normal((len(sizes), 768)).astype(np.float32)
    cpu = O.CpuIvf(clusters, lists, centroids, l)
    for q in range(4):
        query = rng.standard_normal(768).astype(np.float32) / 28.0
        seed = bytes([q + 1] * 32)
        for nprobe, k in ((6, 10), (3, 4), (1, 2)):
            want = O.ivf_score(clusters, lists, centroids, l, query, nprobe, seed, False, k)
            for threads in (1, 3, 8):
                got = cpu.score(query, nprobe, seed, k, threads, simd)
                assert got[0].tolist() == want[0].tolist(), (q, nprobe, threads)
                assert got[1].tolist() == want[1].tolist(), (q, nprobe, threads)


# ------------------------------------------------------ protocol invariants
def _toy_cluster(seed_byte, m, n1=128):
    rng = np.random.default_rng(seed_byte)
    m_plain = _field(rng, (m, 1024))
    l = O.generate_subspace(bytes([seed_byte + 1] * 32), 1024, n1)
    h, rowptr, idx, sign = O.regenerate_h_s(bytes([seed_byte + 2] * 32), m, 1024, n1)
    m_enc, a_l = O.encrypt_cluster(m_plain, m, l, h, rowptr, idx, sign, 1024, n1)
    return dict(m=m, m_plain=m_plain.reshape(-1), m_enc=m_enc, a_l=a_l, h=h, s_rowptr=rowptr,
                s_idx=idx, s_sign=sign), l, rng


@pytest.mark.parametrize("n1", [512, 256, 128, 64])
def test_protocol_roundtrip_recovers_mv_exactly(n1):
    # crypto.rs:1232-1266 and the delta sweep :1306-1343
    cl, l, rng = _toy_cluster(7 + n1 % 13, 16, n1)
    v = _field(rng, 1024)
    v_enc, g, t_idx, t_sign = O.encode_query(v, l, bytes([9] * 32), 1024, n1)
    r = O.matvec(cl["m_enc"], 16, 1024, v_enc)
    mv = O.decode_scores(r, 16, cl["a_l"], cl["m_plain"], cl["h"], cl["s_rowptr"], cl["s_idx"],
                         cl["s_sign"], l, v_enc, g, t_idx, t_sign, 1024, n1)
    want = _exact_matvec(cl["m_plain"].reshape(16, 1024), v)
    assert [int(x) for x in mv] == want


def test_protocol_two_independent_queries():
    # crypto.rs:1349-1376
    cl, l, rng = _toy_cluster(11, 16)
    for q_seed in (bytes([13] * 32), bytes([17] * 32)):
        v = _field(rng, 1024)
        v_enc, g, ti, ts = O.encode_query(v, l, q_seed)
        r = O.matvec(cl["m_enc"], 16, 1024, v_enc)
        mv = O.decode_scores(r, 16, cl["a_l"], cl["m_plain"], cl["h"], cl["s_rowptr"],
                             cl["s_idx"], cl["s_sign"], l, v_enc, g, ti, ts)
        assert [int(x) for x in mv] == _exact_matvec(cl["m_plain"].reshape(16, 1024), v)


def test_encrypt_cluster_closed_form():
    # M_enc = M + H L^T + S elementwise, checked in exact big ints on 2 rows.
    cl, l, _ = _toy_cluster(23, 2)
    L = l.reshape(1024, 128)
    H = cl["h"].reshape(2, 128)
    M = cl["m_plain"].reshape(2, 1024)
    for i in range(2):
        s_row = {int(j): int(s) for j, s in zip(cl["s_idx"][cl["s_rowptr"][i]:cl["s_rowptr"][i + 1]],
                                                cl["s_sign"][cl["s_rowptr"][i]:cl["s_rowptr"][i + 1]])}
        for j in (0, 5, 777, 1023):
            hl = sum(int(H[i, k]) * int(L[j, k]) for k in range(128))
            want = (int(M[i, j]) + hl + s_row.get(j, 0)) % P
            assert int(cl["m_enc"][i * 1024 + j]) == want
        for k in (0, 127):
            assert int(cl["a_l"][i * 128 + k]) == sum(int(M[i, j]) * int(L[j, k])
                                                      for j in range(1024)) % P


def test_h_s_regeneration_deterministic_and_density():
    # crypto.rs:1379-1409 (mu = 0.125 at Sec128, params.rs:242-244)
    a = O.regenerate_h_s(bytes([42] * 32), 64)
    b = O.regenerate_h_s(bytes([42] * 32), 64)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    mean = a[2].size / 64
    assert abs(mean - 128) < 64
    # rows are sorted with +-1 signs
    for i in range(64):
        row = a[2][a[1][i]:a[1][i + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0)
    assert set(np.unique(a[3]).tolist()) <= {-1, 1}


def test_verification_passes_on_honest_server():
    # crypto.rs:1415-1439
    cl, l, rng = _toy_cluster(21, 16)
    for b in range(16):
        v = _field(rng, 1024)
        v_enc, *_ = O.encode_query(v, l, bytes([b] * 32))
        r = O.matvec(cl["m_enc"], 16, 1024, v_enc)
        assert O.verify_response(r, cl["m_enc"], 16, v_enc, bytes([b ^ 0xAA] * 32)) == -1


def test_verification_detects_tampered_response():
    # crypto.rs:1446-1474: one coordinate + 1 -> rejected at trial 0
    cl, l, rng = _toy_cluster(31, 16)
    v = _field(rng, 1024)
    v_enc, *_ = O.encode_query(v, l, bytes([34] * 32))
    r = O.matvec(cl["m_enc"], 16, 1024, v_enc)
    r[8] = O.add_mod(int(r[8]), 1)
    assert O.verify_response(r, cl["m_enc"], 16, v_enc, bytes([35] * 32)) == 0


def test_flat_score_topk_equals_plaintext_ranking():
    # Zero-noise: top-k of decode == top-k of the exact quantised scores.
    cl, l, rng = _toy_cluster(41, 40)
    v = _field(rng, 1024)
    st, ids, scores = O.bntm_flat_score(cl, l, v, bytes([3] * 32), True, 10)
    assert st == 0
    exact = _exact_matvec(cl["m_plain"].reshape(40, 1024), v)
    signed = [x if x < P // 2 else x - P for x in exact]
    order = sorted(range(40), key=lambda i: (-signed[i], i))[:10]
    assert ids.tolist() == order
    assert scores.tolist() == [signed[i] for i in order]


def test_k_capped_at_cluster_size():
    # lib.rs:1069-1083
    cl, l, rng = _toy_cluster(5, 3)
    st, ids, _ = O.bntm_flat_score(cl, l, _field(rng, 1024), bytes([8] * 32), False, 100)
    assert st == 0 and ids.size == 3


# --------------------------------------------------------------- top-k/route
def test_topk_tie_rule(golden):
    t = golden["topk"]
    ids, scores = O.topk(t["ids"], t["scores"], t["k"])
    assert ids.tolist() == t["out_ids"] and scores.tolist() == t["out_scores"]


def test_probe_route_golden(golden):
    g = golden["probe_route"]
    rng = np.random.default_rng(g["seed"])
    cents = rng.standard_normal((20, 24)).astype(np.float32)
    qs = rng.standard_normal((4, 24)).astype(np.float32)
    assert _sha(cents, qs) == g["sha"]
    for q, want in zip(qs, g["out"]):
        assert O.probe_route(cents, q, g["nprobe"]).tolist() == want


# --------------------------------------------------------------------- EMVP
def test_emvp_golden_and_kat(golden):
    e = golden["emvp"]
    rng = np.random.default_rng(e["seed"])
    m_hat = _field(rng, (e["m"], 1292))
    q_hat = _field(rng, 1292)
    assert _sha(m_hat, q_hat) == e["sha"]
    got = O.emvp_compute_products(m_hat, e["m"], q_hat)
    assert [str(int(x)) for x in got] == e["out"]
    # scorer-emvp/src/crypto.rs:463-498
    x = (P - 1) * (P - 1)
    assert O.emvp_fp_reduce(x) == 1 == int(e["kat_mul_pm1"])
    assert O.add_mod(P - 1, 1) == 0 and O.add_mod(P - 1, P - 1) == P - 2


def test_emvp_batch_matches_per_query():
    rng = np.random.default_rng(77)
    m_hat = _field(rng, (5, 1292))
    qs = _field(rng, (4, 1292))
    b = O.emvp_compute_products_batch(m_hat, 5, qs)
    for q in range(4):
        assert np.array_equal(b[q], O.emvp_compute_products(m_hat, 5, qs[q]))


# ------------------------------------------------------------------- Tiptoe
def test_emvp_client_roundtrip_exact_and_reference_smoke():
    """EMVP client side (scorer-emvp/src/crypto.rs): decode(compute(encrypt(
    M0), encode(q, u))) = fl32(signed(M0 . q mod p) / 2^40) for random masks
    u (the D' u and R terms cancel exactly); D = [I | D'] puts the identity
    where encode_query assumes it; and the reference's own
    encode_decode_single_vector (crypto.rs:509-542)."""
    rng = np.random.default_rng(41)
    d = O.emvp_generate_d(bytes([5] * 32))
    assert np.array_equal(d[:, :1024], np.eye(1024, dtype=np.uint64))
    assert (d[:, 1024:] < P).all() and len(set(d[:, 1024:].reshape(-1).tolist())) > 270000
    m = 3
    vecs = rng.standard_normal((m, 768)).astype(np.float32) / 28.0  # d < ell0: zero-padded
    r_seed = bytes(range(32))
    m_hat = O.emvp_encrypt_cluster(d, r_seed, vecs)
    for t in range(2):
        q = rng.standard_normal(1024) / 32.0
        u = rng.integers(0, P, 268, dtype=np.uint64)
        q_hat = O.emvp_encode_query(d, q, u)
        assert np.array_equal(q_hat[1024:], u)
        got = O.emvp_decode_scores(m, r_seed, q_hat, O.emvp_compute_products(m_hat, m, q_hat))
        for i in range(m):
            t_ = sum(O.emvp_quantize(vecs[i][l]) * O.emvp_quantize(np.float32(q[l]))
                     for l in range(768)) % P
            t_ = t_ - P if t_ > P // 2 else t_
            assert got[i] == np.float32(np.float64(t_) / 2.0 ** 40), (t, i)
    # q = 0 with a random mask: the true score is 0, and only the right
    # r_seed cancels R . q_hat
    q_hat = O.emvp_encode_query(d, np.zeros(1024), rng.integers(0, P, 268, dtype=np.uint64))
    parts = O.emvp_compute_products(m_hat, m, q_hat)
    assert (O.emvp_decode_scores(m, r_seed, q_hat, parts) == 0).all()
    assert (O.emvp_decode_scores(m, bytes(32), q_hat, parts) != 0).all()
    dim = 1024
    v = np.array([(i / dim) * 0.01 for i in range(dim)], np.float32)
    qq = np.array([((dim - i) / dim) * 0.01 for i in range(dim)], np.float64)
    d0 = O.emvp_generate_d(bytes(32))
    mh = O.emvp_encrypt_cluster(d0, bytes([1] * 32), v[None, :])
    qh = O.emvp_encode_query(d0, qq, np.zeros(268, np.uint64))
    dec = O.emvp_decode_scores(1, bytes([1] * 32), qh, O.emvp_compute_products(mh, 1, qh))[0]
    ip = np.float32(0)
    for a, b in zip(v, qq):
        ip = np.float32(ip + a * np.float32(b))
    assert abs((dec - ip) / max(abs(ip), 1e-9)) < 1e-3
    # quantize and decode_field at their edges (crypto.rs:47-62)
    assert O.emvp_quantize(0.5) == 524288 and O.emvp_quantize(-0.5) == P - 524288
    assert O.emvp_decode_field(P - 1) == np.float32(-1.0 / 2 ** 40)
    assert O.emvp_decode_field(P // 2) == np.float32(float(P // 2) / 2 ** 40)


# ------------------------------------------------ Tiptoe BFV compression
def _is_prime(n):
    if n < 2:
        return False
    d, r = n - 1, 0
    while d % 2 == 0:
        d, r = d // 2, r + 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(r - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def test_bfv_modulus_ntt_and_roundtrip():
    """The oracle's BFV at tiptoe_text_params' sizes (bfv/mod.rs:50-70,
    N = 2048, t = 65537, one 38-bit modulus): q is the largest prime below
    2^38 that is 1 mod 2N; psi is a primitive 2N-th root; the NTT inverts
    and turns pointwise products into negacyclic products (against the
    schoolbook product); encryption decrypts, and a plaintext x ciphertext
    product decrypts to the negacyclic product mod t."""
    q = O.bfv_modulus()
    assert q < 2 ** 38 and q.bit_length() == 38 and (q - 1) % 4096 == 0 and _is_prime(q)
    assert not any(_is_prime(c) for c in range(q + 4096, 2 ** 38, 4096))
    psi = int(O.lib().or_bfv_psi())
    assert pow(psi, 2048, q) == q - 1
    rng = np.random.default_rng(51)
    a = rng.integers(0, q, 2048, dtype=np.uint64)
    b = rng.integers(0, q, 2048, dtype=np.uint64)
    A, B = O.bfv_ntt(a), O.bfv_ntt(b)
    assert np.array_equal(O.bfv_intt(A), a)
    prod = np.array([int(x) * int(y) % q for x, y in zip(A.tolist(), B.tolist())], np.uint64)
    assert np.array_equal(O.bfv_intt(prod), O.bfv_negacyclic_naive(a, b))
    one = np.zeros(2048, np.uint64)
    one[0] = 1
    assert (O.bfv_ntt(one) == 1).all()  # the constant 1 is 1 in every slot
    sk = O.bfv_keygen(bytes(range(32)))
    m = rng.integers(0, 65537, 2048, dtype=np.uint64)
    c0, c1 = O.bfv_encrypt(sk, m, bytes([3] * 32))
    assert np.array_equal(O.bfv_decrypt(sk, c0, c1), m)
    # m x p for a small plaintext p (limb-sized coefficients)
    p = rng.integers(0, 16, 2048, dtype=np.uint64)
    m2 = np.zeros(2048, np.uint64)
    m2[0], m2[5] = 2, 1
    c0, c1 = O.bfv_encrypt(sk, m2, bytes([4] * 32))
    P_ = O.bfv_ntt(p)
    mul = lambda c: np.array([int(x) * int(y) % q for x, y in zip(c.tolist(), P_.tolist())],
                             np.uint64)
    want = [0] * 2048
    for i in range(2048):
        for j, mv in ((0, 2), (5, 1)):
            k = i + j
            if k < 2048:
                want[k] += mv * int(p[i])
            else:
                want[k - 2048] -= mv * int(p[i])
    assert O.bfv_decrypt(sk, mul(c0), mul(c1)).tolist() == [w % 65537 for w in want]


def test_tiptoe_limbs_and_signed_lift():
    """compress.rs:343-370: get_chunk walks the 4-bit limbs low to high;
    from_modulo_p lifts (t/2, t) to negatives."""
    v = 0xFEDCBA9876543210
    assert [O.tiptoe_get_chunk(v, i) for i in range(16)] == list(range(16))
    assert O.tiptoe_get_chunk(0, 15) == 0 and O.tiptoe_get_chunk(2 ** 64 - 1, 0) == 15
    t = 65537
    assert O.tiptoe_from_modulo_t(t, 0) == 0 and O.tiptoe_from_modulo_t(t, t // 2) == t // 2
    assert O.tiptoe_from_modulo_t(t, t // 2 + 1) == -(t // 2)
    assert O.tiptoe_from_modulo_t(t, t - 1) == -1


def test_tiptoe_decompose_hint_recombines_top_32_bits():
    """compress.rs:372-429 on the oracle: the inverse NTT of each limb
    plaintext holds that limb of its column, and the 8 limbs recombine to
    entry & 0xFFFF_FFFF_0000_0000."""
    hint = np.array([0xFEDCBA9876543210, 0x0123456789ABCDEF, 0x00000000FFFFFFFF,
                     0xFFFFFFFF00000000, 0xDEADBEEFCAFEBABE, 0x1111222233334444,
                     0xAAAA5555AAAA5555, 0, 2 ** 64 - 1, 1, 0x8000000000000000,
                     0x0FFFFFFFFFFFFFFF], np.uint64)
    r, n_lwe = 3, 4
    pt = O.tiptoe_decompose_hint(hint, r, n_lwe)
    assert pt.shape == (8, 1, n_lwe, 2048)
    h = hint.reshape(r, n_lwe)
    for col in range(n_lwe):
        coeffs = [O.bfv_intt(pt[b, 0, col]) for b in range(8)]
        for row in range(r):
            re = 0
            for b in range(8):
                assert int(coeffs[b][row]) == O.tiptoe_get_chunk(h[row, col], 15 - b)
                re |= int(coeffs[b][row]) << (4 * (15 - b))
            assert re == int(h[row, col]) & 0xFFFFFFFF00000000
        assert all((c[r:] == 0).all() for c in coeffs)


def test_tiptoe_apply_hint_recovers_top_32_bits_of_h_times_s():
    """compress.rs:445-476 on the oracle (n_lwe 512 here; 2048 in the GPU
    test): H with zero low halves, s in {0,1,2}; decrypting apply_hint's
    output recovers H s exactly; an out-of-range secret is refused
    (compress.rs:211-216)."""
    rng = np.random.default_rng(52)
    r, n_lwe = 4, 512
    hint = ((np.arange(r * n_lwe, dtype=np.uint64) * 17 + 3) & 0xFFFF) << np.uint64(32)
    s = rng.integers(0, 3, n_lwe).astype(np.uint8)
    sk = O.bfv_keygen(bytes([8] * 32))
    ct = O.tiptoe_encrypt_secret(s, sk, bytes([9] * 32))
    ans = O.tiptoe_apply_hint(O.tiptoe_decompose_hint(hint, r, n_lwe), ct)
    got = O.tiptoe_recover_h_s(ans, r, sk)
    h = hint.reshape(r, n_lwe)
    want = [sum(int(h[i, j]) * int(s[j]) for j in range(n_lwe)) % 2 ** 64 for i in range(r)]
    assert got.tolist() == want
    bad = s.copy()
    bad[7] = 3
    with pytest.raises(ValueError):
        O.tiptoe_encrypt_secret(bad, sk, bytes(32))


def test_tiptoe_kat(golden):
    t = golden["tiptoe_kat"]
    m = np.array(t["m"], np.uint64)
    x = np.array(t["x"], np.uint64)
    for col in range(2):
        ans = O.tiptoe_answer_u64(m, 2, 3, np.ascontiguousarray(x[:, col]))
        assert ans.tolist() == [row[col] for row in t["mx"]]


def test_tiptoe_golden_wrapping(golden):
    t = golden["tiptoe"]
    rng = np.random.default_rng(t["seed"])
    mdb = rng.integers(0, 32, size=(t["rows"], t["cols"]), dtype=np.uint64)
    ct = rng.integers(0, 1 << 64, size=t["cols"], dtype=np.uint64)
    assert _sha(mdb, ct) == t["sha"]
    assert [str(int(x)) for x in O.tiptoe_answer_u64(mdb, t["rows"], t["cols"], ct)] == t["ans64"]
    a32 = O.tiptoe_answer_u32(mdb.astype(np.uint32), t["rows"], t["cols"],
                              (ct & np.uint64(0xFFFFFFFF)).astype(np.uint32))
    assert [str(int(x)) for x in a32] == t["ans32"]
