"""GPU parity of EMVP's score GEMV and client decode and Tiptoe's SimplePIR
answer against the oracle (bit-exact). Mirrors scorer-emvp/src/crypto.rs
:463-542 and scorer-tiptoe/src/pir/simplepir.rs:293-376."""
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


def test_emvp_partials_match_oracle(ctx, golden):
    from svs.schemes import EmvpCluster
    e = golden["emvp"]
    rng = np.random.default_rng(e["seed"])
    m_hat = rng.integers(0, P, size=(e["m"], 1292), dtype=np.uint64)
    q_hat = rng.integers(0, P, size=1292, dtype=np.uint64)
    say("emvp golden")
    cl = EmvpCluster(m_hat, ctx)
    assert [str(int(x)) for x in cl.compute_products(q_hat)] == e["out"]
    for m in (1, 7, 63, 64, 65, 257):
        say("emvp m =", m)
        mh = rng.integers(0, P, size=(m, 1292), dtype=np.uint64)
        mh[0, :] = P - 1
        cl = EmvpCluster(mh, ctx)
        qs = rng.integers(0, P, size=(5, 1292), dtype=np.uint64)
        qs[0] = P - 1
        got = cl.compute_products_batch(qs)
        want = O.emvp_compute_products_batch(mh, m, qs)
        assert np.array_equal(got, want)
    assert EmvpCluster(np.zeros((0, 1292), np.uint64), ctx).compute_products(q_hat) is None


def _emvp_plain_scores(vecs, queries):
    """fl32(signed(M0 . q mod p) / 2^40) with M0, q quantised (params.rs Q =
    2^20): what decode must return when R cancels exactly."""
    out = []
    for q in queries:
        qq = [int(O.emvp_quantize(np.float32(x))) for x in q]
        row = []
        for v in vecs:
            t = sum(int(O.emvp_quantize(x)) * y for x, y in zip(v, qq)) % P
            t = t - P if t > P // 2 else t
            row.append(np.float32(np.float64(t) / 2.0 ** 40))
        out.append(row)
    return np.array(out, np.float32)


def test_emvp_decode_roundtrip_exact(ctx):
    """Encrypt with the oracle (generate_d, encrypt_cluster), encode with a
    random mask u, then score on the device (product + decode, partials kept
    on the device): every score equals the plaintext product, exactly. Also
    the reference's encode_decode_single_vector (crypto.rs:509-542)."""
    from svs.schemes import EmvpCluster
    rng = np.random.default_rng(31)
    d = O.emvp_generate_d(bytes(32))
    m = 24
    vecs = rng.standard_normal((m, 1024)).astype(np.float32)
    vecs /= np.linalg.norm(vecs, axis=1, keepdims=True)
    r_seed = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
    say("emvp round trip: encrypting", m, "rows with the oracle")
    m_hat = O.emvp_encrypt_cluster(d, r_seed, vecs)
    queries = rng.standard_normal((3, 1024))
    queries /= np.linalg.norm(queries, axis=1, keepdims=True)
    q_hats = np.stack([O.emvp_encode_query(d, q, rng.integers(0, P, 268, dtype=np.uint64))
                       for q in queries])
    cl = EmvpCluster(m_hat, ctx)
    got = cl.score_batch(r_seed, q_hats)
    want = _emvp_plain_scores(vecs, queries)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    for j in range(3):
        parts = O.emvp_compute_products(m_hat, m, q_hats[j])
        assert np.array_equal(O.emvp_decode_scores(m, r_seed, q_hats[j], parts), want[j])
    cl.close()
    # the reference's single-vector test: key [0; 32], r_seed [1; 32], u = 0
    dim = 1024
    v = np.array([(i / dim) * 0.01 for i in range(dim)], np.float32)
    q = np.array([((dim - i) / dim) * 0.01 for i in range(dim)], np.float64)
    m_hat1 = O.emvp_encrypt_cluster(d, bytes([1] * 32), v[None, :])
    q_hat1 = O.emvp_encode_query(d, q, np.zeros(268, np.uint64))
    one = EmvpCluster(m_hat1, ctx)
    dec = one.score_batch(bytes([1] * 32), q_hat1[None, :])[0, 0]
    ip = np.float32(0)
    for a, b in zip(v, q):
        ip = np.float32(ip + a * np.float32(b))
    assert abs((dec - ip) / max(abs(ip), 1e-9)) < 1e-3
    one.close()


def test_emvp_decode_20000_rows_matches_oracle(ctx):
    """decode_scores on the device at 20,000 x 1292 against the oracle's
    decode_scores (crypto.rs:435-453), bit for bit (f32 bits): fused with
    the product, from host partials (with all-(p-1) and zero partials), and
    from device pointers; an empty cluster and non-canonical inputs."""
    from concurrent.futures import ThreadPoolExecutor
    import svs
    from svs import _lib as L
    from svs.schemes import EmvpCluster
    rng = np.random.default_rng(32)
    m, nq = 20000, 4
    mh = rng.integers(0, P, size=(m, 1292), dtype=np.uint64)
    mh[m - 1, :] = P - 1
    qs = rng.integers(0, P, size=(nq, 1292), dtype=np.uint64)
    qs[nq - 1] = P - 1
    r_seed = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
    cl = EmvpCluster(mh, ctx)
    say("emvp decode: device, m =", m)
    got = cl.score_batch(r_seed, qs)
    parts = cl.compute_products_batch(qs)
    parts[1, :m] = P - 1           # block 0 of every row at its maximum
    parts[2, 5 * m:6 * m] = 0
    with ThreadPoolExecutor(nq) as ex:
        want_fused = list(ex.map(
            lambda j: O.emvp_decode_scores(m, r_seed, qs[j],
                                           O.emvp_compute_products(mh, m, qs[j])), range(nq)))
        want_host = list(ex.map(lambda j: O.emvp_decode_scores(m, r_seed, qs[j], parts[j]),
                                range(nq)))
    say("emvp decode: oracle done")
    for j in range(nq):
        assert np.array_equal(got[j].view(np.uint32), want_fused[j].view(np.uint32)), j
    host = cl.decode_scores_batch(r_seed, qs, parts)
    for j in range(nq):
        assert np.array_equal(host[j].view(np.uint32), want_host[j].view(np.uint32)), j
    qd = svs.DeviceArray.from_host(ctx, qs)
    pd = svs.DeviceArray.from_host(ctx, parts)
    sd = svs.DeviceArray(ctx, (nq, m), np.float32)
    assert cl.decode_scores_dev(nq, r_seed, qd.ptr(), pd.ptr(), sd.ptr())
    assert np.array_equal(sd.to_host().view(np.uint32), host.view(np.uint32))
    for a in (qd, pd, sd):
        a.close()
    bad = parts.copy()
    bad[0, 3] = P
    with pytest.raises(L.GpuError):
        cl.decode_scores_batch(r_seed, qs, bad)
    cl.close()
    empty = EmvpCluster(np.zeros((0, 1292), np.uint64), ctx)
    assert empty.score_batch(r_seed, qs[:1]) is None
    empty.close()


def test_emvp_dev_grid_stride_matches_oracle(ctx):
    """Device-pointer entry at a size whose work units outnumber the grid
    (every block takes several row tiles x column groups), a ragged last
    row tile, and the host-buffer entry on the same inputs."""
    import svs
    from svs.schemes import EmvpCluster
    rng = np.random.default_rng(12)
    m, nq = 5001, 16
    mh = rng.integers(0, P, size=(m, 1292), dtype=np.uint64)
    mh[m - 1, :] = P - 1
    qs = rng.integers(0, P, size=(nq, 1292), dtype=np.uint64)
    qs[nq - 1] = P - 1
    cl = EmvpCluster(mh, ctx)
    want = O.emvp_compute_products_batch(mh, m, qs)
    say("emvp dev m =", m, "nq =", nq)
    qd = svs.DeviceArray.from_host(ctx, qs)
    out = svs.DeviceArray(ctx, (nq, 76 * m), np.uint64)
    assert cl.compute_products_dev(nq, qd.ptr(), out.ptr())
    assert np.array_equal(out.to_host(), want)
    assert np.array_equal(cl.compute_products_batch(qs[:3]), want[:3])
    qd.close()
    out.close()
    cl.close()


def test_tiptoe_kat(ctx, golden):
    from svs.schemes import TiptoeDatabase
    t = golden["tiptoe_kat"]
    db = TiptoeDatabase(np.array(t["m"], np.uint64), ctx)
    x = np.array(t["x"], np.uint64)
    for col in range(2):
        assert db.server_answer(np.ascontiguousarray(x[:, col])).tolist() == \
            [row[col] for row in t["mx"]]


def test_tiptoe_u64_and_u32_match_oracle(ctx, golden):
    from svs.schemes import TiptoeDatabase
    t = golden["tiptoe"]
    rng = np.random.default_rng(t["seed"])
    mdb = rng.integers(0, 32, size=(t["rows"], t["cols"]), dtype=np.uint64)
    ct = rng.integers(0, 1 << 64, size=t["cols"], dtype=np.uint64)
    say("tiptoe golden u64")
    assert [str(int(x)) for x in TiptoeDatabase(mdb, ctx).server_answer(ct)] == t["ans64"]
    db32 = TiptoeDatabase(mdb.astype(np.uint32), ctx)
    assert [str(int(x)) for x in db32.server_answer((ct & np.uint64(0xFFFFFFFF)).astype(np.uint32))] \
        == t["ans32"]


@pytest.mark.parametrize("rows,cols,nq", [(16, 64, 1), (100, 192, 3), (257, 1024, 17),
                                          (64, 8192, 64)])
def test_tiptoe_mod32_kernels(ctx, rows, cols, nq):
    from svs.schemes import TiptoeDatabase
    from svs._lib import KERNEL_MFMA, KERNEL_VALU
    rng = np.random.default_rng(rows * 7 + nq)
    mdb = rng.integers(0, 32, size=(rows, cols), dtype=np.uint32)
    mdb[0] = 31
    ct = rng.integers(0, 1 << 32, size=(nq, cols), dtype=np.uint32)
    ct[0] = 0xFFFFFFFF
    db = TiptoeDatabase(mdb, ctx)
    assert db.stored() == (rows * cols, 1)  # 5-bit entries stored one byte each
    want = np.stack([O.tiptoe_answer_u32(mdb, rows, cols, c) for c in ct])
    for kernel in (KERNEL_VALU, KERNEL_MFMA, -1):
        say("tiptoe", rows, cols, nq, "kernel", kernel)
        assert np.array_equal(db.server_answer(ct, kernel), want)


def test_tiptoe_wide_entries_use_u32_store(ctx):
    from svs.schemes import TiptoeDatabase
    rng = np.random.default_rng(3)
    mdb = rng.integers(0, 1 << 32, size=(40, 100), dtype=np.uint32)
    ct = rng.integers(0, 1 << 32, size=(2, 100), dtype=np.uint32)
    db = TiptoeDatabase(mdb, ctx)
    assert db.stored() == (40 * 100 * 4, 4)
    want = np.stack([O.tiptoe_answer_u32(mdb, 40, 100, c) for c in ct])
    assert np.array_equal(db.server_answer(ct), want)


# ------------------------------------------------ Tiptoe BFV compression
def test_tiptoe_apply_hint_matches_oracle(ctx):
    """decompose_hint + apply_hint on the device (compress.rs:146-190,
    241-266) against the oracle, bit for bit in the NTT domain: a ragged
    second row chunk, random full-width hint entries and random tokens
    (1 and 5 per call: the 4-token pass and its remainder)."""
    from svs.schemes import TiptoeHint
    from svs import _lib as L
    rng = np.random.default_rng(61)
    r, n_lwe = 2048 + 37, 24
    hint = rng.integers(0, 2 ** 64, size=(r, n_lwe), dtype=np.uint64)
    hint[0, :] = 2 ** 64 - 1
    th = TiptoeHint(hint, r, n_lwe, ctx)
    q = O.bfv_modulus()
    assert th.q == q and th.chunks == 2 and th.stored_bytes == 8 * 2 * n_lwe * 2048 * 8
    pt = O.tiptoe_decompose_hint(hint, r, n_lwe)
    for nq in (1, 5):
        say("tiptoe apply_hint vs oracle, tokens =", nq)
        ct = rng.integers(0, q, size=(nq, n_lwe, 2, 2048), dtype=np.uint64)
        ct[0, 0, :, :] = q - 1
        got = th.apply_hint(ct)
        for i in range(nq):
            assert np.array_equal(got[i], O.tiptoe_apply_hint(pt, ct[i])), i
    bad = rng.integers(0, q, size=(n_lwe, 2, 2048), dtype=np.uint64)
    bad[3, 1, 7] = q
    with pytest.raises(L.GpuError):
        th.apply_hint(bad)
    th.close()


def test_tiptoe_apply_hint_recovers_h_times_s(ctx):
    """The reference's apply_hint_recovers_top_32_bits_of_h_times_s_small
    (compress.rs:445-476) with the device doing decompose_hint and
    apply_hint: r = 4, n_lwe = 2048, H[k] = ((17 k + 3) & 0xFFFF) << 32, a
    ternary LWE secret encrypted under BFV; decrypting the 8 answers
    recovers H s exactly."""
    from svs.schemes import TiptoeHint
    rng = np.random.default_rng(62)
    r, n_lwe = 4, 2048
    hint = ((np.arange(r * n_lwe, dtype=np.uint64) * np.uint64(17) + np.uint64(3))
            & np.uint64(0xFFFF)) << np.uint64(32)
    s = rng.integers(0, 3, n_lwe).astype(np.uint8)
    sk = O.bfv_keygen(bytes([21] * 32))
    say("tiptoe: encrypting the secret (oracle)")
    ct = O.tiptoe_encrypt_secret(s, sk, bytes([22] * 32))
    th = TiptoeHint(hint, r, n_lwe, ctx)
    got = O.tiptoe_recover_h_s(th.apply_hint(ct), r, sk)
    h = hint.reshape(r, n_lwe)
    want = [sum(int(h[i, j]) * int(s[j]) for j in range(n_lwe)) % 2 ** 64 for i in range(r)]
    assert got.tolist() == want
    th.close()


def test_tiptoe_apply_hint_full_size_32768x2048(ctx):
    """The SimplePIR hint of the 32768 x 8192 database (r = 32768,
    n_lwe = 2048: 16 chunks, 4 GiB of limb plaintexts): entries with random
    top halves and zero low halves, one BFV-encrypted ternary secret; every
    one of the 32768 recovered values equals H s mod 2^64 (the limb budget
    2048 x 2 x 15 < t and the noise margin hold at full size)."""
    from svs.schemes import TiptoeHint
    rng = np.random.default_rng(63)
    r, n_lwe = 32768, 2048
    hint = rng.integers(0, 2 ** 32, size=(r, n_lwe), dtype=np.uint64) << np.uint64(32)
    s = rng.integers(0, 3, n_lwe).astype(np.uint8)
    sk = O.bfv_keygen(bytes([31] * 32))
    ct = O.tiptoe_encrypt_secret(s, sk, bytes([32] * 32))
    say("tiptoe full size: decompose on the device")
    th = TiptoeHint(hint, r, n_lwe, ctx)
    assert th.chunks == 16 and th.stored_bytes == 8 * 16 * 2048 * 2048 * 8
    say("tiptoe full size: apply_hint")
    ans = th.apply_hint(ct)
    th.close()
    got = O.tiptoe_recover_h_s(ans, r, sk)
    want = np.zeros(r, np.uint64)
    s64 = s.astype(np.uint64)
    for a in range(0, r, 4096):
        want[a:a + 4096] = (hint[a:a + 4096] * s64[None, :]).sum(axis=1, dtype=np.uint64)
    assert np.array_equal(got, want)
