"""ctypes bindings for the CPU oracle (oracle/svs_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline. The
product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

P = (1 << 61) - 1
N = 1024
N1 = 128
LAMBDA = 3
EMVP_N, EMVP_S, EMVP_B = 1292, 76, 17

u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile the oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(variant: str = "") -> C.CDLL:
    global _LIB
    name = "liboracle%s.so" % variant
    path = os.path.join(_HERE, "_build", name)
    if variant:
        return _bind(C.CDLL(path))
    if _LIB is None:
        if not os.path.exists(path):
            build()
        _LIB = _bind(C.CDLL(path))
    return _LIB


def _bind(L: C.CDLL) -> C.CDLL:
    sz = C.c_size_t
    L.or_mersenne_reduce.restype = C.c_uint64
    L.or_mersenne_reduce.argtypes = [C.c_uint64, C.c_uint64]
    for f in ("or_add_mod", "or_sub_mod", "or_mul_mod"):
        getattr(L, f).restype = C.c_uint64
        getattr(L, f).argtypes = [C.c_uint64, C.c_uint64]
    L.or_chacha20_block.argtypes = [u32p, C.c_uint64, u32p]
    L.or_domain_separate.argtypes = [u8p, C.c_char_p, sz, u8p]
    L.or_sparse_threshold.restype = C.c_uint32
    L.or_quantise_vector.argtypes = [f32p, sz, sz, C.c_uint64, u64p]
    L.or_quantise_rows.argtypes = [f32p, sz, sz, sz, C.c_uint64, u64p]
    L.or_fp_to_signed.restype = C.c_int64
    L.or_fp_to_signed.argtypes = [C.c_uint64]
    L.or_generate_subspace.argtypes = [u8p, sz, sz, u64p]
    L.or_regenerate_h_s.restype = sz
    L.or_regenerate_h_s.argtypes = [u8p, sz, sz, sz, u64p, u32p, u32p, i8p, sz]
    L.or_encrypt_cluster.argtypes = [u64p, sz, sz, sz, u64p, u64p, u32p, u32p, i8p, u64p, u64p]
    L.or_encode_query.restype = sz
    L.or_encode_query.argtypes = [u64p, u64p, sz, sz, u8p, u64p, u64p, u32p, i8p]
    L.or_matvec.argtypes = [u64p, sz, sz, u64p, u64p]
    L.or_matvec_mt.argtypes = [u64p, sz, sz, u64p, u64p, C.c_int]
    L.or_matvec_transpose.argtypes = [u64p, sz, sz, u64p, u64p]
    L.or_matvec_batch.argtypes = [u64p, sz, sz, u64p, sz, u64p]
    L.or_dot_product.restype = C.c_uint64
    L.or_dot_product.argtypes = [u64p, u64p, sz]
    L.or_verify_response.restype = C.c_int
    L.or_verify_response.argtypes = [u64p, u64p, sz, sz, u64p, C.c_int, u8p]
    L.or_decode_scores.argtypes = [u64p, sz, sz, sz, u64p, u64p, u64p, u32p, u32p, i8p,
                                   u64p, u64p, u64p, u32p, i8p, sz, u64p]
    L.or_topk.restype = sz
    L.or_topk.argtypes = [u32p, i64p, sz, sz, u32p, i64p]
    L.or_l2_f32.restype = C.c_float
    L.or_l2_f32.argtypes = [f32p, f32p, sz]
    L.or_probe_route.argtypes = [f32p, sz, sz, f32p, sz, u32p]
    L.or_emvp_fp_reduce.restype = C.c_uint64
    L.or_emvp_fp_reduce.argtypes = [C.c_uint64, C.c_uint64]
    L.or_emvp_compute_products.argtypes = [u64p, sz, u64p, u64p]
    L.or_emvp_compute_products_batch.argtypes = [u64p, sz, u64p, sz, u64p]
    L.or_emvp_quantize.restype = C.c_uint64
    L.or_emvp_quantize.argtypes = [C.c_float]
    L.or_emvp_decode_field.restype = C.c_float
    L.or_emvp_decode_field.argtypes = [C.c_uint64]
    L.or_emvp_generate_d.argtypes = [u8p, u64p]
    L.or_emvp_encrypt_cluster.argtypes = [u64p, u8p, f32p, sz, sz, u64p]
    L.or_emvp_encode_query.argtypes = [u64p, f64p, u64p, u64p]
    L.or_emvp_decode_scores.argtypes = [sz, u8p, u64p, u64p, f32p]
    L.or_tiptoe_server_answer_u64.argtypes = [u64p, sz, sz, u64p, u64p]
    L.or_tiptoe_server_answer_u32.argtypes = [u32p, sz, sz, u32p, u32p]
    L.or_bntm_flat_score.restype = C.c_int
    L.or_bntm_flat_score.argtypes = [u64p, u64p, u64p, u64p, u32p, u32p, i8p, u64p, sz, u64p,
                                     u8p, C.c_int, sz, u32p, i64p, C.POINTER(C.c_size_t)]
    L.or_bfv_modulus.restype = C.c_uint64
    L.or_bfv_psi.restype = C.c_uint64
    L.or_bfv_ntt.argtypes = [u64p]
    L.or_bfv_intt.argtypes = [u64p]
    L.or_bfv_negacyclic_naive.argtypes = [u64p, u64p, u64p]
    L.or_bfv_keygen.argtypes = [u8p, u64p]
    L.or_bfv_encrypt.argtypes = [u64p, u64p, u8p, u64p, u64p]
    L.or_bfv_decrypt.argtypes = [u64p, u64p, u64p, u64p]
    L.or_tiptoe_get_chunk.restype = C.c_uint64
    L.or_tiptoe_get_chunk.argtypes = [C.c_uint64, sz]
    L.or_tiptoe_from_modulo_t.restype = C.c_int64
    L.or_tiptoe_from_modulo_t.argtypes = [C.c_uint64, C.c_uint64]
    L.or_tiptoe_decompose_hint.argtypes = [u64p, sz, sz, u64p]
    L.or_tiptoe_encrypt_secret.restype = C.c_int
    L.or_tiptoe_encrypt_secret.argtypes = [u8p, sz, u64p, u8p, u64p]
    L.or_tiptoe_apply_hint.argtypes = [u64p, sz, sz, u64p, u64p]
    L.or_tiptoe_recover_h_s.argtypes = [u64p, sz, u64p, u64p]
    L.or_has_avx512f.restype = C.c_int
    L.or_matvec_avx512.restype = C.c_int
    L.or_matvec_avx512.argtypes = [u64p, sz, sz, u64p, u64p]
    L.or_ivf_rows_mt.restype = C.c_int
    L.or_ivf_rows_mt.argtypes = [C.c_void_p, sz, sz, sz, u64p, u64p, u64p, u32p, i8p, sz,
                                 C.c_int, C.c_int, i64p]
    return L


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


# ---------------------------------------------------------------- field ops
def mersenne_reduce(x: int) -> int:
    return lib().or_mersenne_reduce(x & ((1 << 64) - 1), x >> 64)


def add_mod(a, b):
    return lib().or_add_mod(a, b)


def sub_mod(a, b):
    return lib().or_sub_mod(a, b)


def mul_mod(a, b):
    return lib().or_mul_mod(a, b)


def chacha20_block(key_words, counter: int) -> np.ndarray:
    out = np.zeros(16, np.uint32)
    lib().or_chacha20_block(np.ascontiguousarray(key_words, np.uint32), counter, out)
    return out


def seed32(b) -> np.ndarray:
    a = np.frombuffer(bytes(b), np.uint8).copy() if not isinstance(b, np.ndarray) else b
    assert a.size == 32
    return np.ascontiguousarray(a, np.uint8)


def domain_separate(seed, tag: bytes) -> np.ndarray:
    out = np.zeros(32, np.uint8)
    lib().or_domain_separate(seed32(seed), tag, len(tag), out)
    return out


def sparse_threshold() -> int:
    return lib().or_sparse_threshold()


def quantise_vector(v, n=N, q=1 << 20) -> np.ndarray:
    v = np.ascontiguousarray(v, np.float32)
    out = np.zeros(n, np.uint64)
    lib().or_quantise_vector(v, v.size, n, q, out)
    return out


def quantise_rows(vectors, n=N, q=1 << 20) -> np.ndarray:
    """quantise_vector of every row of a (rows, d) float32 matrix."""
    v = np.ascontiguousarray(vectors, np.float32)
    rows, d = v.shape
    out = np.zeros((max(rows, 1), n), np.uint64)
    if rows:
        lib().or_quantise_rows(v.reshape(-1), rows, d, n, q, out.reshape(-1))
    return out[:rows]


def fp_to_signed(s: int) -> int:
    return lib().or_fp_to_signed(s)


def generate_subspace(key_seed, n=N, n1=N1) -> np.ndarray:
    out = np.zeros(n * n1, np.uint64)
    lib().or_generate_subspace(seed32(key_seed), n, n1, out)
    return out


def regenerate_h_s(h_seed, m, n=N, n1=N1):
    h = np.zeros(max(m * n1, 1), np.uint64)
    rowptr = np.zeros(m + 1, np.uint32)
    cap = m * n + 1
    idx = np.zeros(cap, np.uint32)
    sign = np.zeros(cap, np.int8)
    nnz = lib().or_regenerate_h_s(seed32(h_seed), m, n, n1, h, rowptr, idx, sign, cap)
    return h[: m * n1], rowptr, idx[:nnz].copy(), sign[:nnz].copy()


def encrypt_cluster(m_plain, m, l, h, rowptr, idx, sign, n=N, n1=N1):
    m_enc = np.zeros(max(m * n, 1), np.uint64)
    a_l = np.zeros(max(m * n1, 1), np.uint64)
    lib().or_encrypt_cluster(_u64(m_plain).reshape(-1) if m else np.zeros(1, np.uint64), m, n,
                             n1, _u64(l), _u64(h) if m else np.zeros(1, np.uint64), rowptr,
                             idx if idx.size else np.zeros(1, np.uint32),
                             sign if sign.size else np.zeros(1, np.int8), m_enc, a_l)
    return m_enc[: m * n], a_l[: m * n1]


def encode_query(v, l, q_seed, n=N, n1=N1):
    v_enc = np.zeros(n, np.uint64)
    g = np.zeros(n1, np.uint64)
    t_idx = np.zeros(n, np.uint32)
    t_sign = np.zeros(n, np.int8)
    nnz = lib().or_encode_query(_u64(v), _u64(l), n, n1, seed32(q_seed), v_enc, g, t_idx, t_sign)
    return v_enc, g, t_idx[:nnz].copy(), t_sign[:nnz].copy()


def matvec(mat, rows, cols, vec) -> np.ndarray:
    out = np.zeros(max(rows, 1), np.uint64)
    if rows:
        lib().or_matvec(_u64(mat).reshape(-1), rows, cols, _u64(vec), out)
    return out[:rows]


def matvec_mt(mat, rows, cols, vec, nthreads, variant="") -> np.ndarray:
    out = np.zeros(max(rows, 1), np.uint64)
    if rows:
        lib(variant).or_matvec_mt(_u64(mat).reshape(-1), rows, cols, _u64(vec), out, nthreads)
    return out[:rows]


def matvec_transpose(mat, rows, cols, vec) -> np.ndarray:
    out = np.zeros(cols, np.uint64)
    lib().or_matvec_transpose(_u64(mat).reshape(-1), rows, cols, _u64(vec), out)
    return out


def matvec_batch(mat, rows, cols, vecs) -> np.ndarray:
    vecs = _u64(vecs).reshape(-1, cols)
    out = np.zeros((vecs.shape[0], max(rows, 1)), np.uint64)
    lib().or_matvec_batch(_u64(mat).reshape(-1), rows, cols, vecs.reshape(-1), vecs.shape[0],
                          out.reshape(-1))
    return out[:, :rows]


def dot_product(a, b) -> int:
    a, b = _u64(a), _u64(b)
    return lib().or_dot_product(a, b, a.size)


def verify_response(r, m_enc, m, v_enc, rng_seed, trials=LAMBDA, n=N) -> int:
    """Returns -1 (accept) or the first failing trial."""
    return lib().or_verify_response(_u64(r), _u64(m_enc).reshape(-1), m, n, _u64(v_enc),
                                    trials, seed32(rng_seed))


def decode_scores(r, m, a_l, m_plain, h, rowptr, idx, sign, l, v_enc, g, t_idx, t_sign,
                  n=N, n1=N1) -> np.ndarray:
    out = np.zeros(max(m, 1), np.uint64)
    z64 = np.zeros(1, np.uint64)
    lib().or_decode_scores(_u64(r) if m else z64, m, n, n1, _u64(a_l) if m else z64,
                           _u64(m_plain).reshape(-1) if m else z64, _u64(h) if m else z64,
                           rowptr, idx if idx.size else np.zeros(1, np.uint32),
                           sign if sign.size else np.zeros(1, np.int8), _u64(l), _u64(v_enc),
                           _u64(g), t_idx if t_idx.size else np.zeros(1, np.uint32),
                           t_sign if t_sign.size else np.zeros(1, np.int8), t_idx.size, out)
    return out[:m]


def topk(ids, scores, k):
    ids = np.ascontiguousarray(ids, np.uint32)
    scores = np.ascontiguousarray(scores, np.int64)
    oi = np.zeros(max(k, 1), np.uint32)
    os_ = np.zeros(max(k, 1), np.int64)
    c = lib().or_topk(ids, scores, ids.size, k, oi, os_)
    return oi[:c], os_[:c]


def fp_to_signed_vec(mv) -> np.ndarray:
    mv = _u64(mv)
    s = mv.astype(np.int64)  # canonical values are < 2^61
    s[mv >= np.uint64(P // 2)] -= np.int64(P)
    return s


def l2_f32(a, b) -> float:
    a = np.ascontiguousarray(a, np.float32)
    return lib().or_l2_f32(a, np.ascontiguousarray(b, np.float32), a.size)


def probe_route(centroids, query, nprobe) -> np.ndarray:
    centroids = np.ascontiguousarray(centroids, np.float32)
    nc, d = centroids.shape
    nprobe = min(nprobe, nc)
    out = np.zeros(max(nprobe, 1), np.uint32)
    lib().or_probe_route(centroids.reshape(-1), nc, d, np.ascontiguousarray(query, np.float32),
                         nprobe, out)
    return out[:nprobe]


def emvp_fp_reduce(x: int) -> int:
    return lib().or_emvp_fp_reduce(x & ((1 << 64) - 1), x >> 64)


def emvp_compute_products(m_hat, m, q_hat) -> np.ndarray:
    out = np.zeros(EMVP_S * max(m, 1), np.uint64)
    lib().or_emvp_compute_products(_u64(m_hat).reshape(-1), m, _u64(q_hat), out)
    return out[: EMVP_S * m]


def emvp_compute_products_batch(m_hat, m, q_hats) -> np.ndarray:
    q_hats = _u64(q_hats).reshape(-1, EMVP_N)
    out = np.zeros(q_hats.shape[0] * EMVP_S * max(m, 1), np.uint64)
    lib().or_emvp_compute_products_batch(_u64(m_hat).reshape(-1), m, q_hats.reshape(-1),
                                         q_hats.shape[0], out)
    return out.reshape(q_hats.shape[0], EMVP_S * max(m, 1))[:, : EMVP_S * m]


EMVP_N, EMVP_S, EMVP_B, EMVP_ELL0, EMVP_K = 1292, 76, 17, 1024, 268


def emvp_generate_d(key_seed) -> np.ndarray:
    """generate_d (scorer-emvp/src/crypto.rs:72-97), ell0 x n."""
    d = np.zeros(EMVP_ELL0 * EMVP_N, np.uint64)
    lib().or_emvp_generate_d(seed32(key_seed), d)
    return d.reshape(EMVP_ELL0, EMVP_N)


def emvp_encrypt_cluster(d_mat, r_seed, vectors) -> np.ndarray:
    """encrypt_cluster (crypto.rs:124-163): m x n M_hat."""
    v = np.ascontiguousarray(vectors, np.float32)
    m, d = v.shape
    out = np.zeros(max(m, 1) * EMVP_N, np.uint64)
    if m:
        lib().or_emvp_encrypt_cluster(_u64(d_mat).reshape(-1), seed32(r_seed), v.reshape(-1), m,
                                      d, out)
    return out[: m * EMVP_N].reshape(m, EMVP_N)


def emvp_encode_query(d_mat, query, u) -> np.ndarray:
    """encode_query (crypto.rs:171-190): q_hat = (q - D'u | u)."""
    out = np.zeros(EMVP_N, np.uint64)
    lib().or_emvp_encode_query(_u64(d_mat).reshape(-1), np.ascontiguousarray(query, np.float64),
                               _u64(u), out)
    return out


def emvp_decode_scores(m, r_seed, q_hat, partials) -> np.ndarray:
    """decode_scores (crypto.rs:435-453): f32 score per row."""
    out = np.zeros(max(m, 1), np.float32)
    if m:
        lib().or_emvp_decode_scores(m, seed32(r_seed), _u64(q_hat), _u64(partials).reshape(-1),
                                    out)
    return out[:m]


def emvp_quantize(v) -> int:
    """quantize (scorer-emvp/src/crypto.rs:47-52)."""
    return lib().or_emvp_quantize(float(v))


def emvp_decode_field(v: int) -> float:
    return lib().or_emvp_decode_field(v)


def tiptoe_answer_u64(m_db, r, c, ct) -> np.ndarray:
    out = np.zeros(r, np.uint64)
    lib().or_tiptoe_server_answer_u64(_u64(m_db).reshape(-1), r, c, _u64(ct), out)
    return out


def tiptoe_answer_u32(m_db, r, c, ct) -> np.ndarray:
    out = np.zeros(r, np.uint32)
    lib().or_tiptoe_server_answer_u32(np.ascontiguousarray(m_db, np.uint32).reshape(-1), r, c,
                                      np.ascontiguousarray(ct, np.uint32), out)
    return out


def bntm_flat_score(cl, l, v_q, q_seed, verify, k):
    """BnTmScorer::score on one cluster dict (keys: m, m_enc, a_l, m_plain, h,
    s_rowptr, s_idx, s_sign). Returns (status, ids, scores)."""
    m = cl["m"]
    oi = np.zeros(max(k, 1), np.uint32)
    os_ = np.zeros(max(k, 1), np.int64)
    cnt = C.c_size_t(0)
    st = lib().or_bntm_flat_score(_u64(cl["m_enc"]), _u64(cl["a_l"]), _u64(cl["m_plain"]),
                                  _u64(cl["h"]), cl["s_rowptr"],
                                  cl["s_idx"] if cl["s_idx"].size else np.zeros(1, np.uint32),
                                  cl["s_sign"] if cl["s_sign"].size else np.zeros(1, np.int8),
                                  _u64(l), m, _u64(v_q), seed32(q_seed), int(verify), k, oi,
                                  os_, C.byref(cnt))
    return st, oi[: cnt.value], os_[: cnt.value]


def build_cluster(m_plain_u64, m, l, h_seed):
    """regenerate_h_s + encrypt_cluster -> dict with the client+server state."""
    h, rowptr, idx, sign = regenerate_h_s(h_seed, m)
    m_enc, a_l = encrypt_cluster(m_plain_u64, m, l, h, rowptr, idx, sign)
    return dict(m=m, m_enc=m_enc, a_l=a_l, m_plain=_u64(m_plain_u64).reshape(-1), h=h,
                s_rowptr=rowptr, s_idx=idx, s_sign=sign)


def derive_h_seed(upload_seed: int, cluster: int) -> np.ndarray:
    """derive_h_seed (scorer-bntm/src/ivf.rs:258-264) from the oracle's
    ChaCha20 block: first 32 stream words, one byte each."""
    seed = int(upload_seed).to_bytes(8, "little") + int(cluster).to_bytes(8, "little") + bytes(16)
    key = np.frombuffer(seed, "<u4").astype(np.uint32)
    words = np.concatenate([chacha20_block(key, 0), chacha20_block(key, 1)])
    return (words & 0xFF).astype(np.uint8)


def u_draws(rng_seed, count: int) -> list:
    """The first `count` sample_fp draws of ChaCha20Rng(rng_seed) (the
    Freivalds u vectors, crypto.rs:1028), via generate_subspace's stream."""
    n1 = 1
    out = np.zeros(count, np.uint64)
    lib().or_generate_subspace(seed32(rng_seed), count, n1, out)
    return [int(x) for x in out]


def ivf_score(clusters, ids_per_list, centroids, l, query_f32, nprobe, q_seed, verify, k):
    """BnTmIvfScorer::score (scorer-bntm/src/ivf.rs:746-822) on the oracle:
    route, encode, then per probed non-empty list compute / verify / decode,
    concatenate (global id, signed score) in probe order, top-k.
    Returns (hits_ids, hits_scores) or raises VerificationFailedOracle."""
    probes = probe_route(centroids, query_f32, nprobe)
    v_q = quantise_vector(query_f32)
    v_enc, g, t_idx, t_sign = encode_query(v_q, l, q_seed)
    all_ids, all_scores = [], []
    for ci in probes.tolist():
        cl = clusters[ci]
        m = cl["m"]
        if m == 0:
            continue
        r = matvec(cl["m_enc"], m, N, v_enc)
        if verify:
            t = verify_response(r, cl["m_enc"], m, v_enc, domain_separate(q_seed, b"protocol2"))
            if t >= 0:
                raise VerificationFailedOracle(ci, t)
        mv = decode_scores(r, m, cl["a_l"], cl["m_plain"], cl["h"], cl["s_rowptr"], cl["s_idx"],
                           cl["s_sign"], l, v_enc, g, t_idx, t_sign)
        all_ids.extend(ids_per_list[ci].tolist())
        all_scores.extend(fp_to_signed_vec(mv).tolist())
    return topk(np.array(all_ids, np.uint32), np.array(all_scores, np.int64), k)


class VerificationFailedOracle(Exception):
    def __init__(self, cluster, trial):
        super().__init__(f"cluster {cluster} trial {trial}")
        self.cluster, self.trial = cluster, trial


# ------------------------------------------------------- timed CPU baseline
def has_avx512f() -> bool:
    return bool(lib().or_has_avx512f())


def matvec_avx512(mat, rows, cols, vec) -> np.ndarray:
    """mersenne_matvec_avx512 (crypto.rs:643-739); None without AVX-512F."""
    out = np.zeros(max(rows, 1), np.uint64)
    if rows and lib().or_matvec_avx512(_u64(mat).reshape(-1), rows, cols, _u64(vec), out) != 0:
        return None
    return out[:rows]


class _ClusterView(C.Structure):
    _fields_ = [("m_enc", C.c_void_p), ("a_l", C.c_void_p), ("m_plain", C.c_void_p),
                ("h", C.c_void_p), ("s_rowptr", C.c_void_p), ("s_idx", C.c_void_p),
                ("s_sign", C.c_void_p), ("m", C.c_uint64)]


class CpuIvf:
    """BnTmIvfScorer::score (ivf.rs:746-822) as the reference runs it in its
    closed-loop B=1 protocol (eval.rs:2338-2344): one query at a time; route,
    quantise and encode on one thread, then compute_products and
    decode_scores over the probed lists with the rows split across
    `nthreads` threads (rayon's row-parallel matvec, crypto.rs:609-625).
    Verification off. `clusters` maps list id -> cluster dict (the keys of
    build_cluster); lists are viewed, not copied."""

    def __init__(self, clusters, ids_per_list, centroids, l):
        self.clusters, self.ids, self.l = clusters, ids_per_list, _u64(l)
        self.centroids = np.ascontiguousarray(centroids, np.float32)
        self._views = {}

    def _view(self, ci):
        v = self._views.get(ci)
        if v is None:
            cl = self.clusters[ci]
            keep = [_u64(cl["m_enc"]), _u64(cl["a_l"]), _u64(cl["m_plain"]), _u64(cl["h"]),
                    np.ascontiguousarray(cl["s_rowptr"], np.uint32),
                    np.ascontiguousarray(cl["s_idx"], np.uint32) if cl["s_idx"].size
                    else np.zeros(1, np.uint32),
                    np.ascontiguousarray(cl["s_sign"], np.int8) if cl["s_sign"].size
                    else np.zeros(1, np.int8)]
            v = (_ClusterView(*[a.ctypes.data for a in keep], cl["m"]), keep)
            self._views[ci] = v
        return v[0]

    def score(self, query, nprobe, q_seed, k, nthreads, simd=False):
        probes = probe_route(self.centroids, query, nprobe)
        v_q = quantise_vector(query)
        v_enc, g, t_idx, t_sign = encode_query(v_q, self.l, q_seed)
        lt = matvec_transpose(self.l, N, N1, v_enc)
        live = [ci for ci in probes.tolist() if self.clusters[ci]["m"] > 0]
        arr = (_ClusterView * max(len(live), 1))(*[self._view(ci) for ci in live])
        total = sum(int(self.clusters[ci]["m"]) for ci in live)
        scores = np.zeros(max(total, 1), np.int64)
        rc = lib().or_ivf_rows_mt(C.addressof(arr), len(live), N, N1, v_enc, g, lt,
                                  t_idx if t_idx.size else np.zeros(1, np.uint32),
                                  t_sign if t_sign.size else np.zeros(1, np.int8), t_idx.size,
                                  int(nthreads), int(bool(simd)), scores)
        if rc != 0:
            raise RuntimeError("AVX-512F is not available on this CPU")
        ids = np.concatenate([self.ids[ci] for ci in live]) if live else np.zeros(0, np.uint32)
        return topk(ids.astype(np.uint32), scores[:total], k)


# ------------------------------------------------- Tiptoe BFV compression
BFV_N, BFV_T, TIPTOE_LIMBS = 2048, 65537, 8


def bfv_modulus() -> int:
    return int(lib().or_bfv_modulus())


def bfv_ntt(a) -> np.ndarray:
    x = _u64(a).copy()
    lib().or_bfv_ntt(x)
    return x


def bfv_intt(a) -> np.ndarray:
    x = _u64(a).copy()
    lib().or_bfv_intt(x)
    return x


def bfv_negacyclic_naive(a, b) -> np.ndarray:
    out = np.zeros(BFV_N, np.uint64)
    lib().or_bfv_negacyclic_naive(_u64(a), _u64(b), out)
    return out


def bfv_keygen(seed) -> np.ndarray:
    s = np.zeros(BFV_N, np.uint64)
    lib().or_bfv_keygen(seed32(seed), s)
    return s


def bfv_encrypt(sk_ntt, m, seed):
    c0, c1 = np.zeros(BFV_N, np.uint64), np.zeros(BFV_N, np.uint64)
    lib().or_bfv_encrypt(_u64(sk_ntt), _u64(m), seed32(seed), c0, c1)
    return c0, c1


def bfv_decrypt(sk_ntt, c0, c1) -> np.ndarray:
    m = np.zeros(BFV_N, np.uint64)
    lib().or_bfv_decrypt(_u64(sk_ntt), _u64(c0), _u64(c1), m)
    return m


def tiptoe_get_chunk(v, idx) -> int:
    return int(lib().or_tiptoe_get_chunk(int(v), idx))


def tiptoe_from_modulo_t(t, v) -> int:
    return int(lib().or_tiptoe_from_modulo_t(t, v))


def tiptoe_decompose_hint(hint, r, n_lwe) -> np.ndarray:
    """decompose_hint (compress.rs:146-190): (8, chunks, n_lwe, N) plaintexts
    in the NTT domain."""
    chunks = (r + BFV_N - 1) // BFV_N
    pt = np.zeros(TIPTOE_LIMBS * chunks * n_lwe * BFV_N, np.uint64)
    lib().or_tiptoe_decompose_hint(_u64(hint).reshape(-1), r, n_lwe, pt)
    return pt.reshape(TIPTOE_LIMBS, chunks, n_lwe, BFV_N)


def tiptoe_encrypt_secret(s, sk_ntt, seed) -> np.ndarray:
    """encrypt_secret_under_bfv (compress.rs:201-227): (n_lwe, 2, N)."""
    s = np.ascontiguousarray(s, np.uint8)
    ct = np.zeros(s.size * 2 * BFV_N, np.uint64)
    rc = lib().or_tiptoe_encrypt_secret(s, s.size, _u64(sk_ntt), seed32(seed), ct)
    if rc != 0:
        raise ValueError(f"LWE secret value out of {{0,1,2}} at index {-1 - rc}")
    return ct.reshape(s.size, 2, BFV_N)


def tiptoe_apply_hint(pt, ct) -> np.ndarray:
    """apply_hint (compress.rs:241-266): (8, chunks, 2, N)."""
    pt = _u64(pt)
    limbs, chunks, n_lwe, _ = pt.shape
    out = np.zeros(limbs * chunks * 2 * BFV_N, np.uint64)
    lib().or_tiptoe_apply_hint(pt.reshape(-1), chunks, n_lwe, _u64(ct).reshape(-1), out)
    return out.reshape(limbs, chunks, 2, BFV_N)


def tiptoe_recover_h_s(answers, r, sk_ntt) -> np.ndarray:
    """recover_h_s (compress.rs:277-318): H s mod 2^64, r values."""
    out = np.zeros(max(r, 1), np.uint64)
    lib().or_tiptoe_recover_h_s(_u64(answers).reshape(-1), r, _u64(sk_ntt), out)
    return out[:r]
