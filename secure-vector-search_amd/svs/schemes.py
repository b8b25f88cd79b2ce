"""EMVP's score GEMV and Tiptoe's SimplePIR answer on the gfx950 backend.

EmvpCluster mirrors the EMVP GPU state (scorer-emvp/src/gpu/mod.rs:217-380):
a resident M_hat and compute_products / compute_products_batch
(scorer-emvp/src/crypto.rs:200-222, 369-415) returning the s x m block-major
partials, plus the client's decode_scores (crypto.rs:435-453) on the device
and the two fused (score_batch). TiptoeDatabase mirrors server_answer
(scorer-tiptoe/src/pir/simplepir.rs:191-204) over a resident database, in the
mod-2^32 regime (u32 database) and the reference's mod-2^64 regime.
TiptoeHint holds the BFV-compressed hint (compress.rs) and runs apply_hint.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check, ptr, u64
from .bntm import Context, _live_indexes, default_context

EMVP_N, EMVP_S, EMVP_B = 1292, 76, 17


def _seed(r_seed) -> np.ndarray:
    a = np.frombuffer(bytes(r_seed), np.uint8).copy() if not isinstance(r_seed, np.ndarray) \
        else np.ascontiguousarray(r_seed, np.uint8)
    if a.size != 32:
        raise ValueError("r_seed must be 32 bytes")
    return a


class EmvpCluster:
    def __init__(self, m_hat, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        a = u64(m_hat).reshape(-1, EMVP_N)
        self.m = a.shape[0]
        h = C.c_void_p()
        check(L.lib().svs_emvp_create(self.ctx.h, self.m, ptr(a) if self.m else None, C.byref(h)))
        self.h = h
        _live_indexes.add(self)

    def compute_products(self, q_hat):
        """s x m partials (block-major); None for an empty cluster."""
        out = self.compute_products_batch(u64(q_hat).reshape(1, EMVP_N))
        return None if out is None else out[0]

    def compute_products_batch(self, q_hats):
        q = u64(q_hats).reshape(-1, EMVP_N)
        out = np.zeros((q.shape[0], EMVP_S * max(self.m, 1)), np.uint64)
        rc = check(L.lib().svs_emvp_compute_products(self.h, q.shape[0], ptr(q), ptr(out)))
        return None if rc == L.SVS_NONE else out[:, : EMVP_S * self.m]

    def compute_products_dev(self, n_queries: int, q_hat_ptr: int, partials_ptr: int):
        """Device pointers (canonical q_hat, n_queries x 1292 u64; partials
        n_queries x 76 x m u64), asynchronous on the context stream. False
        for an empty cluster."""
        rc = check(L.lib().svs_emvp_compute_products_dev(self.h, n_queries, C.c_void_p(q_hat_ptr),
                                                         C.c_void_p(partials_ptr)))
        return rc != L.SVS_NONE

    def decode_scores_batch(self, r_seed, q_hats, partials):
        """decode_scores (scorer-emvp/src/crypto.rs:435-453) on the device:
        n_queries x m f32 scores; None for an empty cluster."""
        q = u64(q_hats).reshape(-1, EMVP_N)
        pa = u64(partials).reshape(q.shape[0], -1)
        if pa.shape[1] != EMVP_S * self.m:
            raise ValueError(f"partials must be n_queries x {EMVP_S * self.m}")
        out = np.zeros((q.shape[0], max(self.m, 1)), np.float32)
        seed = _seed(r_seed)
        rc = check(L.lib().svs_emvp_decode_scores(self.h, q.shape[0], ptr(seed), ptr(q), ptr(pa),
                                                  ptr(out)))
        return None if rc == L.SVS_NONE else out[:, : self.m]

    def decode_scores_dev(self, n_queries: int, r_seed, q_hat_ptr: int, partials_ptr: int,
                          scores_ptr: int):
        """Device pointers, asynchronous on the context stream."""
        seed = _seed(r_seed)
        rc = check(L.lib().svs_emvp_decode_scores_dev(self.h, n_queries, ptr(seed),
                                                      C.c_void_p(q_hat_ptr),
                                                      C.c_void_p(partials_ptr),
                                                      C.c_void_p(scores_ptr)))
        return rc != L.SVS_NONE

    def score_batch(self, r_seed, q_hats):
        """compute_products + decode_scores on the device (the partials stay
        there): n_queries x m f32 scores; None for an empty cluster."""
        q = u64(q_hats).reshape(-1, EMVP_N)
        out = np.zeros((q.shape[0], max(self.m, 1)), np.float32)
        seed = _seed(r_seed)
        rc = check(L.lib().svs_emvp_score(self.h, q.shape[0], ptr(seed), ptr(q), ptr(out)))
        return None if rc == L.SVS_NONE else out[:, : self.m]

    def close(self):
        if getattr(self, "h", None):
            L.lib().svs_emvp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TiptoeDatabase:
    def __init__(self, m_db, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        a = np.ascontiguousarray(m_db)
        if a.dtype not in (np.uint32, np.uint64):
            raise TypeError("database entries must be uint32 or uint64")
        self.rows, self.cols = a.shape
        self.elem = a.dtype.itemsize
        h = C.c_void_p()
        check(L.lib().svs_tiptoe_create(self.ctx.h, self.rows, self.cols, self.elem, ptr(a),
                                        C.byref(h)))
        self.h = h
        _live_indexes.add(self)

    def stored(self):
        b, e = C.c_uint64(0), C.c_int(0)
        check(L.lib().svs_tiptoe_stored_bytes(self.h, C.byref(b), C.byref(e)))
        return b.value, e.value

    def server_answer(self, ct, kernel: int = -1):
        """ans = M_db . ct, wrapping at 2^32 (u32 database) or 2^64 (u64)."""
        if self.elem == 8:
            c = u64(ct).reshape(-1, self.cols)
            ans = np.zeros((c.shape[0], self.rows), np.uint64)
            check(L.lib().svs_tiptoe_answer_u64(self.h, c.shape[0], ptr(c), ptr(ans)))
        else:
            c = np.ascontiguousarray(ct, np.uint32).reshape(-1, self.cols)
            ans = np.zeros((c.shape[0], self.rows), np.uint32)
            check(L.lib().svs_tiptoe_answer_u32(self.h, c.shape[0], ptr(c), ptr(ans), kernel))
        return ans[0] if np.ndim(ct) == 1 else ans

    def close(self):
        if getattr(self, "h", None):
            L.lib().svs_tiptoe_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TiptoeHint:
    """Tiptoe's compressed hint (scorer-tiptoe/src/bfv/compress.rs
    CompressedHint) resident on the device: decompose_hint (:146-190) at
    construction, apply_hint (:241-266) per batch of query tokens. Tokens
    and answers are in the negacyclic NTT domain mod q (see svs_gpu.h)."""

    N, LIMBS = 2048, 8

    def __init__(self, hint, r: int, n_lwe: int, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        a = u64(hint).reshape(-1)
        if a.size != r * n_lwe:
            raise ValueError(f"hint length {a.size} != r * n_lwe = {r * n_lwe}")
        self.r, self.n_lwe = r, n_lwe
        h = C.c_void_p()
        check(L.lib().svs_tiptoe_hint_create(self.ctx.h, r, n_lwe, ptr(a), C.byref(h)))
        self.h = h
        q, ch, nb = C.c_uint64(0), C.c_uint32(0), C.c_uint64(0)
        check(L.lib().svs_tiptoe_hint_info(self.h, C.byref(q), C.byref(ch), C.byref(nb)))
        self.q, self.chunks, self.stored_bytes = q.value, ch.value, nb.value
        _live_indexes.add(self)

    def apply_hint(self, ct):
        """ct: (n_lwe, 2, N) for one token or (nq, n_lwe, 2, N); returns
        (8, chunks, 2, N) per token."""
        c = u64(ct)
        one = c.ndim == 3
        c = c.reshape(-1, self.n_lwe, 2, self.N)
        out = np.zeros((c.shape[0], self.LIMBS, self.chunks, 2, self.N), np.uint64)
        check(L.lib().svs_tiptoe_apply_hint(self.h, c.shape[0], ptr(c), ptr(out)))
        return out[0] if one else out

    def apply_hint_dev(self, n_queries: int, ct_ptr: int, out_ptr: int):
        check(L.lib().svs_tiptoe_apply_hint_dev(self.h, n_queries, C.c_void_p(ct_ptr),
                                                 C.c_void_p(out_ptr)))

    def close(self):
        if getattr(self, "h", None):
            L.lib().svs_tiptoe_hint_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
