#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from first principles.

Nothing here imports or runs the reference (it is Rust; there is no toolchain
in this image). Expected values come from:
  * published test vectors (RFC 7539 A.1 ChaCha20 keystream blocks);
  * known-answer constants that the reference's own unit tests assert
    (scorer-bntm/src/crypto.rs:1079-1101, scorer-bntm/src/params.rs:209-285,
    scorer-emvp/src/crypto.rs:463-498, scorer-tiptoe/src/pir/simplepir.rs:
    293-310);
  * exact Python big-int arithmetic on small, seeded inputs (closed forms:
    sums of products mod p, wrapping sums mod 2^64 / 2^32);
  * an independent pure-Python restatement of the rand_chacha stream
    conventions (so the C oracle's ChaCha consumer is cross-checked by a
    second implementation; the conventions themselves stay unpinned).

Inputs that are too large to store are regenerated in the tests from
numpy's PCG64 with the recorded seed; each fixture stores a SHA-256 of its
inputs so a drift in the generator fails loudly.

Run:  python3 tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import math
import os
import struct
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
P = (1 << 61) - 1
M64 = (1 << 64) - 1

# --------------------------------------------------------------------------
# RFC 7539, Appendix A.1 (ChaCha20 block function). All have a zero nonce,
# which is the only stream rand_chacha's ChaCha20Rng uses by default.
RFC7539_A1 = [
    dict(key="00" * 32, counter=0,
         keystream="76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
                   "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586"),
    dict(key="00" * 32, counter=1,
         keystream="9f07e7be5551387a98ba977c732d080dcb0f29a048e3656912c6533e32ee7aed"
                   "29b721769ce64e43d57133b074d839d531ed1f28510afb45ace10a1f4b794d6f"),
    dict(key="00" * 31 + "01", counter=1,
         keystream="3aeb5224ecf849929b9d828db1ced4dd832025e8018b8160b82284f3c949aa5a"
                   "8eca00bbb4a73bdad192b5c42f73f2fd4e273644c8b36125a64addeb006c13a0"),
    dict(key="00ff" + "00" * 30, counter=2,
         keystream="72d54dfbf12ec44b362692df94137f328fea8da73990265ec1bbbea1ae9af0ca"
                   "13b25aa26cb4a648cb9b9d1be65b2c0924a66c54d545ec1b7374f4872e99f096"),
]


def _rotl(v, c):
    return ((v << c) & 0xFFFFFFFF) | (v >> (32 - c))


def chacha_block(key: bytes, counter: int) -> list[int]:
    st = [0x61707865, 0x3320646E, 0x79622D32, 0x6B206574]
    st += list(struct.unpack("<8I", key))
    st += [counter & 0xFFFFFFFF, counter >> 32, 0, 0]
    x = st[:]

    def qr(a, b, c, d):
        x[a] = (x[a] + x[b]) & 0xFFFFFFFF; x[d] = _rotl(x[d] ^ x[a], 16)
        x[c] = (x[c] + x[d]) & 0xFFFFFFFF; x[b] = _rotl(x[b] ^ x[c], 12)
        x[a] = (x[a] + x[b]) & 0xFFFFFFFF; x[d] = _rotl(x[d] ^ x[a], 8)
        x[c] = (x[c] + x[d]) & 0xFFFFFFFF; x[b] = _rotl(x[b] ^ x[c], 7)

    for _ in range(10):
        qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15)
        qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14)
    return [(x[i] + st[i]) & 0xFFFFFFFF for i in range(16)]


class Stream:
    """rand_chacha ChaCha20Rng word stream (assumed conventions)."""

    def __init__(self, seed: bytes):
        self.key, self.block, self.buf = seed, 0, []

    def u32(self):
        if not self.buf:
            self.buf = chacha_block(self.key, self.block)
            self.block += 1
        return self.buf.pop(0)

    def u64(self):
        lo = self.u32()
        return lo | (self.u32() << 32)

    def fp(self):
        while True:
            x = self.u64() & P
            if x < P:
                return x


def domain_separate(seed: bytes, tag: bytes) -> bytes:
    s = bytearray(seed)
    for i, b in enumerate(tag):
        s[i % 32] ^= b
    st = Stream(bytes(s))
    return bytes(st.u32() & 0xFF for _ in range(32))


def sparse_threshold() -> int:
    mu = math.pow(1024.0, 0.7 - 1.0)
    return min(int(mu * 4294967296.0), 0xFFFFFFFF)


def round_half_away(fr: Fraction) -> int:
    if fr >= 0:
        return int(math.floor(fr + Fraction(1, 2)))
    return -int(math.floor(-fr + Fraction(1, 2)))


def quantise(v: np.ndarray, n: int, q: int) -> list[int]:
    out = [0] * n
    for i, x in enumerate(v.astype(np.float32)):
        k = round_half_away(Fraction(float(x)) * q)
        out[i] = k % P
    return out


def sha(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def field_rand(rng: np.random.Generator, size) -> np.ndarray:
    return rng.integers(0, P, size=size, dtype=np.uint64)


def matvec_exact(mat: np.ndarray, vec: np.ndarray) -> list[int]:
    rows, cols = mat.shape
    v = [int(x) for x in vec]
    return [sum(int(a) * b for a, b in zip(mat[i].tolist(), v)) % P for i in range(rows)]


def main() -> None:
    fx = {}

    # ---- ChaCha20 block function vs RFC 7539 (pins the block function)
    for tv in RFC7539_A1:
        words = chacha_block(bytes.fromhex(tv["key"]), tv["counter"])
        assert struct.pack("<16I", *words).hex() == tv["keystream"], tv
    fx["chacha20_rfc7539_a1"] = RFC7539_A1

    # ---- field KATs from crypto.rs:1079-1101 plus exact big-int cases
    kat = [
        ("reduce", 0, 0), ("reduce", P, 0), ("reduce", 2 * P, 0), ("reduce", P - 1, P - 1),
        ("reduce", (P - 7) * (P - 3), ((P - 7) * (P - 3)) % P),
    ]
    for a, b in [(0, 0), (1, 1), (P - 1, P - 1), (12345, 67890)]:
        kat += [("add", a, b, (a + b) % P), ("sub", a, b, (a - b) % P), ("mul", a, b, (a * b) % P)]
    rng = np.random.default_rng(0xF1E1D)
    for _ in range(64):
        x = int(rng.integers(0, 1 << 63)) << 65 | int(rng.integers(0, 1 << 63)) << 2 | 3
        x &= (1 << 128) - 1
        kat.append(("reduce", x, x % P))
    for x in [(1 << 128) - 1, (1 << 127), (1 << 122) - 1, 64 * (P - 1) ** 2 % (1 << 128)]:
        kat.append(("reduce", x, x % P))
    fx["field_kat"] = [[str(v) for v in k] for k in kat]

    # ---- params KATs (params.rs:209-285) and the sparse threshold
    fx["params"] = dict(n=1024, n1_by_k={"1": 512, "2": 256, "3": 128, "4": 64}, mu=0.125,
                        lambda_prime=3, floor=0.0078125, sparse_threshold=sparse_threshold())
    assert fx["params"]["sparse_threshold"] == (1 << 29) - 1

    # ---- quantise_vector (crypto.rs:105-122): half-away-from-zero cases
    qv = np.array([0.5, -0.5, 2.0 ** -21, -(2.0 ** -21), 3 * 2.0 ** -21, -3 * 2.0 ** -21,
                   0.0, -0.0, 1.0, -1.0, 0.1234567, -0.7654321, 2.0 ** -22, -(2.0 ** -22)],
                  dtype=np.float32)
    fx["quantise"] = dict(v=[float(x) for x in qv], n=16, q=1 << 20,
                          out=[str(x) for x in quantise(qv, 16, 1 << 20)])
    # EMVP quantize(0.5) = 524288 (scorer-emvp/src/crypto.rs:480-488)
    assert quantise(np.array([0.5], np.float32), 1, 1 << 20)[0] == 524288

    # ---- fp_to_signed (crypto.rs:128-135)
    fx["fp_to_signed"] = [[str(s), str(s if s < P // 2 else s - P)]
                          for s in [0, 1, P // 2 - 1, P // 2, P // 2 + 1, P - 1, 1 << 20,
                                    P - (1 << 20)]]

    # ---- domain_separate / derive_h_seed / first draws (convention cross-check)
    seeds = [bytes([42] * 32), bytes(range(32)), bytes([0] * 32)]
    fx["streams"] = []
    for s in seeds:
        st = Stream(s)
        first_fp = [str(st.fp()) for _ in range(8)]
        st2 = Stream(s)
        first_u32 = [st2.u32() for _ in range(8)]
        fx["streams"].append(dict(seed=s.hex(), first_fp=first_fp, first_u32=first_u32,
                                  protocol2=domain_separate(s, b"protocol2").hex()))

    # ---- matvec closed forms (crypto.rs:606-626): m mirrors lib.rs:1267
    fx["matvec"] = []
    for (rows, cols, seed) in [(1, 1024, 1), (7, 1024, 2), (63, 1024, 3), (64, 1024, 4),
                               (65, 1024, 5), (100, 512, 6), (33, 128, 7), (5, 1, 8)]:
        r = np.random.default_rng(seed)
        mat = field_rand(r, (rows, cols))
        vec = field_rand(r, cols)
        fx["matvec"].append(dict(rows=rows, cols=cols, seed=seed, sha=sha(mat, vec),
                                 out=[str(x) for x in matvec_exact(mat, vec)]))
    # max-accumulator stress: all entries P-1 -> each row = cols (mod p)
    fx["matvec_all_pm1"] = dict(rows=3, cols=1024, out=str((1024 * (P - 1) ** 2) % P))

    # ---- transpose matvec (crypto.rs:753-769)
    r = np.random.default_rng(11)
    mat = field_rand(r, (37, 128))
    vec = field_rand(r, 37)
    tout = [sum(int(vec[i]) * int(mat[i, j]) for i in range(37)) % P for j in range(128)]
    fx["matvec_transpose"] = dict(rows=37, cols=128, seed=11, sha=sha(mat, vec),
                                  out=[str(x) for x in tout])

    # ---- EMVP block partials (scorer-emvp/src/crypto.rs:200-222), m = 3
    r = np.random.default_rng(13)
    m_hat = field_rand(r, (3, 1292))
    q_hat = field_rand(r, 1292)
    parts = []
    for si in range(76):
        for i in range(3):
            parts.append(sum(int(m_hat[i, si * 17 + l]) * int(q_hat[si * 17 + l])
                             for l in range(17)) % P)
    fx["emvp"] = dict(m=3, seed=13, sha=sha(m_hat, q_hat), out=[str(x) for x in parts],
                      kat_mul_pm1=str(((P - 1) * (P - 1)) % P), kat_add=[str((P - 1 + 1) % P),
                                                                      str((2 * P - 2) % P)])

    # ---- Tiptoe (simplepir.rs:191-204 / :293-310)
    fx["tiptoe_kat"] = dict(m=[[1, 2, 3], [4, 5, 6]], x=[[1, 0], [0, 1], [1, 1]],
                            mx=[[4, 5], [10, 11]])
    r = np.random.default_rng(17)
    mdb = r.integers(0, 32, size=(9, 40), dtype=np.uint64)
    ct = r.integers(0, 1 << 64, size=40, dtype=np.uint64)
    ans64 = [sum(int(a) * int(b) for a, b in zip(mdb[i].tolist(), ct.tolist())) & M64
             for i in range(9)]
    fx["tiptoe"] = dict(rows=9, cols=40, seed=17, sha=sha(mdb, ct), ans64=[str(x) for x in ans64],
                        ans32=[str(x & 0xFFFFFFFF) for x in ans64])

    # ---- top-k tie rule: descending score, ties keep input order
    ids = [10, 11, 12, 13, 14, 15, 16]
    scores = [5, 7, 7, -1, 7, 5, 9]
    fx["topk"] = dict(ids=ids, scores=scores, k=4, out_ids=[16, 11, 12, 14],
                      out_scores=[9, 7, 7, 7])

    # ---- probe_route: float32 sequential L2, numpy float32 ops (no FMA)
    r = np.random.default_rng(19)
    cents = r.standard_normal((20, 24)).astype(np.float32)
    qs = r.standard_normal((4, 24)).astype(np.float32)
    routes = []
    for q in qs:
        d = []
        for c in cents:
            acc = np.float32(-0.0)
            for a, b in zip(q, c):
                diff = np.float32(a - b)
                acc = np.float32(acc + np.float32(diff * diff))
            d.append(np.sqrt(acc))
        order = sorted(range(20), key=lambda i: (d[i], i))
        routes.append(order[:5])
    fx["probe_route"] = dict(seed=19, sha=sha(cents, qs), nprobe=5, out=routes)

    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(fx, f, indent=0, sort_keys=True)
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
