"""Seeded synthetic corpora for the BN workloads.

A Gaussian mixture in d = 768: component centres ~ N(0, I), members =
centre + sigma * N(0, I), then L2-normalised; queries are drawn the same
way. Component weights are uniform or Zipf(s) (w_i ~ i^-s). The IVF lists
of a corpus are its mixture components and the routing centroids are the
normalised component centres: for blobs this well separated (sigma = 0.3,
d = 768) every member's nearest centre is its own, which is the fixed point
k-means converges to, without a float k-means in the loop.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Corpus:
    vectors: np.ndarray      # n x d float32, unit norm
    component: np.ndarray    # n int32, generating component
    centres: np.ndarray      # c x d float32, normalised centres (routing centroids)
    weights: np.ndarray      # c float64


def zipf_weights(c: int, s: float) -> np.ndarray:
    w = np.arange(1, c + 1, dtype=np.float64) ** (-s)
    return w / w.sum()


def _draw(rng, centres_raw, comp, sigma, d, chunk=1 << 16):
    out = np.empty((comp.size, d), np.float32)
    for a in range(0, comp.size, chunk):
        b = min(comp.size, a + chunk)
        x = centres_raw[comp[a:b]] + sigma * rng.standard_normal((b - a, d), dtype=np.float32)
        x /= np.linalg.norm(x, axis=1, keepdims=True)
        out[a:b] = x
    return out


def make_corpus(n: int, d: int = 768, components: int = 64, sigma: float = 0.3,
                zipf_s: float | None = None, seed: int = 42) -> Corpus:
    rng = np.random.default_rng(seed)
    centres_raw = rng.standard_normal((components, d), dtype=np.float32)
    w = zipf_weights(components, zipf_s) if zipf_s else np.full(components, 1.0 / components)
    comp = rng.choice(components, size=n, p=w).astype(np.int32)
    vecs = _draw(rng, centres_raw, comp, sigma, d)
    centres = centres_raw / np.linalg.norm(centres_raw, axis=1, keepdims=True)
    return Corpus(vecs, comp, centres.astype(np.float32), w)


def make_queries(corpus: Corpus, nq: int, sigma: float = 0.3, seed: int = 4242) -> np.ndarray:
    rng = np.random.default_rng(seed)
    c = corpus.centres.shape[0]
    d = corpus.centres.shape[1]
    # re-derive the raw centres' directions: queries use the normalised
    # centres scaled back to the typical raw norm sqrt(d)
    centres_raw = corpus.centres * np.float32(np.sqrt(d))
    comp = rng.choice(c, size=nq, p=corpus.weights).astype(np.int32)
    return _draw(rng, centres_raw, comp, sigma, d)


def ivf_lists(corpus: Corpus) -> list[np.ndarray]:
    """Members of each component, ascending ids (ivf-index/src/ivf.rs:16-27
    pushes members in corpus order)."""
    order = np.argsort(corpus.component, kind="stable")
    bounds = np.searchsorted(corpus.component[order], np.arange(corpus.centres.shape[0] + 1))
    return [order[bounds[i]:bounds[i + 1]].astype(np.uint32) for i in range(len(bounds) - 1)]
