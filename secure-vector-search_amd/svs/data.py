"""Seeded synthetic corpora for the BN workloads.

A Gaussian mixture in d = 768: component centres ~ N(0, I), members =
centre + sigma * N(0, I), then L2-normalised; queries are drawn the same
way. Component weights are uniform or Zipf(s) (w_i ~ i^-s).

Two ways to cut a corpus into IVF lists:
* ivf_lists: the mixture components, routed by the normalised component
  centres (every member's nearest centre is its own at sigma = 0.3, d = 768);
* kmeans_lists: a seeded k-means of the corpus in the shape of the
  reference's ivf-index (kmeans.rs:15-133: k-means++ initialisation with
  D^2 sampling, then Lloyd iterations until the centroids stop changing, an
  empty list's centroid set to zero; ivf.rs:16-27: every vector to its
  nearest centroid, members in corpus order). k-means++ splits the heavy
  Zipf components over several lists and merges light ones, so list sizes
  are those of a trained IVF index, not of the generator.
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


def _nearest(x: np.ndarray, c: np.ndarray, chunk: int = 1 << 15) -> np.ndarray:
    """Index of the nearest centroid (squared L2, first on ties) per row."""
    cc = (c.astype(np.float32) ** 2).sum(axis=1)
    out = np.empty(x.shape[0], np.int64)
    for a in range(0, x.shape[0], chunk):
        xb = x[a:a + chunk]
        d = cc[None, :] - 2.0 * (xb @ c.T)  # + |x|^2, constant per row
        out[a:a + chunk] = np.argmin(d, axis=1)
    return out


def kmeans(vectors: np.ndarray, k: int, seed: int = 42, max_iter: int = 25,
           init_sample: int = 8192, train_sample: int = 40 * 1024) -> np.ndarray:
    """Seeded k-means (float32), the shape of ivf-index/src/kmeans.rs:15-133.
    k-means++ runs on `init_sample` rows drawn without replacement, Lloyd on
    `train_sample` rows (the whole corpus when it is smaller); returns the
    k x d centroids."""
    x_all = np.ascontiguousarray(vectors, np.float32)
    n = x_all.shape[0]
    if not 0 < k <= n:
        raise ValueError("need 0 < k <= number of vectors")
    rng = np.random.default_rng(seed)
    train = np.sort(rng.choice(n, min(n, max(train_sample, k)), replace=False))
    x = x_all[train]
    init = x[np.sort(rng.choice(x.shape[0], min(x.shape[0], max(init_sample, k)),
                                replace=False))]
    xx = (init ** 2).sum(axis=1)
    # k-means++: the first centroid uniform, each next one with probability
    # proportional to the squared distance to the nearest chosen one,
    # kept incrementally (kmeans.rs:84-133)
    cents = np.empty((k, x.shape[1]), np.float32)
    first = int(rng.integers(0, init.shape[0]))
    cents[0] = init[first]
    min_d = np.maximum(xx - 2.0 * (init @ cents[0]) + (cents[0] ** 2).sum(), 0.0)
    for ci in range(1, k):
        cum = np.cumsum(min_d, dtype=np.float64)
        thr = rng.random() * cum[-1]
        chosen = min(int(np.searchsorted(cum, thr, side="left")), init.shape[0] - 1)
        cents[ci] = init[chosen]
        d = np.maximum(xx - 2.0 * (init @ cents[ci]) + (cents[ci] ** 2).sum(), 0.0)
        np.minimum(min_d, d, out=min_d)
    # Lloyd (kmeans.rs:26-33, 57-78)
    for _ in range(max_iter):
        a = _nearest(x, cents)
        counts = np.bincount(a, minlength=k)
        order = np.argsort(a, kind="stable")
        starts = np.searchsorted(a[order], np.arange(k))
        nz = counts > 0
        sums = np.add.reduceat(x[order].astype(np.float64), starts[nz], axis=0)
        new = np.zeros_like(cents)
        new[nz] = (sums / counts[nz, None]).astype(np.float32)
        if np.array_equal(new, cents):
            break
        cents = new
    return cents


def kmeans_lists(corpus: Corpus, k: int, seed: int = 42, max_iter: int = 25,
                 **kw) -> tuple[np.ndarray, list[np.ndarray]]:
    """(centroids, lists): seeded k-means centroids of the corpus and its
    IVF lists, every vector in the list of its nearest centroid, ascending
    ids (ivf-index/src/ivf.rs:16-27)."""
    cents = kmeans(corpus.vectors, k, seed=seed, max_iter=max_iter, **kw)
    a = _nearest(corpus.vectors, cents)
    order = np.argsort(a, kind="stable")
    bounds = np.searchsorted(a[order], np.arange(k + 1))
    return cents, [order[bounds[i]:bounds[i + 1]].astype(np.uint32) for i in range(k)]
