"""CPU restatement of TrajectoryClusterPlan's host algorithms -- TEST
INFRASTRUCTURE ONLY (imported by tests/ and bench.py's CPU baseline as the
checker, never by the product package).

* ``dbscan``        -- dbscan_cluster + neighbors_within,
  ``crates/traj-engine/src/plans/analysis/clustering.rs:290-382``
* ``kmeans``        -- kmeans_cluster, sq_dist, fisher_yates_shuffle,
  xorshift64, ``clustering.rs:420-550``
* ``aligned_features`` -- build_aligned_features / align_to_reference,
  ``clustering.rs:384-418`` (Kabsch by numpy.linalg.svd with the reflection
  fix of geometry_math.rs:284-295)

Pinned by the reference's trajectory_cluster_dbscan_basic and
trajectory_cluster_kmeans_seed_deterministic (geometry_matrix_tests.rs:398-480)
and by the structure of well-separated inputs (tests/test_oracle_cluster.py).
"""
from __future__ import annotations

from collections import deque

import numpy as np

_MASK = (1 << 64) - 1


def xorshift64(state: list) -> int:
    x = state[0]
    if x == 0:
        x = 0x9E3779B97F4A7C15
    x ^= (x << 13) & _MASK
    x ^= x >> 7
    x ^= (x << 17) & _MASK
    state[0] = x
    return x


def dbscan(dist: np.ndarray, eps: float, min_samples: int):
    """dist: full (n, n) matrix.  Returns (labels int32, centroids uint32, sizes uint32)."""
    n = dist.shape[0]
    d = lambda i, j: 0.0 if i == j else float(dist[i, j])
    UNCL, NOISE = -2, -1
    labels = [UNCL] * n
    visited = [False] * n
    cid = 0

    def neighbors(p):
        return [j for j in range(n) if d(p, j) <= eps]

    for i in range(n):
        if visited[i]:
            continue
        visited[i] = True
        neigh = neighbors(i)
        if len(neigh) < min_samples:
            labels[i] = NOISE
            continue
        labels[i] = cid
        queued = [False] * n
        queue = deque()
        for p in neigh:
            if p != i and not queued[p]:
                queue.append(p)
                queued[p] = True
        while queue:
            p = queue.popleft()
            if not visited[p]:
                visited[p] = True
                nb = neighbors(p)
                if len(nb) >= min_samples:
                    for q in nb:
                        if not queued[q]:
                            queue.append(q)
                            queued[q] = True
            if labels[p] in (UNCL, NOISE):
                labels[p] = cid
        cid += 1
    cents, sizes = [], []
    for c in range(cid):
        members = [i for i in range(n) if labels[i] == c]
        sizes.append(len(members))
        best, best_sum = members[0], float("inf")
        for a in members:
            s = 0.0
            for b in members:
                s += d(a, b)
            if s < best_sum:
                best_sum, best = s, a
        cents.append(best)
    return (np.asarray(labels, np.int32), np.asarray(cents, np.uint32),
            np.asarray(sizes, np.uint32))


def _kabsch(frame: np.ndarray, ref: np.ndarray):
    x = frame.astype(np.float64)
    y = ref.astype(np.float64)
    cx = x.sum(0) / len(x)
    cy = y.sum(0) / len(y)
    h = (x - cx).T @ (y - cy)
    u, s, vt = np.linalg.svd(h)
    r = vt.T @ u.T
    if np.linalg.det(r) < 0.0:
        vt = vt.copy()
        vt[2] *= -1.0
        r = vt.T @ u.T
    return r, cx, cy


def aligned_features(frames: np.ndarray) -> np.ndarray:
    """frames (F, n, 3) -> (F, 3 n) f64: R (x - cx) + cy onto frame 0."""
    ref = frames[0]
    out = []
    for f in frames:
        if len(f) == 0:
            out.append(np.zeros(0))
            continue
        r, cx, cy = _kabsch(f, ref)
        out.append(((f.astype(np.float64) - cx) @ r.T + cy).reshape(-1))
    return np.asarray(out)


def kmeans(features: np.ndarray, n_clusters: int, max_iter: int, tol: float, seed: int):
    n = features.shape[0]
    k = min(max(n_clusters, 1), n)
    state = [seed if seed != 0 else 0x9E3779B97F4A7C15]
    order = list(range(n))
    for i in range(n - 1, 0, -1):
        j = xorshift64(state) % (i + 1)
        order[i], order[j] = order[j], order[i]
    centers = [features[order[c]].copy() for c in range(k)]
    labels = np.zeros(n, np.int32)
    sq = lambda a, b: float(np.sum((a - b) ** 2))
    for _ in range(max(max_iter, 1)):
        counts = [0] * k
        sums = [np.zeros(features.shape[1]) for _ in range(k)]
        for i in range(n):
            dists = [sq(features[i], centers[c]) for c in range(k)]
            best = int(np.argmin(dists))
            labels[i] = best
            counts[best] += 1
            sums[best] += features[i]
        shift = 0.0
        for c in range(k):
            if counts[c] == 0:
                nc = features[xorshift64(state) % n].copy()
            else:
                nc = sums[c] * (1.0 / counts[c])
            shift = max(shift, np.sqrt(sq(centers[c], nc)))
            centers[c] = nc
        if shift <= max(tol, 0.0):
            break
    sizes = np.bincount(labels, minlength=k).astype(np.uint32)
    cents = np.zeros(k, np.uint32)
    for c in range(k):
        best_i, best_d = 0, float("inf")
        for i in range(n):
            if labels[i] != c:
                continue
            dd = sq(features[i], centers[c])
            if dd < best_d:
                best_d, best_i = dd, i
        cents[c] = best_i
    return labels, cents, sizes
