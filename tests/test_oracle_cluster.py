"""Pins oracle/cluster.py (TrajectoryClusterPlan's DBSCAN and k-means,
clustering.rs:290-550) with the reference's own tests and closed forms."""
import numpy as np

import cluster
import oracle


def _frames():
    # trajectory_cluster_dbscan_basic (geometry_matrix_tests.rs:398-425)
    return np.array([[[0, 0, 0], [1, 0, 0]], [[0, 0, 0], [1.1, 0, 0]], [[0, 0, 0], [5, 0, 0]]], np.float32)


def test_dbscan_reference_known_answer():
    m, _ = oracle.pairwise_rmsd(_frames(), [0, 1], "rms")
    labels, cents, sizes = cluster.dbscan(m, np.float64(np.float32(0.1)), 2)
    assert labels.tolist() == [0, 0, -1]
    assert cents.tolist() == [0] and sizes.tolist() == [2]


def test_kmeans_reference_known_answer():
    # trajectory_cluster_kmeans_seed_deterministic (:427-480)
    fr = np.array([[[0, 0, 0], [1, 0, 0]], [[0, 0, 0], [1.2, 0, 0]],
                   [[0, 0, 0], [5, 0, 0]], [[0, 0, 0], [5.2, 0, 0]]], np.float32)
    feats = cluster.aligned_features(fr)
    a = cluster.kmeans(feats, 2, 64, 1e-6, 42)
    b = cluster.kmeans(feats, 2, 64, 1e-6, 42)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert sorted(a[2].tolist()) == [2, 2]
    assert a[0][0] == a[0][1] and a[0][2] == a[0][3] and a[0][0] != a[0][2]


def test_xorshift_closed_form():
    st = [0]
    v = cluster.xorshift64(st)
    x = 0x9E3779B97F4A7C15
    x ^= (x << 13) % 2 ** 64
    x ^= x >> 7
    x ^= (x << 17) % 2 ** 64
    assert v == x and st[0] == x


def test_blobs_recovered():
    rng = np.random.default_rng(0)
    base = rng.normal(scale=3.0, size=(20, 3))
    shapes = [base, base * np.array([1.6, 0.6, 1.0]), base + np.linspace(0, 6, 20)[:, None]]
    fr = np.stack([shapes[i % 3] + rng.normal(scale=0.05, size=base.shape) for i in range(18)]).astype(np.float32)
    m, _ = oracle.pairwise_rmsd(fr, np.arange(20), "rms")
    labels, _, sizes = cluster.dbscan(m, 0.5, 3)
    assert sorted(sizes.tolist()) == [6, 6, 6]
    assert all(labels[i] == labels[i % 3] for i in range(18))
    kl, _, ks = cluster.kmeans(cluster.aligned_features(fr), 3, 100, 1e-6, 7)
    assert sorted(ks.tolist()) == [6, 6, 6]
    assert all(kl[i] == kl[i % 3] for i in range(18))
