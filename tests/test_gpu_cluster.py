"""TrajectoryClusterPlan on the GPU against oracle/cluster.py (clustering.rs):
DBSCAN labels, medoids and sizes identical on the oracle's pairwise matrix,
k-means labels, centroids and sizes identical on the oracle's superposed
features (seeded, deterministic), and the reference's own known answers."""
import numpy as np
import pytest

import cluster
import oracle
from warp_md_amd import System, Trajectory, TrajectoryClusterPlan

pytestmark = pytest.mark.gpu


def _sys(n):
    return System.from_arrays({"name": ["CA"] * n})


def _conformers(F, n, n_shapes, seed, noise=0.08):
    rng = np.random.default_rng(seed)
    base = rng.normal(scale=4.0, size=(n, 3))
    shapes = [base + rng.normal(scale=1.5, size=base.shape) for _ in range(n_shapes)]
    out = []
    for i in range(F):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        s = shapes[rng.integers(n_shapes)]
        out.append(s @ R.T + rng.uniform(-20, 20, 3) + rng.normal(scale=noise, size=s.shape))
    return np.asarray(out, np.float32)


def test_reference_known_answers():
    fr = np.array([[[0, 0, 0], [1, 0, 0]], [[0, 0, 0], [1.1, 0, 0]], [[0, 0, 0], [5, 0, 0]]], np.float32)
    out = TrajectoryClusterPlan([0, 1], method="dbscan", eps=0.1, min_samples=2).run(
        Trajectory.from_numpy(fr), _sys(2), device="cuda:0")
    assert out["labels"].tolist() == [0, 0, -1]
    assert out["centroids"].tolist() == [0] and out["sizes"].tolist() == [2]
    assert out["method"] == "dbscan" and out["n_frames"] == 3
    fr = np.array([[[0, 0, 0], [1, 0, 0]], [[0, 0, 0], [1.2, 0, 0]],
                   [[0, 0, 0], [5, 0, 0]], [[0, 0, 0], [5.2, 0, 0]]], np.float32)
    a = TrajectoryClusterPlan([0, 1], method="kmeans", n_clusters=2, max_iter=64, tol=1e-6, seed=42).run(
        Trajectory.from_numpy(fr), _sys(2), device="cuda:0")
    b = TrajectoryClusterPlan([0, 1], method="kmeans", n_clusters=2, max_iter=64, tol=1e-6, seed=42).run(
        Trajectory.from_numpy(fr), _sys(2), device="cuda:0")
    assert np.array_equal(a["labels"], b["labels"]) and sorted(a["sizes"].tolist()) == [2, 2]
    with pytest.raises(RuntimeError, match="eps must be > 0"):
        TrajectoryClusterPlan([0, 1], eps=0.0).run(Trajectory.from_numpy(fr), _sys(2), device="cuda:0")
    with pytest.raises(RuntimeError, match="memory_budget_bytes"):
        TrajectoryClusterPlan([0, 1], memory_budget_bytes=8).run(Trajectory.from_numpy(fr), _sys(2),
                                                                 device="cuda:0")


@pytest.mark.parametrize("F,n,shapes", [(60, 30, 3), (257, 50, 5)])
def test_dbscan_matches_oracle(F, n, shapes):
    fr = _conformers(F, n, shapes, seed=F)
    out = TrajectoryClusterPlan(np.arange(n), method="dbscan", eps=0.6, min_samples=4).run(
        Trajectory.from_numpy(fr), _sys(n), chunk_frames=50, device="cuda:0")
    m, _ = oracle.pairwise_rmsd(fr, np.arange(n), "rms")
    labels, cents, sizes = cluster.dbscan(m, np.float64(np.float32(0.6)), 4)
    assert np.array_equal(out["labels"], labels)
    assert np.array_equal(out["centroids"], cents) and np.array_equal(out["sizes"], sizes)
    assert len(sizes) == shapes


@pytest.mark.parametrize("F,n,k,seed", [(60, 30, 3, 0), (200, 40, 6, 11)])
def test_kmeans_matches_oracle(F, n, k, seed):
    fr = _conformers(F, n, k, seed=F + 1, noise=0.2)
    out = TrajectoryClusterPlan(np.arange(n), method="kmeans", n_clusters=k, max_iter=50, tol=1e-6,
                                seed=seed).run(Trajectory.from_numpy(fr), _sys(n), device="cuda:0")
    labels, cents, sizes = cluster.kmeans(cluster.aligned_features(fr), k, 50, np.float64(np.float32(1e-6)), seed)
    assert np.array_equal(out["labels"], labels)
    assert np.array_equal(out["centroids"], cents) and np.array_equal(out["sizes"], sizes)
