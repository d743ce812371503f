"""Pins the oracle's PairwiseRmsdPlan (rmsd/plan.rs:586-675; orc_pairwise_rmsd):
the reference's own known answer, agreement with the separately pinned
RmsdPlan restatement pair by pair, an independent numpy SVD Kabsch, and what
the mathematics fixes (rigid motions, symmetry, metric properties)."""
import numpy as np
import pytest

import oracle


def _frames(F, n, seed=0, noise=0.3):
    rng = np.random.default_rng(seed)
    base = rng.normal(scale=6.0, size=(n, 3))
    out = []
    for _ in range(F):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        out.append(base @ R.T + rng.uniform(-30, 30, 3) + rng.normal(scale=noise, size=(n, 3)))
    return np.asarray(out, np.float32)


def test_reference_known_answer_nofit():
    # pairwise_rmsd_plan_basic (traj-engine/src/tests/geometry_matrix_tests.rs:1438-1460)
    c = np.array([[[0, 0, 0], [1, 0, 0]], [[0, 0, 0], [2, 0, 0]]], np.float32)
    out, _ = oracle.pairwise_rmsd(c, [0, 1], "nofit")
    assert out.shape == (2, 2)
    assert abs(out[0, 1] - np.sqrt(np.float32(0.5))) < 1e-6
    assert abs(out[1, 0] - np.sqrt(np.float32(0.5))) < 1e-6
    assert out[0, 0] == 0 and out[1, 1] == 0


def test_two_atom_rms_closed_form():
    # centred segments of lengths 1 and 2 on one line: residual 0.5 per atom
    c = np.array([[[0, 0, 0], [1, 0, 0]], [[0, 0, 0], [0, 2, 0]]], np.float32)
    out, out64 = oracle.pairwise_rmsd(c, [0, 1], "rms")
    assert abs(out64[0, 1] - 0.5) < 1e-12


@pytest.mark.parametrize("metric", ["rms", "nofit"])
def test_pairs_equal_rmsd_plan(metric):
    c = _frames(6, 40, seed=1)
    sel = np.arange(3, 37)
    out, out64 = oracle.pairwise_rmsd(c, sel, metric)
    for j in range(6):
        ref = c[j][sel]
        _, r64 = oracle.rmsd_plan(c, sel, ref, align=(metric == "rms"))
        for i in range(6):
            if i == j:
                continue
            # RmsdPlan's centroid is sum * (1/n), PairwiseRmsdPlan's sum / n
            assert abs(out64[i, j] - r64[i]) <= 1e-10 * max(1.0, r64[i]), (i, j)


def test_numpy_kabsch_restatement():
    c = _frames(5, 25, seed=2, noise=0.8).astype(np.float64)
    out, out64 = oracle.pairwise_rmsd(c.astype(np.float32), np.arange(25), "rms")
    c = c.astype(np.float32).astype(np.float64)
    for i in range(5):
        for j in range(i + 1, 5):
            x = c[i] - c[i].mean(0)
            y = c[j] - c[j].mean(0)
            u, s, vt = np.linalg.svd(x.T @ y)
            d = np.sign(np.linalg.det(vt.T @ u.T))
            R = vt.T @ np.diag([1, 1, d]) @ u.T
            v = np.sqrt(np.mean(np.sum((x @ R.T - y) ** 2, axis=1)))
            assert abs(out64[i, j] - v) < 1e-10
            assert out64[j, i] == out64[i, j]


def test_rigid_motion_and_metric_properties():
    c = _frames(8, 30, seed=3, noise=0.0)
    rms, _ = oracle.pairwise_rmsd(c, np.arange(30), "rms")
    nofit, _ = oracle.pairwise_rmsd(c, np.arange(30), "nofit")
    assert np.max(rms) < 1e-4  # rigid copies superpose exactly
    assert np.min(nofit[~np.eye(8, dtype=bool)]) > 1.0
    c = _frames(9, 30, seed=4, noise=1.0)
    rms, _ = oracle.pairwise_rmsd(c, np.arange(30), "rms")
    nofit, n64 = oracle.pairwise_rmsd(c, np.arange(30), "nofit")
    assert np.array_equal(rms, rms.T) and np.array_equal(nofit, nofit.T)
    assert np.all(np.diag(rms) == 0) and np.all(np.diag(nofit) == 0)
    assert np.all(rms <= nofit + 1e-6)
    # nofit is a Euclidean distance (scaled): triangle inequality
    for i in range(9):
        for j in range(9):
            for k in range(9):
                assert n64[i, k] <= n64[i, j] + n64[j, k] + 1e-9


def test_fewer_than_two_atoms_gives_zeros():
    c = _frames(4, 5, seed=5)
    out, _ = oracle.pairwise_rmsd(c, [2], "rms")
    assert np.all(out == 0) and out.shape == (4, 4)


# ---- metric "dme" (rmsd/plan.rs:612-650, pair_distances_compact :1279-1302) ----

def test_dme_three_atom_closed_form():
    # frame 0: 3-4-5 triangle; frame 1: the same triangle doubled (6-8-10)
    c = np.array([[[0, 0, 0], [3, 0, 0], [0, 4, 0]],
                  [[0, 0, 0], [6, 0, 0], [0, 8, 0]]], np.float32)
    out, out64 = oracle.pairwise_dme(c, [0, 1, 2])
    want = np.sqrt((3.0 ** 2 + 4.0 ** 2 + 5.0 ** 2) / 3.0)
    assert out64[0, 1] == want and out64[1, 0] == want
    assert out64[0, 0] == 0 and out64[1, 1] == 0
    assert out[0, 1] == np.float32(want)


def test_dme_exact_doubling_is_rms_of_distances():
    # doubling every coordinate doubles every f32 difference and every distance
    # exactly, so DME = sqrt(mean(d^2)) of the first frame, with d from scipy's pdist
    from scipy.spatial.distance import pdist

    rng = np.random.default_rng(7)
    # dyadic coordinates: the f32 differences are exact, as pdist's f64 ones
    x = (rng.integers(-1280, 1280, size=(37, 3)) / 64.0).astype(np.float32)
    c = np.stack([x, 2 * x])
    out, out64 = oracle.pairwise_dme(c, np.arange(37))
    d = pdist(x.astype(np.float64))
    assert abs(out64[0, 1] - np.sqrt(np.mean(d * d))) <= 1e-12 * out64[0, 1]


def test_dme_selection_order_and_subset():
    rng = np.random.default_rng(8)
    c = rng.uniform(-10, 10, size=(4, 20, 3)).astype(np.float32)
    sel = np.array([5, 2, 17, 9, 11, 3])
    out, out64 = oracle.pairwise_dme(c, sel)
    out2, out64b = oracle.pairwise_dme(c[:, sel], np.arange(sel.size))
    np.testing.assert_array_equal(out64, out64b)
    # the DME of a frame pair is independent of the other frames and symmetric
    for i in range(4):
        for j in range(4):
            if i != j:
                _, pair = oracle.pairwise_dme(c[[i, j]], sel)
                assert pair[0, 1] == out64[i, j] == out64[j, i]


def test_dme_rigid_motion_is_near_zero():
    rng = np.random.default_rng(9)
    x = rng.uniform(-15, 15, size=(50, 3))
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, a, b, d = q
    R = np.array([[1 - 2 * (b * b + d * d), 2 * (a * b - d * w), 2 * (a * d + b * w)],
                  [2 * (a * b + d * w), 1 - 2 * (a * a + d * d), 2 * (b * d - a * w)],
                  [2 * (a * d - b * w), 2 * (b * d + a * w), 1 - 2 * (a * a + b * b)]])
    c = np.stack([x, x @ R.T + 7.5]).astype(np.float32)
    out, out64 = oracle.pairwise_dme(c, np.arange(50))
    assert out64[0, 1] < 1e-4  # f32 rounding of the moved coordinates only
    noisy = c.copy()
    noisy[1] += np.float32(0.3) * rng.normal(size=(50, 3)).astype(np.float32)
    assert oracle.pairwise_dme(noisy, np.arange(50))[1][0, 1] > 0.05


def test_dme_orthorhombic_wrap():
    # two atoms 2 A apart across the boundary of a 10 A box, then 4 A apart inside it
    c = np.array([[[1, 0, 0], [9, 0, 0]], [[1, 0, 0], [5, 0, 0]]], np.float32)
    box = np.array([[10, 10, 10], [10, 10, 10]], np.float32)
    _, wrapped = oracle.pairwise_dme(c, [0, 1], pbc="orthorhombic", boxes=box)
    _, plain = oracle.pairwise_dme(c, [0, 1])
    assert wrapped[0, 1] == 2.0 and plain[0, 1] == 4.0
    # dyadic coordinates: moving single atoms by whole box vectors is exact and
    # leaves every wrapped distance, hence the DME, at exactly 0
    rng = np.random.default_rng(10)
    x = (rng.integers(0, 128, size=(30, 3)) / 8.0).astype(np.float32)
    y = x.copy()
    y[::3, 0] += 16.0
    y[1::4, 2] -= 32.0
    c = np.stack([x, y])
    box = np.array([[16, 16, 16], [16, 16, 16]], np.float32)
    _, w = oracle.pairwise_dme(c, np.arange(30), pbc="orthorhombic", boxes=box)
    assert w[0, 1] == 0.0
    assert oracle.pairwise_dme(c, np.arange(30))[1][0, 1] > 1.0


def test_dme_fewer_than_two_atoms_is_zero():
    c = np.ones((3, 4, 3), np.float32)
    out, out64 = oracle.pairwise_dme(c, [2])
    assert not out.any() and not out64.any()
