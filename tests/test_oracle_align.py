"""Pins the oracle's AlignPlan / SuperposePlan restatement (orc_align,
align/plan_exec.rs:215-243, 422-456): the reference's known answers, an
independent numpy weighted Kabsch, and closed forms."""
import numpy as np

import oracle


def _rot(q):
    q = np.asarray(q, np.float64)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_superpose_plan_basic():
    # analysis_smoke_tests.rs:171-195: positions0 (0,0,0), (1,0,0); frame (1,1,0), (1,2,0)
    ref = np.array([[0, 0, 0], [1, 0, 0]], np.float32)
    c = np.array([[[1, 1, 0], [1, 2, 0]]], np.float32)
    _, _, xyz = oracle.align(c, [0, 1], ref, superpose=True)
    assert np.linalg.norm(xyz[0, 0] - [0, 0, 0]) < 1e-5
    assert np.linalg.norm(xyz[0, 1] - [1, 0, 0]) < 1e-5


def test_rotation_and_translation_recovered():
    # python/warp_md/tests/test_superpose.py:41-67
    ref = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64)
    rot = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], np.float64)
    frame1 = ref @ rot.T + np.array([2.0, 3.0, 0.0])
    c = np.stack([ref, frame1]).astype(np.float32)
    rt, rt64, xyz = oracle.align(c, [0, 1, 2], ref.astype(np.float32), superpose=True)
    assert np.allclose(xyz[0], ref, atol=1e-5) and np.allclose(xyz[1], ref, atol=1e-5)
    R = rt64[1, :9].reshape(3, 3)
    assert np.allclose(R, rot.T, atol=1e-12)  # undoes the rotation
    assert np.allclose(R @ frame1[0] + rt64[1, 9:], ref[0], atol=1e-12)


def test_numpy_weighted_kabsch():
    rng = np.random.default_rng(3)
    ref = rng.normal(scale=5, size=(30, 3)).astype(np.float32)
    c = np.stack([ref @ _rot(rng.normal(size=4)).T + rng.uniform(-9, 9, 3)
                  + rng.normal(scale=0.4, size=(30, 3)) for _ in range(4)]).astype(np.float32)
    sel = np.arange(2, 28)
    w = rng.uniform(1, 16, size=sel.size).astype(np.float32)
    for weights in (None, w):
        rt, rt64, _ = oracle.align(c, sel, ref, weights=weights)
        ww = np.ones(sel.size) if weights is None else w.astype(np.float64)
        for f in range(4):
            x = c[f][sel].astype(np.float64)
            y = ref[sel].astype(np.float64)
            cx = (x * ww[:, None]).sum(0) / ww.sum()
            cy = (y * ww[:, None]).sum(0) / ww.sum()
            H = ((x - cx) * ww[:, None]).T @ (y - cy)
            u, s, vt = np.linalg.svd(H)
            d = np.sign(np.linalg.det(vt.T @ u.T))
            R = vt.T @ np.diag([1, 1, d]) @ u.T
            assert np.allclose(rt64[f, :9].reshape(3, 3), R, atol=1e-10)
            assert np.allclose(rt64[f, 9:], cy - R @ cx, atol=1e-9)


def test_norotate_is_centroid_shift():
    rng = np.random.default_rng(4)
    ref = rng.normal(size=(10, 3)).astype(np.float32)
    c = rng.normal(size=(3, 10, 3)).astype(np.float32) + 5
    rt, rt64, xyz = oracle.align(c, np.arange(10), ref, norotate=True, superpose=True)
    for f in range(3):
        assert np.array_equal(rt64[f, :9], np.eye(3).reshape(9))
        t = ref.astype(np.float64).mean(0) - c[f].astype(np.float64).mean(0)
        assert np.allclose(rt64[f, 9:], t, atol=1e-12)
        assert np.allclose(xyz[f], c[f] + t, atol=1e-5)
