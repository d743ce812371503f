"""Pins the CPU oracle's RMSD restatement.

The reference's SVD is nalgebra 0.32.6 (not vendored), so the oracle's own 3x3
SVD is pinned against numpy.linalg.svd (LAPACK), and the Kabsch RMSD against an
independent numpy Kabsch, closed forms (rigid motions give 0, raw RMSD is the
textbook formula) and the reference's own known answers.
"""
import json
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))


def np_kabsch(x, y):
    """Independent numpy Kabsch: R minimising sum |R(x-cx) - (y-cy)|^2 over SO(3)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    cx, cy = x.mean(0), y.mean(0)
    h = (x - cx).T @ (y - cy)
    u, s, vt = np.linalg.svd(h)
    d = np.sign(np.linalg.det(vt.T @ u.T))
    D = np.diag([1.0, 1.0, d])
    r = vt.T @ D @ u.T
    return r, cx, cy


def np_rmsd(x, y):
    r, cx, cy = np_kabsch(x, y)
    diff = (x - cx) @ r.T - (y - cy)
    return float(np.sqrt((diff ** 2).sum() / len(x)))


@pytest.mark.parametrize("case", GOLD["rmsd"], ids=lambda c: c["name"])
def test_reference_known_answers(case):
    frames = np.asarray(case["frames"], np.float32)
    ref = np.asarray(case["reference"], np.float32)
    out, out64 = oracle.rmsd_plan(frames, case["sel"], ref[case["sel"]], align=case["align"])
    for v, lim in zip(out, case["expect_abs_below"]):
        assert abs(v) < lim
    if "expect_rotation_frame1" in case:
        _, _, rot, _, _ = oracle.fit_rmsd(frames, case["sel"], case["sel"], ref)
        np.testing.assert_allclose(rot[1], case["expect_rotation_frame1"], atol=1e-6)
        np.testing.assert_allclose(rot[0], np.eye(3), atol=1e-6)


def test_svd3_matches_lapack():
    rng = np.random.default_rng(0)
    mats = [rng.normal(size=(3, 3)) for _ in range(200)]
    mats.append(np.outer([1.0, 2.0, 3.0], [0.5, -1.0, 2.0]))  # rank 1
    a = rng.normal(size=(3, 2))
    mats.append(a @ rng.normal(size=(2, 3)))  # rank 2
    mats.append(np.diag([3.0, 3.0, 1.0]))  # repeated singular value
    mats.append(np.zeros((3, 3)))
    for m in mats:
        u, s, v = oracle.svd3(m)
        s_ref = np.linalg.svd(m, compute_uv=False)
        np.testing.assert_allclose(s, s_ref, atol=1e-12 * max(1.0, s_ref[0]))
        np.testing.assert_allclose(u @ np.diag(s) @ v.T, m, atol=1e-12 * max(1.0, s_ref[0]))
        np.testing.assert_allclose(v.T @ v, np.eye(3), atol=1e-12)
        np.testing.assert_allclose(u.T @ u, np.eye(3), atol=1e-10)


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def test_kabsch_rmsd_matches_numpy_and_closed_forms():
    rng = np.random.default_rng(1)
    n, F = 300, 12
    ref = rng.normal(scale=8.0, size=(n, 3)).astype(np.float32)
    frames = []
    for f in range(F):
        R = _random_rotation(rng)
        t = rng.uniform(-20, 20, size=3)
        noise = 0.0 if f == 0 else 0.5
        frames.append(ref @ R.T + t + rng.normal(scale=noise, size=(n, 3)))
    frames = np.asarray(frames, np.float32)
    sel = np.arange(n)
    out, out64 = oracle.rmsd_plan(frames, sel, ref, align=True)
    for f in range(F):
        assert abs(out64[f] - np_rmsd(frames[f], ref)) < 1e-10
    assert out64[0] < 1e-5  # rigid motion of the reference (f32 rounding of coordinates only)
    raw, raw64 = oracle.rmsd_plan(frames, sel, ref, align=False)
    expect = np.sqrt(((frames.astype(np.float64) - ref.astype(np.float64)) ** 2).sum(-1).mean(-1))
    np.testing.assert_allclose(raw64, expect, rtol=1e-14)


def test_reflection_uses_proper_rotation():
    rng = np.random.default_rng(2)
    ref = rng.normal(scale=5.0, size=(50, 3)).astype(np.float32)
    mirrored = ref.copy()
    mirrored[:, 2] *= -1
    out, out64 = oracle.rmsd_plan(mirrored[None], np.arange(50), ref, align=True)
    assert out64[0] > 0.1
    assert abs(out64[0] - np_rmsd(mirrored, ref)) < 1e-10


def test_fit_rmsd_split_selections():
    rng = np.random.default_rng(4)
    n = 400
    ref = rng.normal(scale=10.0, size=(n, 3)).astype(np.float32)
    R = _random_rotation(rng)
    frames = np.asarray([ref @ R.T + 3.0 + rng.normal(scale=0.4, size=(n, 3)) for _ in range(3)],
                        np.float32)
    fit = np.arange(0, n, 10)
    allat = np.arange(n)
    out, out64, rot, cx, cy = oracle.fit_rmsd(frames, fit, allat, ref)
    for f in range(3):
        r, cxx, cyy = np_kabsch(frames[f, fit], ref[fit])
        diff = (frames[f].astype(np.float64) - cxx) @ r.T - (ref.astype(np.float64) - cyy)
        expect = np.sqrt((diff ** 2).sum() / n)
        assert abs(out64[f] - expect) < 1e-10
        np.testing.assert_allclose(rot[f], r, atol=1e-12)
    # fit == rmsd selection reproduces the RmsdPlan numbers
    o2, o2_64, _, _, _ = oracle.fit_rmsd(frames, allat, allat, ref)
    p, p64 = oracle.rmsd_plan(frames, allat, ref, align=True)
    np.testing.assert_allclose(o2_64, p64, rtol=1e-12)
