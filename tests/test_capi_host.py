"""CPU-side checks of the engine library: it loads, exports every symbol the
header declares, and its host-only routines (RDF normalisation, QCP solver)
agree with the oracle / numpy.  No GPU compute is called here."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from warp_md_amd import Device, RdfPlan, RmsdPlan, System, Trajectory, _native, finalize_rdf
from warp_md_amd.executor import normalize_frame_indices, select_frames

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_symbols_exported():
    header = open(os.path.join(ROOT, "include", "wmd_hip.h")).read()
    declared = set(re.findall(r"\b(wmd_[a-z0-9_]+)\s*\(", header))
    assert declared == set(_native.SYMBOLS)
    L = _native.lib()
    for name in declared:
        assert hasattr(L, name), name


def test_device_count_without_gpu_is_zero_or_more():
    assert _native.device_count() >= 0


@pytest.mark.parametrize("mode", ["rdf", "volume", "number_density", "raw"])
@pytest.mark.parametrize("integral", [False, True])
def test_finalize_matches_oracle(mode, integral):
    rng = np.random.default_rng(8)
    counts = rng.integers(0, 50_000, size=500).astype(np.uint64)
    kw = dict(r_max=15.0, n_a=1e5, n_b=1e5, common=1e5, frames=1000,
              volume_sum=1000 * 144.3 ** 3, density=0.033456, use_volume=(mode == "volume"),
              raw=(mode == "raw"), number_density=(mode == "number_density"), integral=integral)
    a = finalize_rdf(counts, kw["r_max"], kw["n_a"], kw["n_b"], kw["common"], kw["frames"],
                     kw["volume_sum"], kw["density"], kw["use_volume"], kw["raw"],
                     kw["number_density"], integral)
    b = oracle.rdf_finalize(counts, **kw)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_finalize_volume_error():
    with pytest.raises(RuntimeError, match="mismatch: rdf box normalization"):
        finalize_rdf(np.zeros(3, np.uint64), 3.0, 2, 2, 2, 1, 0.0, 1.0, True, False, False, False)


def _qcp(m, gx, gy, n):
    L = _native.lib()
    out = np.zeros(1)
    rot = np.zeros(9)
    lam = np.zeros(1)
    it = ctypes.c_int()
    _native.check(L.wmd_qcp_host(_native.ptr(np.ascontiguousarray(m, np.float64).reshape(9)),
                                 gx, gy, n, _native.ptr(out), _native.ptr(rot), _native.ptr(lam),
                                 ctypes.byref(it)))
    return out[0], rot.reshape(3, 3), lam[0], it.value


def test_qcp_matches_kabsch_and_eigen():
    rng = np.random.default_rng(21)
    for trial in range(300):
        n = int(rng.integers(3, 60))
        y = rng.normal(scale=rng.uniform(0.5, 20), size=(n, 3))
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        from warp_md_amd.workloads import random_rotations

        R0 = random_rotations(rng, 1)[0]
        x = (y @ R0.T) + rng.normal(scale=rng.uniform(0, 2), size=(n, 3))
        if trial % 7 == 0:
            x[:, 2] *= -1  # reflections: best proper rotation
        cx, cy = x.mean(0), y.mean(0)
        M = (x - cx).T @ (y - cy)
        gx = ((x - cx) ** 2).sum()
        gy = ((y - cy) ** 2).sum()
        rmsd, R, lam, it = _qcp(M, gx, gy, n)
        # independent: numpy SVD Kabsch
        u, s, vt = np.linalg.svd(M)
        d = np.sign(np.linalg.det(vt.T @ u.T))
        Rk = vt.T @ np.diag([1, 1, d]) @ u.T
        diff = (x - cx) @ Rk.T - (y - cy)
        expect = np.sqrt((diff ** 2).sum() / n)
        assert abs(rmsd - expect) < 1e-9 * max(1.0, np.sqrt(gx / n))
        np.testing.assert_allclose(R, Rk, atol=1e-7)
        # lambda is the top eigenvalue of the key matrix = sum of (signed) singular values
        assert abs(lam - (s[0] + s[1] + d * s[2])) < 1e-9 * max(1.0, s[0])
        assert it != 0


def test_qcp_degenerate_inputs():
    # identical points: zero moments
    rmsd, R, lam, _ = _qcp(np.zeros((3, 3)), 0.0, 0.0, 4)
    assert rmsd == 0.0 and abs(np.linalg.det(R) - 1) < 1e-12
    # collinear structures (degenerate top eigenvalue) still give a proper rotation
    x = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0.0]])
    y = np.array([[0, 0, 0], [0, 1, 0], [0, 2, 0.0]])
    cx, cy = x.mean(0), y.mean(0)
    M = (x - cx).T @ (y - cy)
    rmsd, R, lam, it = _qcp(M, ((x - cx) ** 2).sum(), ((y - cy) ** 2).sum(), 3)
    assert rmsd < 1e-7
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-10)
    np.testing.assert_allclose((x - cx) @ R.T, y - cy, atol=1e-7)


def test_device_spec_parsing():
    assert Device.from_spec("cpu").kind == "cpu"
    assert Device.from_spec(" AUTO ").kind == "hip"
    assert Device.from_spec("cuda:1").index == 1
    assert Device.from_spec("hip").index == 0
    with pytest.raises(RuntimeError, match="unsupported: unknown device spec"):
        Device.from_spec("tpu")
    with pytest.raises(RuntimeError, match="parse error"):
        Device.from_spec("cuda:x")
    with pytest.raises(RuntimeError, match="unsupported"):
        Device.from_spec("cpu").context()


def test_frame_index_semantics():
    assert normalize_frame_indices([1, -1, 99], 2) == [1, 1]
    coords = np.arange(4 * 2 * 3, dtype=np.float32).reshape(4, 2, 3)
    t = select_frames(Trajectory.from_numpy(coords), [3, 0, 3, 7])
    np.testing.assert_array_equal(t.coords[:, 0, 0], [18, 0, 18])


def test_plan_argument_errors():
    with pytest.raises(RuntimeError, match="pbc must be"):
        RdfPlan([0], [1], bins=2, r_max=1.0, pbc="cubic")
    with pytest.raises(ValueError, match="dimension"):
        RdfPlan([0], [1], bins=2, r_max=1.0, dimension="4d")
    with pytest.raises(RuntimeError, match="reference must be"):
        RmsdPlan([0], reference="frame9")
    sysm = System.from_arrays({"name": ["A", "B"]})
    tr = Trajectory.from_numpy(np.zeros((1, 3, 3), np.float32))
    with pytest.raises(RuntimeError, match="mismatch: trajectory atom count"):
        RdfPlan([0], [1], bins=2, r_max=1.0).run(tr, sysm, device="cpu")
    tr2 = Trajectory.from_numpy(np.zeros((1, 2, 3), np.float32))
    with pytest.raises(RuntimeError, match="unsupported"):
        RdfPlan([0], [1], bins=2, r_max=1.0, pbc="none").run(tr2, sysm, device="cpu")
    with pytest.raises(RuntimeError, match="unsupported: grouped and xy"):
        RdfPlan([0], [1], bins=2, r_max=1.0, pbc="triclinic", center1=True).run(tr2, sysm)


def test_system_selection_subset():
    sysm = System.from_arrays({"name": ["OW", "HW1", "HW2", "OW"], "resid": [1, 1, 1, 2]})
    np.testing.assert_array_equal(sysm.select("name OW").indices, [0, 3])
    np.testing.assert_array_equal(sysm.select("name OW HW1 and resid 1").indices, [0, 1])
    np.testing.assert_array_equal(sysm.select("index 1-2").indices, [1, 2])
    with pytest.raises(RuntimeError, match="invalid selection"):
        sysm.select("within 5 of name OW")


def test_autotune_cache_text_round_trip():
    """The tuning cache's text form: nine fields per line (k, home block, units
    per task after the key); seven-field lines keep the defaults; malformed or
    out-of-range lines are refused without changing the cache."""
    from warp_md_amd.batch import export_tuning, import_tuning

    import_tuning("gfx950 1 300 1094713344 60 -14 2 0 4\n\ngfx950 1 300 1094713344 61 -14 -1\n")
    lines = [l.split() for l in export_tuning().splitlines()]
    mine = {tuple(l[:6]): l[6:] for l in lines}
    assert mine[("gfx950", "1", "300", "1094713344", "60", "-14")] == ["2", "0", "4"]
    assert mine[("gfx950", "1", "300", "1094713344", "61", "-14")] == ["-1", "0", "0"]
    before = export_tuning()
    for bad in ["gfx950 1 300 1094713344 62 -14 -1 128",      # eight fields
                "gfx950 1 300 1094713344 62 -14 4 0 0",       # k out of range
                "gfx950 1 300 1094713344 62 -14 -1 96 0",     # home block not a multiple of 64
                "gfx950 1 300 1094713344 62 -14 -1 64 65",    # task size out of range
                "gfx950 1 300 1094713344 62 -14 -1 64 2 9"]:  # trailing field
        with pytest.raises(RuntimeError, match="parse"):
            import_tuning(bad)
    assert export_tuning() == before
