"""PairwiseRmsdPlan on the GPU (wmd_pairwise_*) against the oracle's
restatement of rmsd/plan.rs:586-675: the f64 value before the cast within
1e-9 A of the oracle's (same f64 arithmetic up to summation order), the f32
matrix within 1e-5 A, symmetric with a zero diagonal.  At the full
5,000-frame chain size: symmetry, zero diagonal and a sample of pairs against
the oracle's RmsdPlan per pair."""
import ctypes

import numpy as np
import pytest

import oracle
from warp_md_amd import PairwiseRmsdPlan, System, Trajectory
from warp_md_amd import _native, workloads
from warp_md_amd._native import check, lib, ptr

pytestmark = pytest.mark.gpu


def _chainish(F, n, seed, noise=0.5):
    c = workloads.chain(n_ca=max(2, n // 2), per_ca=2, n_frames=F, seed=seed, noise=noise)
    return c["coords"][:, :n]


def _device_run(coords, sel, metric):
    import torch

    ctx = _native.context(0)
    h = ctypes.c_void_p()
    check(lib().wmd_pairwise_create(ctx.handle, ptr(sel), sel.size,
                                    {"rms": 0, "nofit": 1, "dme": 2}[metric], ctypes.byref(h)))
    try:
        F = coords.shape[0]
        stride = (sel.size + 3) // 4 * 4
        store = torch.zeros((F, 3, stride), dtype=torch.float32, device="cuda:0")
        store[:, :, :sel.size] = torch.as_tensor(coords[:, sel, :]).cuda().permute(0, 2, 1)
        out = torch.empty((F, F), dtype=torch.float32, device="cuda:0")
        out64 = torch.empty((F, F), dtype=torch.float64, device="cuda:0")
        stream = torch.cuda.current_stream().cuda_stream
        check(lib().wmd_pairwise_run_device(h, ctypes.c_void_p(store.data_ptr()), stride, F,
                                            ctypes.c_void_p(out.data_ptr()),
                                            ctypes.c_void_p(out64.data_ptr()), ctypes.c_void_p(stream)))
        torch.cuda.synchronize()
        return out.cpu().numpy(), out64.cpu().numpy()
    finally:
        lib().wmd_pairwise_destroy(h)


@pytest.mark.parametrize("metric", ["rms", "nofit"])
@pytest.mark.parametrize("F,n", [(5, 37), (70, 61), (33, 200), (96, 1000)])
def test_matches_oracle(metric, F, n):
    c = _chainish(F, n + 3, seed=F + n)
    sel = np.arange(2, n + 2, dtype=np.uint32)
    print(f"pairwise {metric} F={F} n={n}", flush=True)
    out, out64 = _device_run(c, sel, metric)
    ref, ref64 = oracle.pairwise_rmsd(c, sel, metric)
    assert np.max(np.abs(out64 - ref64)) <= 1e-9
    assert np.max(np.abs(out.astype(np.float64) - ref64)) <= 1e-5
    assert np.array_equal(out, out.T) and np.all(np.diag(out) == 0)
    # the cast of the same f64: equal except where the two f64 straddle a rounding point
    assert np.mean(out == ref) > 0.99


@pytest.mark.parametrize("metric", ["rms", "nofit"])
def test_plan_run_matches_oracle(metric):
    c = _chainish(45, 120, seed=9)
    sel = np.arange(5, 117, 2)
    plan = PairwiseRmsdPlan(sel, metric=metric)
    sysm = System.from_arrays({"name": np.array(["X"] * c.shape[1])})
    out = plan.run(Trajectory.from_numpy(c), sysm, chunk_frames=16, device="cuda:0")
    ref, _ = oracle.pairwise_rmsd(c, sel, metric)
    assert out.shape == (45, 45)
    assert np.max(np.abs(out.astype(np.float64) - ref)) <= 1e-5


def test_reference_known_answer_and_edges():
    # pairwise_rmsd_plan_basic (geometry_matrix_tests.rs:1438-1460)
    c = np.array([[[0, 0, 0], [1, 0, 0]], [[0, 0, 0], [2, 0, 0]]], np.float32)
    sysm = System.from_arrays({"name": np.array(["A", "B"])})
    out = PairwiseRmsdPlan([0, 1], metric="nofit").run(Trajectory.from_numpy(c), sysm,
                                                       device="cuda:0")
    assert abs(out[0, 1] - np.sqrt(0.5)) < 1e-6 and abs(out[1, 0] - np.sqrt(0.5)) < 1e-6
    # one atom: zero matrix (plan.rs:597-603); one frame: [[0]]
    out = PairwiseRmsdPlan([1], metric="rms").run(Trajectory.from_numpy(c), sysm, device="cuda:0")
    assert out.shape == (2, 2) and np.all(out == 0)
    out = PairwiseRmsdPlan([0, 1]).run(Trajectory.from_numpy(c[:1]), sysm, device="cuda:0")
    assert out.shape == (1, 1) and out[0, 0] == 0
    with pytest.raises(RuntimeError, match="metric must be"):
        PairwiseRmsdPlan([0, 1], metric="xyz")


def test_full_chain_5000_frames():
    """5,000 frames of the 20,000-atom chain, fit on its 2,000 CA atoms."""
    import torch

    c = workloads.chain(n_frames=0, seed=2)
    ca = c["ca"]
    F = 5000
    store = workloads.chain_store_device(c["ref"], ca, F, seed=2, device="cuda:0")
    ctx = _native.context(0)
    sel = np.arange(ca.size, dtype=np.uint32)
    h = ctypes.c_void_p()
    check(lib().wmd_pairwise_create(ctx.handle, ptr(sel), sel.size, 0, ctypes.byref(h)))
    try:
        out = torch.empty((F, F), dtype=torch.float32, device="cuda:0")
        check(lib().wmd_pairwise_run_device(h, ctypes.c_void_p(store.data_ptr()), store.shape[2], F,
                                            ctypes.c_void_p(out.data_ptr()), None, None))
        torch.cuda.synchronize()
    finally:
        lib().wmd_pairwise_destroy(h)
    m = out.cpu().numpy()
    assert np.array_equal(m, m.T) and np.all(np.diag(m) == 0)
    frames = np.transpose(store.cpu().numpy(), (0, 2, 1))  # (F, 2000, 3)
    rng = np.random.default_rng(0)
    js = rng.choice(F, 4, replace=False)
    worst = 0.0
    for j in js:
        _, r64 = oracle.rmsd_plan(frames, np.arange(ca.size), frames[j], align=True)
        d = np.abs(m[:, j].astype(np.float64) - r64)
        d[j] = 0.0
        worst = max(worst, float(d.max()))
    print(f"5000 x 5000: max |gpu - oracle| over 4 columns {worst:.2e} A", flush=True)
    assert worst <= 1e-5


# ---- metric "dme" (plan.rs:612-650): the reference's k-ordered f64 sums ----

@pytest.mark.parametrize("F,n", [(2, 2), (5, 37), (70, 61), (130, 200)])
def test_dme_matches_oracle_bit_for_bit(F, n):
    c = _chainish(F, n + 3, seed=3 * F + n)
    sel = np.arange(2, n + 2, dtype=np.uint32)[::-1].copy()  # selection order matters for k
    print(f"pairwise dme F={F} n={n}", flush=True)
    out, out64 = _device_run(c, sel, "dme")
    ref, ref64 = oracle.pairwise_dme(c, sel)
    np.testing.assert_array_equal(out64, ref64)
    np.testing.assert_array_equal(out, ref)


def _dme_plan(c, sel, pbc="none", box=None, chunk=None):
    sysm = System.from_arrays({"name": np.array(["X"] * c.shape[1])})
    return PairwiseRmsdPlan(sel, metric="dme", pbc=pbc).run(Trajectory.from_numpy(c, box=box), sysm,
                                                            chunk_frames=chunk, device="cuda:0")


def test_dme_plan_orthorhombic_matches_oracle():
    rng = np.random.default_rng(31)
    F, n, L = 40, 90, 12.0
    c = rng.uniform(0, L, size=(F, n, 3)).astype(np.float32)
    box = np.stack([np.array([L, L * 1.1, L * 0.9], np.float32) * np.float32(1 + 0.01 * f) for f in range(F)])
    sel = np.arange(3, 88, dtype=np.uint32)
    print("pairwise dme plan orthorhombic", flush=True)
    out = _dme_plan(c, sel, "orthorhombic", box, chunk=7)
    ref, ref64 = oracle.pairwise_dme(c, sel, pbc="orthorhombic", boxes=box)
    np.testing.assert_array_equal(out, ref)
    plain = _dme_plan(c, sel, chunk=7)
    assert np.max(np.abs(plain - out)) > 0.1  # the wrap matters for this box
    np.testing.assert_array_equal(plain, oracle.pairwise_dme(c, sel)[0])


def test_dme_errors():
    c = np.zeros((2, 3, 3), np.float32)
    c[1, 1, 0] = 1.0
    sysm = System.from_arrays({"name": np.array(["X"] * 3)})
    with pytest.raises(RuntimeError, match="orthorhombic box required"):
        PairwiseRmsdPlan([0, 1, 2], metric="dme", pbc="orthorhombic").run(Trajectory.from_numpy(c), sysm,
                                                                          device="cuda:0")
    tri = np.array([[10, 0, 0], [1, 10, 0], [0, 0, 10]], np.float32)
    with pytest.raises(RuntimeError, match="orthorhombic box required"):
        PairwiseRmsdPlan([0, 1, 2], metric="dme", pbc="orthorhombic").run(
            Trajectory.from_numpy(c, box=tri), sysm, device="cuda:0")
    with pytest.raises(RuntimeError, match="pbc must be"):
        PairwiseRmsdPlan([0, 1, 2], metric="dme", pbc="triclinic")
    # one atom: the zero matrix (plan.rs:597-603)
    assert not _dme_plan(c, [1]).any()


def test_dme_many_pair_blocks():
    """2,048 frames x 600 atoms (179,700 pair distances in 11 blocks of 16,384):
    a sample of frame pairs against the oracle, bit for bit."""
    F, n = 2048, 600
    c = _chainish(F, n, seed=41, noise=0.8)
    sel = np.arange(n, dtype=np.uint32)
    print("pairwise dme 2048 x 600", flush=True)
    out, out64 = _device_run(c, sel, "dme")
    assert np.array_equal(out, out.T) and np.all(np.diag(out) == 0)
    rng = np.random.default_rng(42)
    for _ in range(12):
        i, j = rng.choice(F, 2, replace=False)
        _, r64 = oracle.pairwise_dme(c[[i, j]], sel)
        assert out64[i, j] == r64[0, 1], (i, j)
