"""AlignPlan / SuperposePlan on the GPU (wmd_align_*) against the oracle's
restatement of the reference CPU path (align/plan_exec.rs:215-243, 422-456):
R within 1e-9 and t within 1e-6 A of the f64 values (QCP and SVD agree to
rounding), the f32 rows within 1e-5, superposed coordinates within 1e-5 A
or one f32 ulp.  At the full chain size (5,000 frames x 20,000 atoms, fit on
2,000 CA) a sample of frames is compared the same way."""
import ctypes

import numpy as np
import pytest

import oracle
from warp_md_amd import AlignPlan, SuperposePlan, System, Trajectory
from warp_md_amd import _native, workloads
from warp_md_amd._native import check, lib, ptr

pytestmark = pytest.mark.gpu


def _close_xyz(got, want):
    d = np.abs(got.astype(np.float64) - want.astype(np.float64))
    ulp = np.spacing(np.abs(want)).astype(np.float64)
    return bool(np.all(d <= np.maximum(1e-5, 1.01 * ulp)))


def _system(n, positions0=None, masses=None):
    atoms = {"name": np.array(["CA"] * n)}
    if masses is not None:
        atoms["mass"] = masses
    return System.from_arrays(atoms, positions0)


def _chain(F, seed):
    c = workloads.chain(n_ca=60, per_ca=5, n_frames=F, seed=seed)
    return c["ref"], c["coords"], c["ca"]


@pytest.mark.parametrize("mass", [False, True])
@pytest.mark.parametrize("norotate", [False, True])
@pytest.mark.parametrize("reference", ["topology", "frame0"])
def test_align_matches_oracle(mass, norotate, reference):
    ref, c, ca = _chain(23, seed=1)
    n = c.shape[1]
    masses = np.random.default_rng(2).uniform(1, 16, n).astype(np.float32) if mass else None
    sysm = _system(n, positions0=ref, masses=masses)
    traj = Trajectory.from_numpy(c)
    print(f"align mass={mass} norotate={norotate} reference={reference}", flush=True)
    rt = AlignPlan(ca, reference=reference, mass=mass, norotate=norotate).run(
        traj, sysm, chunk_frames=7, device="cuda:0")
    xyz = SuperposePlan(ca, reference=reference, mass=mass, norotate=norotate).run(
        traj, sysm, chunk_frames=7, device="cuda:0")
    ref_all = ref if reference == "topology" else c[0]
    w = None if masses is None else masses[ca]
    ort, ort64, oxyz = oracle.align(c, ca, ref_all, weights=w, norotate=norotate, superpose=True)
    assert rt.shape == (23, 12) and xyz.shape == (23 * n * 3,)
    assert np.max(np.abs(rt[:, :9].astype(np.float64) - ort64[:, :9])) <= 1e-6
    assert np.max(np.abs(rt[:, 9:].astype(np.float64) - ort64[:, 9:])) <= 1e-5
    assert _close_xyz(xyz.reshape(23, n, 3), oxyz)


def test_device_rt64_matches_oracle_f64():
    import torch

    ref, c, ca = _chain(40, seed=3)
    n = c.shape[1]
    ref4 = np.ones((n, 4), np.float32)
    ref4[:, :3] = ref
    ctx = _native.context(0)
    h = ctypes.c_void_p()
    check(lib().wmd_align_create(ctx.handle, ptr(ca), ca.size, None, ptr(ref4), n, 0, ctypes.byref(h)))
    try:
        x4 = np.ones((40, n, 4), np.float32)
        x4[..., :3] = c
        d = torch.as_tensor(x4).cuda()
        rt64 = torch.zeros((40, 12), dtype=torch.float64, device="cuda:0")
        check(lib().wmd_align_run_device(h, ctypes.c_void_p(d.data_ptr()), n, 40,
                                         ctypes.c_void_p(rt64.data_ptr()), None, None, None))
        torch.cuda.synchronize()
        got = rt64.cpu().numpy()
    finally:
        lib().wmd_align_destroy(h)
    _, ort64, _ = oracle.align(c, ca, ref)
    assert np.max(np.abs(got[:, :9] - ort64[:, :9])) <= 1e-9
    assert np.max(np.abs(got[:, 9:] - ort64[:, 9:])) <= 1e-6


def test_reference_known_answers():
    # analysis_smoke_tests.rs:171-195
    sysm = _system(2, positions0=np.array([[0, 0, 0], [1, 0, 0]], np.float32))
    c = np.array([[[1, 1, 0], [1, 2, 0]]], np.float32)
    out = SuperposePlan([0, 1]).run(Trajectory.from_numpy(c), sysm, device="cuda:0")
    assert out.shape == (6,)
    assert np.linalg.norm(out[:3] - [0, 0, 0]) < 1e-5 and np.linalg.norm(out[3:] - [1, 0, 0]) < 1e-5
    # test_superpose.py:41-67 (frame 0 as the reference)
    ref = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64)
    rot = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], np.float64)
    c = np.stack([ref, ref @ rot.T + [2.0, 3.0, 0.0]]).astype(np.float32)
    out = SuperposePlan([0, 1, 2], reference="frame0").run(Trajectory.from_numpy(c), _system(3),
                                                          device="cuda:0")
    out = out.reshape(2, 3, 3)
    assert np.allclose(out[0], ref, atol=1e-5) and np.allclose(out[1], ref, atol=1e-5)
    with pytest.raises(RuntimeError, match="no topology reference"):
        AlignPlan([0, 1, 2]).run(Trajectory.from_numpy(c), _system(3), device="cuda:0")


def _device_superpose(c, sel, ref, masses=None, norotate=False):
    """wmd_align_run_device with every atom transformed: (rt64 [F][12], xyz [F][N][3])."""
    import torch

    F, n, _ = c.shape
    ref4 = np.ones((n, 4), np.float32)
    ref4[:, :3] = ref
    x4 = np.ones((F, n, 4), np.float32)
    x4[..., :3] = c
    sel = np.ascontiguousarray(sel, np.uint32)
    m = None if masses is None else np.ascontiguousarray(masses, np.float32)
    ctx = _native.context(0)
    h = ctypes.c_void_p()
    check(lib().wmd_align_create(ctx.handle, ptr(sel), sel.size, ptr(m), ptr(ref4), n, int(norotate),
                                 ctypes.byref(h)))
    try:
        d = torch.as_tensor(x4).cuda()
        rt64 = torch.zeros((F, 12), dtype=torch.float64, device="cuda:0")
        xyz = torch.zeros((F, n, 3), dtype=torch.float32, device="cuda:0")
        check(lib().wmd_align_run_device(h, ctypes.c_void_p(d.data_ptr()), n, F,
                                         ctypes.c_void_p(rt64.data_ptr()), None,
                                         ctypes.c_void_p(xyz.data_ptr()), None))
        torch.cuda.synchronize()
        return rt64.cpu().numpy(), xyz.cpu().numpy()
    finally:
        lib().wmd_align_destroy(h)


@pytest.mark.parametrize("n", [4097, 12288, 12289, 20000, 27648, 27649])
def test_superpose_frame_sizes(n):
    """Frames of 4,097 to 27,649 atoms with masses, a selection that repeats
    atoms and spans the frame, and norotate, against the oracle."""
    rng = np.random.default_rng(n)
    F = 5
    ref = rng.normal(scale=15.0, size=(n, 3)).astype(np.float32)
    c = np.empty((F, n, 3), np.float32)
    for f in range(F):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, a, b, d = q
        R = np.array([[1 - 2 * (b * b + d * d), 2 * (a * b - d * w), 2 * (a * d + b * w)],
                      [2 * (a * b + d * w), 1 - 2 * (a * a + d * d), 2 * (b * d - a * w)],
                      [2 * (a * d - b * w), 2 * (b * d + a * w), 1 - 2 * (a * a + b * b)]])
        c[f] = (ref @ R.T + rng.uniform(-20, 20, 3) + rng.normal(scale=0.5, size=(n, 3))).astype(np.float32)
    sel = np.concatenate([np.arange(3, n, 7), [5, 5, n - 1, n - 1, n // 2]])
    masses = rng.uniform(1, 16, sel.size).astype(np.float32)
    print(f"superpose n={n}", flush=True)
    for m, norot in ((None, False), (masses, False), (None, True)):
        rt64, xyz = _device_superpose(c, sel, ref, masses=m, norotate=norot)
        _, ort64, oxyz = oracle.align(c, sel, ref, weights=m, norotate=norot, superpose=True)
        assert np.max(np.abs(rt64[:, :9] - ort64[:, :9])) <= 1e-9
        assert np.max(np.abs(rt64[:, 9:] - ort64[:, 9:])) <= 1e-6
        assert _close_xyz(xyz, oxyz)


def test_full_chain_superpose():
    """5,000 frames x 20,000 atoms, fit on 2,000 CA: every atom superposed."""
    import torch

    c = workloads.chain(n_frames=0, seed=2)
    ref, ca = c["ref"], c["ca"]
    n = ref.shape[0]
    F = 5000
    atoms = np.arange(n, dtype=np.uint32)
    store = workloads.chain_store_device(ref, atoms, F, seed=2, device="cuda:0")  # (F, 3, n)
    x4 = torch.ones((F, n, 4), dtype=torch.float32, device="cuda:0")
    x4[:, :, :3] = store.permute(0, 2, 1)
    del store
    ref4 = np.ones((n, 4), np.float32)
    ref4[:, :3] = ref
    ctx = _native.context(0)
    h = ctypes.c_void_p()
    check(lib().wmd_align_create(ctx.handle, ptr(ca), ca.size, None, ptr(ref4), n, 0, ctypes.byref(h)))
    try:
        rt64 = torch.zeros((F, 12), dtype=torch.float64, device="cuda:0")
        xyz = torch.zeros((F, n, 3), dtype=torch.float32, device="cuda:0")
        check(lib().wmd_align_run_device(h, ctypes.c_void_p(x4.data_ptr()), n, F,
                                         ctypes.c_void_p(rt64.data_ptr()), None,
                                         ctypes.c_void_p(xyz.data_ptr()), None))
        torch.cuda.synchronize()
    finally:
        lib().wmd_align_destroy(h)
    frames = [0, 1, 2499, 4999]
    sample = x4[frames, :, :3].cpu().numpy()
    _, ort64, oxyz = oracle.align(sample, ca, ref, superpose=True)
    got = xyz[frames].cpu().numpy()
    assert np.max(np.abs(rt64[frames].cpu().numpy() - ort64)) <= 1e-6
    assert _close_xyz(got, oxyz)
    # superposed frames sit on the reference up to the noise (sigma 0.5 A per coordinate)
    d = np.sqrt(np.mean(np.sum((got - ref[None]) ** 2, axis=2)))
    print(f"5000 x 20000 superpose: sample residual {d:.3f} A", flush=True)
    assert 0.6 < d < 1.2
