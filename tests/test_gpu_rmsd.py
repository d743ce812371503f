"""RMSD parity of the HIP engine against the FP64 CPU oracle.

Bar: |RMSD_gpu - RMSD_oracle| <= 1e-5 A for the f32 outputs, and the engine's
FP64 value within its own reported error bound (plus 1e-9 A) of the oracle's
FP64 value, for every frame.  Cases: the reference's known answers through
RmsdPlan, RmsdPlan (same selection, topology / frame0 references, raw mode,
duplicates), fit-on-CA / RMSD-over-all-atoms for the 20,000-atom chain, and the
full 5,000-frame chain on the device (oracle on a frame sample + chunking
invariance over all frames).
"""
import json
import os

import numpy as np
import pytest

import oracle
from warp_md_amd import RmsdPlan, System, Trajectory, workloads

import _gpu

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))
TOL = 1e-5


@pytest.mark.parametrize("case", GOLD["rmsd"], ids=lambda c: c["name"])
def test_reference_known_answers_through_plan(case):
    frames = np.asarray(case["frames"], np.float32)
    ref = np.asarray(case["reference"], np.float32)
    n = frames.shape[1]
    sysm = System.from_arrays({"name": ["CA"] * n}, positions0=ref)
    out = RmsdPlan(case["sel"], reference="topology", align=case["align"]).run(
        Trajectory.from_numpy(frames), sysm)
    for v, lim in zip(out, case["expect_abs_below"]):
        assert abs(v) < lim
    if "expect_rotation_frame1" in case:
        _, _, _, rot = _gpu.rmsd_device(frames, np.asarray(case["sel"]), None, ref)
        np.testing.assert_allclose(rot[1], case["expect_rotation_frame1"], atol=1e-6)
        np.testing.assert_allclose(rot[0], np.eye(3), atol=1e-6)


def _check(gpu32, gpu_err, ref32, ref64, gpu64=None):
    d = np.abs(gpu32.astype(np.float64) - ref32.astype(np.float64))
    assert d.max() <= TOL, f"max |dRMSD| {d.max():.3e} A"
    if gpu64 is not None:
        d64 = np.abs(gpu64 - ref64)
        assert np.all(d64 <= gpu_err + 1e-9), f"f64 error {d64.max():.3e} beyond bound"
    print(f"max |dRMSD| f32 {d.max():.3e} A" +
          ("" if gpu64 is None else f", f64 {np.abs(gpu64 - ref64).max():.3e} A, "
                                    f"bound max {gpu_err.max():.3e} A"))


def _chain_small(n_frames=40, seed=12):
    return workloads.chain(n_ca=300, per_ca=10, n_frames=n_frames, seed=seed)


def test_rmsdplan_same_selection_topology():
    c = _chain_small()
    sel = np.arange(c["ref"].shape[0])
    sysm = System.from_arrays({"name": c["names"]}, positions0=c["ref"])
    plan = RmsdPlan(sel, reference="topology", align=True)
    out = plan.run(Trajectory.from_numpy(c["coords"]), sysm)
    ref32, ref64 = oracle.rmsd_plan(c["coords"], sel, c["ref"][sel], align=True)
    _check(out, plan.last_error_bound, ref32, ref64)


def test_rmsdplan_frame0_subset_and_raw():
    c = _chain_small()
    sel = c["ca"]
    sysm = System.from_arrays({"name": c["names"]}, positions0=c["ref"])
    out = RmsdPlan(sel, reference="frame0", align=True).run(Trajectory.from_numpy(c["coords"]), sysm)
    ref32, ref64 = oracle.rmsd_plan(c["coords"], sel, c["coords"][0][sel], align=True)
    assert abs(out[0]) < 1e-5
    _check(out, np.zeros_like(ref64), ref32, ref64)
    raw = RmsdPlan(sel, reference="topology", align=False).run(Trajectory.from_numpy(c["coords"]), sysm)
    r32, r64 = oracle.rmsd_plan(c["coords"], sel, c["ref"][sel], align=False)
    _check(raw, np.zeros_like(r64), r32, r64)


def test_rmsdplan_duplicates_and_frame_indices():
    c = _chain_small(n_frames=10)
    sel = np.concatenate([c["ca"][:50], c["ca"][:5], [3, 3]])
    sysm = System.from_arrays({"name": c["names"]}, positions0=c["ref"])
    out = RmsdPlan(sel).run(Trajectory.from_numpy(c["coords"]), sysm, frame_indices=[4, -1, 4, 99])
    frames = c["coords"][[4, 9, 4]]
    ref32, ref64 = oracle.rmsd_plan(frames, sel, c["ref"][sel], align=True)
    assert out.shape == (3,)
    _check(out, np.zeros_like(ref64), ref32, ref64)


def test_fit_on_ca_rmsd_over_all_chain():
    """The 20,000-atom chain (2,000 CA), 60 frames: fit on CA, RMSD over all atoms."""
    c = workloads.chain(n_frames=60, seed=2)
    ca = c["ca"]
    allat = np.arange(c["ref"].shape[0])
    g32, gerr = _gpu.rmsd(c["coords"], ca, allat, c["ref"])
    o32, o64, rot, _, _ = oracle.fit_rmsd(c["coords"], ca, allat, c["ref"])
    _check(g32, gerr, o32, o64)
    d32, d64, derr, drot = _gpu.rmsd_device(c["coords"], ca, allat, c["ref"])
    _check(d32, derr, o32, o64, gpu64=d64)
    np.testing.assert_allclose(drot, rot, atol=1e-9)
    # RmsdPlan API with fit_selection
    sysm = System.from_arrays({"name": c["names"]}, positions0=c["ref"])
    p = RmsdPlan(allat, fit_selection=ca).run(Trajectory.from_numpy(c["coords"]), sysm)
    _check(p, gerr, o32, o64)


def test_same_selection_device_f64_within_bound():
    c = workloads.chain(n_frames=30, seed=5)
    allat = np.arange(c["ref"].shape[0])
    d32, d64, derr, drot = _gpu.rmsd_device(c["coords"], allat, None, c["ref"])
    o32, o64 = oracle.rmsd_plan(c["coords"], allat, c["ref"], align=True)
    _check(d32, derr, o32, o64, gpu64=d64)


def test_full_chain_workload_on_device():
    """All 5,000 frames x 20,000 atoms resident on the GPU (generated in the plan's
    store layout): oracle on every 25th frame, and chunking invariance (bitwise)
    over all frames."""
    import torch

    c = workloads.chain(n_frames=0, seed=2)
    ca = c["ca"]
    n = c["ref"].shape[0]
    allat = np.arange(n)
    h = _gpu.rmsd_plan_handle(ca, allat, c["ref"])
    try:
        atoms = _gpu.rmsd_layout(h)
        store = workloads.chain_store_device(c["ref"], atoms, 5000, seed=2)
        d32, d64, derr, _ = _gpu.rmsd_run_device(h, store)
        e32, e64, _, _ = _gpu.rmsd_run_device(h, store[1234:1300].contiguous())
        np.testing.assert_array_equal(d64[1234:1300], e64)
        sample = np.arange(0, 5000, 25)
        host = np.stack([store[int(f)].cpu().numpy() for f in sample])
    finally:
        lib_destroy(h)
    frames = workloads.store_to_frames(host, atoms, n)
    o32, o64, _, _, _ = oracle.fit_rmsd(frames, ca, allat, c["ref"])
    _check(d32[sample], derr[sample], o32, o64, gpu64=d64[sample])
    assert np.all(np.isfinite(d64)) and np.all(d64 > 0.3) and np.all(d64 < 1.5)
    del store
    torch.cuda.empty_cache()


def lib_destroy(h):
    from warp_md_amd._native import lib

    lib().wmd_rmsd_destroy(h)


def test_empty_selection_gives_zero():
    frames = np.random.default_rng(0).normal(size=(3, 5, 3)).astype(np.float32)
    g, _ = _gpu.rmsd(frames, np.zeros(0, np.uint32), None, frames[0])
    np.testing.assert_array_equal(g, np.zeros(3, np.float32))
