"""RDF parity of the HIP engine against the CPU oracle (bit-exact counts).

Covers the reference's known answers through the RdfPlan API, seeded random
inputs over every pbc mode / selection shape / enumeration (brute force and
cell list), the full 1,000-water x 100-frame lattice workload, frames of the
100,000-water box (orthorhombic and triclinic) against the oracle's cell-list
enumeration, and size-independent properties of the full 1,000-frame run.
"""
import json
import math
import os
import zlib

import numpy as np
import pytest

import oracle
from warp_md_amd import RdfPlan, System, Trajectory, rdf, workloads

import _gpu

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))


def _system(n, names=None, positions0=None):
    names = names or ["X"] * n
    return System.from_arrays({"name": names}, positions0=positions0)


@pytest.mark.parametrize("case", GOLD["rdf"], ids=lambda c: c["name"])
def test_reference_known_answers_through_plan(case):
    frames = np.asarray(case["frames"], np.float32)
    n = frames.shape[1]
    sysm = _system(n)
    opts = case.get("options", {})
    plan = RdfPlan(case["sel_a"], case["sel_b"], bins=case["bins"], r_max=case["r_max"],
                   pbc=case["pbc"], density=opts.get("density", 0.033456),
                   volume=opts.get("volume", False), raw_rdf=opts.get("raw_rdf", False),
                   intrdf=opts.get("intrdf", False),
                   number_density=opts.get("number_density", False))
    out = plan.run(Trajectory.from_numpy(frames), sysm, device="auto",
                   frame_indices=case.get("frame_indices"))
    counts, g = out[2], out[1]
    if "expect_counts" in case:
        np.testing.assert_array_equal(counts, np.asarray(case["expect_counts"], np.uint64))
    for k, v in case.get("expect_counts_at", {}).items():
        assert counts[int(k)] == v
    tol = case.get("tol", 0.0)
    for k, v in case.get("expect_g_at", {}).items():
        assert abs(g[int(k)] - v) <= tol
    if "expect_g" in case:
        np.testing.assert_allclose(g, case["expect_g"], atol=tol, rtol=0)
    for k, v in case.get("expect_integral_at", {}).items():
        assert abs(out[3][int(k)] - v) <= tol
    if case.get("expect_no_integral"):
        assert len(out) == 3


def test_rdf_wrapper_index_masks():
    coords = np.array([[[0.0, 0.0, 0.0], [0.5, 0.0, 0.0]]], np.float32)
    sysm = System.from_arrays({"name": ["A", "B"], "resname": ["RES", "RES"], "resid": [1, 2],
                               "chain_id": [0, 0], "mass": [1.0, 1.0]}, positions0=coords[0])
    out = rdf(Trajectory.from_numpy(coords), sysm, solvent_mask="@1", solute_mask="@2",
              maximum=2.0, bin_spacing=1.0, image=False, density=1.0, raw_rdf=True, dtype="dict")
    np.testing.assert_array_equal(out["counts"], np.array([1, 0], np.uint64))


def _random_case(kind, n, F, seed, L=24.0):
    rng = np.random.default_rng(seed)
    frames = (rng.random((F, n, 3)) * L).astype(np.float32)
    frames[:, :3] += np.float32(L * 1.2)  # raw coordinates outside the primary box
    frames[:, 3:6] -= np.float32(L * 2.3)
    if kind == "orthorhombic":
        box = np.array([L, L * 1.07, L * 0.93], np.float32)
    elif kind == "triclinic":
        box = np.array([[L, 0, 0], [L * math.tan(math.radians(10)), L, 0], [1.5, -2.0, L]], np.float32)
    else:
        box = None
    return frames, box


SEL_SHAPES = {
    "same": lambda n: (np.arange(n), np.arange(n)),
    "overlap": lambda n: (np.arange(0, 2 * n // 3), np.arange(n // 3, n)),
    "disjoint": lambda n: (np.arange(0, n // 2), np.arange(n // 2, n)),
    "duplicates": lambda n: (np.concatenate([np.arange(n // 2), [1, 1, 7]]), np.arange(n)[::-1]),
}


@pytest.mark.parametrize("kind", ["none", "orthorhombic", "triclinic"])
@pytest.mark.parametrize("shape", list(SEL_SHAPES))
@pytest.mark.parametrize("force_brute", [False, True])
def test_random_parity_exact(kind, shape, force_brute):
    n, F, bins, r_max = 1500, 3, 97, 7.3
    frames, box = _random_case(kind, n, F, seed=zlib.crc32(f"{kind}/{shape}".encode()) % 1000, L=32.0)
    a, b = SEL_SHAPES[shape](n)
    flags = 0x2 if force_brute else 0
    got, _, path = _gpu.rdf_counts(frames, a, b, bins, r_max, pbc=kind, box=box, flags=flags)
    ref = oracle.rdf_counts(frames, a, b, bins, r_max, pbc=kind,
                            boxes=_gpu.boxes9(box, F, kind))
    if kind != "none" and not force_brute:
        assert path == 2, "expected the cell-list path"
    assert ref.sum() > 0
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("kind", ["orthorhombic", "triclinic"])
def test_exact_wrap_small_box(kind):
    # r_max above half the box: the engine must take the IEEE-division wrap
    n, F, bins = 300, 2, 50
    frames, box = _random_case(kind, n, F, seed=4, L=12.0)
    r_max = 9.0
    got, _, path = _gpu.rdf_counts(frames, np.arange(n), np.arange(n), bins, r_max, pbc=kind, box=box)
    ref = oracle.rdf_counts(frames, np.arange(n), np.arange(n), bins, r_max, pbc=kind,
                            boxes=_gpu.boxes9(box, F, kind))
    assert path == 1
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("kind", ["orthorhombic", "triclinic"])
@pytest.mark.parametrize("L,r_max,cell_k,want", [(34.0, 12.0, 2, 5), (40.0, 12.0, 2, 6),
                                                   (42.0, 12.0, 1, 3), (23.5, 9.0, 3, 7)])
def test_small_grid_shift_path(kind, L, r_max, cell_k, want):
    # grids of 2k+1 and 2k+2 cells per side: the stencil spans more than half
    # the box, yet every counted pair takes its image from the cells
    n, F, bins = 3000, 3, 120
    frames, box = _random_case(kind, n, F, seed=int(L * 10 + cell_k), L=L)
    sel = np.arange(n)
    got, _, path, nc = _gpu.rdf_counts(frames, sel, sel, bins, r_max, pbc=kind, box=box,
                                       cell_k=cell_k, cells=True)
    ref = oracle.rdf_counts(frames, sel, sel, bins, r_max, pbc=kind,
                            boxes=_gpu.boxes9(box, F, kind))
    assert path == 2 and min(nc) == want, (path, nc)
    assert ref.sum() > 0
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("cell_k", [1, 2, 3])
def test_cell_shapes_agree(cell_k):
    w = workloads.water_lattice(n_side=14, box=14 * 3.104, n_frames=3, seed=9, with_h=False)
    n = w["coords"].shape[1]
    got, _, path = _gpu.rdf_counts(w["coords"], np.arange(n), np.arange(n), 200, 10.0,
                                   box=w["box"], cell_k=cell_k)
    ref = oracle.rdf_counts(w["coords"], np.arange(n), np.arange(n), 200, 10.0, boxes=w["box"],
                            method="cells")
    assert path == 2
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("cell_k", [1, 2, 3])
def test_pair_evaluation_count(cell_k):
    """wmd_rdf_pair_evaluations under edge diagnostics: counts unchanged, every
    in-range unordered pair evaluated, no more evaluations than unordered
    pairs, zero without diagnostics, identical on a repeat."""
    w = workloads.water_lattice(n_side=16, box=16 * 3.104, n_frames=2, seed=4, with_h=False)
    n = w["coords"].shape[1]
    sel = np.arange(n)
    plain, _, path, e0 = _gpu.rdf_counts(w["coords"], sel, sel, 200, 10.0, box=w["box"],
                                         cell_k=cell_k, evals=True)
    diag, _, _, ev = _gpu.rdf_counts(w["coords"], sel, sel, 200, 10.0, box=w["box"], flags=0x1,
                                     cell_k=cell_k, evals=True)
    _, _, _, ev2 = _gpu.rdf_counts(w["coords"], sel, sel, 200, 10.0, box=w["box"], flags=0x1,
                                   cell_k=cell_k, evals=True)
    assert path == 2 and e0 == 0 and ev == ev2
    np.testing.assert_array_equal(plain, diag)
    in_range = int(diag.sum()) // 2
    print(f"k={cell_k}: {ev} evaluations, {in_range} in range, efficiency {in_range / ev:.3f}")
    assert in_range <= ev <= 2 * n * (n - 1) // 2


def test_lattice_workload_full_exact_with_edges():
    """1,000 rigid waters (3,000 atoms), L = 31.04 A, 100 frames, OW-OW, 200 bins, r_max 10."""
    w = workloads.water_lattice(n_side=10, box=31.04, n_frames=100, seed=0)
    coords, ow = w["coords"], w["ow"]
    got, edge, _ = _gpu.rdf_counts(coords, ow, ow, 200, 10.0, box=w["box"], flags=0x1)
    ref, ref_edge = oracle.rdf_counts(coords, ow, ow, 200, 10.0, boxes=w["box"], edges=True)
    np.testing.assert_array_equal(got, ref)
    assert edge == int(ref_edge[0])
    print(f"edge pairs (f32 within 1e-5 A of an edge): {edge}; f64-distance edge pairs "
          f"{int(ref_edge[1])}; f32-vs-f64 bin mismatches {int(ref_edge[2])}")
    # g(r) through the plan equals the oracle normalisation
    sysm = System.from_arrays({"name": w["names"]})
    plan = RdfPlan(ow, ow, bins=200, r_max=10.0, pbc="orthorhombic")
    r, g, counts = plan.run(Trajectory.from_numpy(coords, box=w["box"]), sysm)
    np.testing.assert_array_equal(counts, ref)
    r_o, g_o, _ = oracle.rdf_finalize(ref, 10.0, 1000, 1000, 1000, 100)
    np.testing.assert_array_equal(r, r_o)
    np.testing.assert_allclose(g, g_o, rtol=1e-6, atol=0)


@pytest.mark.parametrize("tilt", [None, 10.0])
def test_large_box_frames_exact(tilt):
    """Frames of the 100,000-water box (L = 144.3 A, 500 bins, r_max 15) vs the oracle's
    cell-list enumeration of the identical f32 pair expression."""
    w = workloads.water_lattice(n_side=47, box=144.3, n_frames=2, n_waters=100000, seed=1,
                                with_h=False, tilt_deg=tilt)
    n = w["coords"].shape[1]
    kind = "triclinic" if tilt else "orthorhombic"
    got, _, path = _gpu.rdf_counts(w["coords"], np.arange(n), np.arange(n), 500, 15.0, pbc=kind,
                                   box=w["box"])
    ref = oracle.rdf_counts(w["coords"], np.arange(n), np.arange(n), 500, 15.0, pbc=kind,
                            boxes=_gpu.boxes9(w["box"], 2, kind), method="cells")
    assert path == 2
    np.testing.assert_array_equal(got, ref)


def test_full_size_properties():
    """1,000 frames of the 100,000-water box on the device: chunking invariance,
    even totals (A == B) and total-count bounds."""
    import torch

    store, n, stride, M = workloads.water_lattice_device(47, 144.3, 1000, n_waters=100000, seed=1)
    box = np.array([144.3, 144.3, 144.3], np.float32)
    c_all, path = _gpu.rdf_counts_device(store, n, n, 500, 15.0, "orthorhombic", box)
    c_a, _ = _gpu.rdf_counts_device(store[:400].contiguous(), n, n, 500, 15.0, "orthorhombic", box)
    c_b, _ = _gpu.rdf_counts_device(store[400:].contiguous(), n, n, 500, 15.0, "orthorhombic", box)
    assert path == 2
    np.testing.assert_array_equal(c_all, c_a + c_b)
    total = int(c_all.sum())
    assert total % 2 == 0
    # ~ n * rho * 4/3 pi r^3 pairs per frame (ideal-gas count within 5 %)
    expect = 1000 * n * (n / 144.3 ** 3) * 4.0 / 3.0 * math.pi * 15.0 ** 3
    assert abs(total / expect - 1.0) < 0.05
    # frame 7 alone equals the oracle
    one = store[7:8].contiguous()
    c_one, _ = _gpu.rdf_counts_device(one, n, n, 500, 15.0, "orthorhombic", box)
    xyz = one[0, :, :n].T.contiguous().cpu().numpy()[None]
    ref = oracle.rdf_counts(xyz, np.arange(n), np.arange(n), 500, 15.0, boxes=box, method="cells")
    np.testing.assert_array_equal(c_one, ref)
    del store
    torch.cuda.empty_cache()


@pytest.mark.parametrize("kind", ["orthorhombic", "triclinic"])
def test_far_and_non_finite_atoms_exact(kind):
    """Raw coordinates up to 1,000 box lengths away (frames with them take the
    checked wrap) and NaN / inf coordinates (never counted), cell path and
    brute force, against the oracle."""
    n, F, bins, r_max = 1200, 3, 80, 7.3
    frames, box = _random_case(kind, n, F, seed=17, L=32.0)
    L = np.float32(32.0)
    frames[0, 10:20, 0] += np.float32(20) * L      # 20 box lengths
    frames[1, 30:40, 1] -= np.float32(150) * L     # 150 box lengths
    frames[1, 50:55, 2] += np.float32(1000) * L    # 1,000 box lengths
    frames[2, 60, :] = np.nan
    frames[2, 61, 0] = np.inf
    frames[2, 62, 1] = -np.inf
    sel = np.arange(n)
    for force_brute in (False, True):
        got, _, path = _gpu.rdf_counts(frames, sel, sel, bins, r_max, pbc=kind, box=box,
                                       flags=0x2 if force_brute else 0)
        ref = oracle.rdf_counts(frames, sel, sel, bins, r_max, pbc=kind,
                                boxes=_gpu.boxes9(box, F, kind))
        assert path == (1 if force_brute else 2)
        assert ref.sum() > 0
        np.testing.assert_array_equal(got, ref)


def test_empty_and_degenerate_inputs():
    frames = np.zeros((2, 4, 3), np.float32)
    got, _, _ = _gpu.rdf_counts(frames, [], [0, 1], 10, 5.0, pbc="none")
    assert got.sum() == 0
    got, _, _ = _gpu.rdf_counts(frames, [0], [0], 10, 5.0, pbc="none")  # self pair only
    assert got.sum() == 0
    # all atoms on one point: every distinct pair at r = 0 -> bin 0
    got, _, _ = _gpu.rdf_counts(frames, [0, 1, 2, 3], [0, 1, 2, 3], 10, 5.0, pbc="none")
    assert got[0] == 2 * 12 and got.sum() == 24
    sysm = _system(4)
    with pytest.raises(RuntimeError, match="mismatch"):
        RdfPlan([0, 1], [0, 1], bins=4, r_max=4.0).run(Trajectory.from_numpy(frames), sysm)
    with pytest.raises(RuntimeError, match="parse error"):
        RdfPlan([0, 1], [0, 1], bins=4, r_max=4.0, pbc="none", density=-1.0).run(
            Trajectory.from_numpy(frames), sysm)


@pytest.mark.parametrize("bins", [9000, 20000, 65536])
@pytest.mark.parametrize("path", ["cells", "brute"])
def test_histograms_larger_than_lds(bins, path):
    """Bin counts beyond what LDS holds (the reference accepts any bins,
    rdf.rs:141-236): the engine counts them slice by slice; counts and edge
    pairs bit-exact against the oracle, on both enumerations."""
    w = workloads.water_lattice(n_side=10, box=31.04, n_frames=3, seed=5)
    coords, ow = w["coords"], w["ow"]
    flags = 0x1 | (0x2 if path == "brute" else 0)  # edge diagnostics (+ brute force)
    print(f"large histogram bins={bins} path={path}", flush=True)
    got, edge, used = _gpu.rdf_counts(coords, ow, ow, bins, 10.0, box=w["box"], flags=flags)
    ref, ref_edge = oracle.rdf_counts(coords, ow, ow, bins, 10.0, boxes=w["box"], edges=True)
    assert used == (2 if path == "cells" else 1)
    np.testing.assert_array_equal(got, ref)
    assert edge == int(ref_edge[0])
    assert int(got.sum()) > 0


def test_large_histogram_through_plan():
    w = workloads.water_lattice(n_side=10, box=31.04, n_frames=4, seed=6)
    sysm = System.from_arrays({"name": w["names"]})
    r, g, counts = RdfPlan(w["ow"], w["ow"], bins=20000, r_max=10.0).run(
        Trajectory.from_numpy(w["coords"], box=w["box"]), sysm, device="cuda:0")
    ref = oracle.rdf_counts(w["coords"], w["ow"], w["ow"], 20000, 10.0, boxes=w["box"])
    np.testing.assert_array_equal(counts, ref)
    rr, gg, _ = oracle.rdf_finalize(ref, 10.0, 1000, 1000, 1000, 4)
    np.testing.assert_array_equal(r, rr)
    np.testing.assert_array_equal(g, gg)
