"""Uneven batches on the GPU (wmd_batch_*): every trajectory's counts must equal
the oracle's (the reference pair loop, rdf.rs:882-941) bit for bit and its
RMSD-to-frame-0 series the oracle's RmsdPlan(frame0, align=True)
(rmsd/plan.rs:336-392, 1161-1248) within 1e-5 A (f32) and within the engine's
own error bound (f64); graph replays must equal eager launches; the
autotuner must not change results.  At the full 256-trajectory size the
per-trajectory results are compared with the single-trajectory engine paths,
which the other GPU tests pin to the oracle."""
import math

import numpy as np
import pytest

import oracle
from warp_md_amd.batch import UnevenBatch, export_tuning, import_tuning

pytestmark = pytest.mark.gpu

BINS, RMAX = 300, 12.0


def _case(sizes, n_frames=3, density=0.1, seed=0, spread=0.0, triclinic=False):
    """Uniform positions (optionally spread beyond the box by `spread` x L)."""
    rng = np.random.default_rng(seed)
    coords, boxes = [], []
    for n in sizes:
        L = (n / density) ** (1.0 / 3.0)
        c = rng.uniform(-spread * L, (1.0 + spread) * L, size=(n_frames, n, 3)).astype(np.float32)
        if triclinic:
            M = np.array([[L, 0, 0], [L * math.tan(math.radians(10.0)), L, 0], [0, 0, L]], np.float32)
            frac = rng.random((n_frames, n, 3))
            c = (frac @ M.astype(np.float64)).astype(np.float32)
            boxes.append(M)
        else:
            boxes.append(np.array([L, L, L], np.float32))
        coords.append(c)
    return coords, boxes


def _oracle_counts(c, box, pbc):
    n = c.shape[1]
    allat = np.arange(n)
    method = "omp" if n > 3000 else "brute"
    return oracle.rdf_counts(c, allat, allat, BINS, RMAX, pbc=pbc, boxes=box, method=method)


def _batch(coords, boxes, pbc="orthorhombic", graph=True, rmsd=True):
    b = UnevenBatch([c.shape[1] for c in coords], [c.shape[0] for c in coords], bins=BINS,
                    r_max=RMAX, pbc=pbc, rmsd=rmsd, graph=graph, device="cuda:0")
    b.set_boxes(boxes)
    store = b.store_from_numpy(coords)
    return b, store


# sizes chosen to cover every enumeration: n < 64 (brute), r_max > L/2 (brute,
# checked roundf wrap), L/2 > r_max but too small a box for cells (brute),
# cells with the rint image mode (nc < 2k + 3), cells with image shifts
MIXED = [40, 300, 700, 1500, 3000, 8000, 1, 20000]


@pytest.mark.parametrize("graph", [True, False])
def test_mixed_batch_matches_oracle(graph):
    coords, boxes = _case(MIXED, seed=1)
    b, store = _batch(coords, boxes, graph=graph)
    kinds, units = b.plan()
    print("plan", list(kinds), units, flush=True)
    assert kinds[0] == 0 and kinds[1] == 0 and kinds[-1] > 0
    r = b.run(store)
    counts = b.counts()
    rm = r.cpu().numpy()
    for t, (c, box) in enumerate(zip(coords, boxes)):
        ref = _oracle_counts(c, box, "orthorhombic")
        assert np.array_equal(counts[t], ref), f"trajectory {t} (n={c.shape[1]}) counts differ"
        o32, o64 = oracle.rmsd_plan(c, np.arange(c.shape[1]), c[0], align=True)
        got = b.rmsd_of(rm, t)
        assert np.max(np.abs(got.astype(np.float64) - o32)) <= 1e-5, f"trajectory {t} RMSD"


def test_batch_rmsd_f64_within_error_bound():
    import torch
    coords, boxes = _case([500, 2500, 9000], n_frames=4, seed=2, spread=0.3)
    b, store = _batch(coords, boxes)
    F = sum(c.shape[0] for c in coords)
    r32 = torch.zeros(F, dtype=torch.float32, device="cuda:0")
    r64 = torch.zeros(F, dtype=torch.float64, device="cuda:0")
    err = torch.zeros(F, dtype=torch.float64, device="cuda:0")
    b.run(store, r32, r64, err)
    r64 = r64.cpu().numpy()
    err = err.cpu().numpy()
    for t, c in enumerate(coords):
        _, o64 = oracle.rmsd_plan(c, np.arange(c.shape[1]), c[0], align=True)
        d = np.abs(b.rmsd_of(r64, t) - o64)
        assert np.all(d <= b.rmsd_of(err, t) + 1e-9), (t, d, b.rmsd_of(err, t))
        assert b.rmsd_of(r64, t)[0] <= 1e-6  # frame 0 against itself


def test_spread_coordinates_and_images():
    # atoms up to half a box outside [0, L): image indices != 0 everywhere
    coords, boxes = _case([900, 5000, 12000], n_frames=2, seed=3, spread=0.5)
    b, store = _batch(coords, boxes, rmsd=False)
    b.run(store)
    counts = b.counts()
    for t, (c, box) in enumerate(zip(coords, boxes)):
        assert np.array_equal(counts[t], _oracle_counts(c, box, "orthorhombic")), t


def test_triclinic_batch():
    coords, boxes = _case([600, 4000, 15000], n_frames=2, seed=4, triclinic=True)
    b, store = _batch(coords, boxes, pbc="triclinic", rmsd=False)
    b.run(store)
    counts = b.counts()
    for t, (c, box) in enumerate(zip(coords, boxes)):
        assert np.array_equal(counts[t], _oracle_counts(c, box, "triclinic")), t


def test_no_pbc_batch():
    coords, _ = _case([100, 1200, 2500], n_frames=2, seed=5)
    b, store = _batch(coords, None, pbc="none", rmsd=False)
    b.run(store)
    counts = b.counts()
    for t, c in enumerate(coords):
        assert np.array_equal(counts[t], _oracle_counts(c, None, "none")), t


def test_graph_replays_match_eager():
    import torch
    coords, boxes = _case([350, 4000, 10000], n_frames=3, seed=6)
    be, se = _batch(coords, boxes, graph=False)
    re = be.run(se).clone()
    ce = be.counts()
    bg, sg = _batch(coords, boxes, graph=True)
    out = torch.zeros_like(re)
    for k in range(1, 4):
        bg.run(sg, out)
        # eager work of the program between replays
        junk = torch.randn(1 << 20, device="cuda:0").sum()
        torch.cuda.synchronize()
        assert np.array_equal(bg.counts(), k * ce), f"replay {k}"
        assert torch.equal(out, re), f"replay {k} RMSD"
        out.zero_()
        del junk
    bg.reset()
    bg.run(sg, out)
    assert np.array_equal(bg.counts(), ce)


def test_autotune_keeps_results_and_exports():
    coords, boxes = _case([300, 2000, 6000, 16000], n_frames=2, seed=7)
    b, store = _batch(coords, boxes, rmsd=False)
    b.run(store)
    before = b.counts()
    b.autotune(store, reps=1)
    kinds, _ = b.plan()
    print("tuned", list(kinds), flush=True)
    b.reset()
    b.run(store)
    assert np.array_equal(b.counts(), before)
    text = export_tuning()
    assert text.strip(), "autotuner cache is empty"
    import_tuning(text)
    assert export_tuning() == text


def test_forced_enumerations_agree():
    coords, boxes = _case([5000], n_frames=2, seed=8)
    ref = _oracle_counts(coords[0], boxes[0], "orthorhombic")
    for k in (-1, 1, 2, 3):
        b, store = _batch(coords, boxes, rmsd=False)
        b.set_tuning([k])
        kinds, _ = b.plan()
        b.run(store)
        assert np.array_equal(b.counts()[0], ref), (k, kinds)


def test_home_blocks_and_task_sizes_agree():
    """Brute-force home blocks of 64..512 atoms (sizes that are and are not
    multiples of the block, and one below a single wave) and 1..64 units per
    task, set per trajectory, all give the oracle's counts."""
    sizes = [37, 300, 700, 1100, 5000]
    coords, boxes = _case(sizes, n_frames=2, seed=21)
    refs = [_oracle_counts(c, box, "orthorhombic") for c, box in zip(coords, boxes)]
    for bh, g in [(64, 1), (128, 3), (256, 8), (512, 64), (192, 2)]:
        b, store = _batch(coords, boxes, rmsd=False)
        b.set_tuning([-1, -1, -1, -1, 0], [bh] * len(sizes), [g] * len(sizes))
        kinds, homes, grabs = b.plan_tuning()
        assert list(kinds[:4]) == [0, 0, 0, 0] and kinds[4] > 0
        assert list(homes[:4]) == [bh] * 4 and list(grabs) == [g] * len(sizes)
        b.run(store)
        counts = b.counts()
        for t in range(len(sizes)):
            assert np.array_equal(counts[t], refs[t]), (bh, g, t)


def test_slice_times_are_consistent():
    coords, boxes = _case([3000, 8000, 20000, 900], n_frames=4, seed=22)
    b, store = _batch(coords, boxes, rmsd=False)
    b.run(store)
    st = b.slice_times()
    print("slice times", st.tolist(), flush=True)
    assert st.shape == (8, 4)
    started = st[:, 0] >= 0
    assert started.all()
    assert np.all(st[:, 2] >= st[:, 0])
    assert np.all((st[:, 1] < 0) | (st[:, 1] <= st[:, 2] + 1e-3))
    assert np.all(st[:, 3] >= 0)


def test_full_uneven_batch_matches_single_plans():
    """256 trajectories, log-uniform N in [500, 50000], density 0.1, 50 frames:
    the batch equals the single-trajectory engine paths on every trajectory."""
    import ctypes

    import torch

    from warp_md_amd import _native, workloads
    from warp_md_amd._native import check, lib, ptr

    spec = workloads.uneven_batch(sizes_only=True)
    F = 50
    b = UnevenBatch([s["n"] for s in spec], [F] * len(spec), bins=BINS, r_max=RMAX,
                    device="cuda:0")
    b.set_boxes([s["box"] for s in spec])
    store = workloads.uneven_store_device(b, spec, seed=3)
    r = b.run(store)
    torch.cuda.synchronize()
    counts = b.counts()
    rm = r.cpu().numpy()
    ctx = _native.context(0)
    stream = torch.cuda.current_stream().cuda_stream
    worst = 0.0
    for t, s in enumerate(spec):
        n = s["n"]
        sel = np.arange(n, dtype=np.uint32)
        view = b.store_view(store, t)
        boxes = np.repeat(np.diag(s["box"]).reshape(1, 9), F, 0).astype(np.float32)
        h = ctypes.c_void_p()
        check(lib().wmd_rdf_create(ctx.handle, ptr(sel), n, ptr(sel), n, BINS,
                                   ctypes.c_float(RMAX), 1, 0, ctypes.byref(h)))
        try:
            check(lib().wmd_rdf_accumulate_device(h, ctypes.c_void_p(view.data_ptr()),
                                                  view.shape[2], F, ptr(boxes),
                                                  ctypes.c_void_p(stream)))
            single = np.zeros(BINS, np.uint64)
            check(lib().wmd_rdf_read(h, ptr(single), None))
        finally:
            lib().wmd_rdf_destroy(h)
        assert np.array_equal(counts[t], single), f"trajectory {t} (n={n})"
        # RMSD to frame 0 through the single RmsdPlan engine path
        x0 = view[0, :, :n].t().cpu().numpy()
        r4 = np.ones((n, 4), np.float32)
        r4[:, :3] = x0
        hr = ctypes.c_void_p()
        check(lib().wmd_rmsd_create(ctx.handle, ptr(sel), n, None, 0, ptr(r4), n, 1,
                                    ctypes.byref(hr)))
        try:
            ns, st = ctypes.c_size_t(), ctypes.c_size_t()
            atoms = np.zeros(view.shape[2], np.uint32)
            check(lib().wmd_rmsd_store_layout(hr, ptr(atoms), atoms.size, ctypes.byref(ns),
                                              ctypes.byref(st)))
            assert st.value == view.shape[2] and np.array_equal(atoms[:n], sel)
            out = torch.zeros(F, dtype=torch.float32, device="cuda:0")
            check(lib().wmd_rmsd_run_device(hr, ctypes.c_void_p(view.data_ptr()), st.value, F,
                                            ctypes.c_void_p(out.data_ptr()), None, None, None, None,
                                            ctypes.c_void_p(stream)))
            single_r = out.cpu().numpy()
        finally:
            lib().wmd_rmsd_destroy(hr)
        d = float(np.max(np.abs(b.rmsd_of(rm, t).astype(np.float64) - single_r)))
        worst = max(worst, d)
    print(f"256 trajectories: counts identical, RMSD max diff {worst:.2e} A", flush=True)
    assert worst <= 1e-5


@pytest.mark.parametrize("graph", [True, False])
def test_batch_histogram_larger_than_lds(graph):
    """5,000 bins: the per-wave LDS histograms cannot hold them, so the pair
    launch runs one pass per slice of bins; counts bit-exact per trajectory
    (cells and brute force)."""
    sizes = [300, 1500, 8000]
    coords, boxes = _case(sizes, n_frames=2, seed=9)
    bins = 5000
    b = UnevenBatch(sizes, [2] * len(sizes), bins=bins, r_max=RMAX, pbc="orthorhombic", rmsd=False,
                    graph=graph, device="cuda:0")
    b.set_boxes(boxes)
    store = b.store_from_numpy(coords)
    print(f"batch bins={bins} graph={graph} plan {list(b.plan()[0])}", flush=True)
    b.run(store)
    b.run(store)  # a replay adds the same counts again
    counts = b.counts()
    for t, c in enumerate(coords):
        allat = np.arange(c.shape[1])
        ref = oracle.rdf_counts(c, allat, allat, bins, RMAX, pbc="orthorhombic", boxes=boxes[t],
                                method="omp" if c.shape[1] > 3000 else "brute")
        assert np.array_equal(counts[t], 2 * ref), f"trajectory {t}"
