"""The host-chunk ingest pipeline across many staging batches.

Every host-chunk entry point streams its chunk through two pinned staging
buffers, a batch of frames at a time: host copy into staging buffer b, its
H2D copy on the copy stream, the device gather and kernels on the context's
stream, with events ordering the reuse of each buffer (the engine's side of
the reference's chunk loop, crates/traj-engine/src/executor.rs:486-509).  The
alternating-buffer logic only runs from the second batch on, so these tests
shrink the staging buffers (wmd_ctx_set_staging_bytes) until a chunk spans 8
or more batches, and check each entry point against

* the oracle (bit-exact counts; RMSD within 1e-5 A; superposition within
  1e-5 A or one f32 ulp; the pairwise matrix within 1e-5 A), and
* the same inputs through the device-resident path or in one batch
  (bit-exact counts, identical superposition and pairwise results).

At the default 48 MiB staging, 100 frames of the 300,000-atom water system
(480 MB of chunk, 10 batches) go through RdfPlan's full-chunk path and are
compared with the device-resident store and, on two frames, with the oracle.
"""
import contextlib
import ctypes

import numpy as np
import pytest

import oracle
from warp_md_amd import RdfPlan, System, Trajectory, _native, workloads
from warp_md_amd._native import check, lib, ptr

import _gpu

pytestmark = pytest.mark.gpu


@contextlib.contextmanager
def staging(nbytes):
    ctx = _gpu.ctx()
    prev = ctx.set_staging_bytes(nbytes)
    try:
        yield
    finally:
        ctx.set_staging_bytes(prev)


def _batches(n_frames, frame_bytes, nbytes):
    per = max(1, nbytes // frame_bytes)
    return -(-n_frames // per)


def _xyzw(c):
    x4 = np.ones(c.shape[:2] + (4,), np.float32)
    x4[..., :3] = c
    return x4


def _store(frames):
    """(F, n, 3) host frames -> (F, 3, stride) torch store on the GPU."""
    import torch

    F, n, _ = frames.shape
    stride = (n + 3) // 4 * 4
    s = np.zeros((F, 3, stride), np.float32)
    s[:, :, :n] = np.transpose(frames, (0, 2, 1))
    return torch.tensor(s, device="cuda:0")


def _water(F, seed, tilt=None):
    w = workloads.water_lattice(n_side=10, box=31.04, n_frames=F, seed=seed, tilt_deg=tilt)
    return w["coords"], w["box"], w["ow"]


@pytest.mark.parametrize("kind", ["orthorhombic", "triclinic"])
def test_rdf_full_chunk_many_batches(kind):
    F = 200
    coords, box, ow = _water(F, seed=11, tilt=10.0 if kind == "triclinic" else None)
    if kind == "orthorhombic":
        # a different box every frame: each batch uploads its own boxes
        box = np.stack([box * np.float32(1.0 + 0.0005 * f) for f in range(F)]).astype(np.float32)
    nbytes = 1 << 20
    assert _batches(F, coords.shape[1] * 16, nbytes) >= 8
    print(f"ingest rdf full chunk {kind}", flush=True)
    with staging(nbytes):
        got, _, path = _gpu.rdf_counts(coords, ow, ow, 200, 10.0, pbc=kind, box=box)
    bx = _gpu.boxes9(box, F, kind)
    ref = oracle.rdf_counts(coords, ow, ow, 200, 10.0, pbc=kind, boxes=bx)
    np.testing.assert_array_equal(got, ref)
    dev, _ = _gpu.rdf_counts_device(_store(coords[:, ow]), ow.size, ow.size, 200, 10.0, kind, box)
    np.testing.assert_array_equal(got, dev)
    assert int(got.sum()) > 0


def test_rdf_selected_chunk_many_batches():
    F = 150
    coords, box, ow = _water(F, seed=12)
    sel_a = ow
    sel_b = np.arange(1, coords.shape[1], 3, dtype=np.uint32)  # HW1
    h = ctypes.c_void_p()
    check(lib().wmd_rdf_create(_gpu.ctx().handle, ptr(sel_a), sel_a.size, ptr(sel_b), sel_b.size, 150,
                               ctypes.c_float(9.0), _gpu.PBC["orthorhombic"], 0, ctypes.byref(h)))
    try:
        n_io = ctypes.c_size_t()
        check(lib().wmd_rdf_io_selection(h, None, 0, ctypes.byref(n_io)))
        io = np.zeros(n_io.value, np.uint32)
        check(lib().wmd_rdf_io_selection(h, ptr(io), io.size, ctypes.byref(n_io)))
        sel_chunk = np.ascontiguousarray(_xyzw(coords)[:, io.astype(np.int64)])
        bx = _gpu.boxes9(box, F, "orthorhombic")
        nbytes = 256 << 10
        assert _batches(F, io.size * 16, nbytes) >= 8
        print("ingest rdf selected chunk", flush=True)
        with staging(nbytes):
            check(lib().wmd_rdf_accumulate_selected(h, ptr(sel_chunk), F, ptr(bx)))
        got = np.zeros(150, np.uint64)
        check(lib().wmd_rdf_read(h, ptr(got), None))
    finally:
        lib().wmd_rdf_destroy(h)
    ref = oracle.rdf_counts(coords, sel_a, sel_b, 150, 9.0, pbc="orthorhombic", boxes=bx)
    np.testing.assert_array_equal(got, ref)
    one, _, _ = _gpu.rdf_counts(coords, sel_a, sel_b, 150, 9.0, pbc="orthorhombic", box=box)
    np.testing.assert_array_equal(got, one)


def _molecules(n_mol, L, F, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(0, L, size=(F, n_mol, 3))
    h1 = o + rng.normal(scale=0.6, size=o.shape)
    h2 = o + rng.normal(scale=0.6, size=o.shape)
    c = np.stack([o, h1, h2], axis=2).reshape(F, 3 * n_mol, 3).astype(np.float32)
    n = 3 * n_mol
    atoms = {"name": ["OW", "HW1", "HW2"] * n_mol, "resname": ["SOL"] * n,
             "resid": [1 + i // 3 for i in range(n)], "chain_id": [0] * n,
             "mass": [15.999, 1.008, 1.008] * n_mol, "molecule_id": [1 + i // 3 for i in range(n)]}
    return c, System.from_arrays(atoms)


@pytest.mark.parametrize("mode", ["byres12", "nointramol"])
def test_rdf_grouped_many_batches(mode):
    from grouping import rdf_groups

    L, F = 18.0, 40
    c, system = _molecules(300, L, F, seed=13)
    n = c.shape[1]
    sel_a, sel_b = np.arange(0, n, 3), np.arange(n)
    kw = dict(byres1=True, byres2=True) if mode == "byres12" else dict(no_intramol=True)
    box = np.array([L, L, L], np.float32)
    nbytes = 64 << 10
    assert _batches(F, n * 16, nbytes) >= 8
    print(f"ingest rdf grouped {mode}", flush=True)
    with staging(nbytes):
        _, _, got = RdfPlan(sel_a, sel_b, 90, 9.0, pbc="orthorhombic", **kw).run(
            Trajectory.from_numpy(c, box=box), system, device="cuda:0")
    _, _, one = RdfPlan(sel_a, sel_b, 90, 9.0, pbc="orthorhombic", **kw).run(
        Trajectory.from_numpy(c, box=box), system, device="cuda:0")
    ga = rdf_groups(sel_a, system, False, kw.get("byres1", False), False)
    gb = rdf_groups(sel_b, system, False, kw.get("byres2", False), False)
    masses = np.asarray(system.atoms["mass"], np.float32)  # RdfPlan's default mass_weighted=True
    ref = oracle.rdf_grouped(c, ga, gb, 90, 9.0, "orthorhombic", box, masses, kw.get("no_intramol", False))
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(got, one)


def _chain(F, seed):
    c = workloads.chain(n_ca=2000, per_ca=10, n_frames=F, seed=seed)
    return c["ref"], c["coords"], c["ca"]


@pytest.mark.parametrize("split", [False, True])
def test_rmsd_run_many_batches(split):
    F = 60
    ref, coords, ca = _chain(F, seed=14)
    n = coords.shape[1]
    nbytes = 1 << 20
    assert _batches(F, n * 16, nbytes) >= 8
    fit = ca if split else np.arange(n, dtype=np.uint32)
    rsel = np.arange(n, dtype=np.uint32) if split else None
    print(f"ingest rmsd split={split}", flush=True)
    with staging(nbytes):
        got, err = _gpu.rmsd(coords, fit, rsel, ref)
    if split:
        want, want64 = oracle.fit_rmsd(coords, fit, rsel, ref)[:2]
    else:
        want, want64 = oracle.rmsd_plan(coords, fit, ref, align=True)
    assert np.max(np.abs(got.astype(np.float64) - want64)) <= 1e-5
    dev = _gpu.rmsd_device(coords, fit, rsel, ref)[0]
    assert np.max(np.abs(got.astype(np.float64) - dev)) <= 2e-6
    assert np.all(err >= 0) and np.all(err < 1e-6)


def _align_run(coords, sel, ref_all, nbytes=None):
    F, N, _ = coords.shape
    r4 = _xyzw(ref_all[None])[0]
    h = ctypes.c_void_p()
    check(lib().wmd_align_create(_gpu.ctx().handle, ptr(sel), sel.size, None, ptr(r4), r4.shape[0], 0,
                                 ctypes.byref(h)))
    try:
        rt = np.zeros((F, 12), np.float32)
        xyz = np.zeros((F, N, 3), np.float32)
        x4 = _xyzw(coords)
        if nbytes:
            with staging(nbytes):
                check(lib().wmd_align_run(h, ptr(x4), N, F, ptr(rt), ptr(xyz)))
        else:
            check(lib().wmd_align_run(h, ptr(x4), N, F, ptr(rt), ptr(xyz)))
    finally:
        lib().wmd_align_destroy(h)
    return rt, xyz


def test_align_run_many_batches():
    F = 40
    ref, coords, ca = _chain(F, seed=15)
    n = coords.shape[1]
    nbytes = 1 << 20
    assert _batches(F, n * 16, nbytes) >= 8
    print("ingest align/superpose", flush=True)
    rt, xyz = _align_run(coords, ca, ref, nbytes)
    ort, ort64, oxyz = oracle.align(coords, ca, ref, superpose=True)
    assert np.max(np.abs(rt[:, :9].astype(np.float64) - ort64[:, :9])) <= 1e-6
    assert np.max(np.abs(rt[:, 9:].astype(np.float64) - ort64[:, 9:])) <= 1e-5
    d = np.abs(xyz.astype(np.float64) - oxyz.astype(np.float64))
    assert np.all(d <= np.maximum(1e-5, 1.01 * np.spacing(np.abs(oxyz)).astype(np.float64)))
    rt1, xyz1 = _align_run(coords, ca, ref)
    np.testing.assert_array_equal(rt, rt1)
    np.testing.assert_array_equal(xyz, xyz1)


def _pairwise(coords, sel, parts, nbytes=None):
    F, N, _ = coords.shape
    h = ctypes.c_void_p()
    check(lib().wmd_pairwise_create(_gpu.ctx().handle, ptr(sel), sel.size, 0, ctypes.byref(h)))
    try:
        x4 = _xyzw(coords)
        for f0, f1 in parts:
            chunk = np.ascontiguousarray(x4[f0:f1])
            if nbytes:
                with staging(nbytes):
                    check(lib().wmd_pairwise_append(h, ptr(chunk), N, f1 - f0))
            else:
                check(lib().wmd_pairwise_append(h, ptr(chunk), N, f1 - f0))
        out = np.zeros((F, F), np.float32)
        check(lib().wmd_pairwise_finalize(h, ptr(out)))
    finally:
        lib().wmd_pairwise_destroy(h)
    return out


def test_pairwise_append_many_batches():
    F = 48
    ref, coords, ca = _chain(F, seed=16)
    n = coords.shape[1]
    nbytes = 1 << 20
    assert _batches(30, n * 16, nbytes) >= 8
    print("ingest pairwise append", flush=True)
    got = _pairwise(coords, ca, [(0, 30), (30, F)], nbytes)
    want, want64 = oracle.pairwise_rmsd(coords, ca, "rms")
    assert np.max(np.abs(got.astype(np.float64) - want64)) <= 1e-5
    one = _pairwise(coords, ca, [(0, F)])
    np.testing.assert_array_equal(got, one)


def test_rdf_300k_atoms_default_staging():
    """100 frames x 300,000 atoms (480 MB of chunk) at the default staging size."""
    F = 100
    w = workloads.water_lattice(n_side=47, box=144.3, n_frames=F, n_waters=100000, seed=17)
    coords, ow = w["coords"], w["ow"]
    n_atoms = coords.shape[1]
    assert _batches(F, n_atoms * 16, 48 << 20) >= 8
    system = System.from_arrays({"name": w["names"]})
    box = w["box"]
    print("ingest rdf 300k atoms x 100 frames", flush=True)
    _, _, got = RdfPlan(ow, ow, 500, 15.0, raw_rdf=True).run(Trajectory.from_numpy(coords, box=box),
                                                          system, chunk_frames=F, device="cuda:0")
    dev, path = _gpu.rdf_counts_device(_store(coords[:, ow]), ow.size, ow.size, 500, 15.0,
                                       "orthorhombic", box)
    assert path == 2
    np.testing.assert_array_equal(got, dev)
    # the first and the last frame against the oracle's cell-list enumeration
    two = np.ascontiguousarray(coords[[0, F - 1]][:, ow])
    sel = np.arange(ow.size)
    ref = oracle.rdf_counts(two, sel, sel, 500, 15.0, pbc="orthorhombic",
                            boxes=_gpu.boxes9(box, 2, "orthorhombic"), method="cells")
    dev2, _ = _gpu.rdf_counts_device(_store(two), ow.size, ow.size, 500, 15.0, "orthorhombic", box)
    np.testing.assert_array_equal(dev2, ref)


def test_bad_box_in_a_later_batch_leaves_counts_untouched():
    F = 60
    coords, box, ow = _water(F, seed=18)
    bx = _gpu.boxes9(box, F, "orthorhombic").copy()
    bx[F - 1, 0] = 0.0  # the last batch's box is invalid
    h = ctypes.c_void_p()
    check(lib().wmd_rdf_create(_gpu.ctx().handle, ptr(ow), ow.size, ptr(ow), ow.size, 100,
                               ctypes.c_float(8.0), _gpu.PBC["orthorhombic"], 0, ctypes.byref(h)))
    try:
        with staging(256 << 10):
            st = lib().wmd_rdf_accumulate(h, ptr(_xyzw(coords)), coords.shape[1], F, ptr(bx))
        assert st == _native.WMD_ERR_MISMATCH
        got = np.zeros(100, np.uint64)
        check(lib().wmd_rdf_read(h, ptr(got), None))
        assert int(got.sum()) == 0
    finally:
        lib().wmd_rdf_destroy(h)
