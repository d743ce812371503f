"""Pins the CPU oracle's RDF restatement before it is trusted as the checker.

Checks: the reference's own known answers (tests/golden), a pure-Python
brute-force restatement on tiny inputs (orthorhombic, triclinic, no PBC), a
closed-form lattice shell count, symmetry invariants, cell-list == brute force,
and the finalize normalisation against an independent numpy restatement.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))


def _frames_for(case):
    frames = np.asarray(case["frames"], np.float32)
    if "frame_indices" in case:
        n = frames.shape[0]
        raw = case["frame_indices"]
        if any(i < 0 for i in raw):  # executor.rs:210-222 normalize_frame_indices
            idx = [i + n if i < 0 else i for i in raw]
        else:
            idx = list(raw)
        idx = [i for i in idx if 0 <= i < n]
        frames = frames[idx]
    return frames


@pytest.mark.parametrize("case", GOLD["rdf"], ids=lambda c: c["name"])
def test_reference_known_answers(case):
    frames = _frames_for(case)
    counts = oracle.rdf_counts(frames, case["sel_a"], case["sel_b"], case["bins"], case["r_max"],
                               pbc=case["pbc"])
    if "expect_counts" in case:
        np.testing.assert_array_equal(counts, np.asarray(case["expect_counts"], np.uint64))
    for k, v in case.get("expect_counts_at", {}).items():
        assert counts[int(k)] == v
    opts = case.get("options", {})
    n_a, n_b = len(case["sel_a"]), len(case["sel_b"])
    common = len(set(case["sel_a"]) & set(case["sel_b"]))
    out = oracle.rdf_finalize(counts, case["r_max"], n_a, n_b, common, frames.shape[0],
                              density=opts.get("density", 0.033456),
                              use_volume=opts.get("volume", False), raw=opts.get("raw_rdf", False),
                              number_density=opts.get("number_density", False),
                              integral=opts.get("intrdf", False))
    g = out[1]
    tol = case.get("tol", 0.0)
    for k, v in case.get("expect_g_at", {}).items():
        assert abs(g[int(k)] - v) <= tol
    if "expect_g" in case:
        np.testing.assert_allclose(g, case["expect_g"], atol=tol, rtol=0)
    for k, v in case.get("expect_integral_at", {}).items():
        assert abs(out[3][int(k)] - v) <= tol
    if case.get("expect_no_integral"):
        assert len(out) == 3


def _round_half_away(q):
    q = float(q)
    return np.float32(math.copysign(math.floor(abs(q) + 0.5), q))


def _inv_rows_f32(m):
    m = np.asarray(m, np.float64).reshape(3, 3)
    inv = np.linalg.inv(m)  # independent of the oracle's cofactor formula
    return m.astype(np.float32), inv.astype(np.float32)


def _py_pair_r(pa, pb, kind, box):
    f32 = np.float32
    dx = f32(pb[0] - pa[0]); dy = f32(pb[1] - pa[1]); dz = f32(pb[2] - pa[2])
    if kind == "orthorhombic":
        lx, ly, lz = f32(box[0]), f32(box[1]), f32(box[2])
        dx = f32(dx - f32(_round_half_away(f32(dx / lx)) * lx))
        dy = f32(dy - f32(_round_half_away(f32(dy / ly)) * ly))
        dz = f32(dz - f32(_round_half_away(f32(dz / lz)) * lz))
    elif kind == "triclinic":
        c, r = box
        n = []
        for k in range(3):
            v = f32(f32(f32(r[0][k] * dx) + f32(r[1][k] * dy)) + f32(r[2][k] * dz))
            n.append(f32(_round_half_away(v)))
        d = [dx, dy, dz]
        for k in range(3):
            shift = f32(f32(f32(n[0] * c[0][k]) + f32(n[1] * c[1][k])) + f32(n[2] * c[2][k]))
            d[k] = f32(d[k] - shift)
        dx, dy, dz = d
    s = f32(f32(f32(dx * dx) + f32(dy * dy)) + f32(dz * dz))
    return f32(np.sqrt(s))


def _py_counts(frames, sel_a, sel_b, bins, r_max, kind, boxes):
    f32 = np.float32
    w = f32(f32(r_max) / f32(bins))
    counts = np.zeros(bins, np.uint64)
    for fi, fr in enumerate(frames):
        box = boxes[fi] if boxes is not None else None
        for a in sel_a:
            for b in sel_b:
                if a == b:
                    continue
                r = _py_pair_r(fr[a], fr[b], kind, box)
                if r < f32(r_max):
                    k = int(f32(r / w))
                    if k < bins:
                        counts[k] += 1
    return counts


@pytest.mark.parametrize("kind", ["none", "orthorhombic", "triclinic"])
def test_brute_matches_python_restatement(kind):
    rng = np.random.default_rng(11)
    n, F, bins, r_max = 40, 3, 17, 4.7
    L = 10.3
    frames = (rng.random((F, n, 3)) * L * 1.3 - 0.15 * L).astype(np.float32)  # some outside the box
    sel_a = list(range(0, 30)) + [3]  # a duplicate entry
    sel_b = list(range(10, 40))
    if kind == "orthorhombic":
        boxes_py = [np.array([L, L * 1.1, L * 0.95], np.float32)] * F
        boxes = np.stack([np.diag(b).reshape(9) for b in boxes_py])
    elif kind == "triclinic":
        m = np.array([[L, 0, 0], [L * math.tan(math.radians(10)), L, 0], [0.7, -0.4, L]], np.float32)
        # oracle uses the f64 cofactor inverse cast to f32; check that equals numpy's inverse cast
        c, r = _inv_rows_f32(m)
        boxes_py = [(c.tolist(), r.tolist())] * F
        boxes = np.repeat(m.reshape(1, 9), F, axis=0)
    else:
        boxes_py = None
        boxes = None
    ref = _py_counts(frames, sel_a, sel_b, bins, r_max, kind, boxes_py)
    got = oracle.rdf_counts(frames, sel_a, sel_b, bins, r_max, pbc=kind, boxes=boxes)
    assert ref.sum() > 0
    np.testing.assert_array_equal(got, ref)


def _lattice_shells(a, r_max, bins):
    """Closed form for a jitter-free simple cubic lattice: neighbour vectors a*(i,j,k)."""
    w = r_max / bins
    out = np.zeros(bins, np.int64)
    R = int(math.ceil(r_max / a)) + 1
    for i in range(-R, R + 1):
        for j in range(-R, R + 1):
            for k in range(-R, R + 1):
                if i == j == k == 0:
                    continue
                d = a * math.sqrt(i * i + j * j + k * k)
                if d < r_max:
                    q = d / w
                    assert abs(q - round(q)) * w > 1e-3, "shell too close to a bin edge"
                    out[int(q)] += 1
    return out


def test_lattice_closed_form():
    n, a = 7, 2.1
    L = n * a
    r_max, bins = 6.6, 21  # r_max < L/2 = 7.35; no shell within 1e-3 A of an edge
    g = (np.arange(n) + 0.5) * a
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)[None]
    N = pos.shape[1]
    counts = oracle.rdf_counts(pos, np.arange(N), np.arange(N), bins, r_max, pbc="orthorhombic",
                               boxes=[L, L, L])
    expect = _lattice_shells(a, r_max, bins) * N
    np.testing.assert_array_equal(counts.astype(np.int64), expect)


def test_symmetry_and_permutation_invariance():
    rng = np.random.default_rng(5)
    n, L = 120, 14.0
    frames = (rng.random((2, n, 3)) * L).astype(np.float32)
    a = np.arange(0, 80)
    b = np.arange(40, 120)
    c1 = oracle.rdf_counts(frames, a, b, 40, 6.5, boxes=[L, L, L])
    c2 = oracle.rdf_counts(frames, b, a, 40, 6.5, boxes=[L, L, L])
    np.testing.assert_array_equal(c1, c2)
    perm = rng.permutation(n)
    inv = np.argsort(perm)
    c3 = oracle.rdf_counts(frames[:, perm], inv[a], inv[b], 40, 6.5, boxes=[L, L, L])
    np.testing.assert_array_equal(c1, c3)
    same = oracle.rdf_counts(frames, a, a, 40, 6.5, boxes=[L, L, L])
    assert int(same.sum()) % 2 == 0


@pytest.mark.parametrize("kind", ["orthorhombic", "triclinic"])
def test_cells_equal_brute(kind):
    rng = np.random.default_rng(7)
    n, F, L = 900, 2, 30.0
    frames = (rng.random((F, n, 3)) * L).astype(np.float32)
    frames[0, :5] -= L * 0.02  # a few raw coordinates outside the primary box
    if kind == "orthorhombic":
        boxes = np.repeat(np.diag([L, L * 1.05, L * 0.97]).reshape(1, 9), F, 0)
    else:
        m = np.array([[L, 0, 0], [L * math.tan(math.radians(10)), L, 0], [1.0, 2.0, L]])
        boxes = np.repeat(m.reshape(1, 9), F, 0)
    sel_a = np.arange(0, 700)
    sel_b = np.arange(200, 900)
    b1 = oracle.rdf_counts(frames, sel_a, sel_b, 60, 9.0, pbc=kind, boxes=boxes)
    b2 = oracle.rdf_counts(frames, sel_a, sel_b, 60, 9.0, pbc=kind, boxes=boxes, method="cells")
    b3 = oracle.rdf_counts(frames, sel_a, sel_b, 60, 9.0, pbc=kind, boxes=boxes, method="omp",
                           threads=4)
    np.testing.assert_array_equal(b1, b2)
    np.testing.assert_array_equal(b1, b3)


@pytest.mark.parametrize("mode", ["rdf", "volume", "number_density", "raw"])
@pytest.mark.parametrize("integral", [False, True])
def test_finalize_matches_numpy_restatement(mode, integral):
    rng = np.random.default_rng(3)
    counts = rng.integers(0, 10_000, size=200).astype(np.uint64)
    kw = dict(r_max=10.0, n_a=1000.0, n_b=1000.0, common=1000.0, frames=100,
              volume_sum=100 * 31.04 ** 3, density=0.033456, integral=integral,
              use_volume=(mode == "volume"), raw=(mode == "raw"),
              number_density=(mode == "number_density"))
    a = oracle.rdf_finalize(counts, **kw)
    b = oracle.rdf_finalize_py(counts, **kw)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_finalize_volume_requires_boxes():
    with pytest.raises(RuntimeError):
        oracle.rdf_finalize(np.zeros(4, np.uint64), 4.0, 2, 2, 2, 1, volume_sum=0.0,
                            use_volume=True)


def test_edge_diagnostics_flag_near_edge_pairs():
    # two atoms 1.0 A apart with bin width 0.5: the pair sits exactly on an edge
    fr = np.array([[[0, 0, 0], [1.0, 0, 0], [0, 3.21, 0]]], np.float32)
    counts, edge = oracle.rdf_counts(fr, [0, 1, 2], [0, 1, 2], 8, 4.0, pbc="none", edges=True)
    assert edge[0] == 2 and edge[1] == 2 and edge[2] == 0
    assert counts[2] == 2


@pytest.mark.parametrize("case", GOLD["min_image"], ids=lambda c: c["name"])
def test_min_image_known_vectors(case):
    box = case["box"]
    out = oracle.min_image_f64(case["a"], case["b"], box=box, kind=case["kind"],
                               length_scale=case["scale"])
    np.testing.assert_allclose(out, case["expect"], atol=case["tol"], rtol=0)


@pytest.mark.parametrize("shear", [10.0, 25.0])
def test_triclinic_f32_agrees_with_f64_apply_pbc(shear):
    # the f32 triclinic order (n = round(inv . d), d -= (n0 a + n1 b) + n2 c) must
    # bin every pair as the f64 apply_pbc_triclinic (pbc_math.rs:92-108) does,
    # except pairs within 1e-5 A of a bin edge; a wrong image, sign or index
    # in the lattice shift moves thousands of pairs
    rng = np.random.default_rng(int(shear))
    n, F, L = 1500, 2, 32.0
    m = np.array([[L, 0, 0], [L * math.tan(math.radians(shear)), L, 0], [1.5, -2.0, L]], np.float64)
    frac = rng.random((F, n, 3)) * 1.2 - 0.1  # some atoms outside the primary cell
    frames = (frac @ m).astype(np.float32)
    boxes = np.repeat(m.reshape(1, 9), F, 0).astype(np.float32)
    counts, edge = oracle.rdf_counts(frames, np.arange(n), np.arange(n), 120, 12.0,
                                     pbc="triclinic", boxes=boxes, edges=True)
    assert counts.sum() > 100_000
    assert int(edge[2]) <= int(edge[1]) + int(edge[0])
    assert int(edge[2]) <= 1e-4 * counts.sum()
