"""Pins of the PairDistPlan and StructureFactorPlan restatements (oracle/wmd_oracle.c
orc_pairdist, orc_sf_counts, orc_sf_finalize) and of the host statistics."""
import math

import numpy as np
import pytest

import oracle
from warp_md_amd.plans import pairdist_centers, pairdist_stats


def test_pairdist_reference_distribution_stats():
    # correlation_structure_tests.rs:2193-2231: two atoms, unique pairs,
    # distribution output, compact; r = 1 then 2 over two frames
    c = np.array([[[0, 0, 0], [1, 0, 0]], [[0, 0, 0], [2, 0, 0]]], np.float32)
    counts, fc = oracle.pairdist_counts(c, [0, 1], [0, 1], 4, 4.0, "none", same_sel=True,
                                        unique_pairs=True, per_frame=True)
    sums = fc.sum(0).astype(np.float64)
    sumsq = (fc.astype(np.float64) ** 2).sum(0)
    centers, prob, std, cnt, frames = pairdist_stats(counts, sums, sumsq, 2, np.float32(1.0), True)
    assert frames == 2
    assert centers.tolist() == [1.5, 2.5]
    assert cnt.tolist() == [1, 1]
    assert prob.tolist() == [0.5, 0.5]
    assert np.allclose(std, math.sqrt(0.5), atol=1e-6)


def test_pairdist_selected_chunk_reference_counts():
    # correlation_structure_tests.rs:2313-2422: atoms 0 and 2 selected (1 far away)
    frames = np.array([[[0, 0, 0], [1000, 1000, 1000], [1, 0, 0]],
                       [[0, 0, 0], [1000, 1000, 1000], [2, 0, 0]]], np.float32)
    counts = oracle.pairdist_counts(frames, [0, 2], [0, 2], 4, 4.0, "none", same_sel=True,
                                    unique_pairs=True)
    assert counts.tolist() == [0, 1, 1, 0]


def _random(n, L, F, seed):
    rng = np.random.default_rng(seed)
    return (rng.random((F, n, 3)) * L).astype(np.float32)


@pytest.mark.parametrize("pbc", ["none", "orthorhombic"])
def test_pairdist_rules_against_rdf(pbc):
    c = _random(120, 9.0, 3, seed=1)
    box = np.array([9.0, 9.0, 9.0], np.float32)
    sel = np.arange(0, 120, 2)
    rdf = oracle.rdf_counts(c, sel, sel, 40, 4.0, pbc=pbc, boxes=box)
    # same object: a == b skipped, exactly RdfPlan's pair set
    same = oracle.pairdist_counts(c, sel, sel, 40, 4.0, pbc, box, same_sel=True)
    assert np.array_equal(same, rdf)
    # equal lists held by different objects: self pairs counted at r = 0
    other = oracle.pairdist_counts(c, sel, sel, 40, 4.0, pbc, box, same_sel=False)
    exp = rdf.copy()
    exp[0] += sel.size * 3
    assert np.array_equal(other, exp)
    # unique pairs: every unordered pair once
    uniq = oracle.pairdist_counts(c, sel, sel, 40, 4.0, pbc, box, same_sel=True, unique_pairs=True)
    assert np.array_equal(uniq * 2, rdf)


def _sf_counts_numpy(c, sel, bins, r_max, pbc, box, scale):
    """float64 numpy restatement of structure_factor.rs:160-198."""
    counts = np.zeros(bins, np.int64)
    w = r_max / bins
    for f in range(c.shape[0]):
        x = c[f, sel]
        d32 = (x[None, :, :] - x[:, None, :]).astype(np.float32)  # b - a in f32
        d = d32.astype(np.float64) * scale
        if pbc == "orthorhombic":
            L = np.asarray(box, np.float32).astype(np.float64) * scale
            q = d / L
            d = d - np.copysign(np.floor(np.abs(q) + 0.5), q) * L
        r = np.sqrt(d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2])
        np.fill_diagonal(r, np.inf)
        ok = r < r_max
        b = (r[ok] / w).astype(np.int64)
        b = b[b < bins]
        counts += np.bincount(b, minlength=bins)
    return counts.astype(np.uint64)


@pytest.mark.parametrize("pbc,scale", [("none", 1.0), ("orthorhombic", 1.0), ("orthorhombic", 0.1)])
def test_sf_counts_match_numpy(pbc, scale):
    c = _random(90, 12.0, 2, seed=2)
    box = np.array([12.0, 12.0, 12.0], np.float32)
    sel = np.arange(90)
    counts, vol = oracle.sf_counts(c, sel, 30, 6.0 * scale, pbc, box, scale)
    assert np.array_equal(counts, _sf_counts_numpy(c, sel, 30, 6.0 * scale, pbc, box, scale))
    if pbc == "orthorhombic":
        assert vol == pytest.approx(2 * (12.0 ** 3) * scale ** 3, rel=1e-12)


def test_sf_lattice_edges_exact():
    # simple cubic lattice, spacing 1: shells at 1, sqrt 2, sqrt 3, 2 with r_max 2.5 / 5 bins:
    # r = 1 and r = 2 sit exactly on bin edges (bins 2 and 4 of width 0.5)
    g = np.arange(6, dtype=np.float32)
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(1, -1, 3)
    counts, _ = oracle.sf_counts(pos, np.arange(216), 5, 2.5, "orthorhombic",
                                 np.array([6.0, 6.0, 6.0], np.float32), 1.0)
    # per atom: 6 at 1 (bin 2), 12 at sqrt 2 = 1.414 (bin 2), 8 at sqrt 3 = 1.732 (bin 3),
    # 6 at 2, 24 at sqrt 5 = 2.236 and 24 at sqrt 6 = 2.449 (bin 4)
    assert counts.tolist() == [0, 0, 216 * 18, 216 * 8, 216 * 54]


def test_sf_finalize_matches_numpy_and_shapes():
    rng = np.random.default_rng(5)
    counts = rng.integers(0, 500, size=40).astype(np.uint64)
    r_max, n, frames, vol, qb, qmax = 8.0, 300.0, 4, 4 * 1000.0, 25, 6.0
    r, g, q, sq = oracle.sf_finalize(counts, r_max, n, frames, vol, "orthorhombic", qb, qmax)
    # structure_factor_shapes (correlation_structure_tests.rs:759-780)
    assert r.shape == (40,) and g.shape == (40,) and q.shape == (qb,) and sq.shape == (qb,)
    w = r_max / 40
    r0 = (np.arange(40) + 0.5) * w
    rho = n / (vol / frames)
    g_ref = counts / (rho * 4 * np.pi * r0 * r0 * w * n * frames)
    assert np.array_equal(r, r0.astype(np.float32))
    assert np.array_equal(g, g_ref.astype(np.float32))
    qv = qmax / (qb - 1) * np.arange(qb)
    gd = g.astype(np.float64) - 1.0
    with np.errstate(invalid="ignore", divide="ignore"):
        wt = np.where(qv[:, None] == 0, 1.0, np.sin(qv[:, None] * r0) / (qv[:, None] * r0))
    s_ref = 1.0 + 4 * np.pi * rho * (gd * wt * r0 * r0 * w).sum(1)
    assert np.allclose(sq, s_ref.astype(np.float32), rtol=1e-6, atol=1e-6)
    assert np.array_equal(q, qv.astype(np.float32))
