"""PairDistPlan and StructureFactorPlan on the GPU against the oracle
(rdf.rs:1019-1287 and structure_factor.rs:160-258 restated in
oracle/wmd_oracle.c): counts bit-exact, statistics and S(q) from the same
counts."""
import numpy as np
import pytest

import oracle
from warp_md_amd import PairDistPlan, StructureFactorPlan, System, Trajectory
from warp_md_amd.plans import pairdist_stats

pytestmark = pytest.mark.gpu


def _system(n):
    return System.from_arrays({"name": np.array(["X"] * n)})


def _random(n, L, F, seed, spread=0.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(-spread * L, (1 + spread) * L, size=(F, n, 3)).astype(np.float32)


@pytest.mark.parametrize("pbc", ["none", "orthorhombic"])
@pytest.mark.parametrize("rule", ["same", "other", "unique", "cross"])
def test_pairdist_counts_match_oracle(pbc, rule):
    n, L, F = 2000, 30.0, 3
    c = _random(n, L, F, seed=3, spread=0.2)
    box = np.array([L, L, L], np.float32)
    traj = Trajectory.from_numpy(c, box=box if pbc != "none" else None)
    sel_a = np.arange(0, n, 2)
    if rule == "cross":
        sel_b = np.arange(n // 4, n)           # overlapping, different objects
        plan = PairDistPlan(sel_a, sel_b, 120, 12.0, pbc=pbc)
        same, uniq = False, False
    elif rule == "same":
        sel_b = sel_a
        plan = PairDistPlan(sel_a, sel_a, 120, 12.0, pbc=pbc)
        same, uniq = True, False
    elif rule == "other":
        sel_b = sel_a.copy()
        plan = PairDistPlan(sel_a, sel_b, 120, 12.0, pbc=pbc)
        same, uniq = False, False
    else:
        sel_b = sel_a
        plan = PairDistPlan(sel_a, sel_a, 120, 12.0, pbc=pbc, unique_pairs=True)
        same, uniq = True, True
    centers, counts = plan.run(traj, _system(n), device="cuda:0")
    ref = oracle.pairdist_counts(c, sel_a, sel_b, 120, 12.0, pbc, box, same_sel=same,
                                 unique_pairs=uniq)
    assert np.array_equal(counts, ref)
    w = np.float32(12.0) / np.float32(120)
    assert np.array_equal(centers, (np.arange(120, dtype=np.float32) * w + np.float32(0.5) * w))


def test_pairdist_distribution_matches_oracle():
    n, L, F = 1500, 28.0, 4
    c = _random(n, L, F, seed=4)
    box = np.array([L, L, L], np.float32)
    sel = np.arange(n)
    plan = PairDistPlan(sel, sel, 80, 10.0, output_distribution=True, unique_pairs=True,
                        compact_output=True)
    out = plan.run(Trajectory.from_numpy(c, box=box), _system(n), device="cuda:0")
    counts, fc = oracle.pairdist_counts(c, sel, sel, 80, 10.0, "orthorhombic", box, same_sel=True,
                                        unique_pairs=True, per_frame=True)
    sums = fc.sum(0).astype(np.float64)
    sumsq = (fc.astype(np.float64) ** 2).sum(0)
    ref = pairdist_stats(counts, sums, sumsq, F, np.float32(10.0) / np.float32(80), True)
    for a, b in zip(out[:4], ref[:4]):
        assert np.array_equal(a, b)
    assert out[4] == F


def test_pairdist_unique_across_different_selections_refused():
    c = _random(50, 10.0, 1, seed=5)
    plan = PairDistPlan(np.arange(10), np.arange(5, 20), 10, 4.0, pbc="none", unique_pairs=True)
    with pytest.raises(RuntimeError, match="unsupported"):
        plan.run(Trajectory.from_numpy(c), _system(50), device="cuda:0")


@pytest.mark.parametrize("pbc,scale", [("none", 1.0), ("orthorhombic", 1.0), ("orthorhombic", 0.1)])
def test_structure_factor_matches_oracle(pbc, scale):
    n, L, F = 1800, 26.0, 3
    c = _random(n, L, F, seed=6, spread=0.3)
    box = np.array([L, L, L], np.float32)
    sel = np.arange(1, n, 1)
    r_max = 11.0 * scale
    plan = StructureFactorPlan(sel, 150, r_max, 40, 8.0 / scale, pbc=pbc,
                               length_scale=None if scale == 1.0 else scale)
    r, g, q, sq = plan.run(Trajectory.from_numpy(c, box=box if pbc != "none" else None),
                           _system(n), device="cuda:0")
    counts, vol = oracle.sf_counts(c, sel, 150, r_max, pbc, box, scale)
    assert np.array_equal(plan.last_counts, counts)
    rr, gr, qr, sr = oracle.sf_finalize(counts, r_max, sel.size, F, vol, pbc, 40, 8.0 / scale)
    assert np.array_equal(r, rr) and np.array_equal(g, gr) and np.array_equal(q, qr)
    assert np.allclose(sq, sr, rtol=1e-6, atol=1e-6)


def test_structure_factor_lattice_edges():
    # distances exactly on f64 bin edges (simple cubic lattice, spacing 1, width 0.5)
    g = np.arange(10, dtype=np.float32)
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(1, -1, 3)
    pos = np.repeat(pos, 2, 0)
    box = np.array([10.0, 10.0, 10.0], np.float32)
    sel = np.arange(pos.shape[1])
    plan = StructureFactorPlan(sel, 8, 4.0, 5, 3.0)
    plan.run(Trajectory.from_numpy(pos, box=box), _system(pos.shape[1]), device="cuda:0")
    counts, _ = oracle.sf_counts(pos, sel, 8, 4.0, "orthorhombic", box, 1.0)
    assert np.array_equal(plan.last_counts, counts)
