"""RdfPlan grouped modes (centre / byres / bymol / no_intramol) and the xy
dimension on the GPU (wmd_rdf_create_groups, WMD_RDF_PLANAR_XY) against the
oracle's restatement of rdf.rs:808-850: counts bit-exact, g(r) identical
(same host normalisation with the group counts, rdf.rs:943-1016)."""
import numpy as np
import pytest

import oracle
from warp_md_amd import RdfPlan, System, Trajectory
from grouping import common_group_count, rdf_groups

pytestmark = pytest.mark.gpu


def _waters(n_mol, L, F, seed):
    """n_mol 3-site molecules (residues of 3 atoms, molecule ids) in a cubic box."""
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


MODES = {
    "center1": dict(center1=True),
    "center2": dict(center2=True),
    "byres1": dict(byres1=True),
    "byres12": dict(byres1=True, byres2=True),
    "bymol12": dict(bymol1=True, bymol2=True),
    "nointramol": dict(no_intramol=True),
    "bymol1_nointramol": dict(bymol1=True, no_intramol=True),
}


@pytest.mark.parametrize("pbc", ["none", "orthorhombic"])
@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("mass", [True, False])
def test_grouped_modes_match_oracle(mode, pbc, mass):
    L = 18.0
    c, system = _waters(300, L, 3, seed=len(mode))
    n = c.shape[1]
    sel_a = np.arange(0, n, 3)           # oxygens
    sel_b = np.arange(n)                 # every atom
    kw = MODES[mode]
    print(f"grouped {mode} pbc={pbc} mass={mass}", flush=True)
    plan = RdfPlan(sel_a, sel_b, 90, 9.0, pbc=pbc, mass_weighted=mass, volume=pbc != "none", **kw)
    traj = Trajectory.from_numpy(c, box=np.array([L, L, L], np.float32) if pbc != "none" else None)
    r, g, counts = plan.run(traj, system, device="cuda:0")
    ga = rdf_groups(sel_a, system, kw.get("center1", False), kw.get("byres1", False), kw.get("bymol1", False))
    gb = rdf_groups(sel_b, system, kw.get("center2", False), kw.get("byres2", False), kw.get("bymol2", False))
    m = np.asarray(system.atoms["mass"], np.float32) if mass else None
    ref = oracle.rdf_grouped(c, ga, gb, 90, 9.0, pbc, np.array([L, L, L], np.float32), m,
                             kw.get("no_intramol", False))
    assert np.array_equal(counts, ref)
    vol = 3 * L ** 3 if pbc != "none" else 0.0
    rr, gg, cc = oracle.rdf_finalize(ref, 9.0, len(ga), len(gb),
                                     common_group_count(ga, gb, kw.get("no_intramol", False)), 3, vol,
                                     0.033456, use_volume=pbc != "none")
    assert np.array_equal(r, rr) and np.array_equal(g, gg)


@pytest.mark.parametrize("pbc", ["none", "orthorhombic"])
def test_planar_xy_matches_oracle(pbc):
    L = 20.0
    rng = np.random.default_rng(5)
    c = rng.uniform(0, L, size=(2, 3000, 3)).astype(np.float32)
    system = System.from_arrays({"name": ["X"] * 3000})
    sel_a, sel_b = np.arange(0, 2000), np.arange(1000, 3000)
    plan = RdfPlan(sel_a, sel_b, 80, 8.0, pbc=pbc, dimension="xy", volume=pbc != "none")
    traj = Trajectory.from_numpy(c, box=np.array([L, L, L], np.float32) if pbc != "none" else None)
    r, g, counts = plan.run(traj, system, device="cuda:0")
    ga = [((int(i),), None, None) for i in sel_a]
    gb = [((int(i),), None, None) for i in sel_b]
    ref = oracle.rdf_grouped(c, ga, gb, 80, 8.0, pbc, np.array([L, L, L], np.float32), planar=True)
    assert np.array_equal(counts, ref)
    assert counts.sum() > 0


def test_grouped_xy_and_reference_known_answers():
    # rdf_plan_xy_dimension (correlation_structure_tests.rs:1979-2008): raw counts
    s = System.from_arrays({"name": ["A", "A"], "resid": [1, 2]}, np.array([[0, 0, 0], [0, 0, 5]], np.float32))
    plan = RdfPlan([0, 1], [0, 1], 2, 2.0, pbc="none", dimension="xy", density=1.0, raw_rdf=True)
    r, g, counts = plan.run(Trajectory.from_numpy(s.positions0[None]), s, device="cuda:0")
    assert counts.tolist() == [2, 0] and g.tolist() == [2.0, 0.0]
    # rdf_plan_bymol_groups_atoms_by_molecule_id (:2514-2550)
    pos = np.array([[0, 0, 0], [0, 0, 2], [4, 0, 0], [4, 0, 2]], np.float32)
    s = System.from_arrays({"name": ["A"] * 4, "resid": [1, 1, 2, 2], "mass": [12.0] * 4,
                            "molecule_id": [1, 1, 2, 2]}, pos)
    plan = RdfPlan(np.arange(4), np.arange(4), 8, 8.0, pbc="none", bymol1=True, bymol2=True)
    _, _, counts = plan.run(Trajectory.from_numpy(pos[None]), s, device="cuda:0")
    assert int(counts.sum()) == 2 and counts[4] == 2
    # rdf_plan_nointramol_skips_same_residue_pairs (:2496-2512)
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    s = System.from_arrays({"name": ["CA"] * 3, "resid": [1, 1, 1]}, pos)
    _, _, counts = RdfPlan(np.arange(3), np.arange(3), 5, 5.0, pbc="none", no_intramol=True).run(
        Trajectory.from_numpy(pos[None]), s, device="cuda:0")
    assert int(counts.sum()) == 0
    with pytest.raises(RuntimeError, match="unsupported"):
        RdfPlan([0, 1], [0, 1], 5, 5.0, pbc="triclinic", byres1=True).run(
            Trajectory.from_numpy(pos[None], box=np.eye(3, dtype=np.float32) * 10), s, device="cuda:0")
