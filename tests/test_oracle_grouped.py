"""Pins the oracle's grouped / xy RdfPlan restatement (orc_rdf_grouped,
rdf.rs:808-850) and its group construction (oracle/grouping.py: rdf_groups,
rdf.rs:279-351; common_group_count :387-401) with the reference's own known
answers (correlation_structure_tests.rs:1979-2008, 2455-2588), and against the
atom loop where the two coincide."""
import itertools

import numpy as np
import pytest

import oracle
from warp_md_amd import System
from grouping import rdf_groups, same_intramolecular_group


def _sys(pos, resid=None, mol=None, names=None):
    n = len(pos)
    atoms = {"name": names or ["A"] * n, "resname": ["MOL"] * n,
             "resid": resid or [1] * n, "chain_id": [0] * n, "mass": [12.0] * n}
    if mol is not None:
        atoms["molecule_id"] = mol
    return System.from_arrays(atoms, np.asarray(pos, np.float32))


def _counts(system, sel, bins, r_max, center=(False, False), byres=(False, False),
            bymol=(False, False), no_intramol=False, planar=False, mass=True):
    ga = rdf_groups(np.asarray(sel), system, center[0], byres[0], bymol[0])
    gb = rdf_groups(np.asarray(sel), system, center[1], byres[1], bymol[1])
    m = np.asarray(system.atoms["mass"], np.float32) if mass else None
    return oracle.rdf_grouped(system.positions0[None], ga, gb, bins, r_max, "none", None, m,
                              no_intramol, planar)


def test_planar_xy_known_answer():
    # rdf_plan_xy_dimension (correlation_structure_tests.rs:1979-2008)
    s = _sys([[0, 0, 0], [0, 0, 5]], resid=[1, 2])
    c = _counts(s, [0, 1], 2, 2.0, planar=True)
    assert c.tolist() == [2, 0]


def test_center1_known_answer():
    # rdf_plan_center1_counts_center_to_atoms (:2455-2474): build_system
    s = _sys([[0, 0, 0], [1, 0, 0]], names=["CA", "CA"])
    c = _counts(s, [0, 1], 5, 5.0, center=(True, False))
    assert int(c.sum()) == 2 and c[0] == 2 and c[1] == 0


def test_byres1_and_nointramol_known_answers():
    # build_plane_system: three CA of one residue
    s = _sys([[0, 0, 0], [1, 0, 0], [0, 1, 0]], names=["CA"] * 3)
    c = _counts(s, [0, 1, 2], 5, 5.0, byres=(True, False))  # :2476-2494
    assert int(c.sum()) == 3 and c[0] == 3
    c = _counts(s, [0, 1, 2], 5, 5.0, no_intramol=True)  # :2496-2512
    assert int(c.sum()) == 0


def test_bymol_and_molecule_before_residue_known_answers():
    # rdf_plan_bymol_groups_atoms_by_molecule_id (:2514-2550)
    s = _sys([[0, 0, 0], [0, 0, 2], [4, 0, 0], [4, 0, 2]], resid=[1, 1, 2, 2], mol=[1, 1, 2, 2])
    c = _counts(s, [0, 1, 2, 3], 8, 8.0, bymol=(True, True))
    assert int(c.sum()) == 2 and c[4] == 2
    # rdf_plan_nointramol_uses_molecule_ids_before_residue_keys (:2552-2588)
    s = _sys([[0, 0, 0], [2, 0, 0]], resid=[1, 1], mol=[1, 2])
    c = _counts(s, [0, 1], 5, 5.0, no_intramol=True)
    assert int(c.sum()) == 2 and c[2] == 2


@pytest.mark.parametrize("pbc", ["none", "orthorhombic"])
def test_single_atom_groups_equal_atom_loop(pbc):
    rng = np.random.default_rng(1)
    c = rng.uniform(0, 12, size=(2, 60, 3)).astype(np.float32)
    box = np.array([12.0, 12.0, 12.0], np.float32)
    sel_a, sel_b = np.arange(0, 40), np.arange(20, 60)
    ga = [((int(i),), None, None) for i in sel_a]
    gb = [((int(i),), None, None) for i in sel_b]
    got = oracle.rdf_grouped(c, ga, gb, 50, 5.0, pbc, box)
    ref = oracle.rdf_counts(c, sel_a, sel_b, 50, 5.0, pbc=pbc, boxes=box)
    assert np.array_equal(got, ref)


def test_python_restatement_small():
    rng = np.random.default_rng(2)
    pos = rng.uniform(0, 6, size=(12, 3)).astype(np.float32)
    s = _sys(pos, resid=[1, 1, 1, 2, 2, 3, 3, 3, 4, 4, 5, 5], mol=[1, 1, 1, 1, 1, 2, 2, 2, 3, 3, 3, 3])
    s.atoms["mass"] = [float(x) for x in rng.uniform(1, 16, 12)]
    sel = np.arange(12)
    for center, byres, bymol, nointra, planar in itertools.product(
            [(False, False), (True, False)], [(False, False), (True, False), (False, True)],
            [(False, False), (True, True)], [False, True], [False, True]):
        ga = rdf_groups(sel, s, center[0], byres[0], bymol[0])
        gb = rdf_groups(sel, s, center[1], byres[1], bymol[1])
        m = np.asarray(s.atoms["mass"], np.float32)
        got = oracle.rdf_grouped(pos[None], ga, gb, 20, 6.0, masses=m, no_intramol=nointra, planar=planar)
        # python loop: f64 centres cast to f32, f32 distances
        def cen(g):
            idx = list(g[0])
            if len(idx) == 1:
                return pos[idx[0]]
            w = m[idx].astype(np.float64)
            return ((pos[idx].astype(np.float64) * w[:, None]).sum(0) / w.sum()).astype(np.float32)
        ref = np.zeros(20, np.uint64)
        w = np.float32(6.0) / np.float32(20)
        for a in ga:
            for b in gb:
                if a[0] == b[0] or (nointra and same_intramolecular_group(a, b)):
                    continue
                d = (cen(b) - cen(a)).astype(np.float32)
                r = np.sqrt(np.float32(d[0] * d[0] + d[1] * d[1]) + (np.float32(0) if planar else np.float32(d[2] * d[2])))
                r = np.float32(r)
                if r < np.float32(6.0):
                    k = int(np.float32(r / w))
                    if k < 20:
                        ref[k] += 1
        assert np.array_equal(got, ref), (center, byres, bymol, nointra, planar)
