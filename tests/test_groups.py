"""The product's vectorised grouping (warp_md_amd.groups) against the plain
restatement of rdf_groups / common_group_count (rdf.rs:279-401) in
oracle/grouping.py: identical groups in identical order, identical molecule
keys, residue keys that are the ranks of the reference's (chain, resid,
resname id) keys, identical common-pair counts, and a union whose merged
groups reproduce both lists."""
import itertools

import numpy as np
import pytest

import grouping
from warp_md_amd import System
from warp_md_amd import groups


def _system(n, seed, with_mol=True):
    rng = np.random.default_rng(seed)
    chain = np.sort(rng.integers(0, 3, n))
    resid = rng.integers(1, 6, n)
    names = rng.choice(["SOL", "sol", "LIG", "ALA", "Ala"], n)
    atoms = {"name": ["X"] * n, "resname": list(names), "resid": list(resid),
             "chain_id": list(chain), "mass": list(rng.uniform(1, 16, n))}
    if with_mol:
        atoms["molecule_id"] = list(rng.integers(-3, 5, n))
    return System.from_arrays(atoms)


def _check_groups(vec, ref, system):
    keys = grouping.residue_keys(system)
    rank = {k: i for i, k in enumerate(sorted(set(keys)))}
    got = vec.as_tuples()
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        assert g[0] == r[0]
        assert g[2] == r[2]
        assert g[1] == (None if r[1] is None else rank[r[1]])


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("with_mol", [True, False])
def test_groups_match_restatement(seed, with_mol):
    s = _system(40, seed, with_mol)
    rng = np.random.default_rng(seed + 10)
    sel_a = np.sort(rng.choice(40, 25, replace=False))
    sel_b = rng.choice(40, 30, replace=False)  # unsorted: order of first appearance matters
    for (c1, r1, m1), (c2, r2, m2) in itertools.product(
            [(False, False, False), (True, False, False), (False, True, False), (False, False, True),
             (False, True, True)], repeat=2):
        va = groups.rdf_groups(sel_a, s, c1, r1, m1)
        vb = groups.rdf_groups(sel_b, s, c2, r2, m2)
        ta = grouping.rdf_groups(sel_a, s, c1, r1, m1)
        tb = grouping.rdf_groups(sel_b, s, c2, r2, m2)
        _check_groups(va, ta, s)
        _check_groups(vb, tb, s)
        for nointra in (False, True):
            assert groups.common_group_count(va, vb, nointra) == grouping.common_group_count(ta, tb, nointra)
        offsets, atoms, ida, idb, mol, res = groups.union(va, vb)
        U = offsets.size - 1
        merged = [tuple(int(x) for x in atoms[offsets[u]:offsets[u + 1]]) for u in range(U)]
        assert len(set(merged)) == U
        assert [merged[i] for i in ida] == [g[0] for g in ta]
        assert [merged[i] for i in idb] == [g[0] for g in tb]
        # first-appearance order over A then B
        seen = []
        for g in ta + tb:
            if g[0] not in seen:
                seen.append(g[0])
        assert merged == seen
        np.testing.assert_array_equal(mol[ida], va.mol)
        np.testing.assert_array_equal(mol[U + idb], vb.mol)
        np.testing.assert_array_equal(res[ida], va.res)
        np.testing.assert_array_equal(res[U + idb], vb.res)


def test_long_groups_and_empty_selections():
    n = 300
    s = System.from_arrays({"name": ["X"] * n, "resname": ["BIG"] * n, "resid": [1 + i // 100 for i in range(n)],
                            "chain_id": [0] * n, "molecule_id": [i // 150 for i in range(n)]})
    sel = np.arange(n)
    for args in [(False, True, False), (False, False, True), (True, False, False)]:
        va = groups.rdf_groups(sel, s, *args)
        vb = groups.rdf_groups(sel[::-1], s, *args)
        ta = grouping.rdf_groups(sel, s, *args)
        tb = grouping.rdf_groups(sel[::-1], s, *args)
        _check_groups(va, ta, s)
        _check_groups(vb, tb, s)
        for nointra in (False, True):
            assert groups.common_group_count(va, vb, nointra) == grouping.common_group_count(ta, tb, nointra)
        offsets, atoms, ida, idb, _, _ = groups.union(va, vb)
        assert int(offsets[-1]) == atoms.size
    empty = np.zeros(0, np.int64)
    for args in [(False, False, False), (False, True, False), (False, False, True), (True, False, False)]:
        v = groups.rdf_groups(empty, s, *args)
        t = grouping.rdf_groups(empty, s, *args)
        _check_groups(v, t, s)
