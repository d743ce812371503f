"""Groups of RdfPlan's grouped modes, vectorised.

Restates rdf_groups (crates/traj-engine/src/plans/rdf.rs:279-351),
same_intramolecular_group (:368-373) and common_group_count (:387-401) over
numpy arrays, so systems of 10^5-10^6 atoms are grouped in milliseconds.  A
``Groups`` holds the groups as CSR (``atoms``, ``offsets``) with one molecule
and one residue key per group (``NO_KEY`` = none).
"""
from __future__ import annotations

import numpy as np

NO_KEY = np.int64(-(2 ** 63))


class Groups:
    """CSR list of atom groups with per-group residue / molecule keys."""

    def __init__(self, atoms, offsets, res, mol):
        self.atoms = np.ascontiguousarray(atoms, np.uint32)
        self.offsets = np.ascontiguousarray(offsets, np.int64)
        self.res = np.ascontiguousarray(res, np.int64)
        self.mol = np.ascontiguousarray(mol, np.int64)

    def __len__(self) -> int:
        return int(self.offsets.size - 1)

    def singletons(self) -> bool:
        return bool(np.all(np.diff(self.offsets) == 1))

    def as_tuples(self):
        """[(atom tuple, residue key or None, molecule key or None)]."""
        out = []
        for g in range(len(self)):
            a = tuple(int(x) for x in self.atoms[self.offsets[g]:self.offsets[g + 1]])
            r = None if self.res[g] == NO_KEY else int(self.res[g])
            m = None if self.mol[g] == NO_KEY else int(self.mol[g])
            out.append((a, r, m))
        return out


def residue_ids(system) -> np.ndarray:
    """residue_key (rdf.rs:264-277) per atom as a rank in (chain_id, resid,
    resname id) order; the resname id is the first-seen order of the upper-cased
    name (the interner), so ranks sort like the reference's BTreeMap keys."""
    names = np.char.upper(np.asarray(system.atoms["resname"], dtype=str))
    uniq, first, inv = np.unique(names, return_index=True, return_inverse=True)
    order_id = np.empty(uniq.size, np.int64)
    order_id[np.argsort(first, kind="stable")] = np.arange(uniq.size)
    rn = order_id[inv.reshape(-1)]
    chain = np.asarray(system.atoms["chain_id"], np.int64)
    resid = np.asarray(system.atoms["resid"], np.int64)
    n = chain.size
    ids = np.zeros(n, np.int64)
    if n:
        o = np.lexsort((rn, resid, chain))
        new = np.ones(n, bool)
        new[1:] = (chain[o][1:] != chain[o][:-1]) | (resid[o][1:] != resid[o][:-1]) | (rn[o][1:] != rn[o][:-1])
        ids[o] = np.cumsum(new) - 1
    return ids


def _split(atoms: np.ndarray, keys: np.ndarray):
    """Stable grouping of atoms by key, groups in ascending key order."""
    order = np.argsort(keys, kind="stable")
    ks, at = keys[order], atoms[order]
    starts = np.flatnonzero(np.r_[True, ks[1:] != ks[:-1]])
    return at, np.r_[starts, ks.size].astype(np.int64), ks[starts]


def rdf_groups(sel, system, center: bool, byres: bool, bymol: bool) -> Groups:
    """rdf_groups (rdf.rs:325-351): centre = one group of the whole selection;
    bymol = one group per molecule id (atoms without one dropped), else byres =
    one group per residue, else one group per atom (with its residue and
    molecule keys)."""
    sel = np.asarray(sel, np.int64)
    n = sel.size
    if center:
        return Groups(sel, [0, n], [NO_KEY], [NO_KEY])
    mol_ids = np.asarray(system.molecule_id, np.int64)
    if bymol and mol_ids.size:
        s = sel[sel < mol_ids.size]
        if s.size:
            at, off, ks = _split(s, mol_ids[s])
            return Groups(at, off, np.full(ks.size, NO_KEY), ks)
    rid = residue_ids(system)
    if byres and n:
        at, off, ks = _split(sel, rid[sel])
        return Groups(at, off, ks, np.full(ks.size, NO_KEY))
    mol = np.full(n, NO_KEY)
    ok = sel < mol_ids.size
    mol[ok] = mol_ids[sel[ok]]
    return Groups(sel, np.arange(n + 1), rid[sel] if n else np.zeros(0, np.int64), mol)


def _count_lookup(keys_b: np.ndarray, keys_a: np.ndarray) -> np.ndarray:
    """For every a, the number of b with the same key (NO_KEY never matches)."""
    kb = keys_b[keys_b != NO_KEY]
    if kb.size == 0 or keys_a.size == 0:
        return np.zeros(keys_a.size, np.int64)
    u, c = np.unique(kb, return_counts=True)
    pos = np.searchsorted(u, keys_a)
    pos = np.minimum(pos, u.size - 1)
    hit = (u[pos] == keys_a) & (keys_a != NO_KEY)
    return np.where(hit, c[pos], 0)


def _content_ids(ga: Groups, gb: Groups):
    """One id per distinct atom list over A then B, numbered in first-appearance
    order: (ids of the A groups, ids of the B groups, number of ids, first
    occurrence of each id in the concatenated list)."""
    na, nb = len(ga), len(gb)
    la, lb = np.diff(ga.offsets), np.diff(gb.offsets)
    lmax = int(max(la.max(initial=0), lb.max(initial=0)))
    if lmax <= 64:
        # rows of the atom lists padded with a sentinel (lengths are implied)
        rows = np.full((na + nb, max(lmax, 1)), -1, np.int64)
        for g, r0, ln in ((ga, 0, la), (gb, na, lb)):
            if len(g):
                grp = np.repeat(np.arange(len(g)), ln)
                col = np.arange(g.atoms.size) - np.repeat(g.offsets[:-1], ln)
                rows[r0 + grp, col] = g.atoms
        _, first, inv = np.unique(rows, axis=0, return_index=True, return_inverse=True)
        inv = inv.reshape(-1)
    else:
        seen, first_l, inv = {}, [], []
        for j, (g, i) in enumerate([(ga, i) for i in range(na)] + [(gb, i) for i in range(nb)]):
            k = g.atoms[g.offsets[i]:g.offsets[i + 1]].tobytes()
            if k not in seen:
                seen[k] = len(first_l)
                first_l.append(j)
            inv.append(seen[k])
        first, inv = np.asarray(first_l, np.int64), np.asarray(inv, np.int64)
    rank = np.empty(first.size, np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(first.size)
    ids = rank[inv]
    return ids[:na], ids[na:], int(first.size), np.sort(first)


def common_group_count(ga: Groups, gb: Groups, no_intramol: bool) -> int:
    """common_group_count (rdf.rs:387-401): pairs with the same molecule /
    residue (no_intramol) or with identical atom lists."""
    if no_intramol:
        # same_intramolecular_group: molecules when both have one, else residues
        a_mol, a_res = ga.mol, ga.res
        n_mol = _count_lookup(gb.mol, a_mol)
        n_res = _count_lookup(gb.res, a_res)
        n_res_mol = _count_lookup(gb.res[gb.mol != NO_KEY], a_res)
        has_mol = a_mol != NO_KEY
        return int(np.sum(np.where(has_mol, n_mol + (n_res - n_res_mol), n_res)))
    ida, idb, _, _ = _content_ids(ga, gb)
    return int(np.sum(_count_lookup(idb, ida)))


def union(ga: Groups, gb: Groups):
    """Union of the two group lists with identical atom lists merged
    (rdf.rs:819-821): (offsets [U + 1] u64, atoms u32, ids_a, ids_b u32,
    mol [2 U], res [2 U] keys per role: [u] as a sel_a group, [U + u] as sel_b)."""
    ida, idb, U, first = _content_ids(ga, gb)
    # the merged groups are the first occurrences, gathered from A ++ B as one CSR
    all_atoms = np.concatenate([ga.atoms, gb.atoms])
    all_start = np.concatenate([ga.offsets[:-1], gb.offsets[:-1] + ga.atoms.size])
    all_len = np.concatenate([np.diff(ga.offsets), np.diff(gb.offsets)])
    lens = all_len[first]
    offsets = np.zeros(U + 1, np.uint64)
    offsets[1:] = np.cumsum(lens)
    src = (np.repeat(all_start[first] - offsets[:-1].astype(np.int64), lens)
           + np.arange(int(offsets[-1]), dtype=np.int64))
    atoms = all_atoms[src].astype(np.uint32)
    mol = np.full(2 * U, NO_KEY)
    res = np.full(2 * U, NO_KEY)
    mol[ida] = ga.mol
    res[ida] = ga.res
    mol[U + idb] = gb.mol
    res[U + idb] = gb.res
    return (offsets, np.ascontiguousarray(atoms), np.ascontiguousarray(ida, np.uint32),
            np.ascontiguousarray(idb, np.uint32), mol, res)
