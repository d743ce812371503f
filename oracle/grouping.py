"""Group construction of RdfPlan's grouped modes, restated as plain Python.

TEST INFRASTRUCTURE ONLY: used by ``tests/`` as the checker of the product's
vectorised grouping (``warp_md_amd.groups``) and to build the group lists the
oracle's ``rdf_grouped`` takes; never imported by the product package.

Groups are lists of ``(atom index tuple, residue key or None, molecule key or
None)``; a residue key is ``(chain_id, resid, resname id)``.

* ``residue_keys``               -- ``residue_key``, rdf.rs:264-277
* ``rdf_groups``                 -- rdf.rs:279-351
* ``same_intramolecular_group``  -- rdf.rs:368-373
* ``common_group_count``         -- rdf.rs:387-401, as its double loop
"""
from __future__ import annotations


def residue_keys(system):
    """(chain_id, resid, resname id) per atom; the resname id is the
    first-seen order of the upper-cased name (the reference's interner)."""
    names = {}
    keys = []
    for c, r, rn in zip(system.atoms["chain_id"], system.atoms["resid"], system.atoms["resname"]):
        u = str(rn).upper()
        if u not in names:
            names[u] = len(names)
        keys.append((int(c), int(r), names[u]))
    return keys


def _mol(system, i):
    m = system.molecule_id
    return int(m[i]) if i < len(m) else None


def rdf_groups(sel, system, center: bool, byres: bool, bymol: bool):
    """centre: one group of the whole selection; bymol: one group per
    molecule id in ascending order (atoms without one dropped; none at all
    falls through); byres: one group per residue key in ascending order; else
    one group per atom."""
    idx = [int(i) for i in sel]
    if center:
        return [(tuple(idx), None, None)]
    rkeys = residue_keys(system)
    if bymol:
        grouped = {}
        for i in idx:
            m = _mol(system, i)
            if m is not None:
                grouped.setdefault(m, []).append(i)
        if grouped:
            return [(tuple(grouped[k]), None, k) for k in sorted(grouped)]
    if byres:
        grouped = {}
        for i in idx:
            grouped.setdefault(rkeys[i], []).append(i)
        if grouped:
            return [(tuple(grouped[k]), k, None) for k in sorted(grouped)]
    return [((i,), rkeys[i], _mol(system, i)) for i in idx]


def same_intramolecular_group(a, b) -> bool:
    """Molecule keys decide when both groups have one, else residue keys."""
    if a[2] is not None and b[2] is not None:
        return a[2] == b[2]
    return a[1] is not None and b[1] is not None and a[1] == b[1]


def common_group_count(groups_a, groups_b, no_intramol: bool) -> int:
    """Pairs (a, b) that the pair loop skips: same intramolecular group
    (no_intramol) or identical atom lists."""
    if no_intramol:
        return sum(1 for a in groups_a for b in groups_b if same_intramolecular_group(a, b))
    return sum(1 for a in groups_a for b in groups_b if a[0] == b[0])
