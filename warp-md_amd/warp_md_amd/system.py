"""System / Selection / Trajectory: the host-side data the plans consume.

Mirrors the parts of warp-md's PySystem / PySelection / PyTrajectory the RDF
and RMSD path needs (crates/traj-py/src/io.rs:104-…, 661-686, 1954-2115;
crates/traj-core/src/{system.rs:29-40, selection.rs:9-13, frame.rs:3-20}).
Selections arrive as atom-index lists; the selection grammar here is a small
subset (``all``, ``name``, ``resname``, ``resid``, ``index``, joined by ``and``)
because the full warp-md selection language is not on this path.
"""
from __future__ import annotations

from typing import Iterable, Optional, Sequence

import numpy as np


class Selection:
    """An ordered list of atom indices (``Selection{indices}``, selection.rs:9-13)."""

    def __init__(self, indices: Iterable[int], expr: str = ""):
        arr = np.asarray(list(indices) if not isinstance(indices, np.ndarray) else indices)
        self.indices = np.ascontiguousarray(arr, dtype=np.uint32)
        self.expr = expr

    def __len__(self):
        return int(self.indices.size)

    def __repr__(self):
        return f"Selection({self.expr!r}, n={len(self)})"


def _molecule_ids_from_bonds(n: int, bonds) -> list:
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for a, b in bonds:
        a, b = int(a), int(b)
        if not (0 <= a < n and 0 <= b < n):
            raise RuntimeError("invalid bond: atom index out of range")
        ra, rb = find(a), find(b)
        if ra != rb:
            parent[rb] = ra
    ids, out = {}, []
    for atom in range(n):
        root = find(atom)
        if root not in ids:
            ids[root] = len(ids) + 1
        out.append(ids[root])
    return out


class System:
    """Atom table plus optional topology coordinates ``positions0``."""

    def __init__(self, atoms: dict, positions0: Optional[np.ndarray] = None):
        names = list(atoms.get("name", []))
        n = len(names)
        self.atoms = {
            "name": [str(x) for x in names],
            "resname": [str(x) for x in atoms.get("resname", ["UNK"] * n)],
            "resid": [int(x) for x in atoms.get("resid", list(range(1, n + 1)))],
            "chain_id": [int(x) for x in atoms.get("chain_id", [0] * n)],
            "mass": [float(x) for x in atoms.get("mass", [1.0] * n)],
        }
        for key, col in self.atoms.items():
            if len(col) != n:
                raise RuntimeError(f"mismatch: atom column '{key}' has {len(col)} entries, expected {n}")
        # molecule ids (traj-core/src/system.rs:73-91): given per atom, else the
        # connected components of the bonds numbered from 1 in atom order
        # (molecule_ids_from_bonds, :226-247), else none
        mol = atoms.get("molecule_id")
        bonds = atoms.get("bonds")
        if mol is not None and len(mol):
            if len(mol) != n:
                raise RuntimeError("mismatch: atom_table.molecule_id must match atom count")
            self.molecule_id = np.asarray(mol, dtype=np.int64)
        elif bonds is not None and len(bonds) and n:
            self.molecule_id = np.asarray(_molecule_ids_from_bonds(n, bonds), dtype=np.int64)
        else:
            self.molecule_id = np.zeros(0, np.int64)
        self.positions0 = None
        if positions0 is not None:
            p = np.asarray(positions0, dtype=np.float32)
            if p.shape != (n, 3):
                raise RuntimeError("mismatch: positions0 must have shape (n_atoms, 3)")
            self.positions0 = np.ascontiguousarray(p)

    @staticmethod
    def from_arrays(atoms: dict, positions0=None) -> "System":
        return System(atoms, positions0)

    def n_atoms(self) -> int:
        return len(self.atoms["name"])

    def molecule_id_for_atom(self, atom: int):
        """system.rs:96-98: the atom's molecule id, or None."""
        return int(self.molecule_id[atom]) if atom < len(self.molecule_id) else None

    def atom_table(self) -> dict:
        return self.atoms

    def select_indices(self, indices: Sequence[int]) -> Selection:
        idx = np.asarray(list(indices), dtype=np.int64)
        if idx.size and (idx.min() < 0 or idx.max() >= self.n_atoms()):
            raise RuntimeError("invalid selection: atom index out of range")
        return Selection(idx, expr="indices")

    def select(self, expr: str) -> Selection:
        n = self.n_atoms()
        mask = np.ones(n, dtype=bool)
        for clause in [c.strip() for c in expr.split(" and ")]:
            toks = clause.split()
            if not toks:
                raise RuntimeError(f"invalid selection: empty clause in '{expr}'")
            key, vals = toks[0].lower(), toks[1:]
            if key == "all" and not vals:
                m = np.ones(n, dtype=bool)
            elif key in ("name", "resname") and vals:
                col = np.asarray(self.atoms[key])
                m = np.isin(col, vals)
            elif key in ("resid", "index") and vals:
                col = np.asarray(self.atoms["resid"]) if key == "resid" else np.arange(n)
                m = np.zeros(n, dtype=bool)
                for v in vals:
                    if "-" in v[1:]:
                        lo, hi = v.split("-", 1) if not v.startswith("-") else (v, v)
                        m |= (col >= int(lo)) & (col <= int(hi))
                    else:
                        m |= col == int(v)
            else:
                raise RuntimeError(f"invalid selection: unsupported clause '{clause}'")
            mask &= m
        return Selection(np.nonzero(mask)[0], expr=expr)


def _boxes_from(box, n_frames: int) -> Optional[np.ndarray]:
    """Box input -> [F][9] row-major box vectors.

    (3,) or (F, 3): orthorhombic lengths (io.rs:2062-2115); (3, 3) or
    (F, 3, 3): triclinic box vectors as rows (an addition of this engine).
    """
    if box is None:
        return None
    b = np.asarray(box, dtype=np.float32)
    if b.shape == (3,):
        out = np.diag(b).reshape(1, 9)
    elif b.ndim == 2 and b.shape == (n_frames, 3) and not (b.shape == (3, 3) and n_frames != 3):
        out = np.zeros((n_frames, 9), np.float32)
        out[:, 0], out[:, 4], out[:, 8] = b[:, 0], b[:, 1], b[:, 2]
    elif b.shape == (3, 3):
        out = b.reshape(1, 9)
    elif b.ndim == 3 and b.shape[1:] == (3, 3) and b.shape[0] == n_frames:
        out = b.reshape(n_frames, 9)
    else:
        raise RuntimeError("box must have shape (3,), (frames, 3), (3, 3) or (frames, 3, 3)")
    if out.shape[0] == 1 and n_frames != 1:
        out = np.repeat(out, n_frames, axis=0)
    return np.ascontiguousarray(out, dtype=np.float32)


class Trajectory:
    """In-memory trajectory (``Trajectory.from_numpy``, io.rs:669-686).

    coords: (frames, atoms, 3) or (atoms, 3), float32 or float64 (cast to f32).
    The frames are held once in the reference's FrameChunk layout,
    (frames, atoms, 4) float32 with w = 1 (io.rs:1995), so plans hand chunks
    to the engine without a copy; ``coords`` is the (frames, atoms, 3) view.
    """

    def __init__(self, coords: np.ndarray, boxes: Optional[np.ndarray], time_ps=None):
        c = np.asarray(coords, dtype=np.float32)
        if c.shape[-1] == 3:
            c4 = np.empty(c.shape[:-1] + (4,), np.float32)
            c4[..., :3] = c
            c4[..., 3] = 1.0
        else:
            c4 = np.ascontiguousarray(c)
        self.coords4 = c4
        self.boxes = boxes
        self.time_ps = time_ps
        self._cursor = 0

    @property
    def coords(self) -> np.ndarray:
        return self.coords4[..., :3]

    @staticmethod
    def from_numpy(coords, box=None, time_ps=None) -> "Trajectory":
        c = np.asarray(coords)
        if c.dtype not in (np.float32, np.float64):
            raise RuntimeError("coords must be a numpy array with shape (frames, atoms, 3) or (atoms, 3)")
        if c.ndim == 2:
            c = c[None]
        if c.ndim != 3 or c.shape[2] != 3:
            raise RuntimeError("coords must have final dimension 3")
        boxes = _boxes_from(box, c.shape[0])
        t = None if time_ps is None else np.asarray(time_ps, dtype=np.float32)
        return Trajectory(c, boxes, t)

    def n_atoms(self) -> int:
        return int(self.coords4.shape[1])

    def n_frames(self) -> int:
        return int(self.coords4.shape[0])

    def reset(self):
        self._cursor = 0

    def count_frames(self, chunk_frames=None) -> int:
        return self.n_frames()

    def subset(self, frames: Sequence[int]) -> "Trajectory":
        idx = np.asarray(list(frames), dtype=np.int64)
        boxes = None if self.boxes is None else self.boxes[idx]
        t = None if self.time_ps is None else self.time_ps[idx]
        return Trajectory(self.coords4[idx], boxes, t)

    def _advance(self, max_frames: int):
        if self._cursor >= self.n_frames():
            return None
        a = self._cursor
        b = min(self.n_frames(), a + max(1, int(max_frames)))
        self._cursor = b
        return a, b

    def read_chunk(self, max_frames: int):
        """Next (coords[f, N, 3], boxes[f, 9] or None); None at the end."""
        r = self._advance(max_frames)
        if r is None:
            return None
        a, b = r
        boxes = None if self.boxes is None else self.boxes[a:b]
        return self.coords4[a:b, :, :3], boxes

    def read_chunk4(self, max_frames: int):
        """Next (xyzw[f, N, 4] contiguous, boxes[f, 9] or None); None at the end."""
        r = self._advance(max_frames)
        if r is None:
            return None
        a, b = r
        boxes = None if self.boxes is None else self.boxes[a:b]
        return self.coords4[a:b], boxes
