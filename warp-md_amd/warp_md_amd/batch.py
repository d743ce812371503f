"""Uneven batches: many independent trajectories through one launch.

warp-md runs one trajectory per ``Plan.run()`` with a fresh context
(crates/traj-py/src/pack.rs:1320-1323).  ``UnevenBatch`` runs, for every
trajectory of a batch at once, what

    RdfPlan(all, all, bins, r_max, pbc).run(traj, system)       # counts, g(r)
    RmsdPlan(all, reference="frame0", align=True).run(traj, system)

return per trajectory (rdf.rs:882-1016, rmsd/plan.rs:336-392): one cell-list /
brute-force pair launch over all trajectories' frames, one RMSD launch, the
whole step replayed as a captured HIP graph (``wmd_batch_*`` in
``include/wmd_hip.h``).  Inputs live in one device store; the autotuner picks
the enumeration per class of trajectories and its choices can be saved and
loaded across processes.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

from . import _native
from ._native import check, lib, ptr
from .executor import Device
from .plans import _parse_pbc, finalize_rdf


def _boxes9(box, n_frames: int, pbc: int = _native.PBC_ORTHO) -> np.ndarray:
    """(F, 9) row-major boxes from lengths (3,), (F, 3), a (3, 3) matrix or (F, 3, 3).

    A (3, 3) array is per-frame lengths when there are three frames and the
    boxes are orthorhombic, else one box matrix shared by every frame."""
    b = np.asarray(box, np.float32)
    if b.shape == (3,):
        b = np.diag(b).reshape(1, 9)
    elif b.shape == (n_frames, 3) and pbc != _native.PBC_TRICLINIC:
        b = np.stack([np.diag(r).reshape(9) for r in b])
    elif b.shape == (3, 3):
        b = b.reshape(1, 9)
    elif b.ndim == 2 and b.shape[1] == 3:
        b = np.stack([np.diag(r).reshape(9) for r in b])
    else:
        b = b.reshape(-1, 9)
    if b.shape[0] == 1 and n_frames != 1:
        b = np.repeat(b, n_frames, 0)
    if b.shape[0] != n_frames:
        raise RuntimeError("mismatch: one box per frame required")
    return np.ascontiguousarray(b, np.float32)


class UnevenBatch:
    """RDF histograms and RMSD-to-frame-0 series of many trajectories.

    n_atoms, n_frames: per trajectory.  ``rdf`` / ``rmsd`` choose the parts;
    ``graph`` replays the step as a HIP graph (False: eager launches)."""

    def __init__(self, n_atoms: Sequence[int], n_frames: Sequence[int], bins: int = 300,
                 r_max: float = 12.0, pbc: str = "orthorhombic", rdf: bool = True,
                 rmsd: bool = True, graph: bool = True, device: str = "auto"):
        self.device = Device.from_spec(device)
        self.ctx = self.device.context()
        self.n_atoms = np.ascontiguousarray(n_atoms, np.uint32)
        self.n_frames = np.ascontiguousarray(n_frames, np.uint32)
        if self.n_atoms.shape != self.n_frames.shape or self.n_atoms.ndim != 1:
            raise RuntimeError("mismatch: n_atoms and n_frames must be 1-D of equal length")
        self.n_traj = int(self.n_atoms.size)
        self.bins = int(bins)
        self.r_max = float(np.float32(r_max))
        self.pbc = _parse_pbc(pbc)
        self.rmsd_enabled = bool(rmsd)
        flags = ((_native.BATCH_RDF if rdf else 0) | (_native.BATCH_RMSD if rmsd else 0)
                 | (0 if graph else _native.BATCH_NO_GRAPH))
        h = ctypes.c_void_p()
        check(lib().wmd_batch_create(self.ctx.handle, self.n_traj, ptr(self.n_atoms),
                                     ptr(self.n_frames), self.bins, ctypes.c_float(self.r_max),
                                     self.pbc, flags, ctypes.byref(h)))
        self.handle = h
        self.offsets = np.zeros(self.n_traj, np.uint64)
        self.strides = np.zeros(self.n_traj, np.uint32)
        total = ctypes.c_uint64()
        check(lib().wmd_batch_store_layout(h, ptr(self.offsets), ptr(self.strides),
                                           ctypes.byref(total)))
        self.store_floats = int(total.value)
        self.frame0 = np.concatenate([[0], np.cumsum(self.n_frames.astype(np.int64))])
        self.boxes = None

    # ---- inputs -----------------------------------------------------------
    def new_store(self):
        """A zeroed device store (torch float32) of the batch's layout."""
        import torch
        return torch.zeros(max(self.store_floats, 1), dtype=torch.float32,
                           device=f"cuda:{self.device.index}")

    def store_view(self, store, t: int):
        """Trajectory t's frames in the store as a (F, 3, stride) view."""
        F, s, o = int(self.n_frames[t]), int(self.strides[t]), int(self.offsets[t])
        return store[o:o + 3 * s * F].view(F, 3, s)

    def fill_store(self, store, t: int, coords) -> None:
        """Writes (F, n, 3) coordinates of trajectory t into the store."""
        import torch
        n = int(self.n_atoms[t])
        c = torch.as_tensor(np.ascontiguousarray(coords, np.float32)).to(store.device)
        if tuple(c.shape) != (int(self.n_frames[t]), n, 3):
            raise RuntimeError(f"mismatch: trajectory {t} coords must be (F, {n}, 3)")
        v = self.store_view(store, t)
        v[:, :, :n] = c.permute(0, 2, 1)
        v[:, :, n:] = 0.0

    def store_from_numpy(self, coords_list):
        store = self.new_store()
        for t, c in enumerate(coords_list):
            self.fill_store(store, t, c)
        return store

    def set_boxes(self, boxes_list) -> None:
        """One box spec per trajectory (lengths (3,), (F, 3), (3, 3) or (F, 3, 3))."""
        if self.pbc == _native.PBC_NONE:
            self.boxes = None
            check(lib().wmd_batch_set_boxes(self.handle, None))
            return
        if len(boxes_list) != self.n_traj:
            raise RuntimeError("mismatch: one box spec per trajectory required")
        parts = [_boxes9(b, int(F), self.pbc) for b, F in zip(boxes_list, self.n_frames)]
        self.boxes = (np.ascontiguousarray(np.concatenate(parts)) if parts
                      else np.zeros((0, 9), np.float32))
        check(lib().wmd_batch_set_boxes(self.handle, ptr(self.boxes)))

    # ---- tuning -------------------------------------------------------------
    def set_tuning(self, cell_k: Optional[Sequence[int]] = None,
                   home_block: Optional[Sequence[int]] = None,
                   grab: Optional[Sequence[int]] = None) -> None:
        """Per-trajectory overrides (None or 0 = default): enumeration (-1 brute
        force, 1..3 cells per r_max), brute-force home atoms per unit (a
        multiple of 64 up to 512) and units per queue task (1..64)."""
        arr = lambda v: None if v is None else np.ascontiguousarray(v, np.int32)
        k, bh, g = arr(cell_k), arr(home_block), arr(grab)
        check(lib().wmd_batch_set_tuning(self.handle, ptr(k), ptr(bh), ptr(g)))

    def autotune(self, store, reps: int = 2, cache: Optional[str] = None) -> None:
        """Times the candidate enumerations per trajectory class and applies the
        fastest; ``cache`` (a path) is read before and written after."""
        if cache and os.path.exists(cache):
            load_tuning(cache)
        check(lib().wmd_batch_autotune(self.handle, ctypes.c_void_p(store.data_ptr()), int(reps),
                                       None))
        if cache:
            save_tuning(cache)

    def plan(self):
        """(cell_k per trajectory with 0 = brute force, number of pair units)."""
        k = np.zeros(self.n_traj, np.int32)
        u = ctypes.c_longlong()
        check(lib().wmd_batch_plan_info(self.handle, ptr(k), None, None, ctypes.byref(u)))
        return k, int(u.value)

    def plan_tuning(self):
        """Per trajectory: (cell_k with 0 = brute force, brute-force home block
        (0 for cells), units per task)."""
        k = np.zeros(self.n_traj, np.int32)
        bh = np.zeros(self.n_traj, np.int32)
        g = np.zeros(self.n_traj, np.int32)
        check(lib().wmd_batch_plan_info(self.handle, ptr(k), ptr(bh), ptr(g), None))
        return k, bh, g

    def slice_times(self) -> np.ndarray:
        """[8, 4] per queue slice of the last step's pair kernel: first wave
        start, slice drained, last wave end (ms from the earliest start) and
        the slice's tasks taken by waves of other slices."""
        out = np.zeros((8, 4), np.float64)
        check(lib().wmd_batch_slice_times(self.handle, ptr(out)))
        return out

    # ---- runs ---------------------------------------------------------------
    def run(self, store, rmsd_out=None, rmsd64_out=None, err_out=None, stream=None):
        """One step on the null stream (or ``stream``, a hipStream_t as int):
        adds every trajectory's counts, writes the RMSD of every frame (torch
        f32 [sum F], allocated when not given).  Returns the RMSD tensor."""
        import torch
        if self.rmsd_enabled and rmsd_out is None:
            rmsd_out = torch.zeros(max(int(self.frame0[-1]), 1), dtype=torch.float32,
                                   device=store.device)
        V = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
        check(lib().wmd_batch_run_device(self.handle, V(store), V(rmsd_out), V(rmsd64_out),
                                         V(err_out), None if stream is None else ctypes.c_void_p(stream)))
        return rmsd_out

    def rmsd_of(self, rmsd, t: int):
        return rmsd[int(self.frame0[t]):int(self.frame0[t + 1])]

    def counts(self) -> np.ndarray:
        out = np.zeros((self.n_traj, self.bins), np.uint64)
        check(lib().wmd_batch_read_counts(self.handle, ptr(out)))
        return out

    def reset(self) -> None:
        check(lib().wmd_batch_reset(self.handle))

    def rdf(self, t: int, counts=None, density: float = 0.033456, volume: bool = False,
            raw_rdf: bool = False, number_density: bool = False, intrdf: bool = False):
        """RdfPlan.finalize (rdf.rs:943-1016) of trajectory t: (r, g, counts[, integral])."""
        c = self.counts()[t] if counts is None else counts
        n = float(self.n_atoms[t])
        frames = int(self.n_frames[t])
        volume_sum = 0.0
        if self.boxes is not None:
            for m in self.boxes[self.frame0[t]:self.frame0[t + 1]]:
                if self.pbc == _native.PBC_TRICLINIC:
                    volume_sum += abs(float(np.linalg.det(m.astype(np.float64).reshape(3, 3))))
                else:
                    volume_sum += float(m[0]) * float(m[4]) * float(m[8])
        return finalize_rdf(c, self.r_max, n, n, n, frames, volume_sum, density, volume, raw_rdf,
                            number_density, intrdf)

    def __del__(self):
        h = getattr(self, "handle", None)
        nat = globals().get("_native")  # None while the interpreter shuts down
        if h is not None and nat is not None and nat._lib is not None:
            nat._lib.wmd_batch_destroy(h)
            self.handle = None


def export_tuning() -> str:
    n = ctypes.c_size_t()
    check(lib().wmd_autotune_export(None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    check(lib().wmd_autotune_export(buf, n.value + 1, ctypes.byref(n)))
    return buf.value.decode()


def import_tuning(text: str) -> None:
    check(lib().wmd_autotune_import(text.encode()))


def save_tuning(path: str) -> None:
    """Writes the autotuner's choices (one per line) to ``path``."""
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    with open(path, "w") as fh:
        fh.write(export_tuning())


def load_tuning(path: str) -> None:
    with open(path) as fh:
        import_tuning(fh.read())
