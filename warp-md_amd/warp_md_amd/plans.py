"""RdfPlan, RmsdPlan and their neighbours (PairDistPlan, StructureFactorPlan,
PairwiseRmsdPlan, AlignPlan, SuperposePlan) with warp-md's Python signatures, running on the gfx950
engine through its C ABI.

RdfPlan mirrors PyRdfPlan (crates/traj-py/src/analysis.rs:2298-2399) over
RdfPlan (crates/traj-engine/src/plans/rdf.rs:141-236, 653-1016); RmsdPlan
mirrors PyRmsdPlan (crates/traj-py/src/io.rs:2617-2659) over RmsdPlan
(crates/traj-engine/src/plans/rmsd/plan.rs:164-409).
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _native, groups
from ._native import check, lib, ptr
from .executor import Device, resolve_chunk_frames, select_frames, xyzw
from .system import Selection, System, Trajectory


def _parse_pbc(value: str) -> int:
    # traj-py pack.rs:2665-2673 parse_pbc, plus the triclinic mode of this engine
    table = {"orthorhombic": _native.PBC_ORTHO, "none": _native.PBC_NONE,
             "triclinic": _native.PBC_TRICLINIC}
    if value not in table:
        raise RuntimeError("pbc must be 'orthorhombic', 'triclinic' or 'none'")
    return table[value]


def _parse_dimension(value: str) -> str:
    # analysis.rs:2303-2309 parse_rdf_dimension
    v = str(value).lower()
    if v in ("3d", "xyz"):
        return "3d"
    if v == "xy":
        return "xy"
    raise ValueError("rdf dimension must be '3d' or 'xy'")


def _indices(sel) -> np.ndarray:
    if isinstance(sel, Selection):
        return sel.indices
    return np.ascontiguousarray(np.asarray(sel, dtype=np.int64).astype(np.uint32))


class RdfPlan:
    """Pair-distance histogram g(r) between two selections."""

    def __init__(self, sel_a, sel_b, bins: int, r_max: float, pbc: str = "orthorhombic",
                 center1=False, center2=False, byres1=False, byres2=False, bymol1=False,
                 bymol2=False, no_intramol=False, mass_weighted=True, density=0.033456,
                 volume=False, raw_rdf=False, intrdf=False, dimension="3d",
                 number_density=False, edge_diagnostics=False):
        self.pbc = _parse_pbc(pbc)
        self.dimension = _parse_dimension(dimension)
        self.sel_a = _indices(sel_a)
        self.sel_b = _indices(sel_b)
        self.bins = int(bins)
        if self.bins < 0:
            raise OverflowError("bins must be non-negative")
        self.r_max = float(np.float32(r_max))
        self.grouped = any([center1, center2, byres1, byres2, bymol1, bymol2, no_intramol])
        self.modes = (bool(center1), bool(center2), bool(byres1), bool(byres2), bool(bymol1),
                      bool(bymol2), bool(no_intramol))
        self.mass_weighted = bool(mass_weighted)
        self.density = float(density)
        self.use_volume = bool(volume)
        self.raw = bool(raw_rdf)
        self.intrdf = bool(intrdf)
        self.number_density = bool(number_density)
        self.edge_diagnostics = bool(edge_diagnostics)
        self.last_edge_pairs = 0
        self.last_path = None
        self._cell_k = 0
        self._grab = 0

    # rdf.rs:378-385 common_atom_count
    def _common_count(self) -> int:
        # entries of sel_b (with repeats) whose atom is in sel_a
        return int(np.count_nonzero(np.isin(self.sel_b, self.sel_a)))

    def set_tuning(self, cell_k: int = 0, grab: int = 0):
        """Cell-list shape (cells per r_max, 1..3) and cell units per queue
        grab (1..64); 0 = automatic."""
        self._cell_k = int(cell_k)
        self._grab = int(grab)

    def run(self, traj: Trajectory, system: System, chunk_frames: Optional[int] = None,
            device: str = "auto", frame_indices: Optional[Sequence[int]] = None):
        planar = self.dimension == "xy"
        if (self.grouped or planar) and self.pbc == _native.PBC_TRICLINIC:
            raise RuntimeError("unsupported: grouped and xy RDF modes need pbc none or orthorhombic")
        dev = Device.from_spec(device)
        if traj.n_atoms() != system.n_atoms():
            raise RuntimeError("mismatch: trajectory atom count does not match system")
        # rdf.rs:663-667
        if not np.isfinite(self.density) or self.density <= 0.0:
            raise RuntimeError("parse error: rdf density must be finite and positive")
        ctx = dev.context()
        n_atoms = system.n_atoms()
        for sel in (self.sel_a, self.sel_b):
            if sel.size and int(sel.max()) >= n_atoms:
                raise RuntimeError("invalid selection: atom index out of range")
        flags = _native.RDF_EDGE_DIAGNOSTICS if self.edge_diagnostics else 0
        if planar:
            flags |= _native.RDF_PLANAR_XY
        h = ctypes.c_void_p()
        n_a, n_b, common = float(self.sel_a.size), float(self.sel_b.size), None
        if self.grouped:
            if self.edge_diagnostics:
                raise RuntimeError("unsupported: edge diagnostics in grouped RDF modes")
            c1, c2, r1, r2, m1, m2, nointra = self.modes
            ga = groups.rdf_groups(self.sel_a, system, c1, r1, m1)
            gb = groups.rdf_groups(self.sel_b, system, c2, r2, m2)
            n_a, n_b = float(len(ga)), float(len(gb))
            common = float(groups.common_group_count(ga, gb, nointra))
            self._create_grouped(ctx, system, ga, gb, nointra, flags, h)
        else:
            check(lib().wmd_rdf_create(ctx.handle, ptr(self.sel_a), self.sel_a.size, ptr(self.sel_b),
                                       self.sel_b.size, self.bins, ctypes.c_float(self.r_max),
                                       self.pbc, flags, ctypes.byref(h)))
        try:
            if self._cell_k or self._grab:
                check(lib().wmd_rdf_set_tuning(h, self._cell_k, self._grab))
            tr = select_frames(traj, frame_indices)
            tr.reset()
            frames = 0
            volume_sum = 0.0
            step = resolve_chunk_frames(chunk_frames, max(1, n_atoms))
            need_box = self.pbc != _native.PBC_NONE or self.use_volume
            while True:
                chunk = tr.read_chunk4(step)
                if chunk is None:
                    break
                coords, boxes = chunk
                nf = coords.shape[0]
                if need_box:
                    if boxes is None:
                        raise RuntimeError(
                            "mismatch: RDF requires orthorhombic box for PBC or volume normalization")
                    if self.pbc == _native.PBC_TRICLINIC:
                        for m in boxes:
                            volume_sum += abs(float(np.linalg.det(m.astype(np.float64).reshape(3, 3))))
                    else:
                        if np.any(boxes[:, [1, 2, 3, 5, 6, 7]] != 0) or np.any(boxes[:, [0, 4, 8]] <= 0):
                            raise RuntimeError("mismatch: RDF requires orthorhombic box for PBC or volume normalization")
                        # rdf_box_measure (rdf.rs:480-485): f64 product of the f32
                        # lengths, added frame by frame in order
                        lens = boxes[:, [0, 4, 8]].astype(np.float64)
                        vols = lens[:, 0] * lens[:, 1] if planar else lens[:, 0] * lens[:, 1] * lens[:, 2]
                        for v in vols.tolist():
                            volume_sum += v
                frames += nf
                # the full chunk; the engine gathers the IO selection on the device
                check(lib().wmd_rdf_accumulate(h, ptr(coords), n_atoms, nf, ptr(boxes)))
            counts = np.zeros(self.bins, np.uint64)
            edge = np.zeros(1, np.uint64)
            check(lib().wmd_rdf_read(h, ptr(counts), ptr(edge)))
            path = ctypes.c_int()
            check(lib().wmd_rdf_last_path(h, ctypes.byref(path), None))
            self.last_path = {0: None, 1: "brute", 2: "cells"}[path.value]
            self.last_edge_pairs = int(edge[0])
        finally:
            lib().wmd_rdf_destroy(h)
        if common is None:
            common = float(self._common_count())
        return finalize_rdf(counts, self.r_max, n_a, n_b, common, frames, volume_sum, self.density,
                            self.use_volume, self.raw, self.number_density, self.intrdf,
                            planar=planar)

    def _create_grouped(self, ctx, system, ga, gb, no_intramol, flags, h):
        """Union of the two group lists (identical atom lists merged,
        rdf.rs:819-821) as CSR, per-role molecule / residue keys, and
        wmd_rdf_create_groups."""
        offsets, atoms, ida, idb, mol, res = groups.union(ga, gb)
        n_groups = offsets.size - 1
        masses = None
        if self.mass_weighted:
            mass = np.asarray(system.atoms["mass"], np.float32)
            masses = np.ascontiguousarray(mass[atoms.astype(np.int64)]) if atoms.size else np.zeros(1, np.float32)
        if no_intramol:
            flags |= _native.RDF_NO_INTRAMOL
        else:
            mol = res = None
        check(lib().wmd_rdf_create_groups(ctx.handle, n_groups, ptr(offsets), ptr(atoms), ptr(masses),
                                          ptr(ida), ida.size, ptr(idb), idb.size, ptr(mol), ptr(res),
                                          self.bins, ctypes.c_float(self.r_max), self.pbc, flags,
                                          ctypes.byref(h)))


def finalize_rdf(counts, r_max, n_a, n_b, common, frames, volume_sum, density, use_volume,
                 raw, number_density, intrdf, planar=False):
    """RdfPlan::finalize (rdf.rs:943-1016) through the engine's host routine."""
    counts = np.ascontiguousarray(counts, dtype=np.uint64)
    bins = counts.size
    r = np.zeros(bins, np.float32)
    g = np.zeros(bins, np.float32)
    it = np.zeros(bins, np.float32)
    flags = ((_native.RDF_USE_VOLUME if use_volume else 0) | (_native.RDF_RAW if raw else 0)
             | (_native.RDF_NUMBER_DENSITY if number_density else 0)
             | (_native.RDF_INTEGRAL if intrdf else 0)
             | (_native.RDF_FINALIZE_XY if planar else 0))
    if bins:
        check(lib().wmd_rdf_finalize(ptr(counts), bins, ctypes.c_float(r_max), n_a, n_b, common,
                                     int(frames), float(volume_sum), float(density), flags,
                                     ptr(r), ptr(g), ptr(it)))
    if intrdf:
        return r, g, counts, it
    return r, g, counts


def _parse_reference(value: str) -> str:
    # traj-py pack.rs:2627-2635
    if value not in ("topology", "frame0"):
        raise RuntimeError("reference must be 'topology' or 'frame0'")
    return value


class RmsdPlan:
    """Per-frame RMSD to a reference after optimal superposition.

    ``fit_selection`` (an addition of this engine) fits on one selection and
    takes the RMSD over ``selection`` -- the composition of kabsch_rotation
    and the residual of rms.rs:643-676; with ``align=False`` that composition
    takes the unfitted residual over ``selection`` and leaves the fit set
    unused (rms.rs:664-666).  ``None`` keeps RmsdPlan's behaviour (fit and
    RMSD over the same atoms).
    """

    def __init__(self, selection, reference: str = "topology", align: bool = True,
                 fit_selection=None):
        self.selection = _indices(selection)
        self.reference_mode = _parse_reference(reference)
        self.align = bool(align)
        self.fit_selection = None if fit_selection is None else _indices(fit_selection)
        self.last_error_bound = None

    def run(self, traj: Trajectory, system: System, chunk_frames: Optional[int] = None,
            device: str = "auto", frame_indices: Optional[Sequence[int]] = None):
        dev = Device.from_spec(device)
        if traj.n_atoms() != system.n_atoms():
            raise RuntimeError("mismatch: trajectory atom count does not match system")
        ctx = dev.context()
        n_atoms = system.n_atoms()
        fit = self.fit_selection if self.fit_selection is not None else self.selection
        for sel in (self.selection, fit):
            if sel.size and int(sel.max()) >= n_atoms:
                raise RuntimeError("invalid selection: atom index out of range")
        tr = select_frames(traj, frame_indices)
        tr.reset()
        if self.reference_mode == "topology":
            if system.positions0 is None:
                raise RuntimeError("mismatch: no topology reference coords")
            ref4 = xyzw(np.ascontiguousarray(system.positions0, dtype=np.float32)[None])[0]
        else:
            if tr.n_frames() == 0:
                return np.zeros(0, np.float32)
            ref4 = np.ascontiguousarray(tr.coords4[0])  # rmsd/plan.rs:336-354 first frame seen
        # atom-indexed reference and full chunks: the engine gathers the
        # selections on the device (the executor's selected read, rmsd/plan.rs:224-230)
        h = ctypes.c_void_p()
        if self.fit_selection is None:
            check(lib().wmd_rmsd_create(ctx.handle, ptr(self.selection), self.selection.size, None, 0,
                                        ptr(ref4), n_atoms, int(self.align), ctypes.byref(h)))
        else:
            check(lib().wmd_rmsd_create(ctx.handle, ptr(fit), fit.size, ptr(self.selection),
                                        self.selection.size, ptr(ref4), n_atoms, int(self.align),
                                        ctypes.byref(h)))
        try:
            out = np.zeros(tr.n_frames(), np.float32)
            err = np.zeros(tr.n_frames(), np.float64)
            step = resolve_chunk_frames(chunk_frames, max(1, n_atoms))
            f0 = 0
            while True:
                chunk = tr.read_chunk4(step)
                if chunk is None:
                    break
                x4, _ = chunk
                nf = x4.shape[0]
                o = out[f0:f0 + nf]
                e = err[f0:f0 + nf]
                check(lib().wmd_rmsd_run(h, ptr(x4), n_atoms, nf, ptr(o), ptr(e)))
                f0 += nf
        finally:
            lib().wmd_rmsd_destroy(h)
        self.last_error_bound = err
        return out


class PairDistPlan:
    """Pair-distance histogram between two selections.

    Mirrors PyPairDistPlan (crates/traj-py/src/analysis.rs:2427-2490) over
    PairDistPlan (crates/traj-engine/src/plans/rdf.rs:541-577, 1019-1287), whose
    GPU branch uses the same rdf_accum boundary as RdfPlan.  The pair set is
    the reference CPU path's: ordered pairs (a, b) of sel_a x sel_b, skipping
    a == b only when both selections are the same object (Arc::ptr_eq,
    rdf.rs:1140-1142; ``sel_a is sel_b`` here); with ``unique_pairs`` only the
    pairs whose position in sel_b exceeds the position in sel_a (rdf.rs:1136).
    The engine counts the pairs of distinct atoms; pairs of an atom with itself
    lie at r = 0 and are added to bin 0 on the host.
    """

    def __init__(self, sel_a, sel_b, bins: int, r_max: float, pbc: str = "orthorhombic",
                 output_distribution: bool = False, unique_pairs: bool = False,
                 compact_output: bool = False):
        self.pbc = _parse_pbc(pbc)
        self.same_sel = sel_a is sel_b
        self.sel_a = _indices(sel_a)
        self.sel_b = _indices(sel_b)
        self.bins = int(bins)
        if self.bins < 0:
            raise OverflowError("bins must be non-negative")
        self.r_max = float(np.float32(r_max))
        self.output_distribution = bool(output_distribution)
        self.unique_pairs = bool(unique_pairs)
        self.compact_output = bool(compact_output)
        self.last_path = None

    def _self_pairs_per_frame(self) -> int:
        """Pairs of an atom with itself in one frame (all at r = 0)."""
        if self.unique_pairs:
            # positions i < j holding the same atom
            _, m = np.unique(self.sel_a, return_counts=True)
            return int(np.sum(m * (m - 1) // 2))
        if self.same_sel:
            return 0
        ua, ma = np.unique(self.sel_a, return_counts=True)
        ub, mb = np.unique(self.sel_b, return_counts=True)
        _, ia, ib = np.intersect1d(ua, ub, assume_unique=True, return_indices=True)
        return int(np.sum(ma[ia].astype(np.int64) * mb[ib].astype(np.int64)))

    def run(self, traj: Trajectory, system: System, chunk_frames: Optional[int] = None,
            device: str = "auto", frame_indices: Optional[Sequence[int]] = None):
        dev = Device.from_spec(device)
        if traj.n_atoms() != system.n_atoms():
            raise RuntimeError("mismatch: trajectory atom count does not match system")
        if self.unique_pairs and not np.array_equal(self.sel_a, self.sel_b):
            raise RuntimeError("unsupported: unique_pairs over two different selections runs on "
                               "warp-md's CPU path")
        ctx = dev.context()
        n_atoms = system.n_atoms()
        for sel in (self.sel_a, self.sel_b):
            if sel.size and int(sel.max()) >= n_atoms:
                raise RuntimeError("invalid selection: atom index out of range")
        bins = self.bins
        w = np.float32(self.r_max) / np.float32(bins) if bins else np.float32(0.0)
        h = ctypes.c_void_p()
        check(lib().wmd_rdf_create(ctx.handle, ptr(self.sel_a), self.sel_a.size, ptr(self.sel_b),
                                   self.sel_b.size, bins, ctypes.c_float(self.r_max), self.pbc, 0,
                                   ctypes.byref(h)))
        self_pairs = self._self_pairs_per_frame()
        counts = np.zeros(bins, np.uint64)
        sums = np.zeros(bins, np.float64)
        sumsq = np.zeros(bins, np.float64)
        frames = 0
        try:
            tr = select_frames(traj, frame_indices)
            tr.reset()
            step = 1 if self.output_distribution else resolve_chunk_frames(chunk_frames, max(1, n_atoms))
            while True:
                chunk = tr.read_chunk4(step)
                if chunk is None:
                    break
                coords, boxes = chunk
                nf = coords.shape[0]
                if self.pbc != _native.PBC_NONE:
                    if boxes is None:
                        raise RuntimeError("mismatch: pairdist requires orthorhombic box for PBC")
                    if self.pbc == _native.PBC_ORTHO:
                        for m in boxes:
                            if np.any(m[[1, 2, 3, 5, 6, 7]] != 0) or np.any(m[[0, 4, 8]] <= 0):
                                raise RuntimeError("mismatch: pairdist requires orthorhombic box for PBC")
                check(lib().wmd_rdf_accumulate(h, ptr(coords), n_atoms, nf,
                                               ptr(boxes if self.pbc != _native.PBC_NONE else None)))
                frames += nf
                if self.output_distribution:
                    fc = np.zeros(bins, np.uint64)
                    check(lib().wmd_rdf_read(h, ptr(fc), None))
                    check(lib().wmd_rdf_reset(h))
                    fc = self._pair_counts(fc, 1, self_pairs)
                    # accumulate_pairdist_frame (rdf.rs:420-433)
                    counts += fc
                    sums += fc.astype(np.float64)
                    sumsq += fc.astype(np.float64) ** 2
            if not self.output_distribution:
                raw = np.zeros(bins, np.uint64)
                check(lib().wmd_rdf_read(h, ptr(raw), None))
                counts = self._pair_counts(raw, frames, self_pairs)
            path = ctypes.c_int()
            check(lib().wmd_rdf_last_path(h, ctypes.byref(path), None))
            self.last_path = {0: None, 1: "brute", 2: "cells"}[path.value]
        finally:
            lib().wmd_rdf_destroy(h)
        if not self.output_distribution:
            return pairdist_centers(bins, w), counts
        return pairdist_stats(counts, sums, sumsq, frames, w, self.compact_output)

    def _pair_counts(self, raw, frames, self_pairs):
        """Engine counts (ordered pairs of distinct atoms) -> the plan's pair counts."""
        c = np.asarray(raw, np.uint64).copy()
        if self.unique_pairs:
            c //= np.uint64(2)  # each unordered pair of distinct atoms once
        if self.bins and self_pairs:
            c[0] += np.uint64(self_pairs * frames)
        return c


class StructureFactorPlan:
    """g(r) and S(q) of one selection.

    Mirrors PyStructureFactorPlan (crates/traj-py/src/correlation.rs:2638-2684)
    over StructureFactorPlan (crates/traj-engine/src/plans/analysis/
    structure_factor.rs).  The pair counts follow the reference's CPU path
    bit for bit: distances in f64 from the f32 differences scaled by
    ``length_scale`` (``wmd_rdf_create_f64``); the normalisation and the sine
    transform are ``wmd_sf_finalize`` (structure_factor.rs:200-258).
    Returns (r, g_r, q, s_q) float32, like structure_factor_to_py
    (traj-py pack.rs:2236-2245).
    """

    def __init__(self, selection, bins: int, r_max: float, q_bins: int, q_max: float,
                 pbc: str = "orthorhombic", length_scale: Optional[float] = None):
        if pbc not in ("orthorhombic", "none"):
            raise RuntimeError("pbc must be 'orthorhombic' or 'none'")  # pack.rs:2665-2673
        self.pbc = _parse_pbc(pbc)
        self.selection = _indices(selection)
        self.bins = int(bins)
        self.q_bins = int(q_bins)
        if self.bins < 0 or self.q_bins < 0:
            raise OverflowError("bins must be non-negative")
        self.r_max = float(r_max)
        self.q_max = float(q_max)
        self.length_scale = 1.0 if length_scale is None else float(length_scale)
        self.last_counts = None

    def run(self, traj: Trajectory, system: System, chunk_frames: Optional[int] = None,
            device: str = "auto"):
        dev = Device.from_spec(device)
        if traj.n_atoms() != system.n_atoms():
            raise RuntimeError("mismatch: trajectory atom count does not match system")
        if self.bins == 0 or self.q_bins == 0:
            # structure_factor.rs:71-75
            raise RuntimeError("invalid selection: structure factor requires bins and q_bins > 0")
        ctx = dev.context()
        n_atoms = system.n_atoms()
        sel = self.selection
        if sel.size and int(sel.max()) >= n_atoms:
            raise RuntimeError("invalid selection: atom index out of range")
        h = ctypes.c_void_p()
        check(lib().wmd_rdf_create_f64(ctx.handle, ptr(sel), sel.size, self.bins,
                                       ctypes.c_double(self.r_max), self.pbc,
                                       ctypes.c_double(self.length_scale), ctypes.byref(h)))
        frames = 0
        volume_sum = 0.0
        s3 = self.length_scale * self.length_scale * self.length_scale  # powi(3)
        try:
            tr = select_frames(traj, None)
            tr.reset()
            step = resolve_chunk_frames(chunk_frames, max(1, n_atoms))
            while True:
                chunk = tr.read_chunk4(step)
                if chunk is None:
                    break
                x4, boxes = chunk
                nf = x4.shape[0]
                if self.pbc == _native.PBC_ORTHO:
                    if boxes is None:
                        raise RuntimeError("mismatch: structure factor requires orthorhombic box for PBC")
                    for m in boxes:
                        if np.any(m[[1, 2, 3, 5, 6, 7]] != 0):
                            raise RuntimeError("mismatch: structure factor requires orthorhombic box for PBC")
                        volume_sum += (float(m[0]) * float(m[4]) * float(m[8])) * s3
                frames += nf
                check(lib().wmd_rdf_accumulate(h, ptr(x4), n_atoms, nf,
                                               ptr(boxes if self.pbc != _native.PBC_NONE else None)))
            counts = np.zeros(self.bins, np.uint64)
            check(lib().wmd_rdf_read(h, ptr(counts), None))
        finally:
            lib().wmd_rdf_destroy(h)
        self.last_counts = counts
        r = np.zeros(self.bins, np.float32)
        g = np.zeros(self.bins, np.float32)
        q = np.zeros(self.q_bins, np.float32)
        sq = np.zeros(self.q_bins, np.float32)
        check(lib().wmd_sf_finalize(ptr(counts), self.bins, ctypes.c_double(self.r_max),
                                    ctypes.c_double(float(sel.size)), frames,
                                    ctypes.c_double(volume_sum), self.pbc, self.q_bins,
                                    ctypes.c_double(self.q_max), ptr(r), ptr(g), ptr(q), ptr(sq)))
        return r, g, q, sq


_PAIRWISE_METRICS = {"rms": 0, "nofit": 1, "dme": 2}


class PairwiseRmsdPlan:
    """F x F RMSD matrix between every two frames of one selection.

    Mirrors PyPairwiseRmsdPlan (crates/traj-py/src/io.rs:2759-2800) over
    PairwiseRmsdPlan (crates/traj-engine/src/plans/rmsd/plan.rs:504-675):
    metric "rms" (after optimal superposition), "nofit", or "dme" (the RMS
    difference of the frames' pair-distance vectors, with the orthorhombic
    minimum image when pbc="orthorhombic", plan.rs:612-650).  Frames go to a
    device store as they are read; "rms" / "nofit" are one FP64 Gram product
    plus per-pair QCP, "dme" blocks of f64 pair distances and frame-pair sums
    (``wmd_pairwise_*``).  Returns a (F, F) float32 array.
    """

    def __init__(self, selection, metric: str = "rms", pbc: str = "none"):
        if metric not in _PAIRWISE_METRICS:
            raise RuntimeError("metric must be 'rms', 'nofit', or 'dme'")  # pack.rs:2675-2684
        if pbc not in ("orthorhombic", "none"):  # parse_pbc, pack.rs:2665-2673
            raise RuntimeError("pbc must be 'orthorhombic' or 'none'")
        self.metric = metric
        self.pbc = _parse_pbc(pbc)
        self.selection = _indices(selection)

    def run(self, traj: Trajectory, system: System, chunk_frames: Optional[int] = None,
            device: str = "auto", frame_indices: Optional[Sequence[int]] = None):
        dev = Device.from_spec(device)
        if traj.n_atoms() != system.n_atoms():
            raise RuntimeError("mismatch: trajectory atom count does not match system")
        ctx = dev.context()
        n_atoms = system.n_atoms()
        sel = self.selection
        if sel.size and int(sel.max()) >= n_atoms:
            raise RuntimeError("invalid selection: atom index out of range")
        h = ctypes.c_void_p()
        dme_pbc = self.metric == "dme" and self.pbc == _native.PBC_ORTHO
        if self.metric == "dme":
            check(lib().wmd_pairwise_create_dme(ctx.handle, ptr(sel), sel.size, self.pbc,
                                                ctypes.byref(h)))
        else:
            check(lib().wmd_pairwise_create(ctx.handle, ptr(sel), sel.size,
                                            _PAIRWISE_METRICS[self.metric], ctypes.byref(h)))
        try:
            tr = select_frames(traj, frame_indices)
            tr.reset()
            step = resolve_chunk_frames(chunk_frames, max(1, n_atoms))
            while True:
                chunk = tr.read_chunk4(step)
                if chunk is None:
                    break
                x4, boxes = chunk
                if dme_pbc:
                    if boxes is None:  # box_lengths of Box3::None (pbc_math.rs:11-15)
                        raise RuntimeError("mismatch: orthorhombic box required for PBC")
                    check(lib().wmd_pairwise_append_boxes(h, ptr(x4), n_atoms, x4.shape[0],
                                                          ptr(np.ascontiguousarray(boxes, np.float32))))
                else:
                    check(lib().wmd_pairwise_append(h, ptr(x4), n_atoms, x4.shape[0]))
            nf = ctypes.c_size_t()
            check(lib().wmd_pairwise_frames(h, ctypes.byref(nf)))
            out = np.zeros((nf.value, nf.value), np.float32)
            check(lib().wmd_pairwise_finalize(h, ptr(out)))
        finally:
            lib().wmd_pairwise_destroy(h)
        return out


class _SuperposeBase:
    """Shared state of AlignPlan and SuperposePlan (align/plan_exec.rs:3-43)."""

    def __init__(self, selection, reference: str = "topology", mass: bool = False,
                 norotate: bool = False):
        self.selection = _indices(selection)
        self.reference_mode = _parse_reference(reference)
        self.mass = bool(mass)
        self.norotate = bool(norotate)

    def _run(self, traj: Trajectory, system: System, chunk_frames, device, frame_indices,
             superpose: bool):
        dev = Device.from_spec(device)
        if traj.n_atoms() != system.n_atoms():
            raise RuntimeError("mismatch: trajectory atom count does not match system")
        ctx = dev.context()
        n_atoms = system.n_atoms()
        sel = self.selection
        if sel.size and int(sel.max()) >= n_atoms:
            raise RuntimeError("invalid selection: atom index out of range")
        masses = None
        if self.mass:
            masses = np.ascontiguousarray(np.asarray(system.atom_table()["mass"], np.float32)[sel])
        ref4 = None
        if self.reference_mode == "topology":
            if system.positions0 is None:
                raise RuntimeError("mismatch: no topology reference coords")  # plan_exec.rs:84-87
            ref4 = xyzw(np.ascontiguousarray(system.positions0, np.float32)[None])[0]
        h = ctypes.c_void_p()
        check(lib().wmd_align_create(ctx.handle, ptr(sel), sel.size, ptr(masses), ptr(ref4),
                                     n_atoms if ref4 is not None else 0, int(self.norotate),
                                     ctypes.byref(h)))
        rts, xyzs = [], []
        try:
            tr = select_frames(traj, frame_indices)
            tr.reset()
            step = resolve_chunk_frames(chunk_frames, max(1, n_atoms))
            first = True
            while True:
                chunk = tr.read_chunk4(step)
                if chunk is None:
                    break
                x4, _ = chunk
                nf = x4.shape[0]
                if first and ref4 is None:
                    # frame0: the first frame read is the reference (plan_exec.rs:126-139)
                    f0 = np.ascontiguousarray(x4[0])
                    check(lib().wmd_align_set_reference(h, ptr(f0), n_atoms))
                first = False
                rt = np.zeros((nf, 12), np.float32)
                xyz = np.zeros((nf, n_atoms, 3), np.float32) if superpose else None
                check(lib().wmd_align_run(h, ptr(x4), n_atoms, nf, ptr(rt), ptr(xyz)))
                rts.append(rt)
                if superpose:
                    xyzs.append(xyz)
        finally:
            lib().wmd_align_destroy(h)
        if superpose:
            return (np.concatenate(xyzs).reshape(-1) if xyzs else np.zeros(0, np.float32))
        return np.concatenate(rts) if rts else np.zeros((0, 12), np.float32)


class AlignPlan(_SuperposeBase):
    """Per-frame superposition transform of a selection onto a reference.

    Mirrors PyAlignPlan (crates/traj-py/src/transform.rs:1798-1836) over
    AlignPlan (crates/traj-engine/src/plans/geometry/align/plan_exec.rs:76-255):
    returns (F, 12) float32 rows R (row-major) then t = cy - R cx, with R the
    optimal rotation of kabsch_rotation_weighted (mass=True weights by mass;
    norotate keeps R = I) from the engine's f64 fit (``wmd_align_*``).
    """

    def run(self, traj: Trajectory, system: System, chunk_frames: Optional[int] = None,
            device: str = "auto"):
        return self._run(traj, system, chunk_frames, device, None, superpose=False)


class SuperposePlan(_SuperposeBase):
    """Superposed coordinates of every atom, R x + t per frame.

    Mirrors PySuperposePlan (crates/traj-py/src/correlation.rs:265-305) over
    SuperposePlan (plan_exec.rs:257-465): returns the flat float32 series
    [F * n_atoms * 3], the transform evaluated in f64 and cast as the
    reference's CPU path does.
    """

    def run(self, traj: Trajectory, system: System, chunk_frames: Optional[int] = None,
            device: str = "auto", frame_indices: Optional[Sequence[int]] = None):
        return self._run(traj, system, chunk_frames, device, frame_indices, superpose=True)


class TrajectoryClusterPlan:
    """Frame clustering by DBSCAN on the pairwise RMSD matrix or k-means on
    superposed coordinates.

    Mirrors PyTrajectoryClusterPlan (crates/traj-py/src/io.rs:2805-2880) over
    TrajectoryClusterPlan (crates/traj-engine/src/plans/analysis/
    clustering.rs): the matrix and the k-means distance and sum passes run on
    the device (``wmd_pairwise_*``), DBSCAN's graph walk on the host
    (``wmd_dbscan``).  Returns dict(labels, centroids, sizes, method,
    n_frames) like clustering_to_py (pack.rs:2614-2625).
    """

    def __init__(self, selection, method: str = "dbscan", eps: float = 2.0, min_samples: int = 5,
                 n_clusters: int = 8, max_iter: int = 100, tol: float = 1.0e-4, seed: int = 0,
                 memory_budget_bytes: Optional[int] = None):
        m = str(method).strip().lower()
        if m not in ("dbscan", "kmeans"):
            raise RuntimeError("method must be 'dbscan' or 'kmeans'")
        self.selection = _indices(selection)
        self.method = m
        self.eps = float(np.float32(eps))
        self.min_samples = int(min_samples)
        self.n_clusters = int(n_clusters)
        self.max_iter = int(max_iter)
        self.tol = float(np.float32(tol))
        self.seed = int(seed)
        self.memory_budget_bytes = memory_budget_bytes

    def run(self, traj: Trajectory, system: System, chunk_frames: Optional[int] = None,
            device: str = "auto", frame_indices: Optional[Sequence[int]] = None):
        dev = Device.from_spec(device)
        if traj.n_atoms() != system.n_atoms():
            raise RuntimeError("mismatch: trajectory atom count does not match system")
        ctx = dev.context()
        n_atoms = system.n_atoms()
        sel = self.selection
        if sel.size and int(sel.max()) >= n_atoms:
            raise RuntimeError("mismatch: trajectory_cluster selection index out of bounds")
        h = ctypes.c_void_p()
        check(lib().wmd_pairwise_create(ctx.handle, ptr(sel), sel.size, 0, ctypes.byref(h)))
        try:
            tr = select_frames(traj, frame_indices)
            tr.reset()
            step = resolve_chunk_frames(chunk_frames, max(1, n_atoms))
            while True:
                chunk = tr.read_chunk4(step)
                if chunk is None:
                    break
                x4, _ = chunk
                check(lib().wmd_pairwise_append(h, ptr(x4), n_atoms, x4.shape[0]))
            nf = ctypes.c_size_t()
            check(lib().wmd_pairwise_frames(h, ctypes.byref(nf)))
            F = nf.value
            out = {"labels": np.zeros(0, np.int32), "centroids": np.zeros(0, np.uint32),
                   "sizes": np.zeros(0, np.uint32), "method": self.method, "n_frames": F}
            if F == 0:  # clustering.rs:152-161
                return out
            if self.method == "dbscan":
                if self.eps <= 0.0:
                    raise RuntimeError("parse error: trajectory_cluster DBSCAN eps must be > 0")
                pairs = F * (F - 1) // 2
                need = pairs * 4
                if self.memory_budget_bytes is not None and need > int(self.memory_budget_bytes):
                    raise RuntimeError(
                        f"unsupported: trajectory_cluster distance matrix requires {need} bytes but "
                        f"memory_budget_bytes={int(self.memory_budget_bytes)}")
                mat = np.zeros((F, F), np.float32)
                check(lib().wmd_pairwise_finalize(h, ptr(mat)))
                labels = np.zeros(F, np.int32)
                cent = np.zeros(F, np.uint32)
                sizes = np.zeros(F, np.uint32)
                nc = ctypes.c_size_t()
                check(lib().wmd_dbscan(ptr(mat), F, ctypes.c_double(self.eps), max(self.min_samples, 1),
                                       ptr(labels), ptr(cent), ptr(sizes), ctypes.byref(nc)))
                out.update(labels=labels, centroids=cent[:nc.value].copy(), sizes=sizes[:nc.value].copy())
            else:
                k = min(max(self.n_clusters, 1), F)
                labels = np.zeros(F, np.int32)
                cent = np.zeros(k, np.uint32)
                sizes = np.zeros(k, np.uint32)
                check(lib().wmd_pairwise_kmeans(h, k, max(self.max_iter, 1), ctypes.c_double(max(self.tol, 0.0)),
                                                ctypes.c_uint64(self.seed), ptr(labels), ptr(cent), ptr(sizes)))
                out.update(labels=labels, centroids=cent, sizes=sizes)
            return out
        finally:
            lib().wmd_pairwise_destroy(h)


def pairdist_centers(bins: int, w) -> np.ndarray:
    """Bin centres i * w + 0.5 * w, each f32 operation rounded (rdf.rs:1276-1279, 459)."""
    w = np.float32(w)
    idx = np.arange(bins, dtype=np.float32)
    return (idx * w + np.float32(0.5) * w).astype(np.float32)


def pairdist_stats(counts, sums, sumsq, frames: int, w, compact: bool):
    """pairdist_output_from_stats (rdf.rs:435-471): (centers, probability, std,
    counts, frames) from the summed per-frame counts and their squares."""
    counts = np.asarray(counts, np.uint64)
    sums = np.asarray(sums, np.float64)
    sumsq = np.asarray(sumsq, np.float64)
    bins = counts.size
    centers = pairdist_centers(bins, w)
    ff = float(max(frames, 1))
    keep = sums > 0.0 if compact else np.ones(bins, bool)
    mean = sums / ff
    if frames > 1:
        var = np.maximum((sumsq - sums * sums / ff) / (ff - 1.0), 0.0)
    else:
        var = np.zeros(bins)
    prob = (mean / float(np.float32(w))).astype(np.float32)
    return (centers[keep], prob[keep], np.sqrt(var).astype(np.float32)[keep], counts[keep],
            frames)
