"""ctypes front end to the CPU oracle (``oracle/libwmd_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg, always as the checker (or the timed CPU
baseline), never by the product package ``warp_md_amd``.

The C file restates warp-md's CPU semantics (citations in ``wmd_oracle.c``):

* ``rdf_counts``      -- ``RdfPlan`` pair loop, ``crates/traj-engine/src/plans/rdf.rs:851-941``
* ``rdf_finalize``    -- ``RdfPlan::finalize``, ``rdf.rs:943-1016``
* ``rmsd_plan``       -- ``rmsd_aligned_selected`` / ``rmsd_raw_selected``,
  ``crates/traj-engine/src/plans/rmsd/plan.rs:1093-1248``
* ``fit_rmsd``        -- ``kabsch_rotation`` + per-group residual,
  ``plans/geometry/geometry_math.rs:169-297``, ``plans/geometry/rms.rs:643-676``
* ``min_image_f64``   -- ``crates/traj-core/src/pbc_math.rs:17-164``
* ``pairdist_counts`` -- ``PairDistPlan`` pair loop, ``rdf.rs:1120-1178``
* ``sf_counts``, ``sf_finalize`` -- ``StructureFactorPlan``,
  ``crates/traj-engine/src/plans/analysis/structure_factor.rs:160-258``
* ``pairwise_rmsd``   -- ``PairwiseRmsdPlan::finalize`` (rms / nofit),
  ``rmsd/plan.rs:586-675, 1081-1159``
* ``rdf_grouped``     -- ``RdfPlan`` grouped modes and the xy dimension, ``rdf.rs:808-850``
* ``align``           -- ``AlignPlan`` / ``SuperposePlan`` CPU path,
  ``plans/geometry/align/plan_exec.rs:215-243, 422-456``

``rdf_finalize_py`` below is a second, independent float64/float32 numpy
restatement of the normalisation used to cross-check the C version.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwmd_oracle.so")

PBC_KINDS = {"none": 0, "orthorhombic": 1, "ortho": 1, "triclinic": 2}

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    cmd = ["make", "-s", "-C", HERE]
    if force:
        cmd.insert(1, "-B")
    subprocess.run(cmd, check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    S = ctypes.c_size_t
    L.orc_rdf_brute.argtypes = [P, S, S, S, P, S, P, S, S, ctypes.c_float, ctypes.c_int, P, P, P]
    L.orc_rdf_brute.restype = ctypes.c_int
    L.orc_rdf_brute_omp.argtypes = [P, S, S, S, P, S, P, S, S, ctypes.c_float, ctypes.c_int, P, P,
                                    ctypes.c_int]
    L.orc_rdf_brute_omp.restype = ctypes.c_int
    L.orc_rdf_cells.argtypes = [P, S, S, S, P, S, P, S, S, ctypes.c_float, ctypes.c_int, P, P]
    L.orc_rdf_cells.restype = ctypes.c_int
    L.orc_rdf_finalize.argtypes = [P, S, ctypes.c_float, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, S, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, P]
    L.orc_rdf_finalize.restype = ctypes.c_int
    L.orc_rmsd_plan.argtypes = [P, S, S, S, P, S, P, S, ctypes.c_int, P, P]
    L.orc_rmsd_plan.restype = ctypes.c_int
    L.orc_fit_rmsd.argtypes = [P, S, S, S, P, S, P, S, P, P, P, P, P, P]
    L.orc_fit_rmsd.restype = ctypes.c_int
    L.orc_svd3.argtypes = [P, P, P, P]
    L.orc_svd3.restype = None
    L.orc_min_image_f64.argtypes = [P, P, ctypes.c_int, P, ctypes.c_double, P]
    L.orc_min_image_f64.restype = ctypes.c_int
    L.orc_pairdist.argtypes = [P, S, S, S, P, S, P, S, S, ctypes.c_float, ctypes.c_int, P,
                               ctypes.c_int, ctypes.c_int, P, P]
    L.orc_pairdist.restype = ctypes.c_int
    L.orc_sf_counts.argtypes = [P, S, S, S, P, S, S, ctypes.c_double, ctypes.c_int, P,
                                ctypes.c_double, P, P]
    L.orc_sf_counts.restype = ctypes.c_int
    L.orc_sf_finalize.argtypes = [P, S, ctypes.c_double, ctypes.c_double, S, ctypes.c_double,
                                  ctypes.c_int, S, ctypes.c_double, P, P, P, P]
    L.orc_sf_finalize.restype = ctypes.c_int
    L.orc_pairwise_rmsd.argtypes = [P, S, S, S, P, S, ctypes.c_int, P, P]
    L.orc_pairwise_rmsd.restype = ctypes.c_int
    L.orc_pairwise_dme.argtypes = [P, S, S, S, P, S, ctypes.c_int, P, P, P]
    L.orc_pairwise_dme.restype = ctypes.c_int
    L.orc_align.argtypes = [P, S, S, S, P, S, P, P, S, ctypes.c_int, P, P, P]
    L.orc_align.restype = ctypes.c_int
    L.orc_rdf_grouped.argtypes = [P, S, S, S, P, P, S, P, P, P, P, S, P, P, P, ctypes.c_int,
                                  ctypes.c_int, P, ctypes.c_int, S, ctypes.c_float, P]
    L.orc_rdf_grouped.restype = ctypes.c_int
    L.orc_omp_max_threads.argtypes = []
    L.orc_omp_max_threads.restype = ctypes.c_int
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _coords(coords):
    c = np.ascontiguousarray(coords, dtype=np.float32)
    if c.ndim == 2:
        c = c[None]
    if c.ndim != 3 or c.shape[2] not in (3, 4):
        raise ValueError("coords must be (F, N, 3) or (F, N, 4)")
    return c


def _boxes(boxes, n_frames, pbc):
    if pbc == 0:
        return np.zeros((max(n_frames, 1), 9), np.float32)
    b = np.asarray(boxes, dtype=np.float32)
    if b.shape == (3,):
        b = np.diag(b).reshape(1, 9)
    elif b.shape == (3, 3):
        b = b.reshape(1, 9)
    elif b.ndim == 2 and b.shape[1] == 3:
        b = np.stack([np.diag(row).reshape(9) for row in b])
    elif b.ndim == 3:
        b = b.reshape(b.shape[0], 9)
    if b.shape[0] == 1 and n_frames > 1:
        b = np.repeat(b, n_frames, axis=0)
    return np.ascontiguousarray(b, dtype=np.float32)


def rdf_counts(coords, sel_a, sel_b, bins, r_max, pbc="orthorhombic", boxes=None,
               method="brute", edges=False, threads=0):
    """Ordered-pair histogram counts (u64) exactly as the reference pair loop.

    method: "brute" (reference loop order), "cells" (cell-list enumeration of
    the same pairs with the same f32 expression), "omp" (brute force over
    OpenMP threads).  edges=True also returns [edge32, edge64, f32-vs-f64 bin
    mismatches] (brute only).
    """
    L = lib()
    c = _coords(coords)
    F, N, stride = c.shape
    kind = PBC_KINDS[pbc] if isinstance(pbc, str) else int(pbc)
    bx = _boxes(boxes, F, kind)
    a = np.ascontiguousarray(sel_a, dtype=np.uint32)
    b = np.ascontiguousarray(sel_b, dtype=np.uint32)
    counts = np.zeros(int(bins), np.uint64)
    edge = np.zeros(3, np.uint64)
    if method == "brute":
        st = L.orc_rdf_brute(_ptr(c), stride, N, F, _ptr(a), a.size, _ptr(b), b.size, int(bins),
                             float(r_max), kind, _ptr(bx), _ptr(counts), _ptr(edge) if edges else None)
    elif method == "omp":
        st = L.orc_rdf_brute_omp(_ptr(c), stride, N, F, _ptr(a), a.size, _ptr(b), b.size,
                                 int(bins), float(r_max), kind, _ptr(bx), _ptr(counts), int(threads))
    elif method == "cells":
        st = L.orc_rdf_cells(_ptr(c), stride, N, F, _ptr(a), a.size, _ptr(b), b.size, int(bins),
                             float(r_max), kind, _ptr(bx), _ptr(counts))
        if st == 1:
            st = 0
    else:
        raise ValueError(method)
    if st != 0:
        raise RuntimeError(f"oracle rdf failed with status {st}")
    return (counts, edge) if edges else counts


def rdf_finalize(counts, r_max, n_a, n_b, common, frames, volume_sum=0.0, density=0.033456,
                 use_volume=False, raw=False, number_density=False, integral=False):
    L = lib()
    counts = np.ascontiguousarray(counts, dtype=np.uint64)
    bins = counts.size
    r = np.zeros(bins, np.float32)
    g = np.zeros(bins, np.float32)
    it = np.zeros(bins, np.float32)
    st = L.orc_rdf_finalize(_ptr(counts), bins, float(r_max), float(n_a), float(n_b),
                            float(common), int(frames), float(volume_sum), float(density),
                            int(use_volume), int(raw), int(number_density), int(integral),
                            _ptr(r), _ptr(g), _ptr(it))
    if st != 0:
        raise RuntimeError("rdf box normalization requires frame box metadata")
    return (r, g, counts, it) if integral else (r, g, counts)


def rdf_finalize_py(counts, r_max, n_a, n_b, common, frames, volume_sum=0.0, density=0.033456,
                    use_volume=False, raw=False, number_density=False, integral=False):
    """Independent numpy restatement of rdf.rs:943-1016 (vectorised f32/f64)."""
    counts = np.asarray(counts, dtype=np.uint64)
    bins = counts.size
    w = np.float32(np.float32(r_max) / np.float32(bins))
    i = np.arange(bins, dtype=np.float32)
    r_in = (i * w).astype(np.float32)
    r_out = ((i + np.float32(1)) * w).astype(np.float32)
    shell32 = (np.float32(4.0) / np.float32(3.0)) * np.float32(np.pi) * (
        (r_out * r_out) * r_out - (r_in * r_in) * r_in)
    shell = shell32.astype(np.float32).astype(np.float64)
    fr = float(max(frames, 1))
    total = max(n_a * n_b - common, 0.0)
    if use_volume:
        if volume_sum <= 0.0:
            raise RuntimeError("rdf box normalization requires frame box metadata")
        dens = total / (volume_sum / fr)
    elif n_a > 0:
        dens = density * total / n_a
    else:
        dens = 0.0
    expected = shell * dens
    norm = np.where(expected > 0, fr * expected, 0.0)
    c = counts.astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        normalized = np.where(norm > 0, c / norm, 0.0)
        ndn = fr * n_a * shell
        nd = np.where(ndn > 0, c / ndn, 0.0)
    r = (r_in + np.float32(0.5) * w).astype(np.float32)
    if raw:
        g = c.astype(np.float32)
    elif number_density:
        g = nd.astype(np.float32)
    else:
        g = normalized.astype(np.float32)
    if integral:
        if number_density:
            inc = nd * shell
        elif n_b > 0:
            inc = normalized * expected / n_b
        else:
            inc = np.zeros(bins)
        it = np.cumsum(inc).astype(np.float32)
        return r, g, counts, it
    return r, g, counts


def rmsd_plan(coords, sel, reference, align=True):
    """RmsdPlan CPU path; returns (f32 per frame, f64 before the cast)."""
    L = lib()
    c = _coords(coords)
    F, N, stride = c.shape
    ref = np.ascontiguousarray(reference, dtype=np.float32)
    if ref.shape[-1] != stride:
        pad = np.ones(ref.shape[:-1] + (stride,), np.float32)
        pad[..., :3] = ref[..., :3]
        ref = pad
    s = np.ascontiguousarray(sel, dtype=np.uint32)
    out = np.zeros(F, np.float32)
    out64 = np.zeros(F, np.float64)
    st = L.orc_rmsd_plan(_ptr(c), stride, N, F, _ptr(s), s.size, _ptr(ref), ref.shape[0],
                         int(bool(align)), _ptr(out), _ptr(out64))
    if st != 0:
        raise RuntimeError(f"oracle rmsd failed {st}")
    return out, out64


def pairwise_rmsd(coords, sel, metric="rms"):
    """PairwiseRmsdPlan CPU path; returns ((F, F) f32, (F, F) f64 before the cast)."""
    L = lib()
    c = _coords(coords)
    F, N, stride = c.shape
    s = np.ascontiguousarray(sel, dtype=np.uint32)
    out = np.zeros((F, F), np.float32)
    out64 = np.zeros((F, F), np.float64)
    m = {"rms": 0, "nofit": 1}[metric]
    st = L.orc_pairwise_rmsd(_ptr(c), stride, N, F, _ptr(s), s.size, m, _ptr(out), _ptr(out64))
    if st != 0:
        raise RuntimeError(f"oracle pairwise rmsd failed {st}")
    return out, out64


def pairwise_dme(coords, sel, pbc="none", boxes=None):
    """PairwiseRmsdPlan metric "dme" CPU path (rmsd/plan.rs:612-650, 1279-1302);
    boxes: (F, 3) box lengths for pbc "orthorhombic".  Returns ((F, F) f32,
    (F, F) f64 before the cast)."""
    L = lib()
    c = _coords(coords)
    F, N, stride = c.shape
    s = np.ascontiguousarray(sel, dtype=np.uint32)
    kind = PBC_KINDS[pbc]
    bx = None
    if kind == 1:
        bx = np.ascontiguousarray(np.broadcast_to(np.asarray(boxes, np.float32).reshape(-1, 3), (F, 3)))
    out = np.zeros((F, F), np.float32)
    out64 = np.zeros((F, F), np.float64)
    st = L.orc_pairwise_dme(_ptr(c), stride, N, F, _ptr(s), s.size, kind, _ptr(bx), _ptr(out), _ptr(out64))
    if st != 0:
        raise RuntimeError(f"oracle pairwise dme failed {st}")
    return out, out64


def rdf_grouped(coords, groups_a, groups_b, bins, r_max, pbc="none", boxes=None, masses=None,
                no_intramol=False, planar=False):
    """RdfPlan grouped / xy counts.  groups: lists of (atom tuple, residue key or
    None, molecule key or None)."""
    L = lib()
    c = _coords(coords)
    F, N, stride = c.shape
    rid = {}
    none = -(2 ** 63)

    def pack(groups):
        off = np.zeros(len(groups) + 1, np.uint64)
        off[1:] = np.cumsum([len(g[0]) for g in groups])
        atoms = np.array([a for g in groups for a in g[0]], np.uint32)
        if atoms.size == 0:
            atoms = np.zeros(1, np.uint32)
        mol = np.array([none if g[2] is None else int(g[2]) for g in groups] or [none], np.int64)
        res = np.array([none if g[1] is None else rid.setdefault(g[1], len(rid)) for g in groups]
                       or [none], np.int64)
        return off, atoms, mol, res

    oa, aa, ma, ra = pack(groups_a)
    ob, ab, mb, rb = pack(groups_b)
    bx = _boxes(boxes, F, PBC_KINDS[pbc])
    m = None if masses is None else np.ascontiguousarray(masses, np.float32)
    counts = np.zeros(bins, np.uint64)
    st = L.orc_rdf_grouped(_ptr(c), stride, N, F, _ptr(oa), _ptr(aa), len(groups_a), _ptr(ma), _ptr(ra),
                           _ptr(ob), _ptr(ab), len(groups_b), _ptr(mb), _ptr(rb), _ptr(m),
                           int(bool(no_intramol)), PBC_KINDS[pbc], _ptr(bx), int(bool(planar)), bins,
                           ctypes.c_float(r_max), _ptr(counts))
    if st != 0:
        raise RuntimeError(f"oracle rdf_grouped failed {st}")
    return counts


def align(coords, sel, ref_all, weights=None, norotate=False, superpose=False):
    """AlignPlan rows (F, 12) f32, their f64 values, and (superpose) the
    (F, n_atoms, 3) f32 transformed coordinates (else None)."""
    L = lib()
    c = _coords(coords)
    F, N, stride = c.shape
    ref = np.ascontiguousarray(ref_all, dtype=np.float32)
    if ref.shape[-1] != stride:
        pad = np.ones(ref.shape[:-1] + (stride,), np.float32)
        pad[..., :3] = ref[..., :3]
        ref = pad
    s = np.ascontiguousarray(sel, dtype=np.uint32)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
    rt = np.zeros((F, 12), np.float32)
    rt64 = np.zeros((F, 12), np.float64)
    xyz = np.zeros((F, N, 3), np.float32) if superpose else None
    st = L.orc_align(_ptr(c), stride, N, F, _ptr(s), s.size, _ptr(w) if w is not None else None,
                     _ptr(ref), ref.shape[0], int(bool(norotate)), _ptr(rt), _ptr(rt64),
                     _ptr(xyz) if xyz is not None else None)
    if st != 0:
        raise RuntimeError(f"oracle align failed {st}")
    return rt, rt64, xyz


def fit_rmsd(coords, fit_sel, rmsd_sel, ref_all):
    """Fit on fit_sel, RMSD over rmsd_sel; returns (f32, f64, R[F,3,3], cx[F,3], cy[F,3])."""
    L = lib()
    c = _coords(coords)
    F, N, stride = c.shape
    ref = np.ascontiguousarray(ref_all, dtype=np.float32)
    if ref.shape[-1] != stride:
        pad = np.ones(ref.shape[:-1] + (stride,), np.float32)
        pad[..., :3] = ref[..., :3]
        ref = pad
    a = np.ascontiguousarray(fit_sel, dtype=np.uint32)
    b = np.ascontiguousarray(rmsd_sel, dtype=np.uint32)
    out = np.zeros(F, np.float32)
    out64 = np.zeros(F, np.float64)
    rot = np.zeros((F, 3, 3), np.float64)
    cx = np.zeros((F, 3), np.float64)
    cy = np.zeros((F, 3), np.float64)
    st = L.orc_fit_rmsd(_ptr(c), stride, N, F, _ptr(a), a.size, _ptr(b), b.size, _ptr(ref),
                        _ptr(out), _ptr(out64), _ptr(rot), _ptr(cx), _ptr(cy))
    if st != 0:
        raise RuntimeError(f"oracle fit_rmsd failed {st}")
    return out, out64, rot, cx, cy


def svd3(a):
    L = lib()
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(9)
    u = np.zeros(9)
    s = np.zeros(3)
    v = np.zeros(9)
    L.orc_svd3(_ptr(a), _ptr(u), _ptr(s), _ptr(v))
    return u.reshape(3, 3), s, v.reshape(3, 3)


def min_image_f64(a, b, box=None, kind="none", length_scale=1.0):
    L = lib()
    aa = np.ascontiguousarray(a, dtype=np.float64)
    bb = np.ascontiguousarray(b, dtype=np.float64)
    k = PBC_KINDS[kind]
    m = np.zeros(9, np.float32) if box is None else _boxes(box, 1, k)[0]
    out = np.zeros(3)
    st = L.orc_min_image_f64(_ptr(aa), _ptr(bb), k, _ptr(m), float(length_scale), _ptr(out))
    if st != 0:
        raise RuntimeError("box matrix not invertible")
    return out


def omp_max_threads() -> int:
    return int(lib().orc_omp_max_threads())


def pairdist_counts(coords, sel_a, sel_b, bins, r_max, pbc="orthorhombic", boxes=None,
                    same_sel=False, unique_pairs=False, per_frame=False):
    """PairDistPlan counts (u64 [bins]); per_frame=True also returns [F][bins]."""
    L = lib()
    c = _coords(coords)
    F, N, stride = c.shape
    kind = PBC_KINDS[pbc] if isinstance(pbc, str) else int(pbc)
    bx = _boxes(boxes, F, kind)
    a = np.ascontiguousarray(sel_a, dtype=np.uint32)
    b = np.ascontiguousarray(sel_b, dtype=np.uint32)
    counts = np.zeros(int(bins), np.uint64)
    fc = np.zeros((F, int(bins)), np.uint64) if per_frame else None
    st = L.orc_pairdist(_ptr(c), stride, N, F, _ptr(a), a.size, _ptr(b), b.size, int(bins),
                        float(r_max), kind, _ptr(bx), int(bool(same_sel)), int(bool(unique_pairs)),
                        _ptr(counts), _ptr(fc))
    if st != 0:
        raise RuntimeError(f"oracle pairdist failed with status {st}")
    return (counts, fc) if per_frame else counts


def sf_counts(coords, sel, bins, r_max, pbc="orthorhombic", boxes=None, length_scale=1.0):
    """StructureFactorPlan CPU counts (f64 distances) and the scaled volume sum."""
    L = lib()
    c = _coords(coords)
    F, N, stride = c.shape
    kind = PBC_KINDS[pbc] if isinstance(pbc, str) else int(pbc)
    bx = _boxes(boxes, F, kind)
    s = np.ascontiguousarray(sel, dtype=np.uint32)
    counts = np.zeros(int(bins), np.uint64)
    vol = ctypes.c_double(0.0)
    st = L.orc_sf_counts(_ptr(c), stride, N, F, _ptr(s), s.size, int(bins), float(r_max), kind,
                         _ptr(bx), float(length_scale), _ptr(counts), ctypes.byref(vol))
    if st != 0:
        raise RuntimeError(f"oracle structure factor failed with status {st}")
    return counts, vol.value


def sf_finalize(counts, r_max, n, frames, volume_sum, pbc, q_bins, q_max):
    """(r, g_r, q, s_q) float32 of StructureFactorPlan::finalize."""
    L = lib()
    counts = np.ascontiguousarray(counts, np.uint64)
    bins = counts.size
    r = np.zeros(bins, np.float32)
    g = np.zeros(bins, np.float32)
    q = np.zeros(int(q_bins), np.float32)
    sq = np.zeros(int(q_bins), np.float32)
    kind = PBC_KINDS[pbc] if isinstance(pbc, str) else int(pbc)
    st = L.orc_sf_finalize(_ptr(counts), bins, float(r_max), float(n), int(frames),
                           float(volume_sum), kind, int(q_bins), float(q_max), _ptr(r), _ptr(g),
                           _ptr(q), _ptr(sq))
    if st != 0:
        raise RuntimeError(f"oracle structure factor finalize failed with status {st}")
    return r, g, q, sq
