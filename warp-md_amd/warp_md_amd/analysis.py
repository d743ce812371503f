"""``rdf()`` convenience wrapper with warp-md's keyword mapping
(python/warp_md/analysis/rdf.py:35-163)."""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from .plans import RdfPlan


def _normalize_rdf_output_mode(norm, raw_rdf: bool, volume: bool):
    # rdf.py:35-48
    number_density = False
    if norm is None:
        return bool(raw_rdf), bool(volume), number_density
    key = str(norm).strip().lower().replace("-", "_").replace(" ", "_")
    if key in ("rdf", "g_r", "gr"):
        return False, bool(volume), number_density
    if key in ("none", "raw", "count", "counts"):
        return True, False, number_density
    if key in ("volume", "box", "box_volume"):
        return False, True, number_density
    if key in ("density", "number_density", "numberdensity"):
        return False, False, True
    raise ValueError("norm must be 'rdf', 'volume', 'number_density', or 'none'")


def _format(dtype, centers, values, counts, integral):
    # rdf.py:15-32 output formatting
    key = str(dtype).lower()
    if key == "dict":
        out = {"bin_centers": centers, "rdf": values, "counts": counts}
        if integral is not None:
            out["integral_rdf"] = np.asarray(integral, np.float32)
        return out
    if key in ("full", "tuple_full"):
        if integral is not None:
            return centers, values, counts, np.asarray(integral, np.float32)
        return centers, values, counts
    return centers, values


def _selection(system, mask):
    if mask in ("", None):
        return None
    if isinstance(mask, str) and mask.startswith("@"):
        # 1-based atom numbers, comma separated (allow_at_indices)
        idx = [int(v) - 1 for v in mask[1:].split(",") if v.strip()]
        return system.select_indices(idx)
    if isinstance(mask, str):
        return system.select(mask)
    return system.select_indices(mask)


def rdf(traj, system, solvent_mask=":WAT@O", solute_mask="", maximum: float = 10.0,
        bin_spacing: float = 0.5, image: bool = True, density: Optional[float] = 0.033456,
        volume: bool = False, center_solvent: bool = False, center_solute: bool = False,
        intramol: bool = True, frame_indices: Optional[Sequence[int]] = None, top=None,
        raw_rdf: bool = False, norm: Optional[str] = None, dtype: str = "tuple",
        byres1: bool = False, byres2: bool = False, bymol1: bool = False, bymol2: bool = False,
        mass: bool = True, intrdf: bool = False, dimension: str = "3d",
        chunk_frames: Optional[int] = None, device: str = "auto", pbc: Optional[str] = None):
    """RDF between two masks.  ``pbc`` (an addition) overrides ``image``, e.g.
    ``pbc="triclinic"``."""
    del top
    maximum = float(maximum)
    bin_spacing = float(bin_spacing)
    if maximum <= 0.0:
        raise ValueError("maximum must be positive")
    if bin_spacing <= 0.0:
        raise ValueError("bin_spacing must be positive")
    density_value = 0.033456 if density is None else float(density)
    if not np.isfinite(density_value) or density_value <= 0.0:
        raise ValueError("density must be finite and positive")
    dimension_value = str(dimension).lower()
    if dimension_value not in ("3d", "xyz", "xy"):
        raise ValueError("dimension must be '3d' or 'xy'")
    raw_value, volume_value, number_density = _normalize_rdf_output_mode(norm, raw_rdf, volume)
    # rdf.py:107-108
    bins = max(1, int(np.ceil(maximum / bin_spacing)))
    r_max = np.float32(bins * bin_spacing)
    pbc_mode = pbc if pbc is not None else ("orthorhombic" if image else "none")
    sel_a = _selection(system, solvent_mask)
    sel_b = sel_a if solute_mask in ("", None) else _selection(system, solute_mask)
    plan = RdfPlan(sel_a, sel_b, bins=bins, r_max=float(r_max), pbc=pbc_mode,
                   center1=bool(center_solvent), center2=bool(center_solute), byres1=bool(byres1),
                   byres2=bool(byres2), bymol1=bool(bymol1), bymol2=bool(bymol2),
                   no_intramol=not bool(intramol), mass_weighted=bool(mass), density=density_value,
                   volume=volume_value, raw_rdf=raw_value, intrdf=bool(intrdf),
                   dimension=dimension_value, number_density=number_density)
    result = plan.run(traj, system, chunk_frames=chunk_frames, device=device,
                      frame_indices=None if frame_indices is None else [int(v) for v in frame_indices])
    if len(result) == 4:
        centers, values, counts, integral = result
    else:
        centers, values, counts = result
        integral = None
    return _format(dtype, np.asarray(centers, np.float32), np.asarray(values, np.float32),
                   np.asarray(counts, np.uint64), integral)
