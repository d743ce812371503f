"""warp-md's RdfPlan / RmsdPlan hot path on AMD Instinct MI355X (gfx950).

The compute runs in hand-written HIP kernels behind the C ABI declared in
``include/wmd_hip.h`` (library ``warp_md_amd/lib/libwmd_hip.so``); this package
is the host-side mirror of warp-md's Python plan API:

    plan = RdfPlan(sel_a, sel_b, bins=200, r_max=10.0, pbc="orthorhombic")
    r, g, counts = plan.run(traj, system, device="auto")
    rmsd = RmsdPlan(sel, reference="topology", align=True).run(traj, system)
"""
from ._native import WmdError, device_count, library_built
from .analysis import rdf
from .executor import Device, normalize_frame_indices
from .batch import UnevenBatch
from .plans import (AlignPlan, PairDistPlan, PairwiseRmsdPlan, RdfPlan, RmsdPlan,
                    StructureFactorPlan, SuperposePlan, TrajectoryClusterPlan, finalize_rdf)
from .system import Selection, System, Trajectory

__all__ = [
    "AlignPlan", "Device", "PairDistPlan", "PairwiseRmsdPlan", "RdfPlan", "RmsdPlan", "StructureFactorPlan",
    "SuperposePlan", "TrajectoryClusterPlan", "Selection", "System", "Trajectory", "UnevenBatch", "WmdError",
    "device_count", "finalize_rdf", "library_built", "normalize_frame_indices", "rdf",
]
__version__ = "0.1.0"
