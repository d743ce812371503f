"""Device dispatch and the chunk loop (crates/traj-engine/src/executor.rs).

``Device.from_spec`` follows executor.rs:21-72.  This package *is* the HIP
path: ``"cpu"`` belongs to warp-md's own CPU plans and is refused here, and
``"auto"`` without a usable gfx950 device raises instead of silently running
something else (in the Rust integration the traj-engine dispatch keeps its CPU
branch for that case).
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from . import _native


class Device:
    def __init__(self, kind: str, index: int = 0):
        self.kind = kind
        self.index = index

    def __repr__(self):
        return f"Device({self.kind}:{self.index})" if self.kind == "hip" else "Device(cpu)"

    @staticmethod
    def from_spec(spec: str) -> "Device":
        s = str(spec).strip()
        low = s.lower()
        if low == "cpu":
            return Device("cpu")
        if low == "auto":
            return Device("hip", 0)
        for head in ("cuda", "hip", "rocm"):
            if low.startswith(head):
                return Device("hip", _parse_index(s, head))
        raise RuntimeError(f"unsupported: unknown device spec '{s}'")

    def context(self) -> "_native.Context":
        if self.kind != "hip":
            raise RuntimeError(
                "unsupported: device 'cpu' is served by warp-md's CPU plans; "
                "this engine implements the HIP device path")
        if _native.device_count() == 0:
            raise RuntimeError("unsupported: no HIP device available for device='auto'/'hip'")
        return _native.context(self.index)


def _parse_index(spec: str, head: str) -> int:
    # executor.rs:54-72 parse_cuda_index
    if ":" in spec:
        h, tail = spec.split(":", 1)
        if h.lower() != head or ":" in tail:
            raise RuntimeError(f"parse error: invalid {head} device spec '{spec}'")
        try:
            return int(tail)
        except ValueError:
            raise RuntimeError(f"parse error: invalid {head} device spec '{spec}'") from None
    if spec.lower() == head:
        return 0
    raise RuntimeError(f"parse error: invalid {head} device spec '{spec}'")


def normalize_frame_indices(frame_indices: Sequence[int], n_frames: int) -> list:
    """executor.rs:210-222: negatives count from the end, out-of-range dropped."""
    out = []
    for raw in frame_indices:
        idx = int(raw)
        if idx < 0:
            idx += n_frames
        if 0 <= idx < n_frames:
            out.append(idx)
    return out


def select_frames(traj, frame_indices: Optional[Sequence[int]]):
    """run_plan_with_frame_subset frame selection (traj-py pack.rs:1259-1358,
    executor.rs:261-353): given order kept, duplicates kept, out-of-range dropped."""
    if frame_indices is None:
        return traj
    raw = [int(v) for v in frame_indices]
    n = traj.n_frames()
    if any(v < 0 for v in raw):
        idx = normalize_frame_indices(raw, n)
    else:
        idx = [v for v in raw if v < n]
    return traj.subset(idx)


def resolve_chunk_frames(chunk_frames: Optional[int], n_atoms_per_frame: int) -> int:
    """Frames per host->device chunk.  Results do not depend on it; the
    default sizes a chunk near 256 MiB of xyzw coordinates so that transfers
    are large (the reference's 8 MiB CPU heuristic, pack.rs:1360-1466, is
    tuned for cache, not PCIe)."""
    if chunk_frames is not None:
        return max(1, int(chunk_frames))
    per = max(1, 16 * int(n_atoms_per_frame))
    return int(max(16, min(65536, (256 << 20) // per)))


def xyzw(coords: np.ndarray) -> np.ndarray:
    """(f, n, 3) float32 -> (f, n, 4) with w = 1 (the FrameChunk layout, io.rs:1995)."""
    f, n, _ = coords.shape
    out = np.empty((f, n, 4), np.float32)
    out[..., :3] = coords
    out[..., 3] = 1.0
    return out
