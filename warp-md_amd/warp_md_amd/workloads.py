"""Seeded synthetic inputs of the headline workloads (float32 coordinates).

* water lattice  -- rigid 3-site waters: OW on an n^3 lattice (optionally a
  seeded subset of its sites) plus per-frame Gaussian jitter, H atoms at
  0.9572 A / 104.52 deg in a random orientation per water and frame; cubic
  box, or the same lattice mapped through a triclinic box with an xy tilt.
* chain          -- a CA random walk (3.8 A steps) with 9 satellites per CA at
  1.5 A; each frame a Haar-random rotation, a U(-20, 20)^3 translation and
  Gaussian noise; the noiseless structure is the reference.
* uneven batch   -- trajectories with log-uniform atom counts, uniform
  positions in a cubic box of fixed density.
* beads          -- uniform base configuration plus per-frame jitter.

Host (numpy) generators serve the parity tests; ``*_device`` variants draw the
same distributions with torch on the GPU (elementwise ops only) so that the
large benchmark inputs are created in HBM directly, already in the engine's
SoA store layout.
"""
from __future__ import annotations

import math

import numpy as np

HOH_D = 0.9572
HOH_ANGLE = math.radians(104.52)


def random_rotations(rng, n):
    """n Haar-random rotation matrices (unit quaternions from 4 Gaussians)."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def triclinic_box(L, tilt_deg=10.0):
    """Rows a=(L,0,0), b=(L tan(tilt), L, 0), c=(0,0,L)."""
    return np.array([[L, 0, 0], [L * math.tan(math.radians(tilt_deg)), L, 0], [0, 0, L]], np.float64)


def water_lattice(n_side=10, box=31.04, n_frames=100, n_waters=None, jitter=0.3, seed=0,
                  with_h=True, tilt_deg=None):
    """Returns dict(coords (F, N, 3) f32, box (3,) or (3, 3), names, ow index array, sites).

    Lattice sites (i + 0.5) * box / n_side per axis (fractional when tilted);
    with n_waters < n_side^3 a seeded subset of sites is used.
    """
    rng = np.random.default_rng(seed)
    g = (np.arange(n_side) + 0.5) / n_side
    F3 = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    if n_waters is not None and n_waters < len(F3):
        keep = np.sort(rng.choice(len(F3), size=n_waters, replace=False))
        F3 = F3[keep]
    M = triclinic_box(box, tilt_deg) if tilt_deg else np.diag([box, box, box]).astype(np.float64)
    sites = F3 @ M
    nw = len(sites)
    per = 3 if with_h else 1
    coords = np.empty((n_frames, nw * per, 3), np.float32)
    h1 = HOH_D * np.array([math.sin(HOH_ANGLE / 2), 0.0, math.cos(HOH_ANGLE / 2)])
    h2 = HOH_D * np.array([-math.sin(HOH_ANGLE / 2), 0.0, math.cos(HOH_ANGLE / 2)])
    for f in range(n_frames):
        ow = sites + rng.normal(scale=jitter, size=sites.shape)
        if with_h:
            R = random_rotations(rng, nw)
            coords[f, 0::3] = ow
            coords[f, 1::3] = ow + R @ h1
            coords[f, 2::3] = ow + R @ h2
        else:
            coords[f] = ow
    names = (["OW", "HW1", "HW2"] * nw) if with_h else (["OW"] * nw)
    ow_idx = np.arange(0, nw * per, per, dtype=np.uint32)
    bx = M.astype(np.float32) if tilt_deg else np.array([box, box, box], np.float32)
    return {"coords": coords, "box": bx, "names": names, "ow": ow_idx, "sites": sites}


def water_lattice_device(n_side, box, n_frames, n_waters=None, jitter=0.3, seed=1, tilt_deg=None,
                         device="cuda"):
    """OW-only store on the GPU: (F, 3, stride) float32 SoA + the box rows (3, 3)."""
    import torch

    rng = np.random.default_rng(seed)
    g = (np.arange(n_side) + 0.5) / n_side
    F3 = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    if n_waters is not None and n_waters < len(F3):
        keep = np.sort(rng.choice(len(F3), size=n_waters, replace=False))
        F3 = F3[keep]
    M = triclinic_box(box, tilt_deg) if tilt_deg else np.diag([box, box, box]).astype(np.float64)
    sites = torch.tensor((F3 @ M).T.astype(np.float32), device=device)  # (3, n)
    n = sites.shape[1]
    stride = (n + 3) // 4 * 4
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    store = torch.zeros((n_frames, 3, stride), dtype=torch.float32, device=device)
    step = 64
    for f0 in range(0, n_frames, step):
        nf = min(step, n_frames - f0)
        noise = torch.randn((nf, 3, n), generator=gen, device=device, dtype=torch.float32) * jitter
        store[f0:f0 + nf, :, :n] = sites[None] + noise
    return store, n, stride, M.astype(np.float32)


def chain(n_ca=2000, per_ca=10, n_frames=5000, step=3.8, sat=1.5, noise=0.5, trans=20.0, seed=2):
    """Returns dict(ref (N, 3) f32, coords (F, N, 3) f32, ca indices, names)."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n_ca, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ca = np.cumsum(d * step, axis=0)
    ca -= ca[0]
    N = n_ca * per_ca
    ref = np.empty((N, 3))
    ref[0::per_ca] = ca
    for k in range(1, per_ca):
        u = rng.normal(size=(n_ca, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        ref[k::per_ca] = ca + sat * u
    coords = None
    if n_frames:
        R = random_rotations(rng, n_frames)
        t = rng.uniform(-trans, trans, size=(n_frames, 1, 3))
        coords = (np.einsum("fij,nj->fni", R, ref) + t
                  + rng.normal(scale=noise, size=(n_frames, N, 3))).astype(np.float32)
    names = ["CA" if i % per_ca == 0 else "X" for i in range(N)]
    return {"ref": ref.astype(np.float32), "coords": coords,
            "ca": np.arange(0, N, per_ca, dtype=np.uint32), "names": names}


def chain_store_device(ref, atoms, n_frames, noise=0.5, trans=20.0, seed=2, device="cuda"):
    """Frames of `chain` drawn on the GPU directly in an RMSD plan's SoA store
    layout: returns a torch (F, 3, stride) float32 tensor with
    store[f, :, k] = R_f ref[atoms[k]] + t_f + noise (zero where atoms[k] is
    padding).  Only elementwise torch ops are used.
    """
    import torch

    atoms = np.asarray(atoms, dtype=np.uint32)
    stride = atoms.size
    valid = atoms != 0xFFFFFFFF
    soa = np.zeros((3, stride), np.float32)
    soa[:, valid] = np.asarray(ref, np.float32)[atoms[valid].astype(np.int64)].T
    mask = torch.tensor(valid.astype(np.float32), device=device)
    r = torch.tensor(soa, device=device)  # (3, stride)
    rng = np.random.default_rng(seed + 1000)
    R = torch.tensor(random_rotations(rng, n_frames), dtype=torch.float32, device=device)
    t = torch.tensor(rng.uniform(-trans, trans, size=(n_frames, 3)), dtype=torch.float32,
                     device=device)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    store = torch.empty((n_frames, 3, stride), dtype=torch.float32, device=device)
    step = 128
    for f0 in range(0, n_frames, step):
        nf = min(step, n_frames - f0)
        Rb = R[f0:f0 + nf]
        for d in range(3):
            v = (Rb[:, d, 0:1] * r[0:1] + Rb[:, d, 1:2] * r[1:2] + Rb[:, d, 2:3] * r[2:3]
                 + t[f0:f0 + nf, d:d + 1]
                 + noise * torch.randn((nf, stride), generator=gen, device=device,
                                       dtype=torch.float32))
            store[f0:f0 + nf, d] = v * mask
    return store


def store_to_frames(store_np, atoms, n_atoms):
    """Inverse of the store layout on the host: (F, 3, stride) -> (F, n_atoms, 3)."""
    atoms = np.asarray(atoms, dtype=np.uint32)
    valid = atoms != 0xFFFFFFFF
    out = np.zeros((store_np.shape[0], n_atoms, 3), np.float32)
    out[:, atoms[valid].astype(np.int64), :] = np.transpose(store_np[:, :, valid], (0, 2, 1))
    return out


def uneven_batch(n_traj=256, n_min=500, n_max=50000, n_frames=50, density=0.1, seed=3,
                 sizes_only=False):
    """List of dict(n, box, coords (F, n, 3) f32) with log-uniform atom counts."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(math.log(n_min), math.log(n_max), size=n_traj))).astype(int)
    out = []
    for k, n in enumerate(sizes):
        L = (n / density) ** (1.0 / 3.0)
        item = {"n": int(n), "box": np.array([L, L, L], np.float32)}
        if not sizes_only:
            item["coords"] = (rng.random((n_frames, n, 3)) * L).astype(np.float32)
        out.append(item)
    return out


def uneven_store_device(batch, spec, seed=3):
    """Uniform positions in each trajectory's cubic box, drawn on the GPU
    directly into an UnevenBatch store (padding slots stay zero)."""
    import torch

    store = batch.new_store()
    gen = torch.Generator(device=store.device)
    gen.manual_seed(seed)
    for t, item in enumerate(spec):
        n = int(item["n"])
        F = int(batch.n_frames[t])
        v = batch.store_view(store, t)
        L = float(item["box"][0])
        v[:, :, :n] = torch.rand((F, 3, n), generator=gen, device=store.device,
                                 dtype=torch.float32) * L
    return store


def beads(n=20000, density=0.08, n_frames=64, jitter=0.2, seed=4):
    rng = np.random.default_rng(seed)
    L = (n / density) ** (1.0 / 3.0)
    base = rng.random((n, 3)) * L
    coords = (base[None] + rng.normal(scale=jitter, size=(n_frames, n, 3))).astype(np.float32)
    return {"coords": coords, "box": np.array([L, L, L], np.float32), "L": L}
