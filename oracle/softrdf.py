"""Gaussian-smoothed RDF and fitted-RMSD adjoint: dense float64 definitions.

TEST INFRASTRUCTURE ONLY.  This module is the checker of the engine's
``wmd_softrdf_*`` and ``wmd_rmsd_backward_device`` entry points; only tests/,
``__graft_entry__.smoke()`` and bench.py's CPU baseline may import it.

The reference (Haayhur/warp-md) has no smoothed RDF and no gradients; warp-cg
optimises without them (crates/warp-cg/src/optimize/bo/mod.rs:237) against
objectives such as crates/warp-cg/src/optimize/objective.rs:271-367.  The
definition restated here is therefore this engine's own, built on the
reference's pair set and minimum-image rule:

* pairs: unordered i < j of one atom set, displacement x_j - x_i wrapped with
  d -= round(d / L) * L per axis, round half away from zero
  (crates/traj-engine/src/plans/rdf.rs:924-928), or f -= round(f) on the
  fractional components of a triclinic box (crates/traj-core/src/pbc_math.rs:92-164);
* soft count of bin k (centre r_k = (k + 1/2) w, w = r_max / bins):
      c_k = sum over ordered pairs of w / (sigma sqrt(2 pi)) exp(-(r_k - r)^2 / (2 sigma^2)),
  restricted to the 17 bins k0 - 8 .. k0 + 8, k0 = floor(r / w);
* normalisation as RdfPlan::finalize (rdf.rs:943-1016) with use_volume:
      g_k = c_k / (frames * shell_k * rho),  rho = N (N - 1) / mean volume,
  shell_k = 4/3 pi ((r_k + w/2)^3 - (r_k - w/2)^3) (evaluated here in float64);
* loss L = sum_k (g_k - g_t(r_k))^2 with g_t(r) = 1 + 0.3 exp(-(r - 4)^2) cos(2 r).

Pins (tests/test_oracle_softrdf.py): the gradient below against central
finite differences of the loss; mass conservation (a pair far from r = 0 and
r_max adds exactly 2 to sum_k c_k up to e^-30, by Poisson summation of the
Gaussian sampled at spacing w <= 2 sigma); the closed form of a single pair.
"""
from __future__ import annotations

import math

import numpy as np

TAPS = 8
SQRT_2PI = math.sqrt(2.0 * math.pi)


def round_half_away(v):
    return np.copysign(np.floor(np.abs(v) + 0.5), v)


def _box_matrix(box):
    b = np.asarray(box, np.float64)
    if b.shape == (3,):
        return np.diag(b)
    return b.reshape(3, 3)


def pair_vectors(x, pbc, box=None):
    """Unordered pairs of one frame: (i, j, d = x_j - x_i minimum image), float64."""
    x = np.asarray(x, np.float64)
    n = len(x)
    i, j = np.triu_indices(n, k=1)
    d = x[j] - x[i]
    if pbc == "orthorhombic":
        L = np.diag(_box_matrix(box))
        d = d - round_half_away(d / L) * L
    elif pbc == "triclinic":
        M = _box_matrix(box)          # rows = box vectors
        inv = np.linalg.inv(M)
        f = d @ inv                   # fractional components
        f = f - round_half_away(f)
        d = f @ M
    return i, j, d


def _taps(r, bins, r_max):
    w = r_max / bins
    k0 = np.floor(r / w).astype(np.int64)
    ks = k0[:, None] + np.arange(-TAPS, TAPS + 1)[None, :]
    valid = (ks >= 0) & (ks < bins)
    rk = (ks + 0.5) * w
    return w, ks, valid, rk


def soft_counts(frames, bins, r_max, sigma, pbc="orthorhombic", boxes=None):
    """c [bins] float64 summed over the frames (F, N, 3)."""
    frames = np.asarray(frames, np.float64)
    c = np.zeros(bins)
    for f in range(len(frames)):
        box = None if boxes is None else boxes[f]
        _, _, d = pair_vectors(frames[f], pbc, box)
        r = np.sqrt((d * d).sum(1))
        w, ks, valid, rk = _taps(r, bins, r_max)
        val = 2.0 * w / (sigma * SQRT_2PI) * np.exp(-(rk - r[:, None]) ** 2 / (2.0 * sigma * sigma))
        np.add.at(c, ks[valid], val[valid])
    return c


def soft_grad(frames, bins, r_max, sigma, dldc, pbc="orthorhombic", boxes=None):
    """dL/dx (F, N, 3) float64 given dL/dc [bins]."""
    frames = np.asarray(frames, np.float64)
    dldc = np.asarray(dldc, np.float64)
    out = np.zeros_like(frames)
    for f in range(len(frames)):
        box = None if boxes is None else boxes[f]
        i, j, d = pair_vectors(frames[f], pbc, box)
        r = np.sqrt((d * d).sum(1))
        w, ks, valid, rk = _taps(r, bins, r_max)
        kk = np.where(valid, ks, 0)
        dd = rk - r[:, None]
        g = 2.0 * w / (sigma * SQRT_2PI) * np.exp(-dd ** 2 / (2.0 * sigma * sigma))
        # d/dr of g(r_k - r) = (r_k - r) / sigma^2 * g
        F = (np.where(valid, dldc[kk], 0.0) * dd / (sigma * sigma) * g).sum(1)
        fr = np.where(r > 0, F / np.where(r > 0, r, 1.0), 0.0)
        gv = fr[:, None] * d          # dL/dx_j; dL/dx_i is its negative
        np.add.at(out[f], j, gv)
        np.add.at(out[f], i, -gv)
    return out


def shells(bins, r_max):
    w = r_max / bins
    ri = np.arange(bins) * w
    ro = ri + w
    return 4.0 / 3.0 * math.pi * (ro ** 3 - ri ** 3)


def g_target(r):
    return 1.0 + 0.3 * np.exp(-(r - 4.0) ** 2) * np.cos(2.0 * r)


def normalisation(n_atoms, n_frames, bins, r_max, volumes):
    """frames * shell_k * rho for g_k = c_k / that."""
    rho = n_atoms * (n_atoms - 1) / (float(np.sum(volumes)) / n_frames)
    return n_frames * shells(bins, r_max) * rho


def rdf_loss(frames, bins, r_max, sigma, volumes, pbc="orthorhombic", boxes=None):
    """(L, dL/dc) of the structure-matching loss."""
    frames = np.asarray(frames, np.float64)
    c = soft_counts(frames, bins, r_max, sigma, pbc, boxes)
    norm = normalisation(frames.shape[1], len(frames), bins, r_max, volumes)
    w = r_max / bins
    rk = (np.arange(bins) + 0.5) * w
    resid = c / norm - g_target(rk)
    return float((resid ** 2).sum()), 2.0 * resid / norm


# ---- fitted RMSD and its adjoint --------------------------------------------
def fit_rmsd_parts(x, y):
    """(rmsd, R, cx, cy) of the Kabsch fit R (x - cx) ~ (y - cy), float64."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    cx = x.mean(0)
    cy = y.mean(0)
    H = (x - cx).T @ (y - cy)
    U, _, Vt = np.linalg.svd(H)
    D = np.eye(3)
    D[2, 2] = np.sign(np.linalg.det(Vt.T @ U.T))
    R = Vt.T @ D @ U.T
    res = (x - cx) @ R.T - (y - cy)
    return math.sqrt((res * res).sum() / len(x)), R, cx, cy


def fit_rmsd_grad(x, y):
    """dRMSD/dx = ((x - cx) - R^T (y - cy)) / (n RMSD) (envelope theorem)."""
    rmsd, R, cx, cy = fit_rmsd_parts(x, y)
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    return ((x - cx) - (y - cy) @ R) / (len(x) * rmsd)
