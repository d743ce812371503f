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
* particle-mesh smoothing: nodes m h with h = w / 4 (w = r_max / bins); every
  ordered pair deposits the quadratic B-spline B(r / h - m) onto its 3 nearest
  nodes, H_m = sum of the deposits; bin k (centre r_k = (k + 1/2) w) collects
      c_k = sum over |m h - r_k| <= 6 sigma of H_m w phi(r_k - m h),
  phi the N(0, sigma^2) density.  The pair kernel is the Gaussian convolved
  with B (variance sigma^2 + h^2 / 4); c is C1 in the positions;
* normalisation as RdfPlan::finalize (rdf.rs:943-1016) with use_volume:
      g_k = c_k / (frames * shell_k * rho),  rho = N (N - 1) / mean volume,
  shell_k = 4/3 pi ((r_k + w/2)^3 - (r_k - w/2)^3) (evaluated here in float64);
* loss L = sum_k (g_k - g_t(r_k))^2 with g_t(r) = 1 + 0.3 exp(-(r - 4)^2) cos(2 r).

Pins (tests/test_oracle_softrdf.py): the gradient below against central
finite differences of the loss; mass conservation (the B-spline is a
partition of unity and the Gaussian sampled at spacing w <= sigma / 4...4 sigma
sums to 1 up to the Poisson-summation alias); the closed form of a single
pair; the effective kernel against the Gaussian of variance sigma^2 + h^2 / 4.
"""
from __future__ import annotations

import math

import numpy as np

NODES_PER_BIN = 4
REACH_SIGMAS = 6.0
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


def geometry(bins, r_max, sigma):
    """(w, h, node count, bin -> node map matrix G [bins, nodes])."""
    w = r_max / bins
    h = w / NODES_PER_BIN
    n_nodes = int(math.ceil(((bins - 0.5) * w + REACH_SIGMAS * sigma) / h)) + 2
    rk = (np.arange(bins) + 0.5) * w
    rm = np.arange(n_nodes) * h
    d = rk[:, None] - rm[None, :]
    G = w / (sigma * SQRT_2PI) * np.exp(-0.5 * d * d / (sigma * sigma))
    G[np.abs(d) > REACH_SIGMAS * sigma] = 0.0
    return w, h, n_nodes, G


def bspline2(u):
    """Nearest node n0 = rint(u), t = u - n0, weights and u-derivatives on n0-1, n0, n0+1."""
    n0 = np.rint(u).astype(np.int64)
    t = u - n0
    wts = np.stack([0.5 * (0.5 - t) ** 2, 0.75 - t * t, 0.5 * (0.5 + t) ** 2], 1)
    der = np.stack([t - 0.5, -2.0 * t, t + 0.5], 1)
    return n0, wts, der


def node_sums(frames, bins, r_max, sigma, pbc="orthorhombic", boxes=None):
    """H [n_nodes] (ordered-pair B-spline deposits) summed over the frames."""
    w, h, n_nodes, _ = geometry(bins, r_max, sigma)
    frames = np.asarray(frames, np.float64)
    H = np.zeros(n_nodes + 4)
    for f in range(len(frames)):
        box = None if boxes is None else boxes[f]
        _, _, d = pair_vectors(frames[f], pbc, box)
        r = np.sqrt((d * d).sum(1))
        n0, wts, _ = bspline2(r / h)
        ok = n0 + 1 < n_nodes + 3
        for o in range(3):
            np.add.at(H, n0[ok] - 1 + o + 1, 2.0 * wts[ok, o])
    return H[1:n_nodes + 1]


def soft_counts(frames, bins, r_max, sigma, pbc="orthorhombic", boxes=None):
    """c [bins] float64 summed over the frames (F, N, 3)."""
    _, _, _, G = geometry(bins, r_max, sigma)
    return G @ node_sums(frames, bins, r_max, sigma, pbc, boxes)


def soft_grad(frames, bins, r_max, sigma, dldc, pbc="orthorhombic", boxes=None):
    """dL/dx (F, N, 3) float64 given dL/dc [bins]."""
    w, h, n_nodes, G = geometry(bins, r_max, sigma)
    D = np.zeros(n_nodes + 4)
    D[1:n_nodes + 1] = np.asarray(dldc, np.float64) @ G
    frames = np.asarray(frames, np.float64)
    out = np.zeros_like(frames)
    for f in range(len(frames)):
        box = None if boxes is None else boxes[f]
        i, j, d = pair_vectors(frames[f], pbc, box)
        r = np.sqrt((d * d).sum(1))
        n0, _, der = bspline2(r / h)
        n0 = np.minimum(n0, n_nodes + 1)
        F = np.zeros_like(r)
        for o in range(3):
            F += D[n0 - 1 + o + 1] * der[:, o]
        F *= 2.0 / h                  # both orders, du/dr = 1/h
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


def split_rmsd_parts(x, y, fit, rmsd_sel):
    """Fit on atoms `fit`, RMSD over atoms `rmsd_sel` (float64), the
    composition of kabsch_rotation (geometry_math.rs:169-297) and the residual
    of rms.rs:656-672: R (x - cx) ~ (y - cy) from the fit set's centroids and
    cross-covariance H; returns (rmsd, R, cx, cy, H)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    fit = np.asarray(fit, np.int64)
    rsel = np.asarray(rmsd_sel, np.int64)
    cx = x[fit].mean(0)
    cy = y[fit].mean(0)
    H = (x[fit] - cx).T @ (y[fit] - cy)
    U, _, Vt = np.linalg.svd(H)
    D = np.eye(3)
    D[2, 2] = np.sign(np.linalg.det(Vt.T @ U.T))
    R = Vt.T @ D @ U.T
    res = (x[rsel] - cx) @ R.T - (y[rsel] - cy)
    return math.sqrt((res * res).sum() / len(rsel)), R, cx, cy, H


def split_rmsd_grad(x, y, fit, rmsd_sel):
    """dRMSD_B/dx for a fit on set A and the RMSD over set B (float64).

    With u_i = x_i - cx, v_i = y_i - cy, r_i = R u_i - v_i and
    E = sum_B |r_i|^2 (RMSD = sqrt(E / n_B)):
      dE/dx_i = 2 R^T r_i [i in B] - (2 / n_A) R^T sum_B r [i in A] + 2 G^T v_i [i in A].
    The last term is the adjoint through the rotation: the optimum keeps
    P = R H symmetric, so a perturbation dR = [k]x R obeys
    (tr(P) I - P) k = -2 axial(skew(R dH)); with W = sum_B u_i r_i^T,
    Q = R W, q = (Q23 - Q32, Q31 - Q13, Q12 - Q21), z = (tr(P) I - P)^-1 q and
    G = [z]x R, the rotation's share of dE is 2 tr(G dH).  For A = B it
    vanishes (q = 0), leaving the envelope-theorem form of fit_rmsd_grad.
    Duplicated indices add their terms."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    fit = np.asarray(fit, np.int64)
    rsel = np.asarray(rmsd_sel, np.int64)
    rmsd, R, cx, cy, H = split_rmsd_parts(x, y, fit, rsel)
    g = np.zeros_like(x)
    if rmsd == 0.0:
        return g
    u = x[rsel] - cx
    v = y[rsel] - cy
    r = u @ R.T - v
    np.add.at(g, rsel, 2.0 * (r @ R))
    Q = R @ (u.T @ r)
    q = np.array([Q[1, 2] - Q[2, 1], Q[2, 0] - Q[0, 2], Q[0, 1] - Q[1, 0]])
    P = R @ H
    P = 0.5 * (P + P.T)
    z = np.linalg.solve(np.trace(P) * np.eye(3) - P, q)
    Z = np.array([[0.0, -z[2], z[1]], [z[2], 0.0, -z[0]], [-z[1], z[0], 0.0]])
    G = Z @ R
    np.add.at(g, fit, 2.0 * ((y[fit] - cy) @ G))
    np.add.at(g, fit, -(2.0 / len(fit)) * (r.sum(0) @ R))
    return g / (2.0 * len(rsel) * rmsd)
