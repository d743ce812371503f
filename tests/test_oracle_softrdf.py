"""Pins of the float64 soft-RDF / RMSD-adjoint definitions (oracle/softrdf.py)."""
import math

import numpy as np
import pytest

import softrdf as S


def _beads(n, L, frames, seed):
    rng = np.random.default_rng(seed)
    base = rng.random((n, 3)) * L
    return base[None] + rng.normal(scale=0.3, size=(frames, n, 3))


def test_single_pair_closed_form():
    # one pair at r: 2 x B-spline deposits on 3 nodes, each spread by the Gaussian
    bins, r_max, sigma = 100, 4.0, 0.05
    w = r_max / bins
    h = w / S.NODES_PER_BIN
    r = 1.2345
    x = np.array([[[0.0, 0.0, 0.0], [r, 0.0, 0.0]]])
    c = S.soft_counts(x, bins, r_max, sigma, pbc="none")
    n0 = int(round(r / h))
    t = r / h - n0
    B = [0.5 * (0.5 - t) ** 2, 0.75 - t * t, 0.5 * (0.5 + t) ** 2]
    for k in range(bins):
        rk = (k + 0.5) * w
        want = 0.0
        for o in range(3):
            d = rk - (n0 - 1 + o) * h
            if abs(d) <= 6 * sigma:
                want += 2 * B[o] * w / (sigma * math.sqrt(2 * math.pi)) * math.exp(-d * d / (2 * sigma ** 2))
        assert c[k] == pytest.approx(want, rel=1e-12, abs=1e-300)


@pytest.mark.parametrize("sigma_over_w", [0.6, 1.25, 2.0])
def test_mass_conservation(sigma_over_w):
    # a pair away from 0 and r_max adds 2 (both orders): the B-spline is a
    # partition of unity and the sampled Gaussian sums to 1 up to the
    # Poisson-summation alias 2 exp(-2 pi^2 sigma^2 / w^2) and the 6 sigma cut
    bins, r_max = 300, 12.0
    w = r_max / bins
    sigma = sigma_over_w * w
    rng = np.random.default_rng(7)
    for r in rng.uniform(1.0, 11.0, size=20):
        x = np.array([[[0.0, 0.0, 0.0], [0.0, r, 0.0]]])
        tot = S.soft_counts(x, bins, r_max, sigma, pbc="none").sum()
        trunc = math.erfc(6.0 / math.sqrt(2)) * 2
        alias = 4 * math.exp(-2 * math.pi ** 2 * sigma ** 2 / w ** 2)
        assert abs(tot - 2.0) <= 3 * (trunc + alias) + 1e-12


def test_effective_kernel_is_gaussian():
    # the deposit-then-spread kernel is phi_sigma convolved with the quadratic
    # B-spline of spacing h: close to the Gaussian of variance sigma^2 + h^2 / 4
    bins, r_max, sigma = 300, 12.0, 0.05
    w = r_max / bins
    h = w / S.NODES_PER_BIN
    se = math.sqrt(sigma ** 2 + h * h / 4)
    for r in (3.01, 5.5437, 8.123):
        x = np.array([[[0.0, 0.0, 0.0], [r, 0.0, 0.0]]])
        c = S.soft_counts(x, bins, r_max, sigma, pbc="none")
        rk = (np.arange(bins) + 0.5) * w
        g = 2 * w / (se * math.sqrt(2 * math.pi)) * np.exp(-(rk - r) ** 2 / (2 * se * se))
        assert np.abs(c - g).max() <= 2e-3 * g.max()


def test_min_image_matches_round_half_away():
    L = 10.0
    x = np.array([[0.0, 0.0, 0.0], [5.0, 0.0, 0.0], [9.9, 0.0, 0.0]])
    _, _, d = S.pair_vectors(x, "orthorhombic", np.array([L, L, L]))
    # 5.0 / 10 = 0.5 rounds away from zero -> -5.0
    assert d[0, 0] == -5.0
    assert d[1, 0] == pytest.approx(-0.1)
    assert d[2, 0] == pytest.approx(4.9)


@pytest.mark.parametrize("pbc", ["none", "orthorhombic", "triclinic"])
def test_gradient_matches_finite_differences(pbc):
    n, L, frames = 40, 7.0, 2
    bins, r_max, sigma = 60, 3.3, 0.07
    x = _beads(n, L, frames, seed=11)
    if pbc == "triclinic":
        M = np.array([[L, 0, 0], [1.2, L, 0], [0.4, -0.7, L]])
        boxes = [M] * frames
        vol = [abs(np.linalg.det(M))] * frames
    else:
        boxes = [np.array([L, L, L])] * frames
        vol = [L ** 3] * frames
    loss, dldc = S.rdf_loss(x, bins, r_max, sigma, vol, pbc, boxes)
    g = S.soft_grad(x, bins, r_max, sigma, dldc, pbc, boxes)
    rng = np.random.default_rng(3)
    h = 1e-5
    for _ in range(25):
        f, a, d = rng.integers(frames), rng.integers(n), rng.integers(3)
        xp = x.copy(); xp[f, a, d] += h
        xm = x.copy(); xm[f, a, d] -= h
        fd = (S.rdf_loss(xp, bins, r_max, sigma, vol, pbc, boxes)[0]
              - S.rdf_loss(xm, bins, r_max, sigma, vol, pbc, boxes)[0]) / (2 * h)
        scale = np.abs(g).max()
        assert abs(fd - g[f, a, d]) <= 1e-6 * scale + 1e-6 * abs(fd), (fd, g[f, a, d])


def test_translation_invariance_periodic():
    n, L = 50, 8.0
    x = _beads(n, L, 1, seed=5)
    bins, r_max, sigma = 50, 3.5, 0.08
    box = [np.array([L, L, L])]
    c0 = S.soft_counts(x, bins, r_max, sigma, "orthorhombic", box)
    c1 = S.soft_counts(x + np.array([3.1, -7.7, 0.4]), bins, r_max, sigma, "orthorhombic", box)
    np.testing.assert_allclose(c0, c1, rtol=1e-9, atol=1e-9)


def test_fit_rmsd_gradient_matches_finite_differences():
    rng = np.random.default_rng(9)
    y = rng.normal(size=(30, 3)) * 4
    R = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    if np.linalg.det(R) < 0:
        R[:, 0] *= -1
    x = y @ R.T + rng.normal(scale=0.4, size=y.shape) + np.array([2.0, -1.0, 0.5])
    g = S.fit_rmsd_grad(x, y)
    h = 1e-6
    for a in range(0, 30, 3):
        for d in range(3):
            xp = x.copy(); xp[a, d] += h
            xm = x.copy(); xm[a, d] -= h
            fd = (S.fit_rmsd_parts(xp, y)[0] - S.fit_rmsd_parts(xm, y)[0]) / (2 * h)
            assert fd == pytest.approx(g[a, d], rel=1e-6, abs=1e-9)


def test_fit_rmsd_zero_for_rigid_transform():
    rng = np.random.default_rng(1)
    y = rng.normal(size=(20, 3))
    R = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    if np.linalg.det(R) < 0:
        R[:, 0] *= -1
    assert S.fit_rmsd_parts(y @ R.T + 3.0, y)[0] < 1e-12


def _torch_loss(x, bins, r_max, sigma, L):
    """The soft-RDF loss as dense float64 torch operations (autograd gives dL/dx).

    Independent of the oracle's hand-written adjoint: pairs by broadcasting,
    round-half-away minimum image, B-spline weights scattered onto the nodes
    with index_add, the node-to-bin map and the normalisation from the
    definitions."""
    import torch

    F, n, _ = x.shape
    w, h, n_nodes, G = S.geometry(bins, r_max, sigma)
    Gt = torch.as_tensor(G)
    iu, ju = np.triu_indices(n, k=1)
    iu = torch.as_tensor(iu)
    ju = torch.as_tensor(ju)
    H = torch.zeros(n_nodes + 4, dtype=torch.float64)
    for f in range(F):
        d = x[f, ju] - x[f, iu]
        q = d / L
        d = d - torch.sign(q) * torch.floor(torch.abs(q) + 0.5) * L
        r = torch.sqrt((d * d).sum(1))
        u = r / h
        n0 = torch.round(u).detach()  # ties have measure zero
        t = u - n0
        idx = n0.long()
        ok = idx + 1 < n_nodes + 3
        wts = (0.5 * (0.5 - t) ** 2, 0.75 - t * t, 0.5 * (0.5 + t) ** 2)
        for o in range(3):
            H = H.index_add(0, idx[ok] - 1 + o + 1, 2.0 * wts[o][ok])
    c = Gt @ H[1:n_nodes + 1]
    norm = torch.as_tensor(S.normalisation(n, F, bins, r_max, [L ** 3] * F))
    rk = torch.as_tensor((np.arange(bins) + 0.5) * w)
    gt = 1.0 + 0.3 * torch.exp(-(rk - 4.0) ** 2) * torch.cos(2.0 * rk)
    return ((c / norm - gt) ** 2).sum()


def test_oracle_gradient_matches_dense_torch_autograd():
    import torch

    n, L, frames = 60, 8.0, 2
    bins, r_max, sigma = 80, 3.9, 0.06
    x = _beads(n, L, frames, seed=21)
    boxes = [np.array([L, L, L])] * frames
    loss, dldc = S.rdf_loss(x, bins, r_max, sigma, [L ** 3] * frames, "orthorhombic", boxes)
    g = S.soft_grad(x, bins, r_max, sigma, dldc, "orthorhombic", boxes)
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    lt = _torch_loss(xt, bins, r_max, sigma, L)
    lt.backward()
    assert float(lt) == pytest.approx(loss, rel=1e-12)
    gt = xt.grad.numpy()
    assert np.max(np.abs(gt - g)) <= 1e-10 * np.max(np.abs(g))


def _chain300(seed, noise=0.5):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(30, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ca = np.cumsum(3.8 * d, axis=0)
    y = np.repeat(ca, 10, axis=0) + 1.5 * rng.normal(size=(300, 3)) * (np.arange(300) % 10 != 0)[:, None]
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, a, b, c = q
    R = np.array([[1 - 2 * (b * b + c * c), 2 * (a * b - c * w), 2 * (a * c + b * w)],
                  [2 * (a * b + c * w), 1 - 2 * (a * a + c * c), 2 * (b * c - a * w)],
                  [2 * (a * c - b * w), 2 * (b * c + a * w), 1 - 2 * (a * a + b * b)]])
    x = y @ R.T + 5.0 + noise * rng.normal(size=(300, 3))
    return x, y


def test_split_rmsd_gradient_matches_finite_differences():
    """300-atom chain fitted on every 10th atom, RMSD over all 300 atoms:
    the adjoint through the rotation against FP64 central differences."""
    x, y = _chain300(11)
    fit = np.arange(0, 300, 10)
    rsel = np.arange(300)
    g = S.split_rmsd_grad(x, y, fit, rsel)
    h = 1e-6
    worst = 0.0
    scale = np.abs(g).max()
    for i in list(range(0, 300, 7)) + [0, 10, 290]:
        for d in range(3):
            xp = x.copy()
            xm = x.copy()
            xp[i, d] += h
            xm[i, d] -= h
            fd = (S.split_rmsd_parts(xp, y, fit, rsel)[0] - S.split_rmsd_parts(xm, y, fit, rsel)[0]) / (2 * h)
            worst = max(worst, abs(fd - g[i, d]))
    assert worst <= 1e-7 * scale, (worst, scale)
    # the rotation term matters here: dropping it fails the same check
    assert np.abs(g - _without_rotation_term(x, y, fit, rsel)).max() > 1e-3 * scale


def _without_rotation_term(x, y, fit, rsel):
    rmsd, R, cx, cy, _ = S.split_rmsd_parts(x, y, fit, rsel)
    r = (x[rsel] - cx) @ R.T - (y[rsel] - cy)
    g = np.zeros_like(x)
    np.add.at(g, rsel, 2.0 * (r @ R))
    np.add.at(g, fit, -(2.0 / len(fit)) * (r.sum(0) @ R))
    return g / (2.0 * len(rsel) * rmsd)


def test_split_rmsd_gradient_matches_torch_autograd():
    """A dense float64 torch restatement (autograd through torch.linalg.svd)."""
    import torch

    x, y = _chain300(12, noise=0.8)
    fit = np.concatenate([np.arange(0, 300, 10), [10, 20]])  # duplicates weight the fit
    rsel = np.concatenate([np.arange(150, 300), [151]])
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    yt = torch.tensor(y, dtype=torch.float64)
    fa = torch.as_tensor(fit)
    rb = torch.as_tensor(rsel)
    cx = xt[fa].mean(0)
    cy = yt[fa].mean(0)
    H = (xt[fa] - cx).T @ (yt[fa] - cy)
    U, _, Vh = torch.linalg.svd(H)
    dsign = torch.sign(torch.linalg.det(Vh.T @ U.T)).detach()
    D = torch.diag(torch.stack([torch.ones((), dtype=torch.float64), torch.ones((), dtype=torch.float64), dsign]))
    R = Vh.T @ D @ U.T
    res = (xt[rb] - cx) @ R.T - (yt[rb] - cy)
    rmsd = torch.sqrt((res * res).sum() / len(rsel))
    rmsd.backward()
    g = S.split_rmsd_grad(x, y, fit, rsel)
    assert abs(rmsd.item() - S.split_rmsd_parts(x, y, fit, rsel)[0]) < 1e-12
    np.testing.assert_allclose(g, xt.grad.numpy(), rtol=0, atol=1e-10 * np.abs(g).max())


def test_split_rmsd_gradient_same_sets_is_envelope_form():
    x, y = _chain300(13)
    sel = np.arange(300)
    np.testing.assert_allclose(S.split_rmsd_grad(x, y, sel, sel), S.fit_rmsd_grad(x, y), rtol=0,
                               atol=1e-12)
