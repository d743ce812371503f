"""Soft RDF and RMSD adjoints on the GPU against the float64 definitions
(oracle/softrdf.py) and FP64 central finite differences."""
import ctypes
import math

import numpy as np
import pytest

import softrdf as S

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _beads(n, L, frames, seed, jitter=0.2):
    rng = np.random.default_rng(seed)
    base = rng.random((n, 3)) * L
    return (base[None] + rng.normal(scale=jitter, size=(frames, n, 3))).astype(np.float32)


def _plan(**kw):
    from warp_md_amd.autograd import SoftRdf
    return SoftRdf(**kw)


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


BINS, RMAX, SIG = 300, 12.0, 0.05


@pytest.fixture(scope="module")
def small():
    # 500 beads, L = 18.42 A < 2 r_max: minimum image over every pair (brute force)
    n = 500
    L = (n / 0.08) ** (1.0 / 3.0)
    x = _beads(n, L, 2, seed=4)
    return x, L


def test_brute_counts_match_oracle(small):
    x, L = small
    plan = _plan(bins=BINS, r_max=RMAX, sigma=SIG)
    c = plan.counts(torch.tensor(x, device="cuda"), np.array([L, L, L], np.float32)).cpu().numpy()
    assert plan.last_path[0] == 1
    ref = S.soft_counts(x, BINS, RMAX, SIG, "orthorhombic", [np.array([L, L, L])] * len(x))
    assert _rel(c, ref) < 1e-5


def test_brute_gradient_matches_oracle(small):
    x, L = small
    plan = _plan(bins=BINS, r_max=RMAX, sigma=SIG)
    rng = np.random.default_rng(1)
    dldc = rng.normal(size=BINS)
    xt = torch.tensor(x, device="cuda", requires_grad=True)
    c = plan.counts(xt, np.array([L, L, L], np.float32))
    c.backward(torch.tensor(dldc, device="cuda"))
    g = xt.grad.cpu().numpy()
    ref = S.soft_grad(x, BINS, RMAX, SIG, dldc, "orthorhombic", [np.array([L, L, L])] * len(x))
    assert _rel(g, ref) < 1e-4


def test_loss_gradient_matches_fp64_finite_differences(small):
    x, L = small
    plan = _plan(bins=BINS, r_max=RMAX, sigma=SIG)
    box = np.array([L, L, L], np.float32)
    xt = torch.tensor(x, device="cuda", requires_grad=True)
    g = plan.g(xt, box)
    t = torch.tensor(S.g_target(plan.bin_centers), device="cuda")
    loss = ((g - t) ** 2).sum()
    loss.backward()
    grad = xt.grad.cpu().numpy().astype(np.float64)
    x64 = x.astype(np.float64)
    boxes = [np.array([L, L, L], np.float64)] * len(x)
    vol = [L ** 3] * len(x)
    # the GPU loss value against the float64 definition
    l64, _ = S.rdf_loss(x64, BINS, RMAX, SIG, vol, "orthorhombic", boxes)
    assert float(loss.detach()) == pytest.approx(l64, rel=1e-5)
    rng = np.random.default_rng(2)
    h = 1e-4
    scale = np.abs(grad).max()
    for _ in range(12):
        f, a, d = rng.integers(len(x)), rng.integers(x.shape[1]), rng.integers(3)
        xp = x64.copy(); xp[f, a, d] += h
        xm = x64.copy(); xm[f, a, d] -= h
        fd = (S.rdf_loss(xp, BINS, RMAX, SIG, vol, "orthorhombic", boxes)[0]
              - S.rdf_loss(xm, BINS, RMAX, SIG, vol, "orthorhombic", boxes)[0]) / (2 * h)
        assert abs(grad[f, a, d] - fd) <= 1e-4 * scale, (f, a, d, grad[f, a, d], fd)


@pytest.mark.parametrize("n,tilt,k,rho", [(6000, None, 0, 0.08), (12000, None, 0, 0.08),
                                          (12000, 10.0, 0, 0.08), (12000, None, 1, 0.08),
                                          (12000, None, 1, 0.2), (20000, None, 0, 0.08)])
def test_cell_path_matches_brute(n, tilt, k, rho):
    """Cell enumeration against brute force; k = 1 gives cells of ~140
    (rho 0.08) and ~350 (rho 0.2) atoms: home atoms beyond the LDS slots."""
    L = (n / rho) ** (1.0 / 3.0)
    x = _beads(n, L, 2, seed=n)
    if tilt:
        M = np.array([[L, 0, 0], [L * math.tan(math.radians(tilt)), L, 0], [0, 0, L]], np.float32)
        x = (x / L) @ M  # fractional sites through the box
        box, pbc = M, "triclinic"
    else:
        box, pbc = np.array([L, L, L], np.float32), "orthorhombic"
    xt = torch.tensor(x, device="cuda")
    cell = _plan(bins=BINS, r_max=RMAX, sigma=SIG, pbc=pbc, cell_k=k)
    brute = _plan(bins=BINS, r_max=RMAX, sigma=SIG, pbc=pbc, cell_k=-1)
    cc = cell.counts(xt, box)
    assert cell.last_path[0] == 2
    cb = brute.counts(xt, box)
    assert brute.last_path[0] == 1
    assert _rel(cc.cpu().numpy(), cb.cpu().numpy()) < 1e-5
    dldc = torch.tensor(np.random.default_rng(5).normal(size=BINS), device="cuda")
    xa = xt.clone().requires_grad_(True)
    cell.counts(xa, box).backward(dldc)
    xb = xt.clone().requires_grad_(True)
    brute.counts(xb, box).backward(dldc)
    assert _rel(xa.grad.cpu().numpy(), xb.grad.cpu().numpy()) < 1e-4


def test_counts_add_over_frame_chunks():
    n = 6000
    L = (n / 0.08) ** (1.0 / 3.0)
    x = torch.tensor(_beads(n, L, 4, seed=8), device="cuda")
    box = np.array([L, L, L], np.float32)
    plan = _plan(bins=BINS, r_max=RMAX, sigma=SIG)
    whole = plan.counts(x, box)
    parts = plan.counts(x[:1], box) + plan.counts(x[1:], box)
    torch.testing.assert_close(whole, parts, rtol=1e-12, atol=1e-9)
    # a frame's gradient depends on that frame alone
    dldc = torch.tensor(np.random.default_rng(9).normal(size=BINS), device="cuda")
    xa = x.clone().requires_grad_(True)
    plan.counts(xa, box).backward(dldc)
    g_parts = []
    for sl in (slice(0, 1), slice(1, 4)):
        xs = x[sl].clone().requires_grad_(True)
        plan.counts(xs, box).backward(dldc)
        g_parts.append(xs.grad)
    g_parts = torch.cat(g_parts)
    assert _rel(g_parts.cpu().numpy(), xa.grad.cpu().numpy()) < 1e-5


def _chain(n, frames, seed):
    rng = np.random.default_rng(seed)
    y = np.cumsum(rng.normal(size=(n, 3)), 0) * 1.5
    out = []
    for _ in range(frames):
        R = np.linalg.qr(rng.normal(size=(3, 3)))[0]
        if np.linalg.det(R) < 0:
            R[:, 0] *= -1
        out.append(y @ R.T + rng.uniform(-5, 5, 3) + rng.normal(scale=0.5, size=y.shape))
    return y.astype(np.float32), np.array(out, np.float32)


@pytest.mark.parametrize("align", [True, False])
def test_rmsd_backward_matches_oracle(align):
    from warp_md_amd.autograd import FittedRmsd
    y, x = _chain(300, 3, seed=6)
    plan = FittedRmsd(y, align=align)
    xt = torch.tensor(x, device="cuda", requires_grad=True)
    r = plan(xt)
    w = torch.tensor([0.5, -1.0, 2.0], dtype=torch.float64, device="cuda")
    (r * w).sum().backward()
    g = xt.grad.cpu().numpy()
    for f in range(3):
        x64, y64 = x[f].astype(np.float64), y.astype(np.float64)
        if align:
            ref_r = S.fit_rmsd_parts(x64, y64)[0]
            ref_g = S.fit_rmsd_grad(x64, y64)
        else:
            diff = x64 - y64
            ref_r = math.sqrt((diff * diff).sum() / len(x64))
            ref_g = diff / (len(x64) * ref_r)
        assert float(r[f]) == pytest.approx(ref_r, abs=1e-5)
        assert _rel(g[f], float(w[f]) * ref_g) < 1e-4


def test_rmsd_backward_fd():
    from warp_md_amd.autograd import FittedRmsd
    y, x = _chain(120, 1, seed=3)
    plan = FittedRmsd(y)
    xt = torch.tensor(x, device="cuda", requires_grad=True)
    plan(xt).sum().backward()
    g = xt.grad.cpu().numpy()[0]
    x64, y64 = x[0].astype(np.float64), y.astype(np.float64)
    h = 1e-5
    for a in range(0, 120, 11):
        for d in range(3):
            xp = x64.copy(); xp[a, d] += h
            xm = x64.copy(); xm[a, d] -= h
            fd = (S.fit_rmsd_parts(xp, y64)[0] - S.fit_rmsd_parts(xm, y64)[0]) / (2 * h)
            assert abs(g[a, d] - fd) <= 1e-4 * np.abs(g).max()


def _chain300(frames, seed):
    """A 300-atom random-walk chain, frames randomly moved and noised (the
    split tests fit on every 10th atom)."""
    y, x = _chain(300, frames, seed=seed)
    return y, x


@pytest.mark.parametrize("align", [True, False])
def test_split_rmsd_backward_matches_oracle(align):
    """Fit on every 10th atom, RMSD over all 300: the adjoint through the
    rotation against the float64 restatement (itself pinned by FP64 central
    differences and torch autograd), and the forward against the oracle."""
    from warp_md_amd.autograd import FittedRmsd
    y, x = _chain300(4, seed=21)
    fit = np.arange(0, 300, 10)
    plan = FittedRmsd(y, align=align, fit_selection=fit)
    xt = torch.tensor(x, device="cuda", requires_grad=True)
    r = plan(xt)
    w = torch.tensor([0.5, -1.0, 2.0, 1.5], dtype=torch.float64, device="cuda")
    (r * w).sum().backward()
    g = xt.grad.cpu().numpy()
    allat = np.arange(300)
    for f in range(4):
        x64, y64 = x[f].astype(np.float64), y.astype(np.float64)
        if align:
            ref_r = S.split_rmsd_parts(x64, y64, fit, allat)[0]
            ref_g = S.split_rmsd_grad(x64, y64, fit, allat)
        else:
            diff = x64 - y64
            ref_r = math.sqrt((diff * diff).sum() / len(x64))
            ref_g = diff / (len(x64) * ref_r)
        assert float(r[f].detach()) == pytest.approx(ref_r, abs=1e-5)
        assert _rel(g[f], float(w[f]) * ref_g) < 1e-5


def test_split_rmsd_backward_fd_and_duplicates():
    """FP64 central differences of the oracle's split RMSD on the same f32
    inputs, with duplicated fit atoms and a partial RMSD set (atoms in both
    sets, fit-only and RMSD-only)."""
    from warp_md_amd import _native as N
    y, x = _chain300(1, seed=22)
    fit = np.concatenate([np.arange(0, 300, 10), [10, 20]]).astype(np.uint32)
    rsel = np.arange(100, 300, dtype=np.uint32)
    ctx = N.context(0)
    xyzw = np.zeros((300, 4), np.float32)
    xyzw[:, :3] = y
    h = ctypes.c_void_p()
    N.check(N.lib().wmd_rmsd_create(ctx.handle, N.ptr(fit), fit.size, N.ptr(rsel), rsel.size,
                                    N.ptr(xyzw), 300, 1, ctypes.byref(h)))
    try:
        import _gpu
        atoms = _gpu.rmsd_layout(h)
        store = torch.tensor(_gpu.host_store(x, atoms), device="cuda")
        stride = atoms.size
        r32 = torch.empty(1, dtype=torch.float32, device="cuda")
        r64 = torch.empty(1, dtype=torch.float64, device="cuda")
        rot = torch.empty((1, 9), dtype=torch.float64, device="cuda")
        cx = torch.empty((1, 3), dtype=torch.float64, device="cuda")
        V = lambda t: ctypes.c_void_p(t.data_ptr())
        N.check(N.lib().wmd_rmsd_run_device(h, V(store), stride, 1, V(r32), V(r64), None, V(rot), V(cx), None))
        gout = torch.ones(1, dtype=torch.float32, device="cuda")
        gs = torch.empty_like(store)
        N.check(N.lib().wmd_rmsd_backward_device(h, V(store), stride, 1, V(r64), V(rot), V(cx), V(gout),
                                                 V(gs), None))
        torch.cuda.synchronize()
        gs = gs.cpu().numpy()[0]
    finally:
        N.lib().wmd_rmsd_destroy(h)
    g = np.zeros((300, 3))
    valid = atoms != 0xFFFFFFFF
    np.add.at(g, atoms[valid].astype(np.int64), gs[:, valid].T.astype(np.float64))
    x64, y64 = x[0].astype(np.float64), y.astype(np.float64)
    assert float(r64.cpu()[0]) == pytest.approx(S.split_rmsd_parts(x64, y64, fit, rsel)[0], abs=1e-9)
    hh = 1e-6
    scale = np.abs(g).max()
    for a in [0, 10, 20, 50, 99, 100, 150, 200, 290, 299]:
        for d in range(3):
            xp = x64.copy(); xp[a, d] += hh
            xm = x64.copy(); xm[a, d] -= hh
            fd = (S.split_rmsd_parts(xp, y64, fit, rsel)[0] - S.split_rmsd_parts(xm, y64, fit, rsel)[0]) / (2 * hh)
            assert abs(g[a, d] - fd) <= 1e-5 * scale, (a, d, g[a, d], fd)


def test_split_unfitted_rmsd_is_rmsd_over_the_rmsd_set():
    """align=False with a fit set: the unfitted residual over the RMSD set
    alone (rms.rs:664-666), through RmsdPlan."""
    from warp_md_amd import RmsdPlan, System, Trajectory
    y, x = _chain300(3, seed=23)
    sysm = System.from_arrays({"name": np.array(["X"] * 300)}, positions0=y)
    rsel = np.arange(50, 300)
    out = RmsdPlan(rsel, align=False, fit_selection=np.arange(0, 300, 10)).run(
        Trajectory.from_numpy(x), sysm, device="cuda:0")
    diff = x[:, rsel].astype(np.float64) - y[rsel].astype(np.float64)
    want = np.sqrt((diff * diff).sum(axis=(1, 2)) / rsel.size)
    assert np.max(np.abs(out.astype(np.float64) - want)) <= 1e-5


def test_sigma_outside_window_rejected():
    with pytest.raises(RuntimeError, match="unsupported"):
        _plan(bins=300, r_max=12.0, sigma=0.2)


def test_node_histogram_beyond_lds_is_refused():
    """8,192 bins put 32,772 f64 nodes (262 KB) in a workgroup's LDS: the
    launch is refused with a clean 'unsupported' error before any kernel runs,
    on the brute-force and the cell-list paths."""
    for n in (500, 6000):
        L = (n / 0.08) ** (1.0 / 3.0)
        x = torch.tensor(_beads(n, L, 2, seed=31), device="cuda")
        plan = _plan(bins=8192, r_max=RMAX, sigma=0.002)
        with pytest.raises(RuntimeError, match="unsupported"):
            plan.counts(x, np.array([L, L, L], np.float32))


def test_run_to_run_spread_is_bounded():
    """The soft kernels add f64 node sums and f32 gradients with atomics whose
    order varies between runs: two identical runs agree to 1e-12 relative in
    the counts and 1e-5 relative in the gradients (the spread the design
    records), and the counts' total is order-independent to 1e-12."""
    n = 6000
    L = (n / 0.08) ** (1.0 / 3.0)
    x = torch.tensor(_beads(n, L, 3, seed=12), device="cuda")
    box = np.array([L, L, L], np.float32)
    plan = _plan(bins=BINS, r_max=RMAX, sigma=SIG)
    dldc = torch.tensor(np.random.default_rng(6).normal(size=BINS), device="cuda")
    out = []
    for _ in range(2):
        xa = x.clone().requires_grad_(True)
        c = plan.counts(xa, box)
        c.backward(dldc)
        out.append((c.detach().cpu().numpy(), xa.grad.cpu().numpy()))
    (c0, g0), (c1, g1) = out
    assert _rel(c1, c0) < 1e-12
    assert _rel(g1, g0) < 1e-5
    print(f"run-to-run: counts {_rel(c1, c0):.2e}, gradients {_rel(g1, g0):.2e} (relative to max)")


def test_backward_reuses_the_forward_cells():
    """A backward on the forward's own store takes its cell lists: the
    gradient equals a plan that sorts the frames again; a second forward on
    other positions in between (another store) makes the backward sort again;
    a changed box does too."""
    n = 6000
    L = (n / 0.08) ** (1.0 / 3.0)
    box = np.array([L, L, L], np.float32)
    xa = torch.tensor(_beads(n, L, 3, seed=21), device="cuda")
    xb = torch.tensor(_beads(n, L, 3, seed=22), device="cuda")
    dldc = torch.tensor(np.random.default_rng(23).normal(size=BINS), device="cuda")
    keep = _plan(bins=BINS, r_max=RMAX, sigma=SIG, keep_cells=True)
    fresh = _plan(bins=BINS, r_max=RMAX, sigma=SIG, keep_cells=False)

    def grad(plan, x, b=box):
        xg = x.clone().requires_grad_(True)
        plan.counts(xg, b).backward(dldc)
        return xg.grad.cpu().numpy()

    ga, gb = grad(fresh, xa), grad(fresh, xb)
    assert _rel(grad(keep, xa), ga) < 1e-5
    assert keep.last_path[0] == 2
    # forward a, forward b, then the backward of a (the cells hold b)
    x1 = xa.clone().requires_grad_(True)
    x2 = xb.clone().requires_grad_(True)
    c1 = keep.counts(x1, box)
    c2 = keep.counts(x2, box)
    c1.backward(dldc)
    c2.backward(dldc)
    assert _rel(x1.grad.cpu().numpy(), ga) < 1e-5
    assert _rel(x2.grad.cpu().numpy(), gb) < 1e-5
    # the same store pointer with another box: raw calls on one store
    stride = (n + 3) // 4 * 4
    soa = torch.zeros((3, 3, stride), dtype=torch.float32, device="cuda")
    soa[:, :, :n] = xa.permute(0, 2, 1)
    b1 = np.repeat(np.diag(box).reshape(1, 9), 3, 0).astype(np.float32)
    b2 = (b1 * np.float32(1.01)).astype(np.float32)
    keep.forward_store(soa, stride, n, b1)
    g_keep = keep.backward_store(soa, stride, n, b2, dldc)
    g_fresh = fresh.backward_store(soa, stride, n, b2, dldc)
    assert _rel(g_keep.cpu().numpy(), g_fresh.cpu().numpy()) < 1e-5
