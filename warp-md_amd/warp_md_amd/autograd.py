"""Differentiable soft RDF and fitted RMSD over the HIP engine.

Two ``torch.autograd.Function``s with hand-written backward kernels:

* ``SoftRdf`` -- the Gaussian-smoothed RDF counts c_k (``wmd_softrdf_forward``)
  and their adjoint dL/dx (``wmd_softrdf_backward``, which enumerates the pairs
  again instead of storing per-pair terms);
* ``FittedRmsd`` -- per-frame RMSD after the QCP superposition
  (``wmd_rmsd_run_device``) and its adjoint (``wmd_rmsd_backward_device``).

The reference has neither (warp-cg optimises its coarse-grained models
without gradients, crates/warp-cg/src/optimize/bo/mod.rs:237); the RDF
normalisation follows RdfPlan::finalize (crates/traj-engine/src/plans/rdf.rs:943-1016)
with use_volume, and the fitted RMSD is RmsdPlan's (rmsd/plan.rs:1161-1248).

Positions are (F, N, 3) float32 tensors on the engine's device.  Torch only
carries memory and streams here: every pair and moment is computed by the
library, on torch's current stream.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _native
from .executor import Device

_PBC = {"none": _native.PBC_NONE, "orthorhombic": _native.PBC_ORTHO,
        "triclinic": _native.PBC_TRICLINIC}


def _torch():
    import torch
    return torch


def _stream(device):
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _boxes9(box, n_frames, pbc):
    """Box argument -> ([F, 9] float32 row-major box vectors, volumes [F]) or (None, None)."""
    if pbc == "none":
        return None, None
    b = np.asarray(box.detach().cpu().numpy() if hasattr(box, "detach") else box, np.float64)
    # per-frame shapes are tested before the shared (3, 3) matrix, so that
    # three frames of box lengths are never read as one box matrix
    if b.shape == (3,):
        m = np.tile(np.diag(b).reshape(1, 9), (n_frames, 1))
    elif b.shape == (n_frames, 3) and pbc != "triclinic":
        m = np.zeros((n_frames, 9))
        m[:, 0], m[:, 4], m[:, 8] = b[:, 0], b[:, 1], b[:, 2]
    elif b.shape == (n_frames, 3, 3):
        m = b.reshape(n_frames, 9)
    elif b.shape == (3, 3):
        m = np.tile(b.reshape(1, 9), (n_frames, 1))
    else:
        raise RuntimeError(f"mismatch: box shape {b.shape} for {n_frames} frames")
    vol = np.abs(np.linalg.det(m.reshape(-1, 3, 3)))
    return np.ascontiguousarray(m, np.float32), vol


def _soa(x):
    """(F, N, 3) -> zero-padded SoA store (F, 3, stride), stride = N rounded up to 4."""
    torch = _torch()
    F, N, _ = x.shape
    stride = (N + 3) // 4 * 4
    soa = torch.zeros((F, 3, stride), dtype=torch.float32, device=x.device)
    soa[:, :, :N] = x.detach().to(torch.float32).permute(0, 2, 1)
    return soa, stride


def _check_positions(x, name="positions"):
    torch = _torch()
    if not isinstance(x, torch.Tensor) or x.dim() != 3 or x.shape[2] != 3:
        raise RuntimeError(f"mismatch: {name} must be a (frames, atoms, 3) tensor")
    if not x.is_cuda:
        raise RuntimeError(f"unsupported: {name} must live on the HIP device")


class SoftRdf:
    """Gaussian-smoothed RDF of one atom set (all atoms of the positions).

    Every ordered pair i != j deposits the quadratic B-spline B(r_ij / h - m)
    on nodes m h (h = w / 4, w = r_max / bins); bin k collects
    c_k = sum over |m h - r_k| <= 6 sigma of H_m w phi_sigma(r_k - m h),
    r_k = (k + 1/2) w; g_k = c_k / (frames shell_k rho) with
    rho = N (N - 1) / mean box volume.
    """

    def __init__(self, bins: int = 300, r_max: float = 12.0, sigma: float = 0.05,
                 pbc: str = "orthorhombic", device="auto", cell_k: int = 0,
                 keep_cells: bool = True):
        if pbc not in _PBC:
            raise RuntimeError(f"unsupported: pbc '{pbc}'")
        self.bins = int(bins)
        self.r_max = float(r_max)
        self.sigma = float(sigma)
        self.pbc = pbc
        self._ctx = Device.from_spec(device).context()
        h = ctypes.c_void_p()
        _native.check(_native.lib().wmd_softrdf_create(
            self._ctx.handle, self.bins, ctypes.c_float(self.r_max), ctypes.c_float(self.sigma),
            _PBC[pbc], ctypes.byref(h)))
        self._h = h
        if cell_k:
            _native.check(_native.lib().wmd_softrdf_set_tuning(h, int(cell_k)))
        # the backward of the forward that just ran on the same store takes its
        # cell lists (counts() owns its store; forward_store / backward_store
        # callers keep the store unchanged in between, or pass keep_cells=False)
        _native.check(_native.lib().wmd_softrdf_keep_cells(h, 1 if keep_cells else 0))
        w = np.float32(self.r_max) / np.float32(self.bins)
        self.bin_width = float(w)
        self.bin_centers = (np.arange(self.bins) + 0.5) * float(w)

    def __del__(self):
        h = getattr(self, "_h", None)
        nat = globals().get("_native")  # None while the interpreter shuts down
        if h is not None and nat is not None and nat._lib is not None:
            nat._lib.wmd_softrdf_destroy(h)
            self._h = None

    @property
    def last_path(self):
        p = ctypes.c_int()
        nc = (ctypes.c_int * 3)()
        _native.check(_native.lib().wmd_softrdf_last_path(self._h, ctypes.byref(p), nc))
        return int(p.value), [int(v) for v in nc]

    def set_tuning(self, cell_k: int):
        _native.check(_native.lib().wmd_softrdf_set_tuning(self._h, int(cell_k)))

    # raw device calls on an SoA store --------------------------------------
    def forward_store(self, soa, stride, n_atoms, boxes9):
        torch = _torch()
        c = torch.zeros(self.bins, dtype=torch.float64, device=soa.device)
        _native.check(_native.lib().wmd_softrdf_forward(
            self._h, ctypes.c_void_p(soa.data_ptr()), stride, n_atoms, soa.shape[0],
            _native.ptr(boxes9), ctypes.c_void_p(c.data_ptr()), _stream(soa.device)))
        return c

    def backward_store(self, soa, stride, n_atoms, boxes9, dcounts):
        torch = _torch()
        dc = dcounts.detach().to(torch.float64).contiguous()
        g = torch.zeros_like(soa)
        _native.check(_native.lib().wmd_softrdf_backward(
            self._h, ctypes.c_void_p(soa.data_ptr()), stride, n_atoms, soa.shape[0],
            _native.ptr(boxes9), ctypes.c_void_p(dc.data_ptr()), ctypes.c_void_p(g.data_ptr()),
            _stream(soa.device)))
        return g

    # differentiable API ------------------------------------------------------
    def counts(self, x, box=None):
        """Soft counts c [bins] (float64), differentiable in x (F, N, 3)."""
        _check_positions(x)
        boxes9, _ = _boxes9(box, x.shape[0], self.pbc)
        return _SoftRdfFn.apply(x, self, boxes9)

    def normalisation(self, n_atoms, n_frames, box=None, density=None):
        """frames * shell_k * rho, the divisor of g (float64 numpy)."""
        w = self.bin_width
        ri = np.arange(self.bins) * w
        shell = 4.0 / 3.0 * math.pi * ((ri + w) ** 3 - ri ** 3)
        if density is None:
            _, vol = _boxes9(box, n_frames, self.pbc)
            if vol is None:
                raise RuntimeError("invalid: without PBC pass the number density")
            rho = n_atoms * (n_atoms - 1) / (float(vol.sum()) / n_frames)
        else:
            rho = float(density) * (n_atoms - 1)
        return n_frames * shell * rho

    def g(self, x, box=None, density=None):
        """Smoothed g(r) [bins] (float64), differentiable in x."""
        torch = _torch()
        c = self.counts(x, box)
        norm = torch.as_tensor(self.normalisation(x.shape[1], x.shape[0], box, density),
                               dtype=torch.float64, device=c.device)
        return c / norm


def _torch_function():
    torch = _torch()

    class _SoftRdf(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, plan, boxes9):
            soa, stride = _soa(x)
            c = plan.forward_store(soa, stride, x.shape[1], boxes9)
            ctx.plan, ctx.soa, ctx.stride, ctx.n, ctx.boxes9 = plan, soa, stride, x.shape[1], boxes9
            ctx.x_dtype = x.dtype
            return c

        @staticmethod
        def backward(ctx, dc):
            g = ctx.plan.backward_store(ctx.soa, ctx.stride, ctx.n, ctx.boxes9, dc)
            gx = g[:, :, :ctx.n].permute(0, 2, 1).to(ctx.x_dtype)
            return gx, None, None

    class _Rmsd(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x, plan):
            store = plan.store(x)
            F = x.shape[0]
            dev = x.device
            r32 = torch.empty(F, dtype=torch.float32, device=dev)
            r64 = torch.empty(F, dtype=torch.float64, device=dev)
            rot = torch.empty((F, 9), dtype=torch.float64, device=dev)
            cx = torch.empty((F, 3), dtype=torch.float64, device=dev)
            _native.check(_native.lib().wmd_rmsd_run_device(
                plan._h, ctypes.c_void_p(store.data_ptr()), plan.stride, F,
                ctypes.c_void_p(r32.data_ptr()), ctypes.c_void_p(r64.data_ptr()), None,
                ctypes.c_void_p(rot.data_ptr()), ctypes.c_void_p(cx.data_ptr()), _stream(dev)))
            ctx.plan, ctx.store, ctx.r64, ctx.rot, ctx.cx = plan, store, r64, rot, cx
            ctx.shape, ctx.x_dtype = x.shape, x.dtype
            return r64

        @staticmethod
        def backward(ctx, gr):
            plan = ctx.plan
            F = ctx.shape[0]
            gout = gr.detach().to(torch.float32).contiguous()
            gs = torch.empty_like(ctx.store)
            _native.check(_native.lib().wmd_rmsd_backward_device(
                plan._h, ctypes.c_void_p(ctx.store.data_ptr()), plan.stride, F,
                ctypes.c_void_p(ctx.r64.data_ptr()), ctypes.c_void_p(ctx.rot.data_ptr()),
                ctypes.c_void_p(ctx.cx.data_ptr()), ctypes.c_void_p(gout.data_ptr()),
                ctypes.c_void_p(gs.data_ptr()), _stream(gs.device)))
            gx = torch.zeros(ctx.shape, dtype=torch.float32, device=gs.device)
            # an atom in both sets, or repeated in one, owns several slots: add them
            gx.index_add_(1, plan.atom_idx, gs[:, :, plan.slot_idx].permute(0, 2, 1))
            return gx.to(ctx.x_dtype), None

    return _SoftRdf, _Rmsd


class _Lazy:
    def __init__(self, i):
        self.i = i
        self.cls = None

    def apply(self, *a):
        if self.cls is None:
            self.cls = _torch_function()[self.i]
        return self.cls.apply(*a)


_SoftRdfFn = _Lazy(0)
_RmsdFn = _Lazy(1)


class FittedRmsd:
    """Per-frame RMSD to a fixed reference after the optimal superposition
    (align=True) or without it (align=False), over all N atoms; differentiable
    in the positions.  The reference is a constant.

    ``fit_selection`` (atom indices) fits on that subset and takes the RMSD
    over all atoms (the composition of rms.rs:643-676); its backward runs the
    adjoint through the rotation."""

    def __init__(self, reference, align: bool = True, device="auto", fit_selection=None):
        ref = np.asarray(reference.detach().cpu().numpy() if hasattr(reference, "detach")
                         else reference, np.float32)
        if ref.ndim != 2 or ref.shape[1] != 3:
            raise RuntimeError("mismatch: reference must be (atoms, 3)")
        self.n = ref.shape[0]
        self._ctx = Device.from_spec(device).context()
        L = _native.lib()
        sel = np.arange(self.n, dtype=np.uint32)
        xyzw = np.zeros((self.n, 4), np.float32)
        xyzw[:, :3] = ref
        h = ctypes.c_void_p()
        if fit_selection is None:
            _native.check(L.wmd_rmsd_create(self._ctx.handle, _native.ptr(sel), self.n, None, 0,
                                            _native.ptr(xyzw), self.n, int(bool(align)),
                                            ctypes.byref(h)))
        else:
            fit = np.ascontiguousarray(np.asarray(fit_selection, np.int64).astype(np.uint32))
            if fit.size and int(fit.max()) >= self.n:
                raise RuntimeError("invalid selection: atom index out of range")
            _native.check(L.wmd_rmsd_create(self._ctx.handle, _native.ptr(fit), fit.size,
                                            _native.ptr(sel), self.n, _native.ptr(xyzw), self.n,
                                            int(bool(align)), ctypes.byref(h)))
        self._h = h
        n_store, stride = ctypes.c_size_t(), ctypes.c_size_t()
        _native.check(L.wmd_rmsd_store_layout(h, None, 0, ctypes.byref(n_store), ctypes.byref(stride)))
        atoms = np.zeros(stride.value, np.uint32)
        _native.check(L.wmd_rmsd_store_layout(h, _native.ptr(atoms), atoms.size,
                                              ctypes.byref(n_store), ctypes.byref(stride)))
        self.stride = int(stride.value)
        valid = atoms != 0xFFFFFFFF
        torch = _torch()
        dev = torch.device("cuda", self._ctx.device)
        self.slot_idx = torch.as_tensor(np.nonzero(valid)[0], dtype=torch.long, device=dev)
        self.atom_idx = torch.as_tensor(atoms[valid].astype(np.int64), dtype=torch.long, device=dev)

    def __del__(self):
        h = getattr(self, "_h", None)
        nat = globals().get("_native")  # None while the interpreter shuts down
        if h is not None and nat is not None and nat._lib is not None:
            nat._lib.wmd_rmsd_destroy(h)
            self._h = None

    def store(self, x):
        """(F, N, 3) positions -> the plan's SoA store (F, 3, stride)."""
        torch = _torch()
        st = torch.zeros((x.shape[0], 3, self.stride), dtype=torch.float32, device=x.device)
        st[:, :, self.slot_idx] = x.detach().to(torch.float32)[:, self.atom_idx, :].permute(0, 2, 1)
        return st

    def __call__(self, x):
        _check_positions(x)
        if x.shape[1] != self.n:
            raise RuntimeError(f"mismatch: {x.shape[1]} atoms for a {self.n}-atom reference")
        return _RmsdFn.apply(x, self)


def g_target(r):
    """Structure-matching target 1 + 0.3 exp(-(r - 4)^2) cos(2 r)."""
    return 1.0 + 0.3 * np.exp(-(np.asarray(r) - 4.0) ** 2) * np.cos(2.0 * np.asarray(r))
