"""Helpers for the GPU tests: direct C-ABI calls on the engine library."""
import ctypes

import numpy as np

from warp_md_amd import _native
from warp_md_amd._native import check, lib, ptr

PBC = {"none": 0, "orthorhombic": 1, "triclinic": 2}


def ctx():
    return _native.context(0)


def boxes9(box, n_frames, pbc):
    if pbc == "none" or box is None:
        return None
    b = np.asarray(box, np.float32)
    if b.shape == (3,):
        b = np.diag(b).reshape(1, 9)
    elif b.shape == (n_frames, 3) and pbc != "triclinic":
        b = np.stack([np.diag(r).reshape(9) for r in b])
    elif b.shape == (3, 3):
        b = b.reshape(1, 9)
    elif b.ndim == 2 and b.shape[1] == 3:
        b = np.stack([np.diag(r).reshape(9) for r in b])
    else:
        b = b.reshape(-1, 9)
    if b.shape[0] == 1:
        b = np.repeat(b, n_frames, 0)
    return np.ascontiguousarray(b, np.float32)


def rdf_counts(coords, sel_a, sel_b, bins, r_max, pbc="orthorhombic", box=None, flags=0,
               cell_k=0, full_chunk=True, split=None, evals=False, cells=False):
    """GPU counts through wmd_rdf_accumulate (full-system xyzw chunk).

    Returns (counts u64, edge, path) where path is 1 brute / 2 cells, with
    evals=True also the pair evaluations (wmd_rdf_pair_evaluations), and with
    cells=True also the cells per side of the last cell grid."""
    c = np.ascontiguousarray(coords, np.float32)
    F, N, _ = c.shape
    x4 = np.ones((F, N, 4), np.float32)
    x4[..., :3] = c
    a = np.ascontiguousarray(sel_a, np.uint32)
    b = np.ascontiguousarray(sel_b, np.uint32)
    bx = boxes9(box, F, pbc)
    h = ctypes.c_void_p()
    check(lib().wmd_rdf_create(ctx().handle, ptr(a), a.size, ptr(b), b.size, int(bins),
                               ctypes.c_float(r_max), PBC[pbc], flags, ctypes.byref(h)))
    try:
        if cell_k:
            check(lib().wmd_rdf_set_tuning(h, cell_k, 0))
        parts = [(0, F)] if not split else [(s, min(F, s + split)) for s in range(0, F, split)]
        for f0, f1 in parts:
            check(lib().wmd_rdf_accumulate(h, ptr(np.ascontiguousarray(x4[f0:f1])), N, f1 - f0,
                                           ptr(None if bx is None else np.ascontiguousarray(bx[f0:f1]))))
        counts = np.zeros(int(bins), np.uint64)
        edge = np.zeros(1, np.uint64)
        check(lib().wmd_rdf_read(h, ptr(counts), ptr(edge)))
        path = ctypes.c_int()
        nc = (ctypes.c_int * 3)()
        check(lib().wmd_rdf_last_path(h, ctypes.byref(path), nc))
        n_eval = ctypes.c_uint64()
        check(lib().wmd_rdf_pair_evaluations(h, ctypes.byref(n_eval)))
    finally:
        lib().wmd_rdf_destroy(h)
    out = (counts, int(edge[0]), path.value)
    if evals:
        out += (int(n_eval.value),)
    if cells:
        out += (tuple(nc),)
    return out


def rdf_counts_device(store, n, sel_len, bins, r_max, pbc, boxes, flags=0, cell_k=0):
    """Counts on a device-resident OW store (torch (F, 3, stride)), A == B == all n atoms."""
    import torch

    F, _, stride = store.shape
    sel = np.arange(n, dtype=np.uint32)
    h = ctypes.c_void_p()
    check(lib().wmd_rdf_create(ctx().handle, ptr(sel), n, ptr(sel), n, int(bins),
                               ctypes.c_float(r_max), PBC[pbc], flags, ctypes.byref(h)))
    try:
        if cell_k:
            check(lib().wmd_rdf_set_tuning(h, cell_k, 0))
        bx = boxes9(boxes, F, pbc)
        stream = torch.cuda.current_stream().cuda_stream
        check(lib().wmd_rdf_accumulate_device(h, ctypes.c_void_p(store.data_ptr()), stride, F,
                                              ptr(bx), ctypes.c_void_p(stream)))
        counts = np.zeros(int(bins), np.uint64)
        check(lib().wmd_rdf_read(h, ptr(counts), None))
        path = ctypes.c_int()
        check(lib().wmd_rdf_last_path(h, ctypes.byref(path), None))
    finally:
        lib().wmd_rdf_destroy(h)
    return counts, path.value


def rmsd(coords, fit_sel, rmsd_sel, ref_all, align=True):
    """GPU RMSD through wmd_rmsd_run: returns (f32 [F], err [F])."""
    c = np.ascontiguousarray(coords, np.float32)
    F, N, _ = c.shape
    x4 = np.ones((F, N, 4), np.float32)
    x4[..., :3] = c
    r4 = np.ones((ref_all.shape[0], 4), np.float32)
    r4[:, :3] = ref_all
    a = np.ascontiguousarray(fit_sel, np.uint32)
    h = ctypes.c_void_p()
    if rmsd_sel is None:
        check(lib().wmd_rmsd_create(ctx().handle, ptr(a), a.size, None, 0, ptr(r4), r4.shape[0],
                                    int(align), ctypes.byref(h)))
    else:
        b = np.ascontiguousarray(rmsd_sel, np.uint32)
        check(lib().wmd_rmsd_create(ctx().handle, ptr(a), a.size, ptr(b), b.size, ptr(r4),
                                    r4.shape[0], int(align), ctypes.byref(h)))
    try:
        out = np.zeros(F, np.float32)
        err = np.zeros(F, np.float64)
        check(lib().wmd_rmsd_run(h, ptr(x4), N, F, ptr(out), ptr(err)))
    finally:
        lib().wmd_rmsd_destroy(h)
    return out, err


def rmsd_plan_handle(fit_sel, rmsd_sel, ref_all, align=True):
    r4 = np.ones((ref_all.shape[0], 4), np.float32)
    r4[:, :3] = ref_all
    a = np.ascontiguousarray(fit_sel, np.uint32)
    h = ctypes.c_void_p()
    if rmsd_sel is None:
        check(lib().wmd_rmsd_create(ctx().handle, ptr(a), a.size, None, 0, ptr(r4), r4.shape[0],
                                    int(align), ctypes.byref(h)))
    else:
        b = np.ascontiguousarray(rmsd_sel, np.uint32)
        check(lib().wmd_rmsd_create(ctx().handle, ptr(a), a.size, ptr(b), b.size, ptr(r4),
                                    r4.shape[0], int(align), ctypes.byref(h)))
    return h


def rmsd_layout(h):
    n_store = ctypes.c_size_t()
    stride = ctypes.c_size_t()
    check(lib().wmd_rmsd_store_layout(h, None, 0, ctypes.byref(n_store), ctypes.byref(stride)))
    atoms = np.zeros(stride.value, np.uint32)
    check(lib().wmd_rmsd_store_layout(h, ptr(atoms), atoms.size, ctypes.byref(n_store),
                                      ctypes.byref(stride)))
    return atoms


def host_store(frames, atoms):
    """(F, N, 3) host frames -> (F, 3, stride) SoA store (numpy) for a plan's layout."""
    F = frames.shape[0]
    store = np.zeros((F, 3, atoms.size), np.float32)
    valid = atoms != 0xFFFFFFFF
    store[:, :, valid] = np.transpose(frames[:, atoms[valid].astype(np.int64), :], (0, 2, 1))
    return store


def rmsd_run_device(h, store_t):
    """Runs a plan on a device store (torch (F, 3, stride)); returns (f32, f64, err, rot)."""
    import torch

    F, _, stride = store_t.shape
    dev = store_t.device
    o32 = torch.zeros(F, dtype=torch.float32, device=dev)
    o64 = torch.zeros(F, dtype=torch.float64, device=dev)
    err = torch.zeros(F, dtype=torch.float64, device=dev)
    rot = torch.zeros((F, 9), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    V = lambda t: ctypes.c_void_p(t.data_ptr())
    check(lib().wmd_rmsd_run_device(h, V(store_t), stride, F, V(o32), V(o64), V(err), V(rot),
                                    None, ctypes.c_void_p(stream)))
    torch.cuda.synchronize()
    return (o32.cpu().numpy(), o64.cpu().numpy(), err.cpu().numpy(),
            rot.cpu().numpy().reshape(-1, 3, 3))


def rmsd_device(frames, fit_sel, rmsd_sel, ref_all):
    """Device path for small host inputs: the SoA store is built on the host and
    copied to the GPU.  Returns (f32, f64, err, rot) numpy arrays."""
    import torch

    h = rmsd_plan_handle(fit_sel, rmsd_sel, ref_all)
    try:
        atoms = rmsd_layout(h)
        store = torch.tensor(host_store(np.asarray(frames, np.float32), atoms), device="cuda")
        return rmsd_run_device(h, store)
    finally:
        lib().wmd_rmsd_destroy(h)
