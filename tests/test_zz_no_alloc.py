"""wmd_rmsd_run_device_noalloc (the allocation-free variant of
wmd_rmsd_run_device, partials in the library's fixed scratch array) against
wmd_rmsd_run_device through the C ABI: identical outputs (same kernels, same
tiling, so the same sums in the same order) on two calls in a row, the second
with a smaller input, for each plan mode; the size limit is rejected before
anything is enqueued."""
import ctypes

import numpy as np
import pytest

from warp_md_amd import workloads
from warp_md_amd._native import WMD_ERR_UNSUPPORTED, check, lib, ptr

import _gpu

pytestmark = pytest.mark.gpu


def _run(fn, h, store, F):
    import torch

    stride = store.shape[2]
    dev = store.device
    o32 = torch.full((F,), -1.0, dtype=torch.float32, device=dev)
    o64 = torch.full((F,), -1.0, dtype=torch.float64, device=dev)
    err = torch.full((F,), -1.0, dtype=torch.float64, device=dev)
    rot = torch.full((F, 9), -1.0, dtype=torch.float64, device=dev)
    cx = torch.full((F, 3), -1.0, dtype=torch.float64, device=dev)
    V = lambda t: ctypes.c_void_p(t.data_ptr())
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(fn(h, V(store), stride, F, V(o32), V(o64), V(err), V(rot), V(cx), stream))
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (o32, o64, err, rot, cx)]


@pytest.mark.parametrize("mode", ["same", "split", "raw"])
def test_noalloc_matches_allocating_entry_point(mode):
    import torch

    c = workloads.chain(n_ca=300, per_ca=10, n_frames=0, seed=21)
    ref = c["ref"]
    n = ref.shape[0]
    fit = c["ca"] if mode == "split" else np.arange(n, dtype=np.uint32)
    rsel = np.arange(n, dtype=np.uint32) if mode == "split" else None
    h = _gpu.rmsd_plan_handle(fit, rsel, ref, align=mode != "raw")
    try:
        atoms = _gpu.rmsd_layout(h)
        for F in (700, 130):  # the second call smaller than the first
            print(f"noalloc mode={mode} frames={F}", flush=True)
            store = workloads.chain_store_device(ref, atoms, F, seed=22 + F)
            want = _run(lib().wmd_rmsd_run_device, h, store, F)
            got = _run(lib().wmd_rmsd_run_device_noalloc, h, store, F)
            for g, w in zip(got, want):
                np.testing.assert_array_equal(g, w)
            assert np.all(got[0] > 0)
            del store
            torch.cuda.empty_cache()
    finally:
        lib().wmd_rmsd_destroy(h)


def test_noalloc_rejects_more_frames_than_its_scratch():
    import torch

    c = workloads.chain(n_ca=4, per_ca=1, n_frames=0, seed=23)
    ref = c["ref"]
    n = ref.shape[0]
    sel = np.arange(n, dtype=np.uint32)
    h = _gpu.rmsd_plan_handle(sel, None, ref)
    try:
        F = 200000  # 200,000 frames x 13 moments x 8 B > 16 MiB; never enqueued
        out = torch.zeros(1, dtype=torch.float32, device="cuda:0")
        st = lib().wmd_rmsd_run_device_noalloc(h, ctypes.c_void_p(out.data_ptr()), 4, F,
                                               ctypes.c_void_p(out.data_ptr()), None, None, None,
                                               None, None)
        assert st == WMD_ERR_UNSUPPORTED
        assert b"scratch" in lib().wmd_last_error()
    finally:
        lib().wmd_rmsd_destroy(h)
