"""ctypes binding of the gfx950 engine (``include/wmd_hip.h``).

Loads ``warp_md_amd/lib/libwmd_hip.so`` built in-tree by ``make -C warp-md_amd``
(or ``__graft_entry__.build()``).  There is no fallback: if the library is
missing, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
# WMD_HIP_LIB names another build of the same library (for example one built
# with different compile-time options to compare kernels side by side).
LIB_PATH = os.environ.get("WMD_HIP_LIB") or os.path.join(HERE, "lib", "libwmd_hip.so")

WMD_OK = 0
WMD_ERR_UNSUPPORTED = 1
WMD_ERR_MISMATCH = 2
WMD_ERR_PARSE = 3
WMD_ERR_DEVICE = 4
WMD_ERR_OOM = 5
WMD_ERR_INVALID = 6

PBC_NONE, PBC_ORTHO, PBC_TRICLINIC = 0, 1, 2

BATCH_RDF = 0x1
BATCH_RMSD = 0x2
BATCH_NO_GRAPH = 0x4

RDF_EDGE_DIAGNOSTICS = 0x1
RDF_PLANAR_XY = 0x4
RDF_NO_INTRAMOL = 0x8
RDF_FORCE_BRUTE = 0x2
RDF_USE_VOLUME = 0x1
RDF_RAW = 0x2
RDF_NUMBER_DENSITY = 0x4
RDF_INTEGRAL = 0x8
RDF_FINALIZE_XY = 0x10

# TrajError Display prefixes (crates/traj-core/src/error.rs:13-24); the
# reference surfaces every TrajError as RuntimeError(str(err)) (traj-py pack.rs:1188-1190).
_PREFIX = {
    WMD_ERR_UNSUPPORTED: "unsupported",
    WMD_ERR_MISMATCH: "mismatch",
    WMD_ERR_PARSE: "parse error",
    WMD_ERR_DEVICE: "device error",
    WMD_ERR_OOM: "device error: out of memory",
    WMD_ERR_INVALID: "invalid",
}

# Every symbol include/wmd_hip.h declares.
SYMBOLS = (
    "wmd_last_error", "wmd_version", "wmd_device_count",
    "wmd_ctx_create", "wmd_ctx_destroy", "wmd_ctx_synchronize", "wmd_ctx_stream",
    "wmd_ctx_cu_count", "wmd_ctx_set_staging_bytes", "wmd_ctx_profile", "wmd_ctx_profile_read",
    "wmd_rdf_create", "wmd_rdf_io_selection", "wmd_rdf_accumulate",
    "wmd_rdf_accumulate_selected", "wmd_rdf_accumulate_device", "wmd_rdf_read", "wmd_rdf_pair_evaluations",
    "wmd_rdf_reset", "wmd_rdf_destroy", "wmd_rdf_last_path", "wmd_rdf_set_tuning",
    "wmd_rdf_finalize", "wmd_rdf_create_f64", "wmd_sf_finalize",
    "wmd_rmsd_create", "wmd_rmsd_store_layout", "wmd_rmsd_run", "wmd_rmsd_run_device",
    "wmd_rmsd_run_device_noalloc",
    "wmd_rmsd_backward_device", "wmd_rmsd_destroy", "wmd_qcp_host",
    "wmd_softrdf_create", "wmd_softrdf_forward", "wmd_softrdf_backward",
    "wmd_softrdf_last_path", "wmd_softrdf_set_tuning", "wmd_softrdf_keep_cells", "wmd_softrdf_destroy",
    "wmd_batch_create", "wmd_batch_store_layout", "wmd_batch_set_boxes", "wmd_batch_set_tuning",
    "wmd_batch_autotune", "wmd_batch_run_device", "wmd_batch_read_counts", "wmd_batch_reset",
    "wmd_batch_plan_info", "wmd_batch_slice_times", "wmd_batch_destroy", "wmd_autotune_export", "wmd_autotune_import",
    "wmd_pairwise_create", "wmd_pairwise_create_dme", "wmd_pairwise_append",
    "wmd_pairwise_append_boxes", "wmd_pairwise_frames", "wmd_pairwise_finalize",
    "wmd_pairwise_store_stride", "wmd_pairwise_run_device", "wmd_pairwise_destroy",
    "wmd_align_create", "wmd_align_set_reference", "wmd_align_run", "wmd_align_run_device",
    "wmd_align_destroy", "wmd_rdf_create_groups", "wmd_pairwise_kmeans", "wmd_dbscan",
)


class WmdError(RuntimeError):
    """Engine error; str() follows the reference's TrajError display."""

    def __init__(self, status: int, message: str):
        self.status = status
        prefix = _PREFIX.get(status, "error")
        super().__init__(f"{prefix}: {message}")


_lib = None
_lock = threading.Lock()


def library_built() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    """The loaded engine library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise WmdError(WMD_ERR_UNSUPPORTED,
                           f"HIP engine library not built ({LIB_PATH}); "
                           "run `make -C warp-md_amd` or __graft_entry__.build()")
        # One HIP runtime per process: torch wheels carry their own
        # libamdhip64.so.7.  Loading torch first makes the engine's dependency
        # (same SONAME) resolve to that already-loaded copy, so device memory
        # and streams are shared with torch instead of a second runtime
        # claiming the GPU.  WMD_NO_TORCH=1 skips this for torch-free hosts.
        if os.environ.get("WMD_NO_TORCH", "0") != "1":
            try:
                import torch  # noqa: F401
            except ImportError:
                pass
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        S = ctypes.c_size_t
        I = ctypes.c_int
        PP = ctypes.POINTER(ctypes.c_void_p)
        St = ctypes.c_int
        sig = {
            "wmd_last_error": ([], ctypes.c_char_p),
            "wmd_version": ([], ctypes.c_char_p),
            "wmd_device_count": ([ctypes.POINTER(I)], St),
            "wmd_ctx_create": ([I, PP], St),
            "wmd_ctx_destroy": ([P], None),
            "wmd_ctx_synchronize": ([P], St),
            "wmd_ctx_stream": ([P], P),
            "wmd_ctx_cu_count": ([P], I),
            "wmd_ctx_set_staging_bytes": ([P, S, ctypes.POINTER(S)], St),
            "wmd_ctx_profile": ([P, I], St),
            "wmd_ctx_profile_read": ([P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(I)], St),
            "wmd_rdf_create": ([P, P, S, P, S, S, ctypes.c_float, I, ctypes.c_uint32, PP], St),
            "wmd_rdf_io_selection": ([P, P, S, ctypes.POINTER(S)], St),
            "wmd_rdf_accumulate": ([P, P, S, S, P], St),
            "wmd_rdf_accumulate_selected": ([P, P, S, P], St),
            "wmd_rdf_accumulate_device": ([P, P, S, S, P, P], St),
            "wmd_rdf_read": ([P, P, P], St),
            "wmd_rdf_pair_evaluations": ([P, P], St),
            "wmd_rdf_reset": ([P], St),
            "wmd_rdf_destroy": ([P], None),
            "wmd_rdf_last_path": ([P, ctypes.POINTER(I), P], St),
            "wmd_rdf_set_tuning": ([P, I, I], St),
            "wmd_rdf_finalize": ([P, S, ctypes.c_float, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, S, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_uint32, P, P, P], St),
            "wmd_rdf_create_f64": ([P, P, S, S, ctypes.c_double, I, ctypes.c_double, PP], St),
            "wmd_sf_finalize": ([P, S, ctypes.c_double, ctypes.c_double, S, ctypes.c_double, I, S,
                                 ctypes.c_double, P, P, P, P], St),
            "wmd_rmsd_create": ([P, P, S, P, S, P, S, I, PP], St),
            "wmd_rmsd_store_layout": ([P, P, S, ctypes.POINTER(S), ctypes.POINTER(S)], St),
            "wmd_rmsd_run": ([P, P, S, S, P, P], St),
            "wmd_rmsd_run_device": ([P, P, S, S, P, P, P, P, P, P], St),
            "wmd_rmsd_run_device_noalloc": ([P, P, S, S, P, P, P, P, P, P], St),
            "wmd_rmsd_backward_device": ([P, P, S, S, P, P, P, P, P, P], St),
            "wmd_rmsd_destroy": ([P], None),
            "wmd_softrdf_create": ([P, S, ctypes.c_float, ctypes.c_float, I, PP], St),
            "wmd_softrdf_forward": ([P, P, S, S, S, P, P, P], St),
            "wmd_softrdf_backward": ([P, P, S, S, S, P, P, P, P], St),
            "wmd_softrdf_last_path": ([P, ctypes.POINTER(I), P], St),
            "wmd_softrdf_set_tuning": ([P, I], St),
            "wmd_softrdf_keep_cells": ([P, I], St),
            "wmd_softrdf_destroy": ([P], None),
            "wmd_qcp_host": ([P, ctypes.c_double, ctypes.c_double, ctypes.c_double, P, P, P,
                              ctypes.POINTER(I)], St),
            "wmd_batch_create": ([P, S, P, P, S, ctypes.c_float, I, ctypes.c_uint32, PP], St),
            "wmd_batch_store_layout": ([P, P, P, ctypes.POINTER(ctypes.c_uint64)], St),
            "wmd_batch_set_boxes": ([P, P], St),
            "wmd_batch_set_tuning": ([P, P, P, P], St),
            "wmd_batch_autotune": ([P, P, I, P], St),
            "wmd_batch_run_device": ([P, P, P, P, P, P], St),
            "wmd_batch_read_counts": ([P, P], St),
            "wmd_batch_reset": ([P], St),
            "wmd_batch_plan_info": ([P, P, P, P, ctypes.POINTER(ctypes.c_longlong)], St),
            "wmd_batch_slice_times": ([P, P], St),
            "wmd_batch_destroy": ([P], None),
            "wmd_autotune_export": ([ctypes.c_char_p, S, ctypes.POINTER(S)], St),
            "wmd_autotune_import": ([ctypes.c_char_p], St),
            "wmd_pairwise_create": ([P, P, S, I, PP], St),
            "wmd_pairwise_create_dme": ([P, P, S, I, PP], St),
            "wmd_pairwise_append": ([P, P, S, S], St),
            "wmd_pairwise_append_boxes": ([P, P, S, S, P], St),
            "wmd_pairwise_frames": ([P, ctypes.POINTER(S)], St),
            "wmd_pairwise_finalize": ([P, P], St),
            "wmd_pairwise_store_stride": ([P, ctypes.POINTER(S)], St),
            "wmd_pairwise_run_device": ([P, P, S, S, P, P, P], St),
            "wmd_pairwise_destroy": ([P], None),
            "wmd_align_create": ([P, P, S, P, P, S, I, PP], St),
            "wmd_align_set_reference": ([P, P, S], St),
            "wmd_align_run": ([P, P, S, S, P, P], St),
            "wmd_align_run_device": ([P, P, S, S, P, P, P, P], St),
            "wmd_align_destroy": ([P], None),
            "wmd_rdf_create_groups": ([P, S, P, P, P, P, S, P, S, P, P, S, ctypes.c_float, I,
                                       ctypes.c_uint32, PP], St),
            "wmd_pairwise_kmeans": ([P, S, S, ctypes.c_double, ctypes.c_uint64, P, P, P], St),
            "wmd_dbscan": ([P, S, ctypes.c_double, S, P, P, P, ctypes.POINTER(S)], St),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
        return L


def check(status: int) -> None:
    if status != WMD_OK:
        msg = lib().wmd_last_error()
        raise WmdError(status, msg.decode() if msg else "")


def ptr(a):
    """Address of a numpy array (or None) as a c_void_p that holds a reference
    to the array, so a temporary passed as ``f(ptr(np.ascontiguousarray(x)))``
    stays alive until the call returns."""
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.c_void_p)


def device_count() -> int:
    n = ctypes.c_int(0)
    check(lib().wmd_device_count(ctypes.byref(n)))
    return int(n.value)


class Context:
    """A HIP context: one device, one stream (the GpuContext replacement)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().wmd_ctx_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = int(device)

    @property
    def stream(self):
        return lib().wmd_ctx_stream(self.handle)

    @property
    def cu_count(self) -> int:
        return int(lib().wmd_ctx_cu_count(self.handle))

    def synchronize(self):
        check(lib().wmd_ctx_synchronize(self.handle))

    def set_staging_bytes(self, nbytes: int) -> int:
        """Sets the size of the host-chunk pipelines' staging buffers (0 = the
        48 MiB default); returns the previous size."""
        prev = ctypes.c_size_t()
        check(lib().wmd_ctx_set_staging_bytes(self.handle, int(nbytes), ctypes.byref(prev)))
        return int(prev.value)

    def profile(self, enable: bool = True):
        """Start (clearing old marks) or stop HIP-event timing of the engine's kernels."""
        check(lib().wmd_ctx_profile(self.handle, int(bool(enable))))

    def profile_read(self, name: str):
        """(total ms, launches) of kernel family `name` since profiling started."""
        ms = ctypes.c_double()
        n = ctypes.c_int()
        check(lib().wmd_ctx_profile_read(self.handle, name.encode(), ctypes.byref(ms),
                                         ctypes.byref(n)))
        return float(ms.value), int(n.value)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib is not None:
            _lib.wmd_ctx_destroy(h)
            self.handle = None


_contexts = {}


def context(device: int = 0) -> Context:
    """Process-wide cached context per device (code objects load once)."""
    with _lock:
        ctx = _contexts.get(device)
    if ctx is None:
        ctx = Context(device)
        with _lock:
            _contexts.setdefault(device, ctx)
            ctx = _contexts[device]
    return ctx
