"""ctypes binding of ``libfastesp_hip.so`` (the C ABI of ``include/fastesp_hip.h``).

The library is loaded from the package directory (built in-tree by
``fastesp_amd.build``).  There is no fallback: if the library is missing or
no HIP device is usable, every entry point raises ``HipBackendError``.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

import numpy as np

# the in-tree build; FASTESP_HIP_LIB names another build of the same ABI
LIB_PATH = Path(os.environ.get("FASTESP_HIP_LIB") or
                Path(__file__).resolve().parent / "libfastesp_hip.so")

FE_OK = 0
FE_PTR_HOST = 0
FE_PTR_DEVICE = 1
FE_PREC_FP64 = 0
FE_PREC_MIXED = 1
FE_PART_NUCLEAR = 1
FE_PART_ELECTRON = 2
FE_PART_ESP = 3
FE_LMAX = 6
FE_BOYS_MMAX = 8
FE_NOALLOC_MAX_POINTS = 65536
FE_BATCH_FIELD = 1
FE_BATCH_DLDP = 2
FE_BATCH_REBUILD = 4
FE_BOUT_V, FE_BOUT_FIELD, FE_BOUT_DLDP, FE_BOUT_LOSS, FE_BOUT_GV = range(5)
FE_BT_ITEMS_PER_WG = 1
FE_BT_DG_ITEMS_PER_WG = 2
FE_BT_CLASS_MASK = 3
FE_BT_PPL0 = 10
FE_BT_MIXED0 = 20

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_vp = C.c_void_p

# name -> argtypes (all return int unless listed in _RESTYPE)
SIGNATURES = {
    "fe_version": [],
    "fe_last_error": [],
    "fe_init": [C.c_int, C.POINTER(_vp)],
    "fe_finalize": [_vp],
    "fe_ctx_trim": [_vp],
    "fe_table_from_pairs": [_vp, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _ip, _ip,
                            _ip, _ip, _ip, _ip, _dp, C.c_int, _dp, _dp, C.POINTER(_vp)],
    "fe_table_free": [_vp],
    "fe_table_info": [_vp, C.POINTER(C.c_int64), C.c_int],
    "fe_esp_forward": [_vp, _vp, C.c_int64, C.c_int, C.c_int, C.c_int, _vp, _vp],
    "fe_esp_forward_noalloc": [_vp, _vp, C.c_int64, C.c_int, C.c_int, _vp, _vp],
    "fe_boys_eval": [_vp, _vp, C.c_int64, C.c_int, C.c_int, _vp, _vp],
    "fe_set_tuning": [_vp, C.c_int, C.c_int64],
    "fe_table_build": [_vp, C.c_int, C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp, _dp,
                       _vp, C.c_int, _vp, C.c_int, _dp, _dp, C.c_double, C.c_int, C.c_int, _vp,
                       C.POINTER(_vp)],
    "fe_table_stats": [_vp, _dp, C.c_int],
    "fe_table_rebuild": [_vp, _vp, C.c_int, _vp, C.c_double, C.c_int, C.c_int, _vp],
    "fe_esp_forward_grad": [_vp, _vp, C.c_int64, C.c_int, C.c_int, _vp, _vp, _vp],
    "fe_esp_backward": [_vp, _vp, C.c_int64, _vp, C.c_int, C.c_int, _vp, _vp, _vp],
    "fe_table_export": [_vp, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _ip,
                        _dp, _dp, _dp],
    "fe_batch_create": [_vp, C.c_int, C.POINTER(_vp), _vp, C.POINTER(C.c_int64), C.c_int, _vp,
                        C.POINTER(_vp)],
    "fe_batch_free": [_vp],
    "fe_batch_set_points": [_vp, _vp, C.c_int, _vp],
    "fe_batch_set_loss": [_vp, _vp, _vp, C.c_int, _vp],
    "fe_batch_set_tuning": [_vp, C.c_int, C.c_int64],
    "fe_batch_set_density": [_vp, _vp, C.c_double, C.c_int, C.c_int, _vp],
    "fe_batch_forward": [_vp, C.c_int, C.c_int, _vp],
    "fe_batch_loss": [_vp, _vp],
    "fe_batch_backward": [_vp, _vp, C.c_int, _vp],
    "fe_batch_step": [_vp, C.c_int, C.c_int, _vp],
    "fe_batch_capture": [_vp, C.c_int, C.c_int],
    "fe_batch_launch": [_vp, _vp],
    "fe_batch_read": [_vp, C.c_int, C.c_int, _vp, _vp],
    "fe_batch_info": [_vp, C.POINTER(C.c_int64), C.c_int],
    "fe_profile": [_vp, C.c_int, _dp, C.POINTER(C.c_int64), C.c_int],
    "fe_eval_work": [_vp, _vp, C.c_int64, C.c_int, C.c_int, _vp, C.POINTER(C.c_int64)],
    "fe_density_create": [_vp, C.c_int, C.c_int, _ip, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _vp,
                          C.c_int, _vp, C.POINTER(_vp)],
    "fe_density_eval": [_vp, _vp, C.c_int64, C.c_int, _vp, _vp],
    "fe_density_info": [_vp, C.POINTER(C.c_int64), C.c_int],
    "fe_density_set_tuning": [_vp, C.c_int, C.c_int64],
    "fe_density_free": [_vp],
    "fe_vdw_candidates": [_vp, C.c_int, _dp, _dp, C.POINTER(C.c_int64), _dp, _ip],
}
_RESTYPE = {"fe_last_error": C.c_char_p}


class HipBackendError(RuntimeError):
    """The HIP backend is unavailable or a library call failed."""


_lib = None
_lock = threading.Lock()


def load():
    """Load the shared library once (raises ``HipBackendError`` if absent)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not LIB_PATH.exists():
            raise HipBackendError(
                f"{LIB_PATH} is not built; run `python -m fastesp_amd.build` first")
        # The PyTorch ROCm wheel ships its own libamdhip64 under the same soname.
        # Loading it first makes it the one HIP runtime of the process; the other
        # order leaves torch bound to a runtime it does not find devices with.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(str(LIB_PATH))
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, C.c_int)
        _lib = lib
        return lib


def check(rc: int, what: str) -> None:
    if rc != FE_OK:
        msg = load().fe_last_error().decode(errors="replace")
        raise HipBackendError(f"{what} failed ({rc}): {msg}")


def ptr(a: np.ndarray, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct))


class Context:
    """One device context (``fe_init``); shared per device within a process."""

    _by_device: dict = {}

    def __init__(self, device: int = 0):
        lib = load()
        h = _vp()
        check(lib.fe_init(device, C.byref(h)), f"fe_init(device={device})")
        self.handle = h
        self.device = device
        self.lib = lib

    @classmethod
    def get(cls, device: int | None = None) -> "Context":
        if device is None:
            device = int(os.environ.get("FASTESP_HIP_DEVICE", "0"))
        with _lock:
            ctx = cls._by_device.get(device)
        if ctx is None:
            ctx = cls(device)
            with _lock:
                cls._by_device[device] = ctx
        return ctx

    def set_tuning(self, key: int, value: int) -> None:
        check(self.lib.fe_set_tuning(self.handle, key, int(value)), "fe_set_tuning")

    def trim(self) -> None:
        """Release the context's workspace and cached basis set-up (``fe_ctx_trim``)."""
        check(self.lib.fe_ctx_trim(self.handle), "fe_ctx_trim")

    def profile_start(self) -> None:
        check(self.lib.fe_profile(self.handle, 1, None, None, 0), "fe_profile")

    def profile_stop(self) -> dict:
        """{phase id: (total ms, launches)} recorded since profile_start()."""
        ms = (C.c_double * 32)()
        cnt = (C.c_int64 * 32)()
        check(self.lib.fe_profile(self.handle, 0, ms, cnt, 32), "fe_profile")
        return {i: (ms[i], int(cnt[i])) for i in range(32) if cnt[i]}

    def boys(self, T, M: int) -> np.ndarray:
        T = np.ascontiguousarray(T, np.float64).ravel()
        F = np.empty((len(T), M + 1))
        check(self.lib.fe_boys_eval(self.handle, T.ctypes.data, len(T), M, FE_PTR_HOST, None,
                                    F.ctypes.data), "fe_boys_eval")
        return F
