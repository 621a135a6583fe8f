"""ctypes binding of libpm.so (the C ABI declared in include/pm.h).

The library is loaded on first use.  A missing library or a missing HIP device raises
``RuntimeError``; there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libpm.so"

PM_OK = 0
PM_FLAG_CIC_DECONVOLVE = 1
PM_FLAG_FIX_F_COSMOLOGY = 2
PM_FLAG_NO_GRAPH = 4
PM_FLAG_FUSED_DEPOSIT = 8
PM_FLAG_DETERMINISTIC = 16
PM_FLAG_FORWARD_ORDER = 32
PM_E_CORRUPT = 6

# N stages timed by pm_sim_get_timing, in order
STAGES = ("deposit", "r2c_x", "c2c_y_fwd", "z_green", "c2c_y_inv", "c2r_x", "gather_kick_drift",
          "bin", "deposit_r2c")

_vp = C.c_void_p
_i = C.c_int
_i64 = C.c_int64
_d = C.c_double
_sz = C.c_size_t

# name -> (restype, argtypes)
SIGNATURES = {
    "pm_init": (_i, [_i]),
    "pm_last_error": (C.c_char_p, []),
    "pm_version": (_i, []),
    "pm_supported_grid": (_i, [_i]),
    "pm_malloc": (_i, [C.POINTER(_vp), _sz]),
    "pm_free": (_i, [_vp]),
    "pm_memcpy_htod": (_i, [_vp, _vp, _sz]),
    "pm_memcpy_dtoh": (_i, [_vp, _vp, _sz]),
    "pm_memcpy_dtod": (_i, [_vp, _vp, _sz]),
    "pm_memset": (_i, [_vp, _i, _sz]),
    "pm_synchronize": (_i, []),
    "pm_mem_info": (_i, [C.POINTER(_sz), C.POINTER(_sz)]),
    "pm_density": (_i, [_vp, _i64, _i, _d, _vp]),
    "pm_fourier_grid": (_i, [_i, _vp]),
    "pm_density_k": (_i, [_vp, _i, _vp]),
    "pm_potential_k": (_i, [_vp, _vp, _i, _d, _d, _vp]),
    "pm_potential_real": (_i, [_vp, _i, _vp]),
    "pm_potential": (_i, [_vp, _vp, _i, _d, _d, _vp]),
    "pm_potential_static": (_i, [_vp, _vp, _i, _d, _d, _vp]),
    "pm_cell_centers": (_i, [_vp, _i64, _i, _vp]),
    "pm_weights": (_i, [_vp, _vp, _i64, _vp]),
    "pm_sweep_one_direction": (_i, [_vp, _vp, _vp, _vp, _vp, _i64, _i, _d, _d, _d, _i]),
    "pm_integrate": (_i, [_vp, _vp, _i64, _i, _d, _d, _d, _vp]),
    "pm_advance_time": (_i, [_vp, _vp, _vp, _vp, _i64, _i, _d, _d, _d, _d, _d, _d]),
    "pm_f": (_d, [_d, _d, _d, _d]),
    "pm_sim_create": (_i, [_i, _i, _i, C.c_uint, C.POINTER(_vp)]),
    "pm_sim_destroy": (_i, [_vp]),
    "pm_sim_set_cosmology": (_i, [_vp, _d, _d, _d, _d]),
    "pm_sim_set_mass": (_i, [_vp, _d]),
    "pm_sim_set_green": (_i, [_vp, _vp]),
    "pm_sim_set_state": (_i, [_vp, _vp, _vp, _i, _i]),
    "pm_sim_get_state": (_i, [_vp, _vp, _vp, _i, _i]),
    "pm_sim_set_density": (_i, [_vp, _vp, _i, _i]),
    "pm_sim_deposit": (_i, [_vp]),
    "pm_sim_get_density": (_i, [_vp, _vp, _i, _i]),
    "pm_sim_get_potential": (_i, [_vp, _vp, _i, _i]),
    "pm_sim_step": (_i, [_vp, _d, _d]),
    "pm_sim_run": (_i, [_vp, _d, _d, _i, C.POINTER(_d)]),
    "pm_sim_set_timing": (_i, [_vp, _i]),
    "pm_sim_get_timing": (_i, [_vp, C.POINTER(_d), C.POINTER(_i64), _i]),
    "pm_sim_sync": (_i, [_vp]),
    "pm_sim_footprint": (_i, [_i, _i, _i, C.POINTER(_sz)]),
    "pm_sim_set_tiles": (_i, [_vp, _i, _i, _i]),
    "pm_sim_get_tiles": (_i, [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)]),
    "pm_sim_set_pencils": (_i, [_vp, _i, _i, _i]),
    "pm_sim_get_pencils": (_i, [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)]),
    "pm_sim_autotune": (_i, [_vp, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_d)]),
    "pm_white_noise": (_i, [_vp, _i64, C.c_uint64]),
    "pm_sim_zeldovich": (_i, [_vp, C.POINTER(_d), _i64, _d, _d, C.c_uint64, _vp, _i]),
    "pm_sim_power_spectrum": (_i, [_vp, _d, _i, C.POINTER(_d), C.POINTER(_d), C.POINTER(_d)]),
    "pm_ensemble_create": (_i, [C.POINTER(_vp)]),
    "pm_ensemble_destroy": (_i, [_vp]),
    "pm_ensemble_add": (_i, [_vp, _vp]),
    "pm_ensemble_remove": (_i, [_vp, _vp]),
    "pm_ensemble_size": (_i, [_vp, C.POINTER(_i)]),
    "pm_ensemble_run": (_i, [_vp, C.POINTER(_d), C.POINTER(_d), _i]),
    "pm_ensemble_set_timing": (_i, [_vp, _i]),
    "pm_ensemble_get_timing": (_i, [_vp, C.POINTER(_d), C.POINTER(_i64), _i]),
    "pm_ensemble_sync": (_i, [_vp]),
}

_lock = threading.Lock()
_lib = None


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libpm.so and attach signatures (no device is touched)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = Path(path) if path else LIB_PATH
        if not p.exists():
            raise RuntimeError(
                f"{p} is missing: build the HIP library first (python -m pmgpu.build); "
                "this package has no CPU fallback")
        lib = C.CDLL(str(p))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def exported_symbols() -> list[str]:
    return list(SIGNATURES)


def check(rc: int, what: str = "") -> None:
    if rc != PM_OK:
        msg = load().pm_last_error()
        msg = msg.decode() if msg else ""
        raise RuntimeError(f"libpm{(' ' + what) if what else ''} failed (code {rc}): {msg}")


def call(name: str, *args) -> None:
    """Call a status-returning entry point and raise on failure."""
    lib = load()
    check(getattr(lib, name)(*args), name)


_init_done = False


def init(device: int = -1) -> None:
    global _init_done
    if not _init_done:
        call("pm_init", device)
        _init_done = True
