"""CPU port of the reference's PM step — TEST INFRASTRUCTURE ONLY (the bench's CPU baseline).

One step follows pmesh.py:60-61 with the reference's data types: density() in C/OpenMP
(pm_port.c), density_k/potential_real as complex128 c2c transforms with scipy.fft standing in
for pyFFTW/FFTW3 (potential.py:17-29), potential_k (potential.py:12-15) and integrate() in
C/OpenMP (integrate.py:15-97), with fa1 = f(a+da, [H0, OL, OK]) (integrate.py:12).
"""
from __future__ import annotations

import ctypes as C
import os
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libpm_port.so"

_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run make -C oracle")
        lib = C.CDLL(str(LIB))
        vp, i64, i, d = C.c_void_p, C.c_int64, C.c_int, C.c_double
        lib.port_set_threads.argtypes = [i]
        lib.port_max_threads.restype = i
        lib.port_density.argtypes = [vp, i64, i, d, vp]
        lib.port_cell_centers.argtypes = [vp, i64, i, vp]
        lib.port_weights.argtypes = [vp, vp, i64, i, vp]
        lib.port_sweep.argtypes = [vp, vp, vp, vp, vp, i64, i, d, d, d, i]
        lib.port_integrate.argtypes = [vp, vp, i64, i, d, d, d, vp]
        lib.port_potential_k.argtypes = [vp, vp, i64, d]
        _lib = lib
    return _lib


def cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def density(pos, mass, n):
    lib = load()
    pos = np.ascontiguousarray(pos, np.float32)
    grid = np.empty((n, n, n), np.float32)
    lib.port_density(pos.ctypes.data, pos.shape[1], n, float(mass), grid.ctypes.data)
    return grid


def potential(rho, fgrid, a, omega_m0, workers):
    import scipy.fft as sf
    lib = load()
    n = rho.shape[0]
    spec = sf.fftn(rho.astype(np.complex128), workers=workers, overwrite_x=True)   # potential.py:19-21
    lib.port_potential_k(spec.ctypes.data, np.ascontiguousarray(fgrid, np.float32).ctypes.data,
                         spec.size, -3 * omega_m0 / 8 / a)                           # potential.py:15
    out = sf.ifftn(spec, workers=workers, overwrite_x=True)                          # potential.py:26-28
    return out.real.astype(np.float32)                                               # potential.py:29


def integrate(pos, vel, a, fa1, da, phi):
    lib = load()
    n = phi.shape[0]
    lib.port_integrate(pos.ctypes.data, vel.ctypes.data, pos.shape[1], n, float(a), float(fa1),
                       float(da), np.ascontiguousarray(phi, np.float32).ctypes.data)


def f(a, c):
    return 1 / ((c[0] + c[2] * a + c[1] * a ** 3) / a) ** 0.5


def step(pos, vel, fgrid, mass, a, da, om, ol, ok, h0, workers, timings=None):
    """One PM step in place (pmesh.py:60-61).  If ``timings`` is a dict, the seconds of each
    stage are added to it: density (density.py), fft_forward, green (potential_k), fft_inverse
    (potential.py:17-29, with the complex128 conversions) and integrate (integrate.py)."""
    import scipy.fft as sf
    lib = load()
    n = fgrid.shape[0]
    t0 = time.perf_counter()
    rho = density(pos, mass, n)
    t1 = time.perf_counter()
    spec = sf.fftn(rho.astype(np.complex128), workers=workers, overwrite_x=True)     # :19-21
    t2 = time.perf_counter()
    lib.port_potential_k(spec.ctypes.data, np.ascontiguousarray(fgrid, np.float32).ctypes.data,
                         spec.size, -3 * om / 8 / a)                                 # :15
    t3 = time.perf_counter()
    phi = sf.ifftn(spec, workers=workers, overwrite_x=True).real.astype(np.float32)  # :26-29
    del spec
    t4 = time.perf_counter()
    integrate(pos, vel, a, f(a + da, [h0, ol, ok]), da, phi)
    t5 = time.perf_counter()
    if timings is not None:
        for k, v in (("density", t1 - t0), ("fft_forward", t2 - t1), ("green", t3 - t2),
                     ("fft_inverse", t4 - t3), ("integrate", t5 - t4)):
            timings[k] = timings.get(k, 0.0) + v
    return rho, phi
