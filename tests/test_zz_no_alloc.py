"""pm_potential_static (the entry point that allocates nothing) against pm_potential.

Both solve potential.py:7-10 for a float32 density and Green's function grid with complex128
FFT arithmetic; the static variant keeps its workspace and FFT tables in a fixed device array
(tables computed on the device, which may differ from the host-computed ones by an ulp).  Bar:
the two potentials agree to 1e-12 norm-wise, on two calls in a row, the second one smaller,
and sizes above the fixed scratch are refused with an error.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import rel, uniform_positions

pytestmark = pytest.mark.gpu


def _potential(name, rho, fgrid, n, a, om):
    from pmgpu import _lib
    from pmgpu.device import DeviceArray
    d_rho = DeviceArray.from_numpy(np.ascontiguousarray(rho, np.float32))
    d_g = DeviceArray.from_numpy(np.ascontiguousarray(fgrid, np.float32))
    d_phi = DeviceArray.empty((n, n, n), np.float32)
    _lib.call(name, C.c_void_p(d_rho.ptr), C.c_void_p(d_g.ptr), n, a, om, C.c_void_p(d_phi.ptr))
    return d_phi.numpy()


def _inputs(n, seed):
    from pmgpu import ops
    pos = uniform_positions(n, (n // 2) ** 3, seed=seed)
    rho = ops.density(pos, 8.0, n)
    return np.asarray(rho, np.float32), np.asarray(ops.fourier_grid(n), np.float32)


def test_static_potential_matches_pm_potential_on_consecutive_calls():
    for n, seed, a in ((64, 3, 0.05), (32, 4, 0.3), (128, 5, 0.7)):   # the second call is smaller
        rho, fgrid = _inputs(n, seed)
        ref = _potential("pm_potential", rho, fgrid, n, a, 0.31)
        got = _potential("pm_potential_static", rho, fgrid, n, a, 0.31)
        e = rel(got, ref)
        print(f"N={n}: pm_potential_static vs pm_potential {e:.2e}")
        assert e < 1e-12


def test_static_potential_rejects_grids_above_its_scratch():
    from pmgpu import _lib
    n = 192
    with pytest.raises(RuntimeError, match="exceeds its fixed scratch"):
        _lib.call("pm_potential_static", None, None, n, 0.1, 0.3, None)
