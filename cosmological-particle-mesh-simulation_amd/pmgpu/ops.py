"""Reference-layout PM operations on the GPU (one C-ABI call each).

Every function accepts NumPy arrays or ``DeviceArray`` objects.  NumPy inputs are copied to
the device and the result is copied back (the function then returns NumPy arrays, and
in-place updates land in the caller's arrays, as in the reference); DeviceArray inputs stay
on the device and results are DeviceArrays.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .device import DeviceArray, as_device


def _out(result: DeviceArray, host: bool):
    return result.numpy() if host else result


def _check_pos(positions):
    shape = tuple(positions.shape)
    if len(shape) != 2 or shape[0] != 3:
        raise ValueError(f"positions must have shape (3, N_p), got {shape}")
    return shape[1]


def density(positions, mass: float, n_cells: int):
    """density.py:7-48 — CIC deposit of float32 positions (3, N_p) into a float32 N^3 grid."""
    n_p = _check_pos(positions)
    host = not isinstance(positions, DeviceArray)
    d_pos, _ = as_device(positions, np.float32)
    rho = DeviceArray((n_cells,) * 3, np.float32)
    _lib.call("pm_density", d_pos.ptr, n_p, n_cells, float(mass), rho.ptr)
    return _out(rho, host)


def fourier_grid(n_cells: int, host: bool = True):
    """fourier_utils.py:5-16 — float32 N^3 grid of 1/sum sin^2(k/2), G(0) = 0."""
    g = DeviceArray((n_cells,) * 3, np.float32)
    _lib.call("pm_fourier_grid", n_cells, g.ptr)
    return _out(g, host)


def density_k(rho):
    """potential.py:17-21 — forward c2c DFT of a float32 grid, complex128 result."""
    n = rho.shape[0]
    host = not isinstance(rho, DeviceArray)
    d_rho, _ = as_device(rho, np.float32)
    out = DeviceArray((n, n, n), np.complex128)
    _lib.call("pm_density_k", d_rho.ptr, n, out.ptr)
    return _out(out, host)


def potential_k(rho_k, fgrid, a: float, omega_m0: float):
    """potential.py:12-15 — -3*Omega_m/8/a * fgrid * density_k."""
    n = rho_k.shape[0]
    host = not isinstance(rho_k, DeviceArray)
    d_k, _ = as_device(rho_k, np.complex128)
    d_g, _ = as_device(fgrid, np.float32)
    out = DeviceArray((n, n, n), np.complex128)
    _lib.call("pm_potential_k", d_k.ptr, d_g.ptr, n, float(a), float(omega_m0), out.ptr)
    return _out(out, host)


def potential_real(phi_k):
    """potential.py:23-29 — Re(inverse DFT)/N^3 as float32."""
    n = phi_k.shape[0]
    host = not isinstance(phi_k, DeviceArray)
    d_k, _ = as_device(phi_k, np.complex128)
    out = DeviceArray((n, n, n), np.float32)
    _lib.call("pm_potential_real", d_k.ptr, n, out.ptr)
    return _out(out, host)


def potential(rho, fgrid, a: float, omega_m0: float):
    """potential.py:7-10 — the three steps above fused on the device (complex128)."""
    n = rho.shape[0]
    host = not isinstance(rho, DeviceArray)
    d_rho, _ = as_device(rho, np.float32)
    d_g, _ = as_device(fgrid, np.float32)
    out = DeviceArray((n, n, n), np.float32)
    _lib.call("pm_potential", d_rho.ptr, d_g.ptr, n, float(a), float(omega_m0), out.ptr)
    return _out(out, host)


def cell_centers(positions, n_cells: int):
    """integrate.py:20 — floor(positions).astype(int64) % N_CELLS."""
    n_p = _check_pos(positions)
    host = not isinstance(positions, DeviceArray)
    d_pos, _ = as_device(positions, np.float32)
    out = DeviceArray((3, n_p), np.int64)
    _lib.call("pm_cell_centers", d_pos.ptr, n_p, n_cells, out.ptr)
    return _out(out, host)


def weights(cells, positions):
    """integrate.py:27-52 — the eight CIC weights, float32 (8, N_p)."""
    n_p = _check_pos(positions)
    host = not isinstance(positions, DeviceArray)
    d_c, _ = as_device(cells, np.int64)
    d_pos, _ = as_device(positions, np.float32)
    out = DeviceArray((8, n_p), np.float32)
    _lib.call("pm_weights", d_c.ptr, d_pos.ptr, n_p, out.ptr)
    return _out(out, host)


def _inplace(arr, dtype):
    """(device array, write-back callable) for an array updated in place by the device."""
    if isinstance(arr, DeviceArray):
        return arr, (lambda: None)
    if not (isinstance(arr, np.ndarray) and arr.dtype == np.dtype(dtype) and arr.flags.c_contiguous):
        raise TypeError(f"in-place argument must be a C-contiguous {np.dtype(dtype)} array")
    d = DeviceArray.from_numpy(arr)

    def back():
        arr[...] = d.numpy()
    return d, back


def sweep_one_direction(cells, positions_d, velocities_d, t, phi, da, f_a1, a_val, direction,
                        n_cells):
    """integrate.py:54-97 — kick and drift of one axis; updates the 1-D arrays in place."""
    n_p = int(positions_d.shape[0])
    d_c, _ = as_device(cells, np.int64)
    d_t, _ = as_device(t, np.float32)
    d_phi, _ = as_device(phi, np.float32)
    d_x, back_x = _inplace(positions_d, np.float32)
    d_v, back_v = _inplace(velocities_d, np.float32)
    _lib.call("pm_sweep_one_direction", d_c.ptr, d_x.ptr, d_v.ptr, d_t.ptr, d_phi.ptr, n_p,
              n_cells, float(da), float(f_a1), float(a_val), int(direction))
    back_x()
    back_v()
    return positions_d, velocities_d


def integrate(positions, velocities, a_val, f_a1, da, phi, n_cells):
    """integrate.py:15-25 — gather, kick and drift of all axes; in place."""
    n_p = _check_pos(positions)
    d_phi, _ = as_device(phi, np.float32)
    d_x, back_x = _inplace(positions, np.float32)
    d_v, back_v = _inplace(velocities, np.float32)
    _lib.call("pm_integrate", d_x.ptr, d_v.ptr, n_p, n_cells, float(a_val), float(f_a1),
              float(da), d_phi.ptr)
    back_x()
    back_v()
    return positions, velocities


def advance_time(rho, positions, velocities, fgrid, a, da, cosmology_f, omega_m0):
    """integrate.py:9-13 — potential, f(a+da, cosmology_f), integrate; in place."""
    n_p = _check_pos(positions)
    n = rho.shape[0]
    d_rho, _ = as_device(rho, np.float32)
    d_g, _ = as_device(fgrid, np.float32)
    d_x, back_x = _inplace(positions, np.float32)
    d_v, back_v = _inplace(velocities, np.float32)
    c0, c1, c2 = (float(c) for c in cosmology_f)
    _lib.call("pm_advance_time", d_rho.ptr, d_x.ptr, d_v.ptr, d_g.ptr, n_p, n, float(a),
              float(da), c0, c1, c2, float(omega_m0))
    back_x()
    back_v()
    return positions, velocities


def f(a, c0, c1, c2) -> float:
    return float(_lib.load().pm_f(float(a), float(c0), float(c1), float(c2)))
