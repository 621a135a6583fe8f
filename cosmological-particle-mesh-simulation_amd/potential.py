"""FFT Poisson solve — the reference's potential.py:7-29 on the GPU.

Same functions and dtypes as the reference: ``density_k`` returns the complex128 forward DFT,
``potential_k`` multiplies by -3*OMEGA_M0/8/a * fgrid, ``potential_real`` returns the float32
real part of the normalised inverse DFT, and ``potential`` chains the three (fused on the
device, complex128 arithmetic).  The transforms are this package's own HIP FFT.
"""
import numpy as np

import configure_me
from fourier_utils import GreenGrid
from pmgpu import ops
from pmgpu.device import DeviceArray
from pmgpu.sim import SimDensity


def _grid(fgrid):
    return fgrid.device() if isinstance(fgrid, GreenGrid) else fgrid


def _dens(density):
    if isinstance(density, SimDensity):
        return np.asarray(density)
    return density


def potential(density, fgrid, a):
    return ops.potential(_dens(density), _grid(fgrid), a, configure_me.OMEGA_M0)


def potential_k(density_k, fgrid, a):
    return ops.potential_k(density_k, _grid(fgrid), a, configure_me.OMEGA_M0)


def density_k(density):
    return ops.density_k(_dens(density))


def potential_real(potential_k):
    return ops.potential_real(potential_k)


__all__ = ["potential", "potential_k", "density_k", "potential_real", "DeviceArray"]
