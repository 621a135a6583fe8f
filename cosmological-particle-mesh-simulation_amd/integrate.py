"""CIC gather + kick/drift — the reference's integrate.py:9-97 on the GPU.

``advance_time(density, positions, velocities, fgrid, a, da)`` solves for the potential,
takes fa1 = f(a+da, [H0, OMEGA_LAMBDA0, OMEGA_K0]) exactly as integrate.py:12 does (set
configure_me.FIX_F_COSMOLOGY for [OMEGA_M0, ...]), gathers the CIC-interpolated central
difference of the potential at the pre-step positions for all three axes, kicks the
velocities and drifts the positions modulo N_CELLS.  ``positions`` and ``velocities`` are
updated in place and returned, as in the reference.

With the ``SimPositions``/``SimVelocities`` views of a device-resident simulation the whole
step runs as one fused device sequence (R2C FFT, k-space pass, C2R FFT, gather+kick+drift,
next deposit) with no host transfer.
"""
import numpy as np

import configure_me
from cosmology import f, step_coefficient_cosmology
from fourier_utils import GreenGrid
from pmgpu import ops
from pmgpu.device import DeviceArray
from pmgpu.sim import SimDensity, SimPositions, SimVelocities


def _density_read_after_step() -> bool:
    """Whether the configured driver writes or draws a step's density after the step."""
    c = configure_me
    return bool((c.SAVE_DATA and c.SAVE_DENSITY) or c.PLOT_STEPS or c.PLOT_PROJECTIONS)


def advance_time(density, positions, velocities, fgrid, a, da):
    if isinstance(positions, SimPositions):
        sim = positions.sim
        if not (isinstance(velocities, SimVelocities) and velocities.sim is sim):
            raise TypeError("positions and velocities must be views of the same simulation")
        if not (isinstance(density, SimDensity) and density.sim is sim
                and density.version == sim.version):
            sim.set_density(np.asarray(density))
        elif _density_read_after_step():
            density.keep()   # the step overwrites the density the caller may still save or plot
        if isinstance(fgrid, GreenGrid):
            sim.set_green(None)
        else:
            sim.set_green(DeviceArray.from_numpy(np.asarray(fgrid, dtype=np.float32)))
        c = configure_me
        sim.set_cosmology(c.OMEGA_M0, c.OMEGA_LAMBDA0, c.OMEGA_K0, c.H0)
        sim.step(a, da)
        return positions, velocities
    if isinstance(density, SimDensity):
        density = np.asarray(density)
    g = fgrid.device() if isinstance(fgrid, GreenGrid) else fgrid
    return ops.advance_time(density, positions, velocities, g, a, da,
                            step_coefficient_cosmology(), configure_me.OMEGA_M0)


def integrate(positions, velocities, a_val, f_a1, da, potentials):
    return ops.integrate(positions, velocities, a_val, f_a1, da, potentials,
                         int(configure_me.N_CELLS))


def weights(cell_centers, positions):
    return ops.weights(cell_centers, positions)


def cell_centers(positions):
    """integrate.py:20 as a function of its own."""
    return ops.cell_centers(positions, int(configure_me.N_CELLS))


def sweep_one_direction(cell_centers, positions, velocities, t, potentials, da, f_a1, a_val,
                        direction):
    return ops.sweep_one_direction(cell_centers, positions, velocities, t, potentials, da, f_a1,
                                   a_val, direction, int(configure_me.N_CELLS))


__all__ = ["advance_time", "integrate", "weights", "sweep_one_direction", "cell_centers", "f"]
