"""CIC mass deposit — the reference's density.py:7-48 on the GPU.

``density(positions, mass)`` keeps the reference's signature and result: a new float32 grid of
shape (N_CELLS, N_CELLS, N_CELLS) indexed [z, y, x], positions[0] on grid axis 2.  Positions
may be a NumPy array (copied to the device and back), a ``DeviceArray`` (stays on the device)
or the ``SimPositions`` view of a device-resident simulation (the deposit is fused into that
simulation's step; the returned ``SimDensity`` copies to the host only when read).
Differences from the reference: the deposit is race-free (density.py:17,37-47 races under
prange), and a float32 position equal to N_CELLS is wrapped to 0 (its CIC weights would
otherwise be +-(N-1)).
"""
import configure_me
from pmgpu import ops
from pmgpu.sim import SimDensity, SimPositions


def density(positions, mass):
    if isinstance(positions, SimPositions):
        sim = positions.sim
        sim.set_mass(mass)
        return SimDensity(sim, sim.version)
    return ops.density(positions, mass, int(configure_me.N_CELLS))
