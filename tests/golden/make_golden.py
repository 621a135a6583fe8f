"""Writes the golden fixtures of this directory from the float64 oracle (run from the repo root:
``python tests/golden/make_golden.py``).  The fixtures freeze small inputs and the oracle's
outputs so that a later change to either side shows up as a diff.

pm_step_n16.npz  one PM step (N_CELLS = 16, N_PARTS = 8, seeded lattice + jitter): positions,
                 velocities, scalars in; density, potential, new positions / velocities and the
                 P(k) of the density out.
white_noise.npz  the first 1027 deviates of the Philox4x32-10 + Box-Muller field for two seeds.
"""
from pathlib import Path
import sys

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent / "oracle"))
import pm_oracle as O  # noqa: E402


def step_fixture():
    n, npart = 16, 8
    rng = np.random.default_rng(2024)
    q = O.lattice_q(npart, n).astype(np.float64)
    pos = np.mod(q + rng.normal(0.0, 0.6, q.shape), n).astype(np.float32)
    pos[pos >= n] = 0.0
    vel = rng.normal(0.0, 0.3, q.shape).astype(np.float32)
    p = O.OracleParams(n_cells=n, n_parts=npart, omega_m0=0.3, omega_lambda0=0.7, h0=0.7)
    a, da = 0.05, 0.01
    rho = O.density(pos.astype(np.float64), p.mass, n)
    x1, v1, phi = O.advance_time(rho, pos.astype(np.float64), vel.astype(np.float64), O.fourier_grid(n),
                                 a, da, p, return_phi=True)
    k, pk, cnt = O.power_spectrum(rho, 100.0)
    np.savez_compressed(HERE / "pm_step_n16.npz", positions=pos, velocities=vel,
                        scalars=np.array([a, da, p.mass, 0.3, 0.7, 0.0, 0.7, 100.0]),
                        rho=rho, phi=phi, positions_out=x1, velocities_out=v1, k=k, pk=pk, counts=cnt)


def noise_fixture():
    np.savez_compressed(HERE / "white_noise.npz", seed42=O.white_noise(1027, 42),
                        seed_big=O.white_noise(1027, 2 ** 40 + 3))


if __name__ == "__main__":
    step_fixture()
    noise_fixture()
    print("fixtures written to", HERE)
