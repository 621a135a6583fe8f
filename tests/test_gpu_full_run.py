"""A whole run of the reference's small configuration against the CPU reference port.

Configuration: 64^3 particles on a 128^3 grid, box 100 Mpc/h, Omega_m 0.3, Omega_L 0.7,
h 0.7, EH no-wiggle P(k) with n_s 0.96 and sigma8 0.8, seed 1, Zel'dovich at a = 0.02, STEPS = 100
(99 executed, pmesh.py:56-63).  The CPU side is the C/OpenMP port of density.py / integrate.py
with complex128 FFTs (the reference's dtypes).  Bar: the matter power spectra of the final
densities agree within 0.5 % for k < k_Nyquist / 2, for the FP32 and the FP64 GPU runs.
"""
import numpy as np
import pytest

import pm_oracle as O
from helpers import make_ics

pytestmark = pytest.mark.gpu


def test_small_config_full_run_power_spectrum_matches_cpu_port():
    import pm_port
    from pmgpu.sim import PMSim
    n, npart, box = 128, 64, 100.0
    a0, a_end, steps_cfg = 0.02, 1.0, 100
    pos, vel = make_ics(n, npart, seed=1, box=box)
    steps = O.step_count(a0, a_end, steps_cfg)
    assert steps == 99
    da = (a_end - a0) / steps_cfg
    mass = (n / npart) ** 3
    workers = max(1, min(pm_port.cores(), 16))
    pm_port.load().port_set_threads(workers)
    x, v = pos.copy(), vel.copy()
    fgrid = O.fourier_grid(n, ref_dtypes=True)
    a = a0
    for _ in range(steps):
        pm_port.step(x, v, fgrid, mass, a, da, 0.3, 0.7, 0.0, 0.7, workers)
        a += da
    _, p_cpu, cnt = O.power_spectrum(O.density(O.wrap_positions(x, n).astype(np.float64), mass, n), box)
    sel = np.arange(1, len(p_cpu) + 1) <= n // 4          # k < k_Nyquist / 2
    for prec in (32, 64):
        sim = PMSim(n, npart, prec, omega_m0=0.3, omega_lambda0=0.7, omega_k0=0.0, h0=0.7)
        sim.set_state(pos, vel)
        a_gpu = sim.run(a0, da, steps)
        assert abs(a_gpu - a) < 1e-12
        _, p_gpu, cnt_gpu = sim.power_spectrum(box)
        np.testing.assert_array_equal(cnt_gpu, cnt)
        err = np.max(np.abs(p_gpu[sel] / p_cpu[sel] - 1))
        print(f"FP{prec} GPU vs CPU port after {steps} steps: max P(k) rel err {err:.2e} "
              f"(k < k_Ny/2); P ratio at the largest scale {p_gpu[0] / p_cpu[0]:.6f}")
        assert err < 5e-3
        sim.close()
