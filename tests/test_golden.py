"""The committed golden fixtures (tests/golden/, written by make_golden.py from the oracle):
the oracle must still reproduce them (CPU), and the HIP path must match them (GPU)."""
from pathlib import Path

import numpy as np
import pytest

import pm_oracle as O
from helpers import periodic_diff, rel

GOLDEN = Path(__file__).resolve().parent / "golden"


def _step():
    return dict(np.load(GOLDEN / "pm_step_n16.npz"))


def test_oracle_reproduces_step_fixture():
    f = _step()
    a, da, mass, om, ol, ok, h0, box = f["scalars"]
    n, npart = 16, 8
    p = O.OracleParams(n_cells=n, n_parts=npart, omega_m0=om, omega_lambda0=ol, omega_k0=ok, h0=h0)
    rho = O.density(f["positions"].astype(np.float64), mass, n)
    np.testing.assert_allclose(rho, f["rho"], rtol=1e-13, atol=1e-13)
    x1, v1, phi = O.advance_time(rho, f["positions"].astype(np.float64), f["velocities"].astype(np.float64),
                                 O.fourier_grid(n), a, da, p, return_phi=True)
    np.testing.assert_allclose(phi, f["phi"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(x1, f["positions_out"], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(v1, f["velocities_out"], rtol=1e-12, atol=1e-14)
    k, pk, cnt = O.power_spectrum(rho, box)
    np.testing.assert_allclose(pk, f["pk"], rtol=1e-12)
    np.testing.assert_array_equal(cnt, f["counts"])


def test_oracle_reproduces_noise_fixture():
    f = np.load(GOLDEN / "white_noise.npz")
    np.testing.assert_array_equal(O.white_noise(1027, 42), f["seed42"])
    np.testing.assert_array_equal(O.white_noise(1027, 2 ** 40 + 3), f["seed_big"])


@pytest.mark.gpu
def test_gpu_step_matches_fixture():
    from pmgpu import ops
    from pmgpu.sim import PMSim
    f = _step()
    a, da, mass, om, ol, ok, h0, box = f["scalars"]
    n, npart = 16, 8
    pos, vel = f["positions"], f["velocities"]
    # reference-layout entry points
    rho = ops.density(pos, mass, n)
    assert rel(rho, f["rho"]) < 1e-6
    x, v = pos.copy(), vel.copy()
    ops.advance_time(f["rho"].astype(np.float32), x, v, ops.fourier_grid(n), a, da, (h0, ol, ok), om)
    assert np.abs(periodic_diff(x, f["positions_out"], n)).max() < 1e-5 * n
    assert rel(v - vel, f["velocities_out"] - vel) < 1e-5
    # the fused resident step, FP32 and FP64
    for prec, tol in ((32, 1e-5), (64, 1e-10)):
        sim = PMSim(n, npart, prec, omega_m0=om, omega_lambda0=ol, omega_k0=ok, h0=h0)
        sim.set_state(pos, vel)
        sim.step(a, da)
        assert rel(sim.potential(np.float64), f["phi"]) < tol
        x1, v1 = sim.get_state(np.float64)
        assert rel(v1 - vel, f["velocities_out"] - vel) < tol
        sim.close()
    sim = PMSim(n, npart, 64, omega_m0=om, omega_lambda0=ol, omega_k0=ok, h0=h0)
    sim.set_state(pos, vel)
    k, pk, cnt = sim.power_spectrum(box)
    np.testing.assert_allclose(pk, f["pk"], rtol=1e-10)
    np.testing.assert_array_equal(cnt, f["counts"])
    sim.close()


@pytest.mark.gpu
def test_gpu_noise_matches_fixture():
    from pmgpu import _lib
    from pmgpu.device import DeviceArray
    f = np.load(GOLDEN / "white_noise.npz")
    for key, seed in (("seed42", 42), ("seed_big", 2 ** 40 + 3)):
        d = DeviceArray((1027,), np.float64)
        _lib.call("pm_white_noise", d.ptr, 1027, seed)
        np.testing.assert_allclose(d.numpy(), f[key], rtol=1e-13, atol=1e-13)
