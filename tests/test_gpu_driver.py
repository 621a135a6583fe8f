"""pmesh.py end to end: the reference's loop (pmesh.py:18-77) runs unchanged on the GPU path,
executes the reference's step count, writes snapshots with the reference's datasets, restarts
from them, and gives the same particles as the resident sim stepped directly.

Tolerances: resident loop vs direct sim runs 1e-4 cells / 1e-5 (the same kernels; only the
deposit's sum order can differ); the non-resident loop (reference-layout entry points, complex128
FFT, float32 storage) vs the resident FP32 loop 2e-3 cells / 1e-3 after 9 steps.
"""
import contextlib
import importlib
import io

import numpy as np
import pytest

import pm_oracle as O
from helpers import periodic_diff, rel

pytestmark = pytest.mark.gpu


@pytest.fixture
def config(tmp_path):
    import configure_me as c
    saved = {k: getattr(c, k) for k in dir(c) if k.isupper()}
    c.N_CELLS, c.N_PARTS, c.BOX_SIZE, c.STEPS = 32, 16, 50.0, 10
    c.A_INIT, c.A_END = 0.01, 1.0
    c.N_SAVE_FILES, c.N_PLOTS = 2, 2
    c.SAVE_DATA, c.SAVE_DENSITY, c.PRINT_STATUS = True, True, False
    c.PLOT_STEPS = c.PLOT_PROJECTIONS = c.PLOT_GRF = False
    c.RESTART, c.RESTART_FROM_N = False, 0
    c.RESIDENT, c.PRECISION = True, 32
    c.DATA_DIR = str(tmp_path)
    yield c
    for k, v in saved.items():
        setattr(c, k, v)


def _run():
    import pmesh
    importlib.reload(pmesh)
    with contextlib.redirect_stdout(io.StringIO()):
        return pmesh.simulator()


def _ics(c):
    from gaussian_random_field import gaussian_random_field
    from initial_conditions import zeldovich_numpy
    with contextlib.redirect_stdout(io.StringIO()):
        return zeldovich_numpy(gaussian_random_field())


def test_pmesh_loop_matches_direct_sim_and_writes_snapshots(config):
    from pmgpu.sim import PMSim
    c = config
    pos, vel, a_end, steps = _run()
    assert steps == O.step_count(c.A_INIT, c.A_END, c.STEPS)
    x, v = np.asarray(pos), np.asarray(vel)
    p0, v0 = _ics(c)
    sim = PMSim(c.N_CELLS, c.N_PARTS, 32, omega_m0=c.OMEGA_M0, omega_lambda0=c.OMEGA_LAMBDA0,
                omega_k0=c.OMEGA_K0, h0=c.H0)
    sim.set_state(p0, v0)
    sim.set_mass((c.N_CELLS / c.N_PARTS) ** 3)
    da = (c.A_END - c.A_INIT) / c.STEPS
    a = sim.run(c.A_INIT, da, steps)
    assert abs(a - a_end) < 1e-12
    xs, vs = sim.get_state()
    assert np.abs(periodic_diff(x, xs, c.N_CELLS)).max() < 1e-4
    assert rel(v, vs) < 1e-5
    sim.close()
    # snapshots: initial conditions (0) and the save cadence, read back with from_file
    from save_data import from_file
    for n in (0, 1):
        xp, vp, a_snap = from_file(n)
        assert xp.shape == (3, c.N_PARTS ** 3) and 0 < a_snap <= 1
    x0, v0r, a0 = from_file(0)
    np.testing.assert_allclose(x0, p0, rtol=2e-6, atol=2e-6 * c.N_CELLS)
    np.testing.assert_allclose(v0r, v0, rtol=2e-6, atol=1e-9)
    assert a0 == np.float32(c.A_INIT)
    z = np.load(f"{c.DATA_DIR}/data.1.npz", allow_pickle=False)
    assert set(z.files) == {"x1", "x2", "x3", "vx1", "vx2", "vx3", "a", "density"}
    assert z["density"].shape == (c.N_CELLS,) * 3


def test_snapshot_density_is_the_density_of_its_step(config):
    # pmesh.py:60-66 writes the density that the step just solved for (deposited from the
    # pre-step positions), not the one deposited from the positions the step produced
    from pmgpu.sim import PMSim
    c = config
    _run()
    da = (c.A_END - c.A_INIT) / c.STEPS
    da_save = (c.A_END - c.A_INIT) / c.N_SAVE_FILES
    a, k, k1 = c.A_INIT, 0, None
    while a < c.A_END - da and k1 is None:      # the step after which snapshot 1 is written
        a += da
        k += 1
        if a >= c.A_INIT + 1 * da_save:
            k1 = k
    assert k1 is not None and k1 >= 2
    snap = np.load(f"{c.DATA_DIR}/data.1.npz", allow_pickle=False)["density"]
    p0, v0 = _ics(c)
    sim = PMSim(c.N_CELLS, c.N_PARTS, 32, omega_m0=c.OMEGA_M0, omega_lambda0=c.OMEGA_LAMBDA0,
                omega_k0=c.OMEGA_K0, h0=c.H0)
    sim.set_state(p0, v0)
    sim.set_mass((c.N_CELLS / c.N_PARTS) ** 3)
    sim.run(c.A_INIT, da, k1 - 1)
    rho_before = sim.density(np.float64)
    sim.step(c.A_INIT + (k1 - 1) * da, da)
    rho_after = sim.density(np.float64)
    sim.close()
    assert rel(snap, rho_before) < 1e-5
    assert rel(snap, rho_after) > 1e-3       # the two candidates are far apart at this size


def test_stale_sim_density_raises(config):
    from density import density
    from pmgpu.sim import PMSim
    c = config
    sim = PMSim(c.N_CELLS, c.N_PARTS, 32)
    p0, v0 = _ics(c)
    sim.set_state(p0, v0)
    rho = density(sim.views()[0], (c.N_CELLS / c.N_PARTS) ** 3)
    kept = density(sim.views()[0], (c.N_CELLS / c.N_PARTS) ** 3)
    kept.keep()
    before = np.asarray(rho).copy()
    sim.step(0.02, 0.01)
    with pytest.raises(RuntimeError):
        np.asarray(rho)
    np.testing.assert_array_equal(np.asarray(kept), before)
    sim.close()


def test_pmesh_restart_continues_from_snapshot(config):
    # restart (pmesh.py:38-42) from snapshot 1 and compare with the resident sim stepped the
    # same way from the same file
    from pmgpu.sim import PMSim
    from save_data import from_file
    c = config
    _run()
    xp, vp, a0 = from_file(1)   # read first: the restarted loop writes its snapshots again
    c.RESTART, c.RESTART_FROM_N = True, 1
    pos2, vel2, a_end2, steps2 = _run()
    sim = PMSim(c.N_CELLS, c.N_PARTS, 32, omega_m0=c.OMEGA_M0, omega_lambda0=c.OMEGA_LAMBDA0,
                omega_k0=c.OMEGA_K0, h0=c.H0)
    sim.set_state(xp, vp)
    da = (c.A_END - c.A_INIT) / c.STEPS
    a, k = float(a0), 0
    while a < c.A_END - da:                     # pmesh.py:56-63 from the restart point
        sim.step(a, da)
        a += da
        k += 1
    assert k == steps2 and abs(a - a_end2) < 1e-12
    xs, vs = sim.get_state()
    assert np.abs(periodic_diff(np.asarray(pos2), xs, c.N_CELLS)).max() < 1e-4
    assert rel(np.asarray(vel2), vs) < 1e-5
    sim.close()


def test_pmesh_non_resident_loop_matches_resident(config):
    c = config
    c.STEPS, c.SAVE_DATA = 10, False
    pos, vel, _, steps = _run()
    x_res, v_res = np.asarray(pos), np.asarray(vel)
    c.RESIDENT = False
    pos2, vel2, _, steps2 = _run()
    assert steps2 == steps
    assert isinstance(pos2, np.ndarray)
    assert np.abs(periodic_diff(pos2, x_res, c.N_CELLS)).max() < 2e-3
    assert rel(vel2, v_res) < 1e-3
