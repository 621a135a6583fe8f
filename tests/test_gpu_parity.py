"""GPU parity: every HIP entry point against the float64 oracle on the same seeded inputs.

Tolerances (norm-wise relative error, fields cross zero):
  complex128 reference-layout path (density_k, potential*, advance_time): 1e-6 for float32
    outputs (one float32 rounding of each output), 1e-12 for complex128 spectra;
  fused FP32 step: rho, phi, accelerations <= 1e-5; FP64 step <= 1e-10;
  cell assignment: bit-exact on equal positions.
"""
import math

import numpy as np
import pytest

import pm_oracle as O
from helpers import make_ics, periodic_diff, rel, uniform_positions

pytestmark = pytest.mark.gpu

SIZES = [(16, 8), (32, 16), (48, 24), (64, 32), (96, 48), (128, 64)]


@pytest.fixture(scope="module")
def ops():
    from pmgpu import ops as _ops
    return _ops


# ------------------------------------------------------------------ reference-layout entry points
@pytest.mark.parametrize("n", [8, 16, 32, 48, 64, 96, 192])
def test_density_k_matches_dft(ops, n):
    rho = np.random.default_rng(n).uniform(0, 2, (n, n, n)).astype(np.float32)
    got = ops.density_k(rho)
    ref = np.fft.fftn(rho.astype(np.float64))
    print(f"N={n} density_k rel err {rel(got, ref):.2e}")
    assert got.dtype == np.complex128 and got.shape == (n, n, n)
    assert rel(got, ref) < 1e-12


@pytest.mark.parametrize("n", [8, 32, 96])
def test_potential_real_of_general_spectrum(ops, n):
    r = np.random.default_rng(n)
    x = r.normal(size=(n, n, n)) + 1j * r.normal(size=(n, n, n))
    got = ops.potential_real(x)
    ref = np.fft.ifftn(x).real
    assert got.dtype == np.float32
    assert rel(got, ref) < 1e-6


@pytest.mark.parametrize("n,npart", SIZES)
def test_density_matches_oracle(ops, n, npart):
    pos, _ = make_ics(n, npart, seed=n)
    got = ops.density(pos, (n / npart) ** 3, n)
    ref = O.density(pos, (n / npart) ** 3, n)
    print(f"N={n} density rel err {rel(got, ref):.2e}")
    assert got.dtype == np.float32 and got.shape == (n, n, n)
    assert rel(got, ref) < 1e-6
    assert abs(got.sum(dtype=np.float64) - npart ** 3 * (n / npart) ** 3) < 1e-3 * n ** 3


def test_density_edge_cases(ops):
    n = 16
    # empty particle set
    got = ops.density(np.zeros((3, 0), np.float32), 1.0, n)
    assert got.shape == (n, n, n) and not got.any()
    # Q4: a float32 position equal to N is the particle at 0; N - eps wraps its +1 corner
    pos = np.array([[16.0, 15.75], [0.0, 3.0], [0.0, 15.0]], np.float32)
    got = ops.density(pos, 1.0, n)
    np.testing.assert_allclose(got, O.density(pos, 1.0, n), atol=1e-7)


def test_fourier_grid_matches(ops):
    n = 32
    got = ops.fourier_grid(n)
    # float32 sin of float32 k: NumPy and the device may differ by a couple of ulps
    np.testing.assert_allclose(got, O.fourier_grid(n, ref_dtypes=True), rtol=5e-7, atol=0)
    np.testing.assert_allclose(got, O.fourier_grid(n), rtol=1e-6, atol=0)
    assert got[0, 0, 0] == 0.0


@pytest.mark.parametrize("n,npart", [(32, 16), (48, 24), (64, 32)])
def test_potential_chain_matches_oracle(ops, n, npart):
    pos, _ = make_ics(n, npart, seed=3)
    rho = O.density(pos, (n / npart) ** 3, n).astype(np.float32)
    g = O.fourier_grid(n, ref_dtypes=True)
    a, om = 0.05, 0.3
    ref = O.potential(rho, g, a, om)
    fused = ops.potential(rho, g, a, om)
    rk = ops.density_k(rho)
    pk = ops.potential_k(rk, g, a, om)
    chain = ops.potential_real(pk)
    np.testing.assert_allclose(pk, O.potential_k(O.density_k(rho), g, a, om), rtol=1e-10,
                               atol=1e-10 * np.abs(pk).max())
    print(f"N={n} potential fused {rel(fused, ref):.2e} chain {rel(chain, ref):.2e}")
    assert rel(fused, ref) < 1e-6 and rel(chain, ref) < 1e-6


@pytest.mark.parametrize("n,npart", [(32, 16), (64, 32)])
def test_cells_weights_bit_exact(ops, n, npart):
    pos, _ = make_ics(n, npart, seed=5)
    # 0, the largest float32 below N, N itself (Q4: wraps to 0) and an integer
    pos[0, :4] = [0.0, np.nextafter(np.float32(n), np.float32(0)), float(n), 7.0]
    cells = ops.cell_centers(pos, n)
    ref_cells = O.cell_centers(pos, n)
    np.testing.assert_array_equal(cells, ref_cells)
    assert cells[0, 1] == n - 1 and cells[0, 2] == 0
    t = ops.weights(cells, pos)
    np.testing.assert_allclose(t, O.weights(ref_cells, O.wrap_positions(pos, n)), rtol=1e-6,
                               atol=1e-7)


@pytest.mark.parametrize("n,npart", [(32, 16), (64, 32)])
def test_sweep_and_integrate_match_oracle(ops, n, npart):
    pos, vel = make_ics(n, npart, seed=7)
    rho = O.density(pos, (n / npart) ** 3, n)
    phi = O.potential(rho, O.fourier_grid(n), 0.05, 0.3).astype(np.float32)
    a, da, fa1 = 0.05, 0.01, 0.9
    rx, rv = O.integrate(pos, vel, a, fa1, da, phi, n)
    # integrate (all axes fused)
    x, v = pos.copy(), vel.copy()
    ops.integrate(x, v, a, fa1, da, phi, n)
    assert rel(v - vel, rv - vel) < 1e-5
    assert np.abs(periodic_diff(x, rx, n)).max() < 4e-6 * n
    # the three single-axis sweeps with the reference's precomputed cells and weights
    cells = ops.cell_centers(pos, n)
    t = ops.weights(cells, pos)
    x2, v2 = pos.copy(), vel.copy()
    for d in range(3):
        xd, vd = x2[d].copy(), v2[d].copy()
        ops.sweep_one_direction(cells, xd, vd, t, phi, da, fa1, a, d, n)
        x2[d], v2[d] = xd, vd
    assert rel(v2 - vel, rv - vel) < 1e-5
    assert np.abs(periodic_diff(x2, rx, n)).max() < 4e-6 * n


@pytest.mark.parametrize("n,npart", [(32, 16), (64, 32)])
def test_advance_time_stateless(ops, n, npart):
    pos, vel = make_ics(n, npart, seed=9)
    p = O.OracleParams(n_cells=n, n_parts=npart, omega_m0=0.3, omega_lambda0=0.7, h0=0.7)
    rho = O.density(pos, p.mass, n).astype(np.float32)
    g = O.fourier_grid(n, ref_dtypes=True)
    a, da = 0.02, 0.0098
    rx, rv = O.advance_time(rho, pos, vel, g, a, da, p)
    x, v = pos.copy(), vel.copy()
    out = ops.advance_time(rho, x, v, g, a, da, [0.7, 0.7, 0.0], 0.3)
    assert out[0] is x and out[1] is v            # in place, like the reference
    assert rel(v - vel, rv - vel) < 1e-5
    assert np.abs(periodic_diff(x, rx, n)).max() < 4e-6 * n


# ------------------------------------------------------------------------- fused resident sim
def _sim(n, npart, prec, **kw):
    from pmgpu.sim import PMSim
    return PMSim(n, npart, prec, omega_m0=0.3, omega_lambda0=0.7, omega_k0=0.0, h0=0.7, **kw)


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("n,npart", SIZES)
def test_sim_one_step_matches_oracle(n, npart, prec):
    pos, vel = make_ics(n, npart, seed=11)
    sim = _sim(n, npart, prec)
    sim.set_state(pos, vel)
    x0, v0 = sim.get_state(np.float64 if prec == 64 else np.float32)
    # the displacement round trip loses at most half an ulp of the stored displacement
    s_max = float(np.abs(periodic_diff(pos, O.lattice_q(npart, n), n)).max())
    ulp = float(np.spacing(np.float32(s_max))) if prec == 32 else 0.0
    assert np.abs(periodic_diff(x0, pos, n)).max() <= ulp / 2
    np.testing.assert_array_equal(np.asarray(v0, np.float32), vel)
    p = O.OracleParams(n_cells=n, n_parts=npart, omega_m0=0.3, omega_lambda0=0.7, h0=0.7)
    rho_ref = O.density(x0, p.mass, n)
    rho = sim.density(np.float64)
    a, da = 0.02, 0.0098
    x1r, v1r, phi_ref = O.advance_time(rho_ref, x0, v0, O.fourier_grid(n), a, da, p,
                                       return_phi=True)
    sim.step(a, da)
    phi = sim.potential(np.float64)
    x1, v1 = sim.get_state(np.float64)
    tol = 1e-5 if prec == 32 else 1e-10
    e_rho, e_phi = rel(rho, rho_ref), rel(phi, phi_ref)
    e_acc = rel(v1 - v0, v1r - v0)
    e_dx = rel(periodic_diff(x1, x0, n), periodic_diff(x1r, x0, n))
    print(f"N={n} fp{prec}: rho {e_rho:.2e} phi {e_phi:.2e} accel {e_acc:.2e} drift {e_dx:.2e}")
    assert e_rho < tol and e_phi < tol and e_acc < tol
    assert e_dx < (1e-4 if prec == 32 else 1e-9)
    sim.close()


def test_sim_cell_assignment_bit_exact():
    n, npart = 64, 32
    pos, vel = make_ics(n, npart, seed=13)
    sim = _sim(n, npart, 32)
    sim.set_state(pos, vel)
    for _ in range(3):
        sim.step(0.05, 0.01)
    x, _ = sim.get_state(np.float32)
    # the sim deposits from its own canonical float32 positions: the oracle's deposit of the
    # materialised positions must then see identical cells (weights differ only by rounding)
    rho = sim.density(np.float64)
    ref = O.density(x, sim.n_cells ** 3 / sim.n_p, n)
    assert rel(rho, ref) < 1e-6
    sim.close()


def test_sim_fp32_vs_fp64_multistep_pk():
    n, npart = 64, 32
    pos, vel = make_ics(n, npart, seed=17, box=64.0)
    sims = [_sim(n, npart, 32), _sim(n, npart, 64)]
    for s in sims:
        s.set_state(pos, vel)
    a0, da = 0.02, 0.0098
    for s in sims:
        s.run(a0, da, 100)
    r32, r64 = (s.density(np.float64) for s in sims)
    k, p32, _ = O.power_spectrum(r32, 64.0)
    _, p64, _ = O.power_spectrum(r64, 64.0)
    sel = np.arange(1, len(k) + 1) <= n // 4   # k < k_Nyquist / 2
    err = np.max(np.abs(p32[sel] / p64[sel] - 1))
    print(f"P(k) FP32 vs FP64 after 100 steps: max rel err {err:.2e}")
    assert err < 5e-3
    for s in sims:
        s.close()


def test_sim_run_equals_repeated_step():
    n, npart = 32, 16
    pos, vel = make_ics(n, npart, seed=19)
    s1, s2 = _sim(n, npart, 32), _sim(n, npart, 32)
    s1.set_state(pos, vel)
    s2.set_state(pos, vel)
    a, da = 0.02, 0.0098
    a_end = s1.run(a, da, 5)
    for k in range(5):
        s2.step(a, da)
        a += da
    assert a_end == a
    x1, v1 = s1.get_state()
    x2, v2 = s2.get_state()
    np.testing.assert_allclose(x1, x2, atol=1e-5)
    np.testing.assert_allclose(v1, v2, rtol=1e-5, atol=1e-6 * np.abs(v1).max())


def test_graph_replays_match_eager_steps():
    # the captured step graph is replayed several times, with eager work and synchronisation
    # between replays; every replay must agree with the same steps enqueued kernel by kernel
    n, npart = 64, 32
    pos, vel = make_ics(n, npart, seed=23)
    g, e = _sim(n, npart, 32), _sim(n, npart, 32, graph=False)
    for s in (g, e):
        s.set_state(pos, vel)
    a, da = 0.02, 0.0098
    for rep in range(4):
        a_g = g.run(a, da, 9)          # one replay of 8 captured steps + 1 table-driven step
        a_e = e.run(a, da, 9)
        assert a_g == a_e
        xg, vg = g.get_state()          # eager work between replays
        xe, ve = e.get_state()
        assert np.abs(periodic_diff(xg, xe, n)).max() < 1e-4, rep
        assert rel(vg, ve) < 1e-5, rep
        assert rel(g.density(np.float64), e.density(np.float64)) < 1e-6, rep
        a = a_g
    for s in (g, e):
        s.close()


def test_sim_deconvolve_and_fix_flags_change_result():
    n, npart = 32, 16
    pos, vel = make_ics(n, npart, seed=21)
    base = _sim(n, npart, 64)
    dec = _sim(n, npart, 64, cic_deconvolve=True)
    fix = _sim(n, npart, 64, fix_f_cosmology=True)
    outs = []
    for s in (base, dec, fix):
        s.set_state(pos, vel)
        s.step(0.02, 0.01)
        outs.append(s.get_state(np.float64)[1])
    assert rel(outs[1] - vel, outs[0] - vel) > 1e-3        # deconvolution boosts small scales
    # FIX_F uses f(a+da, [Om, OL, OK]) instead of [H0, OL, OK]: the kick scales by the ratio
    a1 = 0.03
    ratio = O.f(a1, [0.3, 0.7, 0.0]) / O.f(a1, [0.7, 0.7, 0.0])
    assert rel(outs[2] - vel, ratio * (outs[0] - vel)) < 1e-9


@pytest.mark.parametrize("n", [256, 512])
def test_register_pass_green_variants(n):
    # the FP32 register z pass has three k-space operators: the analytic Green's function, a
    # user grid G (set_green, symmetrised over k and -k) and the CIC deconvolution; the grid
    # holding the analytic G must reproduce the analytic step, and the deconvolved FP32 step
    # the deconvolved FP64 step (LDS Stockham passes)
    from pmgpu.device import DeviceArray
    npart = 32
    pos = uniform_positions(n, npart ** 3, seed=n + 1)
    vel = np.zeros_like(pos)
    plain, grid = _sim(n, npart, 32), _sim(n, npart, 32)
    g = DeviceArray.from_numpy(O.fourier_grid(n).astype(np.float32))
    grid.set_green(g)
    dec32, dec64 = _sim(n, npart, 32, cic_deconvolve=True), _sim(n, npart, 64, cic_deconvolve=True)
    phi = []
    for s in (plain, grid, dec32, dec64):
        s.set_state(pos, vel)
        s.step(0.05, 0.01)
        phi.append(s.potential(np.float64))
        s.close()
    e_grid, e_dec = rel(phi[1], phi[0]), rel(phi[2], phi[3])
    print(f"N={n}: user grid vs analytic {e_grid:.2e}; deconvolved FP32 vs FP64 {e_dec:.2e}")
    assert e_grid < 1e-5 and e_dec < 1e-5
    assert rel(phi[3], phi[0]) > 1e-3  # the deconvolution does change the potential


def test_sim_rejects_bad_input():
    from pmgpu.sim import PMSim
    with pytest.raises(RuntimeError, match="supported grid"):
        PMSim(100, 50, 32)
    with pytest.raises(RuntimeError):
        PMSim(64, 32, 16)
    s = _sim(32, 16, 32)
    with pytest.raises(RuntimeError, match="never set"):
        s.step(0.1, 0.01)
    with pytest.raises(ValueError):
        s.set_state(np.zeros((3, 10), np.float32), np.zeros((3, 10), np.float32))
    s.close()


@pytest.mark.parametrize("n,npart", [(256, 32), (384, 48), (512, 32), (768, 48), (1024, 32)])
def test_large_grid_fp32_step_matches_fp64(n, npart):
    # the FP32 passes at 256 / 512 / 1024 keep each column's DFTs in registers (four-step; at
    # 1024 a z column spans 4 GiB of the spectrum, the limit of the passes' 32-bit offsets),
    # 384 and 768 run LDS Stockham passes in FP32 too; the FP64 sim runs the LDS Stockham
    # passes.  Same state in, one step: potential and update agree to FP32 accuracy.
    pos = uniform_positions(n, npart ** 3, seed=n)
    vel = np.zeros_like(pos)
    sims = [_sim(n, npart, 32), _sim(n, npart, 64)]
    for s in sims:
        s.set_state(pos, vel)
        s.step(0.05, 0.01)
    phi32, phi64 = (s.potential(np.float64) for s in sims)
    (x32, v32), (x64, v64) = (s.get_state(np.float64) for s in sims)
    e_phi, e_v = rel(phi32, phi64), rel(v32, v64)
    # positions: FP32 rounding plus the FP32 force error carried over the step's displacement
    # (isolated particles on a fine grid are kicked by up to ~25 cells at N = 1024)
    err = np.abs(periodic_diff(x32, x64, n))
    step = np.abs(periodic_diff(x64, pos, n))
    print(f"N={n}: phi {e_phi:.2e} accel {e_v:.2e} position {err.max():.2e} (max step {step.max():.2e})")
    assert e_phi < 1e-5 and e_v < 1e-5
    assert np.all(err < 1e-4 + 1e-4 * step)
    for s in sims:
        s.close()


def test_fp32_step_256_matches_oracle():
    n, npart = 256, 32
    pos = uniform_positions(n, npart ** 3, seed=5)
    vel = np.zeros_like(pos)
    sim = _sim(n, npart, 32)
    sim.set_state(pos, vel)
    x0, v0 = sim.get_state(np.float32)
    p = O.OracleParams(n_cells=n, n_parts=npart, omega_m0=0.3, omega_lambda0=0.7, h0=0.7)
    rho_ref = O.density(x0, p.mass, n)
    a, da = 0.05, 0.01
    _, v1r, phi_ref = O.advance_time(rho_ref, x0, v0, O.fourier_grid(n), a, da, p, return_phi=True)
    sim.step(a, da)
    e_phi = rel(sim.potential(np.float64), phi_ref)
    e_acc = rel(sim.get_state(np.float64)[1] - v0, v1r - v0)
    print(f"N=256 fp32 vs oracle: phi {e_phi:.2e} accel {e_acc:.2e}")
    assert e_phi < 1e-5 and e_acc < 1e-5
    sim.close()


@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 16, 8), (32, 8, 8), (16, 16, 8), (64, 8, 8),
                                   (8, 8, 8)])
def test_every_tile_shape_matches_oracle(shape):
    n, npart = 64, 32
    pos, vel = make_ics(n, npart, seed=29)
    sim = _sim(n, npart, 32)
    sim.set_tiles(*shape)
    assert sim.tiles() == shape
    sim.set_state(pos, vel)
    x0, v0 = sim.get_state(np.float32)
    p = O.OracleParams(n_cells=n, n_parts=npart, omega_m0=0.3, omega_lambda0=0.7, h0=0.7)
    rho_ref = O.density(x0, p.mass, n)
    assert rel(sim.density(np.float64), rho_ref) < 1e-6
    a, da = 0.02, 0.0098
    _, v1r, _ = O.advance_time(rho_ref, x0, v0, O.fourier_grid(n), a, da, p, return_phi=True)
    sim.run(a, da, 1)
    sim.run(a + da, da, 9)                      # a captured graph with this shape, replayed
    x1, v1 = sim.get_state(np.float64)
    ref = _sim(n, npart, 32)                    # default tiles, same ten steps
    ref.set_state(pos, vel)
    ref.run(a, da, 1)
    ref.run(a + da, da, 9)
    xr, vr = ref.get_state(np.float64)
    assert np.abs(periodic_diff(x1, xr, n)).max() < 1e-4 and rel(v1, vr) < 1e-5
    with pytest.raises(RuntimeError, match="not instantiated"):
        sim.set_tiles(16, 8, 4)
    sim.close()
    ref.close()


def test_autotune_keeps_state_and_results():
    n, npart = 128, 64
    pos, vel = make_ics(n, npart, seed=37)
    sim, ref = _sim(n, npart, 32), _sim(n, npart, 32)
    for s in (sim, ref):
        s.set_state(pos, vel)
    x0, v0 = sim.get_state()
    shape, ms = sim.autotune(reps=2)
    print(f"autotune N=128: {shape} {ms:.3f} ms")
    assert shape in sim.TILE_SHAPES and sim.tiles() == shape and ms > 0
    x1, v1 = sim.get_state()
    np.testing.assert_array_equal(x0, x1)
    np.testing.assert_array_equal(v0, v1)
    for s in (sim, ref):
        s.run(0.02, 0.01, 12)
    xa, va = sim.get_state(np.float64)
    xb, vb = ref.get_state(np.float64)
    assert np.abs(periodic_diff(xa, xb, n)).max() < 1e-4 and rel(va, vb) < 1e-5
    sim.close()
    ref.close()


# ------------------------------------------------------------ fused deposit + R2C x pass
@pytest.mark.parametrize("n,npart,tiles", [(256, 32, (32, 8, 8)), (384, 48, (64, 8, 8)),
                                           (512, 64, (32, 8, 8))])
def test_fused_deposit_r2c_matches_separate_passes(n, npart, tiles):
    # the row-block kernel that deposits density.py:17-47 into LDS and runs the x transform of
    # potential.py:19-21 from there, against the tile deposit + R2C pass (FP64 LDS sums): both
    # FP32 steps are compared with the FP64 step from the same state (the FP32 FFT's own
    # rounding, ~1e-6 in the potential, dominates either), then over ten graph-replayed steps
    # and with the density read back in the middle of a run
    from pmgpu.sim import PMSim
    pos, vel = make_ics(n, npart, seed=41, box=float(n))
    kw = dict(omega_m0=0.3, omega_lambda0=0.7, omega_k0=0.0, h0=0.7)
    fused, sep = PMSim(n, npart, 32, fused_deposit=True, **kw), PMSim(n, npart, 32, **kw)
    ref = PMSim(n, npart, 64, **kw)
    for s in (fused, sep):
        s.set_tiles(*tiles)
    for s in (fused, sep, ref):
        s.set_state(pos, vel)
    a, da = 0.02, 0.0098
    out = []
    for s in (fused, sep, ref):
        _, v0 = s.get_state(np.float64)
        s.step(a, da)
        out.append((s.potential(np.float64), s.get_state(np.float64)[1] - v0))
    ref.close()
    e_f = (rel(out[0][0], out[2][0]), rel(out[0][1], out[2][1]))
    e_s = (rel(out[1][0], out[2][0]), rel(out[1][1], out[2][1]))
    print(f"N={n} {tiles} vs FP64 step: fused phi {e_f[0]:.2e} kick {e_f[1]:.2e}; "
          f"separate phi {e_s[0]:.2e} kick {e_s[1]:.2e}")
    assert max(e_f) < 1e-5 and max(e_s) < 1e-5
    assert e_f[0] < 3 * e_s[0] + 1e-6 and e_f[1] < 3 * e_s[1] + 1e-6
    for s in (fused, sep):
        s.run(a + da, da, 10)
    e_rho = rel(fused.density(np.float64), sep.density(np.float64))
    (xf, vf), (xs, vs) = fused.get_state(np.float64), sep.get_state(np.float64)
    dx = np.abs(periodic_diff(xf, xs, n)).max()
    print(f"  after 11 steps: rho {e_rho:.2e} max|dx| {dx:.2e} v {rel(vf, vs):.2e}")
    # the trajectories part by ~3e-5 cells, which moves CIC weights (and rho) by about as much
    assert e_rho < 5e-5 and dx < 1e-4 and rel(vf, vs) < 1e-5
    # a density read in the middle of a run (the fused path keeps none) leaves the run unchanged
    for s in (fused, sep):
        s.run(a + 11 * da, da, 3)
    assert rel(fused.density(np.float64), sep.density(np.float64)) < 5e-5
    fused.close()
    sep.close()


def test_fused_deposit_step_matches_oracle():
    from pmgpu.sim import PMSim
    n, npart = 256, 32
    pos, vel = make_ics(n, npart, seed=43, box=256.0)
    sim = PMSim(n, npart, 32, fused_deposit=True, omega_m0=0.3, omega_lambda0=0.7, omega_k0=0.0,
                h0=0.7)
    assert sim.tiles() == (32, 8, 8)             # the FP32 default tile shape at this size
    sim.set_state(pos, vel)
    x0, v0 = sim.get_state(np.float64)
    p = O.OracleParams(n_cells=n, n_parts=npart, omega_m0=0.3, omega_lambda0=0.7, h0=0.7)
    rho_ref = O.density(x0, p.mass, n)
    a, da = 0.02, 0.0098
    _, v1r, phi_ref = O.advance_time(rho_ref, x0, v0, O.fourier_grid(n), a, da, p, return_phi=True)
    sim.step(a, da)
    e_phi = rel(sim.potential(np.float64), phi_ref)
    e_acc = rel(sim.get_state(np.float64)[1] - v0, v1r - v0)
    x1, v1 = sim.get_state(np.float64)
    rho1_ref = O.density(x1, p.mass, n)
    _, v2r, phi2_ref = O.advance_time(rho1_ref, x1, v1, O.fourier_grid(n), a + da, da, p, return_phi=True)
    sim.step(a + da, da)                          # this step's spectrum came from the fused kernel
    e_phi2 = rel(sim.potential(np.float64), phi2_ref)
    e_acc2 = rel(sim.get_state(np.float64)[1] - v1, v2r - v1)
    print(f"N=256 fused: step 1 phi {e_phi:.2e} accel {e_acc:.2e}; step 2 phi {e_phi2:.2e} accel {e_acc2:.2e}")
    assert max(e_phi, e_acc, e_phi2, e_acc2) < 1e-5
    sim.close()


# ------------------------------------------------------------------ deterministic deposit
@pytest.mark.parametrize("n,npart", [(64, 32), (256, 32)])
def test_deterministic_deposit_is_bitwise_repeatable(n, npart):
    # density.py:17,37-47 races (its sums depend on thread timing); with fixed-point LDS sums a
    # run from one state gives identical bits every time, and agrees with the oracle
    from pmgpu.sim import PMSim
    pos, vel = make_ics(n, npart, seed=47, box=float(2 * n))
    kw = dict(omega_m0=0.3, omega_lambda0=0.7, omega_k0=0.0, h0=0.7)
    out = []
    for rep in range(2):
        s = PMSim(n, npart, 32, deterministic=True, **kw)
        s.set_state(pos, vel)
        rho0 = s.density(np.float64)
        s.run(0.02, 0.0098, 20)
        out.append((rho0, *s.get_state(np.float32), s.density(np.float32)))
        s.close()
    for a_, b_ in zip(out[0], out[1]):
        np.testing.assert_array_equal(a_, b_)
    p = O.OracleParams(n_cells=n, n_parts=npart, omega_m0=0.3, omega_lambda0=0.7, h0=0.7)
    x0 = O.wrap_positions(pos.astype(np.float64), n)
    assert rel(out[0][0], O.density(x0, p.mass, n)) < 1e-6
    ref = _sim(n, npart, 32)
    ref.set_state(pos, vel)
    ref.run(0.02, 0.0098, 20)
    xr, vr = ref.get_state(np.float32)
    assert np.abs(periodic_diff(out[0][1], xr, n)).max() < 1e-4 and rel(out[0][2], vr) < 1e-5
    ref.close()


# ------------------------------------------------------------------ traversal directions
@pytest.mark.parametrize("n,npart,prec,tiles", [(512, 256, 32, (32, 8, 8)), (256, 128, 32, (32, 8, 8)),
                                                (128, 64, 64, (16, 16, 16))])
def test_backward_x_pass_traversal_gives_the_bits_of_forward(n, npart, prec, tiles):
    # By default the step's R2C and C2R passes walk their row blocks backwards (per XCD run) so
    # that they read their producer's latest output from the last-level cache; the blocks
    # compute the same values either way, so with the order-independent deposit a 20-step run
    # (two replays of an 8-step graph and 4 direct steps) is bit-identical to all-forward order.
    from pmgpu.sim import PMSim
    pos, vel = make_ics(n, npart, seed=61, box=float(2 * n))
    kw = dict(omega_m0=0.3, omega_lambda0=0.7, omega_k0=0.0, h0=0.7)
    out = []
    for forward in (True, False):
        s = PMSim(n, npart, prec, deterministic=True, forward_order=forward, **kw)
        s.set_tiles(*tiles)
        s.set_state(pos, vel)
        a = s.run(0.02, 0.0098, 20)
        out.append((a, *s.get_state(np.float64), s.density(np.float64), s.potential(np.float64)))
        s.close()
    assert out[0][0] == out[1][0]
    for name, a_, b_ in zip(("x", "v", "rho", "phi"), out[0][1:], out[1][1:]):
        assert np.array_equal(a_, b_), name


# ------------------------------------------------------------------ particles that move far
def test_particles_beyond_the_tile_neighbourhood_are_placed():
    # The tile scatter places a source tile's particles through ranges the gather reserved in
    # the keys of its 4^3 tile neighbourhood; a particle whose new tile lies two or more tiles
    # away is counted apart and placed after those ranges.  Two percent of the particles get
    # +-40 per axis of extra velocity (about 25 cells per step, three 8-cell tiles in y and z):
    # with the order-independent deposit, three steps in 32 x 8 x 8 tiles (tile scatter) give
    # the bits of three steps in 8^3 tiles (particle-parallel scatter).
    from pmgpu.sim import PMSim
    n, npart = 256, 64
    pos, vel = make_ics(n, npart, seed=83, box=float(2 * n))
    rng = np.random.default_rng(83)
    sel = rng.choice(npart ** 3, npart ** 3 // 50, replace=False)
    vel[:, sel] += rng.choice(np.float32([-40.0, 40.0]), size=(3, sel.size))
    kw = dict(omega_m0=0.3, omega_lambda0=0.7, omega_k0=0.0, h0=0.7)
    out = []
    for tiles in ((32, 8, 8), (8, 8, 8)):
        s = PMSim(n, npart, 32, deterministic=True, **kw)
        s.set_tiles(*tiles)
        s.set_state(pos, vel)
        x0, _ = s.get_state(np.float64)
        s.step(0.05, 0.01)
        x1, _ = s.get_state(np.float64)
        moved = np.abs(periodic_diff(x1[:, sel], x0[:, sel], n))
        assert (moved[1:] > 17).mean() > 0.9  # beyond the neighbourhood along y and z
        s.run(0.06, 0.01, 2)
        out.append((*s.get_state(np.float64), s.density(np.float64)))
        s.close()
    for name, a_, b_ in zip(("x", "v", "rho"), out[0], out[1]):
        assert np.array_equal(a_, b_), name


def test_clustered_particles_step_alike_in_both_scatters():
    # 32^3 particles packed into a 12^3-cell cube that straddles tile corners: thousands of
    # particles per tile, so the gather, the deposit and the tile scatter each loop over many
    # batches of a tile.  The first density matches the oracle's; with the order-independent
    # deposit three steps in 32 x 8 x 8 tiles (tile scatter) give the bits of three steps in
    # 8^3 tiles (particle-parallel scatter).
    from pmgpu.sim import PMSim
    n, npart = 128, 32
    rng = np.random.default_rng(91)
    pos = rng.uniform(26.0, 38.0, size=(3, npart ** 3)).astype(np.float32)
    vel = rng.normal(0.0, 2.0, size=(3, npart ** 3)).astype(np.float32)
    kw = dict(omega_m0=0.3, omega_lambda0=0.7, omega_k0=0.0, h0=0.7)
    out = []
    for tiles in ((32, 8, 8), (8, 8, 8)):
        s = PMSim(n, npart, 32, deterministic=True, **kw)
        s.set_tiles(*tiles)
        s.set_state(pos, vel)
        x0, _ = s.get_state(np.float32)
        p = O.OracleParams(n_cells=n, n_parts=npart, omega_m0=0.3, omega_lambda0=0.7, h0=0.7)
        rho_ref = O.density(O.wrap_positions(x0.astype(np.float64), n), p.mass, n)
        assert rel(s.density(np.float64), rho_ref) < 1e-6, tiles
        s.run(0.05, 0.01, 3)
        out.append((*s.get_state(np.float64), s.density(np.float64)))
        s.close()
    for name, a_, b_ in zip(("x", "v", "rho"), out[0], out[1]):
        assert np.array_equal(a_, b_), name


def test_tile_shape_changes_mid_run_keep_the_bits():
    # The tile scatter keeps its arrays (far counts, reserved range starts, the replaced own
    # offsets) in the counts buffer after the current shape's counts, so a change of tile shape
    # moves them.  A deterministic run that switches between shapes with and without the tile
    # scatter and between tile counts gives the bits of a run in one shape.
    from pmgpu.sim import PMSim
    n, npart = 128, 64
    pos, vel = make_ics(n, npart, seed=97, box=float(2 * n))
    kw = dict(omega_m0=0.3, omega_lambda0=0.7, omega_k0=0.0, h0=0.7)
    ref = PMSim(n, npart, 32, deterministic=True, **kw)
    ref.set_tiles(32, 8, 8)
    ref.set_state(pos, vel)
    a_ref = ref.run(0.05, 0.01, 20)
    s = PMSim(n, npart, 32, deterministic=True, **kw)
    s.set_state(pos, vel)
    a = 0.05
    for tiles in ((32, 8, 8), (8, 8, 8), (16, 16, 16), (32, 16, 8), (32, 8, 8)):
        s.set_tiles(*tiles)
        a = s.run(a, 0.01, 4)
    assert a == a_ref
    for name, x, y in zip(("x", "v"), s.get_state(np.float64), ref.get_state(np.float64)):
        assert np.array_equal(x, y), name
    assert np.array_equal(s.density(np.float64), ref.density(np.float64))
    s.close()
    ref.close()


# ------------------------------------------------------------------ corrupt state detection
def test_nonfinite_state_is_reported():
    from pmgpu import _lib
    n, npart = 64, 32
    pos, vel = make_ics(n, npart, seed=53)
    sim = _sim(n, npart, 32)
    bad = pos.copy()
    bad[1, 77] = np.nan
    with pytest.raises(RuntimeError, match=f"code {_lib.PM_E_CORRUPT}.*1 non-finite"):
        sim.set_state(bad, vel)
    # a velocity that overflows the drift: the step reports it instead of binning the particle
    # at a clamped cell silently
    sim.set_state(pos, vel)
    hot = vel.copy()
    hot[0, 5] = 3e38
    sim.set_state(pos, hot)
    with pytest.raises(RuntimeError, match="non-finite"):
        sim.step(0.02, 0.0098)
    # the counters are cleared once reported: a good state steps normally again
    sim.set_state(pos, vel)
    sim.step(0.02, 0.0098)
    sim.run(0.02 + 0.0098, 0.0098, 9)
    sim.close()
