"""The 256^3-particle / 512^3-grid workload at its full size, checked through properties that do
not need the float64 oracle to run at that size:

* the FP32 step against the FP64 step from the same state (the FP64 step is pinned to the
  oracle at the smaller sizes): density, potential and accelerations <= 1e-5 norm-wise;
* mass conservation of the deposit (sum of rho = N^3, the mean density 1 of pmesh.py:28);
* momentum conservation of the PM force (sum over particles of the kicks = 0: CIC deposit and
  CIC gather share their weights and the central difference is odd);
* bit-exact particle-to-cell assignment of the reference-layout entry point (integrate.py:20);
* the potential of the FP32 and FP64 steps against the same solve done with ROCm's FFT library;
* the whole 1000-step run: P(k) of FP32 against FP64 within 0.5 % for k < k_Nyquist/2;
* the whole run twice in the fixed-point deposit mode: bit-identical positions and velocities.

Against the CPU port of the reference (oracle/pm_port: density.py / integrate.py loop for loop
with the reference's float32 storage and complex128 FFTs):

* one step at the full size from the same state: density, potential and accelerations of the
  FP32 and FP64 steps within 1e-5 norm-wise;
* the whole run from the host (NumPy) initial conditions of seed 7: P(k) of the FP32 and FP64
  runs within 0.5 % of the port's for k < k_Nyquist/2, the port's run being the committed
  fixture tests/golden/c2_pk.npz (written by tests/golden/make_c2_pk.py; its initial
  conditions' checksums are checked first).
"""
from pathlib import Path

import numpy as np
import pytest

from helpers import ic_params, periodic_diff, rel

pytestmark = pytest.mark.gpu

N, NP, BOX = 512, 256, 250.0
A0, DA = 0.02, (1.0 - 0.02) / 1000


def _sim(prec):
    from pmgpu.sim import PMSim
    return PMSim(N, NP, prec, omega_m0=0.3, omega_lambda0=0.7, omega_k0=0.0, h0=0.7)


@pytest.fixture(scope="module")
def ics():
    import initial_conditions as ic
    s = _sim(32)
    ic.zeldovich_device(s, ic_params(N, NP, box=BOX, seed=7))
    pos, vel = s.get_state(np.float32)
    s.close()
    return pos, vel


def test_full_size_step_fp32_matches_fp64(ics):
    pos, vel = ics
    sims = [_sim(32), _sim(64)]
    rho, phi, state0, state1 = [], [], [], []
    for s in sims:
        s.set_state(pos, vel)
        state0.append(s.get_state(np.float64))
        s.deposit()
        rho.append(s.density(np.float64))
        s.step(A0, DA)
        phi.append(s.potential(np.float64))
        state1.append(s.get_state(np.float64))
        s.close()
    # mass: every particle deposits m = (N / N_PARTS)^3 in total
    m32, m64 = rho[0].sum() / N ** 3 - 1, rho[1].sum() / N ** 3 - 1
    (x0, v0), (x0d, v0d) = state0
    (x32, v32), (x64, v64) = state1
    acc32, acc64 = v32 - v0, v64 - v0d
    e_rho, e_phi, e_acc = rel(rho[0], rho[1]), rel(phi[0], phi[1]), rel(acc32, acc64)
    err = np.abs(periodic_diff(x32, x64, N))
    step = np.abs(periodic_diff(x64, x0d, N))
    # momentum: the summed kick per axis against the summed kick magnitude
    p32 = np.abs(acc32.sum(axis=1)).max() / np.abs(acc32).sum(axis=1).max()
    p64 = np.abs(acc64.sum(axis=1)).max() / np.abs(acc64).sum(axis=1).max()
    print(f"512^3/256^3: mass {m32:.1e} / {m64:.1e}; rho {e_rho:.2e} phi {e_phi:.2e} accel {e_acc:.2e}; "
          f"position {err.max():.2e} (max step {step.max():.2e}); momentum {p32:.1e} / {p64:.1e}")
    # (measured on MI355X: mass 4.6e-12 / 1.1e-15, momentum 5.8e-9 / 2.2e-16)
    assert abs(m64) < 1e-12 and abs(m32) < 1e-9
    assert e_rho < 1e-5 and e_phi < 1e-5 and e_acc < 1e-5
    assert np.all(err < 1e-4 + 1e-4 * step)
    assert p64 < 1e-12 and p32 < 1e-6


def _rel_chunked(a, b):
    """rel() over the leading axis in slabs (grids of 1024^3 values)."""
    num = den = 0.0
    for i in range(0, a.shape[0], 64):
        x, y = np.asarray(a[i:i + 64], np.float64), np.asarray(b[i:i + 64], np.float64)
        num += float(np.sum((x - y) ** 2))
        den += float(np.sum(y ** 2))
    return (num / den) ** 0.5


def test_largest_run_size_step_fp32_matches_fp64():
    # 512^3 particles on a 1024^3 grid (box 500, seed 42): the FP32 step (register FFT passes on
    # the page-tiled spectrum) against the FP64 step (LDS Stockham passes, plain layout)
    import initial_conditions as ic
    from pmgpu.sim import PMSim
    n, npart = 1024, 512
    s32 = PMSim(n, npart, 32, omega_m0=0.3, omega_lambda0=0.7, omega_k0=0.0, h0=0.7)
    ic.zeldovich_device(s32, ic_params(n, npart, box=500.0, seed=42))
    pos, vel = s32.get_state(np.float32)
    out = {}
    for prec in (32, 64):
        s = s32 if prec == 32 else PMSim(n, npart, 64, omega_m0=0.3, omega_lambda0=0.7, omega_k0=0.0, h0=0.7)
        s.set_state(pos, vel)
        _, v0 = s.get_state(np.float64)
        rho = s.density(np.float32)
        s.step(A0, (1.0 - A0) / 500)
        phi = s.potential(np.float32)
        _, v1 = s.get_state(np.float64)
        v1 -= v0
        del v0
        out[prec] = (rho, phi, v1)
        s.close()
    mass = float(out[32][0].sum(dtype=np.float64)) / n ** 3 - 1
    e_rho = _rel_chunked(out[32][0], out[64][0])
    e_phi = _rel_chunked(out[32][1], out[64][1])
    e_acc = _rel_chunked(out[32][2], out[64][2])
    print(f"1024^3/512^3: mass {mass:.1e}; rho {e_rho:.2e} phi {e_phi:.2e} accel {e_acc:.2e}")
    assert abs(mass) < 1e-6
    assert e_rho < 1e-5 and e_phi < 1e-5 and e_acc < 1e-5


def _potential_rocfft(rho, a, omega_m0=0.3):
    """potential.py:7-10 with ROCm's FFT library (hipFFT over rocFFT, float64 D2Z / Z2D) on
    the GPU: the cross-check of the library's own FFTs.  The k-space factor -3 Om / (8 a) G(k),
    G = 1 / sum sin^2(k_i / 2), G(0) = 0 (fourier_utils.py:5-16), is applied on the host."""
    import ctypes as C
    from pmgpu.device import DeviceArray
    fft = C.CDLL("/opt/rocm/lib/libhipfft.so")
    hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    n = rho.shape[0]
    m = n // 2 + 1
    d_r = DeviceArray.from_numpy(np.ascontiguousarray(rho, np.float64))
    d_k = DeviceArray.empty((n, n, m, 2), np.float64)
    fwd, inv = C.c_void_p(), C.c_void_p()
    assert fft.hipfftPlan3d(C.byref(fwd), n, n, n, 0x6A) == 0  # HIPFFT_D2Z
    assert fft.hipfftPlan3d(C.byref(inv), n, n, n, 0x6C) == 0  # HIPFFT_Z2D
    try:
        assert fft.hipfftExecD2Z(fwd, C.c_void_p(d_r.ptr), C.c_void_p(d_k.ptr)) == 0
        assert hip.hipDeviceSynchronize() == 0
        rk = d_k.numpy().view(np.complex128)[..., 0]
        s2 = np.sin(np.pi * np.fft.fftfreq(n)) ** 2
        s2x = np.sin(np.pi * np.fft.rfftfreq(n)) ** 2
        for z in range(n):  # plane by plane: G of one plane at a time
            den = s2[z] + s2[:, None] + s2x[None, :]
            g = np.divide(1.0, den, out=np.zeros_like(den), where=den > 0)
            rk[z] *= g * (-3.0 * omega_m0 / 8.0 / a)
        d_k.copy_from(rk.view(np.float64).reshape(n, n, m, 2))
        del rk
        assert fft.hipfftExecZ2D(inv, C.c_void_p(d_k.ptr), C.c_void_p(d_r.ptr)) == 0
        assert hip.hipDeviceSynchronize() == 0
        return d_r.numpy() / float(n) ** 3
    finally:
        fft.hipfftDestroy(fwd)
        fft.hipfftDestroy(inv)


def test_full_size_potential_matches_rocfft(ics):
    pos, vel = ics
    for prec, tol in ((32, 1e-5), (64, 1e-12)):
        s = _sim(prec)
        s.set_state(pos, vel)
        rho = s.density(np.float64)
        s.step(A0, DA)
        phi = s.potential(np.float64)
        s.close()
        ref = _potential_rocfft(rho, A0)
        e = rel(phi, ref)
        print(f"512^3 FP{prec} potential vs hipFFT/rocFFT (float64): {e:.2e}")
        assert e < tol


def test_full_size_cell_assignment_bit_exact(ics):
    from pmgpu import ops
    pos, _ = ics
    cells = ops.cell_centers(pos, N)
    np.testing.assert_array_equal(cells, np.floor(pos).astype(np.int64) % N)


def test_full_size_run_power_spectrum_fp32_vs_fp64(ics):
    pos, vel = ics
    sims = [_sim(32), _sim(64)]
    for s in sims:
        s.set_state(pos, vel)
        a = s.run(A0, DA, 1000)
    k, p32, _ = sims[0].power_spectrum(BOX)
    _, p64, _ = sims[1].power_spectrum(BOX)
    for s in sims:
        s.close()
    sel = np.arange(1, len(k) + 1) <= N // 4   # k < k_Nyquist / 2
    err = float(np.max(np.abs(p32[sel] / p64[sel] - 1)))
    print(f"P(k) FP32 vs FP64 after 1000 steps (a = {a:.4f}): max rel err {err:.2e}")
    assert abs(a - 1.0) < 1e-9
    assert err < 5e-3


def test_full_size_deterministic_run_is_bitwise_repeatable(ics):
    # the whole 1000-step run twice in the fixed-point deposit mode: identical positions and
    # velocities (the default FP64 LDS deposit's sums depend on the order its atomics land in,
    # as density.py:17's racy scatter does).  The P(k) estimator sums its shells with FP64
    # atomics, outside the mode: it agrees to rounding of those sums.
    from pmgpu.sim import PMSim
    pos, vel = ics
    out = []
    for rep in range(2):
        s = PMSim(N, NP, 32, deterministic=True, omega_m0=0.3, omega_lambda0=0.7, omega_k0=0.0, h0=0.7)
        s.set_state(pos, vel)
        s.run(A0, DA, 1000)
        out.append((*s.get_state(np.float32), s.power_spectrum(BOX)[1]))
        s.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    np.testing.assert_allclose(out[0][2], out[1][2], rtol=1e-13, atol=0)


def test_full_size_tile_scatter_gives_the_bits_of_the_particle_parallel_scatter(ics):
    # the workload's own shape and tile shape: 40 steps in 32 x 8 x 8 tiles (the tile scatter,
    # its reserved ranges and bin codes) against 40 steps in 8^3 tiles (the particle-parallel
    # scatter with device cursors), both with the order-independent deposit
    from pmgpu.sim import PMSim
    pos, vel = ics
    out = []
    for tiles in ((32, 8, 8), (8, 8, 8)):
        s = PMSim(N, NP, 32, deterministic=True, omega_m0=0.3, omega_lambda0=0.7, omega_k0=0.0, h0=0.7)
        s.set_tiles(*tiles)
        s.set_state(pos, vel)
        s.run(A0, DA, 40)
        out.append(s.get_state(np.float32))
        s.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_full_size_step_matches_cpu_port(ics):
    # density.py:7-48, potential.py:7-29 and integrate.py:9-97 through the port, one step
    import pm_oracle as O
    import pm_port
    pos, vel = ics
    workers = max(1, min(pm_port.cores(), 16))
    pm_port.load().port_set_threads(workers)
    mass = (N / NP) ** 3
    x, v = pos.copy(), vel.copy()
    rho_c, phi_c = pm_port.step(x, v, O.fourier_grid(N, ref_dtypes=True), mass, A0, DA,
                                0.3, 0.7, 0.0, 0.7, workers)
    acc_c = v.astype(np.float64) - vel
    for prec in (32, 64):
        s = _sim(prec)
        s.set_state(pos, vel)
        _, v0 = s.get_state(np.float64)
        rho = s.density(np.float64)
        s.step(A0, DA)
        phi = s.potential(np.float64)
        _, v1 = s.get_state(np.float64)
        s.close()
        e_rho, e_phi, e_acc = rel(rho, rho_c), rel(phi, phi_c), rel(v1 - v0, acc_c)
        print(f"512^3 FP{prec} step vs CPU port: rho {e_rho:.2e} phi {e_phi:.2e} accel {e_acc:.2e}")
        assert e_rho < 1e-5 and e_phi < 1e-5 and e_acc < 1e-5


C2_PK = Path(__file__).resolve().parent / "golden" / "c2_pk.npz"


def test_full_run_power_spectrum_matches_cpu_port_fixture():
    from helpers import make_ics
    if not C2_PK.exists():
        pytest.skip("fixture tests/golden/c2_pk.npz not written (tests/golden/make_c2_pk.py)")
    fx = np.load(C2_PK, allow_pickle=False)
    pos, vel = make_ics(N, NP, seed=7, box=BOX)
    sums = np.concatenate([pos.astype(np.float64).sum(axis=1), vel.astype(np.float64).sum(axis=1)])
    np.testing.assert_allclose(sums, fx["ic_sums"], rtol=1e-9, atol=1e-6)
    steps = int(fx["steps"])
    assert steps == 1000
    sel = np.arange(1, len(fx["k"]) + 1) <= N // 4   # k < k_Nyquist / 2
    for prec in (32, 64):
        s = _sim(prec)
        s.set_state(pos, vel)
        a = s.run(A0, DA, steps)
        k, pk, cnt = s.power_spectrum(BOX)
        s.close()
        np.testing.assert_array_equal(cnt, fx["counts"])
        np.testing.assert_allclose(k, fx["k"], rtol=1e-12)
        err = float(np.max(np.abs(pk[sel] / fx["pk"][sel] - 1)))
        print(f"FP{prec} GPU vs CPU port after {steps} steps (a = {a:.6f}): max P(k) rel err "
              f"{err:.2e} for k < k_Ny/2; largest scale ratio {pk[0] / fx['pk'][0]:.6f}")
        assert abs(a - float(fx["a_final"])) < 1e-12
        assert err < 5e-3
