"""CPU oracle for the particle-mesh (PM) step — TEST INFRASTRUCTURE ONLY.

This module restates, in float64 NumPy, what the reference's PM step computes.  It is the
checker that the HIP library is compared against; nothing in the product package imports it.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
use it.

Each function names the reference file:line it follows (paths are relative to the reference's
``src/`` directory).  Deliberate departures from the reference:

* Q3  ``fourier_grid`` leaves G(k=0) uninitialised (fourier_utils.py:16); here G(0) = 0.
* Q4  a float32 position equal to N after ``% N`` (integrate.py:95) gives CIC weights of
      +-(N-1); positions are wrapped into [0, N) with ``wrap_positions`` before use.
* Q5  the reference deposit races (density.py:17,37-47); here it is ``np.add.at`` (exact).
* Q9  pyFFTW's planner may overwrite its input (potential.py:20,27); here the FFT is the DFT
      definition (``numpy.fft``: forward unnormalised, inverse normalised by 1/N^3).

Reproduced on purpose: Q1 (``f`` receives H0 in the Omega_m slot, integrate.py:12), Q6
(every axis is kicked with forces at pre-step positions, integrate.py:20-25), Q7 (the step
count comes from float64 accumulation of ``a``, pmesh.py:56,63) and Q8 (positions[0] is grid
axis 2, density.py:37).

Every function here is pinned by an independent check in ``tests/test_oracle.py`` (closed
forms, invariants, brute force); see that file for which check pins which function.

Two precisions are provided: ``dtype=np.float64`` (the oracle proper) and a "reference-dtype"
mode (``ref_dtypes=True``) that stores float32 where the reference stores float32 (Q11), so
the gap between the reference's own storage precision and float64 can be reported.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

__all__ = [
    "wrap_positions", "density", "fourier_grid", "density_k", "potential_k",
    "potential_real", "potential", "f", "H", "Dt", "cell_centers", "weights",
    "sweep_one_direction", "integrate", "advance_time", "step_count", "OracleParams",
    "philox4x32_10", "white_noise",
    "run", "lattice_q", "power_spectrum", "shaped_field", "zeldovich_displacement",
    "zeldovich_positions",
]


@dataclass
class OracleParams:
    """The configure_me.py constants the PM step reads (configure_me.py:7-8,34-40)."""
    n_cells: int = 512
    n_parts: int = 256
    omega_m0: float = 0.31
    omega_lambda0: float = 0.69
    omega_k0: float = 0.0
    h0: float = 0.68
    a_init: float = 0.01
    a_end: float = 1.0
    steps: int = 1000

    @property
    def mass(self) -> float:
        # pmesh.py:28 — dens_contrast = (N_CELLS/N_PARTS)**3
        return (self.n_cells / self.n_parts) ** 3


# --------------------------------------------------------------------------------------------
# positions
# --------------------------------------------------------------------------------------------
def wrap_positions(x: np.ndarray, n: int) -> np.ndarray:
    """Canonical periodic wrap into [0, n) in the array's own dtype (fix for Q4).

    One add/subtract of n is applied where needed; a value that rounds to exactly n after
    adding n (a tiny negative input) is mapped to 0.  Inputs must lie in (-n, 2n).
    """
    x = np.array(x, copy=True)
    nn = x.dtype.type(n)
    neg = x < 0
    x[neg] = x[neg] + nn
    hi = x >= nn
    x[hi] = x[hi] - nn
    return x


def lattice_q(n_parts: int, n_cells: int, dtype=np.float32) -> np.ndarray:
    """Unperturbed lattice positions, (3, n_parts**3), in the reference's particle order.

    zeldovich.py:79-83: ``linspace(0, N_CELLS - N_CELLS/N_PARTS, N_PARTS) + 0.5`` per axis,
    ``meshgrid(..., indexing='ij')[d]`` flattened in C order, so particle
    p = (i*N_PARTS + j)*N_PARTS + k sits at (q_i, q_j, q_k) for positions[0..2].
    """
    axis = np.linspace(0, n_cells - n_cells / n_parts, n_parts) + 0.5
    grids = np.meshgrid(axis, axis, axis, indexing="ij")
    return np.stack([g.reshape(-1) for g in grids]).astype(dtype)


# --------------------------------------------------------------------------------------------
# CIC deposit — density.py:7-48
# --------------------------------------------------------------------------------------------
def density(positions: np.ndarray, mass: float, n_cells: int, ref_dtypes: bool = False
            ) -> np.ndarray:
    """CIC mass deposit onto an (N, N, N) grid indexed [z, y, x].

    density.py:19-21 cell = int32(floor(x)) % N; :24-30 d = x - cell, t = 1 - d;
    :33-35 the +1 neighbour wraps mod N; :37-47 the eight products mass*w into
    grid[z, y, x] with positions[0] on grid axis 2 (Q8).  Accumulation is exact-order
    ``np.add.at`` in float64 (Q5 fixed); ``ref_dtypes`` returns the float32 grid the
    reference allocates (density.py:11).
    """
    n = int(n_cells)
    x = np.asarray(positions)
    if x.ndim != 2 or x.shape[0] != 3:
        raise ValueError("positions must have shape (3, N_p)")
    x = wrap_positions(x, n)                       # Q4
    c = np.floor(x).astype(np.int64) % n           # density.py:19-21
    d = x.astype(np.float64) - c                   # density.py:24-26 (float32 - int -> float64)
    t = 1.0 - d                                    # density.py:28-30
    cp = (c + 1) % n                               # density.py:33-35
    grid = np.zeros((n, n, n), dtype=np.float64)
    xc, yc, zc = c
    X, Y, Z = cp
    dx, dy, dz = d
    tx, ty, tz = t
    m = float(mass)
    np.add.at(grid, (zc, yc, xc), m * tx * ty * tz)   # density.py:37
    np.add.at(grid, (zc, yc, X), m * dx * ty * tz)    # :39
    np.add.at(grid, (zc, Y, xc), m * tx * dy * tz)    # :40
    np.add.at(grid, (Z, yc, xc), m * tx * ty * dz)    # :41
    np.add.at(grid, (zc, Y, X), m * dx * dy * tz)     # :43
    np.add.at(grid, (Z, Y, xc), m * tx * dy * dz)     # :44
    np.add.at(grid, (Z, yc, X), m * dx * ty * dz)     # :45
    np.add.at(grid, (Z, Y, X), m * dx * dy * dz)      # :47
    return grid.astype(np.float32) if ref_dtypes else grid


# --------------------------------------------------------------------------------------------
# Green's function and Poisson solve — fourier_utils.py:5-16, potential.py:7-29
# --------------------------------------------------------------------------------------------
def fourier_grid(n_cells: int, ref_dtypes: bool = False) -> np.ndarray:
    """G(k) = 1 / (sin^2(kx/2) + sin^2(ky/2) + sin^2(kz/2)), k = 2*pi*fftfreq(N).

    fourier_utils.py:8-16.  The reference builds k in float32 (:9-11) and evaluates in float32;
    ``ref_dtypes`` reproduces that.  G(0) = 0 (Q3).
    """
    n = int(n_cells)
    if ref_dtypes:
        k = np.array(2 * np.pi * np.fft.fftfreq(n), dtype=np.float32)
    else:
        k = 2 * np.pi * np.fft.fftfreq(n)
    s = np.sin(k / 2) ** 2
    ksq = s[:, None, None] + s[None, :, None] + s[None, None, :]
    out = np.zeros_like(ksq)
    np.divide(1, ksq, out=out, where=ksq != 0)
    return out


def density_k(rho: np.ndarray) -> np.ndarray:
    """Forward, unnormalised 3-D DFT of the density (potential.py:17-21, FFTW_FORWARD)."""
    return np.fft.fftn(np.asarray(rho, dtype=np.complex128), axes=(0, 1, 2))


def potential_k(rho_k: np.ndarray, fgrid: np.ndarray, a: float, omega_m0: float
                ) -> np.ndarray:
    """phi_k = -3*Omega_m/8/a * G * rho_k (potential.py:12-15)."""
    return -3 * omega_m0 / 8 / a * np.asarray(fgrid, dtype=np.float64) * rho_k


def potential_real(phi_k: np.ndarray, ref_dtypes: bool = False) -> np.ndarray:
    """Inverse DFT normalised by 1/N^3, real part (potential.py:23-29)."""
    out = np.fft.ifftn(phi_k, axes=(0, 1, 2)).real
    return out.astype(np.float32) if ref_dtypes else out


def potential(rho: np.ndarray, fgrid: np.ndarray, a: float, omega_m0: float,
              ref_dtypes: bool = False) -> np.ndarray:
    """potential.py:7-10: density_k -> potential_k -> potential_real."""
    return potential_real(potential_k(density_k(rho), fgrid, a, omega_m0), ref_dtypes)


# --------------------------------------------------------------------------------------------
# Cosmology — cosmology.py:11-36
# --------------------------------------------------------------------------------------------
def H(a: float, h0: float, cosmology) -> float:
    """cosmology.py:11-18."""
    om, ol, ok = cosmology
    return math.sqrt(h0 ** 2 * (om / a ** 3 + ok / a ** 2 + ol))


def f(a: float, cosmology) -> float:
    """cosmology.py:20-27: 1/sqrt((c0 + c2*a + c1*a^3)/a) for cosmology = [c0, c1, c2]."""
    om, ol, ok = cosmology
    return 1 / math.sqrt((om + ok * a + ol * a ** 3) / a)


def Dt(a: float, cosmology) -> float:
    """cosmology.py:29-36 (growth-factor approximation; IC use only)."""
    om, ol, ok = cosmology
    return 5 / 2 / om / (om ** (4 / 7) - ol + (1 + om / 2) * (1 + ol / 70)) * a


# --------------------------------------------------------------------------------------------
# CIC gather + kick/drift — integrate.py:9-97
# --------------------------------------------------------------------------------------------
def cell_centers(positions: np.ndarray, n_cells: int) -> np.ndarray:
    """integrate.py:20 — floor(positions).astype(int64) % N_CELLS (after the Q4 wrap)."""
    x = wrap_positions(np.asarray(positions), n_cells)
    return np.floor(x).astype(np.int64) % n_cells


def weights(cells: np.ndarray, positions: np.ndarray, ref_dtypes: bool = False) -> np.ndarray:
    """integrate.py:27-52: the eight CIC weights, rows ordered
    (txtytz, dxtytz, txdytz, txtydz, dxdytz, dxtydz, txdydz, dxdydz)."""
    x = np.asarray(positions)                        # already wrapped into [0, N) by the caller
    d = x.astype(np.float64) - cells                 # integrate.py:36-38
    dx, dy, dz = d
    tx, ty, tz = 1 - dx, 1 - dy, 1 - dz              # :40-42
    t = np.stack([tx * ty * tz, dx * ty * tz, tx * dy * tz, tx * ty * dz,
                  dx * dy * tz, dx * ty * dz, tx * dy * dz, dx * dy * dz])  # :44-51
    return t.astype(np.float32) if ref_dtypes else t


def _gather_gradient(cells, t, phi, direction, n):
    """Sum over the eight corners of t_j * (phi[c_j - e_d] - phi[c_j + e_d]) / 2.

    integrate.py:61-92: cc_n shifts the cell of axis ``direction`` by -1 (numba wraps the -1
    index, :64), cc_p by +1 mod N (:65); the +1 corner offsets wrap mod N (:77-82).
    """
    cp = cells.copy()
    cn = cells.copy()
    cp[direction] = (cells[direction] + 1) % n
    cn[direction] = (cells[direction] - 1) % n
    x, y, z = cp
    x2, y2, z2 = cn
    X, Y, Z = (x + 1) % n, (y + 1) % n, (z + 1) % n
    X2, Y2, Z2 = (x2 + 1) % n, (y2 + 1) % n, (z2 + 1) % n
    p = phi
    g = -p[z, y, x] + p[z2, y2, x2]            # :84
    g_x = -p[z, y, X] + p[z2, y2, X2]          # :85
    g_y = -p[z, Y, x] + p[z2, Y2, x2]          # :86
    g_z = -p[Z, y, x] + p[Z2, y2, x2]          # :87
    g_xy = -p[z, Y, X] + p[z2, Y2, X2]         # :88
    g_xz = -p[Z, y, X] + p[Z2, y2, X2]         # :89
    g_yz = -p[Z, Y, x] + p[Z2, Y2, x2]         # :90
    g_xyz = -p[Z, Y, X] + p[Z2, Y2, X2]        # :91
    return (g * t[0] + g_x * t[1] + g_y * t[2] + g_z * t[3] + g_xy * t[4] + g_xz * t[5]
            + g_yz * t[6] + g_xyz * t[7]) / 2.0   # :92


def sweep_one_direction(cells, positions_d, velocities_d, t, phi, da, f_a1, a_val, direction,
                        n_cells):
    """integrate.py:54-97 for one axis: kick v += da*f_a1*g_p (:94), then drift
    x = (x + da*v/(a+da)^2*f_a1) % N (:95) with the *updated* v.  Returns new arrays."""
    phi = np.asarray(phi, dtype=np.float64)
    g_p = _gather_gradient(cells, np.asarray(t, dtype=np.float64), phi, direction, n_cells)
    v = np.asarray(velocities_d, dtype=np.float64) + da * f_a1 * g_p
    x = np.mod(np.asarray(positions_d, dtype=np.float64) + da * v / (a_val + da) ** 2 * f_a1,
               n_cells)
    return x, v


def integrate(positions, velocities, a_val, f_a1, da, phi, n_cells, ref_dtypes=False):
    """integrate.py:15-25: cells and weights from pre-step positions (Q6), then three sweeps.

    Returns new (positions, velocities) arrays; float64 unless ``ref_dtypes``, in which case
    the reference's float32 stores are mimicked (x, v and t stored as float32).
    """
    x = wrap_positions(np.asarray(positions), n_cells)
    cells = np.floor(x).astype(np.int64) % n_cells
    t = weights(cells, x, ref_dtypes)
    out_x = np.empty(x.shape, np.float32 if ref_dtypes else np.float64)
    out_v = np.empty(x.shape, np.float32 if ref_dtypes else np.float64)
    for d in range(3):
        xd, vd = sweep_one_direction(cells, x[d], np.asarray(velocities)[d], t, phi, da, f_a1,
                                     a_val, d, n_cells)
        if ref_dtypes:
            vd = vd.astype(np.float32)
            # the reference re-reads the float32 velocity for the drift (integrate.py:94-95)
            xd = np.mod(x[d].astype(np.float64) + da * vd.astype(np.float64)
                        / (a_val + da) ** 2 * f_a1, n_cells).astype(np.float32)
        out_x[d], out_v[d] = xd, vd
    return out_x, out_v


def advance_time(rho, positions, velocities, fgrid, a, da, p: OracleParams,
                 ref_dtypes=False, return_phi=False):
    """integrate.py:9-13: phi = potential(rho); fa1 = f(a+da, [H0, OL, OK]) (Q1); integrate."""
    phi = potential(rho, fgrid, a, p.omega_m0, ref_dtypes)
    fa1 = f(a + da, [p.h0, p.omega_lambda0, p.omega_k0])
    x, v = integrate(positions, velocities, a, fa1, da, phi, p.n_cells, ref_dtypes)
    if return_phi:
        return x, v, phi
    return x, v


def step_count(a_init: float, a_end: float, steps: int) -> int:
    """pmesh.py:30,56,63 — how many times ``while a < A_END - da: ...; a += da`` runs (Q7)."""
    da = (a_end - a_init) / steps
    a = a_init
    k = 0
    while a < a_end - da:
        a += da
        k += 1
    return k


def run(positions, velocities, p: OracleParams, n_steps: int | None = None,
        ref_dtypes=False):
    """pmesh.py:54-63 step loop (ICs supplied by the caller).  Returns (x, v, a, steps)."""
    fgrid = fourier_grid(p.n_cells, ref_dtypes)
    da = (p.a_end - p.a_init) / p.steps
    a = p.a_init
    x = np.array(positions, dtype=np.float32 if ref_dtypes else np.float64)
    v = np.array(velocities, dtype=np.float32 if ref_dtypes else np.float64)
    k = 0
    while a < p.a_end - da and (n_steps is None or k < n_steps):
        rho = density(x, p.mass, p.n_cells, ref_dtypes)
        x, v = advance_time(rho, x, v, fgrid, a, da, p, ref_dtypes)
        a += da
        k += 1
    return x, v, a, k


# --------------------------------------------------------------------------------------------
# Initial conditions: shaped Gaussian field and Zel'dovich displacement (IC side of the step)
# --------------------------------------------------------------------------------------------
def _mode_m2(n: int) -> np.ndarray:
    """Integer |m|^2 of every mode of an n^3 DFT (fftfreq order on each axis)."""
    m = np.fft.fftfreq(n, d=1.0 / n)
    return (m[:, None, None] ** 2 + m[None, :, None] ** 2 + m[None, None, :] ** 2).astype(np.int64)


def shaped_field(noise: np.ndarray, amp: np.ndarray) -> np.ndarray:
    """Density contrast on an n^3 lattice from white noise: IDFT(amp[|m|^2] * DFT(noise)).

    gaussian_random_field.py:9-29 multiplies the transformed field by sqrt(P(k)) per mode; the
    amplitude is tabulated per integer |m|^2 (amp[0] multiplies the mean).  ``noise`` is
    (n, n, n) or flat in [i][j][k] order; returns float64 (n, n, n).
    """
    w = np.asarray(noise, dtype=np.float64)
    n = round(w.size ** (1 / 3))
    w = w.reshape(n, n, n)
    return np.fft.ifftn(np.fft.fftn(w) * np.asarray(amp, np.float64)[_mode_m2(n)]).real


def zeldovich_displacement(delta: np.ndarray, n_cells: int, box: float) -> np.ndarray:
    """Psi (3, n^3) in grid cells from the density contrast delta on the n^3 particle lattice.

    zeldovich.py:17-69: phi_k = delta_k / (-k^2) (0 at k = 0, :35-38), displacement along
    direction d = IDFT(-i k_d phi_k) (:56-69 and the backward transform of :45-54) times the
    force resolution N_CELLS/BOX (:46,53), with k = 2 pi fftfreq(n) n / BOX on array axis d
    (:60-67), so positions[d] moves along axis d of delta (Q8).  The Nyquist plane of axis d
    is zeroed: i k_d at k_d = -pi has no real-valued counterpart, and FFTW's real part of it
    depends on the transform's layout.  Psi satisfies div Psi = -delta.
    """
    d = np.asarray(delta, dtype=np.float64)
    n = d.shape[0]
    dk = np.fft.fftn(d)
    k = 2 * np.pi * np.fft.fftfreq(n, d=1.0 / n) / box
    kd = (k[:, None, None], k[None, :, None], k[None, None, :])
    k2 = kd[0] ** 2 + kd[1] ** 2 + kd[2] ** 2
    phi = np.zeros_like(dk)
    np.divide(dk, -k2, out=phi, where=k2 != 0)
    out = np.empty((3, n ** 3))
    for a in range(3):
        comp = -1j * np.broadcast_to(kd[a], dk.shape) * phi
        if n % 2 == 0:
            sl = [slice(None)] * 3
            sl[a] = n // 2
            comp[tuple(sl)] = 0.0
        out[a] = np.fft.ifftn(comp).real.reshape(-1) * (n_cells / box)
    return out


def zeldovich_positions(psi: np.ndarray, n_cells: int) -> np.ndarray:
    """(q + Psi) mod N_CELLS on the lattice of zeldovich.py:79-87, float64, wrapped into
    [0, N).  The reference's unseeded +-2 cell jitter (zeldovich.py:89-91, Q10) is left out."""
    n_parts = round(psi.shape[1] ** (1 / 3))
    x = np.mod(lattice_q(n_parts, n_cells, np.float64) + psi, n_cells)
    x[x >= n_cells] = 0.0
    return x


# --------------------------------------------------------------------------------------------
# Matter power spectrum of a deposited grid (used for full-run comparisons)
# --------------------------------------------------------------------------------------------
def power_spectrum(rho: np.ndarray, box: float, n_bins: int | None = None):
    """Shell-averaged |delta_k|^2 of a density grid, P(k) = V/N^6 <|delta_k|^2>.

    Bins are unit-width shells in integer mode radius |m| (m = 1 .. N/2); k = 2*pi*|m|/box.
    Returns (k_centers, P, counts).  No CIC window correction.
    """
    n = rho.shape[0]
    delta = np.asarray(rho, dtype=np.float64)
    delta = delta / delta.mean() - 1.0
    dk = np.fft.rfftn(delta)
    m = np.fft.fftfreq(n, d=1.0 / n)
    mz = m[:, None, None]
    my = m[None, :, None]
    mx = np.arange(n // 2 + 1)[None, None, :]
    mr = np.sqrt(mx ** 2 + my ** 2 + mz ** 2)
    # half-spectrum weights: kx = 0 and kx = N/2 planes appear once, others twice
    w = np.full(dk.shape, 2.0)
    w[..., 0] = 1.0
    if n % 2 == 0:
        w[..., -1] = 1.0
    nb = n_bins or n // 2
    ib = np.floor(mr + 0.5).astype(np.int64)
    sel = (ib >= 1) & (ib <= nb)
    p2 = np.abs(dk) ** 2 * w
    psum = np.bincount(ib[sel], weights=p2[sel], minlength=nb + 1)
    ksum = np.bincount(ib[sel], weights=(mr * w)[sel], minlength=nb + 1)
    cnt = np.bincount(ib[sel], weights=w[sel], minlength=nb + 1)
    good = cnt > 0
    kc = np.where(good, ksum / np.maximum(cnt, 1), 0) * 2 * np.pi / box
    pk = np.where(good, psum / np.maximum(cnt, 1), 0) * box ** 3 / n ** 6
    return kc[1:], pk[1:], cnt[1:]


# ------------------------------------------------------------------ seeded white noise
# The reference draws its field's deviates from an unseeded parallel RNG
# (gaussian_random_field.py:32-50), which no run can reproduce; the device generator replaces
# it with a counter-based one, restated here.

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11), vectorised: ctr is (4, m) uint32
    counters, key two uint32 words; returns (4, m) uint32.  Each round maps
    (x0, x1, x2, x3) -> (hi(M1 x2) ^ x1 ^ k0, lo(M1 x2), hi(M0 x0) ^ x3 ^ k1, lo(M0 x0)) and
    the key advances by (W0, W1) between rounds."""
    c = [np.asarray(x, dtype=np.uint64) & _MASK for x in ctr]
    k0, k1 = int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF
    for r in range(10):
        if r:
            k0 = (k0 + _W0) & 0xFFFFFFFF
            k1 = (k1 + _W1) & 0xFFFFFFFF
        p0 = _M0 * c[0]
        p1 = _M1 * c[2]
        c = [((p1 >> np.uint64(32)) ^ c[1] ^ np.uint64(k0)) & _MASK, p1 & _MASK,
             ((p0 >> np.uint64(32)) ^ c[3] ^ np.uint64(k1)) & _MASK, p0 & _MASK]
    return np.stack(c).astype(np.uint32)


def white_noise(n: int, seed: int) -> np.ndarray:
    """n standard normal deviates: element e comes from counter (e // 4, 0, 0, 0), key
    (seed low, seed high); uniforms (u + 1/2) 2^-32 and Box-Muller on the word pairs (0, 1)
    -> (r cos, r sin) and (2, 3) -> (r cos, r sin)."""
    groups = (n + 3) // 4
    g = np.arange(groups, dtype=np.uint64)
    ctr = [g & _MASK, g >> np.uint64(32), np.zeros_like(g), np.zeros_like(g)]
    u = philox4x32_10(ctr, (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF)).astype(np.float64)
    s = 2.0 ** -32
    r0 = np.sqrt(-2.0 * np.log((u[0] + 0.5) * s))
    t0 = 2 * np.pi * ((u[1] + 0.5) * s)
    r1 = np.sqrt(-2.0 * np.log((u[2] + 0.5) * s))
    t1 = 2 * np.pi * ((u[3] + 0.5) * s)
    z = np.stack([r0 * np.cos(t0), r0 * np.sin(t0), r1 * np.cos(t1), r1 * np.sin(t1)], axis=1)
    return z.reshape(-1)[:n]
