"""Device-resident PM simulation (the fused step behind pmesh.py's loop).

``PMSim`` owns one simulation's particles and grids in HBM.  Its state is stored as FP32 (or
FP64) displacements from the Lagrangian lattice, in a blocked particle order; positions and
velocities cross the boundary in the reference's layout, float32 (3, N_p) absolute
positions in [0, N_CELLS).

``SimPositions``, ``SimVelocities`` and ``SimDensity`` are views of a sim in the shapes the
reference's functions exchange, so that ``density()`` and ``advance_time()`` can keep every
array on the device; converting a view with ``numpy.asarray`` copies it to the host.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class PMSim:
    def __init__(self, n_cells: int, n_parts: int, precision: int = 32, *,
                 cic_deconvolve: bool = False, fix_f_cosmology: bool = False, graph: bool = True,
                 fused_deposit: bool = False, deterministic: bool = False,
                 forward_order: bool = False,
                 omega_m0: float = 0.31, omega_lambda0: float = 0.69, omega_k0: float = 0.0,
                 h0: float = 0.68):
        _lib.init()
        self.n_cells = int(n_cells)
        self.n_parts = int(n_parts)
        self.precision = int(precision)
        self.n_p = self.n_parts ** 3
        flags = 0
        if cic_deconvolve:
            flags |= _lib.PM_FLAG_CIC_DECONVOLVE
        if fix_f_cosmology:
            flags |= _lib.PM_FLAG_FIX_F_COSMOLOGY
        if not graph:
            flags |= _lib.PM_FLAG_NO_GRAPH
        if fused_deposit:
            flags |= _lib.PM_FLAG_FUSED_DEPOSIT
        if deterministic:
            flags |= _lib.PM_FLAG_DETERMINISTIC
        if forward_order:
            flags |= _lib.PM_FLAG_FORWARD_ORDER
        h = C.c_void_p()
        _lib.call("pm_sim_create", self.n_cells, self.n_parts, self.precision, flags, C.byref(h))
        self._h = h
        self.set_cosmology(omega_m0, omega_lambda0, omega_k0, h0)
        self._green_ref = None
        self.version = 0  # bumped whenever the particle state changes

    @staticmethod
    def footprint(n_cells: int, n_parts: int, precision: int = 32) -> int:
        b = C.c_size_t()
        _lib.call("pm_sim_footprint", n_cells, n_parts, precision, C.byref(b))
        return int(b.value)

    # -- configuration ----------------------------------------------------------------------
    def set_cosmology(self, omega_m0, omega_lambda0, omega_k0, h0):
        self.cosmology = (float(omega_m0), float(omega_lambda0), float(omega_k0), float(h0))
        _lib.call("pm_sim_set_cosmology", self._h, *self.cosmology)

    def set_mass(self, mass: float):
        _lib.call("pm_sim_set_mass", self._h, float(mass))

    def set_green(self, fgrid_device):
        """Use a float32 N^3 device grid (a DeviceArray) as G(k); None restores the analytic one."""
        self._green_ref = fgrid_device
        _lib.call("pm_sim_set_green", self._h, None if fgrid_device is None else fgrid_device.ptr)

    # -- tile shape -------------------------------------------------------------------------
    TILE_SHAPES = ((16, 16, 16), (32, 16, 8), (32, 8, 8), (16, 16, 8), (64, 8, 8), (8, 8, 8))

    def set_tiles(self, tx: int = 0, ty: int = 0, tz: int = 0):
        """Tile shape of the deposit / gather kernels (0, 0, 0: the default)."""
        _lib.call("pm_sim_set_tiles", self._h, int(tx), int(ty), int(tz))

    def tiles(self) -> tuple[int, int, int]:
        t = [C.c_int() for _ in range(3)]
        _lib.call("pm_sim_get_tiles", self._h, *[C.byref(x) for x in t])
        return tuple(x.value for x in t)

    def set_pencils(self, y_fwd: int = 0, z: int = 0, y_inv: int = 0):
        """Columns per block of the forward y, z and inverse y FFT passes (0: the default)."""
        _lib.call("pm_sim_set_pencils", self._h, int(y_fwd), int(z), int(y_inv))

    def pencils(self) -> tuple[int, int, int]:
        t = [C.c_int() for _ in range(3)]
        _lib.call("pm_sim_get_pencils", self._h, *[C.byref(x) for x in t])
        return tuple(x.value for x in t)

    def autotune(self, reps: int = 3):
        """Pick the fastest tile shape on the current particles, then each y / z FFT pass's
        faster pencil shape; returns (tile shape, ms per rep of the particle stages)."""
        t = [C.c_int() for _ in range(3)]
        ms = C.c_double()
        _lib.call("pm_sim_autotune", self._h, int(reps), *[C.byref(x) for x in t], C.byref(ms))
        return tuple(x.value for x in t), ms.value

    # -- state ------------------------------------------------------------------------------
    def set_state(self, positions, velocities):
        pos = np.ascontiguousarray(positions)
        vel = np.ascontiguousarray(velocities)
        if pos.shape != (3, self.n_p) or vel.shape != (3, self.n_p):
            raise ValueError(f"positions/velocities must have shape (3, {self.n_p})")
        if pos.dtype not in (np.float32, np.float64):
            pos = pos.astype(np.float32)
        vel = vel.astype(pos.dtype, copy=False)
        _lib.call("pm_sim_set_state", self._h, pos.ctypes.data, vel.ctypes.data,
                  pos.dtype.itemsize, 0)
        self.version += 1

    def get_state(self, dtype=np.float32):
        dt = np.dtype(dtype)
        pos = np.empty((3, self.n_p), dt)
        vel = np.empty((3, self.n_p), dt)
        _lib.call("pm_sim_get_state", self._h, pos.ctypes.data, vel.ctypes.data, dt.itemsize, 0)
        return pos, vel

    def deposit(self):
        _lib.call("pm_sim_deposit", self._h)

    def zeldovich(self, amp, psi_scale: float, vel_factor: float, seed: int = 0, noise=None):
        """Device initial conditions: s = Psi, v = vel_factor * Psi (see pm_sim_zeldovich).

        ``amp[m2]`` = sqrt(P(k) N_PARTS^3 / V) D(a) per integer |m|^2; ``noise`` is an optional
        DeviceArray of N_PARTS^3 float32/float64 normal deviates in [i][j][k] order (otherwise
        the Philox4x32-10 field of ``seed``)."""
        amp = np.ascontiguousarray(amp, dtype=np.float64)
        ptr, nb = None, 0
        if noise is not None:
            if noise.dtype not in (np.float32, np.float64) or noise.nbytes != self.n_p * noise.dtype.itemsize:
                raise ValueError(f"noise must hold {self.n_p} float32/float64 values")
            ptr, nb = noise.ptr, noise.dtype.itemsize
        _lib.call("pm_sim_zeldovich", self._h, amp.ctypes.data_as(C.POINTER(C.c_double)), amp.size,
                  float(psi_scale), float(vel_factor), int(seed) & (2 ** 64 - 1), ptr, nb)
        self.version += 1

    def power_spectrum(self, box: float, n_bins: int | None = None):
        """(k in h/Mpc, P(k) in (Mpc/h)^3, mode counts) of the current density, unit shells in
        |m| up to N/2 by default; computed on the device (FP64 mean and shell sums)."""
        nb = int(n_bins or self.n_cells // 2)
        k = np.zeros(nb)
        pk = np.zeros(nb)
        cnt = np.zeros(nb)
        dp = C.POINTER(C.c_double)
        _lib.call("pm_sim_power_spectrum", self._h, float(box), nb, k.ctypes.data_as(dp),
                  pk.ctypes.data_as(dp), cnt.ctypes.data_as(dp))
        return k * 2 * np.pi / box, pk, cnt

    def set_density(self, rho):
        rho = np.ascontiguousarray(rho)
        n = self.n_cells
        if rho.shape != (n, n, n):
            raise ValueError(f"density must have shape {(n, n, n)}")
        if rho.dtype not in (np.float32, np.float64):
            rho = rho.astype(np.float32)
        _lib.call("pm_sim_set_density", self._h, rho.ctypes.data, rho.dtype.itemsize, 0)

    def density(self, dtype=np.float32) -> np.ndarray:
        dt = np.dtype(dtype)
        n = self.n_cells
        out = np.empty((n, n, n), dt)
        _lib.call("pm_sim_get_density", self._h, out.ctypes.data, dt.itemsize, 0)
        return out

    def density_device(self):
        """A float32 device copy of the current density grid (a DeviceArray of its own)."""
        from .device import DeviceArray
        n = self.n_cells
        out = DeviceArray((n, n, n), np.float32)
        _lib.call("pm_sim_get_density", self._h, out.ptr, 4, 1)
        return out

    def potential(self, dtype=np.float32) -> np.ndarray:
        dt = np.dtype(dtype)
        n = self.n_cells
        out = np.empty((n, n, n), dt)
        _lib.call("pm_sim_get_potential", self._h, out.ctypes.data, dt.itemsize, 0)
        return out

    # -- stepping ---------------------------------------------------------------------------
    def step(self, a: float, da: float):
        _lib.call("pm_sim_step", self._h, float(a), float(da))
        self.version += 1

    def run(self, a: float, da: float, n_steps: int) -> float:
        out = C.c_double()
        _lib.call("pm_sim_run", self._h, float(a), float(da), int(n_steps), C.byref(out))
        self.version += 1
        return out.value

    def sync(self):
        _lib.call("pm_sim_sync", self._h)

    def set_timing(self, enable: bool):
        _lib.call("pm_sim_set_timing", self._h, 1 if enable else 0)

    def timing(self) -> dict:
        k = len(_lib.STAGES)
        ms = (C.c_double * k)()
        calls = (C.c_int64 * k)()
        _lib.call("pm_sim_get_timing", self._h, ms, calls, k)
        return {name: (ms[i], calls[i]) for i, name in enumerate(_lib.STAGES)}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.load().pm_sim_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reference-shaped views -------------------------------------------------------------
    def views(self):
        return SimPositions(self), SimVelocities(self)


class _SimView:
    """A (3, N_p) float32 array view of a sim's particles; copies to the host on demand."""
    kind = ""

    def __init__(self, sim: PMSim):
        self.sim = sim
        self.shape = (3, sim.n_p)
        self.dtype = np.dtype(np.float32)
        self.ndim = 2

    def _host(self):
        pos, vel = self.sim.get_state(np.float32)
        return pos if self.kind == "pos" else vel

    def __array__(self, dtype=None, copy=None):
        a = self._host()
        return a if dtype is None else a.astype(dtype)

    def __getitem__(self, idx):
        return self._host()[idx]

    def __len__(self):
        return 3

    def __repr__(self):
        return f"{type(self).__name__}(n_p={self.sim.n_p}, n_cells={self.sim.n_cells})"


class SimPositions(_SimView):
    kind = "pos"


class SimVelocities(_SimView):
    kind = "vel"


class SimDensity:
    """The density a sim holds at one version of its particles, as returned by density().

    It reads the sim's density buffer on demand while the sim is still at that version.  The
    next step overwrites that buffer, so a caller that will read the density after stepping
    (a snapshot or plot of the step's density) calls ``keep()`` first, which takes a device
    copy; reading a density that was neither kept nor current raises.
    """

    def __init__(self, sim: PMSim, version: int):
        self.sim = sim
        self.version = version
        self._kept = None
        n = sim.n_cells
        self.shape = (n, n, n)
        self.dtype = np.dtype(np.float32)
        self.ndim = 3

    @property
    def current(self) -> bool:
        return self.sim.version == self.version

    def keep(self) -> None:
        if self._kept is None:
            if not self.current:
                raise RuntimeError("this density belongs to an earlier state of the simulation")
            self._kept = self.sim.density_device()

    def __array__(self, dtype=None, copy=None):
        if self._kept is not None:
            a = self._kept.numpy()
        elif self.current:
            a = self.sim.density(np.float32)
        else:
            raise RuntimeError("the simulation has stepped since this density was taken and it "
                               "was not kept (SimDensity.keep) before the step")
        return a if dtype is None else a.astype(dtype)

    def __getitem__(self, idx):
        return np.asarray(self)[idx]
