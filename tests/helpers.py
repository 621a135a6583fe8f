"""Shared helpers for the parity tests."""
import numpy as np


def rel(a, b):
    """norm-wise relative error ||a - b|| / ||b|| (fields that cross zero)."""
    a = np.asarray(a, dtype=np.float64) if not np.iscomplexobj(a) else np.asarray(a, np.complex128)
    b = np.asarray(b, dtype=np.float64) if not np.iscomplexobj(b) else np.asarray(b, np.complex128)
    nb = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0)


def periodic_diff(x, y, n):
    d = np.asarray(x, np.float64) - np.asarray(y, np.float64)
    return (d + n / 2) % n - n / 2


def ic_params(n_cells, n_parts, box=100.0, seed=1, a_init=0.02, om=0.3, ol=0.7, h=0.7,
              sigma8=0.8, ns=0.96):
    return dict(n_parts=n_parts, n_cells=n_cells, box=box, seed=seed, om=om, ob=0.045,
                ol=ol, ok=0.0, h=h, ns=ns, sigma8=sigma8, a_init=a_init)


def make_ics(n_cells, n_parts, **kw):
    """Seeded Zel'dovich initial conditions, float32 (3, N_p) positions and velocities."""
    import initial_conditions as ic
    p = ic_params(n_cells, n_parts, **kw)
    delta = ic.gaussian_random_field(p)
    return ic.zeldovich_numpy(delta, p)


def uniform_positions(n_cells, n_p, seed=0):
    x = np.random.default_rng(seed).uniform(0, n_cells, (3, n_p)).astype(np.float32)
    x[x >= n_cells] = 0.0
    return x
