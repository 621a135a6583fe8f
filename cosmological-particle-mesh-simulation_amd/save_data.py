"""Snapshots and restart (reference: save_data.py:7-50).

Same datasets and unit conversions as the reference: x1..x3 (Mpc), vx1..vx3 (km/s), a, and
optionally density, in ``Data/data.{step}.hdf5``.  h5py is used when it is installed; without
it the same datasets are written to ``Data/data.{step}.npz``.  Device-resident arrays are copied
to the host here, the only place the step loop does so.
"""
import os

import numpy as np

import configure_me

try:  # optional dependency
    import h5py  # type: ignore
except ImportError:  # pragma: no cover - depends on the environment
    h5py = None


def _units(a):
    c = configure_me
    unit_conv_pos = 7.8 * (c.BOX_SIZE / (c.N_CELLS / 128)) / 10 ** 3           # save_data.py:10
    unit_conv_vel = 0.781 * c.BOX_SIZE * c.H0 / (a * c.N_CELLS / 128)           # save_data.py:11
    return unit_conv_pos, unit_conv_vel


def _path(step, ext):
    return os.path.join(getattr(configure_me, "DATA_DIR", "Data"), f"data.{step}.{ext}")


def save_file(rho, positions, velocities, step, a):
    ext = "hdf5" if h5py is not None else "npz"
    print("Writing to disk: data.{}.{}".format(step, ext))
    up, uv = _units(a)
    pos = np.asarray(positions)
    vel = np.asarray(velocities)
    data = {"x1": pos[0] * up, "x2": pos[1] * up, "x3": pos[2] * up,
            "vx1": vel[0] * uv, "vx2": vel[1] * uv, "vx3": vel[2] * uv, "a": np.float64(a)}
    if configure_me.SAVE_DENSITY:
        data["density"] = np.asarray(rho)
    os.makedirs(getattr(configure_me, "DATA_DIR", "Data"), exist_ok=True)
    if h5py is not None:
        with h5py.File(_path(step, ext), "w") as hf:
            for k, v in data.items():
                hf.create_dataset(k, data=v)
    else:
        np.savez(_path(step, ext), **data)


def from_file(step):
    n_p = configure_me.N_PARTS ** 3
    positions = np.zeros((3, n_p), dtype=np.float32)
    velocities = np.zeros((3, n_p), dtype=np.float32)
    if h5py is not None and os.path.exists(_path(step, "hdf5")):
        with h5py.File(_path(step, "hdf5"), "r") as hf:
            data = {k: np.array(hf.get(k)) for k in hf.keys()}
    else:
        with np.load(_path(step, "npz"), allow_pickle=False) as z:
            data = {k: z[k] for k in z.files}
    a = np.float32(data["a"])                                                    # save_data.py:38
    up, uv = _units(a)
    positions[0] = data["x1"] / up
    positions[1] = data["x2"] / up
    positions[2] = data["x3"] / up
    velocities[0] = data["vx1"] / uv
    velocities[1] = data["vx2"] / uv
    velocities[2] = data["vx3"] / uv
    return positions, velocities, a
