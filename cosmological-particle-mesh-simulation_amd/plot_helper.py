"""Plotting (reference: plot_helper.py) is outside this package's scope.

The functions keep the reference's names so pmesh.py imports unchanged; they raise unless
matplotlib is installed, and then draw a plain slice or projection.
"""
import numpy as np

import configure_me


def _plt():
    try:
        import matplotlib
        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
        return plt
    except ImportError as e:  # pragma: no cover - depends on the environment
        raise RuntimeError("plotting needs matplotlib, which is not installed") from e


def plot_step(rho, savestep):
    plt = _plt()
    c = configure_me
    mass = (c.N_CELLS / c.N_PARTS) ** 3
    fig, ax = plt.subplots()
    ax.imshow(np.asarray(rho)[0], extent=(0, c.BOX_SIZE, 0, c.BOX_SIZE), vmin=0., vmax=mass * 3)
    plt.savefig(f"{c.DATA_DIR}/snapshots_density{savestep}.png", dpi=150, bbox_inches="tight")
    plt.close()


def plot_grf(rho):
    plt = _plt()
    c = configure_me
    fig, ax = plt.subplots()
    ax.imshow(np.asarray(rho)[0], extent=(0, c.BOX_SIZE, 0, c.BOX_SIZE))
    plt.savefig(f"{c.DATA_DIR}/snapshot_grf.png", dpi=150, bbox_inches="tight")
    plt.close()


def plot_projection(rho, savestep, depth):
    plt = _plt()
    c = configure_me
    n_slices = int(c.BOX_SIZE / depth)
    proj = np.asarray(rho)[:n_slices].sum(axis=0, dtype=np.float64)
    fig, ax = plt.subplots()
    ax.imshow(proj / max(n_slices, 1), extent=(0, c.BOX_SIZE, 0, c.BOX_SIZE))
    plt.savefig(f"{c.DATA_DIR}/projection_density{int(savestep)}.png", dpi=150, bbox_inches="tight")
    plt.close()
