"""Run configuration: module-level constants read by every module of the package.

The names and default values are those of the reference's configure_me.py:1-40, so a
configuration written for the reference works here unchanged.  The package reads them at call
time (``configure_me.NAME``) rather than at import, so changing one at run time affects the
next call.  The block at the end holds settings that only this implementation has.
"""

# ---------------------------------------------------------------- resolution and box
N_PARTS          = 256   # particles per box side (N_PARTS^3 in all)
N_CELLS          = 512   # mesh cells per box side
BOX_SIZE         = 100   # comoving box side in Mpc/h

N_CPU            = 16    # worker threads of the reference's CPU code; unused on the GPU
RANDOM_SEED      = 38    # seed of the initial density field

# ---------------------------------------------------------------- time stepping and output
STEPS            = 1000  # steps from A_INIT to A_END (uniform in the scale factor)
N_SAVE_FILES     = 100   # snapshots over the run (they also set the projection cadence)
N_PLOTS          = 100   # density images over the run

PLOT_STEPS       = False  # draw a density slice at every plot mark
PLOT_PROJECTIONS = False  # draw a projected density at every plot mark
PLOT_GRF         = False  # draw the initial density field
SAVE_DATA        = True   # write snapshots
SAVE_DENSITY     = False  # include the density grid in snapshots
PRINT_STATUS     = True   # one progress line per step

RESTART          = False  # resume from snapshot RESTART_FROM_N instead of new initial conditions
RESTART_FROM_N   = 0

# ---------------------------------------------------------------- cosmology
POWER                  = 1.00  # spectral index of the primordial spectrum (n_s)
LCDM_TRANSFER_FUNCTION = True  # shape the initial spectrum with a matter transfer function

OMEGA_M0               = 0.31  # matter density parameter today
OMEGA_B0               = 0.04  # baryon density parameter today
OMEGA_K0               = 0.00  # curvature density parameter today
OMEGA_LAMBDA0          = 0.69  # cosmological-constant density parameter today
H0                     = 0.68  # Hubble constant in units of 100 km/s/Mpc
A_INIT                 = 0.01  # scale factor of the initial conditions
A_END                  = 1.00  # scale factor at which the run stops

# ---------------------------------------------------------------- this implementation only
PRECISION        = 32     # 32: FP32 particles, grids and FFT; 64: FP64 throughout
RESIDENT         = True   # keep particles and grids on the GPU between steps
SIGMA8           = 0.8    # amplitude of the initial power spectrum (top hat, 8 Mpc/h, a=1)
CIC_DECONVOLVE   = False  # divide the Green's function by the CIC window (the reference has none)
FIX_F_COSMOLOGY  = False  # pass [OMEGA_M0, OMEGA_LAMBDA0, OMEGA_K0] to f() in advance_time;
                          # the reference passes [H0, OMEGA_LAMBDA0, OMEGA_K0] (integrate.py:12)
DATA_DIR         = 'Data' # snapshot directory (save_data.py:13)
