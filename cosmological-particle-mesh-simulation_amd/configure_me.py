#Defining Simulation properties
#
# Same names and defaults as the reference's configure_me.py:1-40.  Unlike the reference, the
# modules of this package read these values when they are called (configure_me.NAME), not
# when they are imported, so a value changed at run time takes effect on the next call.
# Settings below the reference's are new and optional.

###################################################
# General simulation settings
###################################################

N_PARTS          = 256 # Number of particles in one dimension
N_CELLS          = 512 # Number of cells in one dimension
BOX_SIZE         = 100 # Box size in Mpc/h in one dimension

N_CPU            = 16 # Number of cpu's used (kept for compatibility; the GPU path ignores it)
RANDOM_SEED      = 38 # random seed to use for the Gaussian Random Field.

STEPS            = 1000 # Number of timesteps
N_SAVE_FILES     = 100  # Number of save files/plotsteps/projections
N_PLOTS          = 100  # Number of plots saved to disk

PLOT_STEPS       = False
PLOT_PROJECTIONS = False
PLOT_GRF         = False
SAVE_DATA        = True
SAVE_DENSITY     = False
PRINT_STATUS     = True

RESTART          = False #If True, restart from savefile number RESTART_FROM_N
RESTART_FROM_N   = 0

###################################################
# Cosmology settings
###################################################
POWER                  = 1.00 # Harrison-Zeldovich spectrum has n~1 (used as n_s)
LCDM_TRANSFER_FUNCTION = True # Specifies whether to use a LCDM transfer function for the GRF

OMEGA_M0               = 0.31 # Mass density
OMEGA_B0               = 0.04 # Baryon density
OMEGA_K0               = 0.00 # Curvature density
OMEGA_LAMBDA0          = 0.69 # Dark energy density
H0                     = 0.68 # Hubble
A_INIT                 = 0.01 # Initial scale factor
A_END                  = 1.00 # Final scale factor

###################################################
# Settings of this implementation (no counterpart in the reference)
###################################################
PRECISION        = 32     # 32: FP32 particles, grids and FFT; 64: FP64 throughout
RESIDENT         = True   # keep particles and grids on the GPU between steps
SIGMA8           = 0.8    # amplitude of the initial power spectrum (top hat, 8 Mpc/h, a=1)
CIC_DECONVOLVE   = False  # divide the Green's function by the CIC window (the reference has none)
FIX_F_COSMOLOGY  = False  # pass [OMEGA_M0, OMEGA_LAMBDA0, OMEGA_K0] to f() in advance_time;
                          # the reference passes [H0, OMEGA_LAMBDA0, OMEGA_K0] (integrate.py:12)
DATA_DIR         = 'Data' # snapshot directory (save_data.py:13)
