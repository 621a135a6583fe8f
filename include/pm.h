/*
 * libpm — particle-mesh (PM) step for AMD Instinct MI355X (gfx950): C ABI.
 *
 * Every entry point returns an int status: 0 (PM_OK) on success, otherwise one of the
 * PM_E_* codes below; pm_last_error() then returns a thread-local message.  No C++
 * exception crosses this boundary.  Pointers named d_* are device pointers (HBM); h_*
 * are host pointers.  Layouts are the reference's NumPy layouts (C order):
 *   positions / velocities : float32 (3, N_p)   row 0 = x (grid axis 2), row 2 = z (axis 0)
 *   grids                  : float32 (N, N, N)  indexed [z, y, x]
 *   spectra                : complex128 (N, N, N) (interleaved re, im doubles)
 *   weights t              : float32 (8, N_p); cell centers int64 (3, N_p)
 * File:line citations refer to the reference's src/ directory.
 */
#ifndef PM_H
#define PM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define PM_API __attribute__((visibility("default")))
#else
#define PM_API
#endif

#define PM_OK 0
#define PM_E_INVALID 1
#define PM_E_HIP 2
#define PM_E_NOMEM 3
#define PM_E_STATE 4
#define PM_E_UNSUPPORTED 5
#define PM_E_CORRUPT 6      /* a step or set_state found non-finite particle values or tile bins
                               inconsistent with the positions; the message gives the counts   */

/* sim flags */
#define PM_FLAG_CIC_DECONVOLVE 1u   /* divide by the CIC window of deposit and gather        */
#define PM_FLAG_FIX_F_COSMOLOGY 2u  /* f() gets [Omega_m, OL, OK] instead of [H0, OL, OK]     */
#define PM_FLAG_NO_GRAPH 4u         /* pm_sim_run enqueues every kernel instead of replaying a
                                       captured graph of consecutive steps                     */
#define PM_FLAG_FUSED_DEPOSIT 8u    /* FP32, N = 256/384/512, 32x8x8 or 64x8x8 tiles: deposit
                                       each row block of 8 x 8 tile rows into LDS and run the R2C
                                       x pass from there (the density grid is never stored);
                                       opt-in: slower than the separate passes at 512^3       */
#define PM_FLAG_DETERMINISTIC 16u   /* deposit with fixed-point LDS sums whose rounding does not
                                       depend on the order of the particles: repeated runs from
                                       one state give bit-identical states (the P(k)
                                       estimator's FP64 shell sums still round in any order)  */
#define PM_FLAG_FORWARD_ORDER 32u   /* every step kernel walks its blocks forward (by default
                                       the R2C and C2R passes walk them backwards, so they and
                                       the kernels after them first read what their producer
                                       wrote last and the last-level cache still holds; the
                                       results are the same)                                   */

/* ---------------------------------------------------------------- library / device memory */
PM_API int pm_init(int device);                 /* select device, create the library stream */
PM_API const char* pm_last_error(void);         /* thread-local message of the last failure */
PM_API int pm_version(void);
PM_API int pm_supported_grid(int n_cells);      /* 1 if N is a supported FFT size            */
PM_API int pm_malloc(void** d_ptr, size_t bytes);
PM_API int pm_free(void* d_ptr);
PM_API int pm_memcpy_htod(void* d_dst, const void* h_src, size_t bytes);
PM_API int pm_memcpy_dtoh(void* h_dst, const void* d_src, size_t bytes);
PM_API int pm_memcpy_dtod(void* d_dst, const void* d_src, size_t bytes);
PM_API int pm_memset(void* d_dst, int value, size_t bytes);
PM_API int pm_synchronize(void);
PM_API int pm_mem_info(size_t* free_bytes, size_t* total_bytes);  /* device HBM, for admission */

/* ------------------------------------------- reference-function mirrors (device pointers) */
/* density.py:7-48  density(positions, mass): CIC deposit into a new float32 N^3 grid.       */
PM_API int pm_density(const float* d_positions, int64_t n_p, int n_cells, double mass,
                      float* d_rho);
/* fourier_utils.py:5-16  fourier_grid(): 1/sum sin^2(k/2), G(0) = 0.                         */
PM_API int pm_fourier_grid(int n_cells, float* d_fgrid);
/* potential.py:17-21  density_k(density): forward c2c DFT of the real grid (complex128).    */
PM_API int pm_density_k(const float* d_rho, int n_cells, void* d_rho_k);
/* potential.py:12-15  potential_k(density_k, fgrid, a) = -3*Omega_m/8/a * fgrid * density_k */
PM_API int pm_potential_k(const void* d_rho_k, const float* d_fgrid, int n_cells, double a,
                          double omega_m0, void* d_phi_k);
/* potential.py:23-29  potential_real(potential_k): Re(IDFT)/N^3 as float32.                 */
PM_API int pm_potential_real(const void* d_phi_k, int n_cells, float* d_phi);
/* potential.py:7-10  potential(density, fgrid, a), fused (complex128 arithmetic).           */
PM_API int pm_potential(const float* d_rho, const float* d_fgrid, int n_cells, double a,
                        double omega_m0, float* d_phi);
/* No counterpart in the reference: pm_potential for callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and result as pm_potential; its complex128 workspace and FFT tables live in a fixed
 * device array of about 17 MiB that the library holds on each device it binds to (found at
 * pm_init), so N_CELLS is limited to 128 (PM_E_UNSUPPORTED above that) and the call allocates
 * nothing.  Calls that overlap on one device share that array and race.                    */
PM_API int pm_potential_static(const float* d_rho, const float* d_fgrid, int n_cells, double a,
                               double omega_m0, float* d_phi);
/* integrate.py:20  cell_centers = floor(positions).astype(int64) % N_CELLS                  */
PM_API int pm_cell_centers(const float* d_positions, int64_t n_p, int n_cells,
                           int64_t* d_cells);
/* integrate.py:27-52  weights(cell_centers, positions) -> float32 (8, N_p)                  */
PM_API int pm_weights(const int64_t* d_cells, const float* d_positions, int64_t n_p,
                      float* d_t);
/* integrate.py:54-97  sweep_one_direction(...): kick + drift of one axis, in place.         */
PM_API int pm_sweep_one_direction(const int64_t* d_cells, float* d_positions_dir,
                                  float* d_velocities_dir, const float* d_t,
                                  const float* d_phi, int64_t n_p, int n_cells, double da,
                                  double f_a1, double a_val, int direction);
/* integrate.py:15-25  integrate(positions, velocities, a, f_a1, da, potentials), in place.  */
PM_API int pm_integrate(float* d_positions, float* d_velocities, int64_t n_p, int n_cells,
                        double a_val, double f_a1, double da, const float* d_phi);
/* integrate.py:9-13  advance_time(density, positions, velocities, fgrid, a, da), in place.  */
PM_API int pm_advance_time(const float* d_rho, float* d_positions, float* d_velocities,
                           const float* d_fgrid, int64_t n_p, int n_cells, double a, double da,
                           double h0, double omega_lambda0, double omega_k0, double omega_m0);
/* cosmology.py:20-27  f(a, [c0, c1, c2]) = 1/sqrt((c0 + c2*a + c1*a^3)/a)  (host scalar)    */
PM_API double pm_f(double a, double c0, double c1, double c2);

/* ------------------------------------- device-resident simulation (fused PM step, pmesh.py) */
typedef struct pm_sim pm_sim;
/* precision: 32 (FP32 state, grids and FFT) or 64 (FP64 throughout).                        */
PM_API int pm_sim_create(int n_cells, int n_parts_side, int precision, unsigned flags,
                         pm_sim** out);
PM_API int pm_sim_destroy(pm_sim* sim);
/* configure_me.py:34-38 constants the step reads                                             */
PM_API int pm_sim_set_cosmology(pm_sim* sim, double omega_m0, double omega_lambda0,
                                double omega_k0, double h0);
/* density.py:8 `mass` argument (default (N_CELLS/N_PARTS)^3, pmesh.py:28)                   */
PM_API int pm_sim_set_mass(pm_sim* sim, double mass);
/* potential.py:15 `fgrid`: use a caller-owned float32 N^3 device grid instead of the
 * analytic Green's function (NULL restores the analytic one).  The grid must outlive use.   */
PM_API int pm_sim_set_green(pm_sim* sim, const float* d_fgrid);
/* positions/velocities in the reference layout (3, N_p), absolute positions in [0, N).
 * dtype_bytes: 4 (float32) or 8 (float64); on_device: 0 host pointers, 1 device pointers.   */
PM_API int pm_sim_set_state(pm_sim* sim, const void* positions, const void* velocities,
                            int dtype_bytes, int on_device);
PM_API int pm_sim_get_state(pm_sim* sim, void* positions, void* velocities, int dtype_bytes,
                            int on_device);
/* density.py:7-48 on the resident state (mass = (N/N_PARTS)^3, pmesh.py:28)                 */
PM_API int pm_sim_deposit(pm_sim* sim);
/* replace the density the next step solves for (integrate.py:9 `density` argument)         */
PM_API int pm_sim_set_density(pm_sim* sim, const void* rho, int dtype_bytes, int on_device);
/* copy the current density / potential grid out (float32 or float64, host or device)       */
PM_API int pm_sim_get_density(pm_sim* sim, void* rho, int dtype_bytes, int on_device);
PM_API int pm_sim_get_potential(pm_sim* sim, void* phi, int dtype_bytes, int on_device);
/* one PM step: pmesh.py:60-61 (density + advance_time) at scale factor a with step da.
 * Returns PM_E_CORRUPT when the step produced non-finite positions or velocities or found its
 * tile bins inconsistent (the kernels count such particles; the reference carries NaN on
 * silently).                                                                                 */
PM_API int pm_sim_step(pm_sim* sim, double a, double da);
/* n_steps steps from a with step da; returns the final a.  Unless PM_FLAG_NO_GRAPH or timing
 * is set, the steps replay one captured graph of consecutive steps whose scalars (Green's
 * coefficient, kick, drift) come from a device table filled for the whole call.             */
PM_API int pm_sim_run(pm_sim* sim, double a, double da, int n_steps, double* a_out);
/* per-stage timing with HIP events on the sim's stream (0 off, 1 on); stage order:
 * deposit, r2c_x, c2c_y_fwd, z_green, c2c_y_inv, c2r_x, gather_kick_drift, bin, deposit_r2c
 * (the fused deposit + R2C x pass, which replaces the first two where it runs)               */
PM_API int pm_sim_set_timing(pm_sim* sim, int enable);
PM_API int pm_sim_get_timing(pm_sim* sim, double* stage_ms, int64_t* stage_calls, int n_stages);
PM_API int pm_sim_sync(pm_sim* sim);
/* bytes of device memory held by the sim                                                     */
PM_API int pm_sim_footprint(int n_cells, int n_parts_side, int precision, size_t* bytes);
/* Eulerian tile shape of the deposit / gather kernels (cells per tile along x, y, z; all 0 =
 * the default 16^3).  Instantiated shapes: 16x16x16, 32x16x8, 32x8x8, 16x16x8, 64x8x8, 8x8x8;
 * each must divide N_CELLS.  No reference counterpart (the reference scatters per particle,
 * density.py:17).                                                                            */
PM_API int pm_sim_set_tiles(pm_sim* sim, int tx, int ty, int tz);
PM_API int pm_sim_get_tiles(pm_sim* sim, int* tx, int* ty, int* tz);
/* Autotuner: times deposit + gather + binning on the sim's own particles for every usable tile
 * shape (reps timed repetitions each, the state is left unchanged) and keeps the fastest;
 * returns the shape and its time per repetition in ms.  It then times each y / z FFT pass that
 * has an alternative pencil shape in both shapes and keeps the faster (see pm_sim_get_pencils). */
PM_API int pm_sim_autotune(pm_sim* sim, int reps, int* tx, int* ty, int* tz, double* best_ms);
/* Pencil shape of the forward y, the z (with the k-space operator) and the inverse y FFT passes:
 * columns per block.  Each pass runs in its default shape or, where it has one (the LDS passes
 * for N_CELLS >= 64), in an alternative with half the columns; 0 = the default.  Results do not
 * depend on the shape.  No reference counterpart (pyFFTW plans its own, potential.py:20,27). */
PM_API int pm_sim_set_pencils(pm_sim* sim, int y_fwd, int z, int y_inv);
PM_API int pm_sim_get_pencils(pm_sim* sim, int* y_fwd, int* z, int* y_inv);

/* ------------------------------------------------ the step's two ends: ICs and the spectrum */
/* gaussian_random_field.py:32-50 draws the field's normal deviates from an unseeded parallel
 * RNG; here: element e = Philox4x32-10(counter e/4, key seed) + Box-Muller, n float64 values.  */
PM_API int pm_white_noise(double* d_out, int64_t n, uint64_t seed);
/* gaussian_random_field.py:9-29 + zeldovich.py:10-100 on the device: the sim's lattice (N_PARTS
 * a supported FFT size below N_CELLS) gets s = Psi and v = vel_factor * Psi with
 * Psi_d = psi_scale * IDFT(i m_d / |m|^2 * h_amp[|m|^2] * W), W the DFT of white noise:
 * d_noise (N_PARTS^3 values in [i][j][k] order, float32 or float64 by noise_dtype_bytes, on
 * the device) or, when NULL, pm_white_noise(seed).  h_amp holds n_amp >= 3 (N_PARTS/2)^2 + 1
 * host doubles sqrt(P(k) N_PARTS^3 / V) D(a) per integer |m|^2.                             */
PM_API int pm_sim_zeldovich(pm_sim* sim, const double* h_amp, int64_t n_amp, double psi_scale,
                            double vel_factor, uint64_t seed, const void* d_noise,
                            int noise_dtype_bytes);
/* matter power spectrum of the sim's density (the analysis the reference leaves to its plots,
 * plot_helper.py): P(k) = V / N^6 <|DFT(rho / <rho> - 1)|^2> over unit shells in |m|
 * (bin floor(|m| + 1/2) = 1 .. n_bins), mean density reduced in FP64; outputs per bin the
 * mean |m|, P and the mode count (host arrays of n_bins doubles).  Overwrites the potential. */
PM_API int pm_sim_power_spectrum(pm_sim* sim, double box, int n_bins, double* h_kmode, double* h_pk,
                                 double* h_counts);

/* ------------------------------------------- batched ensemble (many sims, one launch/stage) */
/* No single reference function corresponds: the reference runs one realisation per process
 * (pmesh.py:54-77).  An ensemble advances many FP32 sims, of grid sizes 32, 48, 64, 96, 128,
 * 192, 256, 384 or 512 that may differ, with one kernel launch per stage of the step for all of
 * them (two for the FFT stages when grids above 256 are mixed with smaller ones); the steps
 * replay a captured graph.  Members stay ordinary sims: they keep their buffers, can be stepped
 * alone between ensemble calls and leave the ensemble when destroyed.                         */
typedef struct pm_ensemble pm_ensemble;
PM_API int pm_ensemble_create(pm_ensemble** out);
PM_API int pm_ensemble_destroy(pm_ensemble* ens);   /* the member sims are left alone          */
/* add a sim (precision 32, a supported size, not in another ensemble); at most 64 members   */
PM_API int pm_ensemble_add(pm_ensemble* ens, pm_sim* sim);
PM_API int pm_ensemble_remove(pm_ensemble* ens, pm_sim* sim);
PM_API int pm_ensemble_size(pm_ensemble* ens, int* n_members);
/* n_steps steps of every member: member i (in the order of adding) starts at a[i] with step
 * da[i] (pmesh.py:56-63 per member); a[i] is advanced in place.  PM_E_CORRUPT names the first
 * member whose state became non-finite ("ensemble member i (after k steps): ..."); the steps
 * done up to then are kept and a[] reflects them.                                           */
PM_API int pm_ensemble_run(pm_ensemble* ens, double* a, const double* da, int n_steps);
/* per-stage timing of the batched launches (same stage order as pm_sim_get_timing); while on,
 * steps are enqueued one by one instead of replaying the graph                               */
PM_API int pm_ensemble_set_timing(pm_ensemble* ens, int enable);
PM_API int pm_ensemble_get_timing(pm_ensemble* ens, double* stage_ms, int64_t* stage_calls,
                                  int n_stages);
PM_API int pm_ensemble_sync(pm_ensemble* ens);

#ifdef __cplusplus
}
#endif
#endif /* PM_H */
