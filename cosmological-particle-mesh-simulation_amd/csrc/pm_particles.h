// Particle-side kernels of the PM step: state conversion, CIC deposit, fused gather+kick+drift,
// and the reference-layout kernels behind the stateless C-ABI entry points.
#pragma once
#include "pm_common.h"

namespace pm {

// Device storage order of the particles of one simulation.
//
// The reference orders particles by their Lagrangian lattice site, p = (i*NP + j)*NP + k with
// i along positions[0] (zeldovich.py:79-83), so consecutive particles step along grid axis 0,
// the slowest one in memory.  On the device the same lattice is stored in cubic blocks of
// B^3 sites (B = 4: one 64-lane wavefront per block), blocks and sites both ordered with i
// (grid x) fastest.  The slot <-> (i, j, k) map is arithmetic, so no id array is stored and
// every particle's lattice point q is known from its slot.
struct Lattice {
  int n;        // grid cells per side (N_CELLS)
  int np_side;  // particles per side (N_PARTS)
  int blk;      // block edge B (4, or 1 when N_PARTS % 4 != 0)
  int nb;       // blocks per side
  long np;      // N_PARTS^3
  double r;     // lattice spacing N_CELLS / N_PARTS
};

__host__ __device__ __forceinline__ void slot_to_ijk(long slot, const Lattice& L, int& i, int& j, int& k) {
  const int B = L.blk;
  const int B3 = B * B * B;
  const long blk = slot / B3;
  const int loc = (int)(slot - blk * B3);
  const int li = loc % B, lj = (loc / B) % B, lk = loc / (B * B);
  const long bi = blk % L.nb;
  const long t = blk / L.nb;
  const long bj = t % L.nb, bk = t / L.nb;
  i = (int)(bi * B + li);
  j = (int)(bj * B + lj);
  k = (int)(bk * B + lk);
}

// Stage timers recorded by the sim (indices into the timing arrays).
enum Stage { ST_DEPOSIT = 0, ST_R2C, ST_Y_FWD, ST_Z_GREEN, ST_Y_INV, ST_C2R, ST_GKD, ST_COUNT };

// ---- resident-state kernels (slot order, displacement storage) -------------------------------
template <typename R, typename Tin>
hipError_t sim_set_state(const Tin* pos, const Tin* vel, R* s, R* v, const Lattice& L, hipStream_t st);
template <typename R, typename Tout>
hipError_t sim_get_state(const R* s, const R* v, Tout* pos, Tout* vel, const Lattice& L, hipStream_t st);
template <typename R>
hipError_t sim_deposit(const R* s, R* rho, double mass, const Lattice& L, hipStream_t st);
template <typename R>
hipError_t sim_gather_kick_drift(R* s, R* v, const R* phi, double kick, double drift, const Lattice& L,
                                 hipStream_t st);

// ---- reference-layout kernels (absolute float32 positions, reference order) ------------------
hipError_t ref_density(const float* pos, long np, int n, double mass, float* rho, hipStream_t st);
hipError_t ref_cell_centers(const float* pos, long np, int n, int64_t* cells, hipStream_t st);
hipError_t ref_weights(const int64_t* cells, const float* pos, long np, float* t, hipStream_t st);
hipError_t ref_sweep(const int64_t* cells, float* xd, float* vd, const float* t, const float* phi, long np,
                     int n, double da, double fa1, double a, int dir, hipStream_t st);
hipError_t ref_integrate(float* pos, float* vel, long np, int n, double a, double fa1, double da,
                         const float* phi, hipStream_t st);
hipError_t ref_fourier_grid(int n, float* g, hipStream_t st);

// ---- spectrum helpers for the complex128 reference path --------------------------------------
// H[z][y][0..N/2) + Q[z][y] (kx = N/2)  ->  full[z][y][x] complex128 using Hermitian symmetry
hipError_t spec_expand_full(const double2* H, const double2* Q, double2* full, int n, hipStream_t st);
// out = c * fgrid * in   (elementwise over N^3, c = -3*Omega_m/8/a)
hipError_t spec_potential_k(const double2* in, const float* fgrid, double c, double2* out, int n,
                            hipStream_t st);
// full complex spectrum -> Hermitian part on the half spectrum, times `scale`
hipError_t spec_hermitian_half(const double2* full, double2* H, double2* Q, int n, double scale,
                               hipStream_t st);
// out[i] = (Tout) in[i]
template <typename Tin, typename Tout>
hipError_t convert_grid(const Tin* in, Tout* out, size_t count, hipStream_t st);

}  // namespace pm
