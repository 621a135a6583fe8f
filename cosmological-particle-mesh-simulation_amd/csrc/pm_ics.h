// Device initial conditions (Gaussian random field + Zel'dovich) and the power spectrum.
#pragma once
#include "pm_particles.h"

namespace pm {

struct U4 {
  uint32_t x, y, z, w;
};

// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror & Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies, the key
// advanced by the Weyl increments between rounds.
__host__ __device__ inline U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;  // round multipliers
  constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;  // key increments
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += W0;
      k1 += W1;
    }
    const uint64_t p0 = (uint64_t)M0 * c.x, p1 = (uint64_t)M1 * c.z;
    const U4 n{(uint32_t)(p1 >> 32) ^ c.y ^ k0, (uint32_t)p1, (uint32_t)(p0 >> 32) ^ c.w ^ k1, (uint32_t)p0};
    c = n;
  }
  return c;
}

// Four standard normal deviates for elements 4g .. 4g+3 of a noise field: counter (g, 0),
// key = seed; uniforms (u + 1/2) 2^-32 in (0, 1), Box-Muller on the pairs (x, y) and (z, w).
__host__ __device__ inline void white4(long g, uint64_t seed, double z[4]) {
  const U4 u = philox4x32_10(U4{(uint32_t)g, (uint32_t)((uint64_t)g >> 32), 0u, 0u}, (uint32_t)seed,
                             (uint32_t)(seed >> 32));
  const double s = 2.3283064365386963e-10;  // 2^-32
  const double two_pi = 6.283185307179586;
  const double r0 = sqrt(-2.0 * log(((double)u.x + 0.5) * s)), t0 = two_pi * (((double)u.y + 0.5) * s);
  const double r1 = sqrt(-2.0 * log(((double)u.z + 0.5) * s)), t1 = two_pi * (((double)u.w + 0.5) * s);
  z[0] = r0 * cos(t0);
  z[1] = r0 * sin(t0);
  z[2] = r1 * cos(t1);
  z[3] = r1 * sin(t1);
}

// n3 normal deviates (element e from white4(e / 4)) into out (float or double per prec)
hipError_t white_noise(void* out, int prec, long n3, uint64_t seed, hipStream_t st);

// Zel'dovich displacements and velocities of a lattice (L.np_side must be a supported FFT
// size): noise = caller field (noise_bytes 4/8, device pointer) or NULL for white_noise(seed);
// d_amp[m2] = sqrt(P(k) n^3 / V) D(a) for m2 = 0 .. 3 (n/2)^2; Psi = scale * IDFT(i m / m^2 A W),
// scale = (N_CELLS / box) / (k_f n^3); s = Psi, v = vel_factor * Psi.  Scratch: wbuf holds
// n^3 + 2 n^2 reals, hbuf n^3 reals, qbuf n^2 complex.
template <typename R>
hipError_t zeldovich_ics(const void* noise, int noise_bytes, uint64_t seed, const double* d_amp, double scale,
                         double vel_factor, const Lattice& L, const FftTables& t, R* s, R* v, R* wbuf, R* hbuf,
                         cpx<R>* qbuf, hipStream_t st);

// *d_out += sum of x[0 .. n) in FP64
template <typename R>
hipError_t grid_sum(const R* x, long n, double* d_out, hipStream_t st);

// d_bins[0 .. 3 (nb + 1)) += shell sums {w |F|^2, w |m|, w} by bin floor(|m| + 1/2) over the
// half spectrum H + Q of an n^3 grid (w = 2 for 0 < kx < n/2, else 1)
template <typename R>
hipError_t shell_sums(const cpx<R>* H, const cpx<R>* Q, int n, int nb, double* d_bins, hipStream_t st);

}  // namespace pm
