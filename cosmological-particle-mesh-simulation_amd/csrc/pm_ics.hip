// The two ends of the PM step on the device: Zel'dovich initial conditions from a Gaussian
// random field (the work of gaussian_random_field.py + zeldovich.py) and the shell-binned
// matter power spectrum of the density grid.
//
// Initial conditions, for an N_PARTS^3 lattice (n below) of a sim on an N_CELLS grid:
//   1. white noise w on the n^3 lattice, [i][j][k] order (i slowest): Philox4x32-10 counters
//      (Salmon, Moraes, Dror & Shaw, SC'11) and the Box-Muller transform, or a caller's field;
//   2. its half spectrum W(k) (forward R2C, the FFT passes of the Poisson solve);
//   3. per displacement component d: Psi_d(k) = i m_d / |m|^2 * A(|m|^2) * W(k) * scale, with
//      A = sqrt(P(k) n^3 / V) D(a) tabulated on the host per integer |m|^2, the Nyquist plane of
//      axis d and k = 0 zeroed; inverse C2R; then s_d = Psi_d and v_d = c_v Psi_d written to
//      the particles in slot order (lattice axis d = positions[d], zeldovich.py:79-83).
// Lattice axis i is the grid's slowest axis here, so component 0 differentiates along the
// spectrum's z index and component 2 along its half (x) axis.
#include "pm_ics.h"

namespace pm {

namespace {

constexpr int kB = 256;

inline unsigned grid_for(long n, long per_thread = 1) {
  const long t = (n + per_thread - 1) / per_thread;
  const long g = (t + kB - 1) / kB;
  return (unsigned)(g < 1 ? 1 : g);
}

template <typename R>
__global__ __launch_bounds__(kB) void k_white_noise(R* out, long n3, uint64_t seed) {
  const long g = blockIdx.x * (long)kB + threadIdx.x;
  if (4 * g >= n3) return;
  double z[4];
  white4(g, seed, z);
#pragma unroll
  for (int r = 0; r < 4; ++r)
    if (4 * g + r < n3) out[4 * g + r] = R(z[r]);
}

template <typename Tin, typename R>
__global__ __launch_bounds__(kB) void k_copy_field(const Tin* __restrict__ in, R* __restrict__ out, long n3) {
  const long i = blockIdx.x * (long)kB + threadIdx.x;
  if (i < n3) out[i] = R(in[i]);
}

// signed mode number of index i on an n-point axis (numpy.fft.fftfreq order)
__device__ __forceinline__ int signed_mode(int i, int n) { return i < n / 2 ? i : i - n; }

// one displacement component's half spectrum from the noise spectrum W (H part + Nyquist plane)
template <typename R>
__global__ __launch_bounds__(kB) void k_zeldovich_component(const cpx<R>* __restrict__ WH, const cpx<R>* __restrict__ WQ,
                                                           cpx<R>* __restrict__ H, cpx<R>* __restrict__ Q,
                                                           const double* __restrict__ amp, int n, int d,
                                                           double scale) {
  const long half = (long)n * n * (n / 2);
  const long total = half + (long)n * n;
  const long idx = blockIdx.x * (long)kB + threadIdx.x;
  if (idx >= total) return;
  int z, y, kx;
  cpx<R> w;
  if (idx < half) {
    kx = (int)(idx % (n / 2));
    y = (int)((idx / (n / 2)) % n);
    z = (int)(idx / ((long)n * (n / 2)));
    w = WH[idx];
  } else {
    const long j = idx - half;
    kx = n / 2;
    y = (int)(j % n);
    z = (int)(j / n);
    w = WQ[j];
  }
  const int mz = signed_mode(z, n), my = signed_mode(y, n), mx = kx;
  const long m2 = (long)mx * mx + (long)my * my + (long)mz * mz;
  // component d: derivative along lattice axis d = spectrum axis z, y, x for d = 0, 1, 2
  const int md = d == 0 ? mz : (d == 1 ? my : mx);
  const int id = d == 0 ? z : (d == 1 ? y : kx);
  double f = 0.0;
  if (m2 > 0 && id != n / 2) f = scale * amp[m2] * (double)md / (double)m2;
  // i * f * w
  const cpx<R> out = cmake<cpx<R>>(R(-f * (double)w.y), R(f * (double)w.x));
  if (idx < half) H[idx] = out;
  else Q[idx - half] = out;
}

// particles of the lattice take component d of the displacement field (real [i][j][k] grid)
template <typename R>
__global__ __launch_bounds__(kB) void k_zeldovich_particles(const R* __restrict__ psi, R* __restrict__ s, R* __restrict__ v,
                                                           Lattice L, int d, double vel_factor) {
  const long slot = blockIdx.x * (long)kB + threadIdx.x;
  if (slot >= L.np) return;
  int i, j, k;
  slot_to_ijk(slot, L, i, j, k);
  const long p = ((long)i * L.np_side + j) * L.np_side + k;  // reference particle index
  const R x = psi[p];
  s[d * L.np + slot] = wrap_disp<R>(x, R(L.n));
  v[d * L.np + slot] = R(vel_factor * (double)x);
}

// ---------------------------------------------------------------- power spectrum
__global__ __launch_bounds__(kB) void k_sum_f32(const float* __restrict__ x, long n, double* out) {
  __shared__ double sh[kB];
  double acc = 0.0;
  for (long i = blockIdx.x * (long)kB + threadIdx.x; i < n; i += (long)gridDim.x * kB) acc += (double)x[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, sh[0]);
}
__global__ __launch_bounds__(kB) void k_sum_f64(const double* __restrict__ x, long n, double* out) {
  __shared__ double sh[kB];
  double acc = 0.0;
  for (long i = blockIdx.x * (long)kB + threadIdx.x; i < n; i += (long)gridDim.x * kB) acc += x[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, sh[0]);
}

// shell sums over the half spectrum: bins[0..nb] of {sum w|F|^2, sum w|m|, sum w}, bin
// floor(|m| + 0.5); weight 2 for 0 < kx < n/2 (each stands for its conjugate), else 1
template <typename R>
__global__ __launch_bounds__(kB) void k_shell_sums(const cpx<R>* __restrict__ H, const cpx<R>* __restrict__ Q, int n,
                                                  int nb, double* __restrict__ bins) {
  extern __shared__ double lb[];  // 3 * (nb + 1)
  for (int i = threadIdx.x; i < 3 * (nb + 1); i += kB) lb[i] = 0.0;
  __syncthreads();
  const long half = (long)n * n * (n / 2);
  const long total = half + (long)n * n;
  for (long idx = blockIdx.x * (long)kB + threadIdx.x; idx < total; idx += (long)gridDim.x * kB) {
    int z, y, kx;
    cpx<R> f;
    if (idx < half) {
      kx = (int)(idx % (n / 2));
      y = (int)((idx / (n / 2)) % n);
      z = (int)(idx / ((long)n * (n / 2)));
      f = H[idx];
    } else {
      const long j = idx - half;
      kx = n / 2;
      y = (int)(j % n);
      z = (int)(j / n);
      f = Q[j];
    }
    const int mz = signed_mode(z, n), my = signed_mode(y, n);
    const double mr = sqrt((double)kx * kx + (double)my * my + (double)mz * mz);
    const int ib = (int)floor(mr + 0.5);
    if (ib < 1 || ib > nb) continue;
    const double w = (kx == 0 || kx == n / 2) ? 1.0 : 2.0;
    const double p2 = (double)f.x * (double)f.x + (double)f.y * (double)f.y;
    __hip_atomic_fetch_add(&lb[ib], w * p2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&lb[(nb + 1) + ib], w * mr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(&lb[2 * (nb + 1) + ib], w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * (nb + 1); i += kB)
    if (lb[i] != 0.0) atomicAdd(&bins[i], lb[i]);
}

}  // namespace

hipError_t white_noise(void* out, int prec, long n3, uint64_t seed, hipStream_t st) {
  if (n3 <= 0) return hipSuccess;
  if (prec == 8) k_white_noise<double><<<grid_for(n3, 4), kB, 0, st>>>((double*)out, n3, seed);
  else k_white_noise<float><<<grid_for(n3, 4), kB, 0, st>>>((float*)out, n3, seed);
  return hipGetLastError();
}

template <typename R>
hipError_t zeldovich_ics(const void* noise, int noise_bytes, uint64_t seed, const double* d_amp, double scale,
                         double vel_factor, const Lattice& L, const FftTables& t, R* s, R* v, R* wbuf, R* hbuf,
                         cpx<R>* qbuf, hipStream_t st) {
  const int n = L.np_side;
  const long n3 = (long)n * n * n;
  // 1. white noise as a real n^3 grid in hbuf
  if (noise) {
    if (noise_bytes == 8) k_copy_field<double, R><<<grid_for(n3), kB, 0, st>>>((const double*)noise, hbuf, n3);
    else k_copy_field<float, R><<<grid_for(n3), kB, 0, st>>>((const float*)noise, hbuf, n3);
  } else {
    k_white_noise<R><<<grid_for(n3, 4), kB, 0, st>>>(hbuf, n3, seed);
  }
  if (hipError_t e = hipGetLastError()) return e;
  // 2. its half spectrum in wbuf (H part, then the Nyquist plane)
  cpx<R>* WH = reinterpret_cast<cpx<R>*>(wbuf);
  cpx<R>* WQ = WH + (long)n * n * (n / 2);
  if (hipError_t e = fft_r2c_x<R, R>(hbuf, WH, WQ, n, t, st)) return e;
  if (hipError_t e = fft_c2c_y<R>(WH, WQ, n, -1, t, st)) return e;
  if (hipError_t e = fft_c2c_z<R>(WH, WQ, n, -1, t, st)) return e;
  // 3. per component: k-space operator, inverse transform, particles
  cpx<R>* H = reinterpret_cast<cpx<R>*>(hbuf);
  const long total = (long)n * n * (n / 2) + (long)n * n;
  for (int d = 0; d < 3; ++d) {
    k_zeldovich_component<R><<<grid_for(total), kB, 0, st>>>(WH, WQ, H, qbuf, d_amp, n, d, scale);
    if (hipError_t e = hipGetLastError()) return e;
    if (hipError_t e = fft_c2c_z<R>(H, qbuf, n, +1, t, st)) return e;
    if (hipError_t e = fft_c2c_y<R>(H, qbuf, n, +1, t, st)) return e;
    if (hipError_t e = fft_c2r_x<R, R>(H, qbuf, hbuf, n, t, st)) return e;
    k_zeldovich_particles<R><<<grid_for(L.np), kB, 0, st>>>(hbuf, s, v, L, d, vel_factor);
    if (hipError_t e = hipGetLastError()) return e;
  }
  return hipSuccess;
}

template <typename R>
hipError_t grid_sum(const R* x, long n, double* d_out, hipStream_t st) {
  const unsigned g = std::min(grid_for(n), 2048u);
  if (sizeof(R) == 8) k_sum_f64<<<g, kB, 0, st>>>((const double*)x, n, d_out);
  else k_sum_f32<<<g, kB, 0, st>>>((const float*)x, n, d_out);
  return hipGetLastError();
}

template <typename R>
hipError_t shell_sums(const cpx<R>* H, const cpx<R>* Q, int n, int nb, double* d_bins, hipStream_t st) {
  const long total = (long)n * n * (n / 2) + (long)n * n;
  const unsigned g = std::min(grid_for(total, 16), 4096u);
  k_shell_sums<R><<<g, kB, 3 * (nb + 1) * sizeof(double), st>>>(H, Q, n, nb, d_bins);
  return hipGetLastError();
}

#define PM_INST(R)                                                                                                \
  template hipError_t zeldovich_ics<R>(const void*, int, uint64_t, const double*, double, double, const Lattice&, \
                                       const FftTables&, R*, R*, R*, R*, cpx<R>*, hipStream_t);                   \
  template hipError_t grid_sum<R>(const R*, long, double*, hipStream_t);                                          \
  template hipError_t shell_sums<R>(const cpx<R>*, const cpx<R>*, int, int, double*, hipStream_t);
PM_INST(float)
PM_INST(double)
#undef PM_INST

}  // namespace pm
