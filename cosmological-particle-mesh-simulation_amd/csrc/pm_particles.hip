// Particle kernels of the PM step (gfx950).
#include "pm_particles.h"

namespace pm {

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(long n, int block = kBlock) {
  long g = (n + block - 1) / block;
  return (unsigned)(g < 1 ? 1 : g);
}



// ------------------------------------------------------------------------------------------
// resident state <-> reference layout
// ------------------------------------------------------------------------------------------
template <typename R, typename Tin>
__global__ __launch_bounds__(kBlock) void k_set_state(const Tin* __restrict__ pos, const Tin* __restrict__ vel,
                                                      R* __restrict__ s, R* __restrict__ v, uint32_t* __restrict__ id,
                                                      Lattice L, uint32_t* __restrict__ health) {
  const long slot = blockIdx.x * (long)kBlock + threadIdx.x;
  if (slot >= L.np) return;
  id[slot] = (uint32_t)slot;
  int ijk[3];
  slot_to_ijk(slot, L, ijk[0], ijk[1], ijk[2]);
  const long p = ((long)ijk[0] * L.np_side + ijk[1]) * L.np_side + ijk[2];
  const double n = L.n, h = 0.5 * L.n;
  bool finite = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double x = (double)pos[d * L.np + p];
    const double vd = (double)vel[d * L.np + p];
    finite = finite && __builtin_isfinite(x) && __builtin_isfinite(vd);
    double sd = x - ((double)ijk[d] * L.r + 0.5);
    if (sd >= h) sd -= n;
    else if (sd < -h) sd += n;
    s[d * L.np + slot] = R(sd);
    v[d * L.np + slot] = R(vd);
  }
  if (!finite && health) atomicAdd(&health[PM_HEALTH_NONFINITE], 1u);
}

// storage index e holds particle id[e]; ids outside the lattice are skipped
template <typename R, typename Tout>
__global__ __launch_bounds__(kBlock) void k_get_state(const R* __restrict__ s, const R* __restrict__ v,
                                                      const uint32_t* __restrict__ id, Tout* __restrict__ pos,
                                                      Tout* __restrict__ vel, Lattice L) {
  const long e = blockIdx.x * (long)kBlock + threadIdx.x;
  if (e >= L.np) return;
  const uint32_t slot = id[e];
  if (slot >= (uint64_t)L.np) return;
  int ijk[3];
  slot_to_ijk(slot, L, ijk[0], ijk[1], ijk[2]);
  const long p = ((long)ijk[0] * L.np_side + ijk[1]) * L.np_side + ijk[2];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const R x = canonical_pos<R>(lattice_coord<R>(ijk[d], L), s[d * L.np + e], R(L.n));
    Tout xo = Tout(x);
    if (xo >= Tout(L.n)) xo = Tout(0);
    pos[d * L.np + p] = xo;
    vel[d * L.np + p] = Tout(v[d * L.np + e]);
  }
}

__global__ __launch_bounds__(kBlock) void k_ids_identity(uint32_t* __restrict__ id, long np) {
  const long e = blockIdx.x * (long)kBlock + threadIdx.x;
  if (e < np) id[e] = (uint32_t)e;
}

// ------------------------------------------------------------------------------------------
// reference-layout kernels (absolute float32 positions in reference particle order)
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float ref_wrap(float x, int n) { return wrap_pos<float>(x, float(n)); }

__global__ __launch_bounds__(kBlock) void k_ref_density(const float* __restrict__ pos, long np, int n, double mass,
                                                        float* __restrict__ rho) {
  const long p = blockIdx.x * (long)kBlock + threadIdx.x;
  if (p >= np) return;
  int c[3];
  double d[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float x = ref_wrap(pos[a * np + p], n);             // Q4: x == N maps to 0
    const int ci = cell_of<float>(x, n);                       // density.py:19-21
    c[a] = ci;
    d[a] = (double)x - (double)ci;                             // density.py:24-26
  }
  const double tx = 1 - d[0], ty = 1 - d[1], tz = 1 - d[2];    // density.py:28-30
  const int X = (c[0] + 1) % n, Y = (c[1] + 1) % n, Z = (c[2] + 1) % n;  // density.py:33-35
  const size_t nn = (size_t)n;
  auto at = [&](int z, int y, int x) { return rho + ((size_t)z * nn + y) * nn + x; };
  atomicAdd(at(c[2], c[1], c[0]), float(mass * tx * ty * tz));   // density.py:37
  atomicAdd(at(c[2], c[1], X), float(mass * d[0] * ty * tz));     // :39
  atomicAdd(at(c[2], Y, c[0]), float(mass * tx * d[1] * tz));     // :40
  atomicAdd(at(Z, c[1], c[0]), float(mass * tx * ty * d[2]));     // :41
  atomicAdd(at(c[2], Y, X), float(mass * d[0] * d[1] * tz));      // :43
  atomicAdd(at(Z, Y, c[0]), float(mass * tx * d[1] * d[2]));      // :44
  atomicAdd(at(Z, c[1], X), float(mass * d[0] * ty * d[2]));      // :45
  atomicAdd(at(Z, Y, X), float(mass * d[0] * d[1] * d[2]));       // :47
}

__global__ __launch_bounds__(kBlock) void k_ref_cells(const float* __restrict__ pos, long np, int n,
                                                      int64_t* __restrict__ cells) {
  const long i = blockIdx.x * (long)kBlock + threadIdx.x;
  if (i >= 3 * np) return;
  const float x = ref_wrap(pos[i], n);
  cells[i] = cell_of<float>(x, n);
}

__global__ __launch_bounds__(kBlock) void k_ref_weights(const int64_t* __restrict__ cells, const float* __restrict__ pos,
                                                        long np, int n_unused, float* __restrict__ t) {
  (void)n_unused;
  const long i = blockIdx.x * (long)kBlock + threadIdx.x;
  if (i >= np) return;
  double d[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    // integrate.py:36-38: positions - cell_centers in float64
    float x = pos[a * np + i];
    const int64_t c = cells[a * np + i];
    // a position stored as N (Q4) has cell 0: use its wrapped value
    if (x >= 0 && (double)x - (double)c >= 1.0) x = 0.0f;
    d[a] = (double)x - (double)c;
  }
  const double tx = 1 - d[0], ty = 1 - d[1], tz = 1 - d[2];
  t[0 * np + i] = float(tx * ty * tz);      // integrate.py:44
  t[1 * np + i] = float(d[0] * ty * tz);    // :45
  t[2 * np + i] = float(tx * d[1] * tz);    // :46
  t[3 * np + i] = float(tx * ty * d[2]);    // :47
  t[4 * np + i] = float(d[0] * d[1] * tz);  // :48
  t[5 * np + i] = float(d[0] * ty * d[2]);  // :49
  t[6 * np + i] = float(tx * d[1] * d[2]);  // :50
  t[7 * np + i] = float(d[0] * d[1] * d[2]);  // :51
}

// Python float modulo: result takes the sign of the divisor (used by integrate.py:95)
__device__ __forceinline__ double py_mod(double x, double n) {
  double r = fmod(x, n);
  if (r != 0.0 && ((r < 0) != (n < 0))) r += n;
  return r;
}

// the gradient of one axis at one particle from its cells and float32 weights
__device__ __forceinline__ float ref_gradient(const int c0[3], const float* t8, const float* __restrict__ phi, int n,
                                              int dir) {
  int cp[3] = {c0[0], c0[1], c0[2]}, cn[3] = {c0[0], c0[1], c0[2]};
  cp[dir] = (c0[dir] + 1) % n;        // integrate.py:65
  cn[dir] = (c0[dir] - 1 + n) % n;    // integrate.py:64 (numba wraps index -1)
  const int x = cp[0], y = cp[1], z = cp[2];
  const int x2 = cn[0], y2 = cn[1], z2 = cn[2];
  const int X = (x + 1) % n, Y = (y + 1) % n, Z = (z + 1) % n;      // :77-79
  const int X2 = (x2 + 1) % n, Y2 = (y2 + 1) % n, Z2 = (z2 + 1) % n;  // :80-82
  const size_t nn = (size_t)n;
  auto P = [&](int zz, int yy, int xx) { return phi[((size_t)zz * nn + yy) * nn + xx]; };
  const float g = -P(z, y, x) + P(z2, y2, x2);        // :84
  const float g_x = -P(z, y, X) + P(z2, y2, X2);      // :85
  const float g_y = -P(z, Y, x) + P(z2, Y2, x2);      // :86
  const float g_z = -P(Z, y, x) + P(Z2, y2, x2);      // :87
  const float g_xy = -P(z, Y, X) + P(z2, Y2, X2);     // :88
  const float g_xz = -P(Z, y, X) + P(Z2, y2, X2);     // :89
  const float g_yz = -P(Z, Y, x) + P(Z2, Y2, x2);     // :90
  const float g_xyz = -P(Z, Y, X) + P(Z2, Y2, X2);    // :91
  return g * t8[0] + g_x * t8[1] + g_y * t8[2] + g_z * t8[3] + g_xy * t8[4] + g_xz * t8[5] + g_yz * t8[6] +
         g_xyz * t8[7];                               // :92 (divided by 2 by the caller)
}

__global__ __launch_bounds__(kBlock) void k_ref_sweep(const int64_t* __restrict__ cells, float* __restrict__ xd,
                                                      float* __restrict__ vd, const float* __restrict__ t,
                                                      const float* __restrict__ phi, long np, int n, double da,
                                                      double fa1, double a, int dir) {
  const long i = blockIdx.x * (long)kBlock + threadIdx.x;
  if (i >= np) return;
  // caller-supplied cells are reduced mod N, so any value indexes inside the grid
  const int64_t nl = n;
  const int c0[3] = {(int)(((cells[i] % nl) + nl) % nl), (int)(((cells[np + i] % nl) + nl) % nl),
                     (int)(((cells[2 * np + i] % nl) + nl) % nl)};
  float t8[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) t8[j] = t[j * np + i];
  const double gp = (double)ref_gradient(c0, t8, phi, n, dir) / 2.;
  const float v = float((double)vd[i] + da * fa1 * gp);                         // integrate.py:94
  vd[i] = v;
  xd[i] = float(py_mod((double)xd[i] + da * (double)v / ((a + da) * (a + da)) * fa1, (double)n));  // :95
}

__global__ __launch_bounds__(kBlock) void k_ref_integrate(float* __restrict__ pos, float* __restrict__ vel, long np,
                                                          int n, double a, double fa1, double da,
                                                          const float* __restrict__ phi) {
  const long i = blockIdx.x * (long)kBlock + threadIdx.x;
  if (i >= np) return;
  int c0[3];
  double d[3];
  float x[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    x[q] = ref_wrap(pos[q * np + i], n);
    const int ci = cell_of<float>(x[q], n);        // integrate.py:20
    c0[q] = ci;
    d[q] = (double)x[q] - (double)ci;              // integrate.py:36-38
  }
  const double tx = 1 - d[0], ty = 1 - d[1], tz = 1 - d[2];
  const float t8[8] = {float(tx * ty * tz),     float(d[0] * ty * tz),   float(tx * d[1] * tz),
                       float(tx * ty * d[2]),   float(d[0] * d[1] * tz), float(d[0] * ty * d[2]),
                       float(tx * d[1] * d[2]), float(d[0] * d[1] * d[2])};  // integrate.py:44-51
#pragma unroll
  for (int dir = 0; dir < 3; ++dir) {             // integrate.py:23-24, all with pre-step cells
    const double gp = (double)ref_gradient(c0, t8, phi, n, dir) / 2.;
    const float v = float((double)vel[dir * np + i] + da * fa1 * gp);
    vel[dir * np + i] = v;
    pos[dir * np + i] =
        float(py_mod((double)x[dir] + da * (double)v / ((a + da) * (a + da)) * fa1, (double)n));
  }
}

__global__ __launch_bounds__(kBlock) void k_ref_fourier_grid(int n, float* __restrict__ g) {
  const long i = blockIdx.x * (long)kBlock + threadIdx.x;
  const long nn = n;
  if (i >= nn * nn * nn) return;
  const int ix = (int)(i % nn), iy = (int)((i / nn) % nn), iz = (int)(i / (nn * nn));
  // fourier_utils.py:8-15: k = float32(2*pi*fftfreq(N)), s = sin(k/2)^2 in float32
  auto s2 = [&](int m) {
    const int f = m < (n + 1) / 2 ? m : m - n;                 // numpy fftfreq ordering
    const float k = float(6.283185307179586 * ((double)f / (double)n));
    const float sf = float(sin((double)(k * 0.5f)));
    return sf * sf;
  };
  const float ksq = (s2(iz) + s2(iy)) + s2(ix);
  g[i] = ksq != 0.0f ? 1.0f / ksq : 0.0f;                      // Q3: G(0) = 0
}

// ------------------------------------------------------------------------------------------
// spectrum helpers
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_expand_full(const double2* __restrict__ H, const double2* __restrict__ Q,
                                                        double2* __restrict__ full, int n) {
  const long i = blockIdx.x * (long)kBlock + threadIdx.x;
  const long nn = n, M = n / 2;
  if (i >= nn * nn * nn) return;
  const long x = i % nn, y = (i / nn) % nn, z = i / (nn * nn);
  double2 val;
  if (x < M) {
    val = H[(z * nn + y) * M + x];
  } else if (x == M) {
    val = Q[z * nn + y];
  } else {
    const long zz = (nn - z) % nn, yy = (nn - y) % nn, xx = nn - x;
    const double2 u = H[(zz * nn + yy) * M + xx];
    val = make_double2(u.x, -u.y);
  }
  full[i] = val;
}

__global__ __launch_bounds__(kBlock) void k_potential_k(const double2* __restrict__ in, const float* __restrict__ fg,
                                                        double c, double2* __restrict__ out, long count) {
  const long i = blockIdx.x * (long)kBlock + threadIdx.x;
  if (i >= count) return;
  const double f = c * (double)fg[i];      // potential.py:15, evaluated left to right
  const double2 u = in[i];
  out[i] = make_double2(f * u.x, f * u.y);
}

__global__ __launch_bounds__(kBlock) void k_hermitian_half(const double2* __restrict__ full, double2* __restrict__ H,
                                                           double2* __restrict__ Q, int n, double scale) {
  const long nn = n, M = n / 2;
  const long i = blockIdx.x * (long)kBlock + threadIdx.x;
  if (i >= nn * nn * (M + 1)) return;
  const long x = i % (M + 1), y = (i / (M + 1)) % nn, z = i / ((M + 1) * nn);
  const double2 a = full[(z * nn + y) * nn + x];
  const long zz = (nn - z) % nn, yy = (nn - y) % nn, xx = (nn - x) % nn;
  const double2 b = full[(zz * nn + yy) * nn + xx];
  const double2 h = make_double2(0.5 * scale * (a.x + b.x), 0.5 * scale * (a.y - b.y));
  if (x < M) H[(z * nn + y) * M + x] = h;
  else Q[z * nn + y] = h;
}

template <typename Tin, typename Tout>
__global__ __launch_bounds__(kBlock) void k_convert(const Tin* __restrict__ in, Tout* __restrict__ out, size_t count) {
  const size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x;
  if (i < count) out[i] = Tout(in[i]);
}

}  // namespace

// ------------------------------------------------------------------------------------------
// host launchers
// ------------------------------------------------------------------------------------------
template <typename R, typename Tin>
hipError_t sim_set_state(const Tin* pos, const Tin* vel, const PState<R>& p, const Lattice& L, hipStream_t st,
                         uint32_t* health) {
  k_set_state<R, Tin><<<grid_for(L.np), kBlock, 0, st>>>(pos, vel, p.s, p.v, p.id, L, health);
  return hipGetLastError();
}
template <typename R, typename Tout>
hipError_t sim_get_state(const PState<R>& p, Tout* pos, Tout* vel, const Lattice& L, hipStream_t st) {
  k_get_state<R, Tout><<<grid_for(L.np), kBlock, 0, st>>>(p.s, p.v, p.id, pos, vel, L);
  return hipGetLastError();
}
hipError_t ids_identity(uint32_t* id, const Lattice& L, hipStream_t st) {
  k_ids_identity<<<grid_for(L.np), kBlock, 0, st>>>(id, L.np);
  return hipGetLastError();
}
#define PM_INST(R, T)                                                                                  \
  template hipError_t sim_set_state<R, T>(const T*, const T*, const PState<R>&, const Lattice&, hipStream_t,   \
                                          uint32_t*);                                                          \
  template hipError_t sim_get_state<R, T>(const PState<R>&, T*, T*, const Lattice&, hipStream_t);
PM_INST(float, float)
PM_INST(float, double)
PM_INST(double, float)
PM_INST(double, double)
#undef PM_INST

hipError_t ref_density(const float* pos, long np, int n, double mass, float* rho, hipStream_t st) {
  hipError_t e = hipMemsetAsync(rho, 0, (size_t)n * n * n * sizeof(float), st);
  if (e != hipSuccess) return e;
  if (np > 0) k_ref_density<<<grid_for(np), kBlock, 0, st>>>(pos, np, n, mass, rho);
  return hipGetLastError();
}
hipError_t ref_cell_centers(const float* pos, long np, int n, int64_t* cells, hipStream_t st) {
  if (np > 0) k_ref_cells<<<grid_for(3 * np), kBlock, 0, st>>>(pos, np, n, cells);
  return hipGetLastError();
}
hipError_t ref_weights(const int64_t* cells, const float* pos, long np, float* t, hipStream_t st) {
  if (np > 0) k_ref_weights<<<grid_for(np), kBlock, 0, st>>>(cells, pos, np, 0, t);
  return hipGetLastError();
}
hipError_t ref_sweep(const int64_t* cells, float* xd, float* vd, const float* t, const float* phi, long np, int n,
                     double da, double fa1, double a, int dir, hipStream_t st) {
  if (np > 0) k_ref_sweep<<<grid_for(np), kBlock, 0, st>>>(cells, xd, vd, t, phi, np, n, da, fa1, a, dir);
  return hipGetLastError();
}
hipError_t ref_integrate(float* pos, float* vel, long np, int n, double a, double fa1, double da, const float* phi,
                         hipStream_t st) {
  if (np > 0) k_ref_integrate<<<grid_for(np), kBlock, 0, st>>>(pos, vel, np, n, a, fa1, da, phi);
  return hipGetLastError();
}
hipError_t ref_fourier_grid(int n, float* g, hipStream_t st) {
  k_ref_fourier_grid<<<grid_for((long)n * n * n), kBlock, 0, st>>>(n, g);
  return hipGetLastError();
}
hipError_t spec_expand_full(const double2* H, const double2* Q, double2* full, int n, hipStream_t st) {
  k_expand_full<<<grid_for((long)n * n * n), kBlock, 0, st>>>(H, Q, full, n);
  return hipGetLastError();
}
hipError_t spec_potential_k(const double2* in, const float* fgrid, double c, double2* out, int n, hipStream_t st) {
  const long cnt = (long)n * n * n;
  k_potential_k<<<grid_for(cnt), kBlock, 0, st>>>(in, fgrid, c, out, cnt);
  return hipGetLastError();
}
hipError_t spec_hermitian_half(const double2* full, double2* H, double2* Q, int n, double scale, hipStream_t st) {
  k_hermitian_half<<<grid_for((long)n * n * (n / 2 + 1)), kBlock, 0, st>>>(full, H, Q, n, scale);
  return hipGetLastError();
}
template <typename Tin, typename Tout>
hipError_t convert_grid(const Tin* in, Tout* out, size_t count, hipStream_t st) {
  if (count) k_convert<Tin, Tout><<<grid_for((long)count), kBlock, 0, st>>>(in, out, count);
  return hipGetLastError();
}
template hipError_t convert_grid<float, double>(const float*, double*, size_t, hipStream_t);
template hipError_t convert_grid<double, float>(const double*, float*, size_t, hipStream_t);
template hipError_t convert_grid<float, float>(const float*, float*, size_t, hipStream_t);
template hipError_t convert_grid<double, double>(const double*, double*, size_t, hipStream_t);

}  // namespace pm
