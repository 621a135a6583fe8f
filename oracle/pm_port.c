/*
 * CPU port of the reference's PM step loops — TEST INFRASTRUCTURE ONLY.
 *
 * Used as the CPU baseline that bench.py times beside the GPU (and by tests as a second
 * restatement of the oracle).  It follows the reference loop for loop, with its dtypes:
 *   port_density        density.py:7-48     (float32 grid, float64 weights; race-free: the
 *                                             reference's prange scatter races, here each
 *                                             add is an OpenMP atomic)
 *   port_cell_centers   integrate.py:20     (int64 cells)
 *   port_weights        integrate.py:27-52  (float32 (8, N_p))
 *   port_sweep          integrate.py:54-97  (one axis; cc_n/cc_p copies as in :61-65)
 *   port_integrate      integrate.py:15-25  (cells, weights, then three sweeps)
 * The FFTs of potential.py are done by the caller (scipy.fft complex128 standing in for FFTW).
 * A float32 position equal to N is wrapped to 0 (Q4), as in the oracle.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline float wrapf(float x, int n) {
  if (x < 0) x += (float)n;
  if (x >= (float)n) x -= (float)n;
  return x;
}

void port_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}

int port_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void port_density(const float* pos, int64_t np, int n, double mass, float* grid) {
  const int64_t nn = n;
  memset(grid, 0, sizeof(float) * nn * nn * nn);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < np; ++i) {
    const float x = wrapf(pos[i], n), y = wrapf(pos[np + i], n), z = wrapf(pos[2 * np + i], n);
    const int64_t xc = ((int64_t)floorf(x)) % n, yc = ((int64_t)floorf(y)) % n, zc = ((int64_t)floorf(z)) % n;
    const double dx = x - xc, dy = y - yc, dz = z - zc;
    const double tx = 1 - dx, ty = 1 - dy, tz = 1 - dz;
    const int64_t X = (xc + 1) % n, Y = (yc + 1) % n, Z = (zc + 1) % n;
#define ADD(zz, yy, xx, w)                                   \
  do {                                                       \
    float* p_ = &grid[((zz) * nn + (yy)) * nn + (xx)];       \
    const float v_ = (float)(mass * (w));                    \
    _Pragma("omp atomic") * p_ += v_;                        \
  } while (0)
    ADD(zc, yc, xc, tx * ty * tz);
    ADD(zc, yc, X, dx * ty * tz);
    ADD(zc, Y, xc, tx * dy * tz);
    ADD(Z, yc, xc, tx * ty * dz);
    ADD(zc, Y, X, dx * dy * tz);
    ADD(Z, Y, xc, tx * dy * dz);
    ADD(Z, yc, X, dx * ty * dz);
    ADD(Z, Y, X, dx * dy * dz);
#undef ADD
  }
}

void port_cell_centers(const float* pos, int64_t np, int n, int64_t* cells) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < 3 * np; ++i) cells[i] = ((int64_t)floorf(wrapf(pos[i], n))) % n;
}

void port_weights(const int64_t* cells, const float* pos, int64_t np, int n, float* t) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < np; ++i) {
    const double dx = wrapf(pos[i], n) - cells[i];
    const double dy = wrapf(pos[np + i], n) - cells[np + i];
    const double dz = wrapf(pos[2 * np + i], n) - cells[2 * np + i];
    const double tx = 1 - dx, ty = 1 - dy, tz = 1 - dz;
    t[0 * np + i] = (float)(tx * ty * tz);
    t[1 * np + i] = (float)(dx * ty * tz);
    t[2 * np + i] = (float)(tx * dy * tz);
    t[3 * np + i] = (float)(tx * ty * dz);
    t[4 * np + i] = (float)(dx * dy * tz);
    t[5 * np + i] = (float)(dx * ty * dz);
    t[6 * np + i] = (float)(tx * dy * dz);
    t[7 * np + i] = (float)(dx * dy * dz);
  }
}

static inline double py_mod(double x, double n) {
  double r = fmod(x, n);
  if (r != 0.0 && ((r < 0) != (n < 0))) r += n;
  return r;
}

void port_sweep(const int64_t* cells, float* xd, float* vd, const float* t, const float* phi, int64_t np, int n,
                double da, double fa1, double a, int dir) {
  const int64_t nn = n;
  /* integrate.py:61-65: the reference copies the whole cell array twice per axis */
  int64_t* cc_n = (int64_t*)malloc(sizeof(int64_t) * 3 * np);
  int64_t* cc_p = (int64_t*)malloc(sizeof(int64_t) * 3 * np);
  memcpy(cc_n, cells, sizeof(int64_t) * 3 * np);
  memcpy(cc_p, cells, sizeof(int64_t) * 3 * np);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < np; ++i) {
    cc_n[dir * np + i] = (cells[dir * np + i] - 1 + n) % n;
    cc_p[dir * np + i] = (cells[dir * np + i] + 1) % n;
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < np; ++i) {
    const int64_t x = cc_p[i], y = cc_p[np + i], z = cc_p[2 * np + i];
    const int64_t x2 = cc_n[i], y2 = cc_n[np + i], z2 = cc_n[2 * np + i];
    const int64_t X = (x + 1) % n, Y = (y + 1) % n, Z = (z + 1) % n;
    const int64_t X2 = (x2 + 1) % n, Y2 = (y2 + 1) % n, Z2 = (z2 + 1) % n;
#define P(zz, yy, xx) phi[((zz) * nn + (yy)) * nn + (xx)]
    const float g = -P(z, y, x) + P(z2, y2, x2);
    const float g_x = -P(z, y, X) + P(z2, y2, X2);
    const float g_y = -P(z, Y, x) + P(z2, Y2, x2);
    const float g_z = -P(Z, y, x) + P(Z2, y2, x2);
    const float g_xy = -P(z, Y, X) + P(z2, Y2, X2);
    const float g_xz = -P(Z, y, X) + P(Z2, y2, X2);
    const float g_yz = -P(Z, Y, x) + P(Z2, Y2, x2);
    const float g_xyz = -P(Z, Y, X) + P(Z2, Y2, X2);
#undef P
    const double gp = (double)(g * t[i] + g_x * t[np + i] + g_y * t[2 * np + i] + g_z * t[3 * np + i] +
                               g_xy * t[4 * np + i] + g_xz * t[5 * np + i] + g_yz * t[6 * np + i] +
                               g_xyz * t[7 * np + i]) / 2.;
    const float v = (float)((double)vd[i] + da * fa1 * gp);
    vd[i] = v;
    xd[i] = (float)py_mod((double)xd[i] + da * (double)v / ((a + da) * (a + da)) * fa1, (double)n);
  }
  free(cc_n);
  free(cc_p);
}

void port_integrate(float* pos, float* vel, int64_t np, int n, double a, double fa1, double da, const float* phi) {
  int64_t* cells = (int64_t*)malloc(sizeof(int64_t) * 3 * np);
  float* t = (float*)malloc(sizeof(float) * 8 * np);
  port_cell_centers(pos, np, n, cells);
  port_weights(cells, pos, np, n, t);
  for (int d = 0; d < 3; ++d) port_sweep(cells, pos + d * np, vel + d * np, t, phi, np, n, da, fa1, a, d);
  free(cells);
  free(t);
}

/* potential.py:15 on a complex128 spectrum in place: X *= c * fgrid */
void port_potential_k(double* spec, const float* fgrid, int64_t count, double c) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < count; ++i) {
    const double f = c * (double)fgrid[i];
    spec[2 * i] *= f;
    spec[2 * i + 1] *= f;
  }
}
