// 3-D real FFT passes for the PM Poisson solve (gfx950).  See pm_fft.h for the layout.
//
// Pass structure of one Poisson solve (5 launches, each one read + one write of the grid):
//   r2c_x   : rho rows (N reals)        -> H rows (N/2 complex) + Nyquist plane Q
//   c2c_y   : forward along y on H and Q (in place)
//   z_green : forward along z, multiply by the k-space operator, inverse along z (in place)
//   c2c_y   : inverse along y (in place)
//   c2r_x   : H rows + Q -> real rows, written over H in place
// The real-to-complex step packs each row as N/2 complex samples z[n] = x[2n] + i x[2n+1],
// transforms them with an N/2-point FFT and untangles the even/odd spectra.
#include <algorithm>

#include "pm_batch.h"

namespace pm {

namespace {

constexpr int kThreads = 256;

template <typename R> constexpr int pencil_width(int L) {
  // 128-byte row segments for L <= 512 (16 float2 or 8 double2); half that above.
  return sizeof(R) == 4 ? (L <= 512 ? 16 : 8) : (L <= 512 ? 8 : 4);
}
template <typename R, int N> constexpr int pencil_w() {
  // largest power of two <= the preferred width that divides the N/2 kx columns
  int w = pencil_width<R>(N);
  while ((N / 2) % w != 0) w /= 2;
  return w;
}
template <typename R, int N> constexpr int rows_per_block() {
  // about 32 KiB of LDS per block for the x passes
  constexpr int M = N / 2;
  constexpr int budget = 32768 / (M * (int)sizeof(cpx<R>));
  int rb = 1;
  while (rb * 2 <= budget && rb * 2 <= 64) rb *= 2;
  return rb;
}

// ------------------------------------------------------------------------------------------
// x pass, real -> complex
// ------------------------------------------------------------------------------------------
// LDS of the x passes: M*(RB+1) complex for the rows plus N twiddles
template <typename R, int N, int RB> constexpr int xpass_lds() { return ((N / 2) * (RB + 1) + N) * (int)sizeof(cpx<R>); }

template <typename R, typename Tin, int N, int RB>
__device__ __forceinline__ void r2c_body(const Tin* __restrict__ rho, cpx<R>* __restrict__ H, cpx<R>* __restrict__ Q,
                                         const cpx<R>* __restrict__ twg, long blk, cpx<R>* lds, cpx<R>* tw) {
  using T = cpx<R>;
  constexpr int M = N / 2;
  constexpr int SN = RB + 1;
  constexpr int NT = kThreads;
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += NT) tw[i] = twg[i];
  const size_t row0 = (size_t)blk * RB;
  const cpx<Tin>* src = reinterpret_cast<const cpx<Tin>*>(rho) + row0 * M;
  // the block's loads are issued in chunks of CH per thread before they are stored
  constexpr int LPT = (RB * M + NT - 1) / NT;
  constexpr int CH = sizeof(R) == 4 ? 32 : 8;
#pragma unroll
  for (int i0 = 0; i0 < LPT; i0 += CH) {
    cpx<Tin> u[CH];
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const int e = (i0 + i) * NT + tid;
      if (i0 + i < LPT && e < RB * M) u[i] = src[e];
    }
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const int e = (i0 + i) * NT + tid;
      if (i0 + i < LPT && e < RB * M) {
        const int w = e / M, n = e - w * M;
        lds[n * SN + w] = cmake<T>(R(u[i].x), R(u[i].y));
      }
    }
  }
  __syncthreads();
  fft_passes<T, M, 1, -1, RB, SN, NT, 2>(lds, tw, tid);

  constexpr int NPAIR = M / 2 + 1;
  constexpr int TOTP = NPAIR * RB;
  constexpr int PERP = (TOTP + NT - 1) / NT;
  T zk[PERP], zm[PERP];
#pragma unroll
  for (int i = 0; i < PERP; ++i) {
    const int b = tid + i * NT;
    if (b < TOTP) {
      const int w = b % RB, k = b / RB;
      zk[i] = lds[k * SN + w];
      zm[i] = lds[((M - k) % M) * SN + w];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PERP; ++i) {
    const int b = tid + i * NT;
    if (b < TOTP) {
      const int w = b % RB, k = b / RB;
      // E = (Z[k] + conj Z[M-k]) / 2 ; O = (Z[k] - conj Z[M-k]) / (2i)
      const T a = zk[i], c = zm[i];
      const T E = cmake<T>(R(0.5) * (a.x + c.x), R(0.5) * (a.y - c.y));
      const T O = cmake<T>(R(0.5) * (a.y + c.y), R(-0.5) * (a.x - c.x));
      const T wO = cmul(O, tw[k]);
      if (k == 0) {
        lds[w] = cmake<T>(E.x + wO.x, R(0));
        Q[row0 + w] = cmake<T>(E.x - wO.x, R(0));
      } else {
        lds[k * SN + w] = cadd(E, wO);                       // X[k]
        if (2 * k != M) lds[(M - k) * SN + w] = cconj(csub(E, wO));  // X[M-k]
      }
    }
  }
  __syncthreads();
  T* dst = H + row0 * M;
#pragma unroll 4
  for (int e = tid; e < RB * M; e += NT) {
    const int w = e / M, n = e - w * M;
    dst[e] = lds[n * SN + w];
  }
}

template <typename R, typename Tin, int N, int RB>
__global__ __launch_bounds__(kThreads) void k_r2c_x(const Tin* __restrict__ rho, cpx<R>* __restrict__ H,
                                                    cpx<R>* __restrict__ Q, const cpx<R>* __restrict__ twg) {
  __shared__ cpx<R> lds[(N / 2) * (RB + 1)];
  __shared__ cpx<R> tw[N];
  r2c_body<R, Tin, N, RB>(rho, H, Q, twg, blockIdx.x, lds, tw);
}

// ------------------------------------------------------------------------------------------
// x pass, complex -> real (in place allowed: H and out may be the same memory)
// ------------------------------------------------------------------------------------------
template <typename R, typename Tout, int N, int RB>
__device__ __forceinline__ void c2r_body(const cpx<R>* H, const cpx<R>* __restrict__ Q, Tout* out,
                                         const cpx<R>* __restrict__ twg, long blk, cpx<R>* lds, cpx<R>* tw,
                                         cpx<R>* qx) {
  using T = cpx<R>;
  constexpr int M = N / 2;
  constexpr int SN = RB + 1;
  constexpr int NT = kThreads;
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += NT) tw[i] = twg[i];
  const size_t row0 = (size_t)blk * RB;
  const T* src = H + row0 * M;
  constexpr int LPT = (RB * M + NT - 1) / NT;
  constexpr int CH = sizeof(R) == 4 ? 32 : 8;
  if (tid < RB) qx[tid] = Q[row0 + tid];
#pragma unroll
  for (int i0 = 0; i0 < LPT; i0 += CH) {
    T u[CH];
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const int e = (i0 + i) * NT + tid;
      if (i0 + i < LPT && e < RB * M) u[i] = src[e];
    }
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const int e = (i0 + i) * NT + tid;
      if (i0 + i < LPT && e < RB * M) {
        const int w = e / M, n = e - w * M;
        lds[n * SN + w] = u[i];
      }
    }
  }
  __syncthreads();

  constexpr int NPAIR = M / 2 + 1;
  constexpr int TOTP = NPAIR * RB;
  constexpr int PERP = (TOTP + NT - 1) / NT;
  T xk[PERP], xm[PERP];
#pragma unroll
  for (int i = 0; i < PERP; ++i) {
    const int b = tid + i * NT;
    if (b < TOTP) {
      const int w = b % RB, k = b / RB;
      xk[i] = lds[k * SN + w];
      xm[i] = (k == 0) ? qx[w] : lds[(M - k) * SN + w];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PERP; ++i) {
    const int b = tid + i * NT;
    if (b < TOTP) {
      const int w = b % RB, k = b / RB;
      const T A = cadd(xk[i], cconj(xm[i]));
      const T B = csub(xk[i], cconj(xm[i]));
      const T wk = tw[k];
      // Z[k] = A + i conj(w^k) B
      const T cb = cmulc(B, wk);
      lds[k * SN + w] = cmake<T>(A.x - cb.y, A.y + cb.x);
      if (k != 0 && 2 * k != M) {
        // Z[M-k] = conj(A) + i w^k conj(B)
        const T wb = cmul(cconj(B), wk);
        lds[(M - k) * SN + w] = cmake<T>(A.x - wb.y, -A.y + wb.x);
      }
    }
  }
  __syncthreads();
  fft_passes<T, M, 1, +1, RB, SN, NT, 2>(lds, tw, tid);
  cpx<Tout>* dst = reinterpret_cast<cpx<Tout>*>(out) + row0 * M;
#pragma unroll 4
  for (int e = tid; e < RB * M; e += NT) {
    const int w = e / M, n = e - w * M;
    const T u = lds[n * SN + w];
    dst[e] = cmake<cpx<Tout>>(Tout(u.x), Tout(u.y));
  }
}

template <typename R, typename Tout, int N, int RB>
__global__ __launch_bounds__(kThreads) void k_c2r_x(const cpx<R>* H, const cpx<R>* __restrict__ Q, Tout* out,
                                                    const cpx<R>* __restrict__ twg) {
  __shared__ cpx<R> lds[(N / 2) * (RB + 1)];
  __shared__ cpx<R> tw[N];
  __shared__ cpx<R> qx[RB];
  c2r_body<R, Tout, N, RB>(H, Q, out, twg, blockIdx.x, lds, tw, qx);
}

// ------------------------------------------------------------------------------------------
// pencil passes along y or z
// ------------------------------------------------------------------------------------------

template <typename R>
__device__ __forceinline__ R green_factor(int kx, int ky, int kz, int n, const GreenArgs& g, R coef, const R* s2,
                                          const R* wd) {
  if (g.from_grid) {
    const size_t nn = (size_t)n;
    const size_t i1 = ((size_t)kz * nn + ky) * nn + kx;
    const size_t i2 = ((size_t)((n - kz) % n) * nn + (n - ky) % n) * nn + (n - kx) % n;
    return coef * R(0.5) * (R(g.fgrid[i1]) + R(g.fgrid[i2]));
  }
  const R den = s2[kx] + s2[ky] + s2[kz];
  R f = den > R(0) ? coef / den : R(0);
  if (g.deconvolve) f *= wd[kx] * wd[ky] * wd[kz];
  return f;
}

// LDS of a pencil pass: L*(W+1) complex + N twiddles (+ 2N reals of k-space tables)
template <typename R, int N> constexpr int pencil_lds() {
  return (N * (pencil_w<R, N>() + 1) + N) * (int)sizeof(cpx<R>) + 2 * N * (int)sizeof(R);
}

// MODE: -1 forward only, +1 inverse only, 0 forward + k-space operator + inverse
template <typename R, int N, int MODE>
__device__ __forceinline__ void pencil_body(cpx<R>* H, cpx<R>* Q, const PencilSeg& s0, const PencilSeg& s1,
                                            const GreenArgs& g, const cpx<R>* __restrict__ twg,
                                            const R* __restrict__ s2g, const R* __restrict__ wdg, long blk,
                                            cpx<R>* lds, cpx<R>* tw, R* s2, R* wd) {
  using T = cpx<R>;
  constexpr int L = N;
  constexpr int W = pencil_w<R, N>();
  constexpr int SN = W + 1;  // room for either layout
  constexpr int NT = kThreads;
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += NT) {
    tw[i] = twg[i];
    if constexpr (MODE == 0) {
      s2[i] = s2g[i];
      wd[i] = wdg[i];
    }
  }
  const bool first = blk < s0.nblk;
  const PencilSeg& s = first ? s0 : s1;
  if (!first) blk -= s0.nblk;
  T* data = first ? H : Q;
  T* base;
  int kx0 = 0, dkx = 0, ky0 = 0, dky = 0;
  if (s.contig) {
    const long o = blk / s.ncol_blk, c = blk - o * s.ncol_blk;
    base = data + o * s.outer_stride + c * W;
    if (s.green_kind == 0) { kx0 = (int)(c * W); dkx = 1; ky0 = (int)o; }
    else { kx0 = N / 2; ky0 = (int)(c * W); dky = 1; }
  } else {
    base = data + blk * (long)W * s.stride_w;
  }
  // load: the block's loads are issued in chunks of CH per thread before they are stored
  constexpr int LPT = (L * W + NT - 1) / NT;
  constexpr int CH = sizeof(R) == 4 ? 32 : 8;
  if (s.contig) {
#pragma unroll
    for (int i0 = 0; i0 < LPT; i0 += CH) {
      T u[CH];
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const int e = (i0 + i) * NT + tid;
        const int n = e / W, w = e - n * W;
        if (i0 + i < LPT && e < L * W) u[i] = base[n * s.stride_n + w];
      }
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const int e = (i0 + i) * NT + tid;
        const int n = e / W, w = e - n * W;
        if (i0 + i < LPT && e < L * W) lds[n * SN + w] = u[i];
      }
    }
  } else {
#pragma unroll
    for (int i0 = 0; i0 < LPT; i0 += CH) {
      T u[CH];
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const int e = (i0 + i) * NT + tid;
        const int w = e / L, n = e - w * L;
        if (i0 + i < LPT && e < L * W) u[i] = base[n + w * s.stride_w];
      }
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        const int e = (i0 + i) * NT + tid;
        const int w = e / L, n = e - w * L;
        if (i0 + i < LPT && e < L * W) lds[n * SN + w] = u[i];
      }
    }
  }
  __syncthreads();
  if constexpr (MODE == -1) {
    fft_passes<T, L, 1, -1, W, SN, NT, 1>(lds, tw, tid);
  } else if constexpr (MODE == 1) {
    fft_passes<T, L, 1, +1, W, SN, NT, 1>(lds, tw, tid);
  } else {
    fft_passes<T, L, 1, -1, W, SN, NT, 1>(lds, tw, tid);
    const R coef = R(g.table ? g.table[3 * (*g.step)] : g.coef);
#pragma unroll 4
    for (int e = tid; e < L * W; e += NT) {
      const int n = e / W, w = e - n * W;
      const R f = green_factor<R>(kx0 + w * dkx, ky0 + w * dky, n, N, g, coef, s2, wd);
      lds[n * SN + w] = cscale(lds[n * SN + w], f);
    }
    __syncthreads();
    fft_passes<T, L, 1, +1, W, SN, NT, 1>(lds, tw, tid);
  }
  // store
  if (s.contig) {
#pragma unroll 4
    for (int e = tid; e < L * W; e += NT) {
      const int n = e / W, w = e - n * W;
      base[n * s.stride_n + w] = lds[n * SN + w];
    }
  } else {
#pragma unroll 4
    for (int e = tid; e < L * W; e += NT) {
      const int w = e / L, n = e - w * L;
      base[n + w * s.stride_w] = lds[n * SN + w];
    }
  }
}

template <typename R, int N, int MODE>
__global__ __launch_bounds__(kThreads) void k_pencil(cpx<R>* H, cpx<R>* Q, PencilSeg s0, PencilSeg s1, GreenArgs g,
                                                     const cpx<R>* __restrict__ twg, const R* __restrict__ s2g,
                                                     const R* __restrict__ wdg) {
  __shared__ cpx<R> lds[N * (pencil_w<R, N>() + 1)];
  __shared__ cpx<R> tw[N];
  __shared__ R s2[MODE == 0 ? N : 1];
  __shared__ R wd[MODE == 0 ? N : 1];
  pencil_body<R, N, MODE>(H, Q, s0, s1, g, twg, s2g, wdg, blockIdx.x, lds, tw, s2, wd);
}

// ------------------------------------------------------------------------------------------
// host-side segment descriptions
// ------------------------------------------------------------------------------------------
template <typename R, int N> void y_segments(PencilSeg& a, PencilSeg& b) {
  constexpr int W = pencil_w<R, N>();
  constexpr long M = N / 2;
  a = {};
  a.ncol_blk = M / W;
  a.nblk = (long)N * a.ncol_blk;
  a.outer_stride = (long)N * M;  // z
  a.stride_n = M;                // y
  a.contig = 1;
  b = {};
  b.nblk = N / W;                // column blocks of z
  b.stride_w = N;                // Q[z][y]: next z
  b.stride_n = 1;
  b.contig = 0;
}
template <typename R, int N> void z_segments(PencilSeg& a, PencilSeg& b) {
  constexpr int W = pencil_w<R, N>();
  constexpr long M = N / 2;
  a = {};
  a.ncol_blk = M / W;
  a.nblk = (long)N * a.ncol_blk;
  a.outer_stride = M;            // y
  a.stride_n = (long)N * M;      // z
  a.contig = 1;
  a.green_kind = 0;
  b = {};
  b.ncol_blk = N / W;            // column blocks of y
  b.nblk = b.ncol_blk;
  b.outer_stride = 0;
  b.stride_n = N;                // Q[z][y]: next z
  b.contig = 1;
  b.green_kind = 1;
}

template <typename R, typename Tin, int N>
hipError_t launch_r2c(const Tin* rho, cpx<R>* H, cpx<R>* Q, const FftTables& t, hipStream_t st) {
  constexpr int RB = rows_per_block<R, N>();
  static_assert(((long)N * N) % RB == 0, "rows per block must divide N^2");
  k_r2c_x<R, Tin, N, RB><<<(long)N * N / RB, kThreads, 0, st>>>(rho, H, Q, (const cpx<R>*)t.tw);
  return hipGetLastError();
}
template <typename R, typename Tout, int N>
hipError_t launch_c2r(cpx<R>* H, const cpx<R>* Q, Tout* out, const FftTables& t, hipStream_t st) {
  constexpr int RB = rows_per_block<R, N>();
  k_c2r_x<R, Tout, N, RB><<<(long)N * N / RB, kThreads, 0, st>>>(H, Q, out, (const cpx<R>*)t.tw);
  return hipGetLastError();
}
template <typename R, int N, int MODE>
hipError_t launch_pencil(cpx<R>* H, cpx<R>* Q, bool zaxis, const GreenArgs& g, const FftTables& t,
                         hipStream_t st) {
  PencilSeg a, b;
  if (zaxis) z_segments<R, N>(a, b);
  else y_segments<R, N>(a, b);
  k_pencil<R, N, MODE><<<a.nblk + b.nblk, kThreads, 0, st>>>(H, Q, a, b, g, (const cpx<R>*)t.tw,
                                                            (const R*)t.sin2, (const R*)t.wdec);
  return hipGetLastError();
}

}  // namespace

bool fft_size_supported(int n) {
  switch (n) {
#define PM_CASE(NN) case NN:
    PM_FOR_EACH_N(PM_CASE)
#undef PM_CASE
    return true;
    default:
      return false;
  }
}

template <typename R, typename Tin>
hipError_t fft_r2c_x(const Tin* rho, cpx<R>* H, cpx<R>* Q, int n, const FftTables& t, hipStream_t st) {
  switch (n) {
#define PM_CASE(NN) case NN: return launch_r2c<R, Tin, NN>(rho, H, Q, t, st);
    PM_FOR_EACH_N(PM_CASE)
#undef PM_CASE
  }
  return hipErrorInvalidValue;
}
template <typename R, typename Tout>
hipError_t fft_c2r_x(cpx<R>* H, const cpx<R>* Q, Tout* out, int n, const FftTables& t, hipStream_t st) {
  switch (n) {
#define PM_CASE(NN) case NN: return launch_c2r<R, Tout, NN>(H, Q, out, t, st);
    PM_FOR_EACH_N(PM_CASE)
#undef PM_CASE
  }
  return hipErrorInvalidValue;
}
template <typename R>
hipError_t fft_c2c_y(cpx<R>* H, cpx<R>* Q, int n, int sign, const FftTables& t, hipStream_t st) {
  GreenArgs g{};
  switch (n) {
#define PM_CASE(NN)                                                                   \
  case NN:                                                                            \
    return sign < 0 ? launch_pencil<R, NN, -1>(H, Q, false, g, t, st)                 \
                    : launch_pencil<R, NN, 1>(H, Q, false, g, t, st);
    PM_FOR_EACH_N(PM_CASE)
#undef PM_CASE
  }
  return hipErrorInvalidValue;
}
template <typename R>
hipError_t fft_c2c_z(cpx<R>* H, cpx<R>* Q, int n, int sign, const FftTables& t, hipStream_t st) {
  GreenArgs g{};
  switch (n) {
#define PM_CASE(NN)                                                                   \
  case NN:                                                                            \
    return sign < 0 ? launch_pencil<R, NN, -1>(H, Q, true, g, t, st)                  \
                    : launch_pencil<R, NN, 1>(H, Q, true, g, t, st);
    PM_FOR_EACH_N(PM_CASE)
#undef PM_CASE
  }
  return hipErrorInvalidValue;
}
template <typename R>
hipError_t fft_z_green(cpx<R>* H, cpx<R>* Q, int n, const GreenArgs& g, const FftTables& t, hipStream_t st) {
  switch (n) {
#define PM_CASE(NN) case NN: return launch_pencil<R, NN, 0>(H, Q, true, g, t, st);
    PM_FOR_EACH_N(PM_CASE)
#undef PM_CASE
  }
  return hipErrorInvalidValue;
}

template hipError_t fft_r2c_x<float, float>(const float*, cpx<float>*, cpx<float>*, int, const FftTables&,
                                             hipStream_t);
template hipError_t fft_r2c_x<double, double>(const double*, cpx<double>*, cpx<double>*, int, const FftTables&,
                                               hipStream_t);
template hipError_t fft_r2c_x<double, float>(const float*, cpx<double>*, cpx<double>*, int, const FftTables&,
                                              hipStream_t);
template hipError_t fft_c2r_x<float, float>(cpx<float>*, const cpx<float>*, float*, int, const FftTables&,
                                             hipStream_t);
template hipError_t fft_c2r_x<double, double>(cpx<double>*, const cpx<double>*, double*, int, const FftTables&,
                                               hipStream_t);
template hipError_t fft_c2r_x<double, float>(cpx<double>*, const cpx<double>*, float*, int, const FftTables&,
                                              hipStream_t);
#define PM_INST(R)                                                                                   \
  template hipError_t fft_c2c_y<R>(cpx<R>*, cpx<R>*, int, int, const FftTables&, hipStream_t);       \
  template hipError_t fft_c2c_z<R>(cpx<R>*, cpx<R>*, int, int, const FftTables&, hipStream_t);       \
  template hipError_t fft_z_green<R>(cpx<R>*, cpx<R>*, int, const GreenArgs&, const FftTables&, hipStream_t);
PM_INST(float)
PM_INST(double)
#undef PM_INST

// ------------------------------------------------------------------------------------------
// batched passes: the blocks of every member of an ensemble in one launch (FP32)
// ------------------------------------------------------------------------------------------
namespace {

template <int N> constexpr int x_rb() { return rows_per_block<float, N>(); }

// the cases of size class CLS (the kernel's registers are the maximum over its cases, so the
// large grids get a kernel of their own)
#define PM_BATCH_CLASS_SWITCH(CLS, NV, CASE)                \
  if constexpr (CLS == 0) {                                 \
    switch (NV) { PM_FOR_EACH_BATCH_N_SMALL(CASE) default: break; } \
  } else {                                                  \
    switch (NV) { PM_FOR_EACH_BATCH_N_LARGE(CASE) default: break; } \
  }

template <int CLS>
__global__ __launch_bounds__(kThreads) void k_batch_r2c(const BatchDesc* __restrict__ D, long base) {
  extern __shared__ __align__(16) unsigned char smem[];
  const long b = base + blockIdx.x;
  const int k = batch_member(D, BS_R2C, b);
  const BatchMember& m = D->m[k];
  const long blk = b - D->off[BS_R2C][k];
  cpx<float>* lds = reinterpret_cast<cpx<float>*>(smem);
#define PM_CASE(NN)                                                                              \
  case NN:                                                                                       \
    r2c_body<float, float, NN, x_rb<NN>()>(m.rho, m.H, m.Q, m.tw, blk, lds,                      \
                                           lds + (NN / 2) * (x_rb<NN>() + 1));                   \
    break;
  PM_BATCH_CLASS_SWITCH(CLS, m.n, PM_CASE)
#undef PM_CASE
}

template <int CLS>
__global__ __launch_bounds__(kThreads) void k_batch_c2r(const BatchDesc* __restrict__ D, long base) {
  extern __shared__ __align__(16) unsigned char smem[];
  const long b = base + blockIdx.x;
  const int k = batch_member(D, BS_C2R, b);
  const BatchMember& m = D->m[k];
  const long blk = b - D->off[BS_C2R][k];
  cpx<float>* lds = reinterpret_cast<cpx<float>*>(smem);
#define PM_CASE(NN)                                                                              \
  case NN: {                                                                                     \
    cpx<float>* tw = lds + (NN / 2) * (x_rb<NN>() + 1);                                          \
    c2r_body<float, float, NN, x_rb<NN>()>(m.H, m.Q, reinterpret_cast<float*>(m.H), m.tw, blk,   \
                                           lds, tw, tw + NN);                                    \
    break;                                                                                       \
  }
  PM_BATCH_CLASS_SWITCH(CLS, m.n, PM_CASE)
#undef PM_CASE
}

template <int CLS, int MODE>
__global__ __launch_bounds__(kThreads) void k_batch_pencil(const BatchDesc* __restrict__ D, long base) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int ST = MODE == 0 ? BS_Z : BS_Y;
  const long b = base + blockIdx.x;
  const int k = batch_member(D, ST, b);
  const BatchMember& m = D->m[k];
  const long blk = b - D->off[ST][k];
  GreenArgs g{};
  if (MODE == 0) {
    g.coef = D->table[3 * ((long)(*D->step) * D->n_members + k)];
    g.deconvolve = m.deconvolve;
    g.from_grid = m.from_grid;
    g.fgrid = m.fgrid;
  }
  cpx<float>* lds = reinterpret_cast<cpx<float>*>(smem);
#define PM_CASE(NN)                                                                              \
  case NN: {                                                                                     \
    cpx<float>* tw = lds + NN * (pencil_w<float, NN>() + 1);                                     \
    float* s2 = reinterpret_cast<float*>(tw + NN);                                               \
    pencil_body<float, NN, MODE>(m.H, m.Q, MODE == 0 ? m.za : m.ya, MODE == 0 ? m.zb : m.yb, g,  \
                                 m.tw, m.sin2, m.wdec, blk, lds, tw, s2, s2 + NN);               \
    break;                                                                                       \
  }
  PM_BATCH_CLASS_SWITCH(CLS, m.n, PM_CASE)
#undef PM_CASE
}

inline hipError_t launch_dyn(const void* fn, long base, long count, int lds, const BatchDesc* d, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  if (count > 0x7fffffffL) return hipErrorInvalidValue;
  void* args[] = {(void*)&d, (void*)&base};
  return hipLaunchKernel(fn, dim3((unsigned)count), dim3(kThreads), args, (size_t)lds, st);
}

}  // namespace

bool batch_size_supported(int n) {
  switch (n) {
#define PM_CASE(NN) case NN:
    PM_FOR_EACH_BATCH_N(PM_CASE)
#undef PM_CASE
    return true;
    default:
      return false;
  }
}

void batch_fft_segments(int n, PencilSeg& ya, PencilSeg& yb, PencilSeg& za, PencilSeg& zb) {
  switch (n) {
#define PM_CASE(NN)                    \
  case NN:                             \
    y_segments<float, NN>(ya, yb);     \
    z_segments<float, NN>(za, zb);     \
    return;
    PM_FOR_EACH_BATCH_N(PM_CASE)
#undef PM_CASE
  }
  ya = yb = za = zb = PencilSeg{};
}

long batch_fft_blocks(int n, int stage) {
  switch (n) {
#define PM_CASE(NN)                                                          \
  case NN: {                                                                 \
    PencilSeg a, b;                                                          \
    if (stage == BS_R2C || stage == BS_C2R) return (long)NN * NN / x_rb<NN>(); \
    if (stage == BS_Y) y_segments<float, NN>(a, b);                          \
    else z_segments<float, NN>(a, b);                                        \
    return a.nblk + b.nblk;                                                  \
  }
    PM_FOR_EACH_BATCH_N(PM_CASE)
#undef PM_CASE
  }
  return 0;
}

int batch_fft_lds(int n, int stage) {
  switch (n) {
#define PM_CASE(NN)                                                                                  \
  case NN:                                                                                           \
    if (stage == BS_R2C) return xpass_lds<float, NN, x_rb<NN>()>();                                  \
    if (stage == BS_C2R) return xpass_lds<float, NN, x_rb<NN>()>() + x_rb<NN>() * 8;                 \
    return pencil_lds<float, NN>();
    PM_FOR_EACH_BATCH_N(PM_CASE)
#undef PM_CASE
  }
  return 0;
}

int batch_class(int n) { return n <= 256 ? 0 : 1; }

hipError_t batch_fft_init() {
  // blocks above 64 KiB of LDS are legal on gfx950: allow the batched FFT kernels the largest
  // request any member size makes (set once, outside any stream capture)
  int lds = 0;
  for (int st = BS_R2C; st <= BS_C2R; ++st) lds = std::max(lds, batch_fft_lds(512, st));
  const void* fns[] = {(const void*)k_batch_r2c<0>,         (const void*)k_batch_r2c<1>,
                       (const void*)k_batch_c2r<0>,         (const void*)k_batch_c2r<1>,
                       (const void*)k_batch_pencil<0, -1>, (const void*)k_batch_pencil<1, -1>,
                       (const void*)k_batch_pencil<0, 0>,  (const void*)k_batch_pencil<1, 0>,
                       (const void*)k_batch_pencil<0, 1>,  (const void*)k_batch_pencil<1, 1>};
  for (const void* f : fns)
    if (hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds)) return e;
  return hipSuccess;
}

hipError_t batch_r2c(const BatchDesc* d, int cls, long base, long count, int lds, hipStream_t st) {
  return launch_dyn(cls ? (const void*)k_batch_r2c<1> : (const void*)k_batch_r2c<0>, base, count, lds, d, st);
}
hipError_t batch_c2r(const BatchDesc* d, int cls, long base, long count, int lds, hipStream_t st) {
  return launch_dyn(cls ? (const void*)k_batch_c2r<1> : (const void*)k_batch_c2r<0>, base, count, lds, d, st);
}
hipError_t batch_pencil(const BatchDesc* d, int cls, int mode, long base, long count, int lds, hipStream_t st) {
  const void* fn;
  if (mode < 0) fn = cls ? (const void*)k_batch_pencil<1, -1> : (const void*)k_batch_pencil<0, -1>;
  else if (mode > 0) fn = cls ? (const void*)k_batch_pencil<1, 1> : (const void*)k_batch_pencil<0, 1>;
  else fn = cls ? (const void*)k_batch_pencil<1, 0> : (const void*)k_batch_pencil<0, 0>;
  return launch_dyn(fn, base, count, lds, d, st);
}

}  // namespace pm
