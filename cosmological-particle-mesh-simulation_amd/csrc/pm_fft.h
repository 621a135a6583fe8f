// Hand-written 3-D real FFT for the PM Poisson solve: Stockham passes staged through LDS.
//
// Layout of the half spectrum of an N^3 real grid rho[z][y][x] (x fastest):
//   H[z][y][kx]  kx in [0, N/2)   (row pitch N/2 complex = N reals, so the inverse C2R can
//                                  overwrite H in place with a dense real N^3 grid)
//   Q[z][y]      kx = N/2         (the Nyquist plane, stored separately)
// The forward transform is unnormalised (FFTW_FORWARD); the inverse is unnormalised too, the
// 1/N^3 factor being folded into the k-space operator.
#pragma once
#include "pm_common.h"

namespace pm {

// ------------------------------------------------------------------------------------------
// small DFT kernels, sign S = -1 (forward) or +1 (inverse)
// ------------------------------------------------------------------------------------------
template <int S, typename T> __device__ __forceinline__ void dft2(T* v) {
  T a = v[0];
  v[0] = cadd(a, v[1]);
  v[1] = csub(a, v[1]);
}

template <int S, typename T> __device__ __forceinline__ void dft3(T* v) {
  using R = decltype(T::x);
  const R h = R(0.86602540378443864676);  // sqrt(3)/2
  T t = cadd(v[1], v[2]);
  T d = csub(v[1], v[2]);
  T y0 = cadd(v[0], t);
  T m = cmake<T>(v[0].x - R(0.5) * t.x, v[0].y - R(0.5) * t.y);
  // S*i*h*d = (-S*h*d.y, S*h*d.x)
  T r = cmake<T>(-R(S) * h * d.y, R(S) * h * d.x);
  v[0] = y0;
  v[1] = cadd(m, r);
  v[2] = csub(m, r);
}

template <int S, typename T> __device__ __forceinline__ void dft4(T* v) {
  T t0 = cadd(v[0], v[2]);
  T t1 = csub(v[0], v[2]);
  T t2 = cadd(v[1], v[3]);
  T t3 = cmul_si<S>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}

template <int S, typename T> __device__ __forceinline__ void dft8(T* v) {
  using R = decltype(T::x);
  const R c = R(0.70710678118654752440);
  T e[4] = {v[0], v[2], v[4], v[6]};
  T o[4] = {v[1], v[3], v[5], v[7]};
  dft4<S>(e);
  dft4<S>(o);
  // o1 *= w8 = (1 + S i)/sqrt2 ; o2 *= S i ; o3 *= (-1 + S i)/sqrt2
  T o1 = cmake<T>(c * (o[1].x - R(S) * o[1].y), c * (o[1].y + R(S) * o[1].x));
  T o2 = cmul_si<S>(o[2]);
  T o3 = cmake<T>(c * (-o[3].x - R(S) * o[3].y), c * (-o[3].y + R(S) * o[3].x));
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

template <int RAD, int S, typename T> __device__ __forceinline__ void dftR(T* v) {
  if constexpr (RAD == 2) dft2<S>(v);
  else if constexpr (RAD == 3) dft3<S>(v);
  else if constexpr (RAD == 4) dft4<S>(v);
  else if constexpr (RAD == 8) dft8<S>(v);
  else static_assert(RAD == 2, "unsupported radix");
}

// ------------------------------------------------------------------------------------------
// DFTs held entirely in registers (sizes 2^a 3^b up to a few dozen), for the four-step pencil
// transform.  Twiddle constants are evaluated at compile time.
// ------------------------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846;

// sin / cos of x in [-pi, pi] by their Taylor series (converged to double precision)
constexpr double c_sin(double x) {
  double term = x, sum = x;
  for (int k = 1; k < 30; ++k) {
    term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double c_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 30; ++k) {
    term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}

// exp(-2 pi i m / A) for m in [0, A)
template <int A> struct RegTwiddles {
  double re[A], im[A];
  constexpr RegTwiddles() : re(), im() {
    for (int m = 0; m < A; ++m) {
      double ang = 2.0 * kPi * m / A;
      if (ang > kPi) ang -= 2.0 * kPi;
      re[m] = c_cos(ang);
      im[m] = -c_sin(ang);
    }
  }
};

// DFT of A values in registers, sign S: decimation in time, A = M * RR with the RR-point
// butterflies last: X[k1 + M k2] = sum_r W_RR^{r k2} W_A^{r k1} DFT_M(x[m RR + r])[k1].
template <int S, int A, typename T> __device__ __forceinline__ void reg_dft(T (&v)[A]) {
  using R = decltype(T::x);
  if constexpr (A == 1) {
  } else if constexpr (A == 2 || A == 3 || A == 4 || A == 8) {
    dftR<A, S>(v);
  } else {
    constexpr int RR = (A % 8 == 0) ? 8 : (A % 4 == 0) ? 4 : (A % 3 == 0) ? 3 : 2;
    constexpr int M = A / RR;
    constexpr RegTwiddles<A> tw{};
    T sub[RR][M];
#pragma unroll
    for (int r = 0; r < RR; ++r)
#pragma unroll
      for (int m = 0; m < M; ++m) sub[r][m] = v[m * RR + r];
#pragma unroll
    for (int r = 0; r < RR; ++r) reg_dft<S, M>(sub[r]);
#pragma unroll
    for (int r = 1; r < RR; ++r)
#pragma unroll
      for (int k1 = 1; k1 < M; ++k1) {
        const int idx = (r * k1) % A;
        const T w = cmake<T>(R(tw.re[idx]), R(S < 0 ? tw.im[idx] : -tw.im[idx]));
        sub[r][k1] = cmul(sub[r][k1], w);
      }
#pragma unroll
    for (int k1 = 0; k1 < M; ++k1) {
      T col[RR];
#pragma unroll
      for (int r = 0; r < RR; ++r) col[r] = sub[r][k1];
      dftR<RR, S>(col);
#pragma unroll
      for (int k2 = 0; k2 < RR; ++k2) v[k1 + M * k2] = col[k2];
    }
  }
}

// largest preferred radix dividing the remaining length
template <int L> struct first_radix {
  static constexpr int value = (L % 8 == 0) ? 8 : (L % 4 == 0) ? 4 : (L % 2 == 0) ? 2 : (L % 3 == 0) ? 3 : 0;
};

template <int L> struct fft_len_ok {
  static constexpr bool value = [] {
    int l = L;
    while (l % 2 == 0) l /= 2;
    while (l % 3 == 0) l /= 3;
    return l == 1;
  }();
};

// ------------------------------------------------------------------------------------------
// One Stockham radix-RAD pass over W independent length-L sequences held in LDS.
// Element (n, w) lives at lds[n*SN + w].  Butterfly b -> (w = b % W, j = b / W), so the
// lanes of a wavefront walk along w (contiguous LDS addresses).
// tw[m*TS] = exp(-2 pi i m / L) for m in [0, L).
// ------------------------------------------------------------------------------------------
template <typename T, int L, int RAD, int NS, int S, int W, int SN, int NT, int TS>
__device__ __forceinline__ void stockham_pass(T* lds, const T* tw, int tid) {
  constexpr int NB = L / RAD;
  constexpr int TOT = NB * W;
  constexpr int PER = (TOT + NT - 1) / NT;
  T v[PER][RAD];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int b = tid + i * NT;
    if (TOT % NT == 0 || b < TOT) {
      const int w = b % W, j = b / W;
#pragma unroll
      for (int r = 0; r < RAD; ++r) v[i][r] = lds[(j + r * NB) * SN + w];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int b = tid + i * NT;
    if (TOT % NT == 0 || b < TOT) {
      const int w = b % W, j = b / W;
      const int k = j % NS;
      if constexpr (NS > 1) {
        constexpr int STEP = L / (NS * RAD);
#pragma unroll
        for (int r = 1; r < RAD; ++r) {
          const T t = tw[(r * k * STEP) * TS];
          v[i][r] = (S < 0) ? cmul(v[i][r], t) : cmulc(v[i][r], t);
        }
      }
      dftR<RAD, S>(v[i]);
      const int base = (j / NS) * NS * RAD + k;
#pragma unroll
      for (int r = 0; r < RAD; ++r) lds[(base + r * NS) * SN + w] = v[i][r];
    }
  }
  __syncthreads();
}

template <typename T, int L, int NS, int S, int W, int SN, int NT, int TS>
__device__ __forceinline__ void fft_passes(T* lds, const T* tw, int tid) {
  if constexpr (NS < L) {
    constexpr int RAD = first_radix<L / NS>::value;
    static_assert(RAD != 0, "FFT length must be 2^a 3^b");
    stockham_pass<T, L, RAD, NS, S, W, SN, NT, TS>(lds, tw, tid);
    fft_passes<T, L, NS * RAD, S, W, SN, NT, TS>(lds, tw, tid);
  }
}

// ------------------------------------------------------------------------------------------
// k-space operator applied between the forward and inverse z transforms
// ------------------------------------------------------------------------------------------
struct GreenArgs {
  double coef;          // -3 Omega_m / (8 a) / N^3 (or any scalar prefactor)
  int from_grid;        // 0: analytic 1/sum sin^2(k/2); 1: from a float32 N^3 grid
  int deconvolve;       // divide by the CIC window of deposit and gather (prod sinc^4)
  const float* fgrid;   // dense [z][y][x] float32 grid when from_grid = 1
  // graph replay: when set, coef is table[3 * (*step) + 0] (the host value is ignored)
  const double* table = nullptr;
  const int* step = nullptr;
};

// A set of pencils of one pass: `nblk` blocks of W columns.  Contiguous segments hold W
// adjacent columns per block (outer index, column block); strided ones hold W columns with the
// FFT axis contiguous.
struct PencilSeg {
  long nblk;      // blocks in this segment
  long ncol_blk;  // column blocks per outer index (contiguous segments only)
  long outer_stride;
  long stride_n;
  long stride_w;  // strided segments: distance between columns
  int contig;     // 1: W columns contiguous; 0: the FFT axis is contiguous
  int green_kind; // 0: main H z-pass (kx = col, ky = outer); 1: Q z-pass (kx = N/2, ky = col)
  // page-tiled spectrum (kPencilTile x kPencilTile blocks of rows; 0: plain layout): outer index
  // o at (o / kPencilTile) outer_stride_hi + (o % kPencilTile) outer_stride, and row r + T of a
  // column (T = rows per load step, kPencilTile when tiled) stride_blk elements after row r
  long outer_stride_hi = 0;
  long stride_blk = 0;
};
constexpr int kPencilTile = 32;

// per-step scalars of a captured step: {green coefficient, kick, drift} per step
struct StepTable {
  const double* table = nullptr;
  const int* step = nullptr;
};

// per-size constant tables held in device memory
struct FftTables {
  int n = 0;
  void* tw = nullptr;     // N complex: exp(-2 pi i m / N), in the plan precision
  void* sin2 = nullptr;   // N reals: sin^2(pi m / N)
  void* wdec = nullptr;   // N reals: 1 / sinc^4(pi m / N) (m folded to [-N/2, N/2))
  int prec = 0;           // 4 or 8 (bytes per real)
};

// Supported grid sizes (2^a and 3 2^a from 8 to 1024); every kernel is instantiated for these.
#define PM_FOR_EACH_N(X) X(8) X(16) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)

bool fft_size_supported(int n);

// Host launchers (defined in pm_fft.hip).  All work on stream `st`.  rev (x passes): the row
// blocks are dealt to the XCDs in contiguous runs, each walked from its end, so the pass first
// reads what a producer walking the same runs forward wrote last; otherwise dispatch order.
template <typename R, typename Tin>
hipError_t fft_r2c_x(const Tin* rho, cpx<R>* H, cpx<R>* Q, int n, const FftTables& t, hipStream_t st,
                     bool rev = false);
// work (optional): an N^2 (N/2)-element array the step's y / z passes may use (at N >= 1024 the
// forward y pass leaves the spectrum there in a page-tiled layout for the z pass, and the inverse
// y pass brings it back); both y passes and the z pass of one transform get the same work array
template <typename R>
hipError_t fft_c2c_y(cpx<R>* H, cpx<R>* Q, int n, int sign, const FftTables& t, hipStream_t st,
                     cpx<R>* work = nullptr, bool alt = false);
template <typename R>
hipError_t fft_c2c_z(cpx<R>* H, cpx<R>* Q, int n, int sign, const FftTables& t, hipStream_t st);
template <typename R>
hipError_t fft_z_green(cpx<R>* H, cpx<R>* Q, int n, const GreenArgs& g, const FftTables& t, hipStream_t st,
                       cpx<R>* work = nullptr, bool alt = false);
// Columns per block of a pencil pass (mode -1: forward y, 0: z with the k-space operator, +1:
// inverse y) in its default shape, or with alt in its alternative one (0: none; the y / z
// launchers take `alt` to run it; results are those of the default shape)
int fft_pencil_cols(int n, int prec, int mode, bool alt);
template <typename R, typename Tout>
// edge (FP32 only, N % 32 == 0; may be null): also write the x margins of 32-wide tiles, see c2r_body
hipError_t fft_c2r_x(cpx<R>* H, const cpx<R>* Q, Tout* out, int n, const FftTables& t, hipStream_t st,
                     float4* edge = nullptr, bool rev = false);

}  // namespace pm
