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
  // the block's loads are issued in chunks of CH per thread before they are stored (the bound
  // test is compiled out when the rows fill whole chunks: a branch around each load would keep
  // the chunk in scratch memory, as it did for the FP64 C2R pass)
  constexpr int LPT = (RB * M + NT - 1) / NT;
  constexpr int CH = sizeof(R) == 4 ? 32 : 8;
#pragma unroll
  for (int i0 = 0; i0 < LPT; i0 += CH) {
    cpx<Tin> u[CH];
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const int e = (i0 + i) * NT + tid;
      if (i0 + i < LPT && ((RB * M) % NT == 0 || e < RB * M)) u[i] = src[e];
    }
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const int e = (i0 + i) * NT + tid;
      if (i0 + i < LPT && ((RB * M) % NT == 0 || e < RB * M)) {
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

// rev: the row blocks are dealt to the XCDs in contiguous runs, each walked from its end (the
// pass then reads first what a producer walking the same runs forward wrote last, which the
// last-level cache still holds); otherwise in dispatch order
template <typename R, typename Tin, int N, int RB>
__global__ __launch_bounds__(kThreads) void k_r2c_x(const Tin* __restrict__ rho, cpx<R>* __restrict__ H,
                                                    cpx<R>* __restrict__ Q, const cpx<R>* __restrict__ twg,
                                                    int rev) {
  __shared__ cpx<R> lds[(N / 2) * (RB + 1)];
  __shared__ cpx<R> tw[N];
  r2c_body<R, Tin, N, RB>(rho, H, Q, twg, rev ? xcd_block(blockIdx.x, gridDim.x, 1) : (int)blockIdx.x, lds, tw);
}

// ------------------------------------------------------------------------------------------
// x pass, complex -> real (in place allowed: H and out may be the same memory)
// ------------------------------------------------------------------------------------------
// edge (optional, FP32): the x margins of the gather's 32-wide tiles.  For every row (z, y) and
// every multiple of 32, b * 32, edge[(z * N/32 + b) * N + y] = (phi(32b - 1), phi(32b),
// phi(32b + 1), 0) (x periodic), so the tile b gather reads its left margin from entry b and its
// two right margins from entry b + 1 instead of three partial cache lines per potential row.
template <typename R, typename Tout, int N, int RB>
__device__ __forceinline__ void c2r_body(const cpx<R>* H, const cpx<R>* __restrict__ Q, Tout* out,
                                         const cpx<R>* __restrict__ twg, long blk, cpx<R>* lds, cpx<R>* tw,
                                         cpx<R>* qx, float4* __restrict__ edge = nullptr) {
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
      if (i0 + i < LPT && ((RB * M) % NT == 0 || e < RB * M)) u[i] = src[e];
    }
#pragma unroll
    for (int i = 0; i < CH; ++i) {
      const int e = (i0 + i) * NT + tid;
      if (i0 + i < LPT && ((RB * M) % NT == 0 || e < RB * M)) {
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
  if constexpr (sizeof(R) == 4 && sizeof(Tout) == 4 && N % 32 == 0) {
    if (edge) {
      constexpr int NB = N / 32;
      auto phi_at = [&](int x, int w) {
        const T v = lds[(x >> 1) * SN + w];
        return (x & 1) ? v.y : v.x;
      };
      for (int i = tid; i < RB * NB; i += NT) {
        const int w = i % RB, b = i / RB;
        const size_t row = row0 + w;
        const size_t z = row / N, y = row - z * N;
        const int x0 = 32 * b;
        edge[(z * NB + b) * N + y] = make_float4(phi_at(x0 == 0 ? N - 1 : x0 - 1, w), phi_at(x0, w),
                                                 phi_at(x0 + 1, w), 0.f);
      }
    }
  }
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
                                                    const cpx<R>* __restrict__ twg, float4* __restrict__ edge,
                                                    int rev) {
  __shared__ cpx<R> lds[(N / 2) * (RB + 1)];
  __shared__ cpx<R> tw[N];
  __shared__ cpx<R> qx[RB];
  c2r_body<R, Tout, N, RB>(H, Q, out, twg, rev ? xcd_block(blockIdx.x, gridDim.x, 1) : (int)blockIdx.x, lds, tw,
                           qx, edge);
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
template <typename R, int N, int W = pencil_w<R, N>()> constexpr int pencil_lds() {
  return (N * (W + 1) + N) * (int)sizeof(cpx<R>) + 2 * N * (int)sizeof(R);
}

// the k-space operator's scalar prefactor of this step
__device__ __forceinline__ double coef_of(const GreenArgs& g) { return g.table ? g.table[3 * (*g.step)] : g.coef; }

// v[m] *= W_N^{S f m} for m < M, from 8 + M/8 table reads: W^{f (8 hi + lo)} = W^{8 f hi} W^{f lo}
// (each factor is one table entry, so every twiddle carries one extra rounding at most)
template <int S, int N, int M, int G, typename T>
__device__ __forceinline__ void twiddle_row(T (&v)[M], const T* tw, int f) {
  constexpr int LO = (M < G ? M : G);
  T lo[LO];
#pragma unroll
  for (int r = 1; r < LO; ++r) lo[r] = tw[(f * r) % N];
#pragma unroll
  for (int hi = 0; hi < M; hi += LO) {
    const T h = tw[(f * hi) % N];
#pragma unroll
    for (int r = 0; r < LO; ++r) {
      if (hi == 0 && r == 0) continue;
      const T c = hi == 0 ? lo[r] : (r == 0 ? h : cmul(h, lo[r]));
      v[hi + r] = S < 0 ? cmul(v[hi + r], c) : cmulc(v[hi + r], c);
    }
  }
}

// Where pencil `blk` of a pass lives: its first element, the strides, and (for the k-space
// operator of the z pass) the kx / ky of its columns.
template <typename R> struct PencilAddr {
  cpx<R>* base;
  long stride_n, stride_w;
  long stride_blk;  // from row r to row r + kThreads / W of a column
  int contig;
  int kx0, dkx, ky0, dky;
  long tail;  // elements of the array from base to its end
};

template <typename R, int N, int W = pencil_w<R, N>()>
__device__ __forceinline__ PencilAddr<R> pencil_addr(cpx<R>* H, cpx<R>* Q, const PencilSeg& s0, const PencilSeg& s1,
                                                     long blk) {
  const bool first = blk < s0.nblk;
  const PencilSeg& s = first ? s0 : s1;
  if (!first) blk -= s0.nblk;
  cpx<R>* data = first ? H : Q;
  PencilAddr<R> p{};
  p.contig = s.contig;
  p.stride_n = s.stride_n;
  p.stride_w = s.stride_w;
  p.stride_blk = s.stride_blk ? s.stride_blk : (long)(kThreads / W) * (s.contig ? s.stride_n : 1);
  if (s.contig) {
    const long o = blk / s.ncol_blk, c = blk - o * s.ncol_blk;
    const long oo = s.outer_stride_hi ? (o / kPencilTile) * s.outer_stride_hi + (o % kPencilTile) * s.outer_stride
                                      : o * s.outer_stride;
    p.base = data + oo + c * W;
    if (s.green_kind == 0) {
      p.kx0 = (int)(c * W);
      p.dkx = 1;
      p.ky0 = (int)o;
    } else {
      p.kx0 = N / 2;
      p.ky0 = (int)(c * W);
      p.dky = 1;
    }
  } else {
    p.base = data + blk * (long)W * s.stride_w;
  }
  p.tail = (first ? (long)N * N * (N / 2) : (long)N * N) - (p.base - data);
  return p;
}

// Raw buffer access for the register pass: the resource holds the pencil's base and the bytes
// to the end of its array (out-of-range accesses read 0 and write nothing), so each element's
// address is a per-thread VGPR offset plus a block-uniform SGPR offset.
using u32x2 = unsigned int __attribute__((ext_vector_type(2)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const void* base, long bytes) {
  // 0x00020000: the buffer-resource format word of the gfx9 family (raw 32-bit dwords)
  // num_records is an unsigned 32-bit byte count; the register pass keeps every offset below 4 GiB
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0,
                                           (int)(uint32_t)(bytes < 0xffffffffL ? bytes : 0xffffffffL), 0x00020000);
}
__device__ __forceinline__ cpx<float> buffer_load(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(cpx<float>, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ void buffer_store(__amdgpu_buffer_rsrc_t r, int voff, int soff, cpx<float> v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
}

// elements of one pencil per thread
template <typename R, int N, int W = pencil_w<R, N>()> constexpr int pencil_lpt() { return (N * W + kThreads - 1) / kThreads; }

// element i of this thread: (row n, column w) of the pencil; contiguous pencils walk along w
template <typename R, int N, int W = pencil_w<R, N>()>
__device__ __forceinline__ bool pencil_elem(int contig, int i, int tid, int& n, int& w) {
  const int e = i * kThreads + tid;
  if (contig) {
    n = e / W;
    w = e - n * W;
  } else {
    w = e / N;
    n = e - w * N;
  }
  return (N * W) % kThreads == 0 || e < N * W;  // (a constant when the pencil fills whole rounds)
}

// global -> LDS, in chunks of CH loads per thread issued before any is stored.  In a
// contiguous pencil thread tid always handles column tid % W, and element i sits (NT / W) * i
// rows below its first.
template <typename R, int N, int W = pencil_w<R, N>()>
__device__ __forceinline__ void pencil_load(const PencilAddr<R>& p, cpx<R>* lds, int tid) {
  constexpr int SN = W + 1;
  constexpr int LPT = pencil_lpt<R, N, W>();
  constexpr int CH = sizeof(R) == 4 ? 32 : 8;
  if (p.contig) {
    const cpx<R>* src = p.base + (tid / W) * p.stride_n + (tid % W);
    cpx<R>* dst = lds + (tid / W) * SN + (tid % W);
    const long step = p.stride_blk;
#pragma unroll
    for (int i0 = 0; i0 < LPT; i0 += CH) {
      cpx<R> u[CH];
#pragma unroll
      for (int i = 0; i < CH; ++i)
        if (i0 + i < LPT && ((N * W) % kThreads == 0 || (i0 + i) * kThreads + tid < N * W))
          u[i] = src[(i0 + i) * step];
#pragma unroll
      for (int i = 0; i < CH; ++i)
        if (i0 + i < LPT && ((N * W) % kThreads == 0 || (i0 + i) * kThreads + tid < N * W))
          dst[(i0 + i) * (kThreads / W) * SN] = u[i];
    }
  } else {
#pragma unroll
    for (int i0 = 0; i0 < LPT; i0 += CH) {
      cpx<R> u[CH];
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        int n, w;
        if (i0 + i < LPT && pencil_elem<R, N, W>(0, i0 + i, tid, n, w)) u[i] = p.base[n + w * p.stride_w];
      }
#pragma unroll
      for (int i = 0; i < CH; ++i) {
        int n, w;
        if (i0 + i < LPT && pencil_elem<R, N, W>(0, i0 + i, tid, n, w)) lds[n * SN + w] = u[i];
      }
    }
  }
}

// Four-step transform of the staged pencil with the columns' DFTs in registers: N = A * TT,
// TT = threads per column.  Thread t of a column takes x[a TT + t] (a < A), does the A-point
// DFT and the twiddle W_N^{t ka}; after one exchange through LDS it takes, for each of its
// J = A / TT values ka = t + j TT, the TT values Y[t'][ka] and does the TT-point DFT, giving
// X[ka + A kb].  The inverse of the fused z pass runs the same steps backwards from there.
template <typename R, int N> constexpr bool reg_transform() {
  constexpr int W = pencil_w<R, N>();
  constexpr int TT = kThreads / W;
  // (the raw-buffer offsets of the register pass are unsigned 32-bit: a column spans at most
  // (N - 1) N (N/2) 8 bytes < 4 GiB for N <= 1024)
  return sizeof(R) == 4 && kThreads % W == 0 && N % TT == 0 && (N / TT) % TT == 0 && (TT == 16 || TT == 32) &&
         N / TT <= 32 && N <= 1024;
}

template <typename R, int N, int MODE>
__device__ __forceinline__ void pencil_transform_reg(const PencilAddr<R>& p, const GreenArgs& g, R coef, cpx<R>* lds,
                                                     const cpx<R>* tw, const R* s2, const R* wd, int tid) {
  using T = cpx<R>;
  constexpr int W = pencil_w<R, N>();
  constexpr int SN = W + 1;
  constexpr int TT = kThreads / W;
  constexpr int A = N / TT;
  constexpr int J = A / TT;
  constexpr int S1 = MODE == 1 ? +1 : -1;
  const int w = tid % W, t = tid / W;
  T v[A];
#pragma unroll
  for (int a = 0; a < A; ++a) v[a] = lds[(a * TT + t) * SN + w];
  reg_dft<S1, A>(v);
#pragma unroll
  for (int ka = 1; ka < A; ++ka) {
    const T c = tw[(t * ka) % N];
    v[ka] = S1 < 0 ? cmul(v[ka], c) : cmulc(v[ka], c);
  }
  __syncthreads();
#pragma unroll
  for (int ka = 0; ka < A; ++ka) lds[(ka * TT + t) * SN + w] = v[ka];
  __syncthreads();
  T u[J][TT];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int ka = t + j * TT;
#pragma unroll
    for (int q = 0; q < TT; ++q) u[j][q] = lds[(ka * TT + q) * SN + w];
    reg_dft<S1, TT>(u[j]);
  }
  if constexpr (MODE != 0) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
      for (int kb = 0; kb < TT; ++kb) lds[(t + j * TT + A * kb) * SN + w] = u[j][kb];
    __syncthreads();
  } else {
    // k-space operator on X[k], k = ka + A kb the z mode, then the inverse back to x
    const int kx = p.kx0 + w * p.dkx, ky = p.ky0 + w * p.dky;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int ka = t + j * TT;
#pragma unroll
      for (int kb = 0; kb < TT; ++kb) u[j][kb] = cscale(u[j][kb], green_factor<R>(kx, ky, ka + A * kb, N, g, coef, s2, wd));
      reg_dft<+1, TT>(u[j]);
      twiddle_row<+1, N, TT, 8>(u[j], tw, ka);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
      for (int q = 0; q < TT; ++q) lds[((t + j * TT) * TT + q) * SN + w] = u[j][q];
    __syncthreads();
#pragma unroll
    for (int ka = 0; ka < A; ++ka) v[ka] = lds[(ka * TT + t) * SN + w];
    reg_dft<+1, A>(v);
    __syncthreads();
#pragma unroll
    for (int a = 0; a < A; ++a) lds[(a * TT + t) * SN + w] = v[a];
    __syncthreads();
  }
}

// the transform(s) of a staged pencil, in LDS (ends with the LDS holding the result)
// Passes that run the register transform: every y and z pass up to N = 512; above that only
// the fused z pass (its LDS slice takes 82-90 KB, one workgroup per CU: the y passes are faster
// as LDS Stockham passes there, the z pass with its 4 MiB column stride is not)
template <typename R, int N, int MODE> constexpr bool reg_pass() { return reg_transform<R, N>() && (N <= 512 || MODE == 0); }

template <typename R, int N, int MODE, int W = pencil_w<R, N>()>
__device__ __forceinline__ void pencil_transform(const PencilAddr<R>& p, const GreenArgs& g, R coef, cpx<R>* lds,
                                                 const cpx<R>* tw, const R* s2, const R* wd, int tid) {
  using T = cpx<R>;
  constexpr int L = N;
  constexpr int SN = W + 1;
  constexpr int NT = kThreads;
  if constexpr (W == pencil_w<R, N>() && reg_pass<R, N, MODE>()) {
    pencil_transform_reg<R, N, MODE>(p, g, coef, lds, tw, s2, wd, tid);
  } else if constexpr (MODE == -1) {
    fft_passes<T, L, 1, -1, W, SN, NT, 1>(lds, tw, tid);
  } else if constexpr (MODE == 1) {
    fft_passes<T, L, 1, +1, W, SN, NT, 1>(lds, tw, tid);
  } else {
    fft_passes<T, L, 1, -1, W, SN, NT, 1>(lds, tw, tid);
#pragma unroll 4
    for (int e = tid; e < L * W; e += NT) {
      const int n = e / W, w = e - n * W;
      const R f = green_factor<R>(p.kx0 + w * p.dkx, p.ky0 + w * p.dky, n, N, g, coef, s2, wd);
      lds[n * SN + w] = cscale(lds[n * SN + w], f);
    }
    __syncthreads();
    fft_passes<T, L, 1, +1, W, SN, NT, 1>(lds, tw, tid);
  }
}

// LDS -> global
template <typename R, int N, int W = pencil_w<R, N>()>
__device__ __forceinline__ void pencil_store(const PencilAddr<R>& p, const cpx<R>* lds, int tid) {
  constexpr int SN = W + 1;
  constexpr int LPT = pencil_lpt<R, N, W>();
  if (p.contig) {
    cpx<R>* dst = p.base + (tid / W) * p.stride_n + (tid % W);
    const cpx<R>* src = lds + (tid / W) * SN + (tid % W);
    const long step = p.stride_blk;
#pragma unroll 8
    for (int i = 0; i < LPT; ++i)
      if ((N * W) % kThreads == 0 || i * kThreads + tid < N * W) dst[i * step] = src[i * (kThreads / W) * SN];
  } else {
#pragma unroll 4
    for (int i = 0; i < LPT; ++i) {
      int n, w;
      if (pencil_elem<R, N, W>(0, i, tid, n, w)) p.base[n + w * p.stride_w] = lds[n * SN + w];
    }
  }
}

template <typename R, int N, int MODE>
__device__ __forceinline__ void pencil_tables(const cpx<R>* __restrict__ twg, const R* __restrict__ s2g,
                                              const R* __restrict__ wdg, cpx<R>* tw, R* s2, R* wd, int tid) {
  for (int i = tid; i < N; i += kThreads) {
    tw[i] = twg[i];
    if constexpr (MODE == 0) {
      s2[i] = s2g[i];
      wd[i] = wdg[i];
    }
  }
}

// One pencil: fetch, transform, store.  MODE: -1 forward only, +1 inverse only, 0 forward +
// k-space operator + inverse.
// Hout / s0out (optional): the first segment's pencils are written to Hout with the layout s0out
// instead of back in place (the passes at N >= 1024 move the spectrum between layouts).
// TAB_LDS: the tables are copied to LDS (tw, s2, wd) or are the global tables themselves.
template <typename R, int N, int MODE, bool TAB_LDS = true, int W = pencil_w<R, N>()>
__device__ __forceinline__ void pencil_body(cpx<R>* H, cpx<R>* Q, const PencilSeg& s0, const PencilSeg& s1,
                                            const GreenArgs& g, const cpx<R>* __restrict__ twg,
                                            const R* __restrict__ s2g, const R* __restrict__ wdg, long blk,
                                            cpx<R>* lds, const cpx<R>* tw, const R* s2, const R* wd,
                                            cpx<R>* Hout = nullptr, const PencilSeg* s0out = nullptr) {
  const int tid = threadIdx.x;
  if constexpr (TAB_LDS)
    pencil_tables<R, N, MODE>(twg, s2g, wdg, const_cast<cpx<R>*>(tw), const_cast<R*>(s2), const_cast<R*>(wd), tid);
  const PencilAddr<R> p = pencil_addr<R, N, W>(H, Q, s0, s1, blk);
  pencil_load<R, N, W>(p, lds, tid);
  __syncthreads();
  const R coef = MODE == 0 ? R(g.table ? g.table[3 * (*g.step)] : g.coef) : R(0);
  pencil_transform<R, N, MODE, W>(p, g, coef, lds, tw, s2, wd, tid);
  if (Hout && blk < s0.nblk) pencil_store<R, N, W>(pencil_addr<R, N, W>(Hout, Q, *s0out, s1, blk), lds, tid);
  else pencil_store<R, N, W>(p, lds, tid);
}

template <typename R, int N, int MODE, int W = pencil_w<R, N>()>
__global__ __launch_bounds__(kThreads) void k_pencil(cpx<R>* H, cpx<R>* Q, PencilSeg s0, PencilSeg s1, GreenArgs g,
                                                     const cpx<R>* __restrict__ twg, const R* __restrict__ s2g,
                                                     const R* __restrict__ wdg, cpx<R>* Hout, PencilSeg s0out) {
  __shared__ cpx<R> lds[N * (W + 1)];
  __shared__ cpx<R> tw[N];  // (the Stockham passes read the twiddles in every stage: kept in LDS)
  __shared__ R s2[MODE == 0 ? N : 1];
  __shared__ R wd[MODE == 0 ? N : 1];
  pencil_body<R, N, MODE, true, W>(H, Q, s0, s1, g, twg, s2g, wdg, xcd_block(blockIdx.x, gridDim.x), lds, tw, s2, wd,
                                   Hout, &s0out);
}

// Register pass: the four-step transform with the columns read from and written to global
// memory directly, and only the exchange between its two DFT stages going through LDS, one
// slice of TT ka values at a time (TT*TT rows of W columns), so a block holds a fraction of its
// pencil in LDS and more blocks stay resident.  Row r of a slice sits at r*SN + r/TT + w: the
// extra word per TT rows spreads both the writers' and the readers' lanes over the banks.
template <typename R, int N> constexpr int xslice_elems() {
  constexpr int W = pencil_w<R, N>();
  constexpr int TT = kThreads / W;
  return TT * TT * (W + 1) + TT;
}

// TAB_LDS: the twiddle and k-space tables are copied to LDS (tw, s2, wd); otherwise tw, s2, wd
// are the global tables themselves (read through the caches), which keeps the block's LDS to the
// exchange slice
template <typename R, int N, int MODE, bool TAB_LDS = true>
__device__ __forceinline__ void pencil_reg_body(cpx<R>* H, cpx<R>* Q, const PencilSeg& s0, const PencilSeg& s1,
                                                const GreenArgs& g, const cpx<R>* __restrict__ twg,
                                                const R* __restrict__ s2g, const R* __restrict__ wdg, long blk,
                                                cpx<R>* xb, const cpx<R>* tw, const R* s2, const R* wd) {
  using T = cpx<R>;
  constexpr int W = pencil_w<R, N>();
  constexpr int SN = W + 1;
  constexpr int TT = kThreads / W;
  constexpr int A = N / TT;
  constexpr int J = A / TT;
  constexpr int S1 = MODE == 1 ? +1 : -1;
  const int tid = threadIdx.x;
  const int w = tid % W, t = tid / W;
  if constexpr (TAB_LDS)
    pencil_tables<R, N, MODE>(twg, s2g, wdg, const_cast<cpx<R>*>(tw), const_cast<R*>(s2), const_cast<R*>(wd), tid);
  const PencilAddr<R> p = pencil_addr<R, N>(H, Q, s0, s1, blk);
  // element (n, w) of the pencil at byte n * sn8 + (w * sw) * 8 from its base
  const uint32_t sn8 = (uint32_t)((p.contig ? p.stride_n : 1) * (long)sizeof(T));
  const __amdgpu_buffer_rsrc_t buf = buffer_of(p.base, p.tail * (long)sizeof(T));
  const int voff = (int)((uint32_t)t * sn8 + (uint32_t)w * (uint32_t)((p.contig ? 1 : p.stride_w) * (long)sizeof(T)));
  // byte offset of row r (a multiple of TT) of the column (unsigned: up to 4 GiB at N = 1024)
  const uint32_t blk8 = (uint32_t)(p.stride_blk * (long)sizeof(T));
  auto roff = [&](int r) { return (int)((uint32_t)(r / TT) * blk8); };
  auto xrow = [&](int r) { return r * SN + r / TT + w; };
  T v[A];
#pragma unroll
  for (int a = 0; a < A; ++a) v[a] = buffer_load(buf, voff, roff(a * TT));
  __syncthreads();  // tables in LDS
  reg_dft<S1, A>(v);
  twiddle_row<S1, N, A, 8>(v, tw, t);
  T u[J][TT];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    if (j) __syncthreads();
#pragma unroll
    for (int q = 0; q < TT; ++q) xb[xrow(q * TT + t)] = v[j * TT + q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TT; ++q) u[j][q] = xb[xrow(t * TT + q)];
    reg_dft<S1, TT>(u[j]);  // u[j][kb] = X[ka + A kb], ka = j TT + t
  }
  if constexpr (MODE != 0) {
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
      for (int kb = 0; kb < TT; ++kb) buffer_store(buf, voff, roff(j * TT + A * kb), u[j][kb]);
  } else {
    // k-space operator on X[k], k = ka + A kb the z mode, then the same steps backwards
    const int kx = p.kx0 + w * p.dkx, ky = p.ky0 + w * p.dky;
    const R coef = R(coef_of(g));
    if (g.from_grid || g.deconvolve) {
#pragma unroll
      for (int j = 0; j < J; ++j)
#pragma unroll
        for (int kb = 0; kb < TT; ++kb)
          u[j][kb] = cscale(u[j][kb], green_factor<R>(kx, ky, j * TT + t + A * kb, N, g, coef, s2, wd));
    } else {  // the plain Green's function: green_factor() without its optional terms
      const R sxy = s2[kx] + s2[ky];
#pragma unroll
      for (int j = 0; j < J; ++j)
#pragma unroll
        for (int kb = 0; kb < TT; ++kb) {
          const R den = sxy + s2[j * TT + t + A * kb];
          u[j][kb] = cscale(u[j][kb], den > R(0) ? coef / den : R(0));
        }
    }
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const int ka = j * TT + t;
      reg_dft<+1, TT>(u[j]);
#pragma unroll
      for (int q = 1; q < TT; ++q) u[j][q] = cmulc(u[j][q], tw[(q * ka) % N]);
    }
#pragma unroll
    for (int j = 0; j < J; ++j) {
      __syncthreads();
#pragma unroll
      for (int q = 0; q < TT; ++q) xb[xrow(t * TT + q)] = u[j][q];
      __syncthreads();
#pragma unroll
      for (int r = 0; r < TT; ++r) v[j * TT + r] = xb[xrow(r * TT + t)];
    }
    reg_dft<+1, A>(v);
#pragma unroll
    for (int a = 0; a < A; ++a) buffer_store(buf, voff, roff(a * TT), v[a]);
  }
}

template <typename R, int N, int MODE>
__global__ __launch_bounds__(kThreads) void k_pencil_reg(cpx<R>* H, cpx<R>* Q, PencilSeg s0, PencilSeg s1, GreenArgs g,
                                                         const cpx<R>* __restrict__ twg, const R* __restrict__ s2g,
                                                         const R* __restrict__ wdg) {
  __shared__ cpx<R> xb[xslice_elems<R, N>()];
  // above N = 512 the tables stay in global memory: with them the block's LDS (90 KB at
  // N = 1024) allows one block per CU, without them two
  constexpr bool kTabLds = N <= 512;
  __shared__ cpx<R> tw[kTabLds ? N : 1];
  __shared__ R s2[kTabLds && MODE == 0 ? N : 1];
  __shared__ R wd[kTabLds && MODE == 0 ? N : 1];
  if constexpr (kTabLds)
    pencil_reg_body<R, N, MODE, true>(H, Q, s0, s1, g, twg, s2g, wdg, xcd_block(blockIdx.x, gridDim.x), xb, tw, s2,
                                      wd);
  else
    pencil_reg_body<R, N, MODE, false>(H, Q, s0, s1, g, twg, s2g, wdg, xcd_block(blockIdx.x, gridDim.x), xb, twg,
                                       s2g, wdg);
}

// ------------------------------------------------------------------------------------------
// host-side segment descriptions
// ------------------------------------------------------------------------------------------
template <typename R, int N, int W = pencil_w<R, N>()> void y_segments(PencilSeg& a, PencilSeg& b) {
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
template <typename R, int N, int W = pencil_w<R, N>()> void z_segments(PencilSeg& a, PencilSeg& b) {
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
hipError_t launch_r2c(const Tin* rho, cpx<R>* H, cpx<R>* Q, const FftTables& t, hipStream_t st, bool rev) {
  constexpr int RB = rows_per_block<R, N>();
  static_assert(((long)N * N) % RB == 0, "rows per block must divide N^2");
  k_r2c_x<R, Tin, N, RB><<<(long)N * N / RB, kThreads, 0, st>>>(rho, H, Q, (const cpx<R>*)t.tw, rev ? 1 : 0);
  return hipGetLastError();
}
template <typename R, typename Tout, int N>
hipError_t launch_c2r(cpx<R>* H, const cpx<R>* Q, Tout* out, const FftTables& t, hipStream_t st, float4* edge,
                      bool rev) {
  constexpr int RB = rows_per_block<R, N>();
  k_c2r_x<R, Tout, N, RB><<<(long)N * N / RB, kThreads, 0, st>>>(H, Q, out, (const cpx<R>*)t.tw, edge, rev ? 1 : 0);
  return hipGetLastError();
}
// Page-tiled spectrum for N >= 1024 (FP32).  The plain layout puts consecutive z of a column
// 4 MiB apart, so a z pencil touches one 2 MiB page per row and the pass is bound by address
// translation (2/3 of its translations missed).  In the tiled layout, blocks of T x T rows
// (T = kPencilTile) are contiguous: element (z, y, kx) sits at
//   (z / T) N T M + (y / T) T^2 M + (y % T) T M + (z % T) M + kx        (M = N / 2),
// so a z column spans N / T pages and a y column 2 N / T.  The step's forward y pass reads the
// plain spectrum and writes the tiled one to a work array, the z pass runs there, and the
// inverse y pass writes the plain layout back; the x passes and the Nyquist plane keep theirs.
template <typename R, int N> constexpr bool tiled_spectrum() {
  return sizeof(R) == 4 && N >= 1024 && kThreads / pencil_w<R, N>() == kPencilTile;
}
template <typename R, int N> PencilSeg y_tiled(PencilSeg a) {  // outer z, rows y
  constexpr long M = N / 2, T = kPencilTile;
  a.outer_stride = M;
  a.outer_stride_hi = (long)N * T * M;
  a.stride_n = T * M;
  a.stride_blk = T * T * M;
  return a;
}
template <typename R, int N> PencilSeg z_tiled(PencilSeg a) {  // outer y, rows z
  constexpr long M = N / 2, T = kPencilTile;
  a.outer_stride = T * M;
  a.outer_stride_hi = T * T * M;
  a.stride_n = M;
  a.stride_blk = (long)N * T * M;
  return a;
}

// The alternative shape of a Stockham pencil pass: half its default columns per block (less LDS,
// more blocks per CU), for N >= 64 in the plain layout; 0 where there is none (register passes,
// the page-tiled passes, widths below 2).
template <typename R, int N, int MODE> constexpr int pencil_alt_w() {
  constexpr int W2 = pencil_w<R, N>() / 2;
  if constexpr (reg_pass<R, N, MODE>() || tiled_spectrum<R, N>() || N < 64 || W2 < 2) return 0;
  else return ((N / 2) % W2 == 0 && N % W2 == 0 && kThreads % W2 == 0) ? W2 : 0;
}

// work (optional, N^2 (N/2) elements): the tiled work array of the step's y / z passes
// alt: run the pass with its alternative shape (pencil_alt_w; hipErrorInvalidValue if none)
template <typename R, int N, int MODE>
hipError_t launch_pencil(cpx<R>* H, cpx<R>* Q, bool zaxis, const GreenArgs& g, const FftTables& t,
                         hipStream_t st, cpx<R>* work = nullptr, bool alt = false) {
  const cpx<R>* tw = (const cpx<R>*)t.tw;
  const R* s2 = (const R*)t.sin2;
  const R* wd = (const R*)t.wdec;
  if (alt) {
    constexpr int W2 = pencil_alt_w<R, N, MODE>();
    if constexpr (W2 > 0) {
      PencilSeg a2, b2;
      if (zaxis) z_segments<R, N, W2>(a2, b2);
      else y_segments<R, N, W2>(a2, b2);
      k_pencil<R, N, MODE, W2><<<(unsigned)(a2.nblk + b2.nblk), kThreads, 0, st>>>(H, Q, a2, b2, g, tw, s2, wd,
                                                                                   nullptr, a2);
      return hipGetLastError();
    } else {
      return hipErrorInvalidValue;
    }
  }
  PencilSeg a, b;
  if (zaxis) z_segments<R, N>(a, b);
  else y_segments<R, N>(a, b);
  const unsigned nb = (unsigned)(a.nblk + b.nblk);
  if constexpr (tiled_spectrum<R, N>()) {
    if (work) {
      if (zaxis) {  // in place on the tiled array
        const PencilSeg at = z_tiled<R, N>(a);
        if constexpr (reg_pass<R, N, MODE>())
          k_pencil_reg<R, N, MODE><<<nb, kThreads, 0, st>>>(work, Q, at, b, g, tw, s2, wd);
        else
          k_pencil<R, N, MODE><<<nb, kThreads, 0, st>>>(work, Q, at, b, g, tw, s2, wd, nullptr, at);
      } else if (MODE < 0) {  // plain H -> tiled work
        k_pencil<R, N, MODE><<<nb, kThreads, 0, st>>>(H, Q, a, b, g, tw, s2, wd, work, y_tiled<R, N>(a));
      } else {  // tiled work -> plain H
        k_pencil<R, N, MODE><<<nb, kThreads, 0, st>>>(work, Q, y_tiled<R, N>(a), b, g, tw, s2, wd, H, a);
      }
      return hipGetLastError();
    }
  }
  if constexpr (reg_pass<R, N, MODE>()) {
    k_pencil_reg<R, N, MODE><<<nb, kThreads, 0, st>>>(H, Q, a, b, g, tw, s2, wd);
  } else {
    k_pencil<R, N, MODE><<<nb, kThreads, 0, st>>>(H, Q, a, b, g, tw, s2, wd, nullptr, a);
  }
  return hipGetLastError();
}

}  // namespace

namespace {
template <typename R, int N> int pencil_cols(int mode, bool alt) {
  const int w2 = mode < 0 ? pencil_alt_w<R, N, -1>() : mode == 0 ? pencil_alt_w<R, N, 0>() : pencil_alt_w<R, N, 1>();
  return alt ? w2 : pencil_w<R, N>();
}
}  // namespace

int fft_pencil_cols(int n, int prec, int mode, bool alt) {
  switch (n) {
#define PM_CASE(NN) case NN: return prec == 4 ? pencil_cols<float, NN>(mode, alt) : pencil_cols<double, NN>(mode, alt);
    PM_FOR_EACH_N(PM_CASE)
#undef PM_CASE
  }
  return 0;
}

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
hipError_t fft_r2c_x(const Tin* rho, cpx<R>* H, cpx<R>* Q, int n, const FftTables& t, hipStream_t st, bool rev) {
  switch (n) {
#define PM_CASE(NN) case NN: return launch_r2c<R, Tin, NN>(rho, H, Q, t, st, rev);
    PM_FOR_EACH_N(PM_CASE)
#undef PM_CASE
  }
  return hipErrorInvalidValue;
}
template <typename R, typename Tout>
hipError_t fft_c2r_x(cpx<R>* H, const cpx<R>* Q, Tout* out, int n, const FftTables& t, hipStream_t st,
                     float4* edge, bool rev) {
  switch (n) {
#define PM_CASE(NN) case NN: return launch_c2r<R, Tout, NN>(H, Q, out, t, st, edge, rev);
    PM_FOR_EACH_N(PM_CASE)
#undef PM_CASE
  }
  return hipErrorInvalidValue;
}
template <typename R>
hipError_t fft_c2c_y(cpx<R>* H, cpx<R>* Q, int n, int sign, const FftTables& t, hipStream_t st, cpx<R>* work,
                     bool alt) {
  GreenArgs g{};
  switch (n) {
#define PM_CASE(NN)                                                                   \
  case NN:                                                                            \
    return sign < 0 ? launch_pencil<R, NN, -1>(H, Q, false, g, t, st, work, alt)      \
                    : launch_pencil<R, NN, 1>(H, Q, false, g, t, st, work, alt);
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
hipError_t fft_z_green(cpx<R>* H, cpx<R>* Q, int n, const GreenArgs& g, const FftTables& t, hipStream_t st,
                       cpx<R>* work, bool alt) {
  switch (n) {
#define PM_CASE(NN) case NN: return launch_pencil<R, NN, 0>(H, Q, true, g, t, st, work, alt);
    PM_FOR_EACH_N(PM_CASE)
#undef PM_CASE
  }
  return hipErrorInvalidValue;
}

template hipError_t fft_r2c_x<float, float>(const float*, cpx<float>*, cpx<float>*, int, const FftTables&,
                                             hipStream_t, bool);
template hipError_t fft_r2c_x<double, double>(const double*, cpx<double>*, cpx<double>*, int, const FftTables&,
                                               hipStream_t, bool);
template hipError_t fft_r2c_x<double, float>(const float*, cpx<double>*, cpx<double>*, int, const FftTables&,
                                              hipStream_t, bool);
template hipError_t fft_c2r_x<float, float>(cpx<float>*, const cpx<float>*, float*, int,
                                                   const FftTables&, hipStream_t, float4*, bool);
template hipError_t fft_c2r_x<double, double>(cpx<double>*, const cpx<double>*, double*, int,
                                                   const FftTables&, hipStream_t, float4*, bool);
template hipError_t fft_c2r_x<double, float>(cpx<double>*, const cpx<double>*, float*, int,
                                                   const FftTables&, hipStream_t, float4*, bool);
#define PM_INST(R)                                                                                   \
  template hipError_t fft_c2c_y<R>(cpx<R>*, cpx<R>*, int, int, const FftTables&, hipStream_t, cpx<R>*, bool); \
  template hipError_t fft_c2c_z<R>(cpx<R>*, cpx<R>*, int, int, const FftTables&, hipStream_t);       \
  template hipError_t fft_z_green<R>(cpx<R>*, cpx<R>*, int, const GreenArgs&, const FftTables&, hipStream_t,  \
                                     cpx<R>*, bool);
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

// The batched x passes walk every XCD's run of blocks backwards, as the single-sim step's do:
// each reads first what its producer (deposit, inverse y pass) wrote last (48-member ensemble:
// 5.25 -> 5.23 ms per step)
template <int CLS>
__global__ __launch_bounds__(kThreads) void k_batch_r2c(const BatchDesc* __restrict__ D, long base) {
  extern __shared__ __align__(16) unsigned char smem[];
  const long b = base + xcd_block(blockIdx.x, gridDim.x, 1);
  const int k = batch_member(D, BS_R2C, b);
  const BatchMember& m = D->m[k];
  const long blk = b - D->off[BS_R2C][k];
  cpx<float>* lds = reinterpret_cast<cpx<float>*>(smem);
#define PM_CASE(NN)                                                                              \
  case NN:                                                                                       \
    r2c_body<float, float, NN, x_rb<NN>()>(gmem(m.rho), gmem(m.H), gmem(m.Q), gmem(m.tw), blk, lds, \
                                           lds + (NN / 2) * (x_rb<NN>() + 1));                   \
    break;
  PM_BATCH_CLASS_SWITCH(CLS, m.n, PM_CASE)
#undef PM_CASE
}

template <int CLS>
__global__ __launch_bounds__(kThreads) void k_batch_c2r(const BatchDesc* __restrict__ D, long base) {
  extern __shared__ __align__(16) unsigned char smem[];
  const long b = base + xcd_block(blockIdx.x, gridDim.x, 1);
  const int k = batch_member(D, BS_C2R, b);
  const BatchMember& m = D->m[k];
  const long blk = b - D->off[BS_C2R][k];
  cpx<float>* lds = reinterpret_cast<cpx<float>*>(smem);
#define PM_CASE(NN)                                                                              \
  case NN: {                                                                                     \
    cpx<float>* tw = lds + (NN / 2) * (x_rb<NN>() + 1);                                          \
    c2r_body<float, float, NN, x_rb<NN>()>(gmem(m.H), gmem(m.Q), gmem(reinterpret_cast<float*>(m.H)), \
                                           gmem(m.tw), blk, lds, tw, tw + NN);                   \
    break;                                                                                       \
  }
  PM_BATCH_CLASS_SWITCH(CLS, m.n, PM_CASE)
#undef PM_CASE
}

// the pencil passes walk their blocks in XCD order (adjacent column chunks share cache lines);
// the x passes keep dispatch order, which leaves the spectrum where the next pass reads it
template <int CLS, int MODE>
__global__ __launch_bounds__(kThreads) void k_batch_pencil(const BatchDesc* __restrict__ D, long base) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int ST = MODE == 0 ? BS_Z : BS_Y;
  const long b = base + xcd_block(blockIdx.x, gridDim.x);
  const int k = batch_member(D, ST, b);
  const BatchMember& m = D->m[k];
  const long blk = b - D->off[ST][k];
  GreenArgs g{};
  if (MODE == 0) {
    g.coef = gmem(D->table)[3 * ((long)(*gmem(D->step)) * D->n_members + k)];
    g.deconvolve = m.deconvolve;
    g.from_grid = m.from_grid;
    g.fgrid = gmem(m.fgrid);
  }
  cpx<float>* lds = reinterpret_cast<cpx<float>*>(smem);
#define PM_CASE(NN)                                                                              \
  case NN: {                                                                                     \
    if constexpr (reg_transform<float, NN>()) {                                                  \
      cpx<float>* tw = lds + xslice_elems<float, NN>();                                          \
      float* s2 = reinterpret_cast<float*>(tw + NN);                                             \
      pencil_reg_body<float, NN, MODE>(gmem(m.H), gmem(m.Q), MODE == 0 ? m.za : m.ya,            \
                                       MODE == 0 ? m.zb : m.yb, g, gmem(m.tw), gmem(m.sin2),      \
                                       gmem(m.wdec), blk, lds, tw, s2, s2 + NN);                 \
    } else {                                                                                     \
      cpx<float>* tw = lds + NN * (pencil_w<float, NN>() + 1);                                   \
      float* s2 = reinterpret_cast<float*>(tw + NN);                                             \
      pencil_body<float, NN, MODE>(gmem(m.H), gmem(m.Q), MODE == 0 ? m.za : m.ya,                \
                                   MODE == 0 ? m.zb : m.yb, g, gmem(m.tw), gmem(m.sin2),          \
                                   gmem(m.wdec), blk, lds, tw, s2, s2 + NN);                     \
    }                                                                                            \
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
    if (reg_transform<float, NN>())                                                                  \
      return (xslice_elems<float, NN>() + NN) * 8 + 2 * NN * 4;                                       \
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
