// Shared device/host helpers for libpm (particle-mesh step on gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace pm {

// ------------------------------------------------------------------------------------------
// complex arithmetic on HIP's native 2-vectors (float2 / double2)
// ------------------------------------------------------------------------------------------
template <typename R> struct cpx_of;
template <> struct cpx_of<float> { using type = float2; };
template <> struct cpx_of<double> { using type = double2; };
template <typename R> using cpx = typename cpx_of<R>::type;

template <typename T> __host__ __device__ __forceinline__ T cmake(decltype(T::x) re, decltype(T::x) im) {
  T r; r.x = re; r.y = im; return r;
}
template <typename T> __host__ __device__ __forceinline__ T cadd(T a, T b) { return cmake<T>(a.x + b.x, a.y + b.y); }
template <typename T> __host__ __device__ __forceinline__ T csub(T a, T b) { return cmake<T>(a.x - b.x, a.y - b.y); }
template <typename T> __host__ __device__ __forceinline__ T cmul(T a, T b) {
  return cmake<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename T> __host__ __device__ __forceinline__ T cmulc(T a, T b) {
  return cmake<T>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T> __host__ __device__ __forceinline__ T cconj(T a) { return cmake<T>(a.x, -a.y); }
template <typename T, typename S> __host__ __device__ __forceinline__ T cscale(T a, S s) {
  return cmake<T>(a.x * s, a.y * s);
}
// multiply by S*i, S = +1 or -1:  (x + iy) * (S i) = -S y + i S x
template <int S, typename T> __host__ __device__ __forceinline__ T cmul_si(T a) {
  return S > 0 ? cmake<T>(-a.y, a.x) : cmake<T>(a.y, -a.x);
}

// ------------------------------------------------------------------------------------------
// periodic helpers
// ------------------------------------------------------------------------------------------
__host__ __device__ __forceinline__ int wrapi(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}

// Canonical wrap of a position into [0, n): one add/subtract of n, and a value that
// rounds to exactly n after adding n maps to 0.  Inputs lie in (-n, 2n).
template <typename R> __host__ __device__ __forceinline__ R wrap_pos(R x, R n) {
  if (x < R(0)) x += n;
  if (x >= n) x -= n;
  return x;
}

// Canonical position of a particle stored as lattice point q in [0, n) plus displacement s in
// [-n/2, n/2).  When q + s leaves [0, n) the periodic image of q (q - n or q + n, both exact)
// is added instead, so a particle near 0 whose lattice point is near n keeps the precision of
// its own coordinate rather than that of n.  Every kernel and the host materialisation use
// this one expression, so equal stored states give bit-identical positions and cells.
template <typename R> __host__ __device__ __forceinline__ R canonical_pos(R q, R s, R n) {
  const R u = q + s;
  R x = u;
  if (u >= n) x = (q - n) + s;
  else if (u < R(0)) x = (q + n) + s;
  return wrap_pos<R>(x, n);
}

// Cell index floor(x) clamped to [0, n-1].  The clamp happens in floating point before the
// conversion, so NaN or out-of-range inputs still give an in-bounds index.
template <typename R> __host__ __device__ __forceinline__ int cell_of(R x, int n) {
  const R hi = R(n - 1);
  R c = x > R(0) ? x : R(0);  // NaN -> 0
  c = c < hi ? c : hi;
  return (int)floor(c);
}

// Keep a displacement within [-n/2, n/2): exact in floating point for |s| < 3n/2.
template <typename R> __host__ __device__ __forceinline__ R wrap_disp(R s, R n) {
  const R h = n * R(0.5);
  if (s >= h) s -= n;
  else if (s < -h) s += n;
  return s;
}

}  // namespace pm

// status codes returned by every C-ABI entry point (0 = success)
// XCD-aware block order.  Workgroups are dealt round-robin to the 8 XCDs of the chip, each with
// its own L2; this maps dispatch index b of an nblocks-block launch to a logical block so that
// every XCD works through one contiguous run of logical blocks, and neighbours that share cache
// lines (adjacent tiles, adjacent column chunks of an FFT pass) use one L2.
// A pointer read from device memory (a descriptor) recast as a global-memory pointer: the
// compiler cannot tell such a pointer from one into LDS or scratch and emits FLAT accesses,
// which also count against the LDS wait counter (so waiting on LDS waits on memory too); through
// this cast it emits global accesses.  Only for pointers into device allocations (or null).
template <class T> __device__ __forceinline__ T* gmem(T* p) {
  return (T*)(__attribute__((address_space(1))) T*)(uintptr_t)p;
}

constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int b, int nblocks) {
  const int per = nblocks / kXcds, rem = nblocks % kXcds;
  const int xcd = b % kXcds, local = b / kXcds;
  return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}
// rev: every XCD walks its run from the end instead.  Kernels that consume the previous
// kernel's output alternate directions, so each reads first what its producer (walking the same
// runs the other way) wrote last, which is still in the 256 MB last-level cache.
__device__ __forceinline__ int xcd_block(int b, int nblocks, int rev) {
  const int per = nblocks / kXcds, rem = nblocks % kXcds;
  const int xcd = b % kXcds, local = b / kXcds;
  const int len = xcd < rem ? per + 1 : per;
  const int start = xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
  return start + (rev ? len - 1 - local : local);
}

#define PM_OK 0
#define PM_E_INVALID 1      // bad argument (size, pointer, unsupported grid)
#define PM_E_HIP 2          // a HIP runtime call failed
#define PM_E_NOMEM 3        // device allocation failed
#define PM_E_STATE 4        // call out of order (e.g. step before state is set)
#define PM_E_UNSUPPORTED 5  // feature not available for this configuration
#define PM_E_CORRUPT 6      // the step found non-finite particles or inconsistent tile bins

// Health counters a sim keeps in device memory; kernels add to them only when they find a
// problem, and the host checks them at the synchronisation that ends a step or a run.
#define PM_HEALTH_NONFINITE 0  // particles whose new position or velocity is not finite
#define PM_HEALTH_BINS 1       // own entries whose cell lies outside their tile (stale bins)
#define PM_HEALTH_OVERFLOW 2   // entries that found no place in their bin (count mismatch)
#define PM_HEALTH_COUNT 4
