// Eulerian tiles for the PM step: per-step counting sort of particle slots into grid tiles,
// owner-computes CIC deposit and the tile-staged gather + kick + drift (gfx950).
//
// Each tile of TX x TY x TZ cells owns two segments of particle entries (storage indices):
//   own  : particles whose cell floor(x) lies in the tile;
//   halo : particles of the lower neighbour tiles whose +1 CIC corner falls in this tile
//          (a particle on the upper face of its tile is entered once more for every tile its
//          cloud reaches, at most 8 entries in all).
// All own segments come first, in tile order, then all halo segments: the own entries are a
// permutation of the particles sorted by tile, and the gather writes each particle's new state
// at its own-entry index into the other state buffer, so the stored particles follow the tiles
// from one step to the next (their lattice ids travel with them).
// The deposit of a tile reads own + halo entries, accumulates the in-tile corners in LDS and
// writes every cell of the tile with plain stores: no global float atomics, no zero fill.
// The LDS accumulator is FP64 in both precisions: on gfx950 ds_add_f64 sustains about twice
// the rate of ds_add_f32 for this access pattern (measured 0.33 ms against 0.76 ms per
// 256^3-particle deposit), and the sums are then rounded once instead of at every add.
// The gather reads the own entries with the tile's potential (plus a -1/+2 stencil margin)
// staged in LDS, and counts the new positions through an LDS histogram of each tile's 4^3 tile
// neighbourhood (binning from positions alone counts with a wave-aggregated atomic per distinct
// key).  A two-kernel exclusive scan turns the counts into offsets; after a gather the tile
// scatter writes the entries through LDS cursors from ranges the gather reserved, otherwise the
// particle-parallel scatter through device cursors.
#include <algorithm>
#include <cmath>
#include <vector>

#include "pm_batch.h"

namespace pm {

namespace {

constexpr int kT = 256;  // threads per block
constexpr uint32_t kNone = 0xffffffffu;  // no entry / empty slot

inline unsigned blocks_for(long n, int b = kT) {
  const long g = (n + b - 1) / b;
  return (unsigned)(g < 1 ? 1 : g);
}


template <typename R> __device__ __forceinline__ bool is_finite3(R a, R b, R c) {
  return __builtin_isfinite(a) && __builtin_isfinite(b) && __builtin_isfinite(c);
}

// index in [-n, 2n) folded into [0, n)
__device__ __forceinline__ int fold(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// XCD-aware block -> tile map (tiles that share potential-box lines or straddling particles
// are processed by one L2)
__device__ __forceinline__ int xcd_tile(int b, int nblocks) { return xcd_block(b, nblocks); }

// canonical position of a resident particle from its displacement: cell, CIC fractions and
// lattice point q per axis
template <typename R>
__device__ __forceinline__ void cell_and_frac(R sx, R sy, R sz, uint32_t id, const Lattice& L, int c[3], R d[3],
                                              R q[3]) {
  int ijk[3];
  slot_to_ijk(id, L, ijk[0], ijk[1], ijk[2]);
  const R sv[3] = {sx, sy, sz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    q[a] = lattice_coord<R>(ijk[a], L);
    const R x = canonical_pos<R>(q[a], sv[a], R(L.n));
    c[a] = cell_of<R>(x, L.n);
    d[a] = x - R(c[a]);
  }
}

// particle at storage index e, lattice id `id`
template <typename R>
__device__ __forceinline__ void cell_and_frac(const R* __restrict__ s, long e, uint32_t id, const Lattice& L, int c[3],
                                              R d[3]) {
  R q[3];
  cell_and_frac<R>(s[e], s[L.np + e], s[2 * L.np + e], id, L, c, d, q);
}

// home tile of a cell and whether the cloud reaches the next tile along each axis
template <int TX, int TY, int TZ>
__device__ __forceinline__ void tile_of(const int c[3], const TileGeom& G, int t[3], int up[3]) {
  t[0] = c[0] / TX;
  t[1] = c[1] / TY;
  t[2] = c[2] / TZ;
  // with a single tile along an axis the wrapped corner stays in the home tile
  up[0] = (c[0] % TX == TX - 1) && G.ntx > 1;
  up[1] = (c[1] % TY == TY - 1) && G.nty > 1;
  up[2] = (c[2] % TZ == TZ - 1) && G.ntz > 1;
}

// key of corner-offset o (bit 0: x, 1: y, 2: z) of a particle in tile t: the tile (own
// segment) for o = 0, ntiles + the tile (halo segment) otherwise
__device__ __forceinline__ uint32_t target_key(const int t[3], int o, const TileGeom& G) {
  int x = t[0] + (o & 1), y = t[1] + ((o >> 1) & 1), z = t[2] + (o >> 2);
  if (x == G.ntx) x = 0;
  if (y == G.nty) y = 0;
  if (z == G.ntz) z = 0;
  const uint32_t tile = (uint32_t)((z * G.nty + y) * G.ntx + x);
  return tile + (o != 0 ? (uint32_t)G.ntiles : 0u);
}

__device__ __forceinline__ bool target_active(const int up[3], int o) {
  return ((o & 1) <= up[0]) && (((o >> 1) & 1) <= up[1]) && ((o >> 2) <= up[2]);
}

// Neighbourhood bins of a source tile: the 4^3 tiles at offsets -1..2 per axis, x {own, halo}
// keys, 128 bins (bin = 2 * ((rz * 4 + ry) * 4 + rx) + kind).  A particle whose new home tile lies
// at offset -1..1 has every corner target inside the neighbourhood; near_home_bin gives the bin
// of its own key (false: moved further) and corner_bin that of corner offset o.
__device__ __forceinline__ bool near_home_bin(const int tn[3], const int tc[3], const TileGeom& G, int& h0) {
  const int ntc[3] = {G.ntx, G.nty, G.ntz};
  int rb[3];
  bool near = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    int dlt = tn[a] - tc[a];
    if (dlt > 2) dlt -= ntc[a];
    if (dlt < -1) dlt += ntc[a];
    rb[a] = dlt + 1;
    near = near && dlt >= -1 && dlt <= 1;
  }
  h0 = 2 * ((rb[2] * 4 + rb[1]) * 4 + rb[0]);
  return near;
}
__device__ __forceinline__ int corner_bin(int h0, int o) {
  return o == 0 ? h0 : h0 + 2 * (((o >> 2) * 4 + ((o >> 1) & 1)) * 4 + (o & 1)) + 1;
}
// key of neighbourhood bin `bin` of the tile at (tix, tiy, tiz)
__device__ __forceinline__ uint32_t bin_key(int tix, int tiy, int tiz, int bin, const TileGeom& G) {
  const int cell = bin >> 1, kind = bin & 1;
  const int rx = cell & 3, ry = (cell >> 2) & 3, rz = cell >> 4;
  const int x = fold(tix + rx - 1, G.ntx), y = fold(tiy + ry - 1, G.nty), z = fold(tiz + rz - 1, G.ntz);
  return (uint32_t)((z * G.nty + y) * G.ntx + x) + (kind ? (uint32_t)G.ntiles : 0u);
}
constexpr int kBins = 128;  // neighbourhood bins per source tile
// Tile scatter arrays, laid out in the counts buffer after the 2*ntiles counts of the current
// tile shape: 2*ntiles far counts (particles that left their source tile's neighbourhood), 128
// reserved range starts per tile (kNone: empty bin), the own offsets of the binning the gather
// read (ntiles + 1; the scan keeps them before it overwrites them: the source range of tile t,
// where the gather wrote its particles, is [old[t], old[t+1])).  The 16-bit bin code of each
// stored particle follows the N_p ids of the state buffer.
__host__ __device__ __forceinline__ size_t ts_far_at(long ntiles) { return 2 * (size_t)ntiles; }
__host__ __device__ __forceinline__ size_t ts_base_at(long ntiles) { return 4 * (size_t)ntiles; }
__host__ __device__ __forceinline__ size_t ts_old_at(long ntiles) { return (4 + (size_t)kBins) * ntiles; }
__device__ __forceinline__ uint32_t* ts_base(uint32_t* counts, const TileGeom& G) {
  return counts + ts_base_at(G.ntiles);
}
// Bin code of a particle the gather moved: the neighbourhood cell of its new home tile (h0 / 2,
// bits 0-5) and its upper-face flags (bits 6-8), or kCodeFar when it moved beyond the
// neighbourhood.
constexpr uint32_t kCodeFar = 0x8000u;
constexpr int kScatterT = 64;  // threads per block of the tile scatter (one wavefront per tile)
constexpr int kScatterE = 4;   // particles per thread and batch in the tile scatter

// Adds 1 to ctr[key] for every active lane with one atomic per distinct key in the wavefront
// and returns each lane's rank among all additions to its key (old value + lane prefix).
// Must be reached by every lane of the wavefront.
template <bool RESULT>
__device__ __forceinline__ uint32_t wave_add(uint32_t* ctr, uint32_t key, bool active) {
  const int lane = __lane_id();
  unsigned long long pending = __ballot(active);
  uint32_t res = 0;
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const uint32_t lkey = __shfl(key, leader);
    const unsigned long long same = __ballot(active && key == lkey) & pending;
    if (RESULT) {
      uint32_t base = 0;
      if (lane == leader) base = atomicAdd(&ctr[lkey], (uint32_t)__popcll(same));
      base = __shfl(base, leader);
      if ((same >> lane) & 1ull) res = base + (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
    } else {
      if (lane == leader) atomicAdd(&ctr[lkey], (uint32_t)__popcll(same));
    }
    pending &= ~same;
  }
  return res;
}

// ------------------------------------------------------------------------------------------
// counting and scattering slot entries
// ------------------------------------------------------------------------------------------
template <typename R, int TX, int TY, int TZ>
__global__ __launch_bounds__(kT) void k_bin_count(const R* __restrict__ s, const uint32_t* __restrict__ id, Lattice L,
                                                  TileGeom G, uint32_t* __restrict__ counts) {
  const long e = blockIdx.x * (long)kT + threadIdx.x;
  const bool valid = e < L.np;
  int c[3] = {0, 0, 0}, t[3] = {0, 0, 0}, up[3] = {0, 0, 0};
  R d[3];
  if (valid) {
    cell_and_frac<R>(s, e, id[e], L, c, d);
    tile_of<TX, TY, TZ>(c, G, t, up);
  }
#pragma unroll 1
  for (int o = 0; o < 8; ++o) {
    const bool act = valid && target_active(up, o);
    wave_add<false>(counts, act ? target_key(t, o, G) : 0u, act);
  }
}

// Places reserved for the active lanes' keys: lanes sharing a key get consecutive places
// from one cursor atomic issued by the lowest such lane.  The first U distinct keys of the
// wavefront take straight-line code whose atomics stay in flight until finish() reads them;
// further keys (rare) go through a loop that waits for each.
template <int U> struct WaveReserve {
  uint32_t res[U];  // atomic results, held by the leader lanes
  int grp;          // group of this lane (U: loop path, place in `direct`)
  int lead;         // leader lane of the group
  uint32_t pre;     // rank among the lanes of the group
  uint32_t direct;

  __device__ __forceinline__ void issue(uint32_t* __restrict__ cursors, uint32_t key, bool active) {
    const int lane = __lane_id();
    unsigned long long pending = __ballot(active);
    grp = U;
    lead = lane;
    pre = 0u;
    direct = 0u;
#pragma unroll
    for (int g = 0; g < U; ++g) {
      res[g] = 0u;
      if (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const uint32_t lkey = __shfl(key, leader);
        const unsigned long long same = __ballot(active && key == lkey) & pending;
        if (lane == leader) res[g] = atomicAdd(&cursors[lkey], (uint32_t)__popcll(same));
        if ((same >> lane) & 1ull) {
          grp = g;
          lead = leader;
          pre = (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
        }
        pending &= ~same;
      }
    }
    while (pending) {
      const int leader = __ffsll((unsigned long long)pending) - 1;
      const uint32_t lkey = __shfl(key, leader);
      const unsigned long long same = __ballot(active && key == lkey) & pending;
      uint32_t base = 0u;
      if (lane == leader) base = atomicAdd(&cursors[lkey], (uint32_t)__popcll(same));
      base = __shfl(base, leader);
      if ((same >> lane) & 1ull) direct = base + (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
      pending &= ~same;
    }
  }

  __device__ __forceinline__ uint32_t finish() const {
    uint32_t b = direct;
#pragma unroll
    for (int g = 0; g < U; ++g) {
      const uint32_t v = (uint32_t)__shfl((int)res[g], lead);
      if (grp == g) b = v + pre;
    }
    return b;
  }
};

// Scatter of every particle into its tile entries (home tile, and the halo lists of the
// tiles its cloud reaches), with the cursor atomics of all its keys in flight together.
template <typename R, int TX, int TY, int TZ>
__device__ __forceinline__ void bin_scatter_body(const R* __restrict__ s, const uint32_t* __restrict__ id,
                                                 const Lattice& L, const TileGeom& G,
                                                 const uint32_t* __restrict__ offsets, uint32_t* __restrict__ cursors,
                                                 uint32_t* __restrict__ entries, uint32_t capacity, long blk,
                                                 uint32_t* __restrict__ health) {
  const long slot = blk * (long)kT + threadIdx.x;  // storage index
  const bool valid = slot < L.np;
  int c[3] = {0, 0, 0}, t[3] = {0, 0, 0}, up[3] = {0, 0, 0};
  R d[3];
  if (valid) {
    // a non-finite displacement is binned at the clamped cell (the deposit counts it)
    cell_and_frac<R>(s, slot, id[slot], L, c, d);
    tile_of<TX, TY, TZ>(c, G, t, up);
  }
  // halo keys only in wavefronts that have any (wave-uniform)
  const bool any_halo = __ballot(valid && (up[0] | up[1] | up[2])) != 0ull;
  uint32_t key[8], beg[8], lim[8];
  bool act[8];
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    act[o] = (o == 0 || any_halo) && valid && target_active(up, o);
    key[o] = act[o] ? target_key(t, o, G) : 0u;
    beg[o] = lim[o] = 0u;
    if (act[o]) {  // segment bounds load alongside the atomics
      beg[o] = offsets[key[o]];
      lim[o] = offsets[key[o] + 1];
    }
  }
  WaveReserve<2> home;
  home.issue(cursors, key[0], act[0]);
  WaveReserve<1> halo[7];
  if (any_halo) {
#pragma unroll
    for (int o = 1; o < 8; ++o) halo[o - 1].issue(cursors, key[o], act[o]);
  }
  uint32_t lost = 0;  // entries beyond their segment: the counts disagree with the positions
  if (act[0]) {
    const uint32_t pos = beg[0] + home.finish();
    if (pos < lim[0] && pos < capacity) entries[pos] = (uint32_t)slot;
    else ++lost;
  }
  if (any_halo) {
#pragma unroll
    for (int o = 1; o < 8; ++o) {
      const uint32_t pos = beg[o] + halo[o - 1].finish();
      if (act[o]) {
        if (pos < lim[o] && pos < capacity) entries[pos] = (uint32_t)slot;
        else ++lost;
      }
    }
  }
  if (lost && health) atomicAdd(&health[PM_HEALTH_OVERFLOW], lost);
}

template <typename R, int TX, int TY, int TZ>
__global__ __launch_bounds__(kT) void k_bin_scatter(const R* __restrict__ s, const uint32_t* __restrict__ id,
                                                    Lattice L, TileGeom G, const uint32_t* __restrict__ offsets,
                                                    uint32_t* __restrict__ cursors, uint32_t* __restrict__ entries,
                                                    uint32_t capacity, uint32_t* __restrict__ health) {
  bin_scatter_body<R, TX, TY, TZ>(s, id, L, G, offsets, cursors, entries, capacity, blockIdx.x, health);
}

// Scatter after a gather, one wavefront per source tile: the particles the gather of tile t
// wrote (storage [old[t], old[t+1]) of its output) are placed through LDS cursors.  The gather
// reserved, with the one counter atomic per (tile, neighbourhood bin) it issues anyway, a range
// of every key's count space for the tile's particles (wbase; kNone for empty bins); the cursor
// of a bin starts at offsets[key] + wbase, so the places cost LDS atomics, not device atomics.
// The gather also left each particle's bin code (2 bytes instead of the 16 of its displacement
// and id).  Particles that moved beyond the neighbourhood were counted in the far counts; the
// scan put them after the reserved ranges of each key, and they take a device cursor atomic on
// keys recomputed from their stored state.
template <typename R, int TX, int TY, int TZ>
__device__ __forceinline__ void tile_scatter_body(const R* __restrict__ s, const uint32_t* __restrict__ id,
                                                  const Lattice& L, const TileGeom& G,
                                                  const uint32_t* __restrict__ offsets, uint32_t* __restrict__ cursors,
                                                  uint32_t* __restrict__ entries, uint32_t capacity,
                                                  uint32_t* __restrict__ health, uint32_t* __restrict__ counts,
                                                  int tile, uint32_t* cur, uint32_t* lim) {
  constexpr int NT = kScatterT, E = kScatterE, BPT = kBins / NT;
  static_assert(kBins % NT == 0, "whole bins per thread");
  const long np = L.np;
  const uint32_t* wbase = ts_base(counts, G);
  const uint32_t* old = counts + ts_old_at(G.ntiles);
  const uint32_t end = min(min(old[tile + 1], capacity), (uint32_t)np), beg = min(old[tile], end);
  if (beg == end) return;  // block-uniform
  const int tix = tile % G.ntx, tiy = (tile / G.ntx) % G.nty, tiz = tile / (G.ntx * G.nty);
  const int tc[3] = {tix, tiy, tiz};
  uint32_t w[BPT];
#pragma unroll
  for (int k = 0; k < BPT; ++k) w[k] = wbase[(size_t)tile * kBins + k * NT + threadIdx.x];
  const uint16_t* codes = reinterpret_cast<const uint16_t*>(id) + 2 * np;
  uint32_t code[E];
  auto load = [&](uint32_t base) {
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const uint32_t e = base + u * NT + threadIdx.x;
      code[u] = e < end ? (uint32_t)codes[e] : 0u;
    }
  };
  load(beg);
#pragma unroll
  for (int k = 0; k < BPT; ++k) {
    const int b = k * NT + threadIdx.x;
    uint32_t c = 0u, l = 0u;
    if (w[k] != kNone) {
      const uint32_t key = bin_key(tix, tiy, tiz, b, G);
      c = offsets[key] + w[k];
      l = offsets[key + 1];
    }
    cur[b] = c;
    lim[b] = l;
  }
  __syncthreads();
  uint32_t lost = 0;
  for (uint32_t base = beg; base < end; base += NT * E) {
    if (base != beg) load(base);
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const uint32_t e = base + u * NT + threadIdx.x;
      const bool valid = e < end;
      const bool near = valid && !(code[u] & kCodeFar);
      const int h0 = 2 * (int)(code[u] & 63u);
      const int up[3] = {(int)(code[u] >> 6) & 1, (int)(code[u] >> 7) & 1, (int)(code[u] >> 8) & 1};
      // one LDS atomic per distinct bin of the wavefront (lanes of a tile mostly share their
      // home bin: per-lane atomics on one address serialize)
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        const bool act = near && target_active(up, o);
        if (!__ballot(act)) continue;  // wave-uniform
        const int b = corner_bin(h0, o);
        const uint32_t pos = wave_add<true>(cur, (uint32_t)b, act);
        if (act) {
          if (pos < lim[b] && pos < capacity) entries[pos] = e;
          else ++lost;
        }
      }
      if (valid && (code[u] & kCodeFar)) {
        // moved beyond the neighbourhood (rare): its keys from the stored state
        int c[3], tn[3], up[3];
        R d[3];
        cell_and_frac<R>(s, e, id[e], L, c, d);
        tile_of<TX, TY, TZ>(c, G, tn, up);
#pragma unroll 1
        for (int o = 0; o < 8; ++o) {
          if (!target_active(up, o)) continue;
          const uint32_t key = target_key(tn, o, G);
          const uint32_t pos = offsets[key] + atomicAdd(&cursors[key], 1u);  // after the reserved ranges
          if (pos < offsets[key + 1] && pos < capacity) entries[pos] = e;
          else ++lost;
        }
      }
    }
  }
  if (lost && health) atomicAdd(&health[PM_HEALTH_OVERFLOW], lost);
}

template <typename R, int TX, int TY, int TZ>
__global__ __launch_bounds__(kScatterT) void k_tile_scatter(const R* __restrict__ s, const uint32_t* __restrict__ id,
                                                            Lattice L, TileGeom G, const uint32_t* __restrict__ offsets,
                                                            uint32_t* __restrict__ cursors,
                                                            uint32_t* __restrict__ entries, uint32_t capacity,
                                                            uint32_t* __restrict__ health,
                                                            uint32_t* __restrict__ counts) {
  __shared__ uint32_t cur[kBins], lim[kBins];
  // each XCD walks its run of tiles backwards, so it starts on what the gather (walking the same
  // runs forward) wrote last and the last-level cache still holds
  tile_scatter_body<R, TX, TY, TZ>(s, id, L, G, offsets, cursors, entries, capacity, health, counts,
                                   xcd_block(blockIdx.x, gridDim.x, 1), cur, lim);
}

// ------------------------------------------------------------------------------------------
// exclusive scan of the 2*ntiles counts -> offsets (+ total at the end); clears the counters.
// With `far` (after a gather whose tile scatter follows) a key's count is counts + far: the
// ranges the gather reserved come first, and the cursor of the particles counted in `far`
// starts after them.
// ------------------------------------------------------------------------------------------
// 4 counts per thread: 1024-count chunks, 128 blocks over the 2 x 65536 counts of a 512^3 grid
// in 32x8x8 tiles (16 per thread left three quarters of the CUs idle: bin stage 183 -> 175 us;
// 2 per thread the same, 1 slower)
constexpr int kScanPer = 4;
constexpr int kScanChunk = kT * kScanPer;

// Exclusive scan of one value per thread over the block; `total` gets the block's sum.  Each
// wavefront scans its 64 values with lane shuffles, the wavefront sums meet in LDS (two
// barriers in all); `sh` may be reused once the call returns.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) sh[wv] = x;
  __syncthreads();
  uint32_t base = 0, sum = 0;
#pragma unroll
  for (int k = 0; k < kT / 64; ++k) {
    const uint32_t w = sh[k];
    base += k < wv ? w : 0u;
    sum += w;
  }
  total = sum;
  __syncthreads();
  return base + x - v;
}

__device__ __forceinline__ void scan_reduce_body(const uint32_t* __restrict__ counts, long m,
                                                 uint32_t* __restrict__ bsum, long blk, uint32_t* sh,
                                                 const uint32_t* __restrict__ far) {
  const long base = blk * (long)kScanChunk;
  uint32_t v = 0;
  for (int i = 0; i < kScanPer; ++i) {
    const long j = base + (long)i * kT + threadIdx.x;
    if (j < m) v += counts[j] + (far ? far[j] : 0u);
  }
  uint32_t total;
  block_exclusive_scan(v, sh, total);
  if (threadIdx.x == 0) bsum[blk] = total;
}

__device__ __forceinline__ void scan_spine_body(uint32_t* __restrict__ bsum, int nb, uint32_t* sh) {
  uint32_t carry = 0;
  for (int base = 0; base < nb; base += kT) {
    const int j = base + threadIdx.x;
    const uint32_t v = j < nb ? bsum[j] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(v, sh, total);
    if (j < nb) bsum[j] = carry + ex;
    carry += total;
  }
}

// Chunk sums, then the last block to finish scans them (the spine): each block publishes its
// sum, fences, and takes a ticket; the block holding the last ticket fences again, scans the
// sums in place and resets the ticket counter for the next scan.  No block waits on another.
__device__ __forceinline__ void scan_reduce_spine_body(const uint32_t* __restrict__ counts, long m,
                                                       uint32_t* __restrict__ bsum, int nb, uint32_t* done,
                                                       long blk, uint32_t* sh, uint32_t* last,
                                                       const uint32_t* __restrict__ far = nullptr) {
  scan_reduce_body(counts, m, bsum, blk, sh, far);
  if (threadIdx.x == 0) {
    __threadfence();
    *last = atomicAdd(done, 1u) == (uint32_t)(nb - 1) ? 1u : 0u;
  }
  __syncthreads();
  if (*last) {
    __threadfence();
    scan_spine_body(bsum, nb, sh);
    if (threadIdx.x == 0) *done = 0u;
  }
}

__global__ __launch_bounds__(kT) void k_scan_reduce_spine(const uint32_t* __restrict__ counts, long m,
                                                          uint32_t* __restrict__ bsum, int nb, uint32_t* done,
                                                          const uint32_t* __restrict__ far) {
  __shared__ uint32_t sh[kT];
  __shared__ uint32_t last;
  scan_reduce_spine_body(counts, m, bsum, nb, done, blockIdx.x, sh, &last, far);
}

__device__ __forceinline__ void scan_apply_body(uint32_t* __restrict__ counts, long m,
                                                const uint32_t* __restrict__ bsum, uint32_t* __restrict__ offsets,
                                                uint32_t* __restrict__ cursors, long blk, uint32_t* sh,
                                                uint32_t* __restrict__ far = nullptr,
                                                uint32_t* __restrict__ old = nullptr) {
  const long base = blk * (long)kScanChunk + (long)threadIdx.x * kScanPer;
  uint32_t v[kScanPer], near[kScanPer];
  uint32_t sum = 0;
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    const long j = base + i;
    near[i] = j < m ? counts[j] : 0u;
    v[i] = near[i] + (far && j < m ? far[j] : 0u);
    sum += v[i];
  }
  uint32_t total;
  uint32_t run = bsum[blk] + block_exclusive_scan(sum, sh, total);
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    const long j = base + i;
    if (j < m) {
      if (old && j <= m / 2) old[j] = offsets[j];  // own offsets of the binning being replaced
      offsets[j] = run;
      counts[j] = 0u;
      cursors[j] = far ? near[i] : 0u;
      if (far) far[j] = 0u;
    }
    run += v[i];
    if (j == m - 1) offsets[m] = run;
  }
}

__global__ __launch_bounds__(kT) void k_scan_apply(uint32_t* __restrict__ counts, long m,
                                                   const uint32_t* __restrict__ bsum, uint32_t* __restrict__ offsets,
                                                   uint32_t* __restrict__ cursors, uint32_t* __restrict__ far,
                                                   uint32_t* __restrict__ old) {
  __shared__ uint32_t sh[kT];
  scan_apply_body(counts, m, bsum, offsets, cursors, blockIdx.x, sh, far, old);
}

// ------------------------------------------------------------------------------------------
// owner-computes CIC deposit (density.py:17-47 for the cells of one tile)
//
// The entry and displacement loads of E particles per thread are issued together before any
// is used, so each pass over a tile's entries costs two memory latencies, not 2E.
// ------------------------------------------------------------------------------------------
// row pitch TX + 3: consecutive lattice particles (2 cells apart) then spread over the LDS
// banks instead of sharing the few that a power-of-two pitch leaves them
template <int TX, int TY, int TZ> constexpr int deposit_lds_cells() { return (TX + 3) * TY * TZ; }

// DET: the LDS sums are 64-bit fixed point with `fix` units per unit mass (a power of two chosen
// by the host so that the whole mass fits in 62 bits); integer adds are associative, so the
// cell values do not depend on the order the atomics land in.
template <typename R, int TX, int TY, int TZ, bool DET = false>
__device__ __forceinline__ void deposit_body(const R* __restrict__ s, const uint32_t* __restrict__ id,
                                             R* __restrict__ rho, R mass, const Lattice& L, const TileGeom& G,
                                             const uint32_t* __restrict__ offsets,
                                             const uint32_t* __restrict__ entries, uint32_t capacity, int tile,
                                             double* acc, uint32_t* __restrict__ health, double fix = 0.0) {
  unsigned long long* iacc = reinterpret_cast<unsigned long long*>(acc);
  constexpr int CELLS = TX * TY * TZ;
  static_assert(CELLS % kT == 0, "tile must hold a multiple of the block size");
  constexpr int E = 4;
  constexpr int PX = TX + 3, PLANE = PX * TY, PADDED = deposit_lds_cells<TX, TY, TZ>();
  const int tix = tile % G.ntx, tiy = (tile / G.ntx) % G.nty, tiz = tile / (G.ntx * G.nty);
  const int o0[3] = {tix * TX, tiy * TY, tiz * TZ};
  const int T[3] = {TX, TY, TZ};
  const int n = L.n;
  const long np = L.np;
  // own entries [ob, oe) and halo entries [hb, he), read as one list of n_own + n_halo
  const uint32_t nt = (uint32_t)G.ntiles;
  const uint32_t oe = min(offsets[tile + 1], capacity), ob = min(offsets[tile], oe);
  const uint32_t he = min(offsets[nt + tile + 1], capacity), hb = min(offsets[nt + tile], he);
  const uint32_t n_own = oe - ob, total = n_own + (he - hb);
  uint32_t sl[E], pid[E];
  R sv[E][3];
  auto load = [&](uint32_t base) {
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const uint32_t i = base + u * kT + threadIdx.x;
      const uint32_t e = i < n_own ? ob + i : hb + (i - n_own);
      sl[u] = i < total ? min(entries[e], (uint32_t)(np - 1)) : kNone;
    }
#pragma unroll
    for (int u = 0; u < E; ++u) {
      pid[u] = sl[u] != kNone ? id[sl[u]] : 0u;
#pragma unroll
      for (int a = 0; a < 3; ++a) sv[u][a] = sl[u] != kNone ? s[a * np + sl[u]] : R(0);
    }
  };
  load(0);  // the first batch is in flight while the accumulator is cleared
  for (int i = threadIdx.x; i < PADDED; i += kT) {
    if (DET) iacc[i] = 0ull;
    else acc[i] = 0.0;
  }
  __syncthreads();
  uint32_t bad = 0;        // own entries whose cell lies outside the tile (stale bins)
  uint32_t nonfinite = 0;  // own entries with a non-finite displacement (binned at a clamped cell)
  for (uint32_t base = 0; base < total; base += kT * E) {
    if (base != 0) load(base);
#pragma unroll
    for (int u = 0; u < E; ++u) {
      if (sl[u] == kNone) continue;
      int c[3];
      R d[3], q[3];
      cell_and_frac<R>(sv[u][0], sv[u][1], sv[u][2], pid[u], L, c, d, q);
      if (base + u * kT + threadIdx.x < n_own) {  // each particle is an own entry once
        bad += (c[0] / TX != tix || c[1] / TY != tiy || c[2] / TZ != tiz) ? 1u : 0u;
        nonfinite += is_finite3(sv[u][0], sv[u][1], sv[u][2]) ? 0u : 1u;
      }
      int r0[3], r1[3];  // tile-relative coordinates of the two corners per axis (-1: outside)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        int lo = c[a] - o0[a];
        if (lo < 0) lo += n;
        int hi = (c[a] + 1 == n ? 0 : c[a] + 1) - o0[a];
        if (hi < 0) hi += n;
        r0[a] = lo < T[a] ? lo : -1;
        r1[a] = hi < T[a] ? hi : -1;
      }
      // density.py:24-47: weights t = 1 - d and d, products with the mass in FP64
      const double dd[3] = {(double)d[0], (double)d[1], (double)d[2]};
      const double w0[3] = {1.0 - dd[0], 1.0 - dd[1], 1.0 - dd[2]};
#pragma unroll
      for (int oz = 0; oz < 2; ++oz) {
        const int rz = oz ? r1[2] : r0[2];
        if (rz < 0) continue;
        const double wz = (double)mass * (oz ? dd[2] : w0[2]);
#pragma unroll
        for (int oy = 0; oy < 2; ++oy) {
          const int ry = oy ? r1[1] : r0[1];
          if (ry < 0) continue;
          const double wyz = wz * (oy ? dd[1] : w0[1]);
#pragma unroll
          for (int ox = 0; ox < 2; ++ox) {
            const int rx = ox ? r1[0] : r0[0];
            if (rx < 0) continue;
            const double w = wyz * (ox ? dd[0] : w0[0]);
            if (DET)
              __hip_atomic_fetch_add(&iacc[rz * PLANE + ry * PX + rx], (unsigned long long)__double2ll_rn(w * fix),
                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            else
              __hip_atomic_fetch_add(&acc[rz * PLANE + ry * PX + rx], w, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_WORKGROUP);
          }
        }
      }
    }
  }
  if (bad && health) atomicAdd(&health[PM_HEALTH_BINS], bad);
  if (nonfinite && health) atomicAdd(&health[PM_HEALTH_NONFINITE], nonfinite);
  __syncthreads();
  const size_t nn = (size_t)n;
#pragma unroll
  for (int i0 = 0; i0 < CELLS / kT; ++i0) {
    const int i = i0 * kT + threadIdx.x;
    const int rx = i % TX, ry = (i / TX) % TY, rz = i / (TX * TY);
    const int li = rz * PLANE + ry * PX + rx;
    rho[((size_t)(o0[2] + rz) * nn + (o0[1] + ry)) * nn + (o0[0] + rx)] =
        DET ? R((double)iacc[li] / fix) : R(acc[li]);
  }
}

// step: when set, the last kernel of a table-driven step advances the table's step counter
// (no kernel of the step reads it after the gather)
template <typename R, int TX, int TY, int TZ, bool DET>
__global__ __launch_bounds__(kT) void k_deposit_tiles(const R* __restrict__ s, const uint32_t* __restrict__ id,
                                                      R* __restrict__ rho, R mass, Lattice L, TileGeom G,
                                                      const uint32_t* __restrict__ offsets,
                                                      const uint32_t* __restrict__ entries, uint32_t capacity,
                                                      int* step, uint32_t* __restrict__ health, double fix) {
  __shared__ double acc[deposit_lds_cells<TX, TY, TZ>()];
  deposit_body<R, TX, TY, TZ, DET>(s, id, rho, mass, L, G, offsets, entries, capacity,
                                   xcd_tile(blockIdx.x, gridDim.x), acc, health, fix);
  if (step && blockIdx.x == 0 && threadIdx.x == 0) *step += 1;
}

// ------------------------------------------------------------------------------------------
// fused deposit + real-to-complex x transform (density.py:17-47, then potential.py:19-21 along
// x), FP32
//
// One workgroup owns a row block: the NTX tiles of one (tile y, tile z) pair, i.e. TY * TZ full
// grid rows of N cells.  It deposits the own and halo entries of those tiles (each entry's
// corners that fall in the tile whose list holds it, as the tile deposit does) into an LDS image
// of the rows, then transforms the rows along x exactly as the R2C pass does (N/2-point complex
// FFT of x[2n] + i x[2n+1] and the even/odd untangling) and writes the half spectrum and the
// Nyquist plane.  The density grid is never written: the pass saves one write and one read of
// N^3 reals per step.  The LDS image is the transform's own layout (complex n of row w at
// n * (RB + 1) + w), so the accumulation needs no reshuffle; it is FP32 (the FP64 accumulator
// of the tile deposit would not fit the rows' LDS).
// ------------------------------------------------------------------------------------------
constexpr int kDrT = 1024;  // threads per row-block workgroup

template <int N, int RB> constexpr int deposit_r2c_lds() { return ((N / 2) * (RB + 1) + N) * 8 + 4 * 80; }

template <int TX, int TY, int TZ, int N>
__global__ __launch_bounds__(kDrT) void k_deposit_r2c(const float* __restrict__ s, const uint32_t* __restrict__ id,
                                                      float mass, Lattice L, TileGeom G,
                                                      const uint32_t* __restrict__ offsets,
                                                      const uint32_t* __restrict__ entries, uint32_t capacity,
                                                      const float2* __restrict__ twg, float2* __restrict__ H,
                                                      float2* __restrict__ Q, int* step, uint32_t* __restrict__ health) {
  constexpr int M = N / 2, RB = TY * TZ, SN = RB + 1, NTX = N / TX, NT = kDrT;
  constexpr int E = 2;  // entries per thread and batch (all loads issued before use)
  static_assert(NTX <= 64, "row block of at most 64 tiles");
  __shared__ float2 lds[M * SN];
  __shared__ float2 tw[N];
  __shared__ uint32_t hoff[NTX + 1];
  float* acc = reinterpret_cast<float*>(lds);
  const int tid = threadIdx.x;
  const int rb = xcd_block(blockIdx.x, gridDim.x);  // = tile z * nty + tile y
  const int ty = rb % G.nty, tz = rb / G.nty;
  const int y0 = ty * TY, z0 = tz * TZ;
  const uint32_t t0 = (uint32_t)rb * NTX, nt = (uint32_t)G.ntiles;
  const uint32_t oe = min(offsets[t0 + NTX], capacity), ob = min(offsets[t0], oe);
  const uint32_t he = min(offsets[nt + t0 + NTX], capacity), hb = min(offsets[nt + t0], he);
  const uint32_t n_own = oe - ob, total = n_own + (he - hb);
  if (tid <= NTX) hoff[tid] = offsets[nt + t0 + tid];
  for (int i = tid; i < N; i += NT) tw[i] = twg[i];
  for (int i = tid; i < M * SN; i += NT) lds[i] = make_float2(0.f, 0.f);
  const long np = L.np;
  const int n = L.n;
  uint32_t bad = 0;        // own entries whose cell lies outside the row block (stale bins)
  uint32_t nonfinite = 0;  // own entries with a non-finite displacement
  __syncthreads();
  for (uint32_t base = 0; base < total; base += NT * E) {
    uint32_t sl[E], pid[E];
    float sv[E][3];
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const uint32_t i = base + u * NT + tid;
      const uint32_t e = i < n_own ? ob + i : hb + (i - n_own);
      sl[u] = i < total ? min(entries[e], (uint32_t)(np - 1)) : kNone;
    }
#pragma unroll
    for (int u = 0; u < E; ++u) {
      pid[u] = sl[u] != kNone ? id[sl[u]] : 0u;
#pragma unroll
      for (int a = 0; a < 3; ++a) sv[u][a] = sl[u] != kNone ? s[a * np + sl[u]] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < E; ++u) {
      if (sl[u] == kNone) continue;
      const uint32_t i = base + u * NT + tid;
      int c[3];
      float d[3], q[3];
      cell_and_frac<float>(sv[u][0], sv[u][1], sv[u][2], pid[u], L, c, d, q);
      // the tile whose list holds the entry: the home tile for own entries, the halo list's
      // tile (largest j with hoff[j] <= entry) for halo entries
      int tx;
      if (i < n_own) {
        tx = c[0] / TX;
        bad += (c[1] / TY != ty || c[2] / TZ != tz) ? 1u : 0u;
        nonfinite += is_finite3(sv[u][0], sv[u][1], sv[u][2]) ? 0u : 1u;
      } else {
        const uint32_t e = hb + (i - n_own);
        int lo = 0, hi = NTX - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (hoff[mid] <= e) lo = mid;
          else hi = mid - 1;
        }
        tx = lo;
      }
      const int o0[3] = {tx * TX, y0, z0};
      const int T[3] = {TX, TY, TZ};
      int r0[3], r1[3];  // tile-relative coordinates of the two corners per axis (-1: outside)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        int lo = c[a] - o0[a];
        if (lo < 0) lo += n;
        int hi = (c[a] + 1 == n ? 0 : c[a] + 1) - o0[a];
        if (hi < 0) hi += n;
        r0[a] = lo < T[a] ? lo : -1;
        r1[a] = hi < T[a] ? hi : -1;
      }
      // density.py:24-47: weights t = 1 - d and d times the mass
      const float w0[3] = {1.f - d[0], 1.f - d[1], 1.f - d[2]};
#pragma unroll
      for (int oz = 0; oz < 2; ++oz) {
        const int rz = oz ? r1[2] : r0[2];
        if (rz < 0) continue;
        const float wz = mass * (oz ? d[2] : w0[2]);
#pragma unroll
        for (int oy = 0; oy < 2; ++oy) {
          const int ry = oy ? r1[1] : r0[1];
          if (ry < 0) continue;
          const float wyz = wz * (oy ? d[1] : w0[1]);
          const int w = rz * TY + ry;
#pragma unroll
          for (int ox = 0; ox < 2; ++ox) {
            const int rx = ox ? r1[0] : r0[0];
            if (rx < 0) continue;
            const int x = o0[0] + rx;
            __hip_atomic_fetch_add(&acc[2 * ((x >> 1) * SN + w) + (x & 1)], wyz * (ox ? d[0] : w0[0]),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          }
        }
      }
    }
  }
  if (bad && health) atomicAdd(&health[PM_HEALTH_BINS], bad);
  if (nonfinite && health) atomicAdd(&health[PM_HEALTH_NONFINITE], nonfinite);
  __syncthreads();
  // rows along x: N/2-point complex FFT, then the even/odd untangling (as r2c_body)
  fft_passes<float2, M, 1, -1, RB, SN, NT, 2>(lds, tw, tid);
  constexpr int NPAIR = M / 2 + 1;
  constexpr int TOTP = NPAIR * RB;
  constexpr int PERP = (TOTP + NT - 1) / NT;
  float2 zk[PERP], zm[PERP];
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
  const size_t nn = (size_t)N;
#pragma unroll
  for (int i = 0; i < PERP; ++i) {
    const int b = tid + i * NT;
    if (b < TOTP) {
      const int w = b % RB, k = b / RB;
      const float2 a = zk[i], c = zm[i];
      const float2 Ev = make_float2(0.5f * (a.x + c.x), 0.5f * (a.y - c.y));
      const float2 Ov = make_float2(0.5f * (a.y + c.y), -0.5f * (a.x - c.x));
      const float2 wO = cmul(Ov, tw[k]);
      if (k == 0) {
        lds[w] = make_float2(Ev.x + wO.x, 0.f);
        Q[(size_t)(z0 + w / TY) * nn + (y0 + w % TY)] = make_float2(Ev.x - wO.x, 0.f);
      } else {
        lds[k * SN + w] = cadd(Ev, wO);
        if (2 * k != M) lds[(M - k) * SN + w] = cconj(csub(Ev, wO));
      }
    }
  }
  __syncthreads();
#pragma unroll 4
  for (int e = tid; e < RB * M; e += NT) {
    const int w = e / M, k = e - w * M;
    H[((size_t)(z0 + w / TY) * nn + (y0 + w % TY)) * M + k] = lds[k * SN + w];
  }
  if (step && blockIdx.x == 0 && threadIdx.x == 0) *step += 1;
}

// ------------------------------------------------------------------------------------------
// tile-staged gather + kick + drift (integrate.py:15-97), counting the new tiles
//
// For axis a the reference gathers g_a = 1/2 sum_corners t_j [phi(c_j - e_a) - phi(c_j + e_a)]
// (integrate.py:61-92) at pre-step positions for all three axes (integrate.py:20-24), kicks
// v_a += da*f(a+da)*g_a (:94) and drifts x_a += da*v_a/(a+da)^2*f(a+da) (:95).  Here the 32
// distinct phi samples of the 4x4x4 stencil come from the LDS copy of the tile, all three
// axes are updated in one pass, the displacement is kept in [-N/2, N/2), and the entries of
// the new position are counted for the next binning.  The potential box is fetched into
// registers (whole z planes per wavefront) while the particle state loads are in flight, then
// written to LDS.
// ------------------------------------------------------------------------------------------
template <int TX, int TY, int TZ> constexpr int gkd_box_cells() { return (TX + 3) * (TY + 3) * (TZ + 3); }

constexpr int kGkdE = 2;  // particles per thread and batch in the gather

// own-entry range [beg, end) of a tile, and a thread's first batch of its slot entries
__device__ __forceinline__ void tile_range(const uint32_t* __restrict__ offsets, uint32_t capacity, int tile,
                                           uint32_t& beg, uint32_t& end) {
  end = min(offsets[tile + 1], capacity);
  beg = min(offsets[tile], end);
}
__device__ __forceinline__ void first_entries(const uint32_t* __restrict__ entries, uint32_t beg, uint32_t end,
                                              long np, uint32_t (&sl)[kGkdE]) {
#pragma unroll
  for (int u = 0; u < kGkdE; ++u) {
    const uint32_t e = beg + u * kT + threadIdx.x;
    sl[u] = e < end ? min(entries[e], (uint32_t)(np - 1)) : kNone;
  }
}

// Particle state the gather reads (at storage index) and writes (at own-entry index)
template <typename R> struct GkdState {
  const R* s_in;
  const R* v_in;
  const uint32_t* id_in;
  R* s_out;
  R* v_out;
  uint32_t* id_out;
};

// Gather + kick + drift of one tile's own particles.  [beg, end) is the tile's own-entry range
// and sl0 the thread's first batch of entries (fetched by the caller before the tile's
// potential box, so the two dependent load chains overlap).  Own entry e's particle is
// written to index e of the output state.
template <typename R, int TX, int TY, int TZ, bool LOCAL_HIST, bool TS>
__device__ __forceinline__ void gkd_tile(const GkdState<R>& ps, const R* __restrict__ phi, R kick, R drift,
                                         const Lattice& L, const TileGeom& G,
                                         const uint32_t* __restrict__ entries, uint32_t* __restrict__ counts,
                                         int tile, uint32_t beg, uint32_t end, const uint32_t (&sl0)[kGkdE], R* box,
                                         uint32_t* hist, const float4* __restrict__ edge) {
  constexpr int BX = TX + 3, BY = TY + 3, BZ = TZ + 3;
  constexpr int NW = kT / 64;                       // wavefronts per block
  constexpr int PLANE = BX * BY;                    // values per box plane
  constexpr int PL = (PLANE + 63) / 64;             // wavefront loads per plane
  constexpr int ZIT = (BZ + NW - 1) / NW;           // planes per wavefront
  constexpr int E = kGkdE;
  // hist: new-tile counts, offset (-1..2)^3 from this tile x {own, halo}
  static_assert(!TS || LOCAL_HIST, "the tile scatter needs the neighbourhood histogram");
  if (beg == end) return;  // empty tile: the whole block leaves together
  const int wv_s = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave index in the block
  const int tix = tile % G.ntx, tiy = (tile / G.ntx) % G.nty, tiz = tile / (G.ntx * G.nty);
  const int o0[3] = {tix * TX, tiy * TY, tiz * TZ};
  const int n = L.n;
  const long np = L.np;
  const size_t nn = (size_t)n;
  // LOCAL_HIST: counts of the new tiles go through an LDS histogram of the 4x4x4 tile
  // neighbourhood, which needs 4 distinct tiles per axis to name it uniquely (the launcher
  // checks); otherwise each wavefront aggregates its own atomics
  if (LOCAL_HIST && threadIdx.x < kBins) hist[threadIdx.x] = 0u;
  // The first batch of entries is fetched before the potential box, the batch's particle
  // state together with the box, so the two dependent load chains overlap.
  uint32_t sl[E], pid[E];
#pragma unroll
  for (int u = 0; u < E; ++u) sl[u] = sl0[u];
  R sv[E][3], vv[E][3];
  auto load_entries = [&](uint32_t base) {
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const uint32_t e = base + u * kT + threadIdx.x;
      sl[u] = e < end ? min(entries[e], (uint32_t)(np - 1)) : kNone;
    }
  };
  auto load_state = [&]() {
#pragma unroll
    for (int u = 0; u < E; ++u) {
      pid[u] = sl[u] != kNone ? ps.id_in[sl[u]] : 0u;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        sv[u][a] = sl[u] != kNone ? ps.s_in[a * np + sl[u]] : R(0);
        vv[u][a] = sl[u] != kNone ? ps.v_in[a * np + sl[u]] : R(0);
      }
    }
  };
  // potential of the tile and its -1/+2 stencil margin (periodic).  Each wavefront loads whole
  // z planes of the box (BY x BX values in PL loads of 64 lanes): a lane's in-plane offsets are
  // computed once, and a plane adds one wave-uniform base.  For 32-wide FP32 tiles with the
  // C2R pass's edge array, a box row is its tile's aligned 32 values (one cache line) plus the
  // three margin values from the edge entries of the tile's two x boundaries.
  bool staged = false;
  if constexpr (TX == 32 && sizeof(R) == 4) {
    if (edge) {
      staged = true;
      const int lane = threadIdx.x & 63;
      const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
      constexpr int MAIN = 32 * BY;  // box values that lie in the tile's own 32 columns
      const int kb1 = tix + 1 == G.ntx ? 0 : tix + 1;
      uint32_t off[PL];
#pragma unroll
      for (int k = 0; k < PL; ++k) {
        const int e = k * 64 + lane;
        if (e < MAIN) {
          off[k] = (uint32_t)(fold(o0[1] - 1 + (e >> 5), n) * n + o0[0] + (e & 31));
        } else if (e < PLANE) {
          const int j = e - MAIN, by = j / 3, c = j - 3 * by;
          off[k] = (uint32_t)((((c == 0 ? tix : kb1) * n + fold(o0[1] - 1 + by, n)) << 2) + c);
        } else {
          off[k] = 0u;
        }
      }
      R tmp[ZIT][PL];
#pragma unroll
      for (int it = 0; it < ZIT; ++it) {
        const int bz = wv + NW * it;
        if (bz < BZ) {
          const size_t zf = (size_t)fold(o0[2] - 1 + bz, n);
          const R* plane = phi + zf * nn * nn;
          const R* eplane = reinterpret_cast<const R*>(edge) + zf * (size_t)G.ntx * nn * 4;
#pragma unroll
          for (int k = 0; k < PL; ++k) {
            const int e = k * 64 + lane;
            if (e < PLANE) tmp[it][k] = (e < MAIN ? plane : eplane)[off[k]];
          }
        }
      }
      load_state();
#pragma unroll
      for (int it = 0; it < ZIT; ++it) {
        const int bz = wv + NW * it;
        if (bz < BZ) {
#pragma unroll
          for (int k = 0; k < PL; ++k) {
            const int e = k * 64 + lane;
            if (e < PLANE) {
              int dst;
              if (e < MAIN) {
                dst = (e >> 5) * BX + 1 + (e & 31);
              } else {
                const int j = e - MAIN, by = j / 3, c = j - 3 * by;
                dst = by * BX + (c == 0 ? 0 : 32 + c);
              }
              box[bz * PLANE + dst] = tmp[it][k];
            }
          }
        }
      }
    }
  }
  if (!staged) {
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    uint32_t off[PL];
#pragma unroll
    for (int k = 0; k < PL; ++k) {
      const int e = k * 64 + lane;
      const int by = e / BX, x = e - by * BX;
      off[k] = e < PLANE ? (uint32_t)(fold(o0[1] - 1 + by, n) * n + fold(o0[0] - 1 + x, n)) : 0u;
    }
    R tmp[ZIT][PL];
#pragma unroll
    for (int it = 0; it < ZIT; ++it) {
      const int bz = wv + NW * it;
      if (bz < BZ) {
        const R* plane = phi + (size_t)fold(o0[2] - 1 + bz, n) * nn * nn;
#pragma unroll
        for (int k = 0; k < PL; ++k)
          if (k * 64 + lane < PLANE) tmp[it][k] = plane[off[k]];
      }
    }
    load_state();
#pragma unroll
    for (int it = 0; it < ZIT; ++it) {
      const int bz = wv + NW * it;
      if (bz < BZ) {
#pragma unroll
        for (int k = 0; k < PL; ++k)
          if (k * 64 + lane < PLANE) box[bz * PLANE + k * 64 + lane] = tmp[it][k];
      }
    }
  }
  __syncthreads();
  const R nr = R(n);
  for (uint32_t base = beg; base < end; base += kT * E) {
    if (base != beg) {
      load_entries(base);
      load_state();
    }
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const uint32_t eo = base + u * kT + threadIdx.x;  // own-entry index = output index
      const bool valid = sl[u] != kNone && eo < (uint32_t)np;
      int tn[3] = {0, 0, 0}, up[3] = {0, 0, 0};
      if (valid) {
        int c[3];
        R d[3], q[3];
        cell_and_frac<R>(sv[u][0], sv[u][1], sv[u][2], pid[u], L, c, d, q);
        // tile-relative cell (box coordinate of the cell is that + 1).  Own entries were binned
        // from these same positions, so the cell lies in the tile: the deposit that read the
        // same bins counted any entry outside its tile into the health counters, and the clamp
        // only keeps the reads of such an entry inside the box.
        int b[3];
        const int T[3] = {TX, TY, TZ};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          int rel = c[a] - o0[a];
          if (rel < 0) rel += n;
          b[a] = rel < T[a] ? rel : T[a] - 1;
        }
        const R* bp = box + (b[2] * BY + b[1]) * BX + b[0];
        R P[4][4][4];
#pragma unroll
        for (int zz = 0; zz < 4; ++zz)
#pragma unroll
          for (int yy = 0; yy < 4; ++yy)
#pragma unroll
            for (int xx = 0; xx < 4; ++xx) {
              const int outside = (xx == 0 || xx == 3) + (yy == 0 || yy == 3) + (zz == 0 || zz == 3);
              if (outside <= 1) P[zz][yy][xx] = bp[(zz * BY + yy) * BX + xx];
            }
        const R w0[2] = {R(1) - d[0], d[0]};
        const R w1[2] = {R(1) - d[1], d[1]};
        const R w2[2] = {R(1) - d[2], d[2]};
        R g[3] = {R(0), R(0), R(0)};
#pragma unroll
        for (int oz = 0; oz < 2; ++oz)
#pragma unroll
          for (int oy = 0; oy < 2; ++oy)
#pragma unroll
            for (int ox = 0; ox < 2; ++ox) {
              const R w = w0[ox] * w1[oy] * w2[oz];
              g[0] += w * (P[oz + 1][oy + 1][ox] - P[oz + 1][oy + 1][ox + 2]);
              g[1] += w * (P[oz + 1][oy][ox + 1] - P[oz + 1][oy + 2][ox + 1]);
              g[2] += w * (P[oz][oy + 1][ox + 1] - P[oz + 2][oy + 1][ox + 1]);
            }
        int cn[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const R va = vv[u][a] + kick * (R(0.5) * g[a]);
          const R sa = wrap_disp<R>(sv[u][a] + drift * va, nr);
          ps.v_out[a * np + eo] = va;
          ps.s_out[a * np + eo] = sa;
          cn[a] = cell_of<R>(canonical_pos<R>(q[a], sa, nr), n);
        }
        ps.id_out[eo] = pid[u];
        tile_of<TX, TY, TZ>(cn, G, tn, up);
        if (LOCAL_HIST) {
          const int tc[3] = {tix, tiy, tiz};
          int h0;
          const bool near = near_home_bin(tn, tc, G, h0);
          if (TS) {  // the particle's bins for the tile scatter
            uint32_t code = near ? (uint32_t)(h0 >> 1) | (uint32_t)(up[0] | (up[1] << 1) | (up[2] << 2)) << 6 : kCodeFar;
            // after the ids of the output buffer (addressed from the id pointer: no extra base)
            reinterpret_cast<uint16_t*>(ps.id_out)[2 * np + eo] = (uint16_t)code;
          }
          if (near) {
            atomicAdd(&hist[h0], 1u);
            if (up[0] | up[1] | up[2]) {
#pragma unroll
              for (int o = 1; o < 8; ++o)
                if (target_active(up, o)) atomicAdd(&hist[corner_bin(h0, o)], 1u);
            }
          } else {
            // moved beyond the neighbourhood: global counters (TS: the far counts, which follow
            // the 2*ntiles counts)
            const uint32_t fk = TS ? 2u * (uint32_t)G.ntiles : 0u;
#pragma unroll 1
            for (int o = 0; o < 8; ++o)
              if (target_active(up, o)) atomicAdd(&counts[target_key(tn, o, G) + fk], 1u);
          }
        }
      }
      if (!LOCAL_HIST) {
#pragma unroll 1
        for (int o = 0; o < 8; ++o) {
          const bool act = valid && target_active(up, o);
          wave_add<false>(counts, act ? target_key(tn, o, G) : 0u, act);
        }
      }
    }
  }
  if (LOCAL_HIST) {
    __syncthreads();
    // bin of this thread from the lane and the wave-uniform wave index (TS: keeps threadIdx.x
    // from living across the kernel, which would cost a spill); the tile index likewise from
    // values that are live anyway
    const int bin = TS ? wv_s * 64 + (int)__lane_id() : (int)threadIdx.x;
    if (bin < kBins) {
      const uint32_t cnt = hist[bin];
      uint32_t* const ctr = &counts[bin_key(tix, tiy, tiz, bin, G)];
      if (TS) {
        // the tile's range of the key's count space, kept for the tile scatter
        const uint32_t t_here = (uint32_t)((tiz * G.nty + tiy) * G.ntx + tix);
        ts_base(counts, G)[(size_t)t_here * kBins + bin] = cnt ? atomicAdd(ctr, cnt) : kNone;
      } else if (cnt) {
        atomicAdd(ctr, cnt);
      }
    }
  }
}

template <typename R, int TX, int TY, int TZ, bool LOCAL_HIST, bool TS = false>
__device__ __forceinline__ void gkd_body(const GkdState<R>& P, const R* __restrict__ phi, R kick, R drift,
                                         const Lattice& L, const TileGeom& G, const uint32_t* __restrict__ offsets,
                                         const uint32_t* __restrict__ entries, uint32_t* __restrict__ counts,
                                         uint32_t capacity, int tile, R* box, uint32_t* hist,
                                         const float4* __restrict__ edge = nullptr) {
  uint32_t beg, end, sl[kGkdE];
  tile_range(offsets, capacity, tile, beg, end);
  first_entries(entries, beg, end, L.np, sl);
  gkd_tile<R, TX, TY, TZ, LOCAL_HIST, TS>(P, phi, kick, drift, L, G, entries, counts, tile, beg, end, sl, box,
                                          hist, edge);
}

// FP32 with the LDS histogram: a register budget of 7 workgroups per CU (72 VGPRs, no spills;
// the kernel waits on dependent loads, so the seventh workgroup pays: 482 -> 470 us at 512^3)
template <typename R, int TX, int TY, int TZ, bool LOCAL_HIST, bool TS>
__global__ __launch_bounds__(kT, sizeof(R) == 4 && LOCAL_HIST ? 7 : 1) void k_gkd_tiles(GkdState<R> P, const R* __restrict__ phi, R kick, R drift,
                                                  Lattice L, TileGeom G, const uint32_t* __restrict__ offsets,
                                                  const uint32_t* __restrict__ entries, uint32_t* __restrict__ counts,
                                                  uint32_t capacity, StepTable tab,
                                                  const float4* __restrict__ edge) {
  __shared__ R box[gkd_box_cells<TX, TY, TZ>()];
  __shared__ uint32_t hist[kBins];
  if (tab.table) {  // graph replay: this step's kick and drift from the device table
    const int k = *tab.step;
    kick = R(tab.table[3 * k + 1]);
    drift = R(tab.table[3 * k + 2]);
  }
  gkd_body<R, TX, TY, TZ, LOCAL_HIST, TS>(P, phi, kick, drift, L, G, offsets, entries, counts, capacity,
                                          xcd_tile(blockIdx.x, gridDim.x), box, hist, edge);
}

// ------------------------------------------------------------------------------------------
// batched tile kernels: the tiles (or particle blocks) of every member in one launch, FP32,
// 16^3 tiles.  The XCD remap is applied to the ensemble-wide block index, so every XCD works
// on one contiguous run of tiles.
// ------------------------------------------------------------------------------------------
constexpr int kBT = 16;

// TS: the members with the tile scatter (stage BS_GKD_TS), otherwise the others (BS_GKD); each
// launch holds one code path, so the tile scatter's reservations cost the others no registers
template <bool TS>
__global__ __launch_bounds__(kT, TS ? 5 : 1) void k_batch_gkd(const BatchDesc* __restrict__ D) {
  __shared__ float box[gkd_box_cells<kBT, kBT, kBT>()];
  __shared__ uint32_t hist[kBins];
  constexpr int stage = TS ? BS_GKD_TS : BS_GKD;
  const long b = xcd_tile(blockIdx.x, gridDim.x);
  const int k = batch_member(D, stage, b);
  const BatchMember& m = D->m[k];
  const int tile = (int)(b - D->off[stage][k]);
  const int step = *gmem(D->step);
  const double* sc = gmem(D->table) + 3 * ((long)step * D->n_members + k);
  const float kick = float(sc[1]), drift = float(sc[2]);
  const int in = (m.par + step) & 1;
  const GkdState<float> ps{gmem(m.s[in]),     gmem(m.v[in]),     gmem(m.id[in]),
                           gmem(m.s[in ^ 1]), gmem(m.v[in ^ 1]), gmem(m.id[in ^ 1])};
  const float* phi = gmem(reinterpret_cast<const float*>(m.H));
  if (TS)
    gkd_body<float, kBT, kBT, kBT, true, true>(ps, phi, kick, drift, m.L, m.G, gmem(m.offsets), gmem(m.entries),
                                               gmem(m.counts), m.capacity, tile, box, hist);
  else if (m.local_hist)
    gkd_body<float, kBT, kBT, kBT, true>(ps, phi, kick, drift, m.L, m.G, gmem(m.offsets), gmem(m.entries),
                                         gmem(m.counts), m.capacity, tile, box, hist);
  else
    gkd_body<float, kBT, kBT, kBT, false>(ps, phi, kick, drift, m.L, m.G, gmem(m.offsets), gmem(m.entries),
                                          gmem(m.counts), m.capacity, tile, box, hist);
}

// state buffer the binning and the deposit read: the gather's output (the scan has advanced
// the step counter since the gather read it)
__device__ __forceinline__ int batch_state(const BatchDesc* D, const BatchMember& m) {
  return (m.par + *gmem(D->step)) & 1;
}

__global__ __launch_bounds__(kT) void k_batch_deposit(const BatchDesc* __restrict__ D) {
  __shared__ double acc[deposit_lds_cells<kBT, kBT, kBT>()];
  const long b = xcd_tile(blockIdx.x, gridDim.x);
  const int k = batch_member(D, BS_TILES, b);
  const BatchMember& m = D->m[k];
  const int cur = batch_state(D, m);
  const int tile = (int)(b - D->off[BS_TILES][k]);
  if (m.det_fix > 0.0)
    deposit_body<float, kBT, kBT, kBT, true>(gmem(m.s[cur]), gmem(m.id[cur]), gmem(m.rho), m.mass, m.L, m.G,
                                             gmem(m.offsets), gmem(m.entries), m.capacity, tile, acc, gmem(m.health),
                                             m.det_fix);
  else
    deposit_body<float, kBT, kBT, kBT, false>(gmem(m.s[cur]), gmem(m.id[cur]), gmem(m.rho), m.mass, m.L, m.G,
                                              gmem(m.offsets), gmem(m.entries), m.capacity, tile, acc,
                                              gmem(m.health));
}

// particle-parallel scatter of the members without the tile scatter (BS_SCATTER has no blocks
// for the others), and the tile scatter of those with it (BS_TILE_SCATTER: one wavefront per
// tile, each XCD walking its run backwards as in the single-simulation launch)
__global__ __launch_bounds__(kT) void k_batch_scatter(const BatchDesc* __restrict__ D) {
  const long b = blockIdx.x;
  const int k = batch_member(D, BS_SCATTER, b);
  const BatchMember& m = D->m[k];
  const int cur = batch_state(D, m);
  bin_scatter_body<float, kBT, kBT, kBT>(gmem(m.s[cur]), gmem(m.id[cur]), m.L, m.G, gmem(m.offsets),
                                         gmem(m.cursors), gmem(m.entries), m.capacity, b - D->off[BS_SCATTER][k],
                                         gmem(m.health));
}

__global__ __launch_bounds__(kScatterT) void k_batch_tile_scatter(const BatchDesc* __restrict__ D) {
  __shared__ uint32_t cur_bins[kBins], lim_bins[kBins];
  const long b = xcd_block(blockIdx.x, gridDim.x, 1);
  const int k = batch_member(D, BS_TILE_SCATTER, b);
  const BatchMember& m = D->m[k];
  const int cur = batch_state(D, m);
  tile_scatter_body<float, kBT, kBT, kBT>(gmem(m.s[cur]), gmem(m.id[cur]), m.L, m.G, gmem(m.offsets),
                                          gmem(m.cursors), gmem(m.entries), m.capacity, gmem(m.health),
                                          gmem(m.counts), (int)(b - D->off[BS_TILE_SCATTER][k]), cur_bins, lim_bins);
}

__global__ __launch_bounds__(kT) void k_batch_scan_reduce(const BatchDesc* __restrict__ D) {
  __shared__ uint32_t sh[kT];
  __shared__ uint32_t last;
  const long b = blockIdx.x;
  const int k = batch_member(D, BS_SCAN, b);
  const BatchMember& m = D->m[k];
  uint32_t* counts = gmem(m.counts);
  scan_reduce_spine_body(counts, 2L * m.G.ntiles, gmem(m.block_sums), m.scan_nb, gmem(m.scan_done),
                         b - D->off[BS_SCAN][k], sh, &last, m.tile_scatter ? counts + ts_far_at(m.G.ntiles) : nullptr);
}

__global__ __launch_bounds__(kT) void k_batch_scan_apply(const BatchDesc* __restrict__ D) {
  __shared__ uint32_t sh[kT];
  const long b = blockIdx.x;
  const int k = batch_member(D, BS_SCAN, b);
  const BatchMember& m = D->m[k];
  uint32_t* counts = gmem(m.counts);
  scan_apply_body(counts, 2L * m.G.ntiles, gmem(m.block_sums), gmem(m.offsets), gmem(m.cursors),
                  b - D->off[BS_SCAN][k], sh, m.tile_scatter ? counts + ts_far_at(m.G.ntiles) : nullptr,
                  m.tile_scatter ? counts + ts_old_at(m.G.ntiles) : nullptr);
  // the last kernel that reads the step's scalars was the gather: advance the counter here, so
  // the binning and the deposit find the gather's output buffer from it
  if (b == 0 && threadIdx.x == 0) *gmem(D->step) += 1;
}

}  // namespace


bool tile_shape_supported(int tx, int ty, int tz) {
#define PM_SHAPE(X, Y, Z) \
  if (tx == X && ty == Y && tz == Z) return true;
  PM_TILE_SHAPES(PM_SHAPE)
#undef PM_SHAPE
  return false;
}

TileGeom make_tiles(int n, int prec_bytes, int tx, int ty, int tz) {
  TileGeom g{};
  if (tx == 0 && ty == 0 && tz == 0 && prec_bytes == 4 && (n == 256 || n == 384 || n == 512)) {
    // FP32 default where the fused deposit + R2C kernel runs: 32 x 8 x 8
    tx = 32;
    ty = tz = 8;
  } else if (tx == 0 && ty == 0 && tz == 0) {  // default: 16^3 (8^3 on the smallest grids)
    auto pick = [n](int pref) {
      int t = pref;
      while (t > 1 && n % t) t /= 2;
      return t;
    };
    tx = pick(16);
    ty = pick(16);
    tz = pick(16);
  }
  if (!tile_shape_supported(tx, ty, tz) || n % tx || n % ty || n % tz) return g;  // not usable
  g.tx = tx;
  g.ty = ty;
  g.tz = tz;
  g.ntx = n / tx;
  g.nty = n / ty;
  g.ntz = n / tz;
  g.ntiles = g.ntx * g.nty * g.ntz;
  return g;
}

std::vector<TileGeom> tile_candidates(int n, int prec_bytes) {
  std::vector<TileGeom> out;
#define PM_SHAPE(X, Y, Z)                                 \
  {                                                       \
    const TileGeom g = make_tiles(n, prec_bytes, X, Y, Z); \
    if (g.ntiles) out.push_back(g);                       \
  }
  PM_TILE_SHAPES(PM_SHAPE)
#undef PM_SHAPE
  return out;
}

int scan_blocks(const TileGeom& g) { return (int)((2L * g.ntiles + kScanChunk - 1) / kScanChunk); }

bool tile_scatter_ok(const TileGeom& g) {
  return g.ntx >= 4 && g.nty >= 4 && g.ntz >= 4 && g.tx * g.ty * g.tz >= kTileScatterCells;
}
size_t counts_words(int n, const TileGeom& gmax) {
  const size_t ts_tiles = (size_t)n * n * n / kTileScatterCells;  // most tiles of a tile-scatter shape
  return std::max(4 * (size_t)gmax.ntiles, ts_old_at((long)ts_tiles) + ts_tiles + 1);
}

// the tile scatter runs when the bins hold its arrays and the shape allows it
static bool tile_scatter_on(const TileGeom& G, const Bins& B) { return B.tile_scatter && tile_scatter_ok(G); }

// tile shapes the kernels are instantiated for; make_tiles only produces these
#define PM_TILE_DISPATCH(G, CALL)                                                                   \
  do {                                                                                              \
    if ((G).tx == 16 && (G).ty == 16 && (G).tz == 16) CALL(16, 16, 16);                             \
    else if ((G).tx == 32 && (G).ty == 16 && (G).tz == 8) CALL(32, 16, 8);                          \
    else if ((G).tx == 32 && (G).ty == 8 && (G).tz == 8) CALL(32, 8, 8);                            \
    else if ((G).tx == 16 && (G).ty == 16 && (G).tz == 8) CALL(16, 16, 8);                          \
    else if ((G).tx == 64 && (G).ty == 8 && (G).tz == 8) CALL(64, 8, 8);                            \
    else if ((G).tx == 8 && (G).ty == 8 && (G).tz == 8) CALL(8, 8, 8);                              \
    else return hipErrorInvalidValue;                                                               \
  } while (0)

template <typename R>
hipError_t bins_count(const PState<R>& p, const Lattice& L, const TileGeom& G, const Bins& B, hipStream_t st) {
#define PM_L(X, Y, Z) k_bin_count<R, X, Y, Z><<<blocks_for(L.np), kT, 0, st>>>(p.s, p.id, L, G, B.counts)
  PM_TILE_DISPATCH(G, PM_L);
#undef PM_L
  return hipGetLastError();
}

hipError_t bins_scan(const TileGeom& G, const Bins& B, hipStream_t st, bool after_gather) {
  const long m = 2L * G.ntiles;
  const int nb = scan_blocks(G);
  const bool ts = after_gather && tile_scatter_on(G, B);
  uint32_t* far = ts ? B.counts + ts_far_at(G.ntiles) : nullptr;
  uint32_t* old = ts ? B.counts + ts_old_at(G.ntiles) : nullptr;
  k_scan_reduce_spine<<<nb, kT, 0, st>>>(B.counts, m, B.block_sums, nb, B.done, far);
  k_scan_apply<<<nb, kT, 0, st>>>(B.counts, m, B.block_sums, B.offsets, B.cursors, far, old);
  return hipGetLastError();
}

template <typename R>
hipError_t bins_scatter(const PState<R>& p, const Lattice& L, const TileGeom& G, const Bins& B, hipStream_t st,
                        bool after_gather) {
  if (after_gather && tile_scatter_on(G, B)) {
#define PM_L(X, Y, Z)                                                                                           \
  do {                                                                                                          \
    if constexpr ((X) * (Y) * (Z) >= kTileScatterCells)                                                         \
      k_tile_scatter<R, X, Y, Z><<<G.ntiles, kScatterT, 0, st>>>(p.s, p.id, L, G, B.offsets, B.cursors,         \
                                                                 B.entries, B.capacity, B.health, B.counts);    \
  } while (0)
    PM_TILE_DISPATCH(G, PM_L);
#undef PM_L
    return hipGetLastError();
  }
#define PM_L(X, Y, Z)                                                                                          \
  k_bin_scatter<R, X, Y, Z><<<blocks_for(L.np), kT, 0, st>>>(p.s, p.id, L, G, B.offsets, B.cursors, B.entries, \
                                                             B.capacity, B.health)
  PM_TILE_DISPATCH(G, PM_L);
#undef PM_L
  return hipGetLastError();
}

template <typename R>
hipError_t deposit_tiles(const PState<R>& p, R* rho, double mass, const Lattice& L, const TileGeom& G, const Bins& B,
                         hipStream_t st, int* step, double det_fix) {
#define PM_L(X, Y, Z)                                                                                              \
  do {                                                                                                             \
    if (det_fix > 0.0)                                                                                             \
      k_deposit_tiles<R, X, Y, Z, true><<<G.ntiles, kT, 0, st>>>(p.s, p.id, rho, R(mass), L, G, B.offsets,         \
                                                                 B.entries, B.capacity, step, B.health, det_fix);  \
    else                                                                                                           \
      k_deposit_tiles<R, X, Y, Z, false><<<G.ntiles, kT, 0, st>>>(p.s, p.id, rho, R(mass), L, G, B.offsets,        \
                                                                  B.entries, B.capacity, step, B.health, 0.0);     \
  } while (0)
  PM_TILE_DISPATCH(G, PM_L);
#undef PM_L
  return hipGetLastError();
}

template <typename R>
hipError_t gkd_tiles(const PState<R>& in, const PState<R>& out, const R* phi, double kick, double drift,
                     const Lattice& L, const TileGeom& G, const Bins& B, const StepTable& tab, hipStream_t st,
                     const float4* edge) {
  if (in.s == out.s || in.v == out.v || in.id == out.id) return hipErrorInvalidValue;  // never in place
  const bool local_hist = G.ntx >= 4 && G.nty >= 4 && G.ntz >= 4;
  const GkdState<R> ps{in.s, in.v, in.id, out.s, out.v, out.id};
  const bool ts = tile_scatter_on(G, B);
#define PM_L(X, Y, Z)                                                                                             \
  do {                                                                                                            \
    if (ts) {                                                                                                     \
      if constexpr ((X) * (Y) * (Z) >= kTileScatterCells)                                                         \
        k_gkd_tiles<R, X, Y, Z, true, true><<<G.ntiles, kT, 0, st>>>(ps, phi, R(kick), R(drift), L, G, B.offsets, \
                                                                     B.entries, B.counts, B.capacity, tab, edge); \
    } else if (local_hist)                                                                                        \
      k_gkd_tiles<R, X, Y, Z, true, false><<<G.ntiles, kT, 0, st>>>(ps, phi, R(kick), R(drift), L, G, B.offsets,  \
                                                                    B.entries, B.counts, B.capacity, tab, edge);  \
    else                                                                                                          \
      k_gkd_tiles<R, X, Y, Z, false, false><<<G.ntiles, kT, 0, st>>>(ps, phi, R(kick), R(drift), L, G, B.offsets, \
                                                                     B.entries, B.counts, B.capacity, tab, edge); \
  } while (0)
  PM_TILE_DISPATCH(G, PM_L);
#undef PM_L
  return hipGetLastError();
}

// grid sizes and tile shapes (x extent; 8 x 8 in y and z) of the fused deposit + R2C kernel
#define PM_DR2C_CASES(X) X(256, 32) X(256, 64) X(384, 32) X(384, 64) X(512, 32) X(512, 64)

bool deposit_r2c_supported(int n, const TileGeom& G) {
  if (G.ty != 8 || G.tz != 8) return false;
#define PM_C(NN, TXX) \
  if (n == NN && G.tx == TXX) return true;
  PM_DR2C_CASES(PM_C)
#undef PM_C
  return false;
}

hipError_t deposit_r2c(const PState<float>& p, float mass, const Lattice& L, const TileGeom& G, const Bins& B,
                       const FftTables& t, cpx<float>* H, cpx<float>* Q, hipStream_t st, int* step) {
  if (!deposit_r2c_supported(L.n, G) || t.n != L.n || t.prec != 4) return hipErrorInvalidValue;
  const unsigned blocks = (unsigned)(G.nty * G.ntz);
#define PM_C(NN, TXX)                                                                                              \
  if (L.n == NN && G.tx == TXX) {                                                                                  \
    k_deposit_r2c<TXX, 8, 8, NN><<<blocks, kDrT, 0, st>>>(p.s, p.id, mass, L, G, B.offsets, B.entries, B.capacity, \
                                                          (const float2*)t.tw, H, Q, step, B.health);              \
    return hipGetLastError();                                                                                      \
  }
  PM_DR2C_CASES(PM_C)
#undef PM_C
  return hipErrorInvalidValue;
}

double deterministic_fix(double total_mass, long n_particles) {
  // each of the <= 8 corner contributions of a particle rounds by at most half a unit
  const double room = std::ldexp(1.0, 62) - 4.0 * (double)n_particles;
  int k = (int)std::floor(std::log2(room / total_mass));
  k = std::min(k, 60);
  return std::ldexp(1.0, k);
}

long batch_scatter_blocks(long np) { return blocks_for(np); }

static hipError_t batch_grid(long total, unsigned& g) {
  if (total > 0x7fffffffL) return hipErrorInvalidValue;
  g = (unsigned)total;
  return hipSuccess;
}

hipError_t batch_gkd(const BatchDesc* d, long total, long total_ts, hipStream_t st) {
  unsigned g;
  if (total > 0) {
    if (hipError_t e = batch_grid(total, g)) return e;
    k_batch_gkd<false><<<g, kT, 0, st>>>(d);
  }
  if (total_ts > 0) {
    if (hipError_t e = batch_grid(total_ts, g)) return e;
    k_batch_gkd<true><<<g, kT, 0, st>>>(d);
  }
  return hipGetLastError();
}
hipError_t batch_deposit(const BatchDesc* d, long total, hipStream_t st) {
  unsigned g;
  if (total <= 0) return hipSuccess;
  if (hipError_t e = batch_grid(total, g)) return e;
  k_batch_deposit<<<g, kT, 0, st>>>(d);
  return hipGetLastError();
}
hipError_t batch_scatter(const BatchDesc* d, long total, long total_tiles, hipStream_t st) {
  unsigned g;
  if (total > 0) {
    if (hipError_t e = batch_grid(total, g)) return e;
    k_batch_scatter<<<g, kT, 0, st>>>(d);
  }
  if (total_tiles > 0) {
    if (hipError_t e = batch_grid(total_tiles, g)) return e;
    k_batch_tile_scatter<<<g, kScatterT, 0, st>>>(d);
  }
  return hipGetLastError();
}
hipError_t batch_scan(const BatchDesc* d, long total, int n_members, hipStream_t st) {
  unsigned g;
  if (total <= 0 || n_members <= 0) return hipSuccess;
  if (hipError_t e = batch_grid(total, g)) return e;
  k_batch_scan_reduce<<<g, kT, 0, st>>>(d);
  k_batch_scan_apply<<<g, kT, 0, st>>>(d);
  return hipGetLastError();
}

#define PM_INST(R)                                                                                           \
  template hipError_t bins_count<R>(const PState<R>&, const Lattice&, const TileGeom&, const Bins&,         \
                                    hipStream_t);                                                            \
  template hipError_t bins_scatter<R>(const PState<R>&, const Lattice&, const TileGeom&, const Bins&,       \
                                      hipStream_t, bool);                                                    \
  template hipError_t deposit_tiles<R>(const PState<R>&, R*, double, const Lattice&, const TileGeom&,       \
                                       const Bins&, hipStream_t, int*, double);                              \
  template hipError_t gkd_tiles<R>(const PState<R>&, const PState<R>&, const R*, double, double,            \
                                   const Lattice&, const TileGeom&, const Bins&, const StepTable&, hipStream_t, \
                                   const float4*);
PM_INST(float)
PM_INST(double)
#undef PM_INST

}  // namespace pm
