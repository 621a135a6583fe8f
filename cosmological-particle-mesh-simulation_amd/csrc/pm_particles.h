// Particle-side kernels of the PM step: state conversion, CIC deposit, fused gather+kick+drift,
// and the reference-layout kernels behind the stateless C-ABI entry points.
#pragma once
#include <vector>

#include "pm_common.h"
#include "pm_fft.h"

namespace pm {

// Division of 32-bit unsigned n < 2^31 by a fixed d >= 1 with one multiply-high and a shift
// (round-up reciprocal, Granlund & Montgomery 1994).
struct FastDiv {
  uint32_t d = 1, m = 0, s = 0;
  FastDiv() = default;
  explicit FastDiv(uint32_t div) : d(div) {
    s = 0;
    while ((1ull << s) < div) ++s;
    m = (uint32_t)((((1ull << 32) * ((1ull << s) - div)) / div) + 1);
  }
  __host__ __device__ __forceinline__ uint32_t div(uint32_t n) const {
#ifdef __HIP_DEVICE_COMPILE__
    const uint32_t hi = __umulhi(n, m);
#else
    const uint32_t hi = (uint32_t)(((uint64_t)n * m) >> 32);
#endif
    return (uint32_t)(((uint64_t)hi + n) >> s);
  }
};

// Device storage order of the particles of one simulation.
//
// The reference orders particles by their Lagrangian lattice site, p = (i*NP + j)*NP + k with
// i along positions[0] (zeldovich.py:79-83), so consecutive particles step along grid axis 0,
// the slowest one in memory.  On the device the same lattice is stored in cubic blocks of
// B^3 sites (B = 4: one 64-lane wavefront per block), blocks and sites both ordered with i
// (grid x) fastest; the slot <-> (i, j, k) map is arithmetic.  A particle's slot is its id:
// the state starts in slot order, and each step stores the particles in the order of the tile
// bins (PState), carrying the ids along.
struct Lattice {
  int n;        // grid cells per side (N_CELLS)
  int np_side;  // particles per side (N_PARTS)
  int blk;      // block edge B (4, or 1 when N_PARTS % 4 != 0)
  int lb;       // log2(B)
  int nb;       // blocks per side
  FastDiv nbd;  // division by nb
  long np;      // N_PARTS^3
  double r;     // lattice spacing N_CELLS / N_PARTS
  float rf;     // the same spacing in float32
  int r_int;    // 1 when the spacing is an integer (then q is exact in float32 arithmetic)
};

// Lattice coordinate q = idx * N_CELLS/N_PARTS + 0.5 (zeldovich.py:79-81) in precision R.  The
// float32 value is the float32 rounding of the exact one in both branches.
template <typename R> __host__ __device__ __forceinline__ R lattice_coord(int idx, const Lattice& L) {
  if (sizeof(R) == 4 && L.r_int) return R(fmaf((float)idx, L.rf, 0.5f));
  return R((double)idx * L.r + 0.5);
}

__host__ __device__ __forceinline__ void slot_to_ijk(long slot, const Lattice& L, int& i, int& j, int& k) {
  const int lb = L.lb;
  const uint32_t sl = (uint32_t)slot;
  const uint32_t blk = sl >> (3 * lb);
  const uint32_t loc = sl & ((1u << (3 * lb)) - 1u);
  const uint32_t msk = (1u << lb) - 1u;
  const uint32_t li = loc & msk, lj = (loc >> lb) & msk, lk = loc >> (2 * lb);
  const uint32_t t = L.nbd.div(blk);
  const uint32_t bi = blk - t * (uint32_t)L.nb;
  const uint32_t bk = L.nbd.div(t);
  const uint32_t bj = t - bk * (uint32_t)L.nb;
  i = (int)((bi << lb) + li);
  j = (int)((bj << lb) + lj);
  k = (int)((bk << lb) + lk);
}

// Stage timers recorded by the sim (indices into the timing arrays).
enum Stage { ST_DEPOSIT = 0, ST_R2C, ST_Y_FWD, ST_Z_GREEN, ST_Y_INV, ST_C2R, ST_GKD, ST_BIN, ST_DEPOSIT_R2C, ST_COUNT };

// Eulerian tiles of the grid (cells per tile and tiles per axis).
struct TileGeom {
  int tx, ty, tz;
  int ntx, nty, ntz;
  int ntiles;
};

// Storage-index entries of every tile: own segment [offsets[t], offsets[t+1]) (all own
// segments together fill [0, N_p)) and halo segment [offsets[ntiles+t], offsets[ntiles+t+1]);
// at most 8 entries per particle.
struct Bins {
  uint32_t* counts = nullptr;      // 2*ntiles (own keys, then halo keys), zero between steps; + tile scatter arrays
  uint32_t* offsets = nullptr;     // 2*ntiles + 1
  uint32_t* cursors = nullptr;     // 2*ntiles
  uint32_t* entries = nullptr;     // 8*N_p
  uint32_t* block_sums = nullptr;  // scan spine
  uint32_t* done = nullptr;        // ticket counter of the scan's last-block spine (zero between scans)
  uint32_t capacity = 0;           // entries allocated (8*N_p)
  uint32_t* health = nullptr;      // PM_HEALTH_COUNT counters of problems the kernels found
  // Tile scatter after a gather (false: the particle-parallel scatter with device cursors).
  // Its arrays follow the 2*ntiles counts of the current shape in the counts buffer (sized by
  // counts_words): 2*ntiles far counts (zero between binnings, like the counts), 128 reserved
  // range starts per tile, the previous binning's ntiles + 1 own offsets, a bin code per particle.
  bool tile_scatter = false;
};

// Tile shapes (x, y, z cells) the tile kernels are instantiated for; the autotuner picks
// among those that divide the grid.
#define PM_TILE_SHAPES(X) X(16, 16, 16) X(32, 16, 8) X(32, 8, 8) X(16, 16, 8) X(64, 8, 8) X(8, 8, 8)
bool tile_shape_supported(int tx, int ty, int tz);
// tiles of shape (tx, ty, tz) on an n^3 grid (all 0: the default 16^3); ntiles = 0 if unusable
TileGeom make_tiles(int n, int prec_bytes, int tx = 0, int ty = 0, int tz = 0);
std::vector<TileGeom> tile_candidates(int n, int prec_bytes);
int scan_blocks(const TileGeom& g);
// tiles of at least this many cells, 4 or more per axis, take the tile scatter after a gather
constexpr int kTileScatterCells = 2048;
bool tile_scatter_ok(const TileGeom& g);
// words of the counts buffer on an n^3 grid whose tile shape with most tiles is gmax: the counts
// and far counts (4 per tile of any shape), and the tile scatter's 133 words per tile (+ 1) for
// every shape it runs on
size_t counts_words(int n, const TileGeom& gmax);

// Resident particle state: displacements s and velocities v (3 x N_p each, component-major)
// and the id (lattice slot) of the particle at each storage index (the id buffer holds N_p
// 16-bit bin codes after the ids, written by the gather for the tile scatter).  A sim holds two of them;
// the gather reads one and writes the other in tile-bin order.
template <typename R> struct PState {
  R* s;
  R* v;
  uint32_t* id;
};

// ---- resident-state kernels (displacement storage) -------------------------------------------
// set: stores slot order (id = storage index); get: writes every particle at its reference index
// set_state counts non-finite inputs into health[PM_HEALTH_NONFINITE] (health may be null)
template <typename R, typename Tin>
hipError_t sim_set_state(const Tin* pos, const Tin* vel, const PState<R>& p, const Lattice& L, hipStream_t st,
                         uint32_t* health = nullptr);
template <typename R, typename Tout>
hipError_t sim_get_state(const PState<R>& p, Tout* pos, Tout* vel, const Lattice& L, hipStream_t st);
// id = storage index (slot order)
hipError_t ids_identity(uint32_t* id, const Lattice& L, hipStream_t st);

// ---- tile binning, owner-computes deposit, tile-staged gather + kick + drift ------------------
// Binning: count (from positions) or a gather's counts, then scan and scatter.  after_gather:
// the counts came from gkd_tiles, whose tile scatter buffers (when B has them and the shape
// allows, tile_scatter_ok) the scan and the scatter then use.
template <typename R>
hipError_t bins_count(const PState<R>& p, const Lattice& L, const TileGeom& G, const Bins& B, hipStream_t st);
hipError_t bins_scan(const TileGeom& G, const Bins& B, hipStream_t st, bool after_gather = false);
template <typename R>
hipError_t bins_scatter(const PState<R>& p, const Lattice& L, const TileGeom& G, const Bins& B, hipStream_t st,
                        bool after_gather = false);
// step (optional): a table-driven step's counter, advanced by the deposit as the step's last kernel.
// det_fix > 0: order-independent fixed-point sums with det_fix units per unit mass
// (deterministic_fix gives the one for a total mass).
template <typename R>
hipError_t deposit_tiles(const PState<R>& p, R* rho, double mass, const Lattice& L, const TileGeom& G, const Bins& B,
                         hipStream_t st, int* step = nullptr, double det_fix = 0.0);
// the largest power of two f with (total mass + one unit per particle corner) * f < 2^62
double deterministic_fix(double total_mass, long n_particles);
// Fused deposit + R2C x pass (FP32): the density of the binned particles, transformed along x,
// written as the half spectrum H (N^2 rows of N/2 complex) and the Nyquist plane Q; the density
// grid itself is never stored.  Needs tiles of 8 x 8 cells in y and z; supported() tells whether
// the grid size and tile shape have an instantiation.
bool deposit_r2c_supported(int n, const TileGeom& G);
hipError_t deposit_r2c(const PState<float>& p, float mass, const Lattice& L, const TileGeom& G, const Bins& B,
                       const FftTables& t, cpx<float>* H, cpx<float>* Q, hipStream_t st, int* step = nullptr);
// Reads `in` through the bins' own entries and writes the updated particles to `out` at their
// own-entry index (never in place); also counts the entries of the new positions into B.counts.
template <typename R>
// edge (FP32, 32-wide tiles; may be null): the C2R pass's x-margin array of phi, which lets a
// box row be read as one aligned cache line
hipError_t gkd_tiles(const PState<R>& in, const PState<R>& out, const R* phi, double kick, double drift,
                     const Lattice& L, const TileGeom& G, const Bins& B, const StepTable& tab, hipStream_t st,
                     const float4* edge = nullptr);

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
