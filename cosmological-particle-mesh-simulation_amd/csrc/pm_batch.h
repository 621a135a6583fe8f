// Batched ensemble: many independent FP32 simulations, possibly of different grid sizes,
// advanced together with one kernel launch per stage of the PM step.
//
// Every stage kernel of the single-simulation step has a batched twin.  A launch covers the
// blocks of all members back to back; block b finds its member by a binary search in the
// stage's prefix array off[stage][0..n_members] and then runs the same device body as the
// single-simulation kernel on that member's buffers, with the member's grid size selected at
// run time (a block-uniform branch).  The descriptor lives in device memory and is rewritten
// only when the membership changes.
#pragma once
#include "pm_particles.h"

namespace pm {

constexpr int kBatchMax = 64;  // members per ensemble

// stages whose block count differs per member; the forward and inverse y passes share BS_Y
// (BS_TILES: all tiles, for the deposit; the gather's tiles split by whether the member takes the
// tile scatter, BS_GKD_TS, or not, BS_GKD, and so do the scatters)
enum BatchStage {
  BS_R2C = 0, BS_Y, BS_Z, BS_C2R, BS_TILES, BS_GKD, BS_GKD_TS, BS_SCAN, BS_SCATTER, BS_TILE_SCATTER, BS_COUNT
};

struct BatchMember {
  // grids (FP32): density, half spectrum (overwritten by the potential) and Nyquist plane
  float* rho;
  cpx<float>* H;
  cpx<float>* Q;
  const cpx<float>* tw;  // FFT tables of this grid size
  const float* sin2;
  const float* wdec;
  PencilSeg ya, yb, za, zb;
  int n;
  int deconvolve;
  int from_grid;
  int local_hist;    // the gather's LDS neighbourhood histogram applies (>= 4 tiles per axis)
  int tile_scatter;  // the gather reserves the tile scatter's ranges (tile_scatter_ok)
  const float* fgrid;
  // particles: the two state buffers (displacements, velocities: 3 x N_p each; ids), the one
  // holding the current state at step 0 of the call (par), tile bins
  float* s[2];
  float* v[2];
  uint32_t* id[2];
  int par;
  float mass;
  uint32_t capacity;
  Lattice L;
  TileGeom G;
  uint32_t* counts;
  uint32_t* offsets;
  uint32_t* cursors;
  uint32_t* entries;
  uint32_t* block_sums;
  uint32_t* scan_done;
  uint32_t* health;  // the member's PM_HEALTH_* counters
  double det_fix;    // > 0: deterministic fixed-point deposit with this many units per unit mass
  int scan_nb;  // scan chunks of this member
};

struct BatchDesc {
  int n_members;
  int max_n;
  long off[BS_COUNT][kBatchMax + 1];  // block prefix per stage
  // per-step scalars: table[(step * n_members + member) * 3 + {0: Green, 1: kick, 2: drift}]
  const double* table;
  int* step;  // advanced by the scan, between the gather and the binning
  BatchMember m[kBatchMax];
};

// member of logical block b of stage `st` (off[st] is non-decreasing, off[st][0] = 0)
__device__ __forceinline__ int batch_member(const BatchDesc* __restrict__ D, int st, long b) {
  int lo = 0, hi = D->n_members - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (D->off[st][mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// grid sizes the batched FFT kernels are instantiated for, in two size classes with a kernel
// each: members are ordered by class, so a stage is one launch per class present
#define PM_FOR_EACH_BATCH_N_SMALL(X) X(32) X(48) X(64) X(96) X(128) X(192) X(256)
#define PM_FOR_EACH_BATCH_N_LARGE(X) X(384) X(512)
#define PM_FOR_EACH_BATCH_N(X) PM_FOR_EACH_BATCH_N_SMALL(X) PM_FOR_EACH_BATCH_N_LARGE(X)
bool batch_size_supported(int n);
int batch_class(int n);  // 0: N <= 256, 1: larger
hipError_t batch_fft_init();  // once per process, before the first batched FFT launch

// host-side block counts and LDS bytes of the batched FFT stages for grid size n
long batch_fft_blocks(int n, int stage);
int batch_fft_lds(int n, int stage);
void batch_fft_segments(int n, PencilSeg& ya, PencilSeg& yb, PencilSeg& za, PencilSeg& zb);

// launchers.  FFT stages run the blocks [base, base + count) of one size class with `lds`
// bytes of dynamic LDS; tile stages run all `total` blocks of the stage.
hipError_t batch_r2c(const BatchDesc* d, int cls, long base, long count, int lds, hipStream_t st);
// mode -1: forward y, 0: forward z + k-space operator + inverse z, +1: inverse y
hipError_t batch_pencil(const BatchDesc* d, int cls, int mode, long base, long count, int lds, hipStream_t st);
hipError_t batch_c2r(const BatchDesc* d, int cls, long base, long count, int lds, hipStream_t st);
// total: tiles of the members without the tile scatter (BS_GKD), total_ts: with it (BS_GKD_TS)
hipError_t batch_gkd(const BatchDesc* d, long total, long total_ts, hipStream_t st);
hipError_t batch_scan(const BatchDesc* d, long total, int n_members, hipStream_t st);
// total: blocks of the particle-parallel scatter (BS_SCATTER), total_tiles: wavefronts of the
// tile scatter (BS_TILE_SCATTER)
hipError_t batch_scatter(const BatchDesc* d, long total, long total_tiles, hipStream_t st);
hipError_t batch_deposit(const BatchDesc* d, long total, hipStream_t st);
long batch_scatter_blocks(long np);

}  // namespace pm
