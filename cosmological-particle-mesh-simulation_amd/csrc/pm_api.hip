// libpm C ABI: stateless reference-function mirrors and the device-resident PM simulation.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pm.h"
#include "pm_batch.h"
#include "pm_fft.h"
#include "pm_ics.h"
#include "pm_particles.h"

using namespace pm;

// ------------------------------------------------------------------------------------------
// errors
// ------------------------------------------------------------------------------------------
namespace {
thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define PM_HIP(call)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) return fail(PM_E_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

std::mutex g_mu;
bool g_init = false;
int g_device = 0;
hipStream_t g_stream = nullptr;

int bind_scratch(int device);

int ensure_init() {
  if (g_init) return PM_OK;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_init) return PM_OK;
  int ndev = 0;
  PM_HIP(hipGetDeviceCount(&ndev));
  if (ndev <= 0) return fail(PM_E_HIP, "no HIP device visible");
  PM_HIP(hipGetDevice(&g_device));
  PM_HIP(hipStreamCreate(&g_stream));
  if (int rc = bind_scratch(g_device)) return rc;
  g_init = true;
  return PM_OK;
}

#define PM_INIT()                  \
  do {                             \
    int rc_ = ensure_init();       \
    if (rc_ != PM_OK) return rc_;  \
  } while (0)

// ------------------------------------------------------------------------------------------
// per-size FFT tables and complex128 workspaces
// ------------------------------------------------------------------------------------------
std::map<std::pair<int, int>, FftTables> g_tables;

int get_tables(int n, int prec, FftTables** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto key = std::make_pair(n, prec);
  auto it = g_tables.find(key);
  if (it != g_tables.end()) {
    *out = &it->second;
    return PM_OK;
  }
  std::vector<double> tw(2 * (size_t)n), s2(n), wd(n);
  for (int m = 0; m < n; ++m) {
    const double ang = 2.0 * M_PI * (double)m / (double)n;
    tw[2 * m] = std::cos(ang);
    tw[2 * m + 1] = -std::sin(ang);
    const double sp = std::sin(M_PI * (double)m / (double)n);
    s2[m] = sp * sp;
    const int mm = m < n / 2 ? m : m - n;
    const double x = M_PI * (double)mm / (double)n;
    const double sinc = mm == 0 ? 1.0 : std::sin(x) / x;
    wd[m] = 1.0 / (sinc * sinc * sinc * sinc);
  }
  FftTables t;
  t.n = n;
  t.prec = prec;
  const size_t rb = (size_t)prec;
  PM_HIP(hipMalloc(&t.tw, 2 * (size_t)n * rb));
  PM_HIP(hipMalloc(&t.sin2, (size_t)n * rb));
  PM_HIP(hipMalloc(&t.wdec, (size_t)n * rb));
  if (prec == 8) {
    PM_HIP(hipMemcpy(t.tw, tw.data(), tw.size() * 8, hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(t.sin2, s2.data(), s2.size() * 8, hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(t.wdec, wd.data(), wd.size() * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<float> a(tw.begin(), tw.end()), b(s2.begin(), s2.end()), c(wd.begin(), wd.end());
    PM_HIP(hipMemcpy(t.tw, a.data(), a.size() * 4, hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(t.sin2, b.data(), b.size() * 4, hipMemcpyHostToDevice));
    PM_HIP(hipMemcpy(t.wdec, c.data(), c.size() * 4, hipMemcpyHostToDevice));
  }
  auto res = g_tables.emplace(key, t);
  *out = &res.first->second;
  return PM_OK;
}

struct Workspace {
  double2* H = nullptr;
  double2* Q = nullptr;
  float* phi = nullptr;
};
std::map<int, Workspace> g_ws;

int get_workspace(int n, bool need_phi, Workspace** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  Workspace& w = g_ws[n];
  const size_t nn = (size_t)n;
  if (!w.H) {
    if (hipMalloc(&w.H, nn * nn * (nn / 2) * sizeof(double2)) != hipSuccess)
      return fail(PM_E_NOMEM, "cannot allocate the complex128 half spectrum for N=%d", n);
    if (hipMalloc(&w.Q, nn * nn * sizeof(double2)) != hipSuccess)
      return fail(PM_E_NOMEM, "cannot allocate the Nyquist plane for N=%d", n);
  }
  if (need_phi && !w.phi) {
    if (hipMalloc(&w.phi, nn * nn * nn * sizeof(float)) != hipSuccess)
      return fail(PM_E_NOMEM, "cannot allocate the potential scratch for N=%d", n);
  }
  *out = &w;
  return PM_OK;
}

// ------------------------------------------------------------------------------------------
// fixed device scratch of pm_potential_static: the complex128 half spectrum, the Nyquist plane
// and the FFT tables of one N^3 grid, N <= kStaticMaxN (about 17 MiB, held by each device the
// library binds to for the life of the process).  Host code never uses the array's name for
// device memory: the address comes from hipGetSymbolAddress, taken once per device at set-up.
// ------------------------------------------------------------------------------------------
constexpr int kStaticMaxN = 128;
constexpr size_t kStaticH = (size_t)kStaticMaxN * kStaticMaxN * (kStaticMaxN / 2) * sizeof(double2);
constexpr size_t kStaticQ = (size_t)kStaticMaxN * kStaticMaxN * sizeof(double2);
constexpr size_t kStaticTables = 4 * (size_t)kStaticMaxN * sizeof(double);  // tw (2N), sin2, wdec
constexpr size_t kStaticBytes = kStaticH + kStaticQ + kStaticTables;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kStaticBytes];

void* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device

// with `device` current: look up this device's copy of the scratch array
int bind_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(PM_E_INVALID, "device %d out of range", device);
  if (g_scratch[device]) return PM_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(PM_E_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
  g_scratch[device] = ptr;
  return PM_OK;
}

// FFT tables of size n computed on the device (the same values get_tables uploads)
__global__ void k_static_tables(int n, double* __restrict__ tw, double* __restrict__ s2, double* __restrict__ wd) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const double ang = 2.0 * M_PI * (double)m / (double)n;
  tw[2 * m] = cos(ang);
  tw[2 * m + 1] = -sin(ang);
  const double sp = sin(M_PI * (double)m / (double)n);
  s2[m] = sp * sp;
  const int mm = m < n / 2 ? m : m - n;
  const double x = M_PI * (double)mm / (double)n;
  const double sinc = mm == 0 ? 1.0 : sin(x) / x;
  wd[m] = 1.0 / (sinc * sinc * sinc * sinc);
}

int check_grid(int n) {
  if (!fft_size_supported(n))
    return fail(PM_E_INVALID,
                "N_CELLS=%d is not a supported grid size (8, 16, 32, 48, 64, 96, 128, 192, 256, 384, 512, "
                "768, 1024)",
                n);
  return PM_OK;
}

double f_of_a(double a, double c0, double c1, double c2) {
  // cosmology.py:20-27
  return 1.0 / std::sqrt((c0 + c2 * a + c1 * a * a * a) / a);
}

}  // namespace

// ------------------------------------------------------------------------------------------
// library / memory
// ------------------------------------------------------------------------------------------
extern "C" {

PM_API int pm_init(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_init) {
    if (device >= 0 && device != g_device) return fail(PM_E_STATE, "library already bound to device %d", g_device);
    return PM_OK;
  }
  int ndev = 0;
  PM_HIP(hipGetDeviceCount(&ndev));
  if (ndev <= 0) return fail(PM_E_HIP, "no HIP device visible");
  if (device < 0) PM_HIP(hipGetDevice(&device));
  if (device >= ndev) return fail(PM_E_INVALID, "device %d out of range (%d visible)", device, ndev);
  PM_HIP(hipSetDevice(device));
  g_device = device;
  PM_HIP(hipStreamCreate(&g_stream));
  if (int rc = bind_scratch(device)) return rc;
  g_init = true;
  return PM_OK;
}

PM_API const char* pm_last_error(void) { return g_err.c_str(); }
PM_API int pm_version(void) { return 100; }
PM_API int pm_supported_grid(int n_cells) { return fft_size_supported(n_cells) ? 1 : 0; }

PM_API int pm_malloc(void** ptr, size_t bytes) {
  PM_INIT();
  if (!ptr) return fail(PM_E_INVALID, "null output pointer");
  if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) return fail(PM_E_NOMEM, "hipMalloc(%zu) failed", bytes);
  return PM_OK;
}
PM_API int pm_free(void* ptr) {
  PM_INIT();
  if (ptr) PM_HIP(hipFree(ptr));
  return PM_OK;
}
PM_API int pm_memcpy_htod(void* dst, const void* src, size_t bytes) {
  PM_INIT();
  if (bytes) PM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return PM_OK;
}
PM_API int pm_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  PM_INIT();
  if (bytes) PM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return PM_OK;
}
PM_API int pm_memcpy_dtod(void* dst, const void* src, size_t bytes) {
  PM_INIT();
  if (bytes) PM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return PM_OK;
}
PM_API int pm_memset(void* dst, int value, size_t bytes) {
  PM_INIT();
  if (bytes) PM_HIP(hipMemset(dst, value, bytes));
  return PM_OK;
}
PM_API int pm_mem_info(size_t* free_bytes, size_t* total_bytes) {
  PM_INIT();
  if (!free_bytes || !total_bytes) return fail(PM_E_INVALID, "null argument");
  PM_HIP(hipMemGetInfo(free_bytes, total_bytes));
  return PM_OK;
}

PM_API int pm_synchronize(void) {
  PM_INIT();
  PM_HIP(hipDeviceSynchronize());
  return PM_OK;
}

// ------------------------------------------------------------------------------------------
// reference-function mirrors
// ------------------------------------------------------------------------------------------
PM_API int pm_density(const float* pos, int64_t n_p, int n, double mass, float* rho) {
  PM_INIT();
  if (n <= 0 || n_p < 0 || !rho || (n_p > 0 && !pos)) return fail(PM_E_INVALID, "pm_density: bad arguments");
  PM_HIP(ref_density(pos, n_p, n, mass, rho, g_stream));
  PM_HIP(hipStreamSynchronize(g_stream));
  return PM_OK;
}

PM_API int pm_fourier_grid(int n, float* g) {
  PM_INIT();
  if (n <= 0 || !g) return fail(PM_E_INVALID, "pm_fourier_grid: bad arguments");
  PM_HIP(ref_fourier_grid(n, g, g_stream));
  PM_HIP(hipStreamSynchronize(g_stream));
  return PM_OK;
}

PM_API int pm_density_k(const float* rho, int n, void* rho_k) {
  PM_INIT();
  if (int rc = check_grid(n)) return rc;
  if (!rho || !rho_k) return fail(PM_E_INVALID, "pm_density_k: null pointer");
  FftTables* t;
  Workspace* w;
  if (int rc = get_tables(n, 8, &t)) return rc;
  if (int rc = get_workspace(n, false, &w)) return rc;
  PM_HIP((fft_r2c_x<double, float>(rho, w->H, w->Q, n, *t, g_stream)));
  PM_HIP(fft_c2c_y<double>(w->H, w->Q, n, -1, *t, g_stream));
  PM_HIP(fft_c2c_z<double>(w->H, w->Q, n, -1, *t, g_stream));
  PM_HIP(spec_expand_full(w->H, w->Q, (double2*)rho_k, n, g_stream));
  PM_HIP(hipStreamSynchronize(g_stream));
  return PM_OK;
}

PM_API int pm_potential_k(const void* rho_k, const float* fgrid, int n, double a, double omega_m0, void* phi_k) {
  PM_INIT();
  if (n <= 0 || !rho_k || !fgrid || !phi_k || a == 0.0) return fail(PM_E_INVALID, "pm_potential_k: bad arguments");
  const double c = -3 * omega_m0 / 8 / a;  // potential.py:15
  PM_HIP(spec_potential_k((const double2*)rho_k, fgrid, c, (double2*)phi_k, n, g_stream));
  PM_HIP(hipStreamSynchronize(g_stream));
  return PM_OK;
}

PM_API int pm_potential_real(const void* phi_k, int n, float* phi) {
  PM_INIT();
  if (int rc = check_grid(n)) return rc;
  if (!phi_k || !phi) return fail(PM_E_INVALID, "pm_potential_real: null pointer");
  FftTables* t;
  Workspace* w;
  if (int rc = get_tables(n, 8, &t)) return rc;
  if (int rc = get_workspace(n, false, &w)) return rc;
  const double scale = 1.0 / ((double)n * n * n);  // pyFFTW normalise_idft=True
  PM_HIP(spec_hermitian_half((const double2*)phi_k, w->H, w->Q, n, scale, g_stream));
  PM_HIP(fft_c2c_z<double>(w->H, w->Q, n, +1, *t, g_stream));
  PM_HIP(fft_c2c_y<double>(w->H, w->Q, n, +1, *t, g_stream));
  PM_HIP((fft_c2r_x<double, float>(w->H, w->Q, phi, n, *t, g_stream)));
  PM_HIP(hipStreamSynchronize(g_stream));
  return PM_OK;
}

static int potential_into(const float* rho, const float* fgrid, int n, double a, double omega_m0, float* phi) {
  FftTables* t;
  Workspace* w;
  if (int rc = get_tables(n, 8, &t)) return rc;
  if (int rc = get_workspace(n, false, &w)) return rc;
  GreenArgs g{};
  g.coef = (-3 * omega_m0 / 8 / a) / ((double)n * n * n);
  g.from_grid = 1;
  g.fgrid = fgrid;
  PM_HIP((fft_r2c_x<double, float>(rho, w->H, w->Q, n, *t, g_stream)));
  PM_HIP(fft_c2c_y<double>(w->H, w->Q, n, -1, *t, g_stream));
  PM_HIP(fft_z_green<double>(w->H, w->Q, n, g, *t, g_stream));
  PM_HIP(fft_c2c_y<double>(w->H, w->Q, n, +1, *t, g_stream));
  PM_HIP((fft_c2r_x<double, float>(w->H, w->Q, phi, n, *t, g_stream)));
  return PM_OK;
}

PM_API int pm_potential(const float* rho, const float* fgrid, int n, double a, double omega_m0, float* phi) {
  PM_INIT();
  if (int rc = check_grid(n)) return rc;
  if (!rho || !fgrid || !phi || a == 0.0) return fail(PM_E_INVALID, "pm_potential: bad arguments");
  if (int rc = potential_into(rho, fgrid, n, a, omega_m0, phi)) return rc;
  PM_HIP(hipStreamSynchronize(g_stream));
  return PM_OK;
}

// potential.py:7-10 like pm_potential, allocating nothing: the workspace and the FFT tables are
// carved out of this device's `scratch` array.  Calls that overlap on one device race.
PM_API int pm_potential_static(const float* rho, const float* fgrid, int n, double a, double omega_m0, float* phi) {
  PM_INIT();
  if (int rc = check_grid(n)) return rc;
  if (n > kStaticMaxN)
    return fail(PM_E_UNSUPPORTED, "pm_potential_static: N_CELLS=%d exceeds its fixed scratch (N_CELLS <= %d)", n,
                kStaticMaxN);
  if (!rho || !fgrid || !phi || a == 0.0) return fail(PM_E_INVALID, "pm_potential_static: bad arguments");
  unsigned char* base = static_cast<unsigned char*>(g_scratch[g_device]);
  if (!base) return fail(PM_E_STATE, "pm_potential_static: no scratch bound to device %d", g_device);
  const size_t nn = (size_t)n;
  double2* H = reinterpret_cast<double2*>(base);
  double2* Q = reinterpret_cast<double2*>(base + nn * nn * (nn / 2) * sizeof(double2));
  double* tab = reinterpret_cast<double*>(base + nn * nn * (nn / 2) * sizeof(double2) + nn * nn * sizeof(double2));
  FftTables t;
  t.n = n;
  t.prec = 8;
  t.tw = tab;
  t.sin2 = tab + 2 * nn;
  t.wdec = tab + 3 * nn;
  // every input the passes read is rewritten here: the tables, then the whole half spectrum
  k_static_tables<<<(n + 127) / 128, 128, 0, g_stream>>>(n, (double*)t.tw, (double*)t.sin2, (double*)t.wdec);
  PM_HIP(hipGetLastError());
  GreenArgs g{};
  g.coef = (-3 * omega_m0 / 8 / a) / ((double)n * n * n);
  g.from_grid = 1;
  g.fgrid = fgrid;
  PM_HIP((fft_r2c_x<double, float>(rho, H, Q, n, t, g_stream)));
  PM_HIP(fft_c2c_y<double>(H, Q, n, -1, t, g_stream));
  PM_HIP(fft_z_green<double>(H, Q, n, g, t, g_stream));
  PM_HIP(fft_c2c_y<double>(H, Q, n, +1, t, g_stream));
  PM_HIP((fft_c2r_x<double, float>(H, Q, phi, n, t, g_stream)));
  PM_HIP(hipStreamSynchronize(g_stream));
  return PM_OK;
}

PM_API int pm_cell_centers(const float* pos, int64_t n_p, int n, int64_t* cells) {
  PM_INIT();
  if (n <= 0 || n_p < 0 || (n_p && (!pos || !cells))) return fail(PM_E_INVALID, "pm_cell_centers: bad arguments");
  PM_HIP(ref_cell_centers(pos, n_p, n, cells, g_stream));
  PM_HIP(hipStreamSynchronize(g_stream));
  return PM_OK;
}

PM_API int pm_weights(const int64_t* cells, const float* pos, int64_t n_p, float* t) {
  PM_INIT();
  if (n_p < 0 || (n_p && (!pos || !cells || !t))) return fail(PM_E_INVALID, "pm_weights: bad arguments");
  PM_HIP(ref_weights(cells, pos, n_p, t, g_stream));
  PM_HIP(hipStreamSynchronize(g_stream));
  return PM_OK;
}

PM_API int pm_sweep_one_direction(const int64_t* cells, float* xd, float* vd, const float* t, const float* phi,
                                  int64_t n_p, int n, double da, double f_a1, double a_val, int direction) {
  PM_INIT();
  if (direction < 0 || direction > 2) return fail(PM_E_INVALID, "direction must be 0, 1 or 2");
  if (n <= 0 || n_p < 0 || (n_p && (!cells || !xd || !vd || !t || !phi)))
    return fail(PM_E_INVALID, "pm_sweep_one_direction: bad arguments");
  PM_HIP(ref_sweep(cells, xd, vd, t, phi, n_p, n, da, f_a1, a_val, direction, g_stream));
  PM_HIP(hipStreamSynchronize(g_stream));
  return PM_OK;
}

PM_API int pm_integrate(float* pos, float* vel, int64_t n_p, int n, double a_val, double f_a1, double da,
                        const float* phi) {
  PM_INIT();
  if (n <= 0 || n_p < 0 || (n_p && (!pos || !vel || !phi))) return fail(PM_E_INVALID, "pm_integrate: bad arguments");
  PM_HIP(ref_integrate(pos, vel, n_p, n, a_val, f_a1, da, phi, g_stream));
  PM_HIP(hipStreamSynchronize(g_stream));
  return PM_OK;
}

PM_API int pm_advance_time(const float* rho, float* pos, float* vel, const float* fgrid, int64_t n_p, int n,
                           double a, double da, double h0, double omega_lambda0, double omega_k0,
                           double omega_m0) {
  PM_INIT();
  if (int rc = check_grid(n)) return rc;
  if (n_p < 0 || !rho || !fgrid || (n_p && (!pos || !vel)) || a == 0.0)
    return fail(PM_E_INVALID, "pm_advance_time: bad arguments");
  Workspace* w;
  if (int rc = get_workspace(n, true, &w)) return rc;
  if (int rc = potential_into(rho, fgrid, n, a, omega_m0, w->phi)) return rc;
  const double fa1 = f_of_a(a + da, h0, omega_lambda0, omega_k0);  // integrate.py:12 (Q1)
  PM_HIP(ref_integrate(pos, vel, n_p, n, a, fa1, da, w->phi, g_stream));
  PM_HIP(hipStreamSynchronize(g_stream));
  return PM_OK;
}

PM_API double pm_f(double a, double c0, double c1, double c2) { return f_of_a(a, c0, c1, c2); }

}  // extern "C"

// ------------------------------------------------------------------------------------------
// device-resident simulation
// ------------------------------------------------------------------------------------------
// HIP event pair around one timed stage
struct PendingEvent {
  int stage;
  hipEvent_t a, b;
};

struct pm_sim {
  Lattice L{};
  int prec = 4;
  unsigned flags = 0;
  double omega_m0 = 0.31, omega_l0 = 0.69, omega_k0 = 0.0, h0 = 0.68;
  double mass = 1.0;
  // particle state, double-buffered: sb/vb/idb[cur] hold the current state; each gather reads
  // them and writes the other buffers in the order of the tile bins
  void* sb[2] = {nullptr, nullptr};
  void* vb[2] = {nullptr, nullptr};
  uint32_t* idb[2] = {nullptr, nullptr};
  int cur = 0;
  // two N^3 grids: the density slot grid[gb] and the spectrum slot grid[gb ^ 1].  A step
  // transforms the density into the spectrum slot and writes the potential over it there.  With
  // the fused deposit + R2C the next step's x spectrum is written into the density slot instead
  // (spec_valid); that step runs its passes there and swaps the slots, so the potential again
  // ends up in the spectrum slot.
  void* grid[2] = {nullptr, nullptr};
  int gb = 0;
  void* Q = nullptr;
  // FP32, N % 32 == 0: the x margins of phi at every multiple of 32, written by the C2R pass for
  // the gather's 32-wide tiles (N^2 * N/32 float4)
  float4* edge = nullptr;
  const float* fgrid = nullptr;  // optional caller-owned Green's function grid
  FftTables* tab = nullptr;
  TileGeom G{};
  Bins B{};
  hipStream_t stream = nullptr;
  bool has_state = false, rho_valid = false, phi_valid = false, bins_valid = false;
  bool spec_valid = false;      // the density slot holds the x spectrum of the current density
  // the forward y, z and inverse y passes in their alternative pencil shape (fft_pencil_cols;
  // picked by the autotuner or set by the caller)
  bool pencil_alt[3] = {false, false, false};
  uint32_t* h_health = nullptr; // pinned host copy of the health counters (B.health on the device)
  // graph replay of consecutive steps: per-step scalars in a device table read at *d_step
  double* d_table = nullptr;   // kTableSteps x {green coefficient, kick, drift}
  int* d_step = nullptr;
  // graph[c][g]: kGraphSteps (an even count) steps captured starting from state buffer c with
  // density slot g
  hipGraphExec_t graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  bool timing = false;
  std::vector<PendingEvent> pending;
  double stage_ms[ST_COUNT] = {};
  int64_t stage_calls[ST_COUNT] = {};
  pm_ensemble* ens = nullptr;  // ensemble this sim is a member of
};

namespace {

// bins are allocated for the candidate tile shape with the most tiles, so that the tile shape
// can change (autotuning) without reallocation
TileGeom max_tiles(int n, int prec) {
  TileGeom m = make_tiles(n, prec);
  for (const TileGeom& g : tile_candidates(n, prec))
    if (g.ntiles > m.ntiles) m = g;
  return m;
}

size_t bins_bytes(int n, int np_side, int prec) {
  const TileGeom g = max_tiles(n, prec);
  const size_t np = (size_t)np_side * np_side * np_side;
  // counts (+ the tile scatter's arrays), offsets, cursors, scan spine, entries
  return 4 * (counts_words(n, g) + 2 * 2 * (size_t)g.ntiles + 1 + (size_t)scan_blocks(g) + 1) + 4 * 8 * np;
}

// per-sim allocations besides the grids, particles and bins: the step table of graph replays,
// the captured graphs and the health counters (a generous bound)
constexpr size_t kSimOverhead = 1u << 20;

bool has_edge(int n, int prec) { return prec == 4 && n % 32 == 0 && n >= 64; }
size_t edge_bytes(int n) { return (size_t)n * n * (n / 32) * sizeof(float4); }

size_t sim_bytes(int n, int np_side, int prec) {
  const size_t nn = (size_t)n, np = (size_t)np_side * np_side * np_side;
  return 2 * (6 * np * prec + 6 * np) /*two state buffers: s, v, ids + bin codes*/ + nn * nn * nn * prec /*rho*/ +
         nn * nn * nn * prec /*H*/ + nn * nn * 2 * prec /*Q*/ + bins_bytes(n, np_side, prec) + kSimOverhead +
         (has_edge(n, prec) ? edge_bytes(n) : 0);
}

template <typename R> PState<R> state_buf(const pm_sim* s, int k) { return {(R*)s->sb[k], (R*)s->vb[k], s->idb[k]}; }
template <typename R> PState<R> cur_state(const pm_sim* s) { return state_buf<R>(s, s->cur); }
void* rho_slot(const pm_sim* s) { return s->grid[s->gb]; }
// fixed-point units per unit mass of the deterministic deposit (0: floating-point sums)
double det_fix(const pm_sim* s) {
  return (s->flags & PM_FLAG_DETERMINISTIC) ? deterministic_fix(s->mass * (double)s->L.np, s->L.np) : 0.0;
}
void* spec_slot(const pm_sim* s) { return s->grid[s->gb ^ 1]; }

// the step's deposit and R2C x pass run as one row-block kernel when the sim asks for it (FP32,
// tiles of 8 x 8 cells in y and z, grid sizes with an instantiation).  Not the default: at
// 512^3 its 137 KB of LDS per row block leave one workgroup per CU, whose load, accumulate,
// transform and write phases cannot overlap, and it ran 0.90 ms against 0.44 ms for the tile
// deposit and the R2C pass it replaces.
bool fused_deposit(const pm_sim* s) {
  return s->prec == 4 && (s->flags & PM_FLAG_FUSED_DEPOSIT) && !(s->flags & PM_FLAG_DETERMINISTIC) &&
         deposit_r2c_supported(s->L.n, s->G);
}

// Health counters: enqueue their copy to the pinned host mirror on `st` before the caller's
// synchronisation, then check them after it.  A problem clears the counters and fails.
int health_fetch(pm_sim* s, hipStream_t st) {
  PM_HIP(hipMemcpyAsync(s->h_health, s->B.health, sizeof(uint32_t) * PM_HEALTH_COUNT, hipMemcpyDeviceToHost, st));
  return PM_OK;
}
int health_report(pm_sim* s, hipStream_t st) {
  const uint32_t* h = s->h_health;
  if (!(h[PM_HEALTH_NONFINITE] | h[PM_HEALTH_BINS] | h[PM_HEALTH_OVERFLOW])) return PM_OK;
  const uint32_t nf = h[PM_HEALTH_NONFINITE], nb = h[PM_HEALTH_BINS], no = h[PM_HEALTH_OVERFLOW];
  PM_HIP(hipMemsetAsync(s->B.health, 0, sizeof(uint32_t) * PM_HEALTH_COUNT, st));
  PM_HIP(hipStreamSynchronize(st));
  std::memset(s->h_health, 0, sizeof(uint32_t) * PM_HEALTH_COUNT);
  return fail(PM_E_CORRUPT,
              "particle state is corrupt: %u non-finite positions or velocities, %u bin entries outside "
              "their tile, %u entries without a place in their bin",
              nf, nb, no);
}

// records HIP events around a stage on the owner's stream when its timing is on (the owner is
// a sim or an ensemble)
template <class O> struct StageTimerT {
  O* s;
  int stage;
  hipEvent_t a = nullptr, b = nullptr;
  StageTimerT(O* sim, int st) : s(sim), stage(st) {
    if (s->timing) {
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      (void)hipEventRecord(a, s->stream);
    }
  }
  ~StageTimerT() {
    if (s->timing) {
      (void)hipEventRecord(b, s->stream);
      s->pending.push_back({stage, a, b});
    }
  }
};
using StageTimer = StageTimerT<pm_sim>;

template <class O> int collect_timing(O* s) {
  if (s->pending.empty()) return PM_OK;
  PM_HIP(hipStreamSynchronize(s->stream));
  for (auto& p : s->pending) {
    float ms = 0;
    PM_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    s->stage_ms[p.stage] += ms;
    s->stage_calls[p.stage] += 1;
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  s->pending.clear();
  return PM_OK;
}
int sim_collect_timing(pm_sim* s) { return collect_timing(s); }

// bin every particle into its tile entries from the current state
template <typename R> int sim_bin_impl(pm_sim* s) {
  StageTimer t(s, ST_BIN);
  // counts and the tile scatter's far counts (after a change of tile shape they may cover what
  // the previous shape's arrays held)
  PM_HIP(hipMemsetAsync(s->B.counts, 0, 4 * sizeof(uint32_t) * s->G.ntiles, s->stream));
  PM_HIP(bins_count<R>(cur_state<R>(s), s->L, s->G, s->B, s->stream));
  PM_HIP(bins_scan(s->G, s->B, s->stream));
  PM_HIP(bins_scatter<R>(cur_state<R>(s), s->L, s->G, s->B, s->stream));
  s->bins_valid = true;
  return PM_OK;
}

template <typename R> int sim_deposit_impl(pm_sim* s, int* step = nullptr) {
  if (!s->bins_valid)
    if (int rc = sim_bin_impl<R>(s)) return rc;
  StageTimer t(s, ST_DEPOSIT);
  PM_HIP(deposit_tiles<R>(cur_state<R>(s), (R*)rho_slot(s), s->mass, s->L, s->G, s->B, s->stream, step,
                          det_fix(s)));
  s->rho_valid = true;
  s->spec_valid = false;
  return PM_OK;
}

// fused deposit + R2C x pass into the density slot (FP32 only; fused_deposit(s) must hold)
int sim_deposit_spec_impl(pm_sim* s, int* step = nullptr) {
  if (!s->bins_valid)
    if (int rc = sim_bin_impl<float>(s)) return rc;
  StageTimer t(s, ST_DEPOSIT_R2C);
  PM_HIP(deposit_r2c(cur_state<float>(s), (float)s->mass, s->L, s->G, s->B, *s->tab, (cpx<float>*)rho_slot(s),
                     (cpx<float>*)s->Q, s->stream, step));
  s->spec_valid = true;
  s->rho_valid = false;
  return PM_OK;
}

// per-step scalars of the step at scale factor a with step da
struct StepScalars {
  double green, kick, drift;
};

StepScalars step_scalars(const pm_sim* s, double a, double da) {
  StepScalars c;
  const double n3 = (double)s->L.n * s->L.n * s->L.n;
  c.green = (-3 * s->omega_m0 / 8 / a) / n3;  // potential.py:15 with the 1/N^3 of the inverse DFT
  // integrate.py:12 — f(a+da, [H0, OL, OK]) (Q1) unless the fix flag is set
  const double c0 = (s->flags & PM_FLAG_FIX_F_COSMOLOGY) ? s->omega_m0 : s->h0;
  const double fa1 = f_of_a(a + da, c0, s->omega_l0, s->omega_k0);
  c.kick = da * fa1;                          // integrate.py:94
  c.drift = da / ((a + da) * (a + da)) * fa1;  // integrate.py:95
  return c;
}

// One PM step.  With `table` set the step's scalars are read on the device from
// d_table[*d_step] and the counter is advanced at the end, so the enqueued work is the same
// for every step and can be captured once into a graph.
template <typename R> int sim_step_impl(pm_sim* s, double a, double da, bool table = false) {
  const int n = s->L.n;
  const bool fused = sizeof(R) == 4 && fused_deposit(s);
  // Traversal directions: the R2C pass (after the deposit) and the C2R pass (after the inverse
  // y pass) walk every XCD's run of row blocks backwards, so each first reads what its producer,
  // walking the same planes forward, wrote last and the 256 MB last-level cache still holds, and
  // the forward y pass and the gather after them find the same.
  const bool rev = !(s->flags & PM_FLAG_FORWARD_ORDER);
  if (!s->spec_valid && (!s->rho_valid || !s->bins_valid)) {
    if (table) return fail(PM_E_STATE, "a captured step needs a valid density");
    if (fused) {
      if (int rc = sim_deposit_spec_impl(s)) return rc;
    } else if (int rc = sim_deposit_impl<R>(s)) {
      return rc;
    }
  }
  cpx<R>* Q = (cpx<R>*)s->Q;
  const StepScalars sc = step_scalars(s, a, da);
  GreenArgs g{};
  g.coef = sc.green;
  if (table) {
    g.table = s->d_table;
    g.step = s->d_step;
  }
  g.deconvolve = (s->flags & PM_FLAG_CIC_DECONVOLVE) ? 1 : 0;
  if (s->fgrid) {
    g.from_grid = 1;
    g.fgrid = s->fgrid;
  }
  // the spectrum: left in the density slot by the fused deposit, or the R2C pass of the density
  const bool in_rho_slot = s->spec_valid;
  cpx<R>* H = (cpx<R>*)(in_rho_slot ? rho_slot(s) : spec_slot(s));
  if (!in_rho_slot) {
    StageTimer t(s, ST_R2C);
    PM_HIP((fft_r2c_x<R, R>((const R*)rho_slot(s), H, Q, n, *s->tab, s->stream, rev)));
  }
  // the other slot is the y / z passes' work array (the density it may hold has been consumed)
  cpx<R>* work = reinterpret_cast<cpx<R>*>(in_rho_slot ? spec_slot(s) : rho_slot(s));
  {
    StageTimer t(s, ST_Y_FWD);
    PM_HIP(fft_c2c_y<R>(H, Q, n, -1, *s->tab, s->stream, work, s->pencil_alt[0]));
  }
  {
    StageTimer t(s, ST_Z_GREEN);
    PM_HIP(fft_z_green<R>(H, Q, n, g, *s->tab, s->stream, work, s->pencil_alt[1]));
  }
  {
    StageTimer t(s, ST_Y_INV);
    PM_HIP(fft_c2c_y<R>(H, Q, n, +1, *s->tab, s->stream, work, s->pencil_alt[2]));
  }
  // the gather's 32-wide tiles read their x margins from the edge array the C2R pass writes
  float4* edge = (s->edge && s->G.tx == 32) ? s->edge : nullptr;
  {
    StageTimer t(s, ST_C2R);
    PM_HIP((fft_c2r_x<R, R>(H, Q, (R*)H, n, *s->tab, s->stream, edge, rev)));
  }
  if (in_rho_slot) s->gb ^= 1;  // the potential now sits in the spectrum slot
  s->phi_valid = true;
  s->spec_valid = s->rho_valid = false;
  StepTable tab;
  if (table) {
    tab.table = s->d_table;
    tab.step = s->d_step;
  }
  {
    StageTimer t(s, ST_GKD);
    PM_HIP(gkd_tiles<R>(cur_state<R>(s), state_buf<R>(s, s->cur ^ 1), (const R*)spec_slot(s), sc.kick, sc.drift,
                        s->L, s->G, s->B, tab, s->stream, edge));
  }
  s->cur ^= 1;
  {
    // the gather counted the entries of the new positions: finish their binning
    StageTimer t(s, ST_BIN);
    PM_HIP(bins_scan(s->G, s->B, s->stream, true));
    PM_HIP(bins_scatter<R>(cur_state<R>(s), s->L, s->G, s->B, s->stream, true));
  }
  int* step = table ? s->d_step : nullptr;
  if (fused) return sim_deposit_spec_impl(s, step);
  return sim_deposit_impl<R>(s, step);
}

constexpr int kTableSteps = 4096;  // steps per uploaded scalar table
constexpr int kGraphSteps = 8;     // steps captured in one graph

static_assert(kGraphSteps % 2 == 0 && kTableSteps % 2 == 0, "a graph replay must return to its state buffer");

// capture kGraphSteps table-driven steps of this sim, starting from the current state buffer,
// into an executable graph (once per buffer)
template <typename R> int sim_build_graph(pm_sim* s) {
  const int c0 = s->cur, g0 = s->gb;
  if (s->graph[c0][g0]) return PM_OK;
  if (!s->d_table) {
    PM_HIP(hipMalloc(&s->d_table, sizeof(double) * 3 * kTableSteps));
    PM_HIP(hipMalloc(&s->d_step, sizeof(int)));
  }
  hipGraph_t graph = nullptr;
  PM_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  int rc = PM_OK;
  const bool rho_v = s->rho_valid, spec_v = s->spec_valid, phi_v = s->phi_valid;
  for (int k = 0; k < kGraphSteps && rc == PM_OK; ++k) rc = sim_step_impl<R>(s, 1.0, 0.0, true);
  hipError_t e = hipStreamEndCapture(s->stream, &graph);
  // capture enqueued nothing: restore the host view (an even step count ends on the same state
  // buffer and density slot anyway)
  s->cur = c0;
  s->gb = g0;
  s->rho_valid = rho_v;
  s->spec_valid = spec_v;
  s->phi_valid = phi_v;
  if (rc != PM_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail(PM_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&s->graph[c0][g0], graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    s->graph[c0][g0] = nullptr;
    return fail(PM_E_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
  }
  return PM_OK;
}

// n steps from a through replays of the captured graph (remainder steps table-driven, direct)
template <typename R> int sim_run_graph(pm_sim* s, double& a, double da, int n_steps) {
  // a captured step starts from the state its predecessor leaves: the fused deposit's spectrum,
  // or a density and valid bins
  if (sizeof(R) == 4 && fused_deposit(s)) {
    if (!s->spec_valid)
      if (int rc = sim_deposit_spec_impl(s)) return rc;
  } else if (!s->rho_valid || !s->bins_valid) {
    if (int rc = sim_deposit_impl<R>(s)) return rc;
  }
  if (int rc = sim_build_graph<R>(s)) return rc;
  std::vector<double> host(3 * (size_t)kTableSteps);
  int done = 0;
  while (done < n_steps) {
    const int chunk = std::min(n_steps - done, kTableSteps);
    for (int k = 0; k < chunk; ++k) {
      if (!(a > 0.0)) return fail(PM_E_INVALID, "scale factor must be positive");
      const StepScalars c = step_scalars(s, a, da);
      host[3 * k] = c.green;
      host[3 * k + 1] = c.kick;
      host[3 * k + 2] = c.drift;
      a += da;  // pmesh.py:63
    }
    const int zero = 0;
    PM_HIP(hipMemcpyAsync(s->d_table, host.data(), sizeof(double) * 3 * chunk, hipMemcpyHostToDevice, s->stream));
    PM_HIP(hipMemcpyAsync(s->d_step, &zero, sizeof(int), hipMemcpyHostToDevice, s->stream));
    int k = 0;
    for (; k + kGraphSteps <= chunk; k += kGraphSteps) {
      PM_HIP(hipGraphLaunch(s->graph[s->cur][s->gb], s->stream));
      s->phi_valid = true;
    }
    for (; k < chunk; ++k)
      if (int rc = sim_step_impl<R>(s, 1.0, 0.0, true)) return rc;
    // the host buffers must outlive the asynchronous copies above
    if (int rc = health_fetch(s, s->stream)) return rc;
    PM_HIP(hipStreamSynchronize(s->stream));
    if (int rc = health_report(s, s->stream)) return rc;
    done += chunk;
  }
  // replays end where they started (even step counts); the direct remainder steps kept the flags
  s->bins_valid = true;
  return PM_OK;
}

void sim_drop_graph(pm_sim* s) {
  for (auto& row : s->graph)
    for (hipGraphExec_t& g : row) {
      if (g) (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// use tile shape g from now on (the captured graph holds the old shape's launches)
int sim_use_tiles(pm_sim* s, const TileGeom& g) {
  if (!g.ntiles) return fail(PM_E_UNSUPPORTED, "tile shape does not divide N_CELLS=%d", s->L.n);
  PM_HIP(hipStreamSynchronize(s->stream));
  if (g.tx != s->G.tx || g.ty != s->G.ty || g.tz != s->G.tz) {
    s->G = g;
    s->bins_valid = false;
    s->rho_valid = s->spec_valid = false;
    sim_drop_graph(s);
  }
  return PM_OK;
}

// Autotuner: times the particle stages (deposit, gather with zero kick and drift, which leaves
// the state unchanged, and the binning that follows) for every tile shape that divides the
// grid, on the sim's own particles, and keeps the fastest.
template <typename R> int sim_autotune_impl(pm_sim* s, int reps, double* best_ms) {
  const int n = s->L.n;
  const size_t nn = (size_t)n;
  // the no-op gather reads the potential buffer: give it zeros (finite forces times zero)
  PM_HIP(hipMemsetAsync(spec_slot(s), 0, nn * nn * nn * s->prec, s->stream));
  s->phi_valid = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  PM_HIP(hipEventCreate(&e0));
  PM_HIP(hipEventCreate(&e1));
  double best = 1e300;
  TileGeom bg = s->G;
  int rc = PM_OK;
  for (const TileGeom& g : tile_candidates(n, s->prec)) {
    s->G = g;
    if ((rc = sim_bin_impl<R>(s))) break;
    float ms = 0.f;
    hipError_t e = hipSuccess;
    for (int r = -1; r < reps && e == hipSuccess; ++r) {  // r = -1: warm-up
      if (r == 0) e = hipEventRecord(e0, s->stream);
      if (!e) e = deposit_tiles<R>(cur_state<R>(s), (R*)rho_slot(s), s->mass, s->L, s->G, s->B, s->stream, nullptr,
                                   det_fix(s));
      if (!e) {
        e = gkd_tiles<R>(cur_state<R>(s), state_buf<R>(s, s->cur ^ 1), (const R*)spec_slot(s), 0.0, 0.0, s->L, s->G, s->B,
                         StepTable{}, s->stream);
        if (!e) s->cur ^= 1;
      }
      if (!e) e = bins_scan(s->G, s->B, s->stream, true);
      if (!e) e = bins_scatter<R>(cur_state<R>(s), s->L, s->G, s->B, s->stream, true);
    }
    if (!e) e = hipEventRecord(e1, s->stream);
    if (!e) e = hipEventSynchronize(e1);
    if (!e) e = hipEventElapsedTime(&ms, e0, e1);
    if (e) {
      rc = fail(PM_E_HIP, "autotune timing failed: %s", hipGetErrorString(e));
      break;
    }
    const double t = ms / std::max(reps, 1);
    if (t < best) {
      best = t;
      bg = g;
    }
  }
  // pencil shapes: each y / z pass that has an alternative shape is timed in both on the
  // spectrum slot (its values do not matter) and keeps the alternative only when it is faster
  // by more than 2 %
  for (int p = 0; p < 3 && rc == PM_OK; ++p) {
    const int mode = p == 0 ? -1 : (p == 1 ? 0 : 1);
    if (!fft_pencil_cols(n, s->prec, mode, true)) {
      s->pencil_alt[p] = false;
      continue;
    }
    float t[2] = {0.f, 0.f};
    hipError_t e = hipSuccess;
    for (int alt = 0; alt < 2 && e == hipSuccess; ++alt) {
      for (int r = -1; r < reps && e == hipSuccess; ++r) {  // r = -1: warm-up
        if (r == 0) e = hipEventRecord(e0, s->stream);
        cpx<R>* H = (cpx<R>*)spec_slot(s);
        if (!e) {
          if (mode == 0) {
            GreenArgs g{};
            e = fft_z_green<R>(H, (cpx<R>*)s->Q, n, g, *s->tab, s->stream, nullptr, alt != 0);
          } else {
            e = fft_c2c_y<R>(H, (cpx<R>*)s->Q, n, mode, *s->tab, s->stream, nullptr, alt != 0);
          }
        }
      }
      if (!e) e = hipEventRecord(e1, s->stream);
      if (!e) e = hipEventSynchronize(e1);
      if (!e) e = hipEventElapsedTime(&t[alt], e0, e1);
    }
    if (e) {
      rc = fail(PM_E_HIP, "autotune timing failed: %s", hipGetErrorString(e));
      break;
    }
    s->pencil_alt[p] = t[1] < 0.98f * t[0];
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  s->G = bg;
  s->bins_valid = false;
  s->rho_valid = s->spec_valid = false;
  sim_drop_graph(s);
  if (rc) return rc;
  if (best_ms) *best_ms = best;
  return sim_deposit_impl<R>(s);
}

}  // namespace

extern "C" {

PM_API int pm_sim_set_tiles(pm_sim* s, int tx, int ty, int tz) {
  if (!s) return fail(PM_E_INVALID, "null sim");
  if ((tx || ty || tz) && !tile_shape_supported(tx, ty, tz))
    return fail(PM_E_UNSUPPORTED, "tile shape %dx%dx%d is not instantiated", tx, ty, tz);
  if (s->ens && !(tx == 16 && ty == 16 && tz == 16) && (tx || ty || tz))
    return fail(PM_E_STATE, "ensemble members keep 16^3 tiles");
  return sim_use_tiles(s, make_tiles(s->L.n, s->prec, tx, ty, tz));
}

PM_API int pm_sim_set_pencils(pm_sim* s, int y_fwd, int z, int y_inv) {
  if (!s) return fail(PM_E_INVALID, "null sim");
  const int want[3] = {y_fwd, z, y_inv};
  bool alt[3];
  for (int p = 0; p < 3; ++p) {
    const int mode = p == 0 ? -1 : (p == 1 ? 0 : 1);
    const int w0 = fft_pencil_cols(s->L.n, s->prec, mode, false), w1 = fft_pencil_cols(s->L.n, s->prec, mode, true);
    if (want[p] != 0 && want[p] != w0 && (w1 == 0 || want[p] != w1))
      return fail(PM_E_UNSUPPORTED, "%d columns per block is not a shape of pass %d at N_CELLS=%d (%d%s%s)", want[p],
                  p, s->L.n, w0, w1 ? " or " : "", w1 ? std::to_string(w1).c_str() : "");
    alt[p] = want[p] != 0 && want[p] != w0;
  }
  PM_HIP(hipStreamSynchronize(s->stream));
  if (alt[0] != s->pencil_alt[0] || alt[1] != s->pencil_alt[1] || alt[2] != s->pencil_alt[2]) {
    for (int p = 0; p < 3; ++p) s->pencil_alt[p] = alt[p];
    sim_drop_graph(s);
  }
  return PM_OK;
}

PM_API int pm_sim_get_pencils(pm_sim* s, int* y_fwd, int* z, int* y_inv) {
  if (!s || !y_fwd || !z || !y_inv) return fail(PM_E_INVALID, "null argument");
  int* out[3] = {y_fwd, z, y_inv};
  for (int p = 0; p < 3; ++p)
    *out[p] = fft_pencil_cols(s->L.n, s->prec, p == 0 ? -1 : (p == 1 ? 0 : 1), s->pencil_alt[p]);
  return PM_OK;
}

PM_API int pm_sim_get_tiles(pm_sim* s, int* tx, int* ty, int* tz) {
  if (!s || !tx || !ty || !tz) return fail(PM_E_INVALID, "null argument");
  *tx = s->G.tx;
  *ty = s->G.ty;
  *tz = s->G.tz;
  return PM_OK;
}

PM_API int pm_sim_autotune(pm_sim* s, int reps, int* tx, int* ty, int* tz, double* best_ms) {
  if (!s) return fail(PM_E_INVALID, "null sim");
  if (!s->has_state) return fail(PM_E_STATE, "state was never set");
  if (s->ens) return fail(PM_E_STATE, "ensemble members keep 16^3 tiles");
  int rc = s->prec == 4 ? sim_autotune_impl<float>(s, std::max(reps, 1), best_ms)
                        : sim_autotune_impl<double>(s, std::max(reps, 1), best_ms);
  if (rc) return rc;
  PM_HIP(hipStreamSynchronize(s->stream));
  if (tx) *tx = s->G.tx;
  if (ty) *ty = s->G.ty;
  if (tz) *tz = s->G.tz;
  return PM_OK;
}

PM_API int pm_sim_footprint(int n_cells, int n_parts_side, int precision, size_t* bytes) {
  if (!bytes || n_cells <= 0 || n_parts_side <= 0 || (precision != 32 && precision != 64))
    return fail(PM_E_INVALID, "pm_sim_footprint: bad arguments");
  *bytes = sim_bytes(n_cells, n_parts_side, precision / 8);
  return PM_OK;
}

PM_API int pm_sim_create(int n_cells, int n_parts_side, int precision, unsigned flags, pm_sim** out) {
  PM_INIT();
  if (!out) return fail(PM_E_INVALID, "null output handle");
  *out = nullptr;
  if (int rc = check_grid(n_cells)) return rc;
  if (n_parts_side <= 0) return fail(PM_E_INVALID, "N_PARTS must be positive");
  if (n_parts_side > 812) return fail(PM_E_INVALID, "N_PARTS above 812 is not supported (slot index range)");
  if (precision != 32 && precision != 64) return fail(PM_E_INVALID, "precision must be 32 or 64");
  pm_sim* s = new pm_sim();
  s->prec = precision / 8;
  s->flags = flags;
  s->L.n = n_cells;
  s->L.np_side = n_parts_side;
  s->L.blk = (n_parts_side % 4 == 0) ? 4 : 1;
  s->L.lb = s->L.blk == 4 ? 2 : 0;
  s->L.nb = n_parts_side / s->L.blk;
  s->L.nbd = FastDiv((uint32_t)s->L.nb);
  s->L.np = (long)n_parts_side * n_parts_side * n_parts_side;
  s->L.r = (double)n_cells / (double)n_parts_side;
  s->L.rf = (float)s->L.r;
  s->L.r_int = (n_cells % n_parts_side) == 0 ? 1 : 0;
  s->mass = std::pow((double)n_cells / (double)n_parts_side, 3);  // pmesh.py:28
  const size_t nn = (size_t)n_cells, pr = (size_t)s->prec;
  auto cleanup = [&](int rc) {
    pm_sim_destroy(s);
    return rc;
  };
  if (int rc = get_tables(n_cells, s->prec, &s->tab)) return cleanup(rc);
  bool ok = true;
  for (int k = 0; k < 2; ++k)
    ok = ok && hipMalloc(&s->sb[k], 3 * (size_t)s->L.np * pr) == hipSuccess &&
         hipMalloc(&s->vb[k], 3 * (size_t)s->L.np * pr) == hipSuccess &&
         hipMalloc(&s->idb[k], (size_t)s->L.np * 6) == hipSuccess;  // ids, then the tile scatter's bin codes
  if (!ok || hipMalloc(&s->grid[0], nn * nn * nn * pr) != hipSuccess ||
      hipMalloc(&s->grid[1], nn * nn * nn * pr) != hipSuccess || hipMalloc(&s->Q, nn * nn * 2 * pr) != hipSuccess ||
      (has_edge(n_cells, s->prec) && hipMalloc(&s->edge, edge_bytes(n_cells)) != hipSuccess))
    return cleanup(fail(PM_E_NOMEM, "device allocation of %zu bytes failed", sim_bytes(n_cells, n_parts_side, s->prec)));
  s->G = make_tiles(n_cells, s->prec);
  const TileGeom gmax = max_tiles(n_cells, s->prec);
  const size_t m = 2 * (size_t)gmax.ntiles;
  if (8 * (size_t)s->L.np >= (size_t)UINT32_MAX) return cleanup(fail(PM_E_INVALID, "too many particles (N_PARTS > 812)"));
  s->B.capacity = (uint32_t)(8 * (size_t)s->L.np);
  const size_t cw = counts_words(n_cells, gmax);
  if (hipMalloc(&s->B.counts, cw * 4) != hipSuccess || hipMalloc(&s->B.offsets, (m + 1) * 4) != hipSuccess ||
      hipMalloc(&s->B.cursors, m * 4) != hipSuccess || hipMalloc(&s->B.entries, 8 * (size_t)s->L.np * 4) != hipSuccess ||
      hipMalloc(&s->B.block_sums, ((size_t)scan_blocks(gmax) + 1) * 4) != hipSuccess ||
      hipMalloc(&s->B.health, sizeof(uint32_t) * PM_HEALTH_COUNT) != hipSuccess ||
      hipHostMalloc(&s->h_health, sizeof(uint32_t) * PM_HEALTH_COUNT, hipHostMallocDefault) != hipSuccess)
    return cleanup(fail(PM_E_NOMEM, "device allocation of the tile bins failed"));
  std::memset(s->h_health, 0, sizeof(uint32_t) * PM_HEALTH_COUNT);
  s->B.tile_scatter = true;
  // every entry holds a valid slot from the start, counters start at zero
  s->B.done = s->B.block_sums + scan_blocks(gmax);
  if (hipMemset(s->B.entries, 0, 8 * (size_t)s->L.np * 4) != hipSuccess ||
      hipMemset(s->idb[0], 0, (size_t)s->L.np * 4) != hipSuccess ||
      hipMemset(s->idb[1], 0, (size_t)s->L.np * 4) != hipSuccess ||
      hipMemset(s->B.done, 0, 4) != hipSuccess ||
      hipMemset(s->B.counts, 0, cw * 4) != hipSuccess || hipMemset(s->B.cursors, 0, m * 4) != hipSuccess ||
      hipMemset(s->B.offsets, 0, (m + 1) * 4) != hipSuccess ||
      hipMemset(s->B.health, 0, sizeof(uint32_t) * PM_HEALTH_COUNT) != hipSuccess)
    return cleanup(fail(PM_E_HIP, "clearing the tile bins failed"));
  if (hipStreamCreate(&s->stream) != hipSuccess) return cleanup(fail(PM_E_HIP, "hipStreamCreate failed"));
  *out = s;
  return PM_OK;
}

PM_API int pm_sim_destroy(pm_sim* s) {
  if (!s) return PM_OK;
  if (s->ens) (void)pm_ensemble_remove(s->ens, s);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (auto& p : s->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (int k = 0; k < 2; ++k) {
    (void)hipFree(s->sb[k]);
    (void)hipFree(s->vb[k]);
    (void)hipFree(s->idb[k]);
  }
  (void)hipFree(s->grid[0]);
  (void)hipFree(s->grid[1]);
  (void)hipFree(s->edge);
  (void)hipFree(s->B.health);
  if (s->h_health) (void)hipHostFree(s->h_health);
  (void)hipFree(s->Q);
  (void)hipFree(s->B.counts);
  (void)hipFree(s->B.offsets);
  (void)hipFree(s->B.cursors);
  (void)hipFree(s->B.entries);
  (void)hipFree(s->B.block_sums);
  sim_drop_graph(s);
  (void)hipFree(s->d_table);
  (void)hipFree(s->d_step);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return PM_OK;
}

PM_API int pm_sim_set_cosmology(pm_sim* s, double omega_m0, double omega_lambda0, double omega_k0, double h0) {
  if (!s) return fail(PM_E_INVALID, "null sim");
  s->omega_m0 = omega_m0;
  s->omega_l0 = omega_lambda0;
  s->omega_k0 = omega_k0;
  s->h0 = h0;
  return PM_OK;
}

PM_API int pm_sim_set_mass(pm_sim* s, double mass) {
  if (!s) return fail(PM_E_INVALID, "null sim");
  if (!(mass > 0.0)) return fail(PM_E_INVALID, "mass must be positive");
  if (mass != s->mass) {
    s->rho_valid = s->spec_valid = false;
    sim_drop_graph(s);  // the captured deposit holds the old mass
  }
  s->mass = mass;
  return PM_OK;
}

PM_API int pm_sim_set_green(pm_sim* s, const float* fgrid) {
  if (!s) return fail(PM_E_INVALID, "null sim");
  if (fgrid != s->fgrid) sim_drop_graph(s);  // the captured k-space pass holds the old grid
  s->fgrid = fgrid;
  return PM_OK;
}

PM_API int pm_sim_set_density(pm_sim* s, const void* rho, int dtype_bytes, int on_device) {
  if (!s || !rho) return fail(PM_E_INVALID, "null argument");
  if (dtype_bytes != 4 && dtype_bytes != 8) return fail(PM_E_INVALID, "dtype_bytes must be 4 or 8");
  const size_t cnt = (size_t)s->L.n * s->L.n * s->L.n;
  if (dtype_bytes == s->prec) {
    PM_HIP(hipMemcpyAsync(rho_slot(s), rho, cnt * dtype_bytes, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                          s->stream));
  } else {
    const void* src = rho;
    void* tmp = nullptr;
    if (!on_device) {
      if (hipMalloc(&tmp, cnt * dtype_bytes) != hipSuccess) return fail(PM_E_NOMEM, "staging allocation failed");
      if (hipMemcpyAsync(tmp, rho, cnt * dtype_bytes, hipMemcpyHostToDevice, s->stream) != hipSuccess) {
        (void)hipFree(tmp);
        return fail(PM_E_HIP, "host to device copy failed");
      }
      src = tmp;
    }
    hipError_t e = s->prec == 4 ? convert_grid<double, float>((const double*)src, (float*)rho_slot(s), cnt, s->stream)
                                : convert_grid<float, double>((const float*)src, (double*)rho_slot(s), cnt, s->stream);
    hipError_t es = hipStreamSynchronize(s->stream);
    if (tmp) (void)hipFree(tmp);
    if (e != hipSuccess || es != hipSuccess) return fail(PM_E_HIP, "density upload failed");
  }
  PM_HIP(hipStreamSynchronize(s->stream));
  s->rho_valid = true;
  s->spec_valid = false;
  return PM_OK;
}

PM_API int pm_sim_set_state(pm_sim* s, const void* pos, const void* vel, int dtype_bytes, int on_device) {
  if (!s || !pos || !vel) return fail(PM_E_INVALID, "pm_sim_set_state: null argument");
  if (dtype_bytes != 4 && dtype_bytes != 8) return fail(PM_E_INVALID, "dtype_bytes must be 4 or 8");
  const size_t bytes = 3 * (size_t)s->L.np * dtype_bytes;
  const void* dp = pos;
  const void* dv = vel;
  void* tmp = nullptr;
  if (!on_device) {
    if (hipMalloc(&tmp, 2 * bytes) != hipSuccess) return fail(PM_E_NOMEM, "staging allocation failed");
    dp = tmp;
    dv = (char*)tmp + bytes;
    hipError_t e1 = hipMemcpyAsync((void*)dp, pos, bytes, hipMemcpyHostToDevice, s->stream);
    hipError_t e2 = hipMemcpyAsync((void*)dv, vel, bytes, hipMemcpyHostToDevice, s->stream);
    if (e1 != hipSuccess || e2 != hipSuccess) {
      (void)hipFree(tmp);
      return fail(PM_E_HIP, "host to device copy failed");
    }
  }
  hipError_t e;
  uint32_t* hl = s->B.health;
  if (s->prec == 4) {
    e = dtype_bytes == 4
            ? sim_set_state<float, float>((const float*)dp, (const float*)dv, cur_state<float>(s), s->L, s->stream, hl)
            : sim_set_state<float, double>((const double*)dp, (const double*)dv, cur_state<float>(s), s->L, s->stream,
                                           hl);
  } else {
    e = dtype_bytes == 4
            ? sim_set_state<double, float>((const float*)dp, (const float*)dv, cur_state<double>(s), s->L, s->stream,
                                           hl)
            : sim_set_state<double, double>((const double*)dp, (const double*)dv, cur_state<double>(s), s->L,
                                            s->stream, hl);
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(s->h_health, s->B.health, sizeof(uint32_t) * PM_HEALTH_COUNT, hipMemcpyDeviceToHost, s->stream);
  hipError_t es = hipStreamSynchronize(s->stream);
  if (tmp) (void)hipFree(tmp);
  if (e != hipSuccess || es != hipSuccess) return fail(PM_E_HIP, "set_state failed: %s", hipGetErrorString(e ? e : es));
  s->has_state = true;
  s->bins_valid = false;
  s->rho_valid = s->spec_valid = false;
  s->phi_valid = false;
  if (int rc = health_report(s, s->stream)) {
    s->has_state = false;  // a state with non-finite values is not kept
    return rc;
  }
  return PM_OK;
}

PM_API int pm_sim_get_state(pm_sim* s, void* pos, void* vel, int dtype_bytes, int on_device) {
  if (!s || !pos || !vel) return fail(PM_E_INVALID, "pm_sim_get_state: null argument");
  if (!s->has_state) return fail(PM_E_STATE, "state was never set");
  if (dtype_bytes != 4 && dtype_bytes != 8) return fail(PM_E_INVALID, "dtype_bytes must be 4 or 8");
  const size_t bytes = 3 * (size_t)s->L.np * dtype_bytes;
  void* dp = pos;
  void* dv = vel;
  void* tmp = nullptr;
  if (!on_device) {
    if (hipMalloc(&tmp, 2 * bytes) != hipSuccess) return fail(PM_E_NOMEM, "staging allocation failed");
    dp = tmp;
    dv = (char*)tmp + bytes;
  }
  hipError_t e;
  if (s->prec == 4) {
    e = dtype_bytes == 4 ? sim_get_state<float, float>(cur_state<float>(s), (float*)dp, (float*)dv, s->L, s->stream)
                         : sim_get_state<float, double>(cur_state<float>(s), (double*)dp, (double*)dv, s->L, s->stream);
  } else {
    e = dtype_bytes == 4 ? sim_get_state<double, float>(cur_state<double>(s), (float*)dp, (float*)dv, s->L, s->stream)
                         : sim_get_state<double, double>(cur_state<double>(s), (double*)dp, (double*)dv, s->L,
                                                         s->stream);
  }
  if (e == hipSuccess && !on_device) {
    e = hipMemcpyAsync(pos, dp, bytes, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(vel, dv, bytes, hipMemcpyDeviceToHost, s->stream);
  }
  hipError_t es = hipStreamSynchronize(s->stream);
  if (tmp) (void)hipFree(tmp);
  if (e != hipSuccess || es != hipSuccess) return fail(PM_E_HIP, "get_state failed: %s", hipGetErrorString(e ? e : es));
  return PM_OK;
}

PM_API int pm_sim_deposit(pm_sim* s) {
  if (!s) return fail(PM_E_INVALID, "null sim");
  if (!s->has_state) return fail(PM_E_STATE, "state was never set");
  int rc = s->prec == 4 ? sim_deposit_impl<float>(s) : sim_deposit_impl<double>(s);
  if (rc) return rc;
  PM_HIP(hipStreamSynchronize(s->stream));
  return PM_OK;
}

static int copy_grid_out(pm_sim* s, const void* src, void* dst, int dtype_bytes, int on_device) {
  if (dtype_bytes != 4 && dtype_bytes != 8) return fail(PM_E_INVALID, "dtype_bytes must be 4 or 8");
  const size_t cnt = (size_t)s->L.n * s->L.n * s->L.n;
  void* d = dst;
  void* tmp = nullptr;
  if (dtype_bytes == s->prec) {
    if (on_device) PM_HIP(hipMemcpyAsync(dst, src, cnt * dtype_bytes, hipMemcpyDeviceToDevice, s->stream));
    else PM_HIP(hipMemcpyAsync(dst, src, cnt * dtype_bytes, hipMemcpyDeviceToHost, s->stream));
    PM_HIP(hipStreamSynchronize(s->stream));
    return PM_OK;
  }
  if (!on_device) {
    if (hipMalloc(&tmp, cnt * dtype_bytes) != hipSuccess) return fail(PM_E_NOMEM, "staging allocation failed");
    d = tmp;
  }
  hipError_t e = s->prec == 4 ? convert_grid<float, double>((const float*)src, (double*)d, cnt, s->stream)
                              : convert_grid<double, float>((const double*)src, (float*)d, cnt, s->stream);
  if (e == hipSuccess && !on_device) e = hipMemcpyAsync(dst, d, cnt * dtype_bytes, hipMemcpyDeviceToHost, s->stream);
  hipError_t es = hipStreamSynchronize(s->stream);
  if (tmp) (void)hipFree(tmp);
  if (e != hipSuccess || es != hipSuccess) return fail(PM_E_HIP, "grid copy failed");
  return PM_OK;
}

PM_API int pm_sim_get_density(pm_sim* s, void* rho, int dtype_bytes, int on_device) {
  if (!s || !rho) return fail(PM_E_INVALID, "null argument");
  if (!s->has_state) return fail(PM_E_STATE, "state was never set");
  if (!s->rho_valid) {
    int rc = s->prec == 4 ? sim_deposit_impl<float>(s) : sim_deposit_impl<double>(s);
    if (rc) return rc;
  }
  return copy_grid_out(s, rho_slot(s), rho, dtype_bytes, on_device);
}

PM_API int pm_sim_get_potential(pm_sim* s, void* phi, int dtype_bytes, int on_device) {
  if (!s || !phi) return fail(PM_E_INVALID, "null argument");
  if (!s->phi_valid) return fail(PM_E_STATE, "no potential computed yet (run a step first)");
  return copy_grid_out(s, spec_slot(s), phi, dtype_bytes, on_device);
}

PM_API int pm_sim_step(pm_sim* s, double a, double da) {
  if (!s) return fail(PM_E_INVALID, "null sim");
  if (!s->has_state) return fail(PM_E_STATE, "state was never set");
  if (!(a > 0.0) || !(a + da > 0.0)) return fail(PM_E_INVALID, "scale factor must be positive");
  int rc = s->prec == 4 ? sim_step_impl<float>(s, a, da) : sim_step_impl<double>(s, a, da);
  if (rc) return rc;
  if (int rc2 = health_fetch(s, s->stream)) return rc2;
  PM_HIP(hipStreamSynchronize(s->stream));
  return health_report(s, s->stream);
}

PM_API int pm_sim_run(pm_sim* s, double a, double da, int n_steps, double* a_out) {
  if (!s) return fail(PM_E_INVALID, "null sim");
  if (!s->has_state) return fail(PM_E_STATE, "state was never set");
  if (n_steps < 0) return fail(PM_E_INVALID, "n_steps must be >= 0");
  if (!s->timing && !(s->flags & PM_FLAG_NO_GRAPH) && n_steps >= 2) {
    int rc = s->prec == 4 ? sim_run_graph<float>(s, a, da, n_steps) : sim_run_graph<double>(s, a, da, n_steps);
    if (rc) return rc;
    if (a_out) *a_out = a;
    return PM_OK;
  }
  for (int k = 0; k < n_steps; ++k) {
    if (!(a > 0.0)) return fail(PM_E_INVALID, "scale factor must be positive");
    int rc = s->prec == 4 ? sim_step_impl<float>(s, a, da) : sim_step_impl<double>(s, a, da);
    if (rc) return rc;
    a += da;  // pmesh.py:63
  }
  if (int rc = health_fetch(s, s->stream)) return rc;
  PM_HIP(hipStreamSynchronize(s->stream));
  if (int rc = health_report(s, s->stream)) return rc;
  if (a_out) *a_out = a;
  return PM_OK;
}

PM_API int pm_sim_set_timing(pm_sim* s, int enable) {
  if (!s) return fail(PM_E_INVALID, "null sim");
  if (int rc = sim_collect_timing(s)) return rc;
  s->timing = enable != 0;
  for (int i = 0; i < ST_COUNT; ++i) {
    s->stage_ms[i] = 0;
    s->stage_calls[i] = 0;
  }
  return PM_OK;
}

PM_API int pm_sim_get_timing(pm_sim* s, double* ms, int64_t* calls, int n_stages) {
  if (!s || !ms || !calls) return fail(PM_E_INVALID, "null argument");
  if (int rc = sim_collect_timing(s)) return rc;
  for (int i = 0; i < n_stages && i < ST_COUNT; ++i) {
    ms[i] = s->stage_ms[i];
    calls[i] = s->stage_calls[i];
  }
  return PM_OK;
}

PM_API int pm_sim_sync(pm_sim* s) {
  if (!s) return fail(PM_E_INVALID, "null sim");
  PM_HIP(hipStreamSynchronize(s->stream));
  return PM_OK;
}

PM_API int pm_white_noise(double* out, int64_t n, uint64_t seed) {
  PM_INIT();
  if (n < 0 || (n && !out)) return fail(PM_E_INVALID, "pm_white_noise: bad arguments");
  PM_HIP(white_noise(out, 8, n, seed, g_stream));
  PM_HIP(hipStreamSynchronize(g_stream));
  return PM_OK;
}

PM_API int pm_sim_zeldovich(pm_sim* s, const double* amp, int64_t n_amp, double psi_scale, double vel_factor,
                            uint64_t seed, const void* noise, int noise_bytes) {
  if (!s || !amp) return fail(PM_E_INVALID, "pm_sim_zeldovich: null argument");
  const int n = s->L.np_side, N = s->L.n;
  if (!fft_size_supported(n))
    return fail(PM_E_UNSUPPORTED, "N_PARTS=%d is not a supported FFT size for the device generator", n);
  // the noise spectrum lives in the density buffer, the components in the potential buffers
  if ((double)n * n * n + 2.0 * n * n > (double)N * N * N || n > N)
    return fail(PM_E_UNSUPPORTED, "the device generator needs N_PARTS < N_CELLS");
  const int64_t need = 3 * (int64_t)(n / 2) * (n / 2) + 1;
  if (n_amp < need) return fail(PM_E_INVALID, "h_amp needs %lld entries", (long long)need);
  if (noise && noise_bytes != 4 && noise_bytes != 8) return fail(PM_E_INVALID, "noise_dtype_bytes must be 4 or 8");
  FftTables* t;
  if (int rc = get_tables(n, s->prec, &t)) return rc;
  double* d_amp = nullptr;
  if (hipMalloc(&d_amp, sizeof(double) * need) != hipSuccess) return fail(PM_E_NOMEM, "amplitude table allocation failed");
  hipError_t e = hipMemcpyAsync(d_amp, amp, sizeof(double) * need, hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess) {
    if (s->prec == 4)
      e = zeldovich_ics<float>(noise, noise_bytes, seed, d_amp, psi_scale, vel_factor, s->L, *t,
                               (float*)s->sb[s->cur], (float*)s->vb[s->cur], (float*)rho_slot(s), (float*)spec_slot(s),
                               (cpx<float>*)s->Q, s->stream);
    else
      e = zeldovich_ics<double>(noise, noise_bytes, seed, d_amp, psi_scale, vel_factor, s->L, *t,
                                (double*)s->sb[s->cur], (double*)s->vb[s->cur], (double*)rho_slot(s), (double*)spec_slot(s),
                                (cpx<double>*)s->Q, s->stream);
    if (e == hipSuccess) e = ids_identity(s->idb[s->cur], s->L, s->stream);  // the state is in slot order
  }
  hipError_t es = hipStreamSynchronize(s->stream);
  (void)hipFree(d_amp);
  if (e != hipSuccess || es != hipSuccess)
    return fail(PM_E_HIP, "device initial conditions failed: %s", hipGetErrorString(e ? e : es));
  s->has_state = true;
  s->rho_valid = s->bins_valid = s->phi_valid = s->spec_valid = false;
  return PM_OK;
}

PM_API int pm_sim_power_spectrum(pm_sim* s, double box, int n_bins, double* k, double* pk, double* counts) {
  if (!s || !k || !pk || !counts) return fail(PM_E_INVALID, "pm_sim_power_spectrum: null argument");
  if (!s->has_state) return fail(PM_E_STATE, "state was never set");
  if (n_bins < 1 || n_bins > 2048) return fail(PM_E_INVALID, "n_bins must be in 1..2048");
  if (!(box > 0.0)) return fail(PM_E_INVALID, "box must be positive");
  if (!s->rho_valid) {
    int rc = s->prec == 4 ? sim_deposit_impl<float>(s) : sim_deposit_impl<double>(s);
    if (rc) return rc;
  }
  const int n = s->L.n;
  const long cells = (long)n * n * n;
  double* d = nullptr;  // [0]: sum of rho, then 3 (n_bins + 1) shell sums
  const size_t nd = 1 + 3 * (size_t)(n_bins + 1);
  if (hipMalloc(&d, sizeof(double) * nd) != hipSuccess) return fail(PM_E_NOMEM, "spectrum scratch allocation failed");
  std::vector<double> h(nd);
  hipError_t e = hipMemsetAsync(d, 0, sizeof(double) * nd, s->stream);
  if (e == hipSuccess) {
    if (s->prec == 4) {
      cpx<float>* H = (cpx<float>*)spec_slot(s);
      cpx<float>* Q = (cpx<float>*)s->Q;
      e = grid_sum<float>((const float*)rho_slot(s), cells, d, s->stream);
      if (!e) e = fft_r2c_x<float, float>((const float*)rho_slot(s), H, Q, n, *s->tab, s->stream);
      if (!e) e = fft_c2c_y<float>(H, Q, n, -1, *s->tab, s->stream);
      if (!e) e = fft_c2c_z<float>(H, Q, n, -1, *s->tab, s->stream);
      if (!e) e = shell_sums<float>(H, Q, n, n_bins, d + 1, s->stream);
    } else {
      cpx<double>* H = (cpx<double>*)spec_slot(s);
      cpx<double>* Q = (cpx<double>*)s->Q;
      e = grid_sum<double>((const double*)rho_slot(s), cells, d, s->stream);
      if (!e) e = fft_r2c_x<double, double>((const double*)rho_slot(s), H, Q, n, *s->tab, s->stream);
      if (!e) e = fft_c2c_y<double>(H, Q, n, -1, *s->tab, s->stream);
      if (!e) e = fft_c2c_z<double>(H, Q, n, -1, *s->tab, s->stream);
      if (!e) e = shell_sums<double>(H, Q, n, n_bins, d + 1, s->stream);
    }
  }
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, sizeof(double) * nd, hipMemcpyDeviceToHost, s->stream);
  hipError_t es = hipStreamSynchronize(s->stream);
  (void)hipFree(d);
  s->phi_valid = false;  // the potential buffer held the spectrum
  if (e != hipSuccess || es != hipSuccess)
    return fail(PM_E_HIP, "power spectrum failed: %s", hipGetErrorString(e ? e : es));
  const double mean = h[0] / (double)cells;
  if (!(mean > 0.0)) return fail(PM_E_STATE, "mean density is not positive");
  const double norm = box * box * box / ((double)cells * (double)cells) / (mean * mean);
  const double* psum = &h[1];
  const double* ksum = psum + (n_bins + 1);
  const double* cnt = ksum + (n_bins + 1);
  for (int b = 1; b <= n_bins; ++b) {
    counts[b - 1] = cnt[b];
    k[b - 1] = cnt[b] > 0 ? ksum[b] / cnt[b] : 0.0;
    pk[b - 1] = cnt[b] > 0 ? psum[b] / cnt[b] * norm : 0.0;
  }
  return PM_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// batched ensemble
// ------------------------------------------------------------------------------------------
struct pm_ensemble {
  std::vector<pm_sim*> members;  // in the order of adding
  std::vector<int> order;        // descriptor slot -> member index (size class 0 first)
  BatchDesc* host = nullptr;     // host image of the descriptor
  BatchDesc* d_desc = nullptr;
  double* d_table = nullptr;     // kEnsTableSteps x kBatchMax x {Green, kick, drift}
  int* d_step = nullptr;
  std::vector<double> table;
  hipStream_t stream = nullptr;
  hipGraphExec_t graph = nullptr;
  int graph_steps = 0;
  bool shape_dirty = true;  // membership changed: launch shapes and the graph are stale
  // launch shapes: FFT stages per size class, tile stages over all members
  long fft_base[2][BS_COUNT] = {}, fft_count[2][BS_COUNT] = {};
  int fft_lds[2][BS_COUNT] = {};
  long total[BS_COUNT] = {};
  bool timing = false;
  std::vector<PendingEvent> pending;
  double stage_ms[ST_COUNT] = {};
  int64_t stage_calls[ST_COUNT] = {};
};

namespace {

constexpr int kEnsTableSteps = 1024;  // steps per uploaded ensemble scalar table
// the batched kernels find a member's state buffer from the parity of the table's step counter
static_assert(kEnsTableSteps % 2 == 0, "every table chunk but the last must hold an even step count");

using EnsTimer = StageTimerT<pm_ensemble>;

void ens_drop_graph(pm_ensemble* e) {
  if (e->graph) (void)hipGraphExecDestroy(e->graph);
  e->graph = nullptr;
  e->graph_steps = 0;
}

// descriptor of the current members (host image) and the launch shapes
void ens_describe(pm_ensemble* e) {
  const int nm = (int)e->members.size();
  e->order.resize(nm);
  for (int i = 0; i < nm; ++i) e->order[i] = i;
  std::stable_sort(e->order.begin(), e->order.end(), [e](int x, int y) {
    return batch_class(e->members[x]->L.n) < batch_class(e->members[y]->L.n);
  });
  BatchDesc& D = *e->host;
  std::memset(&D, 0, sizeof(BatchDesc));
  D.n_members = nm;
  D.table = e->d_table;
  D.step = e->d_step;
  for (int st = 0; st < BS_COUNT; ++st) D.off[st][0] = 0;
  for (int c = 0; c < 2; ++c)
    for (int st = 0; st < BS_COUNT; ++st) {
      e->fft_base[c][st] = -1;
      e->fft_count[c][st] = 0;
      e->fft_lds[c][st] = 0;
    }
  for (int i = 0; i < nm; ++i) {
    const pm_sim* s = e->members[e->order[i]];
    BatchMember& m = D.m[i];
    const int n = s->L.n;
    m.rho = (float*)rho_slot(s);
    m.H = (cpx<float>*)spec_slot(s);
    m.Q = (cpx<float>*)s->Q;
    m.tw = (const cpx<float>*)s->tab->tw;
    m.sin2 = (const float*)s->tab->sin2;
    m.wdec = (const float*)s->tab->wdec;
    batch_fft_segments(n, m.ya, m.yb, m.za, m.zb);
    m.n = n;
    m.deconvolve = (s->flags & PM_FLAG_CIC_DECONVOLVE) ? 1 : 0;
    m.from_grid = s->fgrid ? 1 : 0;
    m.fgrid = s->fgrid;
    m.local_hist = (s->G.ntx >= 4 && s->G.nty >= 4 && s->G.ntz >= 4) ? 1 : 0;
    m.tile_scatter = (s->B.tile_scatter && tile_scatter_ok(s->G)) ? 1 : 0;
    for (int b = 0; b < 2; ++b) {
      m.s[b] = (float*)s->sb[b];
      m.v[b] = (float*)s->vb[b];
      m.id[b] = s->idb[b];
    }
    m.par = s->cur;
    m.mass = (float)s->mass;
    m.capacity = s->B.capacity;
    m.L = s->L;
    m.G = s->G;
    m.counts = s->B.counts;
    m.offsets = s->B.offsets;
    m.cursors = s->B.cursors;
    m.entries = s->B.entries;
    m.block_sums = s->B.block_sums;
    m.scan_done = s->B.done;
    m.health = s->B.health;
    m.det_fix = det_fix(s);
    m.scan_nb = scan_blocks(s->G);
    D.max_n = std::max(D.max_n, n);
    long blocks[BS_COUNT];
    blocks[BS_R2C] = batch_fft_blocks(n, BS_R2C);
    blocks[BS_Y] = batch_fft_blocks(n, BS_Y);
    blocks[BS_Z] = batch_fft_blocks(n, BS_Z);
    blocks[BS_C2R] = batch_fft_blocks(n, BS_C2R);
    blocks[BS_TILES] = s->G.ntiles;
    blocks[BS_GKD] = m.tile_scatter ? 0 : s->G.ntiles;
    blocks[BS_GKD_TS] = m.tile_scatter ? s->G.ntiles : 0;
    blocks[BS_SCAN] = m.scan_nb;
    blocks[BS_SCATTER] = m.tile_scatter ? 0 : batch_scatter_blocks(s->L.np);
    blocks[BS_TILE_SCATTER] = m.tile_scatter ? s->G.ntiles : 0;
    const int c = batch_class(n);
    for (int st = 0; st < BS_COUNT; ++st) {
      D.off[st][i + 1] = D.off[st][i] + blocks[st];
      if (st <= BS_C2R) {
        if (e->fft_base[c][st] < 0) e->fft_base[c][st] = D.off[st][i];
        e->fft_count[c][st] += blocks[st];
        e->fft_lds[c][st] = std::max(e->fft_lds[c][st], batch_fft_lds(n, st));
      }
    }
  }
  for (int st = 0; st < BS_COUNT; ++st) e->total[st] = D.off[st][nm];
}

// one table-driven step of every member (the enqueued work is identical for every step)
int ens_step(pm_ensemble* e) {
  const BatchDesc* d = e->d_desc;
  const hipStream_t st = e->stream;
  auto fft = [&](int stage, int mode) -> hipError_t {
    for (int c = 0; c < 2; ++c) {
      const long cnt = e->fft_count[c][stage];
      if (cnt <= 0) continue;
      hipError_t r;
      const long base = e->fft_base[c][stage];
      const int lds = e->fft_lds[c][stage];
      if (stage == BS_R2C) r = batch_r2c(d, c, base, cnt, lds, st);
      else if (stage == BS_C2R) r = batch_c2r(d, c, base, cnt, lds, st);
      else r = batch_pencil(d, c, mode, base, cnt, lds, st);
      if (r != hipSuccess) return r;
    }
    return hipSuccess;
  };
  {
    EnsTimer t(e, ST_R2C);
    PM_HIP(fft(BS_R2C, 0));
  }
  {
    EnsTimer t(e, ST_Y_FWD);
    PM_HIP(fft(BS_Y, -1));
  }
  {
    EnsTimer t(e, ST_Z_GREEN);
    PM_HIP(fft(BS_Z, 0));
  }
  {
    EnsTimer t(e, ST_Y_INV);
    PM_HIP(fft(BS_Y, +1));
  }
  {
    EnsTimer t(e, ST_C2R);
    PM_HIP(fft(BS_C2R, 0));
  }
  {
    EnsTimer t(e, ST_GKD);
    PM_HIP(batch_gkd(d, e->total[BS_GKD], e->total[BS_GKD_TS], st));
  }
  {
    EnsTimer t(e, ST_BIN);
    PM_HIP(batch_scan(d, e->total[BS_SCAN], (int)e->members.size(), st));
    PM_HIP(batch_scatter(d, e->total[BS_SCATTER], e->total[BS_TILE_SCATTER], st));
  }
  {
    EnsTimer t(e, ST_DEPOSIT);
    PM_HIP(batch_deposit(d, e->total[BS_TILES], st));
  }
  return PM_OK;
}

int ens_build_graph(pm_ensemble* e) {
  if (e->graph) return PM_OK;
  hipGraph_t graph = nullptr;
  PM_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  int rc = PM_OK;
  for (int k = 0; k < kGraphSteps && rc == PM_OK; ++k) rc = ens_step(e);
  hipError_t err = hipStreamEndCapture(e->stream, &graph);
  if (rc != PM_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (err != hipSuccess) return fail(PM_E_HIP, "ensemble graph capture failed: %s", hipGetErrorString(err));
  err = hipGraphInstantiate(&e->graph, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (err != hipSuccess) {
    e->graph = nullptr;
    return fail(PM_E_HIP, "ensemble graph instantiation failed: %s", hipGetErrorString(err));
  }
  e->graph_steps = kGraphSteps;
  return PM_OK;
}

int ens_check_member(const pm_sim* s) {
  if (s->prec != 4) return fail(PM_E_UNSUPPORTED, "ensemble members must be FP32 sims");
  if (!batch_size_supported(s->L.n))
    return fail(PM_E_UNSUPPORTED, "N_CELLS=%d is not a batched grid size (32, 48, 64, 96, 128, 192, 256, 384, 512)",
                s->L.n);
  if (s->G.tx != 16 || s->G.ty != 16 || s->G.tz != 16)
    return fail(PM_E_UNSUPPORTED, "ensemble members need 16^3 tiles");
  return PM_OK;
}

}  // namespace

extern "C" {

PM_API int pm_ensemble_create(pm_ensemble** out) {
  PM_INIT();
  if (!out) return fail(PM_E_INVALID, "null output handle");
  *out = nullptr;
  static std::once_flag once;
  static hipError_t init_err = hipSuccess;
  std::call_once(once, [] { init_err = batch_fft_init(); });
  if (init_err != hipSuccess) return fail(PM_E_HIP, "batched FFT set-up failed: %s", hipGetErrorString(init_err));
  pm_ensemble* e = new pm_ensemble();
  e->host = new BatchDesc();
  if (hipMalloc(&e->d_desc, sizeof(BatchDesc)) != hipSuccess ||
      hipMalloc(&e->d_table, sizeof(double) * 3 * kBatchMax * kEnsTableSteps) != hipSuccess ||
      hipMalloc(&e->d_step, sizeof(int)) != hipSuccess || hipStreamCreate(&e->stream) != hipSuccess) {
    pm_ensemble_destroy(e);
    return fail(PM_E_NOMEM, "ensemble allocation failed");
  }
  e->table.resize((size_t)3 * kBatchMax * kEnsTableSteps);
  *out = e;
  return PM_OK;
}

PM_API int pm_ensemble_destroy(pm_ensemble* e) {
  if (!e) return PM_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  for (pm_sim* s : e->members) s->ens = nullptr;
  for (auto& p : e->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  ens_drop_graph(e);
  (void)hipFree(e->d_desc);
  (void)hipFree(e->d_table);
  (void)hipFree(e->d_step);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e->host;
  delete e;
  return PM_OK;
}

PM_API int pm_ensemble_add(pm_ensemble* e, pm_sim* s) {
  if (!e || !s) return fail(PM_E_INVALID, "null argument");
  if (s->ens == e) return fail(PM_E_INVALID, "the sim is already a member");
  if (s->ens) return fail(PM_E_STATE, "the sim is a member of another ensemble");
  if ((int)e->members.size() >= kBatchMax) return fail(PM_E_UNSUPPORTED, "an ensemble holds at most %d sims", kBatchMax);
  if (s->G.tx != 16 || s->G.ty != 16 || s->G.tz != 16)  // members share the 16^3 tile kernels
    if (int rc = sim_use_tiles(s, make_tiles(s->L.n, s->prec, 16, 16, 16))) return rc;
  if (int rc = ens_check_member(s)) return rc;
  PM_HIP(hipStreamSynchronize(e->stream));
  e->members.push_back(s);
  s->ens = e;
  e->shape_dirty = true;
  return PM_OK;
}

PM_API int pm_ensemble_remove(pm_ensemble* e, pm_sim* s) {
  if (!e || !s) return fail(PM_E_INVALID, "null argument");
  auto it = std::find(e->members.begin(), e->members.end(), s);
  if (it == e->members.end()) return fail(PM_E_INVALID, "the sim is not a member");
  PM_HIP(hipStreamSynchronize(e->stream));
  e->members.erase(it);
  s->ens = nullptr;
  e->shape_dirty = true;
  return PM_OK;
}

PM_API int pm_ensemble_size(pm_ensemble* e, int* n) {
  if (!e || !n) return fail(PM_E_INVALID, "null argument");
  *n = (int)e->members.size();
  return PM_OK;
}

PM_API int pm_ensemble_run(pm_ensemble* e, double* a, const double* da, int n_steps) {
  if (!e) return fail(PM_E_INVALID, "null ensemble");
  if (n_steps < 0) return fail(PM_E_INVALID, "n_steps must be >= 0");
  const int nm = (int)e->members.size();
  if (nm == 0 || n_steps == 0) return PM_OK;
  if (!a || !da) return fail(PM_E_INVALID, "null a / da");
  for (int i = 0; i < nm; ++i) {
    pm_sim* s = e->members[i];
    if (!s->has_state) return fail(PM_E_STATE, "ensemble member %d has no state", i);
    if (int rc = ens_check_member(s)) return rc;
    // a member's own stream must be idle, and its density and bins valid, before the batch
    if (!s->rho_valid || !s->bins_valid)
      if (int rc = sim_deposit_impl<float>(s)) return rc;
    PM_HIP(hipStreamSynchronize(s->stream));
  }
  if (e->shape_dirty) {
    ens_drop_graph(e);
    e->shape_dirty = false;
  }
  // the descriptor is rewritten every call (member masses or Green's grids may have changed);
  // the launch shapes only change with the membership
  ens_describe(e);
  PM_HIP(hipMemcpyAsync(e->d_desc, e->host, sizeof(BatchDesc), hipMemcpyHostToDevice, e->stream));
  // A graph pays for its capture over many replays; right after a membership change a short
  // call is cheaper enqueued launch by launch.
  const bool use_graph = !e->timing && n_steps >= kGraphSteps && (e->graph || n_steps >= 8 * kGraphSteps);
  if (use_graph)
    if (int rc = ens_build_graph(e)) return rc;
  std::vector<double> acur(a, a + nm);
  int done = 0;
  while (done < n_steps) {
    const int chunk = std::min(n_steps - done, kEnsTableSteps);
    for (int k = 0; k < chunk; ++k)
      for (int slot = 0; slot < nm; ++slot) {
        const int i = e->order[slot];
        if (!(acur[i] > 0.0) || !(acur[i] + da[i] > 0.0))
          return fail(PM_E_INVALID, "scale factor of member %d must stay positive", i);
        const StepScalars c = step_scalars(e->members[i], acur[i], da[i]);
        double* t = &e->table[3 * ((size_t)k * nm + slot)];
        t[0] = c.green;
        t[1] = c.kick;
        t[2] = c.drift;
        acur[i] += da[i];  // pmesh.py:63
      }
    const int zero = 0;
    PM_HIP(hipMemcpyAsync(e->d_table, e->table.data(), sizeof(double) * 3 * (size_t)chunk * nm,
                          hipMemcpyHostToDevice, e->stream));
    PM_HIP(hipMemcpyAsync(e->d_step, &zero, sizeof(int), hipMemcpyHostToDevice, e->stream));
    int k = 0;
    if (use_graph)
      for (; k + e->graph_steps <= chunk; k += e->graph_steps) PM_HIP(hipGraphLaunch(e->graph, e->stream));
    for (; k < chunk; ++k)
      if (int rc = ens_step(e)) return rc;
    // the host table must outlive the asynchronous copies above
    for (pm_sim* s : e->members)
      if (int rc = health_fetch(s, e->stream)) return rc;
    PM_HIP(hipStreamSynchronize(e->stream));
    done += chunk;
    bool sick = false;
    for (pm_sim* s : e->members)
      sick = sick || (s->h_health[PM_HEALTH_NONFINITE] | s->h_health[PM_HEALTH_BINS] | s->h_health[PM_HEALTH_OVERFLOW]);
    if (sick) break;  // the steps done so far are kept; the member is named below
  }
  for (int i = 0; i < nm; ++i) {
    a[i] = acur[i];
    pm_sim* s = e->members[i];
    s->cur ^= done & 1;  // every step's gather switched the member's state buffer
    s->rho_valid = s->bins_valid = s->phi_valid = true;
    s->spec_valid = false;
  }
  for (int i = 0; i < nm; ++i)
    if (int rc = health_report(e->members[i], e->stream)) {
      g_err = "ensemble member " + std::to_string(i) + " (after " + std::to_string(done) + " steps): " + g_err;
      return rc;
    }
  return PM_OK;
}

PM_API int pm_ensemble_set_timing(pm_ensemble* e, int enable) {
  if (!e) return fail(PM_E_INVALID, "null ensemble");
  if (int rc = collect_timing(e)) return rc;
  e->timing = enable != 0;
  for (int i = 0; i < ST_COUNT; ++i) {
    e->stage_ms[i] = 0;
    e->stage_calls[i] = 0;
  }
  return PM_OK;
}

PM_API int pm_ensemble_get_timing(pm_ensemble* e, double* ms, int64_t* calls, int n_stages) {
  if (!e || !ms || !calls) return fail(PM_E_INVALID, "null argument");
  if (int rc = collect_timing(e)) return rc;
  for (int i = 0; i < n_stages && i < ST_COUNT; ++i) {
    ms[i] = e->stage_ms[i];
    calls[i] = e->stage_calls[i];
  }
  return PM_OK;
}

PM_API int pm_ensemble_sync(pm_ensemble* e) {
  if (!e) return fail(PM_E_INVALID, "null ensemble");
  PM_HIP(hipStreamSynchronize(e->stream));
  return PM_OK;
}

}  // extern "C"
