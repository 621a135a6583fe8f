// Host restatement of the PM step's tile binning — TEST INFRASTRUCTURE ONLY.
//
// Restates, in plain C++ with std::thread workers and std::atomic counters, what the device
// kernels of csrc/pm_tiles.hip do each step:
//   count    one count per (key = tile for the home cell, ntiles + tile for each tile the
//            particle's CIC cloud reaches on its upper faces)            k_bin_count / gather
//   scan     exclusive prefix of the 2 * ntiles counts -> offsets           k_scan_*
//   scatter  a cursor atomic per key gives each entry its place            k_bin_scatter
//   deposit  owner computes: a tile takes the corners of its own and halo entries that fall
//            inside it                                                     k_deposit_tiles
//   rows     the fused kernel's row block (all tiles of one tile-y, tile-z pair): own entries
//            are deposited in their home tile, halo entries in the tile whose halo list holds
//            them, found by a binary search over the list offsets          k_deposit_r2c
//   tile scatter (after a move of every particle): per source tile, the new positions are
//            counted in a 4^3 neighbourhood histogram whose bins reserve ranges of their keys
//            with one returning counter add each; particles that moved further go to far
//            counts; each particle leaves a 16-bit bin code at its storage index; the scan
//            places far counts after the reserved ranges and keeps the replaced own offsets;
//            the scatter places each source tile's particles through cursors offsets[key] +
//            range start                                  k_gkd_tiles<.., TS> / k_tile_scatter
// and checks the invariants the kernels rely on: the own segments are a permutation of the
// particles, every entry lies inside its segment and the entry array, both deposits equal a
// direct CIC deposit (density.py:17-47).  Built with -fsanitize=address,undefined and with
// -fsanitize=thread by tests/test_native.py, so an out-of-bounds index or a data race in this
// index arithmetic is reported on the CPU.  Exit status 0: all checks passed.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <random>
#include <thread>
#include <vector>

namespace {

std::atomic<int> g_failures{0};

void check(bool ok, const char* what, long a = 0, long b = 0) {
  if (!ok && g_failures.fetch_add(1) < 20) std::fprintf(stderr, "FAIL: %s (%ld, %ld)\n", what, a, b);
}

void parallel_for(long n, const std::function<void(long, long)>& body) {
  const int nt = 4;
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) {
    const long b = n * t / nt, e = n * (t + 1) / nt;
    th.emplace_back([&body, b, e] { body(b, e); });
  }
  for (auto& x : th) x.join();
}

struct Geom {
  int n, tx, ty, tz, ntx, nty, ntz, ntiles;
};

// canonical float32 position wrapped into [0, n) and its cell (pm_common.h wrap_pos, cell_of)
inline float wrap_pos(float x, float n) {
  if (x < 0.f) x += n;
  if (x >= n) x -= n;
  return x;
}
inline int cell_of(float x, int n) {
  float c = x > 0.f ? x : 0.f;
  const float hi = (float)(n - 1);
  c = c < hi ? c : hi;
  return (int)std::floor(c);
}

struct Particle {
  int c[3];
  double d[3];
  int t[3];
  int up[3];
};

Particle classify(const float x[3], const Geom& G) {
  Particle p{};
  const int T[3] = {G.tx, G.ty, G.tz}, NT[3] = {G.ntx, G.nty, G.ntz};
  for (int a = 0; a < 3; ++a) {
    const float w = wrap_pos(x[a], (float)G.n);
    p.c[a] = cell_of(w, G.n);
    p.d[a] = (double)w - p.c[a];
    p.t[a] = p.c[a] / T[a];
    p.up[a] = (p.c[a] % T[a] == T[a] - 1) && NT[a] > 1;
  }
  return p;
}

inline bool active(const Particle& p, int o) {
  return ((o & 1) <= p.up[0]) && (((o >> 1) & 1) <= p.up[1]) && ((o >> 2) <= p.up[2]);
}

inline uint32_t key_of(const Particle& p, int o, const Geom& G) {
  int x = p.t[0] + (o & 1), y = p.t[1] + ((o >> 1) & 1), z = p.t[2] + (o >> 2);
  if (x == G.ntx) x = 0;
  if (y == G.nty) y = 0;
  if (z == G.ntz) z = 0;
  return (uint32_t)((z * G.nty + y) * G.ntx + x) + (o ? (uint32_t)G.ntiles : 0u);
}

// corners of particle p that fall inside tile `tile`, added to grid (density.py:37-47)
void deposit_in_tile(const Particle& p, int tile, const Geom& G, double mass, std::vector<double>& grid,
                     std::vector<std::atomic<int>>& owner_hits) {
  const int T[3] = {G.tx, G.ty, G.tz};
  const int o0[3] = {(tile % G.ntx) * G.tx, ((tile / G.ntx) % G.nty) * G.ty, (tile / (G.ntx * G.nty)) * G.tz};
  int r0[3], r1[3];
  for (int a = 0; a < 3; ++a) {
    int lo = p.c[a] - o0[a];
    if (lo < 0) lo += G.n;
    int hi = (p.c[a] + 1 == G.n ? 0 : p.c[a] + 1) - o0[a];
    if (hi < 0) hi += G.n;
    r0[a] = lo < T[a] ? lo : -1;
    r1[a] = hi < T[a] ? hi : -1;
  }
  const long n = G.n;
  for (int oz = 0; oz < 2; ++oz)
    for (int oy = 0; oy < 2; ++oy)
      for (int ox = 0; ox < 2; ++ox) {
        const int rz = oz ? r1[2] : r0[2], ry = oy ? r1[1] : r0[1], rx = ox ? r1[0] : r0[0];
        if (rz < 0 || ry < 0 || rx < 0) continue;
        const double w = mass * (oz ? p.d[2] : 1 - p.d[2]) * (oy ? p.d[1] : 1 - p.d[1]) * (ox ? p.d[0] : 1 - p.d[0]);
        const long cell = ((long)(o0[2] + rz) * n + (o0[1] + ry)) * n + (o0[0] + rx);
        check(cell >= 0 && cell < n * n * n, "deposit cell in range", cell);
        grid.at(cell) += w;
        owner_hits.at(cell).fetch_add(1, std::memory_order_relaxed);
      }
}

int run_case(const char* name, const Geom& G, const std::vector<float>& x, double mass) {
  const int before = g_failures.load();
  const long np = (long)x.size() / 3;
  const long m = 2L * G.ntiles;
  std::vector<Particle> P(np);
  parallel_for(np, [&](long b, long e) {
    for (long i = b; i < e; ++i) {
      const float xi[3] = {x[i], x[np + i], x[2 * np + i]};
      P[i] = classify(xi, G);
    }
  });
  // count
  std::vector<std::atomic<uint32_t>> counts(m);
  for (auto& c : counts) c.store(0);
  parallel_for(np, [&](long b, long e) {
    for (long i = b; i < e; ++i)
      for (int o = 0; o < 8; ++o)
        if (active(P[i], o)) counts.at(key_of(P[i], o, G)).fetch_add(1, std::memory_order_relaxed);
  });
  // scan
  std::vector<uint32_t> offsets(m + 1);
  uint64_t run = 0;
  for (long k = 0; k < m; ++k) {
    offsets[k] = (uint32_t)run;
    run += counts[k].load();
  }
  offsets[m] = (uint32_t)run;
  const uint64_t capacity = 8ull * (uint64_t)np;
  check(run <= capacity, "entries fit the capacity", (long)run, (long)capacity);
  check(offsets[G.ntiles] == (uint32_t)np, "own segments hold every particle once", offsets[G.ntiles], np);
  // scatter
  std::vector<uint32_t> entries(capacity ? capacity : 1, 0xffffffffu);
  std::vector<std::atomic<uint32_t>> cursors(m);
  for (auto& c : cursors) c.store(0);
  std::atomic<long> lost{0};
  parallel_for(np, [&](long b, long e) {
    for (long i = b; i < e; ++i)
      for (int o = 0; o < 8; ++o)
        if (active(P[i], o)) {
          const uint32_t k = key_of(P[i], o, G);
          const uint64_t pos = offsets[k] + (uint64_t)cursors[k].fetch_add(1, std::memory_order_relaxed);
          if (pos < offsets[k + 1] && pos < capacity) entries[pos] = (uint32_t)i;
          else lost.fetch_add(1);
        }
  });
  check(lost.load() == 0, "every entry found a place", lost.load());
  std::vector<char> seen(np, 0);
  for (int t = 0; t < G.ntiles; ++t)
    for (uint32_t e = offsets[t]; e < offsets[t + 1]; ++e) {
      const uint32_t i = entries.at(e);
      check(i < (uint32_t)np, "own entry is a particle", e, i);
      if (i >= (uint32_t)np) continue;
      check(!seen[i], "particle entered twice", i);
      seen[i] = 1;
      check(key_of(P[i], 0, G) == (uint32_t)t, "own entry in its home tile", i, t);
    }
  // direct deposit (density.py:17-47)
  const long cells = (long)G.n * G.n * G.n;
  std::vector<double> direct(cells, 0.0);
  for (long i = 0; i < np; ++i) {
    const Particle& p = P[i];
    for (int o = 0; o < 8; ++o) {
      const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
      const long cx = (p.c[0] + ox) % G.n, cy = (p.c[1] + oy) % G.n, cz = (p.c[2] + oz) % G.n;
      direct[(cz * G.n + cy) * G.n + cx] +=
          mass * (oz ? p.d[2] : 1 - p.d[2]) * (oy ? p.d[1] : 1 - p.d[1]) * (ox ? p.d[0] : 1 - p.d[0]);
    }
  }
  // owner-computes deposit per tile (own + halo entries, corners inside the tile)
  std::vector<double> tiles(cells, 0.0);
  std::vector<std::atomic<int>> hits(cells);
  for (auto& h : hits) h.store(0);
  parallel_for(G.ntiles, [&](long b, long e) {
    for (long t = b; t < e; ++t) {
      for (uint32_t k = offsets[t]; k < offsets[t + 1]; ++k) deposit_in_tile(P[entries[k]], (int)t, G, mass, tiles, hits);
      for (uint32_t k = offsets[G.ntiles + t]; k < offsets[G.ntiles + t + 1]; ++k)
        deposit_in_tile(P[entries[k]], (int)t, G, mass, tiles, hits);
    }
  });
  double err = 0.0, norm = 0.0;
  for (long c = 0; c < cells; ++c) {
    err = std::max(err, std::fabs(tiles[c] - direct[c]));
    norm = std::max(norm, std::fabs(direct[c]));
  }
  check(err <= 1e-12 * (norm > 0 ? norm : 1), "tile deposit equals the direct deposit", (long)(err * 1e15));
  // fused row blocks: own entries in their home tile, halo entries in the list's tile
  std::vector<double> rows(cells, 0.0);
  std::vector<std::atomic<int>> hits2(cells);
  for (auto& h : hits2) h.store(0);
  parallel_for((long)G.nty * G.ntz, [&](long b, long e) {
    for (long rb = b; rb < e; ++rb) {
      const uint32_t t0 = (uint32_t)rb * G.ntx, nt = (uint32_t)G.ntiles;
      const uint32_t oe = offsets[t0 + G.ntx], ob = offsets[t0];
      const uint32_t he = offsets[nt + t0 + G.ntx], hb = offsets[nt + t0];
      const uint32_t n_own = oe - ob, total = n_own + (he - hb);
      for (uint32_t i = 0; i < total; ++i) {
        const uint32_t ent = i < n_own ? ob + i : hb + (i - n_own);
        const Particle& p = P[entries.at(ent)];
        int tx;
        if (i < n_own) {
          tx = p.t[0];
          check((long)(p.t[2] * G.nty + p.t[1]) == rb, "own entry in its row block", rb, i);
        } else {
          int lo = 0, hi = G.ntx - 1;
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (offsets.at(nt + t0 + mid) <= ent) lo = mid;
            else hi = mid - 1;
          }
          tx = lo;
          check(ent >= offsets[nt + t0 + tx] && ent < offsets[nt + t0 + tx + 1], "halo search finds the list", ent, tx);
        }
        deposit_in_tile(p, (int)(t0 + tx), G, mass, rows, hits2);
      }
    }
  });
  err = 0.0;
  for (long c = 0; c < cells; ++c) err = std::max(err, std::fabs(rows[c] - direct[c]));
  check(err <= 1e-12 * (norm > 0 ? norm : 1), "row-block deposit equals the direct deposit", (long)(err * 1e15));
  std::printf("%-34s n=%d tiles %dx%dx%d np=%ld entries=%llu: %s\n", name, G.n, G.tx, G.ty, G.tz, np,
              (unsigned long long)run, g_failures.load() == before ? "ok" : "FAILED");
  return g_failures.load() - before;
}

// ---- tile scatter -------------------------------------------------------------------------
// neighbourhood bin of the new home tile tn of a particle from source tile tc (false: beyond
// the 4^3 neighbourhood), the bin of its corner offset o, and the key a bin names
bool near_home_bin(const int tn[3], const int tc[3], const Geom& G, int& h0) {
  const int ntc[3] = {G.ntx, G.nty, G.ntz};
  int rb[3];
  bool near = true;
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
inline int corner_bin(int h0, int o) {
  return o == 0 ? h0 : h0 + 2 * (((o >> 2) * 4 + ((o >> 1) & 1)) * 4 + (o & 1)) + 1;
}
inline int fold(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }
uint32_t bin_key(const int tc[3], int bin, const Geom& G) {
  const int cell = bin >> 1, kind = bin & 1;
  const int rx = cell & 3, ry = (cell >> 2) & 3, rz = cell >> 4;
  const int x = fold(tc[0] + rx - 1, G.ntx), y = fold(tc[1] + ry - 1, G.nty), z = fold(tc[2] + rz - 1, G.ntz);
  return (uint32_t)((z * G.nty + y) * G.ntx + x) + (kind ? (uint32_t)G.ntiles : 0u);
}
constexpr uint32_t kCodeFar = 0x8000u;

// Bins the particles at x0 (count, scan, particle-parallel scatter), moves every particle to x1
// (stored, as the gather does, at its own-entry index) and bins the moved particles with the
// tile scatter; checks the new own segments are a permutation of the storage indices, every
// entry is in the segment of its key, each key holds exactly its particles, and the owner-
// computes deposit of the new bins equals a direct deposit of x1.
int tile_scatter_case(const char* name, const Geom& G, const std::vector<float>& x0, const std::vector<float>& x1,
                      double mass) {
  const int before = g_failures.load();
  const long np = (long)x0.size() / 3;
  const long m = 2L * G.ntiles;
  const uint64_t capacity = 8ull * (uint64_t)np;
  auto classify_all = [&](const std::vector<float>& x, std::vector<Particle>& P) {
    P.resize(np);
    parallel_for(np, [&](long b, long e) {
      for (long i = b; i < e; ++i) {
        const float xi[3] = {x[i], x[np + i], x[2 * np + i]};
        P[i] = classify(xi, G);
      }
    });
  };
  std::vector<Particle> P0, P1;
  classify_all(x0, P0);
  classify_all(x1, P1);
  // first binning (count, scan, particle-parallel scatter) of x0
  std::vector<uint32_t> off0(m + 1);
  std::vector<uint32_t> ent0(capacity ? capacity : 1, 0xffffffffu);
  {
    std::vector<std::atomic<uint32_t>> counts(m), cursors(m);
    for (long k = 0; k < m; ++k) counts[k].store(0), cursors[k].store(0);
    for (long i = 0; i < np; ++i)
      for (int o = 0; o < 8; ++o)
        if (active(P0[i], o)) counts[key_of(P0[i], o, G)].fetch_add(1);
    uint64_t run = 0;
    for (long k = 0; k < m; ++k) off0[k] = (uint32_t)run, run += counts[k].load();
    off0[m] = (uint32_t)run;
    for (long i = 0; i < np; ++i)
      for (int o = 0; o < 8; ++o)
        if (active(P0[i], o)) {
          const uint32_t k = key_of(P0[i], o, G);
          ent0.at(off0[k] + cursors[k].fetch_add(1)) = (uint32_t)i;
        }
  }
  // gather: tile t's own entry e holds particle ent0[e]; its moved state is stored at index e
  std::vector<Particle> S(np);  // moved particle at each new storage index
  std::vector<uint32_t> who(np);
  std::vector<uint16_t> code(np);
  std::vector<std::atomic<uint32_t>> counts(m), far(m);
  for (long k = 0; k < m; ++k) counts[k].store(0), far[k].store(0);
  std::vector<uint32_t> base((size_t)G.ntiles * 128, 0xffffffffu);
  parallel_for(G.ntiles, [&](long b, long e_) {
    for (long t = b; t < e_; ++t) {
      const int tc[3] = {(int)(t % G.ntx), (int)((t / G.ntx) % G.nty), (int)(t / ((long)G.ntx * G.nty))};
      uint32_t hist[128] = {};
      for (uint32_t e = off0[t]; e < off0[t + 1]; ++e) {
        const uint32_t i = ent0.at(e);
        check(i < (uint32_t)np, "old own entry is a particle", e, i);
        S[e] = P1[i];
        who[e] = i;
        int h0;
        if (near_home_bin(S[e].t, tc, G, h0)) {
          check(h0 >= 0 && h0 < 128 && (h0 & 1) == 0, "home bin in the neighbourhood", h0);
          code[e] = (uint16_t)((uint32_t)(h0 >> 1) | (uint32_t)(S[e].up[0] | (S[e].up[1] << 1) | (S[e].up[2] << 2)) << 6);
          for (int o = 0; o < 8; ++o)
            if (active(S[e], o)) {
              const int bn = corner_bin(h0, o);
              check(bn >= 0 && bn < 128, "corner bin in range", bn, o);
              check(bin_key(tc, bn, G) == key_of(S[e], o, G), "bin names the corner's key", bn, o);
              ++hist[bn];
            }
        } else {
          code[e] = (uint16_t)kCodeFar;
          for (int o = 0; o < 8; ++o)
            if (active(S[e], o)) far[key_of(S[e], o, G)].fetch_add(1);
        }
      }
      // end of tile: one returning add per non-empty bin reserves its range of the key
      for (int bn = 0; bn < 128; ++bn)
        base[(size_t)t * 128 + bn] = hist[bn] ? counts[bin_key(tc, bn, G)].fetch_add(hist[bn]) : 0xffffffffu;
    }
  });
  // scan: key count = reserved + far; far cursors start after the reserved ranges; keep the
  // replaced own offsets
  std::vector<uint32_t> off1(m + 1), old(off0.begin(), off0.begin() + G.ntiles + 1);
  std::vector<std::atomic<uint32_t>> cursors(m);
  uint64_t run = 0;
  for (long k = 0; k < m; ++k) {
    off1[k] = (uint32_t)run;
    run += counts[k].load() + far[k].load();
    cursors[k].store(counts[k].load());
  }
  off1[m] = (uint32_t)run;
  check(run <= capacity, "entries fit the capacity", (long)run, (long)capacity);
  check(off1[G.ntiles] == (uint32_t)np, "own segments hold every particle once", off1[G.ntiles], np);
  // tile scatter: per source tile, LDS-style cursors from the reserved range starts
  std::vector<uint32_t> ent1(capacity ? capacity : 1, 0xffffffffu);
  std::atomic<long> lost{0};
  parallel_for(G.ntiles, [&](long b, long e_) {
    for (long t = b; t < e_; ++t) {
      const int tc[3] = {(int)(t % G.ntx), (int)((t / G.ntx) % G.nty), (int)(t / ((long)G.ntx * G.nty))};
      uint32_t cur[128], lim[128];
      for (int bn = 0; bn < 128; ++bn) {
        const uint32_t w = base[(size_t)t * 128 + bn];
        cur[bn] = lim[bn] = 0;
        if (w != 0xffffffffu) {
          const uint32_t k = bin_key(tc, bn, G);
          cur[bn] = off1[k] + w;
          lim[bn] = off1[k + 1];
        }
      }
      for (uint32_t e = old[t]; e < old[t + 1]; ++e) {
        const uint32_t c = code.at(e);
        if (!(c & kCodeFar)) {
          const int h0 = 2 * (int)(c & 63u);
          const Particle& p = S[e];
          for (int o = 0; o < 8; ++o)
            if (active(p, o)) {
              const int bn = corner_bin(h0, o);
              const uint32_t pos = cur[bn]++;
              if (pos < lim[bn] && pos < capacity) ent1[pos] = e;
              else lost.fetch_add(1);
            }
        } else {
          for (int o = 0; o < 8; ++o)
            if (active(S[e], o)) {
              const uint32_t k = key_of(S[e], o, G);
              const uint64_t pos = off1[k] + (uint64_t)cursors[k].fetch_add(1);
              if (pos < off1[k + 1] && pos < capacity) ent1[pos] = e;
              else lost.fetch_add(1);
            }
        }
      }
    }
  });
  check(lost.load() == 0, "every entry found a place", lost.load());
  // each key's segment holds exactly the storage indices whose particle has a corner there
  std::vector<std::vector<uint32_t>> expect(m);
  for (long e = 0; e < np; ++e)
    for (int o = 0; o < 8; ++o)
      if (active(S[e], o)) expect[key_of(S[e], o, G)].push_back((uint32_t)e);
  long nfar = 0;
  for (long e = 0; e < np; ++e) nfar += (code[e] & kCodeFar) ? 1 : 0;
  for (long k = 0; k < m; ++k) {
    std::vector<uint32_t> got(ent1.begin() + off1[k], ent1.begin() + off1[k + 1]);
    std::sort(got.begin(), got.end());
    check(got == expect[k], "key segment holds its particles", k, (long)got.size());
  }
  // deposit of the moved particles from the new bins equals the direct one
  const long cells = (long)G.n * G.n * G.n;
  std::vector<double> direct(cells, 0.0), tiles(cells, 0.0);
  std::vector<std::atomic<int>> hits(cells);
  for (auto& h : hits) h.store(0);
  for (long e = 0; e < np; ++e) {
    const Particle& p = S[e];
    for (int o = 0; o < 8; ++o) {
      const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
      const long cx = (p.c[0] + ox) % G.n, cy = (p.c[1] + oy) % G.n, cz = (p.c[2] + oz) % G.n;
      direct[(cz * G.n + cy) * G.n + cx] +=
          mass * (oz ? p.d[2] : 1 - p.d[2]) * (oy ? p.d[1] : 1 - p.d[1]) * (ox ? p.d[0] : 1 - p.d[0]);
    }
  }
  parallel_for(G.ntiles, [&](long b, long e) {
    for (long t = b; t < e; ++t) {
      for (uint32_t k = off1[t]; k < off1[t + 1]; ++k) deposit_in_tile(S[ent1[k]], (int)t, G, mass, tiles, hits);
      for (uint32_t k = off1[G.ntiles + t]; k < off1[G.ntiles + t + 1]; ++k)
        deposit_in_tile(S[ent1[k]], (int)t, G, mass, tiles, hits);
    }
  });
  double err = 0.0, norm = 0.0;
  for (long c = 0; c < cells; ++c) {
    err = std::max(err, std::fabs(tiles[c] - direct[c]));
    norm = std::max(norm, std::fabs(direct[c]));
  }
  check(err <= 1e-12 * (norm > 0 ? norm : 1), "deposit from the tile-scattered bins equals the direct deposit",
        (long)(err * 1e15));
  std::printf("%-34s n=%d tiles %dx%dx%d np=%ld far=%ld: %s\n", name, G.n, G.tx, G.ty, G.tz, np, nfar,
              g_failures.load() == before ? "ok" : "FAILED");
  return g_failures.load() - before;
}

Geom geom(int n, int tx, int ty, int tz) {
  Geom g{n, tx, ty, tz, n / tx, n / ty, n / tz, 0};
  g.ntiles = g.ntx * g.nty * g.ntz;
  return g;
}

}  // namespace

int main() {
  std::mt19937_64 rng(12345);
  auto uniform = [&](int n, long np) {
    std::vector<float> x(3 * np);
    std::uniform_real_distribution<float> u(0.f, (float)n);
    for (auto& v : x) v = u(rng);
    return x;
  };
  // uniform particles, several tile shapes (including one tile along an axis)
  run_case("uniform 32x8x8", geom(64, 32, 8, 8), uniform(64, 20000), 8.0);
  run_case("uniform 16^3", geom(64, 16, 16, 16), uniform(64, 20000), 8.0);
  run_case("uniform 8x8x8", geom(32, 8, 8, 8), uniform(32, 5000), 1.0);
  run_case("uniform 64x8x8 (one x tile)", geom(64, 64, 8, 8), uniform(64, 8000), 8.0);
  // every particle on the upper faces of tiles: the most halo entries a particle can have
  {
    const Geom G = geom(64, 32, 8, 8);
    const long np = 4096;
    std::vector<float> x(3 * np);
    std::uniform_real_distribution<float> f(0.f, 1.f);
    for (long i = 0; i < np; ++i) {
      x[i] = (float)(32 * (i % 2) + 31) + f(rng);
      x[np + i] = (float)(8 * (i % 8) + 7) + f(rng);
      x[2 * np + i] = (float)(8 * ((i / 8) % 8) + 7) + f(rng);
    }
    run_case("all on upper tile faces", G, x, 8.0);
  }
  // all particles in one tile (the clustered limit), and positions at the wrap edge
  {
    const Geom G = geom(64, 32, 8, 8);
    const long np = 6000;
    std::vector<float> x(3 * np);
    std::uniform_real_distribution<float> f(0.f, 1.f);
    for (long i = 0; i < np; ++i) {
      x[i] = 40.f + 24.f * f(rng);
      x[np + i] = 56.f + 8.f * f(rng);
      x[2 * np + i] = 0.f + 8.f * f(rng);
    }
    run_case("one tile holds every particle", G, x, 8.0);
    const float below = std::nextafter(64.f, 0.f);
    for (long i = 0; i < np; ++i) {
      x[i] = (i % 3 == 0) ? below : (i % 3 == 1 ? 64.f : 0.f);  // N - ulp, N (wraps to 0), 0
      x[np + i] = (i % 5 == 0) ? below : 63.5f;
      x[2 * np + i] = (float)(i % 64) + 0.999999f;
    }
    run_case("wrap edge N-ulp / N / 0", G, x, 8.0);
  }
  // tile scatter: small moves, moves of two or more tiles for some particles (far counts), the
  // smallest neighbourhood that names its tiles uniquely (4 tiles per axis), moves across the
  // periodic edge
  {
    auto moved = [&](const std::vector<float>& x, int n, float small, float far_frac, float far_dist) {
      std::vector<float> y(x);
      std::uniform_real_distribution<float> u(-1.f, 1.f), f(0.f, 1.f);
      for (auto& v : y) {
        const float dv = f(rng) < far_frac ? far_dist * (u(rng) < 0 ? -1.f : 1.f) : small * u(rng);
        v = wrap_pos(v + dv, (float)n);
      }
      return y;
    };
    const std::vector<float> a = uniform(64, 20000);
    run_case("tile scatter: first binning", geom(64, 32, 8, 8), a, 8.0);
    tile_scatter_case("tile scatter 32x8x8, small moves", geom(64, 32, 8, 8), a, moved(a, 64, 3.f, 0.f, 0.f), 8.0);
    tile_scatter_case("tile scatter 32x8x8, 5% far", geom(64, 32, 8, 8), a, moved(a, 64, 3.f, 0.05f, 20.f), 8.0);
    tile_scatter_case("tile scatter 16^3 (4 per axis), far", geom(64, 16, 16, 16), a, moved(a, 64, 6.f, 0.05f, 33.f),
                      8.0);
    const std::vector<float> b = uniform(128, 40000);
    tile_scatter_case("tile scatter 16x16x8, across edges", geom(128, 16, 16, 8), b,
                      moved(b, 128, 9.f, 0.02f, 60.f), 8.0);
  }
  if (g_failures.load()) {
    std::fprintf(stderr, "%d check(s) failed\n", g_failures.load());
    return 1;
  }
  return 0;
}
