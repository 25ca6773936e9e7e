// Tile-binned, deterministic paint and LDS-staged readout (float32 storage).
//
// The force mesh is cut into TILE^3-cell tiles.  Each particle belongs to the
// tile holding the first point of its window (the CIC floor cell, or the TSC
// nearest point minus one), so its deposit lies in that tile's box of
// (TILE + order - 1)^3 cells.
//
//   bin      : count particles per tile (integer atomics), exclusive scan,
//              scatter particle indices into per-tile lists.
//   paint    : one workgroup per tile accumulates its particles' deposits in
//              LDS as 64-bit fixed point (integer adds, so the sum does not
//              depend on the order particles arrive in), writes the tile's own
//              cells once (no clearing pass, no global atomics) and the
//              overflow cells (the box beyond the tile) to a per-tile halo
//              buffer; a second kernel adds the neighbours' halos to each
//              tile's low faces in a fixed neighbour order.  The result is
//              bitwise reproducible from run to run.
//   readout  : one workgroup per tile stages the box of every requested mesh
//              in LDS (each mesh cell read once, plus the (1/TILE) halo) and
//              interpolates for the tile's particles.
#include "fpm_common.h"

#include <cmath>

namespace fpm {

template <typename T>
int make_desc(const fpm_particles* part, const fpm_mesh* m, PartDesc<T>& P);

constexpr int TILE = 16;
constexpr int kTileThreads = 512;

struct TileGeom {
  int nm[3];
  int nt[3];       // tiles per axis
  int64_t ntiles;
  int H;           // order - 1: box overhang
  int TB;          // TILE + H
  int64_t halo;    // overflow cells per tile
};

static TileGeom tile_geom(const fpm_mesh* m, int order) {
  TileGeom g{};
  g.ntiles = 1;
  for (int d = 0; d < 3; ++d) {
    g.nm[d] = m->nmesh[d];
    g.nt[d] = m->nmesh[d] / TILE;
    g.ntiles *= g.nt[d];
  }
  g.H = order - 1;
  g.TB = TILE + g.H;
  g.halo = (int64_t)g.H * g.TB * g.TB + (int64_t)TILE * g.H * g.TB + (int64_t)TILE * TILE * g.H;
  return g;
}

// Index of an overflow cell (at least one local coordinate >= TILE) in the
// tile's halo record: x-slab, then y-slab, then z-slab.
__device__ __forceinline__ int halo_index(int x, int y, int z, int H, int TB) {
  if (x >= TILE) return ((x - TILE) * TB + y) * TB + z;
  if (y >= TILE) return H * TB * TB + (x * H + (y - TILE)) * TB + z;
  return H * TB * TB + TILE * H * TB + (x * TILE + y) * H + (z - TILE);
}

struct Workspace {
  int32_t* key;
  int32_t* list;
  int32_t* count;
  int32_t* offs;
  int32_t* cursor;
  int32_t* misplaced;  // count of particles found outside their binned tile
  void* halo;
  int64_t bytes;
};

static int64_t align256(int64_t x) { return (x + 255) / 256 * 256; }

// Byte offsets of the workspace sections: key, list, count, offs, cursor,
// misplaced, halo; off[7] is the total size.
static void ws_offsets(const TileGeom& g, int64_t n, size_t real, int64_t off[8]) {
  const int64_t sizes[7] = {n * 4, n * 4, g.ntiles * 4, (g.ntiles + 1) * 4, g.ntiles * 4, 4,
                            g.ntiles * g.halo * (int64_t)real};
  int64_t o = 0;
  for (int i = 0; i < 7; ++i) {
    off[i] = o;
    o += align256(sizes[i]);
  }
  off[7] = o;
}

static int64_t layout_ws(const TileGeom& g, int64_t n, size_t real, char* base, Workspace* w) {
  int64_t off[8];
  ws_offsets(g, n, real, off);
  if (w) {
    w->key = reinterpret_cast<int32_t*>(base + off[0]);
    w->list = reinterpret_cast<int32_t*>(base + off[1]);
    w->count = reinterpret_cast<int32_t*>(base + off[2]);
    w->offs = reinterpret_cast<int32_t*>(base + off[3]);
    w->cursor = reinterpret_cast<int32_t*>(base + off[4]);
    w->misplaced = reinterpret_cast<int32_t*>(base + off[5]);
    w->halo = base + off[6];
    w->bytes = off[7];
  }
  return off[7];
}

// Window origin (wrapped) and weights, identical to the plain paint/readout.
template <typename T, int ORDER>
__device__ __forceinline__ void window_of(const PartDesc<T>& P, int64_t p, const int nm[3],
                                          int i0[3], T w[3][ORDER]) {
  int64_t cell[3];
  T f[3];
  particle_cell(P, p, cell, f);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int64_t o;
    if constexpr (ORDER == 2) {
      o = cell[d];
      w[d][0] = (T)1 - f[d];
      w[d][1] = f[d];
    } else {
      const bool up = f[d] >= (T)0.5;
      const T dd = up ? f[d] - (T)1 : f[d];
      o = cell[d] + (up ? 1 : 0) - 1;
      const T a = (T)0.5 - dd, b = (T)0.5 + dd;
      w[d][0] = (T)0.5 * a * a;
      w[d][1] = (T)0.75 - dd * dd;
      w[d][2] = (T)0.5 * b * b;
    }
    i0[d] = (int)wrapi(o, nm[d]);
  }
}

__device__ __forceinline__ int64_t tile_of(const int i0[3], const TileGeom& g) {
  return ((int64_t)(i0[0] / TILE) * g.nt[1] + i0[1] / TILE) * g.nt[2] + i0[2] / TILE;
}

// ---------------------------------------------------------------- binning
__global__ void k_zero_i32(int32_t* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// Lanes of a wave that fall in the same tile are merged before touching the
// counters (consecutive lattice particles mostly share tiles), so each
// distinct tile in a wave costs one atomic.
template <typename T, int ORDER>
__global__ void k_bin_count(PartDesc<T> P, TileGeom g, Workspace w) {
  const int lane = threadIdx.x & 63;
  for (int64_t base = blockIdx.x * (int64_t)blockDim.x; base < P.n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = base + threadIdx.x;
    const bool valid = p < P.n;
    int32_t t = -1;
    if (valid) {
      int i0[3];
      T wt[3][ORDER];
      window_of<T, ORDER>(P, p, g.nm, i0, wt);
      t = (int32_t)tile_of(i0, g);
      w.key[p] = t;
    }
    unsigned long long active = __ballot(valid);
    while (active) {
      const int leader = __ffsll(active) - 1;
      const int32_t tl = __shfl(t, leader);
      const unsigned long long same = __ballot(valid && t == tl);
      if (lane == leader) atomicAdd(&w.count[tl], (int32_t)__popcll(same));
      active &= ~same;
    }
  }
}

// Exclusive scan of the tile counts in one workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void k_bin_scan(Workspace w, int64_t ntiles) {
  __shared__ int32_t part[1024];
  const int tid = threadIdx.x;
  const int64_t per = (ntiles + 1023) / 1024;
  const int64_t b = tid * per;
  const int64_t e = b + per < ntiles ? b + per : ntiles;
  int32_t s = 0;
  for (int64_t i = b; i < e; ++i) s += w.count[i];
  part[tid] = s;
  __syncthreads();
  // Hillis-Steele inclusive scan over the 1024 partial sums
  for (int o = 1; o < 1024; o <<= 1) {
    const int32_t v = tid >= o ? part[tid - o] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int32_t run = tid ? part[tid - 1] : 0;
  for (int64_t i = b; i < e; ++i) {
    w.offs[i] = run;
    w.cursor[i] = run;
    run += w.count[i];
  }
  if (tid == 1023) w.offs[ntiles] = part[1023];
}

__global__ void k_bin_scatter(int64_t n, Workspace w) {
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int64_t base = blockIdx.x * (int64_t)blockDim.x; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = base + threadIdx.x;
    const bool valid = p < n;
    const int32_t t = valid ? w.key[p] : -1;
    unsigned long long active = __ballot(valid);
    while (active) {
      const int leader = __ffsll(active) - 1;
      const int32_t tl = __shfl(t, leader);
      const unsigned long long same = __ballot(valid && t == tl);
      int32_t slot0 = 0;
      if (lane == leader) slot0 = atomicAdd(&w.cursor[tl], (int32_t)__popcll(same));
      slot0 = __shfl(slot0, leader);
      if (valid && t == tl) w.list[slot0 + __popcll(same & below)] = (int32_t)p;
      active &= ~same;
    }
  }
}

// ---------------------------------------------------------------- paint
template <typename T, int ORDER>
__global__ __launch_bounds__(kTileThreads) void k_paint_tiles(PartDesc<T> P, TileGeom g,
                                                               Workspace w, T* mesh, double mass_fx,
                                                               double inv_fx) {
  constexpr int H = ORDER - 1;
  constexpr int TB = TILE + H;
  constexpr int NB = TB * TB * TB;
  __shared__ unsigned long long acc[NB];
  const int64_t t = xcd_remap(blockIdx.x, gridDim.x);
  const int tz = (int)(t % g.nt[2]);
  const int ty = (int)((t / g.nt[2]) % g.nt[1]);
  const int tx = (int)(t / ((int64_t)g.nt[2] * g.nt[1]));
  const int lo[3] = {tx * TILE, ty * TILE, tz * TILE};
  for (int i = threadIdx.x; i < NB; i += kTileThreads) acc[i] = 0ull;
  __syncthreads();
  const int32_t b = w.offs[t], e = w.offs[t + 1];
  for (int32_t s = b + threadIdx.x; s < e; s += kTileThreads) {
    const int64_t p = w.list[s];
    int i0[3];
    T wt[3][ORDER];
    window_of<T, ORDER>(P, p, g.nm, i0, wt);
    int lx = i0[0] - lo[0], ly = i0[1] - lo[1], lz = i0[2] - lo[2];
    if ((unsigned)lx >= (unsigned)TILE || (unsigned)ly >= (unsigned)TILE ||
        (unsigned)lz >= (unsigned)TILE) {
      // the particle was placed in another tile when binned: keep the LDS
      // access in bounds and report it (never expected: same arithmetic)
      atomicAdd(w.misplaced, 1);
      lx = lx < 0 ? 0 : (lx >= TILE ? TILE - 1 : lx);
      ly = ly < 0 ? 0 : (ly >= TILE ? TILE - 1 : ly);
      lz = lz < 0 ? 0 : (lz >= TILE ? TILE - 1 : lz);
    }
#pragma unroll
    for (int a = 0; a < ORDER; ++a) {
#pragma unroll
      for (int c = 0; c < ORDER; ++c) {
        const T wac = wt[0][a] * wt[1][c];
#pragma unroll
        for (int d = 0; d < ORDER; ++d) {
          const double v = (double)(wac * wt[2][d]) * mass_fx;
          const long long q = __double2ll_rn(v);
          atomicAdd(&acc[((lx + a) * TB + (ly + c)) * TB + (lz + d)], (unsigned long long)q);
        }
      }
    }
  }
  __syncthreads();
  T* halo = static_cast<T*>(w.halo) + t * g.halo;
  for (int i = threadIdx.x; i < NB; i += kTileThreads) {
    const int z = i % TB, y = (i / TB) % TB, x = i / (TB * TB);
    const T v = (T)((double)(long long)acc[i] * inv_fx);
    if (x < TILE && y < TILE && z < TILE)
      mesh[((int64_t)(lo[0] + x) * g.nm[1] + (lo[1] + y)) * g.nm[2] + (lo[2] + z)] = v;
    else
      halo[halo_index(x, y, z, H, TB)] = v;
  }
}

// Adds the overflow of the up-to-7 lower neighbours into each tile's cells.
template <typename T, int ORDER>
__global__ __launch_bounds__(kTileThreads) void k_paint_halo(TileGeom g, Workspace w, T* mesh) {
  constexpr int H = ORDER - 1;
  constexpr int TB = TILE + H;
  const int64_t t = xcd_remap(blockIdx.x, gridDim.x);
  const int tz = (int)(t % g.nt[2]);
  const int ty = (int)((t / g.nt[2]) % g.nt[1]);
  const int tx = (int)(t / ((int64_t)g.nt[2] * g.nt[1]));
  const int lo[3] = {tx * TILE, ty * TILE, tz * TILE};
  const T* halo = static_cast<const T*>(w.halo);
  // receiving cells: some local coordinate < H
  for (int i = threadIdx.x; i < TILE * TILE * TILE; i += kTileThreads) {
    const int z = i % TILE, y = (i / TILE) % TILE, x = i / (TILE * TILE);
    if (x >= H && y >= H && z >= H) continue;
    T sum = (T)0;
    bool any = false;
    // neighbour offsets (ox, oy, oz) in {0, -1}^3 \ {0}, fixed order
#pragma unroll
    for (int o = 1; o < 8; ++o) {
      const int ox = (o >> 2) & 1, oy = (o >> 1) & 1, oz = o & 1;
      if ((ox && x >= H) || (oy && y >= H) || (oz && z >= H)) continue;
      const int nx = (tx - ox + g.nt[0]) % g.nt[0];
      const int ny = (ty - oy + g.nt[1]) % g.nt[1];
      const int nz = (tz - oz + g.nt[2]) % g.nt[2];
      const int64_t nt = ((int64_t)nx * g.nt[1] + ny) * g.nt[2] + nz;
      sum += halo[nt * g.halo + halo_index(x + ox * TILE, y + oy * TILE, z + oz * TILE, H, TB)];
      any = true;
    }
    if (any) {
      T* cell = mesh + ((int64_t)(lo[0] + x) * g.nm[1] + (lo[1] + y)) * g.nm[2] + (lo[2] + z);
      *cell = *cell + sum;
    }
  }
}

// ---------------------------------------------------------------- readout
template <typename T>
struct TileRead {
  const T* mesh[4];
  T* out[4];
  int64_t stride[4];
  T scale[4];
};

template <typename T, int ORDER, int NC>
__global__ __launch_bounds__(kTileThreads) void k_readout_tiles(PartDesc<T> P, TileGeom g,
                                                                 Workspace w, TileRead<T> R) {
  constexpr int H = ORDER - 1;
  constexpr int TB = TILE + H;
  constexpr int NB = TB * TB * TB;
  __shared__ T box[NC][NB];
  const int64_t t = xcd_remap(blockIdx.x, gridDim.x);
  const int tz = (int)(t % g.nt[2]);
  const int ty = (int)((t / g.nt[2]) % g.nt[1]);
  const int tx = (int)(t / ((int64_t)g.nt[2] * g.nt[1]));
  const int lo[3] = {tx * TILE, ty * TILE, tz * TILE};
  const int32_t b = w.offs[t], e = w.offs[t + 1];
  if (b == e) return;  // uniform across the workgroup
  // Stage the box: each (component, x, y) row is TILE aligned cells loaded as
  // TILE/4 four-wide vectors plus H wrapped halo cells.
  constexpr int NQ = TILE / 4 + 1;
  for (int i = threadIdx.x; i < NC * TB * TB * NQ; i += kTileThreads) {
    const int q = i % NQ;
    const int r = i / NQ;
    const int y = r % TB, x = (r / TB) % TB, c = r / (TB * TB);
    int gx = lo[0] + x, gy = lo[1] + y;
    if (gx >= g.nm[0]) gx -= g.nm[0];
    if (gy >= g.nm[1]) gy -= g.nm[1];
    const T* row = R.mesh[c] + ((int64_t)gx * g.nm[1] + gy) * g.nm[2];
    T* dst = &box[c][(x * TB + y) * TB];
    if (q < TILE / 4) {
      typedef T V4 __attribute__((ext_vector_type(4)));
      const V4 v = *reinterpret_cast<const V4*>(row + lo[2] + 4 * q);
      dst[4 * q + 0] = v.x;
      dst[4 * q + 1] = v.y;
      dst[4 * q + 2] = v.z;
      dst[4 * q + 3] = v.w;
    } else {
#pragma unroll
      for (int h = 0; h < H; ++h) {
        int gz = lo[2] + TILE + h;
        if (gz >= g.nm[2]) gz -= g.nm[2];
        dst[TILE + h] = row[gz];
      }
    }
  }
  __syncthreads();
  for (int32_t s = b + threadIdx.x; s < e; s += kTileThreads) {
    const int64_t p = w.list[s];
    int i0[3];
    T wt[3][ORDER];
    window_of<T, ORDER>(P, p, g.nm, i0, wt);
    int lx = i0[0] - lo[0], ly = i0[1] - lo[1], lz = i0[2] - lo[2];
    if ((unsigned)lx >= (unsigned)TILE || (unsigned)ly >= (unsigned)TILE ||
        (unsigned)lz >= (unsigned)TILE) {
      atomicAdd(w.misplaced, 1);
      lx = lx < 0 ? 0 : (lx >= TILE ? TILE - 1 : lx);
      ly = ly < 0 ? 0 : (ly >= TILE ? TILE - 1 : ly);
      lz = lz < 0 ? 0 : (lz >= TILE ? TILE - 1 : lz);
    }
    T acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = (T)0;
#pragma unroll
    for (int a = 0; a < ORDER; ++a) {
#pragma unroll
      for (int c2 = 0; c2 < ORDER; ++c2) {
        const T wac = wt[0][a] * wt[1][c2];
        const int row = ((lx + a) * TB + (ly + c2)) * TB + lz;
#pragma unroll
        for (int d = 0; d < ORDER; ++d) {
          const T wv = wac * wt[2][d];
#pragma unroll
          for (int c = 0; c < NC; ++c) acc[c] += wv * box[c][row + d];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) R.out[c][p * R.stride[c]] = acc[c] * R.scale[c];
  }
}

// ---------------------------------------------------------------- host side
static bool tiles_ok(const fpm_mesh* m) {
  if (!m) return false;
  for (int d = 0; d < 3; ++d)
    if (m->nmesh[d] <= 0 || m->nmesh[d] % TILE) return false;
  return true;
}

static unsigned grid_parts(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > (1 << 20)) g = 1 << 20;
  return (unsigned)g;
}

template <typename T>
static int prep(const fpm_particles* part, const fpm_mesh* m, int order, void* ws, int64_t ws_bytes,
                PartDesc<T>& P, TileGeom& g, Workspace& w) {
  int rc = make_desc<T>(part, m, P);
  if (rc) return rc;
  if (!tiles_ok(m) || (order != 2 && order != 3) || !ws) return FPM_ERR_ARG;
  if (P.n >= (int64_t)1 << 31) return FPM_ERR_ARG;
  g = tile_geom(m, order);
  const int64_t need = layout_ws(g, P.n, sizeof(T), static_cast<char*>(ws), &w);
  if (ws_bytes < need) return FPM_ERR_WORKSPACE;
  return FPM_OK;
}

template <typename T>
static int bin_impl(const fpm_particles* part, const fpm_mesh* m, int order, void* ws,
                    int64_t ws_bytes, hipStream_t st) {
  PartDesc<T> P;
  TileGeom g;
  Workspace w;
  int rc = prep<T>(part, m, order, ws, ws_bytes, P, g, w);
  if (rc) return rc;
  hipLaunchKernelGGL(k_zero_i32, dim3(grid_parts(g.ntiles)), dim3(256), 0, st, w.count, g.ntiles);
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, w.misplaced, (int64_t)1);
  if (P.n > 0) {
    if (order == 2)
      hipLaunchKernelGGL((k_bin_count<T, 2>), dim3(grid_parts(P.n)), dim3(256), 0, st, P, g, w);
    else
      hipLaunchKernelGGL((k_bin_count<T, 3>), dim3(grid_parts(P.n)), dim3(256), 0, st, P, g, w);
  }
  hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, st, w, g.ntiles);
  if (P.n > 0)
    hipLaunchKernelGGL(k_bin_scatter, dim3(grid_parts(P.n)), dim3(256), 0, st, P.n, w);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T>
static int paint_tiles_impl(const fpm_particles* part, const fpm_mesh* m, int order, double mass,
                            void* out, void* ws, int64_t ws_bytes, hipStream_t st) {
  PartDesc<T> P;
  TileGeom g;
  Workspace w;
  int rc = prep<T>(part, m, order, ws, ws_bytes, P, g, w);
  if (rc) return rc;
  if (!out || !(std::fabs(mass) > 0) || !std::isfinite(mass)) return FPM_ERR_ARG;
  // 64-bit fixed point: the total |mass| of all particles stays below 2^62.
  const double total = std::fabs(mass) * (double)(P.n > 0 ? P.n : 1);
  int S = 62 - (int)std::ceil(std::log2(total + 1.0));
  if (S > 60) S = 60;
  const double mass_fx = mass * std::ldexp(1.0, S);
  const double inv_fx = std::ldexp(1.0, -S);
  T* mesh = static_cast<T*>(out);
  const dim3 grid((unsigned)g.ntiles), blk(kTileThreads);
  if (order == 2) {
    hipLaunchKernelGGL((k_paint_tiles<T, 2>), grid, blk, 0, st, P, g, w, mesh, mass_fx, inv_fx);
    hipLaunchKernelGGL((k_paint_halo<T, 2>), grid, blk, 0, st, g, w, mesh);
  } else {
    hipLaunchKernelGGL((k_paint_tiles<T, 3>), grid, blk, 0, st, P, g, w, mesh, mass_fx, inv_fx);
    hipLaunchKernelGGL((k_paint_halo<T, 3>), grid, blk, 0, st, g, w, mesh);
  }
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T, int ORDER>
static void launch_rt(const PartDesc<T>& P, const TileGeom& g, const Workspace& w,
                      const TileRead<T>& R, int nc, hipStream_t st) {
  const dim3 grid((unsigned)g.ntiles), blk(kTileThreads);
  switch (nc) {
    case 1: hipLaunchKernelGGL((k_readout_tiles<T, ORDER, 1>), grid, blk, 0, st, P, g, w, R); break;
    case 2: hipLaunchKernelGGL((k_readout_tiles<T, ORDER, 2>), grid, blk, 0, st, P, g, w, R); break;
    case 3: hipLaunchKernelGGL((k_readout_tiles<T, ORDER, 3>), grid, blk, 0, st, P, g, w, R); break;
    default: hipLaunchKernelGGL((k_readout_tiles<T, ORDER, 4>), grid, blk, 0, st, P, g, w, R); break;
  }
}

template <typename T>
static int readout_tiles_impl(const fpm_particles* part, const fpm_mesh* m, int order, int ncomp,
                              const void* const* meshes, void* const* outs, const int64_t* strides,
                              const double* scales, void* ws, int64_t ws_bytes, hipStream_t st) {
  PartDesc<T> P;
  TileGeom g;
  Workspace w;
  int rc = prep<T>(part, m, order, ws, ws_bytes, P, g, w);
  if (rc) return rc;
  if (ncomp < 1 || ncomp > 4 || !meshes || !outs || !strides) return FPM_ERR_ARG;
  TileRead<T> R{};
  for (int c = 0; c < ncomp; ++c) {
    if (!meshes[c] || !outs[c] || strides[c] < 1) return FPM_ERR_ARG;
    R.mesh[c] = static_cast<const T*>(meshes[c]);
    R.out[c] = static_cast<T*>(outs[c]);
    R.stride[c] = strides[c];
    R.scale[c] = scales ? (T)scales[c] : (T)1;
  }
  if (P.n == 0) return FPM_OK;
  if (order == 2) launch_rt<T, 2>(P, g, w, R, ncomp, st);
  else launch_rt<T, 3>(P, g, w, R, ncomp, st);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

}  // namespace fpm

using namespace fpm;

extern "C" {

int fpm_tiles_supported(const fpm_mesh* mesh) { return tiles_ok(mesh) ? 1 : 0; }

int64_t fpm_tile_misplaced_offset(const fpm_mesh* mesh, int order, int64_t n) {
  if (!tiles_ok(mesh) || (order != 2 && order != 3) || n < 0) return -1;
  int64_t off[8];
  ws_offsets(tile_geom(mesh, order), n, sizeof(float), off);
  return off[5];
}

int64_t fpm_tile_workspace_size(const fpm_mesh* mesh, int order, int64_t n) {
  if (!tiles_ok(mesh) || (order != 2 && order != 3) || n < 0) return -1;
  return layout_ws(tile_geom(mesh, order), n, sizeof(float), nullptr, nullptr);
}

int fpm_bin_particles_f32(const fpm_particles* part, const fpm_mesh* mesh, int order, void* ws,
                          int64_t ws_bytes, fpm_stream_t s) {
  return bin_impl<float>(part, mesh, order, ws, ws_bytes, (hipStream_t)s);
}
int fpm_paint_tiled_f32(const fpm_particles* part, const fpm_mesh* mesh, int order, double mass,
                        void* out, void* ws, int64_t ws_bytes, fpm_stream_t s) {
  return paint_tiles_impl<float>(part, mesh, order, mass, out, ws, ws_bytes, (hipStream_t)s);
}
int fpm_readout_tiled_f32(const fpm_particles* part, const fpm_mesh* mesh, int order, int ncomp,
                          const void* const* meshes, void* const* outs, const int64_t* strides,
                          const double* scales, void* ws, int64_t ws_bytes, fpm_stream_t s) {
  return readout_tiles_impl<float>(part, mesh, order, ncomp, meshes, outs, strides, scales, ws,
                                   ws_bytes, (hipStream_t)s);
}

}  // extern "C"
