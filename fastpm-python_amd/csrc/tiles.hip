// Tile-binned, deterministic paint and LDS-staged readout (float32 storage).
//
// The force mesh is cut into TILE^3-cell tiles.  A particle's window starts at
// i0 (the CIC floor cell, or the TSC nearest point minus one) and covers
// i0 .. i0 + order - 1 per axis.  It belongs to the tile holding i0, in one of
// eight sub-bins given by a 3-bit mask: bit d is set when its window crosses
// the tile's upper face along axis d (i0 mod TILE > TILE - order).
//
//   bin      : count particles per (tile, mask) with integer atomics merged
//              per wave (the value an atomic returns gives each particle its
//              slot in the sub-bin), exclusive scan, then an atomic-free
//              scatter of particle indices into the per-sub-bin lists.
//   paint    : one workgroup per tile gathers every deposit that lands in its
//              cells: its own particles, plus the particles of the seven lower
//              neighbour tiles whose mask says they cross into it.  Deposits
//              are summed in LDS as 64-bit fixed point (integer adds, so the
//              sum does not depend on the order particles arrive in) and each
//              cell is written exactly once: no clearing pass, no global
//              atomics, no halo pass, and bitwise reproducible results.
//   readout  : one workgroup per tile stages the (TILE + order - 1)^3 box of
//              every requested mesh in LDS (each cell read once plus the
//              1/TILE overhang) and interpolates for the tile's particles.
#include "batch.h"

#include <cmath>

namespace fpm {

template <typename T>
int make_desc(const fpm_particles* part, const fpm_mesh* m, PartDesc<T>& P);

constexpr int TILE = 16;
constexpr int kTileThreads = 512;
// The CIC stencil readout runs 1024 threads per tile: it fits in 63 VGPRs,
// so two 1024-thread workgroups share a CU (twice the waves in flight for
// the latency-bound staging; 0.634 -> 0.615 ms at 512^3).  The TSC readout
// needs 124 VGPRs and keeps 512.
template <int ORDER> constexpr int grad_threads() { return ORDER == 2 ? 1024 : kTileThreads; }
constexpr int kSub = 8;  // sub-bins per tile

struct TileGeom {
  int nm[3];
  int nt[3];       // tiles per axis
  int64_t ntiles;
  int64_t nbins;   // ntiles * kSub
};

static TileGeom tile_geom(const fpm_mesh* m) {
  TileGeom g{};
  g.ntiles = 1;
  for (int d = 0; d < 3; ++d) {
    g.nm[d] = m->nmesh[d];
    g.nt[d] = m->nmesh[d] / TILE;
    g.ntiles *= g.nt[d];
  }
  g.nbins = g.ntiles * kSub;
  return g;
}

// Binned particle record: the window's offsets within its first cell (CIC:
// the fraction f, TSC: f or f - 1 around the nearest point) and the window
// origin local to the particle's tile (4 bits per axis).
struct PRec {
  float t[3];
  uint32_t cells;  // lx | ly << 4 | lz << 8
};

struct Workspace {
  int32_t* rank;       // each particle's slot within its sub-bin (thread order)
  int32_t* list;       // particle index per slot (ncap slots)
  PRec* rec;           // particle record per slot
  int32_t* count;      // particles per sub-bin
  int32_t* offs;       // first slot of each sub-bin (nbins + 1)
  int32_t* misplaced;  // particles whose slot the exact scatter found outside [0, n)
  int32_t* ovf;        // nonzero: the records are placed exactly (the scatter runs):
                       // 1 with the count's ranks, 2 re-ranked by the scatter
  int64_t* hdr;        // magic, particles, sub-bins of the capacity layout in nxt,
                       // and its wide-margin level (wide when above 0)
  int32_t* btot;       // per-workgroup totals of the count scan
  int32_t* btotc;      // per-workgroup totals of the capacity scan
  int32_t* nxt;        // first slot of each sub-bin in the next call's capacity layout
  int64_t ncap;        // slots of list and rec
  int64_t bytes;
  int32_t allow_cap;   // the caller lets this bin call use the capacity layout
};

// Capacity layout.  A binning call places each record straight at its
// sub-bin's start plus its rank when the workspace holds the capacity
// layout its previous call left (sub-bin b gets cap(count) slots of that
// call, the header says for how many particles and sub-bins): one pass over
// the particles instead of a count pass and a scatter pass.  A record that
// does not fit its sub-bin's capacity (or a workspace without that layout)
// sets ovf, and the scan and the scatter then place every record exactly as
// before.  The scatter needs each particle's rank: a count pass stores the
// ranks when it has no layout, or one with the wide margin (ovf = 1 on an
// overflow; those layouts follow an exact placement closely, while the
// counts change fast); after an overflow of a pass that did not store them
// (ovf = 2) the scan clears the counts and the scatter ranks the particles
// again by counting them back up.  Either way sub-bin b's records are slots offs[b] ..
// offs[b] + count[b] - 1, which is all the paint and the readouts use; the
// results do not depend on the layout.  bin_cap(c) sums to at most ncap.
constexpr int64_t kCapMagic = 0x46504d4341503031LL;
// After a call that had to place its records exactly (a sub-bin outgrew its
// capacity: the counts are changing fast, as in the first steps), the next
// kWideCalls layouts get twice the margin.
__host__ __device__ __forceinline__ int32_t bin_cap(int32_t c, bool wide) {
  return wide ? c + (c >> 1) + 64 : c + (c >> 2) + 32;
}
constexpr int64_t kWideCalls = 4;
// The wide-margin level of the layout a bin call leaves (a scan reads the
// level of the layout the call used; exact: the call placed exactly).
__device__ __forceinline__ int64_t next_wide_level(const Workspace& w, bool exact) {
  const int64_t cur = w.hdr[3];
  return exact ? kWideCalls : (cur > 1 ? cur - 1 : 0);
}

static int64_t capacity_slots(int64_t n, int64_t nbins) {
  const int64_t c = n + n / 2 + 64 * nbins;
  return c < ((int64_t)1 << 31) - 1 ? c : n;
}

static int64_t align256(int64_t x) { return (x + 255) / 256 * 256; }

constexpr int kScanThreads = 256;
constexpr int kScanPer = 8;
constexpr int kScanBlock = kScanThreads * kScanPer;  // sub-bins per scan workgroup

static int64_t scan_blocks(const TileGeom& g) { return (g.nbins + kScanBlock - 1) / kScanBlock; }

// Byte offsets of the workspace sections: rank, list, rec, count, offs,
// control words (misplaced, ovf, header), btot, btotc, nxt; off[9] is the
// total size.
static void ws_offsets(const TileGeom& g, int64_t n, int64_t off[10]) {
  const int64_t ncap = capacity_slots(n, g.nbins);
  const int64_t sizes[9] = {n * 4, ncap * 4, ncap * (int64_t)sizeof(PRec), g.nbins * 4,
                            (g.nbins + 1) * 4, 64, scan_blocks(g) * 4, scan_blocks(g) * 4,
                            (g.nbins + 1) * 4};
  int64_t o = 0;
  for (int i = 0; i < 9; ++i) {
    off[i] = o;
    o += align256(sizes[i]);
  }
  off[9] = o;
}

static int64_t layout_ws(const TileGeom& g, int64_t n, char* base, Workspace* w) {
  int64_t off[10];
  ws_offsets(g, n, off);
  if (w) {
    w->rank = reinterpret_cast<int32_t*>(base + off[0]);
    w->list = reinterpret_cast<int32_t*>(base + off[1]);
    w->rec = reinterpret_cast<PRec*>(base + off[2]);
    w->count = reinterpret_cast<int32_t*>(base + off[3]);
    w->offs = reinterpret_cast<int32_t*>(base + off[4]);
    w->misplaced = reinterpret_cast<int32_t*>(base + off[5]);
    w->ovf = reinterpret_cast<int32_t*>(base + off[5] + 4);
    w->hdr = reinterpret_cast<int64_t*>(base + off[5] + 16);
    w->btot = reinterpret_cast<int32_t*>(base + off[6]);
    w->btotc = reinterpret_cast<int32_t*>(base + off[7]);
    w->nxt = reinterpret_cast<int32_t*>(base + off[8]);
    w->ncap = capacity_slots(n, g.nbins);
    w->bytes = off[9];
    w->allow_cap = 0;
  }
  return off[9];
}

// Window origin (wrapped) and the offset t it is weighted by (CIC: the
// fraction f; TSC: f, or f - 1 when the nearest point is the next cell).
template <typename T, int ORDER>
__device__ __forceinline__ void window_cell(const int64_t cell[3], const T f[3], const int nm[3],
                                            int i0[3], T t[3]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int64_t o;
    if constexpr (ORDER == 2) {
      o = cell[d];
      t[d] = f[d];
    } else {
      const bool up = f[d] >= (T)0.5;
      t[d] = up ? f[d] - (T)1 : f[d];
      o = cell[d] + (up ? 1 : 0) - 1;
    }
    i0[d] = (int)wrapi(o, nm[d]);
  }
}

template <typename T, int ORDER>
__device__ __forceinline__ void window_t(const PartDesc<T>& P, int64_t p, const int nm[3],
                                         int i0[3], T t[3]) {
  int64_t cell[3];
  T f[3];
  particle_cell(P, p, cell, f);
  window_cell<T, ORDER>(cell, f, nm, i0, t);
}

// Window weights from t, identical to the plain paint/readout.
template <typename T, int ORDER>
__device__ __forceinline__ void weights_t(const T t[3], T w[3][ORDER]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if constexpr (ORDER == 2) {
      w[d][0] = (T)1 - t[d];
      w[d][1] = t[d];
    } else {
      const T dd = t[d];
      const T a = (T)0.5 - dd, b = (T)0.5 + dd;
      w[d][0] = (T)0.5 * a * a;
      w[d][1] = (T)0.75 - dd * dd;
      w[d][2] = (T)0.5 * b * b;
    }
  }
}

// A record's window: origin local to its tile and weights.
template <typename T, int ORDER>
__device__ __forceinline__ void window_rec(const PRec& r, int l[3], T w[3][ORDER]) {
  l[0] = (int)(r.cells & 15u);
  l[1] = (int)((r.cells >> 4) & 15u);
  l[2] = (int)((r.cells >> 8) & 15u);
  const T t[3] = {(T)r.t[0], (T)r.t[1], (T)r.t[2]};
  weights_t<T, ORDER>(t, w);
}

template <int ORDER>
__device__ __forceinline__ int32_t bin_of(const int i0[3], const TileGeom& g) {
  // i0 is wrapped (non-negative) and the bin count is below 2^31 (checked
  // by the host), so 32-bit unsigned arithmetic is exact
  const uint32_t u[3] = {(uint32_t)i0[0], (uint32_t)i0[1], (uint32_t)i0[2]};
  const uint32_t t = ((u[0] / TILE) * (uint32_t)g.nt[1] + u[1] / TILE) * (uint32_t)g.nt[2] +
                     u[2] / TILE;
  uint32_t mask = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (u[d] % TILE > (uint32_t)(TILE - ORDER)) mask |= 1u << d;
  return (int32_t)(t * kSub + mask);
}

// Tile t's first cell; the tile count is below 2^28 (the host keeps the
// sub-bin count below 2^31), so 32-bit division is exact.
__device__ __forceinline__ void tile_coords(int64_t t, const TileGeom& g, int lo[3]) {
  const uint32_t u = (uint32_t)t, n2 = (uint32_t)g.nt[2], n1 = (uint32_t)g.nt[1];
  const uint32_t r = u / n2;
  lo[2] = (int)(u - r * n2) * TILE;
  lo[1] = (int)(r % n1) * TILE;
  lo[0] = (int)(r / n1) * TILE;
}

// ---------------------------------------------------------------- binning
__global__ void k_zero_i32(int32_t* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// Thread-to-particle order of the binning passes.  On a lattice whose sides
// are multiples of (kBlockX, 8, 16), a workgroup step's kBinSpan particles
// are a kBlockX x 8 x 16 lattice block (4 x 8 x 16 with two particles per
// thread: a quarter of a 16^3 tile's cells at B = 2 before the particles
// move, so they fall in few sub-bins) and each wave a 4 x 16 (y, z) layer of
// it: rows of 16 particles keep the position loads on whole cache lines (an
// 8^3 cube of 8-particle rows measured 24 us slower in the scatter).  With only the y and z sides multiples of 8, wave w takes the
// w-th 8 x 8 (y, z) patch of the x planes; otherwise threads take particles
// in index order.  key and rank are stored by thread index; the scatter
// applies the same order.  The block or patch gives a blocked particle's
// lattice coordinates, so its placement needs no division of the particle
// index.
enum { kBinIndex = 0, kBinPatch = 1, kBinBlock = 2 };
#ifndef FPM_BIN_ILP
#define FPM_BIN_ILP 2
#endif
constexpr int kBinILP = FPM_BIN_ILP;
constexpr int kBinThreads = 256;
constexpr int kBinSpan = kBinThreads * kBinILP;  // particles per workgroup step
constexpr int kBlockX = kBinSpan / 128;          // block: kBlockX x 8 x 16 particles
static_assert(kBinSpan % 128 == 0 && (kBinSpan & (kBinSpan - 1)) == 0, "block order spans");
struct BinOrder {
  int blocked;
  uint32_t nz, ny;  // patches (8 x 8) or blocks (kBlockX x 8 x 16) along z and y
  uint32_t lat1, lat2;
};

// With kd (kd.S != nullptr, uniform) the particle's kicks and drift are
// applied first and its new displacement is used from registers.
template <typename T, int ORDER>
__device__ __forceinline__ int64_t bin_window(const BinOrder& o, const PartDesc<T>& P,
                                              const int nm[3], int64_t i, int i0[3], T t[3],
                                              const KickDrift<T>* kd = nullptr) {
  int64_t cell[3];
  T f[3];
  int64_t p = i;
  if (o.blocked != kBinIndex) {
    int32_t li[3];
    if (o.blocked == kBinBlock) {
      const uint32_t c = (uint32_t)(i / kBinSpan), j = (uint32_t)(i % kBinSpan);
      const uint32_t u = c / o.nz, cz = c - u * o.nz;
      const uint32_t cx = u / o.ny, cy = u - cx * o.ny;
      li[0] = (int32_t)(cx * kBlockX + (j >> 7));
      li[1] = (int32_t)(cy * 8 + ((j >> 4) & 7));
      li[2] = (int32_t)(cz * 16 + (j & 15));
    } else {
      const uint32_t w = (uint32_t)(i >> 6);
      const uint32_t lane = (uint32_t)(i & 63);
      const uint32_t u = w / o.nz, pz = w - u * o.nz;
      const uint32_t ix = u / o.ny, py = u - ix * o.ny;
      li[0] = (int32_t)ix;
      li[1] = (int32_t)(py * 8 + (lane >> 3));
      li[2] = (int32_t)(pz * 8 + (lane & 7));
    }
    p = ((int64_t)li[0] * o.lat1 + li[1]) * o.lat2 + li[2];
    if (kd && kd->S) {
      T sv[3];
      kick_drift_one(*kd, p, sv);
      lattice_cell_s(P, li, sv, cell, f);
    } else {
      lattice_cell(P, p, li, cell, f);
    }
  } else if (kd && kd->S) {
    T sv[3];
    kick_drift_one(*kd, p, sv);
    particle_cell_s(P, p, sv, cell, f);
  } else {
    particle_cell(P, p, cell, f);
  }
  window_cell<T, ORDER>(cell, f, nm, i0, t);
  return p;
}

// Lanes of a wave that fall in the same sub-bin are merged before touching
// the counters, so each distinct sub-bin in a wave costs one atomic.  The
// count before the add plus the lane's rank among its group is the
// particle's slot within its sub-bin; the results are consumed after the
// merge loops, so all of a wave's atomics are in flight together.  Each
// thread takes kBinILP particles (one per 256-particle stripe of its
// block's span), whose position loads and atomics overlap: the pass is
// bound by their latency, not by bandwidth.

// Merges a wave's sub-bin keys (t[u] < 0: no particle) and adds each group
// to its counter; rank[i] = rk[u] = the counter before the add + the lane's
// rank in its group.
__device__ __forceinline__ void count_ranks(const int32_t t[kBinILP], int64_t base, int64_t n,
                                            const Workspace& w, int32_t rk[kBinILP],
                                            bool store_rank) {
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int32_t old[kBinILP];
  int myleader[kBinILP], myrank[kBinILP];
#pragma unroll
  for (int u = 0; u < kBinILP; ++u) {
    const bool valid = t[u] >= 0;
    unsigned long long active = __ballot(valid);
    myleader[u] = 0;
    myrank[u] = 0;
    old[u] = 0;
    while (active) {
      const int leader = __ffsll(active) - 1;
      const int32_t tl = __shfl(t[u], leader);
      const unsigned long long same = __ballot(valid && t[u] == tl);
      if (lane == leader) old[u] = atomicAdd(&w.count[tl], (int32_t)__popcll(same));
      if (valid && t[u] == tl) {
        myleader[u] = leader;
        myrank[u] = __popcll(same & below);
      }
      active &= ~same;
    }
  }
#pragma unroll
  for (int u = 0; u < kBinILP; ++u) {
    const int64_t i = base + u * kBinThreads + threadIdx.x;
    const int32_t r0 = __shfl(old[u], myleader[u]);
    rk[u] = r0 + myrank[u];
    if (store_rank && i < n) w.rank[i] = rk[u];
  }
}

// Workgroup form of count_ranks for block order: each wave's groups are
// merged again across the workgroup in an LDS table keyed by sub-bin (open
// addressing; one slot per particle, so an insertion always finds its key or
// a free slot), and each distinct sub-bin of the workgroup costs one global
// atomic.  A group's slot is the counter before the workgroup's add, plus
// the group's offset within the workgroup (an LDS atomic), plus the lane's
// rank in its group.  All threads of the workgroup must call it.
constexpr int kBinTable = kBinThreads * kBinILP;
static_assert((kBinTable & (kBinTable - 1)) == 0, "the sub-bin table size is a power of two");

__device__ __forceinline__ void count_ranks_wg(const int32_t t[kBinILP], int64_t base, int64_t n,
                                               const Workspace& w, int32_t* tkey, int32_t* tval,
                                               int32_t rk[kBinILP], bool store_rank) {
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  __syncthreads();  // the previous span's readers of the table are done
  for (int k = threadIdx.x; k < kBinTable; k += kBinThreads) {
    tkey[k] = -1;
    tval[k] = 0;
  }
  __syncthreads();
  int32_t off[kBinILP], slot[kBinILP];
  int myleader[kBinILP], myrank[kBinILP];
#pragma unroll
  for (int u = 0; u < kBinILP; ++u) {
    const bool valid = t[u] >= 0;
    unsigned long long active = __ballot(valid);
    myleader[u] = 0;
    myrank[u] = 0;
    off[u] = 0;
    slot[u] = 0;
    while (active) {
      const int leader = __ffsll(active) - 1;
      const int32_t tl = __shfl(t[u], leader);
      const unsigned long long same = __ballot(valid && t[u] == tl);
      if (lane == leader) {
        uint32_t h = ((uint32_t)tl * 2654435761u) >> 16;
        for (;;) {
          h &= kBinTable - 1;
          const int32_t prev = atomicCAS(&tkey[h], -1, tl);
          if (prev == -1 || prev == tl) break;
          ++h;
        }
        slot[u] = (int32_t)h;
        off[u] = atomicAdd(&tval[h], (int32_t)__popcll(same));
      }
      if (valid && t[u] == tl) {
        myleader[u] = leader;
        myrank[u] = __popcll(same & below);
      }
      active &= ~same;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kBinTable; k += kBinThreads) {
    const int32_t key = tkey[k];
    if (key >= 0) tval[k] = atomicAdd(&w.count[key], tval[k]);
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kBinILP; ++u) {
    const int64_t i = base + u * kBinThreads + threadIdx.x;
    const int32_t h = __shfl(slot[u], myleader[u]);
    const int32_t o = __shfl(off[u], myleader[u]);
    rk[u] = tval[h] + o + myrank[u];
    if (store_rank && i < n && t[u] >= 0) w.rank[i] = rk[u];
  }
}

// A particle's binned record from its window: offsets t, origin i0.
template <typename T>
__device__ __forceinline__ PRec make_rec(const T t[3], const int i0[3]) {
  PRec r;
  r.t[0] = (float)t[0];
  r.t[1] = (float)t[1];
  r.t[2] = (float)t[2];
  r.cells = (uint32_t)(i0[0] % TILE) | (uint32_t)(i0[1] % TILE) << 4 |
            (uint32_t)(i0[2] % TILE) << 8;
  return r;
}

// Whether this binning may use the workspace's capacity layout: the caller
// allows it (the workspace is unchanged since the bin call that left it) and
// the layout is for as many particles and sub-bins.
__device__ __forceinline__ bool capacity_ready(const Workspace& w, int64_t n, int64_t nbins) {
  return w.allow_cap && w.ncap > n && w.hdr[0] == kCapMagic && w.hdr[1] == n &&
         w.hdr[2] == nbins;
}

// Kernel bodies take their block index and count explicitly, so one body
// serves a box's own launch (blockIdx.x of gridDim.x) and a batched launch
// over many boxes (the box's local block of its share of the items).
//
// The count pass: each particle's sub-bin is counted and its rank within
// the sub-bin found; with a capacity layout ready, the record and index go
// straight to the sub-bin's capacity start plus the rank (the layout's
// starts are copied to offs for the paint and readouts), and a record that
// does not fit sets ovf = 2.  Without a layout it sets ovf = 1 at once,
// stores the ranks, and the scatter places the records.
template <typename T, int ORDER>
__device__ __forceinline__ void bin_count_body(const PartDesc<T>& P, const TileGeom& g,
                                               const Workspace& w, const BinOrder& o, int64_t blk,
                                               int64_t nblk, const KickDrift<T>& kd) {
  __shared__ int32_t tkey[kBinTable], tval[kBinTable];
  const bool cap = capacity_ready(w, P.n, g.nbins);  // uniform
  const bool keep = !cap || w.hdr[3] > 0;             // store the ranks (uniform)
  if (!cap) {
    if (blk == 0 && threadIdx.x == 0) *w.ovf = 1;
  } else {
    for (int64_t i = blk * kBinThreads + threadIdx.x; i <= g.nbins; i += nblk * kBinThreads)
      w.offs[i] = w.nxt[i];
  }
  const int64_t span = (int64_t)kBinThreads * kBinILP;
  for (int64_t base = blk * span; base < P.n; base += nblk * span) {
    int32_t t[kBinILP], rk[kBinILP], pp[kBinILP];
    PRec r[kBinILP];
#pragma unroll
    for (int u = 0; u < kBinILP; ++u) {
      const int64_t i = base + u * kBinThreads + threadIdx.x;
      t[u] = -1;
      pp[u] = 0;
      if (i < P.n) {
        int i0[3];
        T tt[3];
        pp[u] = (int32_t)bin_window<T, ORDER>(o, P, g.nm, i, i0, tt, &kd);
        t[u] = bin_of<ORDER>(i0, g);
        r[u] = make_rec<T>(tt, i0);
      }
    }
    // the ranks are kept only for the exact placement's scatter
    if (o.blocked == kBinBlock)
      count_ranks_wg(t, base, P.n, w, tkey, tval, rk, keep);
    else
      count_ranks(t, base, P.n, w, rk, keep);
    if (cap) {
#pragma unroll
      for (int u = 0; u < kBinILP; ++u) {
        if (t[u] >= 0) {
          const int32_t sl = w.nxt[t[u]] + rk[u];
          if (sl >= 0 && sl < w.nxt[t[u] + 1] && sl < w.ncap) {
            w.list[sl] = pp[u];
            w.rec[sl] = r[u];
          } else {
            *w.ovf = keep ? 1 : 2;  // without stored ranks the scatter ranks again
          }
        }
      }
    }
  }
}

template <typename T, int ORDER>
__global__ __launch_bounds__(kBinThreads) void k_bin_count(PartDesc<T> P, TileGeom g, Workspace w,
                                                           BinOrder o, KickDrift<T> kd) {
  bin_count_body<T, ORDER>(P, g, w, o, blockIdx.x, gridDim.x, kd);
}

// Exclusive scans of the sub-bin counts (into offs, when the records are to
// be placed exactly) and of their capacities (into nxt, the next call's
// capacity layout) in two passes over kScanBlock-bin chunks.  Pass one scans
// each chunk locally and records the chunk totals; pass two adds the sums of
// the earlier chunk totals, which every workgroup reduces for itself (a few
// hundred values), and the last workgroup writes the layout's header.
__device__ __forceinline__ int32_t wave_incl_scan(int32_t v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t u = __shfl_up(v, o);
    if (lane >= o) v += u;
  }
  return v;
}

__device__ __forceinline__ void scan_local_body(const Workspace& w, int64_t nbins, int64_t blk) {
  __shared__ int32_t wsum[kScanThreads / 64], wsumc[kScanThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const bool exact = *w.ovf != 0;
  const bool wide = next_wide_level(w, exact) > 0;
  const int64_t i0 = blk * kScanBlock + (int64_t)tid * kScanPer;
  int32_t v[kScanPer];
  int32_t s = 0, sc = 0;
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    v[i] = i0 + i < nbins ? w.count[i0 + i] : 0;
    s += v[i];
    sc += i0 + i < nbins ? bin_cap(v[i], wide) : 0;
  }
  const int32_t incl = wave_incl_scan(s, lane), inclc = wave_incl_scan(sc, lane);
  if (lane == 63) {
    wsum[wid] = incl;
    wsumc[wid] = inclc;
  }
  __syncthreads();
  int32_t run = incl - s, runc = inclc - sc;
  for (int k = 0; k < wid; ++k) {
    run += wsum[k];
    runc += wsumc[k];
  }
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    if (i0 + i < nbins) {
      if (exact) w.offs[i0 + i] = run;
      w.nxt[i0 + i] = runc;
    }
    run += v[i];
    runc += bin_cap(v[i], wide);
  }
  if (tid == kScanThreads - 1) {
    w.btot[blk] = run;
    w.btotc[blk] = runc;
  }
}

__global__ __launch_bounds__(kScanThreads) void k_bin_scan_local(Workspace w, int64_t nbins) {
  scan_local_body(w, nbins, blockIdx.x);
}

__device__ __forceinline__ void scan_add_body(const Workspace& w, int64_t nbins, int64_t blk,
                                              int64_t nblk) {
  __shared__ int32_t wsum[kScanThreads / 64], wsumc[kScanThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const bool exact = *w.ovf != 0;
  int32_t s = 0, sc = 0;
  for (int b = tid; b < (int)blk; b += kScanThreads) {
    s += w.btot[b];
    sc += w.btotc[b];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    s += __shfl_xor(s, o);
    sc += __shfl_xor(sc, o);
  }
  if (lane == 0) {
    wsum[wid] = s;
    wsumc[wid] = sc;
  }
  __syncthreads();
  int32_t base = 0, basec = 0;
#pragma unroll
  for (int k = 0; k < kScanThreads / 64; ++k) {
    base += wsum[k];
    basec += wsumc[k];
  }
  // after an overflow the scatter counts the particles back up to rank them
  const bool rerank = *w.ovf == 2;
  const int64_t c0 = blk * kScanBlock;
  for (int i = tid; i < kScanBlock; i += kScanThreads) {
    const int64_t b = c0 + i;
    if (b < nbins) {
      if (exact) w.offs[b] += base;
      if (rerank) w.count[b] = 0;
      w.nxt[b] += basec;
    }
  }
  if (blk == nblk - 1 && tid == 0) {
    const int32_t n = base + w.btot[blk];
    if (exact) w.offs[nbins] = n;
    w.nxt[nbins] = basec + w.btotc[blk];
    w.hdr[0] = w.ncap > n ? kCapMagic : 0;
    w.hdr[1] = n;
    w.hdr[2] = nbins;
    w.hdr[3] = next_wide_level(w, exact);
  }
}

__global__ __launch_bounds__(kScanThreads) void k_bin_scan_add(Workspace w, int64_t nbins) {
  scan_add_body(w, nbins, blockIdx.x, gridDim.x);
}

// Exact placement (when the count pass set ovf): each particle goes to the
// start of its sub-bin (recomputed from its position with the count's
// arithmetic) plus its rank (stored by the count pass, or found again by
// counting up after an overflow), with its record (window offsets and
// tile-local origin), so paint and readout read their tile's particles
// contiguously and need no positions.
template <typename T, int ORDER>
__device__ __forceinline__ void bin_scatter_body(const PartDesc<T>& P, const TileGeom& g,
                                                 const Workspace& w, const BinOrder& o,
                                                 int64_t blk, int64_t nblk) {
  __shared__ int32_t tkey[kBinTable], tval[kBinTable];
  const int32_t mode = *w.ovf;  // uniform
  if (mode == 0) return;        // the count pass placed every record
  // mode 1: the count stored each particle's rank; mode 2: the scan cleared
  // the counts and the particles are ranked again as they are counted back
  // up (the order within a sub-bin does not change any result)
  const int64_t span = (int64_t)kBinThreads * kBinILP;
  for (int64_t base = blk * span; base < P.n; base += nblk * span) {
    int32_t slot[kBinILP], key[kBinILP], rk[kBinILP];
    int32_t pp[kBinILP];
    PRec r[kBinILP];
#pragma unroll
    for (int u = 0; u < kBinILP; ++u) {
      const int64_t i = base + u * kBinThreads + threadIdx.x;
      key[u] = -1;
      rk[u] = 0;
      if (i < P.n) {
        if (mode == 1) rk[u] = w.rank[i];
        int i0[3];
        T t[3];
        const int64_t p = bin_window<T, ORDER>(o, P, g.nm, i, i0, t);
        key[u] = bin_of<ORDER>(i0, g);
        pp[u] = (int32_t)p;
        r[u] = make_rec<T>(t, i0);
      }
    }
    if (mode != 1) {
      if (o.blocked == kBinBlock)
        count_ranks_wg(key, base, P.n, w, tkey, tval, rk, false);
      else
        count_ranks(key, base, P.n, w, rk, false);
    }
#pragma unroll
    for (int u = 0; u < kBinILP; ++u) {
      slot[u] = -1;
      if (key[u] >= 0) {
        const int32_t sl = w.offs[key[u]] + rk[u];
        // a slot outside [0, n) (count and scatter disagreeing) is counted,
        // not stored
        if (sl >= 0 && sl < P.n) {
          slot[u] = sl;
        } else {
          atomicAdd(w.misplaced, 1);
        }
      }
    }
    // the count placed each particle with the same arithmetic (no FMA
    // contraction in particle_cell), so its slot is in its sub-bin; the guard
    // above keeps the stores in bounds and counts any slot it drops
#pragma unroll
    for (int u = 0; u < kBinILP; ++u) {
      if (slot[u] >= 0) {
        w.list[slot[u]] = pp[u];
        w.rec[slot[u]] = r[u];
      }
    }
  }
}

template <typename T, int ORDER>
__global__ __launch_bounds__(kBinThreads) void k_bin_scatter(PartDesc<T> P, TileGeom g,
                                                             Workspace w, BinOrder o) {
  bin_scatter_body<T, ORDER>(P, g, w, o, blockIdx.x, gridDim.x);
}

// The records of NR / kSub consecutive tiles from tile t: NR sub-bin
// ranges, range k at slots offs[t kSub + k] .. + count[t kSub + k] - 1 (one
// run when the binning placed the records exactly, separated by spare
// capacity otherwise).  The starts and the prefix of the lengths are read
// through the scalar cache into registers every wave holds (the workspace
// is not written while they are read), and record f of the tiles
// (0 <= f < total) is found by comparing f with the prefix.
template <int NR>
struct TileSlots {
  static_assert(NR <= 64, "one range per lane");
  int32_t base0;       // start of range 0
  int32_t pre[NR + 1]; // records before range k
  int32_t gap;         // lane k < NR: slots of spare capacity before range k
  bool contiguous;     // no spare capacity between the ranges (exact placement)
  __device__ __forceinline__ void load(const Workspace& w, int64_t t) {
    typedef const __attribute__((address_space(4))) int32_t* CP;
    const CP offs = (CP)(w.offs + t * kSub);
    const CP cnt = (CP)(w.count + t * kSub);
    pre[0] = 0;
#pragma unroll
    for (int k = 0; k < NR; ++k) pre[k + 1] = pre[k] + cnt[k];
    base0 = offs[0];
    const int lane = threadIdx.x & 63;
    gap = 0;
    contiguous = true;
#pragma unroll
    for (int k = 1; k < NR; ++k) {
      if (lane == k) gap = offs[k] - base0 - pre[k];
      contiguous = contiguous && offs[k] == base0 + pre[k];
    }
  }
  __device__ __forceinline__ int32_t total() const { return pre[NR]; }
  // record f's range r (the count of ranges starting at or before it, less
  // one), then its slot from range 0's start, f and the gap before r (read
  // from lane r, or from a copy of the lanes' gaps in LDS, so that the gap
  // register need not stay live through a kernel's busiest part)
  __device__ __forceinline__ int range(int32_t f) const {
    int r = 0;
#pragma unroll
    for (int k = 1; k < NR; ++k) r += f >= pre[k] ? 1 : 0;
    return r;
  }
  // (a lane shuffle: every lane of the wave must call it, inactive source
  // lanes would read as zero)
  __device__ __forceinline__ int32_t slot(int32_t f) const {
    if (contiguous) return base0 + f;  // uniform
    return base0 + f + __shfl(gap, range(f));
  }
  __device__ __forceinline__ void store_gaps(int32_t* lds) const {
    if (threadIdx.x < NR) lds[threadIdx.x] = gap;
  }
  __device__ __forceinline__ int32_t slot(int32_t f, const int32_t* lds) const {
    if (contiguous) return base0 + f;  // uniform
    return base0 + f + lds[range(f)];
  }
};

// ---------------------------------------------------------------- paint
// Deposits of one record into the target tile.  The record's origin is local
// to its source tile, which lies o_d tiles below the target along axis d
// (o_d in {0, 1}); window points outside the target are skipped, and on an
// axis with a single tile they wrap onto it.
// A deposit w * mass in fixed point: round(w * mass_fs) * 2^sh, with
// |mass_fs| < 2^30 so that the rounding is a single-precision conversion.
template <typename T, int ORDER>
__device__ __forceinline__ void deposit_rec(const PRec& rec, int o, const TileGeom& g,
                                            unsigned long long* acc, float mass_fs, int sh) {
  int l[3];
  T wt[3][ORDER];
  window_rec<T, ORDER>(rec, l, wt);
  const int od[3] = {(o >> 2) & 1, (o >> 1) & 1, o & 1};
  int u[3][ORDER];
#pragma unroll
  for (int d = 0; d < 3; ++d)
#pragma unroll
    for (int a = 0; a < ORDER; ++a) {
      int v = l[d] + a - TILE * od[d];
      if (g.nt[d] == 1 && v >= TILE) v -= TILE;
      u[d][a] = v;
    }
#pragma unroll
  for (int a = 0; a < ORDER; ++a) {
    if ((unsigned)u[0][a] >= (unsigned)TILE) continue;
#pragma unroll
    for (int c = 0; c < ORDER; ++c) {
      if ((unsigned)u[1][c] >= (unsigned)TILE) continue;
      const T wac = wt[0][a] * wt[1][c];
#pragma unroll
      for (int d = 0; d < ORDER; ++d) {
        if ((unsigned)u[2][d] >= (unsigned)TILE) continue;
        const int qi = __float2int_rn((float)(wac * wt[2][d]) * mass_fs);
        const unsigned long long q = (unsigned long long)(long long)qi << sh;
        atomicAdd(&acc[(u[0][a] * TILE + u[1][c]) * TILE + u[2][d]], q);
      }
    }
  }
}

// The up to 64 (source tile, sub-bin) ranges whose particles deposit into the
// tile at lo: lane = o * 8 + m takes source tile t - o (o in {0,1}^3; along an
// axis with a single tile only o = 0, the neighbour being the tile itself)
// and sub-bin m, whose crossing mask must contain the direction o.  Fills
// rstart (first slot), rdir (o) and rpre (exclusive prefix of the lengths,
// rpre[64] = total); the caller synchronises.
__device__ __forceinline__ void gather_ranges(const TileGeom& g, const Workspace& w,
                                              const int lo[3], int32_t* rstart, int32_t* rpre,
                                              int32_t* rdir, bool skip_own0 = false) {
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    const int o = lane >> 3, m = lane & 7;
    const int ox = (o >> 2) & 1, oy = (o >> 1) & 1, oz = o & 1;
    const int need = ox | (oy << 1) | (oz << 2);
    int32_t b = 0, len = 0;
    const bool ok = !((ox && g.nt[0] == 1) || (oy && g.nt[1] == 1) || (oz && g.nt[2] == 1)) &&
                    (m & need) == need && !(skip_own0 && lane == 0);
    if (ok) {
      const int tx = lo[0] / TILE, ty = lo[1] / TILE, tz = lo[2] / TILE;
      const int sx = (tx - ox + g.nt[0]) % g.nt[0];
      const int sy = (ty - oy + g.nt[1]) % g.nt[1];
      const int sz = (tz - oz + g.nt[2]) % g.nt[2];
      const int64_t src = ((int64_t)sx * g.nt[1] + sy) * g.nt[2] + sz;
      b = w.offs[src * kSub + m];
      len = w.count[src * kSub + m];
    }
    // inclusive scan of the lengths across the 64 lanes
    int32_t inc = len;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int32_t v = __shfl_up(inc, d);
      if (lane >= d) inc += v;
    }
    rstart[lane] = b;
    rdir[lane] = o;
    rpre[lane + 1] = inc;
    if (lane == 0) rpre[0] = 0;
  }
}

// The up to 64 (source tile, sub-bin) ranges feeding a tile are gathered
// first (one lane each) and scanned, then all their particles are processed
// as one flat index space, so their loads are issued together.
template <typename T, int ORDER, int NT>
__device__ __forceinline__ void paint_tiles_body(const PartDesc<T>& P, const TileGeom& g,
                                                 const Workspace& w, T* mesh, float mass_fs,
                                                 int sh, double inv_fx, int64_t blk,
                                                 int64_t nblk) {
  constexpr int NB = TILE * TILE * TILE;
  __shared__ unsigned long long acc[NB];
  __shared__ int32_t rstart[64];
  __shared__ int32_t rpre[65];
  __shared__ int32_t rdir[64];
  const int64_t t = xcd_remap(blk, nblk);
  int lo[3];
  tile_coords(t, g, lo);
  // the tile's own sub-bin 0 (windows inside the tile: most of its
  // deposits) is a range known from two offsets, so its first records are
  // loaded at once, in flight while the accumulator is cleared and the other
  // ranges are gathered and scanned
  constexpr int PF = 2;
  const int32_t b0 = w.offs[t * kSub], e0 = b0 + w.count[t * kSub];
  PRec pre[PF];
#pragma unroll
  for (int u = 0; u < PF; ++u) {
    const int32_t sl = b0 + threadIdx.x + u * NT;
    pre[u] = sl < e0 ? w.rec[sl] : PRec{};
  }
  for (int i = threadIdx.x; i < NB; i += NT) acc[i] = 0ull;
  gather_ranges(g, w, lo, rstart, rpre, rdir, true);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PF; ++u)
    if (b0 + (int32_t)threadIdx.x + u * NT < e0) deposit_rec<T, ORDER>(pre[u], 0, g, acc, mass_fs, sh);
  for (int32_t sl = b0 + threadIdx.x + PF * NT; sl < e0; sl += NT)
    deposit_rec<T, ORDER>(w.rec[sl], 0, g, acc, mass_fs, sh);
  const int32_t total = rpre[64];
  for (int32_t f = threadIdx.x; f < total; f += NT) {
    // range r with rpre[r] <= f < rpre[r + 1]
    int r = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1)
      if (rpre[r + step] <= f) r += step;
    deposit_rec<T, ORDER>(w.rec[rstart[r] + (f - rpre[r])], rdir[r], g, acc, mass_fs, sh);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < NB; i += NT) {
    const int z = i % TILE, y = (i / TILE) % TILE, x = i / (TILE * TILE);
    mesh[mesh_index(lo[0] + x, lo[1] + y, lo[2] + z, g.nm)] =
        (T)((double)(long long)acc[i] * inv_fx);
  }
}

template <typename T, int ORDER, int NT = kTileThreads>
__global__ __launch_bounds__(NT) void k_paint_tiles(PartDesc<T> P, TileGeom g, Workspace w,
                                                     T* mesh, float mass_fs, int sh,
                                                     double inv_fx) {
  paint_tiles_body<T, ORDER, NT>(P, g, w, mesh, mass_fs, sh, inv_fx, blockIdx.x, gridDim.x);
}

// Weighted paint of NC per-particle weight arrays (the adjoint of a
// readout): the same gather over source ranges as k_paint_tiles, one
// component after the other through a 64-bit fixed-point LDS accumulator, so
// the sums are exact integer adds in any order and the result is bitwise
// reproducible (float LDS atomics measured 3.5x slower).  The fixed-point
// scale is chosen per tile and component from the count of depositing
// particles and the largest |weight| among them (both independent of
// order): no cell can exceed 2^62.
template <typename T>
struct TilePaintW {
  const T* w[3];
  int64_t wstride[3];
  T* out[3];
  T scale;
};

template <typename T, int ORDER, int NC>
__global__ __launch_bounds__(kTileThreads) void k_paint_tiles_weighted(TileGeom g, Workspace w,
                                                                        TilePaintW<T> A) {
  constexpr int NB = TILE * TILE * TILE;
  constexpr int NW = kTileThreads / 64;
  __shared__ unsigned long long acc[NB];
  __shared__ int32_t rstart[64];
  __shared__ int32_t rpre[65];
  __shared__ int32_t rdir[64];
  __shared__ float wmax[NC][NW];
  __shared__ int ebits[NC];
  const int64_t t = xcd_remap(blockIdx.x, gridDim.x);
  int lo[3];
  tile_coords(t, g, lo);
  gather_ranges(g, w, lo, rstart, rpre, rdir);
  __syncthreads();
  const int32_t total = rpre[64];
  auto slot_of = [&](int32_t f, int& r) {
    r = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1)
      if (rpre[r + step] <= f) r += step;
    return rstart[r] + (f - rpre[r]);
  };
  // largest |weight| of the depositing particles, per component
  float m[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) m[c] = 0.0f;
  for (int32_t f = threadIdx.x; f < total; f += kTileThreads) {
    int r;
    const int64_t p = w.list[slot_of(f, r)];
#pragma unroll
    for (int c = 0; c < NC; ++c) m[c] = fmaxf(m[c], fabsf((float)(A.w[c][p * A.wstride[c]] * A.scale)));
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m[c] = fmaxf(m[c], __shfl_xor(m[c], o));
    if ((threadIdx.x & 63) == 0) wmax[c][threadIdx.x >> 6] = m[c];
  }
  __syncthreads();
  if (threadIdx.x < NC) {
    const int c = threadIdx.x;
    float mc = 0.0f;
    for (int k = 0; k < NW; ++k) mc = fmaxf(mc, wmax[c][k]);
    int e = 0;
    if (mc > 0.0f && total > 0) {
      int k2;
      frexpf(mc, &k2);  // mc < 2^k2
      const int cl = total <= 1 ? 0 : 32 - __clz(total - 1);  // total <= 2^cl
      e = 62 - cl - k2;
    }
    ebits[c] = e;
  }
#pragma unroll 1
  for (int c = 0; c < NC; ++c) {
    for (int i = threadIdx.x; i < NB; i += kTileThreads) acc[i] = 0ull;
    __syncthreads();  // the accumulator is clear and ebits is visible
    const int e = ebits[c];
    for (int32_t f = threadIdx.x; f < total; f += kTileThreads) {
      int r;
      const int32_t sl = slot_of(f, r);
      const int64_t p = w.list[sl];
      const T wc = A.w[c][p * A.wstride[c]] * A.scale;
      int l[3];
      T wt[3][ORDER];
      window_rec<T, ORDER>(w.rec[sl], l, wt);
      const int o = rdir[r];
      const int od[3] = {(o >> 2) & 1, (o >> 1) & 1, o & 1};
      int u[3][ORDER];
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int a = 0; a < ORDER; ++a) {
          int v = l[d] + a - TILE * od[d];
          if (g.nt[d] == 1 && v >= TILE) v -= TILE;
          u[d][a] = v;
        }
#pragma unroll
      for (int a = 0; a < ORDER; ++a) {
        if ((unsigned)u[0][a] >= (unsigned)TILE) continue;
#pragma unroll
        for (int b = 0; b < ORDER; ++b) {
          if ((unsigned)u[1][b] >= (unsigned)TILE) continue;
          const T wab = wt[0][a] * wt[1][b];
#pragma unroll
          for (int d = 0; d < ORDER; ++d) {
            if ((unsigned)u[2][d] >= (unsigned)TILE) continue;
            const double v = ldexp((double)(wab * wt[2][d]) * (double)wc, e);
            atomicAdd(&acc[(u[0][a] * TILE + u[1][b]) * TILE + u[2][d]],
                      (unsigned long long)__double2ll_rn(v));
          }
        }
      }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < NB; i += kTileThreads) {
      const int z = i % TILE, y = (i / TILE) % TILE, x = i / (TILE * TILE);
      const int64_t oo = mesh_index(lo[0] + x, lo[1] + y, lo[2] + z, g.nm);
      A.out[c][oo] = (T)ldexp((double)(long long)acc[i], -e);
    }
    __syncthreads();  // the next component clears the accumulator
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

// LDS cell vector: the NC components of one cell side by side (three are
// padded to four), so a window point is one LDS read for every component.
template <typename T, int NC> struct CellVec;
template <typename T> struct CellVec<T, 1> { typedef T type; };
template <typename T> struct CellVec<T, 2> { typedef T type __attribute__((ext_vector_type(2))); };
template <typename T> struct CellVec<T, 3> { typedef T type __attribute__((ext_vector_type(4))); };
template <typename T> struct CellVec<T, 4> { typedef T type __attribute__((ext_vector_type(4))); };

template <typename T, int NC, typename VC>
__device__ __forceinline__ T cell_comp(const VC& v, int c) {
  if constexpr (NC == 1) return v;
  else return v[c];
}

template <typename T, int NC, typename VC>
__device__ __forceinline__ void set_comp(VC& v, int c, T x) {
  if constexpr (NC == 1) v = x;
  else v[c] = x;
}

// Box staging of one tile into registers: each (x, y) row holds TILE aligned
// cells of every component, loaded as TILE/4 four-wide vectors per component,
// plus H wrapped overhang cells.
template <typename T, int ORDER, int NC>
struct BoxRegs {
  typedef T V4 __attribute__((ext_vector_type(4)));
  static constexpr int H = ORDER - 1;
  static constexpr int TB = TILE + H;
  static constexpr int NQ = TILE / 4;
  static constexpr int NV = TB * TB * NQ;
  static constexpr int PV = (NV + kTileThreads - 1) / kTileThreads;
  static constexpr int NH = TB * TB * H;
  static constexpr int PH = (NH + kTileThreads - 1) / kTileThreads;
  V4 vr[PV][NC];
  T hr[PH][NC];

  __device__ __forceinline__ void load(const TileGeom& g, const TileRead<T>& R, const int lo[3]) {
#pragma unroll
    for (int k = 0; k < PV; ++k) {
      const int i = threadIdx.x + k * kTileThreads;
      if (i < NV) {
        const int q = i % NQ, r = i / NQ;
        const int y = r % TB, x = r / TB;
        int gx = lo[0] + x, gy = lo[1] + y;
        if (gx >= g.nm[0]) gx -= g.nm[0];
        if (gy >= g.nm[1]) gy -= g.nm[1];
        const int64_t off = mesh_index(gx, gy, lo[2] + 4 * q, g.nm);
#pragma unroll
        for (int c = 0; c < NC; ++c) vr[k][c] = *reinterpret_cast<const V4*>(R.mesh[c] + off);
      }
    }
#pragma unroll
    for (int k = 0; k < PH; ++k) {
      const int i = threadIdx.x + k * kTileThreads;
      if (i < NH) {
        const int h = i % H, r = i / H;
        const int y = r % TB, x = r / TB;
        int gx = lo[0] + x, gy = lo[1] + y, gz = lo[2] + TILE + h;
        if (gx >= g.nm[0]) gx -= g.nm[0];
        if (gy >= g.nm[1]) gy -= g.nm[1];
        if (gz >= g.nm[2]) gz -= g.nm[2];
        const int64_t off = mesh_index(gx, gy, gz, g.nm);
#pragma unroll
        for (int c = 0; c < NC; ++c) hr[k][c] = R.mesh[c][off];
      }
    }
  }

  template <typename VC>
  __device__ __forceinline__ void store(VC* box) const {
#pragma unroll
    for (int k = 0; k < PV; ++k) {
      const int i = threadIdx.x + k * kTileThreads;
      if (i < NV) {
        const int q = i % NQ, r = i / NQ;
        VC* dst = &box[r * TB + 4 * q];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          VC v{};
#pragma unroll
          for (int c = 0; c < NC; ++c) set_comp<T, NC>(v, c, vr[k][c][u]);
          dst[u] = v;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < PH; ++k) {
      const int i = threadIdx.x + k * kTileThreads;
      if (i < NH) {
        const int h = i % H, r = i / H;
        VC v{};
#pragma unroll
        for (int c = 0; c < NC; ++c) set_comp<T, NC>(v, c, hr[k][c]);
        box[r * TB + TILE + h] = v;
      }
    }
  }
};

// Interpolation of one particle from the staged box.
template <typename T, int ORDER, int NC, typename VC>
__device__ __forceinline__ void gather_one(const VC* box, const TileRead<T>& R, int64_t p,
                                           const int l[3], const T wt[3][ORDER]) {
  constexpr int TB = TILE + ORDER - 1;
  const int lx = l[0], ly = l[1], lz = l[2];
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
        const VC v = box[row + d];
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c] += wv * cell_comp<T, NC>(v, c);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) R.out[c][p * R.stride[c]] = acc[c] * R.scale[c];
}

// One workgroup per tile: the (TILE + ORDER - 1)^3 box of every component is
// staged in LDS, then the tile's particles are interpolated from it.  The
// indices of each thread's first two particles are fetched before the box and
// their position loads issued before the LDS stores, so those latencies
// overlap the box loads (a tile holds 512 particles on average at B = 2).
// (A persistent variant that prefetched the next tile's box while
// interpolating measured slower at 512^3 on gfx950: 1.27 vs 0.98 ms.)
template <typename T, int ORDER, int NC>
__global__ __launch_bounds__(kTileThreads) void k_readout_tiles(PartDesc<T> P, TileGeom g,
                                                                 Workspace w, TileRead<T> R) {
  constexpr int TB = TILE + ORDER - 1;
  constexpr int NB = TB * TB * TB;
  typedef typename CellVec<T, NC>::type VC;
  __shared__ VC box[NB];
  __shared__ int32_t sgap[kSub];
  const int64_t t = xcd_remap(blockIdx.x, gridDim.x);
  int lo[3];
  tile_coords(t, g, lo);
  TileSlots<kSub> ts;
  ts.load(w, t);
  const int32_t e = ts.total();
  if (e == 0) return;  // uniform across the workgroup
  const int32_t s0 = threadIdx.x, s1 = s0 + kTileThreads;
  const int32_t q0 = ts.slot(s0), q1 = ts.slot(s1);  // used below only when < e
  const int64_t p0 = s0 < e ? w.list[q0] : 0;
  const int64_t p1 = s1 < e ? w.list[q1] : 0;
  const PRec r0 = s0 < e ? w.rec[q0] : PRec{};
  const PRec r1 = s1 < e ? w.rec[q1] : PRec{};
  ts.store_gaps(sgap);
  BoxRegs<T, ORDER, NC> br;
  br.load(g, R, lo);
  br.store(box);
  __syncthreads();
  int l[3];
  T wt[3][ORDER];
  if (s0 < e) {
    window_rec<T, ORDER>(r0, l, wt);
    gather_one<T, ORDER, NC>(box, R, p0, l, wt);
  }
  if (s1 < e) {
    window_rec<T, ORDER>(r1, l, wt);
    gather_one<T, ORDER, NC>(box, R, p1, l, wt);
  }
  for (int32_t s = s1 + kTileThreads; s < e; s += kTileThreads) {
    const int32_t q = ts.slot(s, sgap);
    window_rec<T, ORDER>(w.rec[q], l, wt);
    gather_one<T, ORDER, NC>(box, R, (int64_t)w.list[q], l, wt);
  }
}

// ---------------------------------------------------------------- force readout
// The four-point gradient of the force, i a(k) with
// a = (8 sin kh - sin 2kh) / (6h), is the Fourier symbol of the real-space
// stencil D f(x) = [8 (f(x+h) - f(x-h)) - (f(x+2h) - f(x-2h))] / (12 h), so
// c2r(i a_d pot_k) = D_d c2r(pot_k) exactly: one inverse transform of the
// potential replaces three, and the stencil is applied while interpolating.
// A workgroup per tile stages the density box (lo .. lo+TILE+ORDER-2) and the
// potential box with two more cells on each side, then reads RHO and
// F_d = sum_window w * D_d phi for its particles.
template <typename T>
struct GradRead {
  const T* rho;      // density mesh (nullptr: no RHO output)
  const T* phi;      // potential mesh
  T* rho_out;
  T* f_out;          // f_out[p * f_stride + d]
  int64_t rho_stride, f_stride;
  T rho_scale;
  T f_scale[3];      // force factor / (12 h_d)
};

// Scalar box of (TILE + L + U)^2 x (ZT + L + U) cells starting at lo - L
// (ZT = TILE for one tile, 2 TILE for two tiles adjacent along z), staged
// with the aligned ZT-cell middle of every row as four-wide vectors and
// L + U wrapped edge cells.  With CORE > 0 only the cells a stencil readout
// can reach are loaded: the core [L, L + CORE)^2 x [L, L + CORE + ZT - TILE)
// extended along one axis at a time (the four-point stencil runs along one
// axis while the window stays in the core of the other two); the rest of the
// box is left unwritten and never read.
template <typename T, int L, int U, int CORE = 0, int NT = kTileThreads, int ZT = TILE>
struct ScalarBox {
  typedef T V4 __attribute__((ext_vector_type(4)));
  static constexpr int TB = TILE + L + U;   // x and y
  static constexpr int TBZ = ZT + L + U;
  static constexpr int NQ = ZT / 4;
  static constexpr int NV = TB * TB * NQ;
  static constexpr int PV = (NV + NT - 1) / NT;
  static constexpr int NE = TB * TB * (L + U);
  static constexpr int PE = (NE + NT - 1) / NT;
  V4 vr[PV];
  T er[PE];

  __device__ __forceinline__ static int wrap(int i, int n) {
    return i < 0 ? i + n : (i >= n ? i - n : i);
  }

  // in the core along x or y (box coordinate b), or along z
  __device__ __forceinline__ static bool in_core(int b) { return b >= L && b < L + CORE; }
  __device__ __forceinline__ static bool in_core_z(int b) {
    return b >= L && b < L + CORE + ZT - TILE;
  }

  // Row-middle element i: box row (bx, by) and four-cell group q.
  __device__ __forceinline__ static void mid_of(int i, int& bx, int& by, int& q) {
    q = i % NQ;
    const int r = i / NQ;
    by = r % TB;
    bx = r / TB;
  }
  // Edge element i: box row (bx, by) and box z coordinate bz.
  __device__ __forceinline__ static void edge_of(int i, int& bx, int& by, int& bz) {
    const int h = i % (L + U), r = i / (L + U);
    by = r % TB;
    bx = r / TB;
    bz = h < L ? h : ZT + h;
  }
  __device__ __forceinline__ static bool mid_needed(int bx, int by) {
    // the middle of a row is in the z core: the row is needed unless both
    // x and y are outside the core
    return CORE == 0 || in_core(bx) || in_core(by);
  }
  __device__ __forceinline__ static bool edge_needed(int bx, int by, int bz) {
    // an edge cell outside the z core is needed only on a core row
    return CORE == 0 || (in_core(bx) && in_core(by)) ||
           (in_core_z(bz) && (in_core(bx) || in_core(by)));
  }

  __device__ __forceinline__ void load(const TileGeom& g, const T* mesh, const int lo[3]) {
    // A box inside the mesh (most tiles) is indexed without wrapping: the
    // box origin's cell plus bx * n1 n2 + by * n2 + bz, one 24-bit and one
    // 32 x 32 -> 64-bit multiply-add per element (n1 n2 < 2^32).
    bool inner = true;
#pragma unroll
    for (int d = 0; d < 3; ++d)
      inner = inner && lo[d] >= L && lo[d] + (d == 2 ? ZT : TILE) + U <= g.nm[d];
    if (inner) {
      const T* base = mesh + mesh_index(lo[0] - L, lo[1] - L, lo[2] - L, g.nm);
      const uint32_t n2 = (uint32_t)g.nm[2], n12 = (uint32_t)g.nm[1] * n2;
      auto at = [&](int bx, int by, int bz) {
        return base + ((uint64_t)(uint32_t)bx * n12 +
                       (uint32_t)(__umul24((uint32_t)by, n2) + (uint32_t)bz));
      };
#pragma unroll
      for (int k = 0; k < PV; ++k) {
        const int i = threadIdx.x + k * NT;
        if (i < NV) {
          int bx, by, q;
          mid_of(i, bx, by, q);
          vr[k] = mid_needed(bx, by) ? *reinterpret_cast<const V4*>(at(bx, by, L + 4 * q))
                                     : V4{(T)0, (T)0, (T)0, (T)0};
        }
      }
#pragma unroll
      for (int k = 0; k < PE; ++k) {
        const int i = threadIdx.x + k * NT;
        if (i < NE) {
          int bx, by, bz;
          edge_of(i, bx, by, bz);
          er[k] = edge_needed(bx, by, bz) ? *at(bx, by, bz) : (T)0;
        }
      }
      return;
    }
#pragma unroll
    for (int k = 0; k < PV; ++k) {
      const int i = threadIdx.x + k * NT;
      if (i < NV) {
        int bx, by, q;
        mid_of(i, bx, by, q);
        if (mid_needed(bx, by)) {
          const int gy = wrap(lo[1] - L + by, g.nm[1]);
          const int gx = wrap(lo[0] - L + bx, g.nm[0]);
          vr[k] = *reinterpret_cast<const V4*>(mesh + mesh_index(gx, gy, lo[2] + 4 * q, g.nm));
        } else {
          vr[k] = V4{(T)0, (T)0, (T)0, (T)0};
        }
      }
    }
#pragma unroll
    for (int k = 0; k < PE; ++k) {
      const int i = threadIdx.x + k * NT;
      if (i < NE) {
        int bx, by, bz;
        edge_of(i, bx, by, bz);
        if (edge_needed(bx, by, bz)) {
          const int gy = wrap(lo[1] - L + by, g.nm[1]);
          const int gx = wrap(lo[0] - L + bx, g.nm[0]);
          const int gz = wrap(lo[2] - L + bz, g.nm[2]);
          er[k] = mesh[mesh_index(gx, gy, gz, g.nm)];
        } else {
          er[k] = (T)0;
        }
      }
    }
  }

  __device__ __forceinline__ void store(T* box) const {
#pragma unroll
    for (int k = 0; k < PV; ++k) {
      const int i = threadIdx.x + k * NT;
      if (i < NV) {
        const int q = i % NQ, r = i / NQ;
        T* dst = &box[r * TBZ + L + 4 * q];
        dst[0] = vr[k].x;
        dst[1] = vr[k].y;
        dst[2] = vr[k].z;
        dst[3] = vr[k].w;
      }
    }
#pragma unroll
    for (int k = 0; k < PE; ++k) {
      const int i = threadIdx.x + k * NT;
      if (i < NE) {
        const int h = i % (L + U), r = i / (L + U);
        box[r * TBZ + (h < L ? h : ZT + h)] = er[k];
      }
    }
  }
};

// The density at tile-box coordinates (x, y, z) (box origin at the tile's
// first cell): staged in LDS (a box ZT + ORDER - 1 cells long along z), or
// read from the mesh through the cache.
template <typename T, int ORDER, int ZT = TILE>
struct RhoLds {
  const T* rbox;
  __device__ __forceinline__ T operator()(int x, int y, int z) const {
    constexpr int TR = TILE + ORDER - 1, TRZ = ZT + ORDER - 1;
    return rbox[(x * TR + y) * TRZ + z];
  }
};

template <typename T>
struct RhoMesh {
  const T* rho;
  int lo[3], nm[3];
  __device__ __forceinline__ T operator()(int x, int y, int z) const {
    int g[3] = {lo[0] + x, lo[1] + y, lo[2] + z};
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (g[d] >= nm[d]) g[d] -= nm[d];
    return rho[mesh_index(g[0], g[1], g[2], nm)];
  }
};

// Window origin l in box coordinates of the density box (the potential
// box starts two cells earlier on every axis); the boxes are ZT + ORDER - 1
// and ZT + ORDER + 3 cells long along z.  rho_one writes RHO, force_one the
// force; grad_one does both.
template <typename T, int ORDER, typename RhoAt>
__device__ __forceinline__ void rho_one(const RhoAt& rho_at, const GradRead<T>& G, int64_t p,
                                        const int l[3], const T wt[3][ORDER]) {
  const int lx = l[0], ly = l[1], lz = l[2];
  T acc = (T)0;
#pragma unroll
  for (int a = 0; a < ORDER; ++a)
#pragma unroll
    for (int b = 0; b < ORDER; ++b) {
      const T wab = wt[0][a] * wt[1][b];
#pragma unroll
      for (int c = 0; c < ORDER; ++c) acc += wab * wt[2][c] * rho_at(lx + a, ly + b, lz + c);
    }
  G.rho_out[p * G.rho_stride] = acc * G.rho_scale;
}

template <typename T, int ORDER, int ZT = TILE>
__device__ __forceinline__ void force_one(const T* pbox, const GradRead<T>& G, int64_t p,
                                          const int l[3], const T wt[3][ORDER]) {
  constexpr int TP = TILE + ORDER + 3;   // potential box side (2 more cells each way)
  constexpr int TPZ = ZT + ORDER + 3;
  constexpr int NLN = ORDER + 4;         // stencil line: window points -2 .. +2
  const int lx = l[0], ly = l[1], lz = l[2];
  // potential box index of window point (a, b, c): (lx+2+a, ly+2+b, lz+2+c)
  auto at = [&](int x, int y, int z) { return pbox[(x * TP + y) * TPZ + z]; };
  T f[3] = {(T)0, (T)0, (T)0};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
#pragma unroll
    for (int b = 0; b < ORDER; ++b)
#pragma unroll
      for (int c = 0; c < ORDER; ++c) {
        T line[NLN];
#pragma unroll
        for (int u = 0; u < NLN; ++u) {
          int ix[3];
          ix[d] = (d == 0 ? lx : (d == 1 ? ly : lz)) + u;
          ix[e1] = (e1 == 0 ? lx : (e1 == 1 ? ly : lz)) + 2 + b;
          ix[e2] = (e2 == 0 ? lx : (e2 == 1 ? ly : lz)) + 2 + c;
          line[u] = at(ix[0], ix[1], ix[2]);
        }
        T s = (T)0;
#pragma unroll
        for (int a = 0; a < ORDER; ++a)
          s += wt[d][a] * ((T)8 * (line[a + 3] - line[a + 1]) - (line[a + 4] - line[a]));
        f[d] += wt[e1][b] * wt[e2][c] * s;
      }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) G.f_out[p * G.f_stride + d] = f[d] * G.f_scale[d];
}

template <typename T, int ORDER, int ZT = TILE, typename RhoAt>
__device__ __forceinline__ void grad_one(const RhoAt& rho_at, const T* pbox, const GradRead<T>& G,
                                         int64_t p, const int l[3], const T wt[3][ORDER]) {
  if (G.rho_out) rho_one<T, ORDER>(rho_at, G, p, l, wt);
  force_one<T, ORDER, ZT>(pbox, G, p, l, wt);
}

// RLDS: the density box staged in LDS (true) or the density read through
// the cache (false: a third less LDS, the same values in the same order).
// SPAN: tiles per workgroup along z (1 or 2; the tile count along z is a
// multiple of SPAN).  Two z-adjacent tiles share their boxes: each box row's
// middle is then a whole 128-byte line, and the halo rows are loaded once
// for both, so a particle costs about 40% fewer cache-line requests.  A
// particle of the second tile has its window origin moved by TILE along z
// into the shared boxes; the values read and the arithmetic are the same.
template <typename T, int ORDER, int NT, bool RLDS = true, int SPAN = 1>
__device__ __forceinline__ void readout_grad_body(const PartDesc<T>& P, const TileGeom& g,
                                                  const Workspace& w, const GradRead<T>& G,
                                                  int64_t blk, int64_t nblk) {
  constexpr int ZT = SPAN * TILE;
  constexpr int TR = TILE + ORDER - 1, TRZ = ZT + ORDER - 1;
  constexpr int TP = TILE + ORDER + 3, TPZ = ZT + ORDER + 3;
  // two tiles with the density from LDS: the density box and then the
  // potential box in one buffer (RHO read out in between), so that two
  // workgroups share a CU
  constexpr bool PHASED = SPAN > 1 && RLDS;
  static_assert(TR * TR * TRZ <= TP * TP * TPZ, "the density box fits the potential buffer");
  __shared__ T rbox[RLDS && !PHASED ? TR * TR * TRZ : 1];
  __shared__ T pbox[TP * TP * TPZ];
  __shared__ int32_t sgap[SPAN * kSub];
  // z is the fastest tile axis, so unit u's tiles are u * SPAN + [0, SPAN)
  const int64_t t = xcd_remap(blk, nblk) * SPAN;
  int lo[3];
  tile_coords(t, g, lo);
  // records 0 .. e - 1 of the tiles (index s; slot ts.slot(s)), the
  // second tile's from mid
  TileSlots<SPAN * kSub> ts;
  ts.load(w, t);
  const int32_t e = ts.total();
  const int32_t mid = SPAN == 1 ? e : ts.pre[kSub];
  if (e == 0) return;  // uniform across the workgroup
  // each thread's first two particles, loaded while the first box's loads
  // are in flight (their slots take a few instructions to find)
  const int32_t s0 = threadIdx.x, s1 = s0 + NT;
  int64_t p0 = 0, p1 = 0;
  PRec r0{}, r1{};
  auto preload = [&]() {
    const int32_t q0 = ts.slot(s0), q1 = ts.slot(s1);  // used only when < e
    if (s0 < e) {
      p0 = w.list[q0];
      r0 = w.rec[q0];
    }
    if (s1 < e) {
      p1 = w.list[q1];
      r1 = w.rec[q1];
    }
    ts.store_gaps(sgap);  // read after the box's barrier
  };
  auto window = [&](int32_t s, const PRec& r, int l[3], T wt[3][ORDER]) {
    window_rec<T, ORDER>(r, l, wt);
    if (SPAN > 1 && s >= mid) l[2] += TILE;
  };
  ScalarBox<T, 2, ORDER + 1, TILE + ORDER - 1, NT, ZT> pb;
  if constexpr (PHASED) {
    if (G.rho_out) {
      {
        ScalarBox<T, 0, ORDER - 1, 0, NT, ZT> rb;
        rb.load(g, G.rho, lo);
        preload();
        rb.store(pbox);
      }
      __syncthreads();
      pb.load(g, G.phi, lo);  // in flight while RHO is read out
      const RhoLds<T, ORDER, ZT> rl{pbox};
      auto rone = [&](int32_t s, int64_t p, const PRec& r) {
        int l[3];
        T wt[3][ORDER];
        window(s, r, l, wt);
        rho_one<T, ORDER>(rl, G, p, l, wt);
      };
      if (s0 < e) rone(s0, p0, r0);
      if (s1 < e) rone(s1, p1, r1);
      for (int32_t s = s1 + NT; s < e; s += NT) {
        const int32_t q = ts.slot(s, sgap);
        rone(s, (int64_t)w.list[q], w.rec[q]);
      }
      __syncthreads();  // every read of the density box is done
    } else {
      pb.load(g, G.phi, lo);
      preload();
    }
    pb.store(pbox);
    __syncthreads();
    auto fone = [&](int32_t s, int64_t p, const PRec& r) {
      int l[3];
      T wt[3][ORDER];
      window(s, r, l, wt);
      force_one<T, ORDER, ZT>(pbox, G, p, l, wt);
    };
    if (s0 < e) fone(s0, p0, r0);
    if (s1 < e) fone(s1, p1, r1);
    for (int32_t s = s1 + NT; s < e; s += NT) {
      const int32_t q = ts.slot(s, sgap);
      fone(s, (int64_t)w.list[q], w.rec[q]);
    }
    return;
  }
  pb.load(g, G.phi, lo);
  preload();
  if constexpr (RLDS) {
    ScalarBox<T, 0, ORDER - 1, 0, NT, ZT> rb;
    if (G.rho_out) rb.load(g, G.rho, lo);
    pb.store(pbox);
    if (G.rho_out) rb.store(rbox);
  } else {
    pb.store(pbox);
  }
  __syncthreads();
  const RhoLds<T, ORDER, ZT> rl{rbox};
  const RhoMesh<T> rm{G.rho, {lo[0], lo[1], lo[2]}, {g.nm[0], g.nm[1], g.nm[2]}};
  auto one = [&](int32_t s, int64_t p, const PRec& r) {
    int l[3];
    T wt[3][ORDER];
    window(s, r, l, wt);
    if constexpr (RLDS) grad_one<T, ORDER, ZT>(rl, pbox, G, p, l, wt);
    else grad_one<T, ORDER, ZT>(rm, pbox, G, p, l, wt);
  };
  if (s0 < e) one(s0, p0, r0);
  if (s1 < e) one(s1, p1, r1);
  for (int32_t s = s1 + NT; s < e; s += NT) {
    const int32_t q = ts.slot(s, sgap);
    one(s, (int64_t)w.list[q], w.rec[q]);
  }
}

template <typename T, int ORDER, int NT = grad_threads<ORDER>(), bool RLDS = true, int SPAN = 1>
__global__ __launch_bounds__(NT) void k_readout_grad_tiles(PartDesc<T> P, TileGeom g,
                                                                      Workspace w, GradRead<T> G) {
  readout_grad_body<T, ORDER, NT, RLDS, SPAN>(P, g, w, G, blockIdx.x, gridDim.x);
}

// Position gradient of the stencil force readout (the window-derivative
// part of gravity_gradient, fastpm/operators.py:236-247, with the four-point
// gradient as a stencil): for each particle
//   gx[e] (+)= scale * inv_cell[e] * sum_d gf[d] d/du_e readout(D_d phi)(u)
// with D_d including 1 / (12 h_d), from the same LDS-staged potential box as
// k_readout_grad_tiles, so the three gradient meshes are never formed.
template <typename T>
struct PosGrad {
  const T* phi;
  const T* gf;        // gf[p * gf_stride + d]
  T* gx;              // gx[3 p + e]
  int64_t gf_stride;
  T dscale[3];        // 1 / (12 h_d)
  T oscale[3];        // scale * Nmesh_e / L_e
  int accumulate;
};

template <typename T, int ORDER>
__device__ __forceinline__ void window_dw(const PRec& r, int l[3], T w[3][ORDER], T dw[3][ORDER]) {
  window_rec<T, ORDER>(r, l, w);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const T t = (T)r.t[d];
    if constexpr (ORDER == 2) {
      dw[d][0] = (T)-1;
      dw[d][1] = (T)1;
    } else {
      dw[d][0] = -((T)0.5 - t);
      dw[d][1] = (T)-2 * t;
      dw[d][2] = (T)0.5 + t;
    }
  }
}

template <typename T, int ORDER>
__device__ __forceinline__ void pos_grad_one(const T* pbox, const PosGrad<T>& G, int64_t p,
                                             const int l[3], const T w[3][ORDER],
                                             const T dw[3][ORDER]) {
  constexpr int TP = TILE + ORDER + 3;
  constexpr int NLN = ORDER + 4;
  const int lv[3] = {l[0], l[1], l[2]};
  auto at = [&](int x, int y, int z) { return pbox[(x * TP + y) * TP + z]; };
  T gfv[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) gfv[d] = G.gf[p * G.gf_stride + d] * G.dscale[d];
  T g[3] = {(T)0, (T)0, (T)0};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int e1 = (d + 1) % 3, e2 = (d + 2) % 3;
#pragma unroll
    for (int b = 0; b < ORDER; ++b)
#pragma unroll
      for (int c = 0; c < ORDER; ++c) {
        T line[NLN];
#pragma unroll
        for (int u = 0; u < NLN; ++u) {
          int ix[3];
          ix[d] = lv[d] + u;
          ix[e1] = lv[e1] + 2 + b;
          ix[e2] = lv[e2] + 2 + c;
          line[u] = at(ix[0], ix[1], ix[2]);
        }
        T s = (T)0, sd = (T)0;
#pragma unroll
        for (int a = 0; a < ORDER; ++a) {
          const T st = (T)8 * (line[a + 3] - line[a + 1]) - (line[a + 4] - line[a]);
          s += w[d][a] * st;
          sd += dw[d][a] * st;
        }
        g[d] += gfv[d] * w[e1][b] * w[e2][c] * sd;
        g[e1] += gfv[d] * dw[e1][b] * w[e2][c] * s;
        g[e2] += gfv[d] * w[e1][b] * dw[e2][c] * s;
      }
  }
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const T v = g[e] * G.oscale[e];
    G.gx[3 * p + e] = G.accumulate ? G.gx[3 * p + e] + v : v;
  }
}

template <typename T, int ORDER>
__global__ __launch_bounds__(kTileThreads) void k_pos_grad_tiles(TileGeom g, Workspace w,
                                                                 PosGrad<T> G) {
  constexpr int TP = TILE + ORDER + 3;
  __shared__ T pbox[TP * TP * TP];
  __shared__ int32_t sgap[kSub];
  const int64_t t = xcd_remap(blockIdx.x, gridDim.x);
  int lo[3];
  tile_coords(t, g, lo);
  TileSlots<kSub> ts;
  ts.load(w, t);
  const int32_t e = ts.total();
  if (e == 0) return;  // uniform across the workgroup
  ts.store_gaps(sgap);
  ScalarBox<T, 2, ORDER + 1, TILE + ORDER - 1> pb;
  pb.load(g, G.phi, lo);
  pb.store(pbox);
  __syncthreads();
  for (int32_t s = threadIdx.x; s < e; s += kTileThreads) {
    const int32_t q = ts.slot(s, sgap);
    int l[3];
    T wt[3][ORDER], dwt[3][ORDER];
    window_dw<T, ORDER>(w.rec[q], l, wt, dwt);
    pos_grad_one<T, ORDER>(pbox, G, (int64_t)w.list[q], l, wt, dwt);
  }
}


// ---------------------------------------------------------------- host side
// Tiled meshes: sides multiples of TILE, below 2^24 cells, fewer than 2^32
// rows (mesh_index's 32-bit row arithmetic) and fewer than 2^32 cells per
// x plane (the box loads' 32-bit plane stride).
static bool tiles_ok(const fpm_mesh* m) {
  if (!m) return false;
  for (int d = 0; d < 3; ++d)
    if (m->nmesh[d] <= 0 || m->nmesh[d] % TILE || m->nmesh[d] >= (1 << 24)) return false;
  return (int64_t)m->nmesh[0] * m->nmesh[1] < ((int64_t)1 << 32) &&
         (int64_t)m->nmesh[1] * m->nmesh[2] < ((int64_t)1 << 32);
}

static unsigned grid_parts(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > (1 << 20)) g = 1 << 20;
  return (unsigned)g;
}

static unsigned grid_bin(int64_t n) {
  const int64_t span = (int64_t)kBinThreads * kBinILP;
  int64_t g = (n + span - 1) / span;
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
  g = tile_geom(m);
  if (g.nbins >= (int64_t)1 << 31) return FPM_ERR_ARG;
  const int64_t need = layout_ws(g, P.n, static_cast<char*>(ws), &w);
  if (ws_bytes < need) return FPM_ERR_WORKSPACE;
  return FPM_OK;
}

// The walk over lattice particles of the count and scatter passes: 4 x 8 x 16
// blocks where the lattice divides into them, else 8 x 8 patches, else index
// order (the mesh's FPM_VAR_BIN field can ask for patches or index order).
// Every walk gives the same sub-bins; only the order within a sub-bin differs,
// which the fixed-point paint and the readouts do not see.
static BinOrder bin_order(const fpm_particles* part, const fpm_mesh* m) {
  BinOrder o{};
  const int bv = FPM_VAR_FIELD(m->variant, FPM_VAR_BIN);
  if (bv != 2 && !part->q && part->lattice[2] % 8 == 0 && part->lattice[1] % 8 == 0) {
    // block order needs whole blocks and a workgroup span of one block
    const bool block =
        bv == 0 && part->lattice[0] % kBlockX == 0 && part->lattice[2] % 16 == 0;
    o.blocked = block ? kBinBlock : kBinPatch;
    o.nz = (uint32_t)(part->lattice[2] / (block ? 16 : 8));
    o.ny = (uint32_t)(part->lattice[1] / 8);
    o.lat1 = (uint32_t)part->lattice[1];
    o.lat2 = (uint32_t)part->lattice[2];
  }
  return o;
}

static void bin_clear(const TileGeom& g, const Workspace& w, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_i32, dim3(grid_parts(g.nbins)), dim3(256), 0, st, w.count, g.nbins);
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, w.misplaced, (int64_t)2);  // and ovf
}

// the scan of the counts and the scatter, after a count pass
template <typename T>
static int bin_finish(const PartDesc<T>& P, const TileGeom& g, const Workspace& w,
                      const BinOrder& o, int order, hipStream_t st) {
  const unsigned nsb = (unsigned)scan_blocks(g);
  hipLaunchKernelGGL(k_bin_scan_local, dim3(nsb), dim3(kScanThreads), 0, st, w, g.nbins);
  hipLaunchKernelGGL(k_bin_scan_add, dim3(nsb), dim3(kScanThreads), 0, st, w, g.nbins);
  if (P.n > 0) {
    // the scatter runs only when the count pass set ovf; its workgroups
    // stride the particles, so a grid of a few per CU returns quickly when
    // it has nothing to do
    const unsigned gs = grid_bin(P.n) < 2048u ? grid_bin(P.n) : 2048u;
    if (order == 2)
      hipLaunchKernelGGL((k_bin_scatter<T, 2>), dim3(gs), dim3(kBinThreads), 0, st, P, g, w, o);
    else
      hipLaunchKernelGGL((k_bin_scatter<T, 3>), dim3(gs), dim3(kBinThreads), 0, st, P, g, w, o);
  }
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T>
static int bin_impl(const fpm_particles* part, const fpm_mesh* m, int order, void* ws,
                    int64_t ws_bytes, hipStream_t st, bool rebin,
                    const KickDrift<T>& kd = KickDrift<T>{}) {
  PartDesc<T> P;
  TileGeom g;
  Workspace w;
  int rc = prep<T>(part, m, order, ws, ws_bytes, P, g, w);
  if (rc) return rc;
  w.allow_cap = rebin ? 1 : 0;
  const BinOrder o = bin_order(part, m);
  bin_clear(g, w, st);
  if (P.n > 0) {
    if (order == 2)
      hipLaunchKernelGGL((k_bin_count<T, 2>), dim3(grid_bin(P.n)), dim3(kBinThreads), 0, st, P, g,
                         w, o, kd);
    else
      hipLaunchKernelGGL((k_bin_count<T, 3>), dim3(grid_bin(P.n)), dim3(kBinThreads), 0, st, P, g,
                         w, o, kd);
  }
  return bin_finish<T>(P, g, w, o, order, st);
}

// Threads per tile of the tiled paint on mesh m (FPM_VAR_PAINT).
static int paint_nt(const fpm_mesh* m) {
  switch (FPM_VAR_FIELD(m->variant, FPM_VAR_PAINT)) {
    case 1: return 256;
    case 2: return 1024;
    default: return kTileThreads;
  }
}

// Fixed-point scale of the tiled paint of n particles of `mass`: deposits are
// summed as 64-bit integers in units of 2^-S, with the total |mass| of all
// particles below 2^62, and each deposit is rounded to a multiple of 2^sh,
// the shift that keeps |mass| 2^(S - sh) below 2^30 (a relative step of
// 2^-30 of the mass), so that the rounding is a single-precision conversion.
static int paint_scale(int64_t n, double mass, float& mass_fs, int& sh, double& inv_fx) {
  if (!(std::fabs(mass) > 0) || !std::isfinite(mass)) return FPM_ERR_ARG;
  const double total = std::fabs(mass) * (double)(n > 0 ? n : 1);
  int S = 62 - (int)std::ceil(std::log2(total + 1.0));
  if (S > 60) S = 60;
  int e = 0;
  std::frexp(std::fabs(mass) * std::ldexp(1.0, S), &e);  // |mass| 2^S < 2^e
  sh = e > 30 ? e - 30 : 0;
  mass_fs = (float)(mass * std::ldexp(1.0, S - sh));
  inv_fx = std::ldexp(1.0, -S);
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
  if (!out) return FPM_ERR_ARG;
  float mass_fs;
  int sh;
  double inv_fx;
  if ((rc = paint_scale(P.n, mass, mass_fs, sh, inv_fx))) return rc;
  T* mesh = static_cast<T*>(out);
  const dim3 grid((unsigned)g.ntiles);
  switch (paint_nt(m) * 4 + order) {
#define FPM_PAINT(O_, NT_)                                                                      \
  case NT_ * 4 + O_:                                                                            \
    hipLaunchKernelGGL((k_paint_tiles<T, O_, NT_>), grid, dim3(NT_), 0, st, P, g, w, mesh,       \
                       mass_fs, sh, inv_fx);                                                    \
    break;
    FPM_PAINT(2, 256) FPM_PAINT(3, 256) FPM_PAINT(2, 512) FPM_PAINT(3, 512)
    FPM_PAINT(2, 1024) FPM_PAINT(3, 1024)
#undef FPM_PAINT
    default: return FPM_ERR_ARG;
  }
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T, int ORDER, int NC>
static void launch_rt_nc(const PartDesc<T>& P, const TileGeom& g, const Workspace& w,
                         const TileRead<T>& R, hipStream_t st) {
  hipLaunchKernelGGL((k_readout_tiles<T, ORDER, NC>), dim3((unsigned)g.ntiles),
                     dim3(kTileThreads), 0, st, P, g, w, R);
}

template <typename T, int ORDER>
static void launch_rt(const PartDesc<T>& P, const TileGeom& g, const Workspace& w,
                      const TileRead<T>& R, int nc, hipStream_t st) {
  switch (nc) {
    case 1: launch_rt_nc<T, ORDER, 1>(P, g, w, R, st); break;
    case 2: launch_rt_nc<T, ORDER, 2>(P, g, w, R, st); break;
    case 3: launch_rt_nc<T, ORDER, 3>(P, g, w, R, st); break;
    default: launch_rt_nc<T, ORDER, 4>(P, g, w, R, st); break;
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
    if (!meshes[c] || (P.n > 0 && !outs[c]) || strides[c] < 1) return FPM_ERR_ARG;
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

template <typename T>
static int paint_tiles_w_impl(const fpm_particles* part, const fpm_mesh* m, int order, int ncomp,
                              const void* const* weights, const int64_t* wstrides, double scale,
                              void* const* outs, void* ws, int64_t ws_bytes, hipStream_t st) {
  PartDesc<T> P;
  TileGeom g;
  Workspace w;
  int rc = prep<T>(part, m, order, ws, ws_bytes, P, g, w);
  if (rc) return rc;
  if (ncomp < 1 || ncomp > 3 || !weights || !wstrides || !outs) return FPM_ERR_ARG;
  TilePaintW<T> A{};
  for (int c = 0; c < ncomp; ++c) {
    if ((P.n > 0 && !weights[c]) || !outs[c] || wstrides[c] < 1) return FPM_ERR_ARG;
    A.w[c] = static_cast<const T*>(weights[c]);
    A.wstride[c] = wstrides[c];
    A.out[c] = static_cast<T*>(outs[c]);
  }
  A.scale = (T)scale;
  const dim3 grid((unsigned)g.ntiles), blk(kTileThreads);
#define FPM_PW(O_, C_)                                                                        \
  hipLaunchKernelGGL((k_paint_tiles_weighted<T, O_, C_>), grid, blk, 0, st, g, w, A)
  if (order == 2) {
    if (ncomp == 1) FPM_PW(2, 1);
    else if (ncomp == 2) FPM_PW(2, 2);
    else FPM_PW(2, 3);
  } else {
    if (ncomp == 1) FPM_PW(3, 1);
    else if (ncomp == 2) FPM_PW(3, 2);
    else FPM_PW(3, 3);
  }
#undef FPM_PW
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

// Arguments of the stencil force readout (the force factor carries the
// stencil's 1 / (12 h_d)).
template <typename T>
static int grad_setup(const fpm_mesh* m, int order, int64_t n, const void* rho, const void* phi,
                      void* rho_out, int64_t rho_stride, void* f_out, int64_t f_stride,
                      double rho_scale, double f_scale, GradRead<T>& G) {
  if (!phi || (n > 0 && !f_out) || f_stride < 3 || (rho_out && (!rho || rho_stride < 1)))
    return FPM_ERR_ARG;
  // the potential box reaches two cells past the density box on each side
  for (int d = 0; d < 3; ++d)
    if (m->nmesh[d] < TILE + order + 3) return FPM_ERR_UNSUPPORTED;
  G = GradRead<T>{};
  G.rho = static_cast<const T*>(rho);
  G.phi = static_cast<const T*>(phi);
  G.rho_out = static_cast<T*>(rho_out);
  G.f_out = static_cast<T*>(f_out);
  G.rho_stride = rho_stride;
  G.f_stride = f_stride;
  G.rho_scale = (T)rho_scale;
  for (int d = 0; d < 3; ++d)
    G.f_scale[d] = (T)(f_scale * (double)m->nmesh[d] / (12.0 * m->box[d]));
  return FPM_OK;
}

// Whether the stencil readout stages the density box in LDS (FPM_VAR_RHO 0)
// or reads the density through the cache (1).
static int grad_rho_lds(const fpm_mesh* m) {
  return FPM_VAR_FIELD(m->variant, FPM_VAR_RHO) == 1 ? 0 : 1;
}

// Tiles per workgroup of the stencil readout on mesh m (FPM_VAR_SPAN 1:
// two z-adjacent tiles where the tile count along z is even, else one).
static int grad_span(const fpm_mesh* m, const TileGeom& g) {
  return FPM_VAR_FIELD(m->variant, FPM_VAR_SPAN) == 1 && g.nt[2] % 2 == 0 ? 2 : 1;
}

// Threads per tile of the stencil readout on mesh m (FPM_VAR_READOUT).
static int grad_nt(const fpm_mesh* m, int order) {
  switch (FPM_VAR_FIELD(m->variant, FPM_VAR_READOUT)) {
    case 1: return 256;
    case 2: return 512;
    case 3: return 1024;
    default: return order == 2 ? grad_threads<2>() : grad_threads<3>();
  }
}

template <typename T>
static int readout_grad_impl(const fpm_particles* part, const fpm_mesh* m, int order,
                             const void* rho, const void* phi, void* rho_out, int64_t rho_stride,
                             void* f_out, int64_t f_stride, double rho_scale, double f_scale,
                             void* ws, int64_t ws_bytes, hipStream_t st) {
  PartDesc<T> P;
  TileGeom g;
  Workspace w;
  int rc = prep<T>(part, m, order, ws, ws_bytes, P, g, w);
  if (rc) return rc;
  GradRead<T> G{};
  if ((rc = grad_setup<T>(m, order, P.n, rho, phi, rho_out, rho_stride, f_out, f_stride,
                          rho_scale, f_scale, G)))
    return rc;
  if (P.n == 0) return FPM_OK;
  int nt = grad_nt(m, order);
  const int rl = grad_rho_lds(m);
  const int span = grad_span(m, g);
  if (span == 2 && nt < 512) nt = 512;  // pairs are compiled for 512 and 1024 threads
  const dim3 grid((unsigned)(g.ntiles / span));
  switch (((nt * 4 + order) * 2 + rl) * 4 + span) {
#define FPM_GRAD(O_, NT_, R_, S_)                                                               \
  case ((NT_ * 4 + O_) * 2 + R_) * 4 + S_:                                                      \
    hipLaunchKernelGGL((k_readout_grad_tiles<T, O_, NT_, R_, S_>), grid, dim3(NT_), 0, st, P, g, \
                       w, G);                                                                   \
    break;
    FPM_GRAD(2, 256, 1, 1) FPM_GRAD(2, 512, 1, 1) FPM_GRAD(2, 1024, 1, 1)
    FPM_GRAD(3, 256, 1, 1) FPM_GRAD(3, 512, 1, 1) FPM_GRAD(3, 1024, 1, 1)
    FPM_GRAD(2, 256, 0, 1) FPM_GRAD(2, 512, 0, 1) FPM_GRAD(2, 1024, 0, 1)
    FPM_GRAD(3, 256, 0, 1) FPM_GRAD(3, 512, 0, 1) FPM_GRAD(3, 1024, 0, 1)
    FPM_GRAD(2, 512, 1, 2) FPM_GRAD(2, 1024, 1, 2) FPM_GRAD(3, 512, 1, 2) FPM_GRAD(3, 1024, 1, 2)
    FPM_GRAD(2, 512, 0, 2) FPM_GRAD(2, 1024, 0, 2) FPM_GRAD(3, 512, 0, 2) FPM_GRAD(3, 1024, 0, 2)
#undef FPM_GRAD
    default: return FPM_ERR_ARG;
  }
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}


template <typename T>
static int pos_grad_impl(const fpm_particles* part, const fpm_mesh* m, int order, const void* phi,
                         const void* gf, int64_t gf_stride, void* gx, double scale, int accumulate,
                         void* ws, int64_t ws_bytes, hipStream_t st) {
  PartDesc<T> P;
  TileGeom g;
  Workspace w;
  int rc = prep<T>(part, m, order, ws, ws_bytes, P, g, w);
  if (rc) return rc;
  if (!phi || (P.n > 0 && (!gf || !gx)) || gf_stride < 3) return FPM_ERR_ARG;
  for (int d = 0; d < 3; ++d)
    if (m->nmesh[d] < TILE + order + 3) return FPM_ERR_UNSUPPORTED;
  PosGrad<T> G{};
  G.phi = static_cast<const T*>(phi);
  G.gf = static_cast<const T*>(gf);
  G.gx = static_cast<T*>(gx);
  G.gf_stride = gf_stride;
  for (int d = 0; d < 3; ++d) {
    G.dscale[d] = (T)((double)m->nmesh[d] / (12.0 * m->box[d]));
    G.oscale[d] = (T)(scale * (double)m->nmesh[d] / m->box[d]);
  }
  G.accumulate = accumulate ? 1 : 0;
  if (P.n == 0) return FPM_OK;
  const dim3 grid((unsigned)g.ntiles), blk(kTileThreads);
  if (order == 2)
    hipLaunchKernelGGL((k_pos_grad_tiles<T, 2>), grid, blk, 0, st, g, w, G);
  else
    hipLaunchKernelGGL((k_pos_grad_tiles<T, 3>), grid, blk, 0, st, g, w, G);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}
// ---------------------------------------------------------------- batched boxes
// The binning, paint and stencil readout of many boxes, one launch per
// compiled kernel: each box's arguments (as its own calls would set them up)
// sit in a device table, and the workgroups of a launch stride the boxes'
// items (count-array chunks, particle spans, scan blocks or tiles), one
// workgroup each, through the bodies above.
template <typename T>
struct TileBox {
  PartDesc<T> P;
  TileGeom g;
  Workspace w;
  BinOrder o;
  T* rho;
  float mass_fs;
  int sh;
  double inv_fx;
  GradRead<T> G;
  // the step boundary before the force, applied by the count pass: rows of
  // kick / drift coefficients as fpm_kick_drift_table reads them
  T* S;
  T* Pm;
  const T* F;
  const double* table;
  int rows, stride, close, open, drift;
};

enum { kTbClear = 1, kTbCount, kTbScanLocal, kTbScanAdd, kTbScatter, kTbPaint, kTbReadout };

template <typename T>
__global__ __launch_bounds__(256) void k_batch_bin_clear(const TileBox<T>* __restrict__ tab,
                                                         const int64_t* __restrict__ pre,
                                                         const int64_t* __restrict__ cnt, int nb,
                                                         int64_t total) {
  const int64_t it = blockIdx.x;
  if (it >= total) return;
  const int b = batch_box(pre, nb, it);
  const int64_t j = it - pre[b];
  if (j >= cnt[b]) return;
  const TileBox<T>& x = tab[b];
  const int64_t i = j * 256 + threadIdx.x;
  if (i < x.g.nbins) x.w.count[i] = 0;
  if (j == 0 && threadIdx.x == 0) {
    x.w.misplaced[0] = 0;
    x.w.ovf[0] = 0;
  }
}

template <typename T, int ORDER>
__global__ __launch_bounds__(kBinThreads) void k_batch_bin_count(const TileBox<T>* __restrict__ tab, const int64_t* __restrict__ pre, const int64_t* __restrict__ cnt, int nb, int64_t total, const int32_t* __restrict__ counter) {
  const int64_t it = blockIdx.x;
  if (it >= total) return;
  const int b = batch_box(pre, nb, it);
  const int64_t j = it - pre[b];
  if (j >= cnt[b]) return;
  const TileBox<T>& x = tab[b];
  // step row counter[0]: kicks by the previous row's closing and this row's
  // opening coefficient (only the opening one on row 0), then the drift, as
  // fpm_kick_drift_table; a row outside [0, rows) leaves S and P unchanged
  KickDrift<T> kd{};
  const int32_t row = counter ? counter[0] : -1;
  if (row >= 0 && row < x.rows) {
    kd.S = x.S;
    kd.P = x.Pm;
    kd.F = x.F;
    const T open = (T)x.table[(int64_t)row * x.stride + x.open];
    if (row > 0) {
      kd.k0 = (T)x.table[(int64_t)(row - 1) * x.stride + x.close];
      kd.k1 = open;
      kd.nkick = 2;
    } else {
      kd.k0 = open;
      kd.nkick = 1;
    }
    kd.d = (T)x.table[(int64_t)row * x.stride + x.drift];
  }
  bin_count_body<T, ORDER>(x.P, x.g, x.w, x.o, j, cnt[b], kd);
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_batch_scan_local(const TileBox<T>* __restrict__ tab, const int64_t* __restrict__ pre, const int64_t* __restrict__ cnt, int nb, int64_t total) {
  const int64_t it = blockIdx.x;
  if (it >= total) return;
  const int b = batch_box(pre, nb, it);
  const int64_t j = it - pre[b];
  if (j >= cnt[b]) return;
  scan_local_body(tab[b].w, tab[b].g.nbins, j);
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_batch_scan_add(const TileBox<T>* __restrict__ tab, const int64_t* __restrict__ pre, const int64_t* __restrict__ cnt, int nb, int64_t total) {
  const int64_t it = blockIdx.x;
  if (it >= total) return;
  const int b = batch_box(pre, nb, it);
  const int64_t j = it - pre[b];
  if (j >= cnt[b]) return;
  scan_add_body(tab[b].w, tab[b].g.nbins, j, cnt[b]);
}

template <typename T, int ORDER>
__global__ __launch_bounds__(kBinThreads) void k_batch_bin_scatter(const TileBox<T>* __restrict__ tab, const int64_t* __restrict__ pre, const int64_t* __restrict__ cnt, int nb, int64_t total) {
  const int64_t it = blockIdx.x;
  if (it >= total) return;
  const int b = batch_box(pre, nb, it);
  const int64_t j = it - pre[b];
  if (j >= cnt[b]) return;
  const TileBox<T>& x = tab[b];
  bin_scatter_body<T, ORDER>(x.P, x.g, x.w, x.o, j, cnt[b]);
}

template <typename T, int ORDER, int NT>
__global__ __launch_bounds__(NT) void k_batch_paint(const TileBox<T>* __restrict__ tab, const int64_t* __restrict__ pre, const int64_t* __restrict__ cnt, int nb, int64_t total) {
  const int64_t it = blockIdx.x;
  if (it >= total) return;
  const int b = batch_box(pre, nb, it);
  const int64_t j = it - pre[b];
  if (j >= cnt[b]) return;
  const TileBox<T>& x = tab[b];
  paint_tiles_body<T, ORDER, NT>(x.P, x.g, x.w, x.rho, x.mass_fs, x.sh, x.inv_fx, j, cnt[b]);
}

template <typename T, int ORDER, int NT, bool RLDS, int SPAN>
__global__ __launch_bounds__(NT) void k_batch_readout(const TileBox<T>* __restrict__ tab, const int64_t* __restrict__ pre, const int64_t* __restrict__ cnt, int nb, int64_t total) {
  const int64_t it = blockIdx.x;
  if (it >= total) return;
  const int b = batch_box(pre, nb, it);
  const int64_t j = it - pre[b];
  if (j >= cnt[b]) return;
  const TileBox<T>& x = tab[b];
  readout_grad_body<T, ORDER, NT, RLDS, SPAN>(x.P, x.g, x.w, x.G, j, cnt[b]);
}

int tile_batch_build(int nbox, const fpm_batch_box* boxes, DevArena* arena,
                     std::vector<BatchLaunch>& bin, std::vector<BatchLaunch>& paint,
                     std::vector<BatchLaunch>& readout) {
  std::vector<TileBox<float>> args(nbox);
  std::vector<int> order(nbox), pnt(nbox), rnt(nbox), rlds(nbox), rspan(nbox);
  for (int i = 0; i < nbox; ++i) {
    const fpm_batch_box& bx = boxes[i];
    TileBox<float>& x = args[i];
    x = TileBox<float>{};
    int rc = prep<float>(&bx.part, &bx.mesh, bx.order, bx.ws, bx.ws_bytes, x.P, x.g, x.w);
    if (rc) return rc;
    if (!bx.rho) return FPM_ERR_ARG;
    if (x.P.n > 0 && (!bx.part.s || !bx.P || !bx.force_out || bx.force_stride != 3 ||
                      !bx.table || bx.rows < 1 || bx.stride < 1))
      return FPM_ERR_ARG;
    for (int k : {bx.close, bx.open, bx.drift})
      if (x.P.n > 0 && (k < 0 || k >= bx.stride)) return FPM_ERR_ARG;
    x.S = static_cast<float*>(const_cast<void*>(bx.part.s));
    x.Pm = static_cast<float*>(bx.P);
    x.F = static_cast<const float*>(bx.force_out);
    x.table = bx.table;
    x.rows = bx.rows;
    x.stride = bx.stride;
    x.close = bx.close;
    x.open = bx.open;
    x.drift = bx.drift;
    // the plan owns the box's workspace from here: its first step places
    // the records exactly (no layout header yet), the later ones rebin
    x.w.allow_cap = 1;
    if (arena->base && hipMemset(x.w.hdr, 0, 4 * sizeof(int64_t)) != hipSuccess)
      return FPM_ERR_HIP;
    x.o = bin_order(&bx.part, &bx.mesh);
    x.rho = static_cast<float*>(bx.rho);
    if ((rc = paint_scale(x.P.n, bx.mass, x.mass_fs, x.sh, x.inv_fx))) return rc;
    if ((rc = grad_setup<float>(&bx.mesh, bx.order, x.P.n, bx.rho, bx.phi, bx.rho_out,
                                bx.rho_stride, bx.force_out, bx.force_stride, bx.rho_scale,
                                bx.force_scale, x.G)))
      return rc;
    order[i] = bx.order;
    pnt[i] = paint_nt(&bx.mesh);
    rnt[i] = grad_nt(&bx.mesh, bx.order);
    rlds[i] = grad_rho_lds(&bx.mesh);
    rspan[i] = grad_span(&bx.mesh, x.g);
    if (rspan[i] == 2 && rnt[i] < 512) rnt[i] = 512;  // as in the box's own launch
  }
  // keys carry the box index through the grouping lambdas
  auto idx = [&](const TileBox<float>& x) { return (int)(&x - args.data()); };
  const int64_t span = (int64_t)kBinThreads * kBinILP;
  int rc;
  auto one = [&](const TileBox<float>&) { return BatchKey{0, 0, 0, 0}; };
  auto by_order = [&](const TileBox<float>& x) { return BatchKey{order[idx(x)], 0, 0, 0}; };
  if ((rc = batch_groups(args, kTbClear, one,
                         [&](const TileBox<float>& x) { return (x.g.nbins + 255) / 256; }, arena,
                         bin)))
    return rc;
  auto spans = [&](const TileBox<float>& x) { return (x.P.n + span - 1) / span; };
  if ((rc = batch_groups(args, kTbCount, by_order, spans, arena, bin))) return rc;
  auto sblk = [&](const TileBox<float>& x) { return scan_blocks(x.g); };
  if ((rc = batch_groups(args, kTbScanLocal, one, sblk, arena, bin))) return rc;
  if ((rc = batch_groups(args, kTbScanAdd, one, sblk, arena, bin))) return rc;
  if ((rc = batch_groups(args, kTbScatter, by_order, spans, arena, bin))) return rc;
  if ((rc = batch_groups(
           args, kTbPaint,
           [&](const TileBox<float>& x) { return BatchKey{order[idx(x)], pnt[idx(x)], 0, 0}; },
           [&](const TileBox<float>& x) { return x.g.ntiles; }, arena, paint)))
    return rc;
  if ((rc = batch_groups(
           args, kTbReadout,
           [&](const TileBox<float>& x) {
             return BatchKey{order[idx(x)], rnt[idx(x)], rlds[idx(x)], rspan[idx(x)]};
           },
           [&](const TileBox<float>& x) {
             return x.P.n > 0 ? x.g.ntiles / rspan[idx(x)] : (int64_t)0;
           },
           arena,
           readout)))
    return rc;
  return FPM_OK;
}

int tile_batch_run(const std::vector<BatchLaunch>& ls, hipStream_t st, const int32_t* counter) {
  for (const BatchLaunch& L : ls) {
    const auto* tab = static_cast<const TileBox<float>*>(L.tab);
    const dim3 grid(batch_grid(L.total));
    switch (L.kind) {
      case kTbClear:
        hipLaunchKernelGGL(k_batch_bin_clear<float>, grid, dim3(256), 0, st, tab, L.pre, L.cnt,
                           L.nb, L.total);
        break;
      case kTbCount:
        if (L.a == 2)
          hipLaunchKernelGGL((k_batch_bin_count<float, 2>), grid, dim3(kBinThreads), 0, st, tab,
                             L.pre, L.cnt, L.nb, L.total, counter);
        else
          hipLaunchKernelGGL((k_batch_bin_count<float, 3>), grid, dim3(kBinThreads), 0, st, tab,
                             L.pre, L.cnt, L.nb, L.total, counter);
        break;
      case kTbScanLocal:
        hipLaunchKernelGGL(k_batch_scan_local<float>, grid, dim3(kScanThreads), 0, st, tab, L.pre,
                           L.cnt, L.nb, L.total);
        break;
      case kTbScanAdd:
        hipLaunchKernelGGL(k_batch_scan_add<float>, grid, dim3(kScanThreads), 0, st, tab, L.pre,
                           L.cnt, L.nb, L.total);
        break;
      case kTbScatter:
        if (L.a == 2)
          hipLaunchKernelGGL((k_batch_bin_scatter<float, 2>), grid, dim3(kBinThreads), 0, st, tab,
                             L.pre, L.cnt, L.nb, L.total);
        else
          hipLaunchKernelGGL((k_batch_bin_scatter<float, 3>), grid, dim3(kBinThreads), 0, st, tab,
                             L.pre, L.cnt, L.nb, L.total);
        break;
      case kTbPaint:
        switch (L.b * 4 + L.a) {
#define FPM_BPAINT(O_, NT_)                                                                    \
  case NT_ * 4 + O_:                                                                           \
    hipLaunchKernelGGL((k_batch_paint<float, O_, NT_>), grid, dim3(NT_), 0, st, tab, L.pre, L.cnt, \
                       L.nb, L.total);                                                               \
    break;
          FPM_BPAINT(2, 256) FPM_BPAINT(3, 256) FPM_BPAINT(2, 512) FPM_BPAINT(3, 512)
          FPM_BPAINT(2, 1024) FPM_BPAINT(3, 1024)
#undef FPM_BPAINT
          default: return FPM_ERR_ARG;
        }
        break;
      case kTbReadout:
        switch (((L.b * 4 + L.a) * 2 + L.c) * 4 + L.d) {
#define FPM_BREAD(O_, NT_, R_, S_)                                                             \
  case ((NT_ * 4 + O_) * 2 + R_) * 4 + S_:                                                     \
    hipLaunchKernelGGL((k_batch_readout<float, O_, NT_, R_, S_>), grid, dim3(NT_), 0, st, tab,   \
                       L.pre, L.cnt, L.nb, L.total);                                           \
    break;
          FPM_BREAD(2, 256, 1, 1) FPM_BREAD(3, 256, 1, 1) FPM_BREAD(2, 512, 1, 1)
          FPM_BREAD(3, 512, 1, 1) FPM_BREAD(2, 1024, 1, 1) FPM_BREAD(3, 1024, 1, 1)
          FPM_BREAD(2, 256, 0, 1) FPM_BREAD(3, 256, 0, 1) FPM_BREAD(2, 512, 0, 1)
          FPM_BREAD(3, 512, 0, 1) FPM_BREAD(2, 1024, 0, 1) FPM_BREAD(3, 1024, 0, 1)
          FPM_BREAD(2, 512, 1, 2) FPM_BREAD(3, 512, 1, 2) FPM_BREAD(2, 1024, 1, 2)
          FPM_BREAD(3, 1024, 1, 2) FPM_BREAD(2, 512, 0, 2) FPM_BREAD(3, 512, 0, 2)
          FPM_BREAD(2, 1024, 0, 2) FPM_BREAD(3, 1024, 0, 2)
#undef FPM_BREAD
          default: return FPM_ERR_ARG;
        }
        break;
      default: return FPM_ERR_ARG;
    }
    FPM_LAUNCH_CHECK();
  }
  return FPM_OK;
}

}  // namespace fpm

using namespace fpm;

extern "C" {

int fpm_tiles_supported(const fpm_mesh* mesh) { return tiles_ok(mesh) ? 1 : 0; }

int64_t fpm_tile_workspace_size(const fpm_mesh* mesh, int order, int64_t n) {
  if (!tiles_ok(mesh) || (order != 2 && order != 3) || n < 0) return -1;
  return layout_ws(tile_geom(mesh), n, nullptr, nullptr);
}

int64_t fpm_tile_misplaced_offset(const fpm_mesh* mesh, int order, int64_t n) {
  if (!tiles_ok(mesh) || (order != 2 && order != 3) || n < 0) return -1;
  int64_t off[10];
  ws_offsets(tile_geom(mesh), n, off);
  return off[5];
}

int fpm_bin_particles_f32(const fpm_particles* part, const fpm_mesh* mesh, int order, void* ws,
                          int64_t ws_bytes, fpm_stream_t s) {
  return bin_impl<float>(part, mesh, order, ws, ws_bytes, (hipStream_t)s, false);
}
int fpm_rebin_particles_f32(const fpm_particles* part, const fpm_mesh* mesh, int order, void* ws,
                            int64_t ws_bytes, fpm_stream_t s) {
  return bin_impl<float>(part, mesh, order, ws, ws_bytes, (hipStream_t)s, true);
}
int fpm_kick_drift_rebin_f32(const fpm_particles* part, const fpm_mesh* mesh, int order, void* P,
                             const void* F, const double* kicks, int nkick, double drift,
                             void* ws, int64_t ws_bytes, fpm_stream_t s) {
  if (!part || !part->s || !P || nkick < 0 || nkick > 2 || (nkick > 0 && (!kicks || !F)))
    return FPM_ERR_ARG;
  KickDrift<float> kd{};
  kd.S = static_cast<float*>(const_cast<void*>(part->s));
  kd.P = static_cast<float*>(P);
  kd.F = static_cast<const float*>(F);
  kd.k0 = nkick > 0 ? (float)kicks[0] : 0.0f;
  kd.k1 = nkick > 1 ? (float)kicks[1] : 0.0f;
  kd.d = (float)drift;
  kd.nkick = nkick;
  return bin_impl<float>(part, mesh, order, ws, ws_bytes, (hipStream_t)s, true, kd);
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

int fpm_paint_tiled_weighted_f32(const fpm_particles* part, const fpm_mesh* mesh, int order,
                                 int ncomp, const void* const* weights, const int64_t* wstrides,
                                 double scale, void* const* outs, void* ws, int64_t ws_bytes,
                                 fpm_stream_t s) {
  return paint_tiles_w_impl<float>(part, mesh, order, ncomp, weights, wstrides, scale, outs, ws,
                                   ws_bytes, (hipStream_t)s);
}

int fpm_readout_tiled_grad4_f32(const fpm_particles* part, const fpm_mesh* mesh, int order,
                                const void* rho, const void* phi, void* rho_out,
                                int64_t rho_stride, void* force_out, int64_t force_stride,
                                double rho_scale, double force_scale, void* ws, int64_t ws_bytes,
                                fpm_stream_t s) {
  return readout_grad_impl<float>(part, mesh, order, rho, phi, rho_out, rho_stride, force_out,
                                  force_stride, rho_scale, force_scale, ws, ws_bytes,
                                  (hipStream_t)s);
}

int fpm_window_grad_tiled_grad4_f32(const fpm_particles* part, const fpm_mesh* mesh, int order,
                                    const void* phi, const void* gf, int64_t gf_stride, void* gx,
                                    double scale, int accumulate, void* ws, int64_t ws_bytes,
                                    fpm_stream_t s) {
  return pos_grad_impl<float>(part, mesh, order, phi, gf, gf_stride, gx, scale, accumulate, ws,
                              ws_bytes, (hipStream_t)s);
}

}  // extern "C"
