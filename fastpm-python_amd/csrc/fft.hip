// 3D real <-> half-complex FFT for gfx950: Stockham autosort passes in LDS.
//
// A transform is three passes, one per axis, each reading and writing the mesh
// once:
//   forward : z (real -> half complex, packed as a length n2/2 complex FFT),
//             y (complex, in place), x (complex, in place, times 1/N).
//   inverse : x (complex, reads the input once and applies up to six k-space
//             multipliers, writing one scratch mesh per output), y (in place),
//             z (half complex -> real).
// The strided y/x passes work on tiles of BZ adjacent z columns so each global
// row segment is 128 bytes; the tile (n x BZ complex) lives in LDS and every
// Stockham stage reads its radix-R inputs to registers, synchronises, and writes
// the outputs back (in-place autosort).  Stage indexing follows the Stockham
// formulation of Govindaraju et al., "High performance discrete Fourier
// transforms on graphics processors" (SC08): input j + r N/R, twiddle
// w_N^{r (j mod Ns) N/(Ns R)}, output (j/Ns) Ns R + j mod Ns + r Ns.
// Twiddles come from device tables computed in double on the host.
#include "batch.h"

#include <cmath>
#include <vector>

namespace fpm {

// ---------------------------------------------------------------- radix plans
template <int N> struct Plan;
#define FPM_PLAN(N_, NS_, ...)                       \
  template <> struct Plan<N_> {                      \
    static constexpr int ns = NS_;                   \
    static constexpr int r[6] = {__VA_ARGS__};       \
  };
FPM_PLAN(2, 1, 2)
FPM_PLAN(4, 1, 4)
FPM_PLAN(8, 1, 8)
FPM_PLAN(16, 2, 4, 4)
FPM_PLAN(32, 2, 8, 4)
FPM_PLAN(64, 2, 8, 8)
FPM_PLAN(128, 3, 8, 4, 4)
FPM_PLAN(256, 3, 8, 8, 4)
FPM_PLAN(512, 3, 8, 8, 8)
FPM_PLAN(1024, 4, 8, 8, 4, 4)
FPM_PLAN(3, 1, 3)
FPM_PLAN(6, 2, 3, 2)
FPM_PLAN(12, 2, 3, 4)
FPM_PLAN(24, 2, 3, 8)
FPM_PLAN(48, 3, 3, 4, 4)
FPM_PLAN(96, 3, 3, 8, 4)
FPM_PLAN(192, 3, 3, 8, 8)
FPM_PLAN(384, 4, 3, 8, 4, 4)
FPM_PLAN(768, 4, 3, 8, 8, 4)
#undef FPM_PLAN

// ---------------------------------------------------------------- butterflies
// Multiply by -i (forward) or +i (inverse).
template <bool INV, typename V> __device__ __forceinline__ V rot90(V a) {
  return INV ? cmk<V>(-a.y, a.x) : cmk<V>(a.y, -a.x);
}

template <int R, bool INV, typename V> struct Dft;

template <bool INV, typename V> struct Dft<2, INV, V> {
  __device__ __forceinline__ static void run(V* a) {
    V t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};

template <bool INV, typename V> struct Dft<4, INV, V> {
  __device__ __forceinline__ static void run(V* a) {
    V t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    V t2 = cadd(a[1], a[3]), t3 = rot90<INV>(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};

template <bool INV, typename V> struct Dft<8, INV, V> {
  __device__ __forceinline__ static void run(V* a) {
    using T = decltype(V::x);
    const T h = (T)0.70710678118654752440084436210485;
    V e[4] = {a[0], a[2], a[4], a[6]};
    V o[4] = {a[1], a[3], a[5], a[7]};
    Dft<4, INV, V>::run(e);
    Dft<4, INV, V>::run(o);
    // o1 *= w8, o2 *= w8^2, o3 *= w8^3 with w8 = exp(-+ i pi/4)
    V w1 = INV ? cmk<V>((o[1].x - o[1].y) * h, (o[1].x + o[1].y) * h)
               : cmk<V>((o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h);
    V w2 = rot90<INV>(o[2]);
    V w3i = INV ? cmk<V>((o[3].x - o[3].y) * h, (o[3].x + o[3].y) * h)
                : cmk<V>((o[3].x + o[3].y) * h, (o[3].y - o[3].x) * h);
    V w3 = rot90<INV>(w3i);
    a[0] = cadd(e[0], o[0]);
    a[4] = csub(e[0], o[0]);
    a[1] = cadd(e[1], w1);
    a[5] = csub(e[1], w1);
    a[2] = cadd(e[2], w2);
    a[6] = csub(e[2], w2);
    a[3] = cadd(e[3], w3);
    a[7] = csub(e[3], w3);
  }
};

template <bool INV, typename V> struct Dft<3, INV, V> {
  __device__ __forceinline__ static void run(V* a) {
    using T = decltype(V::x);
    const T s3 = (T)0.86602540378443864676372317075294;
    V t = cadd(a[1], a[2]);
    V s = csub(a[1], a[2]);
    V m = cmk<V>(a[0].x - (T)0.5 * t.x, a[0].y - (T)0.5 * t.y);
    V u = rot90<INV>(cscale(s, s3));
    a[0] = cadd(a[0], t);
    a[1] = cadd(m, u);
    a[2] = csub(m, u);
  }
};

// ---------------------------------------------------------------- one stage
// NL lines of length N in LDS.  Element e of line l lives at
// l*LS + pad(e)*ES with pad(e) = e + (e >> PAD) when PAD > 0: one spare slot
// every 2^PAD elements spreads the stride-R stores of the early stages over
// the LDS banks.  LFAST selects which index varies fastest across threads
// (line index for the column tiles, whose lines are interleaved; element
// index for row tiles).  TS is the stride into the twiddle table (2 when the
// table is for length 2N).
template <int PAD> __device__ __forceinline__ constexpr int padx(int e) {
  if constexpr (PAD > 0) return e + (e >> PAD);
  else return e;
}

template <typename T, int N, int R, int NS, int NL, int NT, bool LFAST, int ES, int LS, int TS,
          bool INV, int PAD = 0>
__device__ __forceinline__ void stage(C<T>* s, const C<T>* tw) {
  using V = C<T>;
  constexpr int NB = NL * (N / R);
  constexpr int PER = (NB + NT - 1) / NT;
  V v[PER][R];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = threadIdx.x + q * NT;
    if ((NB % NT == 0) || b < NB) {
      int l, j;
      if (LFAST) { l = b % NL; j = b / NL; } else { j = b % (N / R); l = b / (N / R); }
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = s[l * LS + padx<PAD>(j + r * (N / R)) * ES];
      if (NS > 1) {
        const int k = j % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          V w = tw[(r * k * (N / (NS * R))) * TS];
          if (INV) w.y = -w.y;
          v[q][r] = cmul(v[q][r], w);
        }
      }
      Dft<R, INV, V>::run(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = threadIdx.x + q * NT;
    if ((NB % NT == 0) || b < NB) {
      int l, j;
      if (LFAST) { l = b % NL; j = b / NL; } else { j = b % (N / R); l = b / (N / R); }
      const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
      for (int r = 0; r < R; ++r) s[l * LS + padx<PAD>(base + r * NS) * ES] = v[q][r];
    }
  }
  __syncthreads();
}

// Stages S..SEND-1 of the plan (LDS to LDS).
template <typename T, int N, int S, int SEND, int NS, int NL, int NT, bool LFAST, int ES, int LS,
          int TS, bool INV, int PAD = 0>
struct StagesTo {
  __device__ __forceinline__ static void run(C<T>* s, const C<T>* tw) {
    if constexpr (S < SEND) {
      constexpr int R = Plan<N>::r[S];
      stage<T, N, R, NS, NL, NT, LFAST, ES, LS, TS, INV, PAD>(s, tw);
      StagesTo<T, N, S + 1, SEND, NS * R, NL, NT, LFAST, ES, LS, TS, INV, PAD>::run(s, tw);
    }
  }
};

constexpr int round64(int x) { return ((x + 63) / 64) * 64; }
constexpr int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

#ifndef FPM_COL_BYTES
#define FPM_COL_BYTES 64  // bytes of each global row segment of a column tile
#endif
#ifndef FPM_EPT_F32
#define FPM_EPT_F32 8  // complex elements per thread in a column tile (float)
#endif
#ifndef FPM_EPT_F64
#define FPM_EPT_F64 8  // (double)
#endif
// Tile shapes chosen by a sweep of row-segment width (64/128/256 bytes) and
// elements per thread (4/8/16) at 512^3 on gfx950.

// Threads for a tile of E complex elements of type T.
template <typename T> constexpr int threads_for(int E) {
  return clampi(round64(E / (sizeof(T) == 4 ? FPM_EPT_F32 : FPM_EPT_F64)), 64, 1024);
}

constexpr int pow2floor(int x) {
  int p = 1;
  while (2 * p <= x) p *= 2;
  return p;
}

// Column tile width (z columns per tile): row segments of FPM_COL_BYTES, or
// twice that when WIDE, capped so the tile stays within 96 KiB of LDS (144 KiB
// for a single-output pass when WIDE: one workgroup per CU, so that
// 1024-point columns still get 128-byte segments; the rest holds twiddles
// and multiplier tables, six outputs' worth for MC).  Both
// widths are compiled; the one used per mesh is chosen at run time (the
// mesh's variant or fpm_fft_set_tiling, set by the autotuner).  A column's
// arithmetic does not depend on the width (up to the compiler's multiply-add
// contraction in each instantiation).
template <typename T, int N, int WIDE = 0, bool MC = false> struct ColW {
  static constexpr int a = (FPM_COL_BYTES << WIDE) / (int)sizeof(C<T>);
  static constexpr int b =
      pow2floor(((WIDE && !MC ? 144 : 96) * 1024 / (int)sizeof(C<T>)) / N);
  static constexpr int BZ = a < b ? a : b;
};

// Run-time tiling table: wide column tiles per (precision, axis length).
static int g_col_wide[2][1025];

static int col_wide(int elem_bytes, int n) {
  return (n >= 0 && n <= 1024) ? g_col_wide[elem_bytes == 8 ? 1 : 0][n] : 0;
}

// The tiling of a strided pass over an axis of length n of mesh m: the mesh's
// FPM_VAR_FFT_COLS field, or the process-wide table when that is 0.
static int col_choice(const fpm_mesh* m, int elem_bytes, int n) {
  const int v = FPM_VAR_FIELD(m->variant, FPM_VAR_FFT_COLS);
  return v == 1 ? 0 : (v == 2 ? 1 : col_wide(elem_bytes, n));
}

#ifndef FPM_Z_ELEMS
#define FPM_Z_ELEMS 2048  // complex elements per z-pass workgroup
#endif
#ifndef FPM_Z_EPT
#define FPM_Z_EPT 8  // complex elements per thread in the z pass
#endif

// Lines per workgroup for the z pass: about FPM_Z_ELEMS complex elements
// (times 2^ZV / 2 for the FPM_VAR_Z_LINES choices 1 and 2: half and double).
constexpr int z_elems(int ZV) { return ZV == 1 ? FPM_Z_ELEMS / 2 : (ZV == 2 ? 2 * FPM_Z_ELEMS : FPM_Z_ELEMS); }
constexpr int z_lines(int M, int ZV = 0) {
  int nl = 1;
  while (nl * 2 * M <= z_elems(ZV)) nl *= 2;
  return nl;
}

// ---------------------------------------------------------------- k-space multipliers
struct KopPack {
  int ncomp;
  int laplace[6];
  int ngrad[6];
  int gdir[6][2];
  double factor[6];
  double r2[6];          // r_split^2
  int grad_order[6];
  int deconv[6];
  int nm[3];
  double box[3];
};

// Per-axis factors for mode index i on an axis of n points and side L.
// Returns k^2, the gradient factor for both gradient orders and sinc(k h/2).
__device__ __forceinline__ void axis_factors(int i, int n, double L, bool half, double& kk,
                                             double& g0, double& g1, double& sn) {
  const int nn = half ? i : (i < (n + 1) / 2 ? i : i - n);
  const bool nyq = (n % 2 == 0) && (i == n / 2);
  const double two_pi = 6.283185307179586476925286766559;
  const double k = two_pi * (double)nn / L;
  const double h = L / (double)n;
  kk = k * k;
  g0 = nyq ? 0.0 : k;
  // 4-point finite difference: (8 sin w - sin 2w) / (6 h), w = 2 pi nn / n.
  const double x = 2.0 * (double)nn / (double)n;
  g1 = nyq ? 0.0 : (8.0 * sinpi(x) - sinpi(2.0 * x)) / (6.0 * h);
  const double hx = (double)nn / (double)n;  // w/2 over pi
  sn = (nn == 0) ? 1.0 : sinpi(hx) / (3.14159265358979323846264338327950 * hx);
}

template <typename T>
struct AxisTab {
  T kk, g0, g1, sn;
};

template <typename T>
__device__ __forceinline__ AxisTab<T> axis_tab(int i, int n, double L, bool half) {
  double kk, g0, g1, sn;
  axis_factors(i, n, L, half, kk, g0, g1, sn);
  return AxisTab<T>{(T)kk, (T)g0, (T)g1, (T)sn};
}

// Multiplier for component c; ax/ay/az are the axis factors of the mode.
// Returns (m, ngrad) with the full multiplier = m * i^ngrad.
template <typename T>
__device__ __forceinline__ T kop_mult(const KopPack& K, int c, const AxisTab<T>& ax,
                                      const AxisTab<T>& ay, const AxisTab<T>& az) {
  T m = (T)K.factor[c];
  const T kk = ax.kk + ay.kk + az.kk;
  if (K.laplace[c]) m = (kk == (T)0) ? (T)0 : m / kk;
  if (K.r2[c] != 0.0) m *= exp(-kk * (T)K.r2[c]);
  const bool o0 = K.grad_order[c] == 0;
  for (int g = 0; g < K.ngrad[c]; ++g) {
    const int d = K.gdir[c][g];
    const T gx = o0 ? ax.g0 : ax.g1, gy = o0 ? ay.g0 : ay.g1, gz = o0 ? az.g0 : az.g1;
    m *= d == 0 ? gx : (d == 1 ? gy : gz);
  }
  const int p = K.deconv[c];
  if (p) {
    const T w = ax.sn * ay.sn * az.sn;
    T wp = w;
    for (int q = 1; q < p; ++q) wp *= w;
    m /= wp;
  }
  return m;
}

// The axis-separable part of component c's multiplier along one mesh axis at
// mode index i: its gradient factors along this axis and the window
// compensation 1/sinc(k h/2)^p of this axis.  Also returns k^2 of the axis.
__device__ __forceinline__ double comp_axis_factor(const KopPack& K, int c, int axis, int i,
                                                   bool half, double& kk) {
  double g0, g1, sn;
  axis_factors(i, K.nm[axis], K.box[axis], half, kk, g0, g1, sn);
  double f = 1.0;
  for (int g = 0; g < K.ngrad[c]; ++g)
    if (K.gdir[c][g] == axis) f *= K.grad_order[c] == 0 ? g0 : g1;
  for (int q = 0; q < K.deconv[c]; ++q) f /= sn;
  return f;
}

// The same axis factors in T precision, read off the axis' twiddle table
// tw[m] = exp(-2 pi i m / n) (m < n): sin w and sin 2w of w = 2 pi nn / n are
// table entries, and sin(w / 2) follows from the half-angle formulas
// sin(t/2) = sin t / sqrt(2 (1 + cos t)) for cos t >= 0, sqrt((1 - cos t)/2)
// otherwise (no cancellation on either branch).
template <typename T>
__device__ __forceinline__ void axis_factors_tw(int i, int n, T two_pi_l, T inv6h, bool half,
                                                const C<T>* tw, T& kk, T& g0, T& g1, T& sn) {
  const int nn = half ? i : (i < (n + 1) / 2 ? i : i - n);
  const bool nyq = (n % 2 == 0) && (i == n / 2);
  const T k = two_pi_l * (T)nn;
  kk = k * k;
  g0 = nyq ? (T)0 : k;
  const T s1 = -tw[i % n].y, s2 = -tw[(2 * i) % n].y;
  g1 = nyq ? (T)0 : ((T)8 * s1 - s2) * inv6h;
  const int a = nn < 0 ? -nn : nn;
  const C<T> t = tw[a % n];
  const T c = t.x, s = -t.y;
  const T sh = c >= (T)0 ? s / sqrt((T)2 * ((T)1 + c)) : sqrt(((T)1 - c) * (T)0.5);
  sn = a == 0 ? (T)1
              : sh * ((T)n / ((T)3.14159265358979323846264338327950 * (T)a));
}

// The same from the twiddle value t = exp(-2 pi i m / n) of this mode's table
// entry m = i (mod n): sin 2w = 2 sin w cos w, and |sin w| with cos w give
// sin(|w| / 2) by the half-angle formulas.
template <typename T>
__device__ __forceinline__ void axis_factors_tv(int i, int n, C<T> t, T two_pi_l, T inv6h,
                                                bool half, T& kk, T& g0, T& g1, T& sn) {
  const int nn = half ? i : (i < (n + 1) / 2 ? i : i - n);
  const bool nyq = (n % 2 == 0) && (i == n / 2);
  const T k = two_pi_l * (T)nn;
  kk = k * k;
  g0 = nyq ? (T)0 : k;
  const T c = t.x, s1 = -t.y;
  g1 = nyq ? (T)0 : ((T)8 * s1 - (T)2 * s1 * c) * inv6h;
  const int a = nn < 0 ? -nn : nn;
  const T sa = s1 < (T)0 ? -s1 : s1;
  const T sh = c >= (T)0 ? sa / sqrt((T)2 * ((T)1 + c)) : sqrt(((T)1 - c) * (T)0.5);
  sn = a == 0 ? (T)1
              : sh * ((T)n / ((T)3.14159265358979323846264338327950 * (T)a));
}

// Component c's separable factor along `axis` from that axis' factors.
template <typename T>
__device__ __forceinline__ T comp_factor_t(const KopPack& K, int c, int axis, T g0, T g1, T sn) {
  T f = (T)1;
  for (int g = 0; g < K.ngrad[c]; ++g)
    if (K.gdir[c][g] == axis) f *= K.grad_order[c] == 0 ? g0 : g1;
  for (int q = 0; q < K.deconv[c]; ++q) f /= sn;
  return f;
}

template <typename T> __device__ __forceinline__ T fast_rcp(T x);
template <> __device__ __forceinline__ float fast_rcp<float>(float x) {
  return __builtin_amdgcn_rcpf(x);
}
template <> __device__ __forceinline__ double fast_rcp<double>(double x) { return 1.0 / x; }

template <typename V>
__device__ __forceinline__ V apply_mult(V v, decltype(V::x) m, int ngrad) {
  switch (ngrad & 3) {
    case 0: return cmk<V>(v.x * m, v.y * m);
    case 1: return cmk<V>(-v.y * m, v.x * m);
    case 2: return cmk<V>(-v.x * m, -v.y * m);
    default: return cmk<V>(v.y * m, -v.x * m);
  }
}

// ---------------------------------------------------------------- column pass
// Transform along a strided axis of a complex mesh.  The tile is the N points
// of the axis times BZ adjacent z columns; `outer` enumerates the remaining
// axis.  Element (e, outer, z) is at outer*ostride + e*estride + z.
//
// Nyquist super-tiles (nyq_split): a half-spectrum row has n2/2 + 1 columns,
// so with n2/2 a multiple of BZ the last tile of every outer row would hold
// the Nyquist column z = n2/2 alone (a workgroup doing a whole tile's work,
// and a whole cache line fetched per element, for one column).  Instead the
// main tiles cover z < n2/2, and nsup workgroups, dispatched before them,
// each take the Nyquist columns of BZ consecutive outer rows as one tile
// (column l of super-tile s is outer row s*BZ + l, at stride ostride).  A
// column's arithmetic is the same in either kind of tile, and the
// multiplier factors are multiplied in the same order (products and sums of
// two factors are commutative in IEEE arithmetic).
//
// Nyquist plane (nyq_plane, with nyq_split): the Nyquist elements are not
// kept at column n2/2 of their own rows (one cache line per element) but 16
// to a line in the row padding: the element of row r (r = x n1 + y) at
// column n2/2 + (r mod 16) of row r / 16, which needs n2/2 and n1 multiples
// of 16 and a row pitch of at least n2/2 + 16.  The z passes write and read
// them there; a y-pass super-tile then takes its elements (y) in blocks of
// 16 per line (nyq_eblk), an x-pass super-tile its columns (y).  The
// spectrum stored to kout keeps the usual layout.
template <typename T>
struct ColArgs {
  const C<T>* in;
  C<T>* out[6];
  int ncomp;
  int64_t estride, ostride;
  int nouter, ncols, ntz;
  int nyq_split;   // request: Nyquist super-tiles where n2/2 divides into tiles
  int nyq_plane;   // request: the Nyquist elements in plane layout (see above)
  int nyq_eblk;    // plane layout: this pass transforms along y (1) or x (0)
  int nyq_col;     // set at launch: the Nyquist column (n2/2), 0 = none
  int64_t nsup;    // set at launch: super-tile workgroups (before the main ones)
  // set at launch: element e, column c of a super-tile at nyq_col plus
  // (sup_eb ? blk(e) : e sup_es) + (sup_cb ? blk(c) : c sup_cs), where
  // blk(i) = (i / 16) sup_p + i mod 16; its kout element at nyq_col + e
  // estride + c ostride (the usual layout)
  int64_t sup_es, sup_cs, sup_p;
  int sup_eb, sup_cb;
  T scale;
  const C<T>* tw;
  int64_t nblk;
  // k-space multipliers, applied before the transform: e runs along mesh axis
  // eaxis, outer along oaxis (x and y passes), the column along z.
  KopPack kop;
  int eaxis, oaxis;
  // fused forward + inverse pass (FWD): the forward transform of the loaded
  // tile, times kscale, is stored to kout before the inverse outputs
  C<T>* kout;
  T kscale;
};


// The first Stockham stage takes its radix-R0 inputs straight from the global
// loads (element j + r N/R0 of column l for butterfly (l, j)) and the last
// stage stores its outputs (element j + r N/RL) straight to global memory, so
// an n-point column costs ns - 1 LDS exchanges.  With several outputs (MC)
// the loaded tile stays in registers and each output re-applies its
// multiplier to a copy.  Measured at 512^3 on
// gfx950 this beats one workgroup per (tile, output), whose repeated tile
// reads mostly miss L2.
template <typename T, int N, int BZ, int NT, bool INV, bool KOP, bool MC, bool FWD>
__device__ __forceinline__ void fft_col_body(const ColArgs<T>& a, int64_t blk) {
  using V = C<T>;
  constexpr int E = N * BZ;
  constexpr int MAXC = MC ? 6 : 1;
  constexpr int NSTG = Plan<N>::ns;
  constexpr int R0 = Plan<N>::r[0];
  constexpr int RL = Plan<N>::r[NSTG - 1];
  constexpr int NB0 = BZ * (N / R0), PER0 = (NB0 + NT - 1) / NT;
  constexpr int NBL = BZ * (N / RL), PERL = (NBL + NT - 1) / NT;
  __shared__ V tile[E];
  __shared__ V tw[N];
  // k-space multiplier tables: per-output separable factors along the e axis
  // (fe) and the z columns (fz) and the outer-axis factor times the output's
  // factor (fo); k^2 along z (kz).  k^2 along e is recomputed from the mode
  // index, which keeps the single-output tile within a quarter of the LDS.
  __shared__ T fe[KOP ? MAXC * N : 1];
  __shared__ T fz[KOP ? MAXC * BZ : 1];
  __shared__ T kz[KOP ? BZ : 1];
  __shared__ T fo[KOP ? MAXC : 1];

  // main tile (outer row, BZ columns from z0) or Nyquist super-tile (outer
  // rows from z0, the Nyquist column of each): base offset, column stride
  // and column count of this workgroup's tile
  // (the super-tiles come first: each gathers single elements from BZ outer
  // rows, so they take longest and run while the main tiles stream)
  const int64_t nmain = a.nblk - a.nsup;
  const bool sup = blk < a.nsup;  // uniform across the workgroup
  int outer, z0, ncols;
  int64_t base, cs, es, kbase, kcs;
  bool eb = false;
  if (!sup) {
    const int64_t bid = xcd_remap(blk - a.nsup, nmain);
    outer = (int)(bid / a.ntz);
    z0 = (int)(bid % a.ntz) * BZ;
    ncols = a.ncols;
    base = (int64_t)outer * a.ostride + z0;
    cs = 1;
    es = a.estride;
    kbase = base;
    kcs = 1;
  } else {
    outer = a.nyq_col;
    z0 = (int)blk * BZ;
    ncols = a.nouter;
    // a tile of BZ <= 16 columns from z0 (a multiple of BZ) stays in one
    // block of 16, so blocked columns are contiguous within it
    base = a.nyq_col + (a.sup_cb ? (int64_t)(z0 >> 4) * a.sup_p + (z0 & 15)
                                 : (int64_t)z0 * a.sup_cs);
    cs = a.sup_cb ? 1 : a.sup_cs;
    es = a.sup_es;
    eb = a.sup_eb != 0;
    kbase = a.nyq_col + (int64_t)z0 * a.ostride;
    kcs = a.ostride;
  }
  // offset of element e of a column (blocked along y in a plane-layout
  // y-pass super-tile)
  auto eoff = [&](int e) -> int64_t {
    return eb ? (int64_t)(e >> 4) * a.sup_p + (e & 15) : (int64_t)e * es;
  };

  // the twiddle loads go out first and the tile loads second, so the tables
  // (built from each thread's own twiddles) wait only for the former
  constexpr int PTW = (N + NT - 1) / NT;
  V twr[PTW];
#pragma unroll
  for (int q = 0; q < PTW; ++q) {
    const int i = threadIdx.x + q * NT;
    twr[q] = ((N % NT == 0) || i < N) ? a.tw[i] : cmk<V>((T)0, (T)0);
  }
  // butterfly b of the first stage: column l = b % BZ, j = b / BZ
  const V* src = a.in + base;
  V v0[PER0][R0];
#pragma unroll
  for (int q = 0; q < PER0; ++q) {
    const int b = threadIdx.x + q * NT;
    const int l = b % BZ, j = b / BZ;
    const bool ok = ((NB0 % NT == 0) || b < NB0) && z0 + l < ncols;
    const int64_t lc = (int64_t)l * cs;
#pragma unroll
    for (int r = 0; r < R0; ++r)
      v0[q][r] = ok ? src[eoff(j + r * (N / R0)) + lc] : cmk<V>((T)0, (T)0);
  }
#pragma unroll
  for (int q = 0; q < PTW; ++q) {
    const int i = threadIdx.x + q * NT;
    if ((N % NT == 0) || i < N) tw[i] = twr[q];
  }
  T ko = (T)0, kstep = (T)0;
  if constexpr (KOP) {
    const KopPack& K = a.kop;
    const int nc = MC ? a.ncomp : 1;
    // the transform axis: factors once per mode from the twiddle table, then
    // each output's product
    const T two_pi_l = (T)(6.283185307179586476925286766559 / K.box[a.eaxis]);
    const T inv6h = (T)((double)N / (6.0 * K.box[a.eaxis]));
#pragma unroll
    for (int q = 0; q < PTW; ++q) {
      const int e = threadIdx.x + q * NT;
      if ((N % NT == 0) || e < N) {
        T kk, g0, g1, sn;
        axis_factors_tv<T>(e, N, twr[q], two_pi_l, inv6h, false, kk, g0, g1, sn);
        for (int c = 0; c < nc; ++c) fe[c * N + e] = comp_factor_t<T>(K, c, a.eaxis, g0, g1, sn);
      }
    }
    // a super-tile's columns run along the outer axis and its outer index is
    // the Nyquist column: the tables swap roles (fo takes the z factor,
    // fz and kz the outer axis' factor and k^2), the products stay the same
    for (int i = threadIdx.x; i < BZ * nc; i += NT) {
      const int c = i / BZ, l = i % BZ;
      double kk;
      if (sup)
        fz[c * BZ + l] = (T)(K.factor[c] * comp_axis_factor(K, c, a.oaxis, z0 + l, false, kk));
      else
        fz[c * BZ + l] = (T)comp_axis_factor(K, c, 2, z0 + l, true, kk);
      if (c == 0) kz[l] = (T)kk;
    }
    if (threadIdx.x < nc) {
      double kk;
      const int c = threadIdx.x;
      fo[c] = sup ? (T)comp_axis_factor(K, c, 2, outer, true, kk)
                  : (T)(K.factor[c] * comp_axis_factor(K, c, a.oaxis, outer, false, kk));
    }
    if (sup) {
      double kk;
      comp_axis_factor(K, 0, 2, outer, true, kk);
      ko = (T)kk;
    } else {
      const int no = K.nm[a.oaxis];
      const int nno = outer < (no + 1) / 2 ? outer : outer - no;
      const double kao = 6.283185307179586476925286766559 * (double)nno / K.box[a.oaxis];
      ko = (T)(kao * kao);
    }
    kstep = (T)(6.283185307179586476925286766559 / K.box[a.eaxis]);
  }

  if constexpr (KOP) __syncthreads();  // multiplier tables ready

  if constexpr (FWD) {
    // forward transform of the tile: first stage in registers, the rest in
    // LDS, the last stage's outputs (element j + r N/RL) back in registers
    constexpr bool SAME = (R0 == RL) && (PER0 == PERL);
    V vf[PERL][RL];
    if constexpr (NSTG == 1) {
#pragma unroll
      for (int q = 0; q < PER0; ++q) {
        Dft<R0, false, V>::run(v0[q]);
#pragma unroll
        for (int r = 0; r < R0; ++r) vf[q][r] = v0[q][r];
      }
    } else {
#pragma unroll
      for (int q = 0; q < PER0; ++q) {
        const int b = threadIdx.x + q * NT;
        if ((NB0 % NT == 0) || b < NB0) {
          const int l = b % BZ, j = b / BZ;
          V v[R0];
#pragma unroll
          for (int r = 0; r < R0; ++r) v[r] = v0[q][r];
          Dft<R0, false, V>::run(v);
#pragma unroll
          for (int r = 0; r < R0; ++r) tile[(j * R0 + r) * BZ + l] = v[r];
        }
      }
      __syncthreads();
      StagesTo<T, N, 1, NSTG - 1, R0, BZ, NT, true, BZ, 1, 1, false>::run(tile, tw);
#pragma unroll
      for (int q = 0; q < PERL; ++q) {
        const int b = threadIdx.x + q * NT;
        if ((NBL % NT == 0) || b < NBL) {
          const int l = b % BZ, j = b / BZ;
#pragma unroll
          for (int r = 0; r < RL; ++r) vf[q][r] = tile[(j + r * (N / RL)) * BZ + l];
#pragma unroll
          for (int r = 1; r < RL; ++r) vf[q][r] = cmul(vf[q][r], tw[r * j]);
          Dft<RL, false, V>::run(vf[q]);
        }
      }
    }
    // scale, store the forward result, and hand it to the inverse
    constexpr int RS = NSTG == 1 ? R0 : RL;
    constexpr int PS = NSTG == 1 ? PER0 : PERL;
    constexpr int NBS = NSTG == 1 ? NB0 : NBL;
    // (kout may be null: the forward result is then only handed on)
#pragma unroll
    for (int q = 0; q < PS; ++q) {
      const int b = threadIdx.x + q * NT;
      if ((NBS % NT == 0) || b < NBS) {
        const int l = b % BZ, j = b / BZ;
#pragma unroll
        for (int r = 0; r < RS; ++r) {
          vf[q][r] = cscale(vf[q][r], a.kscale);
          const int e = NSTG == 1 ? j * R0 + r : j + r * (N / RL);
          if (a.kout && z0 + l < ncols)
            a.kout[kbase + (int64_t)e * a.estride + (int64_t)l * kcs] = vf[q][r];
        }
      }
    }
    if constexpr (SAME || NSTG == 1) {
      // the forward outputs of butterfly (l, j) are the inverse's first-stage
      // inputs of the same butterfly (single stage: element j R0 + r, j = 0)
#pragma unroll
      for (int q = 0; q < PER0; ++q)
#pragma unroll
        for (int r = 0; r < R0; ++r) v0[q][r] = vf[q][r];
      __syncthreads();  // the last forward stage has read the tile
    } else {
      __syncthreads();
#pragma unroll
      for (int q = 0; q < PERL; ++q) {
        const int b = threadIdx.x + q * NT;
        if ((NBL % NT == 0) || b < NBL) {
          const int l = b % BZ, j = b / BZ;
#pragma unroll
          for (int r = 0; r < RL; ++r) tile[(j + r * (N / RL)) * BZ + l] = vf[q][r];
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < PER0; ++q) {
        const int b = threadIdx.x + q * NT;
        if ((NB0 % NT == 0) || b < NB0) {
          const int l = b % BZ, j = b / BZ;
#pragma unroll
          for (int r = 0; r < R0; ++r) v0[q][r] = tile[(j + r * (N / R0)) * BZ + l];
        }
      }
      __syncthreads();
    }
  }

  // output c from the first-stage inputs v (overwritten)
  auto one_output = [&](int c, V (&v)[PER0][R0]) {
    V* dst = a.out[c] + base;
    bool lap = false;
    T r2 = (T)0;
    int ng = 0;
    if constexpr (KOP) {
      lap = a.kop.laplace[c] != 0;
      r2 = (T)a.kop.r2[c];
      ng = a.kop.ngrad[c];
    }
#pragma unroll
    for (int q = 0; q < PER0; ++q) {
      const int b = threadIdx.x + q * NT;
      if ((NB0 % NT == 0) || b < NB0) {
        const int l = b % BZ, j = b / BZ;
        if constexpr (KOP) {
          const T mz = fo[c] * fz[c * BZ + l];
          const T kzo = ko + kz[l];
#pragma unroll
          for (int r = 0; r < R0; ++r) {
            const int e = j + r * (N / R0);
            T m = mz * fe[c * N + e];
            if (lap || r2 != (T)0) {
              const T ka = kstep * (T)(e < (N + 1) / 2 ? e : e - N);
              const T kk = ka * ka + kzo;
              if (lap) m = kk == (T)0 ? (T)0 : m * fast_rcp<T>(kk);
              if (r2 != (T)0) m *= exp(-kk * r2);
            }
            v[q][r] = apply_mult(v[q][r], m, ng);
          }
        }
        Dft<R0, INV, V>::run(v[q]);
        if constexpr (NSTG == 1) {
          if (z0 + l < ncols) {
#pragma unroll
            for (int r = 0; r < R0; ++r)
              dst[eoff(j * R0 + r) + (int64_t)l * cs] = cscale(v[q][r], a.scale);
          }
        } else {
#pragma unroll
          for (int r = 0; r < R0; ++r) tile[(j * R0 + r) * BZ + l] = v[q][r];
        }
      }
    }
    if constexpr (NSTG > 1) {
      __syncthreads();
      StagesTo<T, N, 1, NSTG - 1, R0, BZ, NT, true, BZ, 1, 1, INV>::run(tile, tw);
      // last stage: Ns = N / RL, outputs j + r N/RL straight to global memory
#pragma unroll
      for (int q = 0; q < PERL; ++q) {
        const int b = threadIdx.x + q * NT;
        if ((NBL % NT == 0) || b < NBL) {
          const int l = b % BZ, j = b / BZ;
          V w8[RL];
#pragma unroll
          for (int r = 0; r < RL; ++r) w8[r] = tile[(j + r * (N / RL)) * BZ + l];
#pragma unroll
          for (int r = 1; r < RL; ++r) {
            V w = tw[r * j];
            if (INV) w.y = -w.y;
            w8[r] = cmul(w8[r], w);
          }
          Dft<RL, INV, V>::run(w8);
          if (z0 + l < ncols) {
#pragma unroll
            for (int r = 0; r < RL; ++r)
              dst[eoff(j + r * (N / RL)) + (int64_t)l * cs] = cscale(w8[r], a.scale);
          }
        }
      }
    }
  };

  if constexpr (MC) {
    for (int c = 0; c < a.ncomp; ++c) {
      if (c > 0) __syncthreads();  // the previous output has read the tile
      V v[PER0][R0];
#pragma unroll
      for (int q = 0; q < PER0; ++q)
#pragma unroll
        for (int r = 0; r < R0; ++r) v[q][r] = v0[q][r];
      one_output(c, v);
    }
  } else {
    one_output(0, v0);
  }
}

template <typename T, int N, int BZ, int NT, bool INV, bool KOP, bool MC, bool FWD = false>
__global__ __launch_bounds__(NT) void k_fft_col(ColArgs<T> a) {
  fft_col_body<T, N, BZ, NT, INV, KOP, MC, FWD>(a, blockIdx.x);
}

// ---------------------------------------------------------------- z passes
// Real rows of length 2M are transformed as length-M complex rows
// (z_m = x_{2m} + i x_{2m+1}) and unpacked to M+1 half-spectrum modes:
//   X_k = (Z_k + conj Z_{M-k})/2 - (i/2) w^k (Z_k - conj Z_{M-k}),  w = exp(-2 pi i/2M).
// The inverse packs Z_k = (X_k + conj X_{M-k}) + i w^-k (X_k - conj X_{M-k})
// (imaginary parts of X_0 and X_M ignored) and runs an unnormalised inverse.
template <typename T>
struct ZArgs {
  const T* rin;      // r2c input rows [nrows][2M]
  const C<T>* cin;   // c2r input rows [nrows][pitch2]
  C<T>* cout;        // r2c output rows
  T* rout;           // c2r output rows
  int64_t nrows;
  int64_t pitch2;
  const C<T>* tw;    // table of length 2M
  // c2r only: multiply mode k by i a_z(k) first (zgrad = 1), a_z of order
  // zorder on an axis of n2 points and side L2
  int zgrad, zorder;
  double L2;
  // the Nyquist mode of row r at column M + (r mod 16) of row r / 16 (the
  // column passes' plane layout) instead of column M of row r
  int nyq_plane;
};

// Offset of row `row`'s Nyquist mode (column M).
template <typename T>
__device__ __forceinline__ int64_t z_nyq_at(const ZArgs<T>& a, int64_t row, int M) {
  return a.nyq_plane ? M + (row >> 4) * a.pitch2 + (row & 15) : row * a.pitch2 + M;
}

#ifndef FPM_Z_PAD
#define FPM_Z_PAD 4  // LDS padding shift of the z-pass lines
#endif

// Line pitch in LDS of a padded z-pass line of M elements.
// (odd, so adjacent lines start on different banks).
template <int M> constexpr int z_pitch() {
  const int p = padx<FPM_Z_PAD>(M - 1) + 1;
  return p % 2 ? p : p + 1;
}

// Forward: the first Stockham stage reads its inputs straight from the real
// rows; the unpacked half spectrum is stored straight from registers.
template <typename T, int M, int NL, int NT>
__device__ __forceinline__ void fft_z_r2c_body(const ZArgs<T>& a, int64_t blk) {
  using V = C<T>;
  constexpr int PAD = FPM_Z_PAD;
  constexpr int LP = z_pitch<M>();
  constexpr int NSTG = Plan<M>::ns;
  constexpr int R0 = Plan<M>::r[0];
  constexpr int NB0 = NL * (M / R0), PER0 = (NB0 + NT - 1) / NT;
  // pairs (k, M-k) for k < (M + 1) / 2; for even M the task k = 0 also
  // takes the middle mode, X_{M/2} = conj Z_{M/2}, so that NL * M/2 tasks
  // fill the threads' last round instead of spilling 1 per line into another
  constexpr int NP = (M + 1) / 2;
  constexpr int PERP = (NL * NP + NT - 1) / NT;
  __shared__ V s[NL * LP];
  __shared__ V tw[2 * M];
  const int64_t row0 = blk * NL;
  for (int i = threadIdx.x; i < 2 * M; i += NT) tw[i] = a.tw[i];
  const V* in = reinterpret_cast<const V*>(a.rin);
#pragma unroll
  for (int q = 0; q < PER0; ++q) {
    const int b = threadIdx.x + q * NT;
    if ((NB0 % NT == 0) || b < NB0) {
      const int j = b % (M / R0), l = b / (M / R0);
      const int64_t row = row0 + l;
      V v[R0];
#pragma unroll
      for (int r = 0; r < R0; ++r)
        v[r] = row < a.nrows ? in[row * M + j + r * (M / R0)] : cmk<V>((T)0, (T)0);
      Dft<R0, false, V>::run(v);
#pragma unroll
      for (int r = 0; r < R0; ++r) s[l * LP + padx<PAD>(j * R0 + r)] = v[r];
    }
  }
  __syncthreads();
  StagesTo<T, M, 1, NSTG, R0, NL, NT, false, 1, LP, 2, false, PAD>::run(s, tw);
#pragma unroll
  for (int q = 0; q < PERP; ++q) {
    const int idx = threadIdx.x + q * NT;
    if (idx < NL * NP) {
      const int l = idx / NP, k = idx % NP;
      const V za = s[l * LP + padx<PAD>(k % M)];
      const V zb = s[l * LP + padx<PAD>((M - k) % M)];
      const V cb = cconj(zb), ca = cconj(za);
      V xa, xb;
      {  // X_k
        const V sum = cadd(za, cb), dif = csub(za, cb);
        const V t = cmul_fma(tw[k], dif);  // w^k (Z_k - conj Z_{M-k})
        xa = cmk<V>((T)0.5 * (sum.x + t.y), (T)0.5 * (sum.y - t.x));
      }
      {  // X_{M-k}
        const V sum = cadd(zb, ca), dif = csub(zb, ca);
        const V t = cmul_fma(tw[M - k], dif);
        xb = cmk<V>((T)0.5 * (sum.x + t.y), (T)0.5 * (sum.y - t.x));
      }
      const int64_t row = row0 + l;
      if (row < a.nrows) {
        a.cout[row * a.pitch2 + k] = xa;
        if (k == 0) {
          a.cout[z_nyq_at(a, row, M)] = xb;
          if (M % 2 == 0) a.cout[row * a.pitch2 + M / 2] = cconj(s[l * LP + padx<PAD>(M / 2)]);
        } else {
          a.cout[row * a.pitch2 + (M - k)] = xb;
        }
      }
    }
  }
}

template <typename T, int M, int NL, int NT>
__global__ __launch_bounds__(NT) void k_fft_z_r2c(ZArgs<T> a) {
  fft_z_r2c_body<T, M, NL, NT>(a, blockIdx.x);
}

// Inverse: the packed rows are built straight from the global loads (with the
// optional z-gradient factor) and the last stage stores the real rows.
template <typename T, int M, int NL, int NT>
__device__ __forceinline__ void fft_z_c2r_body(const ZArgs<T>& a, int64_t blk) {
  using V = C<T>;
  constexpr int PAD = FPM_Z_PAD;
  constexpr int LP = z_pitch<M>();
  constexpr int NSTG = Plan<M>::ns;
  constexpr int RL = Plan<M>::r[NSTG - 1];
  constexpr int NBL = NL * (M / RL), PERL = (NBL + NT - 1) / NT;
  // pairs (k, M-k) for k < (M + 1) / 2; for even M the task k = 0 also
  // packs the middle mode, Z_{M/2} = 2 conj X_{M/2}
  constexpr int NP = (M + 1) / 2;
  constexpr int PERP = (NL * NP + NT - 1) / NT;
  __shared__ V s[NL * LP];
  __shared__ V tw[2 * M];
  const int64_t row0 = blk * NL;
  // every pair's loads go out first (they do not need the twiddles), then
  // the twiddle table is staged while they are in flight
  V xa[PERP], xb[PERP], xm[PERP];
#pragma unroll
  for (int q = 0; q < PERP; ++q) {
    const int idx = threadIdx.x + q * NT;
    xa[q] = xb[q] = xm[q] = cmk<V>((T)0, (T)0);
    if (idx < NL * NP) {
      const int l = idx / NP, k = idx % NP;
      const int64_t row = row0 + l;
      if (row < a.nrows) {
        xa[q] = a.cin[row * a.pitch2 + k];
        xb[q] = a.cin[k == 0 ? z_nyq_at(a, row, M) : row * a.pitch2 + (M - k)];
        if (k == 0 && M % 2 == 0) xm[q] = a.cin[row * a.pitch2 + M / 2];
      }
    }
  }
  for (int i = threadIdx.x; i < 2 * M; i += NT) tw[i] = a.tw[i];
  T two_pi_l = (T)0, inv6h = (T)0;
  if (a.zgrad) {
    two_pi_l = (T)(6.283185307179586476925286766559 / a.L2);
    inv6h = (T)((double)(2 * M) / (6.0 * a.L2));
  }
  __syncthreads();  // twiddles staged
#pragma unroll
  for (int q = 0; q < PERP; ++q) {
    const int idx = threadIdx.x + q * NT;
    if (idx < NL * NP) {
      const int l = idx / NP, k = idx % NP;
      V xa_ = xa[q], xb_ = xb[q];
      if (a.zgrad) {
        T kk, g0, g1, sn;
        axis_factors_tw<T>(k, 2 * M, two_pi_l, inv6h, true, tw, kk, g0, g1, sn);
        xa_ = apply_mult(xa_, a.zorder == 0 ? g0 : g1, 1);
        axis_factors_tw<T>(M - k, 2 * M, two_pi_l, inv6h, true, tw, kk, g0, g1, sn);
        xb_ = apply_mult(xb_, a.zorder == 0 ? g0 : g1, 1);
      }
      if (k == 0) { xa_.y = (T)0; xb_.y = (T)0; }
      const V ca = cconj(xa_), cb = cconj(xb_);
      V za, zb;
      {
        const V sum = cadd(xa_, cb), dif = csub(xa_, cb);
        const V t = cmul_fma(cconj(tw[k]), dif);  // w^-k (X_k - conj X_{M-k})
        za = cmk<V>(sum.x - t.y, sum.y + t.x);
      }
      {
        const V sum = cadd(xb_, ca), dif = csub(xb_, ca);
        const V t = cmul_fma(cconj(tw[M - k]), dif);
        zb = cmk<V>(sum.x - t.y, sum.y + t.x);
      }
      s[l * LP + padx<PAD>(k)] = za;
      if (k != 0) {
        s[l * LP + padx<PAD>(M - k)] = zb;
      } else if (M % 2 == 0) {
        V x = xm[q];
        if (a.zgrad) {
          T kk, g0, g1, sn;
          axis_factors_tw<T>(M / 2, 2 * M, two_pi_l, inv6h, true, tw, kk, g0, g1, sn);
          x = apply_mult(x, a.zorder == 0 ? g0 : g1, 1);
        }
        s[l * LP + padx<PAD>(M / 2)] = cmk<V>((T)2 * x.x, (T)-2 * x.y);
      }
    }
  }
  __syncthreads();
  StagesTo<T, M, 0, NSTG - 1, 1, NL, NT, false, 1, LP, 2, true, PAD>::run(s, tw);
  // last stage: Ns = M / RL, outputs j + r M/RL straight to the real rows
  V* out = reinterpret_cast<V*>(a.rout);
#pragma unroll
  for (int q = 0; q < PERL; ++q) {
    const int b = threadIdx.x + q * NT;
    if ((NBL % NT == 0) || b < NBL) {
      const int j = b % (M / RL), l = b / (M / RL);
      V v[RL];
#pragma unroll
      for (int r = 0; r < RL; ++r) v[r] = s[l * LP + padx<PAD>(j + r * (M / RL))];
      if constexpr (NSTG > 1) {
#pragma unroll
        for (int r = 1; r < RL; ++r) v[r] = cmul(v[r], cconj(tw[2 * r * j]));
      }
      Dft<RL, true, V>::run(v);
      const int64_t row = row0 + l;
      if (row < a.nrows) {
#pragma unroll
        for (int r = 0; r < RL; ++r) out[row * M + j + r * (M / RL)] = v[r];
      }
    }
  }
}

template <typename T, int M, int NL, int NT>
__global__ __launch_bounds__(NT) void k_fft_z_c2r(ZArgs<T> a) {
  fft_z_c2r_body<T, M, NL, NT>(a, blockIdx.x);
}

// ---------------------------------------------------------------- dispatch
#define FPM_LENGTHS(X) \
  X(2) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) \
  X(3) X(6) X(12) X(24) X(48) X(96) X(192) X(384) X(768)

static bool supported_len(int n) {
  switch (n) {
#define FPM_CASE(N_) case N_:
    FPM_LENGTHS(FPM_CASE)
#undef FPM_CASE
    return true;
    default: return false;
  }
}

// Tile counts of a column pass with BZ-column tiles: whole-row tiles, or with
// nyq_split (and n2/2 = ncols - 1 a multiple of BZ) tiles of the first
// ncols - 1 columns plus the Nyquist super-tiles.
template <typename T>
static void col_tiles(ColArgs<T>& a, int BZ) {
  a.nyq_col = 0;
  a.nsup = 0;
  if (a.nyq_split && a.ncols > 1 && (a.ncols - 1) % BZ == 0) {
    a.nyq_col = a.ncols - 1;
    a.ncols -= 1;
    a.nsup = (a.nouter + BZ - 1) / BZ;
    // Nyquist column of each row: columns are outer rows, elements along e
    a.sup_es = a.estride;
    a.sup_cs = a.ostride;
    a.sup_eb = a.sup_cb = 0;
    a.sup_p = 0;
    if (a.nyq_plane && a.nyq_eblk) {
      // y pass (estride = pitch, ostride = n1 pitch): rows x n1 + y, y in
      // blocks of 16 per padding line, x at n1 / 16 lines each
      a.sup_p = a.estride;
      a.sup_cs = a.ostride / 16;
      a.sup_eb = 1;
    } else if (a.nyq_plane) {
      // x pass (estride = n1 pitch, ostride = pitch): columns y blocked
      a.sup_p = a.ostride;
      a.sup_es = a.estride / 16;
      a.sup_cb = 1;
    }
  }
  a.ntz = (a.ncols + BZ - 1) / BZ;
  a.nblk = (int64_t)a.ntz * a.nouter + a.nsup;
}

template <typename T, int N, bool INV, bool KOP, bool MC, int WIDE>
static void launch_col_nwm(ColArgs<T> a, hipStream_t st) {
  constexpr int BZ = ColW<T, N, WIDE, MC>::BZ;
  constexpr int NT = threads_for<T>(N * BZ);
  col_tiles(a, BZ);
  if (a.nblk <= 0) return;
  hipLaunchKernelGGL((k_fft_col<T, N, BZ, NT, INV, KOP, MC>), dim3((unsigned)a.nblk), dim3(NT),
                     0, st, a);
}

template <typename T, int N, bool INV, bool KOP, int WIDE>
static void launch_col_nw(ColArgs<T> a, hipStream_t st) {
  if (KOP && a.ncomp > 1) launch_col_nwm<T, N, INV, KOP, KOP, WIDE>(a, st);
  else launch_col_nwm<T, N, INV, KOP, false, WIDE>(a, st);
}

template <typename T, int N, bool INV, bool KOP>
static void launch_col_n(ColArgs<T> a, const fpm_mesh* m, hipStream_t st) {
  if (col_choice(m, (int)sizeof(T), N)) launch_col_nw<T, N, INV, KOP, 1>(a, st);
  else launch_col_nw<T, N, INV, KOP, 0>(a, st);
}

template <typename T, int N, bool MC, int WIDE>
static void launch_col_fused_nwm(ColArgs<T> a, hipStream_t st) {
  constexpr int BZ = ColW<T, N, WIDE, MC>::BZ;
  constexpr int NT = threads_for<T>(N * BZ);
  col_tiles(a, BZ);
  if (a.nblk <= 0) return;
  hipLaunchKernelGGL((k_fft_col<T, N, BZ, NT, true, true, MC, true>), dim3((unsigned)a.nblk),
                     dim3(NT), 0, st, a);
}

template <typename T, int N, int WIDE>
static void launch_col_fused_nw(ColArgs<T> a, hipStream_t st) {
  if (a.ncomp > 1) launch_col_fused_nwm<T, N, true, WIDE>(a, st);
  else launch_col_fused_nwm<T, N, false, WIDE>(a, st);
}

template <typename T, int N>
static void launch_col_fused_n(ColArgs<T> a, const fpm_mesh* m, hipStream_t st) {
  if (col_choice(m, (int)sizeof(T), N)) launch_col_fused_nw<T, N, 1>(a, st);
  else launch_col_fused_nw<T, N, 0>(a, st);
}

// Forward transform along the e axis (stored to a.kout, times a.kscale) fused
// with the multiplier and inverse transform of every output.
template <typename T>
static int launch_col_fused(int n, ColArgs<T> a, const fpm_mesh* m, hipStream_t st) {
  if (a.ncomp < 1 || a.ncomp > 6) return FPM_ERR_ARG;
  for (int c = 0; c < a.ncomp; ++c)
    if (!a.out[c]) return FPM_ERR_ARG;
  switch (n) {
#define FPM_CASE(N_) \
  case N_: launch_col_fused_n<T, N_>(a, m, st); break;
    FPM_LENGTHS(FPM_CASE)
#undef FPM_CASE
    default: return FPM_ERR_UNSUPPORTED;
  }
  return FPM_OK;
}

template <typename T, bool INV, bool KOP>
static int launch_col(int n, ColArgs<T> a, const fpm_mesh* m, hipStream_t st) {
  if (a.ncomp < 1 || a.ncomp > 6 || (!KOP && a.ncomp != 1)) return FPM_ERR_ARG;
  for (int c = 0; c < a.ncomp; ++c)
    if (!a.out[c]) return FPM_ERR_ARG;
  switch (n) {
#define FPM_CASE(N_) \
  case N_: launch_col_n<T, N_, INV, KOP>(a, m, st); break;
    FPM_LENGTHS(FPM_CASE)
#undef FPM_CASE
    default: return FPM_ERR_UNSUPPORTED;
  }
  return FPM_OK;
}

template <typename T, int M, bool INV, int ZV>
static void launch_z_mv(const ZArgs<T>& a, hipStream_t st) {
  constexpr int NL = z_lines(M, ZV);
  constexpr int NT = clampi(round64(NL * M / FPM_Z_EPT), 64, 1024);
  const unsigned nblk = (unsigned)((a.nrows + NL - 1) / NL);
  if (INV)
    hipLaunchKernelGGL((k_fft_z_c2r<T, M, NL, NT>), dim3(nblk), dim3(NT), 0, st, a);
  else
    hipLaunchKernelGGL((k_fft_z_r2c<T, M, NL, NT>), dim3(nblk), dim3(NT), 0, st, a);
}

template <typename T, int M, bool INV>
static void launch_z_m(const ZArgs<T>& a, const fpm_mesh* m, hipStream_t st) {
  switch (FPM_VAR_FIELD(m->variant, FPM_VAR_Z_LINES)) {
    case 1: launch_z_mv<T, M, INV, 1>(a, st); break;
    case 2: launch_z_mv<T, M, INV, 2>(a, st); break;
    default: launch_z_mv<T, M, INV, 0>(a, st); break;
  }
}

template <typename T, bool INV>
static int launch_z(int n2, const ZArgs<T>& a, const fpm_mesh* m, hipStream_t st) {
  if (n2 % 2) return FPM_ERR_UNSUPPORTED;
  if (a.nrows <= 0) return FPM_OK;
  switch (n2 / 2) {
#define FPM_CASE(N_) \
  case N_: launch_z_m<T, N_, INV>(a, m, st); break;
    FPM_LENGTHS(FPM_CASE)
#undef FPM_CASE
    default: return FPM_ERR_UNSUPPORTED;
  }
  return FPM_OK;
}

static int check_mesh(const fpm_mesh* m, int64_t pitch2) {
  if (!m) return FPM_ERR_ARG;
  for (int d = 0; d < 3; ++d)
    if (m->nmesh[d] <= 0 || !(m->box[d] > 0)) return FPM_ERR_ARG;
  if (pitch2 < m->nmesh[2] / 2 + 1) return FPM_ERR_ARG;
  if (!supported_len(m->nmesh[0]) || !supported_len(m->nmesh[1])) return FPM_ERR_UNSUPPORTED;
  if (m->nmesh[2] % 2 || !supported_len(m->nmesh[2] / 2)) return FPM_ERR_UNSUPPORTED;
  return FPM_OK;
}

// Whether a mesh's transform chain can keep its Nyquist modes in plane
// layout (ColArgs): n2/2 and n1 multiples of 16, and 16 columns of padding.
static int nyq_plane_ok(int n1, int n2, int64_t pitch2) {
  const int M = n2 / 2;
  return M % 16 == 0 && n1 % 16 == 0 && pitch2 >= M + 16 ? 1 : 0;
}

template <typename T>
static int r2c_impl(const void* in, void* out, const fpm_mesh* m, int64_t pitch2, const void* tw0,
                    const void* tw1, const void* tw2, hipStream_t st) {
  int rc = check_mesh(m, pitch2);
  if (rc) return rc;
  if (!in || !out || !tw0 || !tw1 || !tw2) return FPM_ERR_ARG;
  const int n0 = m->nmesh[0], n1 = m->nmesh[1], n2 = m->nmesh[2];
  const int nh = n2 / 2 + 1;
  C<T>* cout = static_cast<C<T>*>(out);
  ZArgs<T> z{};
  z.rin = static_cast<const T*>(in);
  z.cout = cout;
  z.nrows = (int64_t)n0 * n1;
  z.pitch2 = pitch2;
  z.tw = static_cast<const C<T>*>(tw2);
  if ((rc = launch_z<T, false>(n2, z, m, st))) return rc;
  FPM_LAUNCH_CHECK();
  // y pass, in place
  ColArgs<T> y{};
  y.in = cout;
  y.out[0] = cout;
  y.ncomp = 1;
  y.estride = pitch2;
  y.ostride = (int64_t)n1 * pitch2;
  y.nouter = n0;
  y.ncols = nh;
  y.scale = (T)1;
  y.tw = static_cast<const C<T>*>(tw1);
  if ((rc = launch_col<T, false, false>(n1, y, m, st))) return rc;
  FPM_LAUNCH_CHECK();
  // x pass, in place, with 1/N
  ColArgs<T> x{};
  x.in = cout;
  x.out[0] = cout;
  x.ncomp = 1;
  x.estride = (int64_t)n1 * pitch2;
  x.ostride = pitch2;
  x.nouter = n1;
  x.ncols = nh;
  x.scale = (T)(1.0 / ((double)n0 * n1 * n2));
  x.tw = static_cast<const C<T>*>(tw0);
  if ((rc = launch_col<T, false, false>(n0, x, m, st))) return rc;
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

static int fill_kop(KopPack& K, int ncomp, const fpm_kop* ops, const fpm_mesh* m) {
  K = KopPack{};
  K.ncomp = ncomp;
  for (int d = 0; d < 3; ++d) {
    K.nm[d] = m->nmesh[d];
    K.box[d] = m->box[d];
  }
  for (int c = 0; c < ncomp; ++c) {
    if (!ops) {
      K.factor[c] = 1.0;
      continue;
    }
    const fpm_kop& o = ops[c];
    if (o.ngrad < 0 || o.ngrad > 2 || o.deconv < 0 || o.deconv > 8) return FPM_ERR_ARG;
    if (o.grad_order != 0 && o.grad_order != 1) return FPM_ERR_ARG;
    for (int g = 0; g < o.ngrad; ++g)
      if (o.grad_dir[g] < 0 || o.grad_dir[g] > 2) return FPM_ERR_ARG;
    K.factor[c] = o.factor;
    K.laplace[c] = o.laplace ? 1 : 0;
    K.ngrad[c] = o.ngrad;
    K.gdir[c][0] = o.grad_dir[0];
    K.gdir[c][1] = o.grad_dir[1];
    K.r2[c] = o.r_split * o.r_split;
    K.grad_order[c] = o.grad_order;
    K.deconv[c] = o.deconv;
  }
  return FPM_OK;
}

template <typename T>
static int c2r_impl(const void* in, int ncomp, const fpm_kop* ops, void* const* outs,
                    void* const* work, const fpm_mesh* m, int64_t pitch2, const void* tw0,
                    const void* tw1, const void* tw2, hipStream_t st) {
  int rc = check_mesh(m, pitch2);
  if (rc) return rc;
  if (ncomp < 1 || ncomp > 6 || !in || !outs || !work || !tw0 || !tw1 || !tw2) return FPM_ERR_ARG;
  for (int c = 0; c < ncomp; ++c) {
    if (!outs[c] || !work[c] || work[c] == in) return FPM_ERR_ARG;
    for (int d = 0; d < c; ++d)
      if (work[c] == work[d]) return FPM_ERR_ARG;
  }
  const int n0 = m->nmesh[0], n1 = m->nmesh[1], n2 = m->nmesh[2];
  const int nh = n2 / 2 + 1;
  ColArgs<T> x{};
  x.in = static_cast<const C<T>*>(in);
  for (int c = 0; c < ncomp; ++c) x.out[c] = static_cast<C<T>*>(work[c]);
  x.ncomp = ncomp;
  x.estride = (int64_t)n1 * pitch2;
  x.ostride = pitch2;
  x.nouter = n1;
  x.ncols = nh;
  x.scale = (T)1;
  x.tw = static_cast<const C<T>*>(tw0);
  x.eaxis = 0;
  x.oaxis = 1;
  if ((rc = fill_kop(x.kop, ncomp, ops, m))) return rc;
  if ((rc = launch_col<T, true, true>(n0, x, m, st))) return rc;
  FPM_LAUNCH_CHECK();
  for (int c = 0; c < ncomp; ++c) {
    C<T>* wc = static_cast<C<T>*>(work[c]);
    ColArgs<T> y{};
    y.in = wc;
    y.out[0] = wc;
    y.ncomp = 1;
    y.estride = pitch2;
    y.ostride = (int64_t)n1 * pitch2;
    y.nouter = n0;
    y.ncols = nh;
    y.scale = (T)1;
    y.tw = static_cast<const C<T>*>(tw1);
    if ((rc = launch_col<T, true, false>(n1, y, m, st))) return rc;
    FPM_LAUNCH_CHECK();
    ZArgs<T> z{};
    z.cin = wc;
    z.rout = static_cast<T*>(outs[c]);
    z.nrows = (int64_t)n0 * n1;
    z.pitch2 = pitch2;
    z.tw = static_cast<const C<T>*>(tw2);
    if ((rc = launch_z<T, true>(n2, z, m, st))) return rc;
    FPM_LAUNCH_CHECK();
  }
  return FPM_OK;
}

// Forward transform of `rin` into `kout` and inverse transforms of ncomp
// multiplier outputs, sharing the x pass: the forward x pass of r2c and the
// first inverse pass of c2r work on the same x columns, so one kernel does
// both (storing the forward result unless kout is null) and the mesh is read
// once less.  Without kout the forward passes run in work[0], which the x
// pass overwrites in place with component 0: each workgroup reads its whole
// column tile before it writes the same columns.
template <typename T>
static int r2c_c2r_impl(const void* rin, void* kout, int ncomp, const fpm_kop* ops,
                        void* const* outs, void* const* work, const fpm_mesh* m, int64_t pitch2,
                        const void* tw0, const void* tw1, const void* tw2, hipStream_t st) {
  int rc = check_mesh(m, pitch2);
  if (rc) return rc;
  if (ncomp < 1 || ncomp > 6 || !rin || !outs || !work || !tw0 || !tw1 || !tw2)
    return FPM_ERR_ARG;
  for (int c = 0; c < ncomp; ++c) {
    if (!outs[c] || !work[c] || work[c] == kout || work[c] == rin) return FPM_ERR_ARG;
    for (int d = 0; d < c; ++d)
      if (work[c] == work[d]) return FPM_ERR_ARG;
  }
  const int n0 = m->nmesh[0], n1 = m->nmesh[1], n2 = m->nmesh[2];
  const int nh = n2 / 2 + 1;
  // Nyquist modes in plane layout where the mesh allows; the forward passes
  // then run in work[0], so that kout keeps the usual layout
  const int plane = nyq_plane_ok(n1, n2, pitch2);
  C<T>* ck = static_cast<C<T>*>(kout && !plane ? kout : work[0]);
  ZArgs<T> z{};
  z.rin = static_cast<const T*>(rin);
  z.cout = ck;
  z.nrows = (int64_t)n0 * n1;
  z.pitch2 = pitch2;
  z.tw = static_cast<const C<T>*>(tw2);
  z.nyq_plane = plane;
  if ((rc = launch_z<T, false>(n2, z, m, st))) return rc;
  FPM_LAUNCH_CHECK();
  ColArgs<T> y{};
  y.in = ck;
  y.out[0] = ck;
  y.ncomp = 1;
  y.estride = pitch2;
  y.ostride = (int64_t)n1 * pitch2;
  y.nouter = n0;
  y.ncols = nh;
  y.nyq_split = 1;
  y.nyq_plane = plane;
  y.nyq_eblk = 1;
  y.scale = (T)1;
  y.tw = static_cast<const C<T>*>(tw1);
  if ((rc = launch_col<T, false, false>(n1, y, m, st))) return rc;
  FPM_LAUNCH_CHECK();
  ColArgs<T> x{};
  x.in = ck;
  x.kout = static_cast<C<T>*>(kout);
  x.kscale = (T)(1.0 / ((double)n0 * n1 * n2));
  for (int c = 0; c < ncomp; ++c) x.out[c] = static_cast<C<T>*>(work[c]);
  x.ncomp = ncomp;
  x.estride = (int64_t)n1 * pitch2;
  x.ostride = pitch2;
  x.nouter = n1;
  x.ncols = nh;
  x.nyq_split = 1;
  x.nyq_plane = plane;
  x.scale = (T)1;
  x.tw = static_cast<const C<T>*>(tw0);
  x.eaxis = 0;
  x.oaxis = 1;
  if ((rc = fill_kop(x.kop, ncomp, ops, m))) return rc;
  if ((rc = launch_col_fused<T>(n0, x, m, st))) return rc;
  FPM_LAUNCH_CHECK();
  for (int c = 0; c < ncomp; ++c) {
    C<T>* wc = static_cast<C<T>*>(work[c]);
    ColArgs<T> yi = y;
    yi.in = wc;
    yi.out[0] = wc;
    if ((rc = launch_col<T, true, false>(n1, yi, m, st))) return rc;
    FPM_LAUNCH_CHECK();
    ZArgs<T> zi{};
    zi.cin = wc;
    zi.rout = static_cast<T*>(outs[c]);
    zi.nrows = (int64_t)n0 * n1;
    zi.pitch2 = pitch2;
    zi.tw = static_cast<const C<T>*>(tw2);
    zi.nyq_plane = plane;
    if ((rc = launch_z<T, true>(n2, zi, m, st))) return rc;
    FPM_LAUNCH_CHECK();
  }
  return FPM_OK;
}

// outs[d] = c2r(i a_d(k_d) * common(k) * in), d = 0, 1, 2, using that a_d
// depends on k_d alone: the x pass reads `in` once and writes the common
// product before and after the x gradient (A, B); the y pass turns A into
// A and C = (y gradient of A); the z pass applies the z gradient to A on the
// fly.  Eight mesh passes instead of ten.
template <typename T>
static int c2r_grad3_impl(const void* in, const fpm_kop* common, void* const* outs,
                          void* const* work, const fpm_mesh* m, int64_t pitch2, const void* tw0,
                          const void* tw1, const void* tw2, hipStream_t st) {
  int rc = check_mesh(m, pitch2);
  if (rc) return rc;
  if (!in || !common || !outs || !work || !tw0 || !tw1 || !tw2) return FPM_ERR_ARG;
  if (common->ngrad != 0) return FPM_ERR_ARG;
  for (int c = 0; c < 3; ++c) {
    if (!outs[c] || !work[c] || work[c] == in) return FPM_ERR_ARG;
    for (int d = 0; d < c; ++d)
      if (work[c] == work[d]) return FPM_ERR_ARG;
  }
  const int n0 = m->nmesh[0], n1 = m->nmesh[1], n2 = m->nmesh[2];
  const int nh = n2 / 2 + 1;
  C<T>* A = static_cast<C<T>*>(work[0]);
  C<T>* B = static_cast<C<T>*>(work[1]);
  C<T>* Cw = static_cast<C<T>*>(work[2]);
  fpm_kop xo[2] = {*common, *common};
  xo[1].ngrad = 1;
  xo[1].grad_dir[0] = 0;
  ColArgs<T> x{};
  x.in = static_cast<const C<T>*>(in);
  x.out[0] = A;
  x.out[1] = B;
  x.ncomp = 2;
  x.estride = (int64_t)n1 * pitch2;
  x.ostride = pitch2;
  x.nouter = n1;
  x.ncols = nh;
  x.scale = (T)1;
  x.tw = static_cast<const C<T>*>(tw0);
  x.eaxis = 0;
  x.oaxis = 1;
  if ((rc = fill_kop(x.kop, 2, xo, m))) return rc;
  if ((rc = launch_col<T, true, true>(n0, x, m, st))) return rc;
  FPM_LAUNCH_CHECK();
  // y pass on A: identity (in place) and the y gradient (into C)
  fpm_kop yo[2] = {};
  yo[0].factor = 1.0;
  yo[1].factor = 1.0;
  yo[1].ngrad = 1;
  yo[1].grad_dir[0] = 1;
  yo[1].grad_order = common->grad_order;
  ColArgs<T> y{};
  y.in = A;
  y.out[0] = A;
  y.out[1] = Cw;
  y.ncomp = 2;
  y.estride = pitch2;
  y.ostride = (int64_t)n1 * pitch2;
  y.nouter = n0;
  y.ncols = nh;
  y.scale = (T)1;
  y.tw = static_cast<const C<T>*>(tw1);
  y.eaxis = 1;
  y.oaxis = 0;
  if ((rc = fill_kop(y.kop, 2, yo, m))) return rc;
  if ((rc = launch_col<T, true, true>(n1, y, m, st))) return rc;
  FPM_LAUNCH_CHECK();
  // y pass on B, in place
  ColArgs<T> yb = y;
  yb.in = B;
  yb.out[0] = B;
  yb.ncomp = 1;
  if ((rc = launch_col<T, true, false>(n1, yb, m, st))) return rc;
  FPM_LAUNCH_CHECK();
  // z passes: B -> F_x, C -> F_y, A with the z gradient -> F_z
  const C<T>* src[3] = {B, Cw, A};
  for (int d = 0; d < 3; ++d) {
    ZArgs<T> z{};
    z.cin = src[d];
    z.rout = static_cast<T*>(outs[d]);
    z.nrows = (int64_t)n0 * n1;
    z.pitch2 = pitch2;
    z.tw = static_cast<const C<T>*>(tw2);
    z.zgrad = d == 2 ? 1 : 0;
    z.zorder = common->grad_order;
    z.L2 = m->box[2];
    if ((rc = launch_z<T, true>(n2, z, m, st))) return rc;
    FPM_LAUNCH_CHECK();
  }
  return FPM_OK;
}

// In-place multiplier on a complex mesh (no transform).
template <typename T>
__global__ void k_kspace_apply(C<T>* f, int64_t pitch2, KopPack K) {
  const int n0 = K.nm[0], n1 = K.nm[1], nh = K.nm[2] / 2 + 1;
  const int64_t total = (int64_t)n0 * n1 * nh;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(t % nh);
    const int64_t r = t / nh;
    const int j = (int)(r % n1), i = (int)(r / n1);
    const AxisTab<T> ax = axis_tab<T>(i, n0, K.box[0], false);
    const AxisTab<T> ay = axis_tab<T>(j, n1, K.box[1], false);
    const AxisTab<T> az = axis_tab<T>(k, K.nm[2], K.box[2], true);
    const T mult = kop_mult<T>(K, 0, ax, ay, az);
    C<T>* p = f + r * pitch2 + k;
    *p = apply_mult(*p, mult, K.ngrad[0]);
  }
}

template <typename T>
static int kspace_apply_impl(void* field, const fpm_kop* op, const fpm_mesh* m, int64_t pitch2,
                             hipStream_t st) {
  if (!m || !field || !op || pitch2 < m->nmesh[2] / 2 + 1) return FPM_ERR_ARG;
  KopPack K;
  int rc = fill_kop(K, 1, op, m);
  if (rc) return rc;
  const int64_t total = (int64_t)m->nmesh[0] * m->nmesh[1] * (m->nmesh[2] / 2 + 1);
  const int64_t nb = (total + 255) / 256;
  hipLaunchKernelGGL(k_kspace_apply<T>, dim3((unsigned)(nb < 65536 ? nb : 65536)), dim3(256), 0, st,
                     static_cast<C<T>*>(field), pitch2, K);
  FPM_LAUNCH_CHECK();
  return FPM_OK;
}

template <typename T>
static int twiddles_impl(int n, T* out) {
  if (!out || n <= 0) return FPM_ERR_ARG;
  const double two_pi = 6.283185307179586476925286766559;
  for (int m = 0; m < n; ++m) {
    // exact values at the quarter points keep the tables symmetric
    const double ang = -two_pi * (double)m / (double)n;
    double c = std::cos(ang), s = std::sin(ang);
    if (4 * m == n) { c = 0.0; s = -1.0; }
    if (2 * m == n) { c = -1.0; s = 0.0; }
    if (4 * m == 3 * n) { c = 0.0; s = 1.0; }
    out[2 * m] = (T)c;
    out[2 * m + 1] = (T)s;
  }
  return FPM_OK;
}

// ---------------------------------------------------------------- batched boxes
// The Force's transform (fpm_fft_r2c_c2r with one multiplier and no stored
// spectrum) of many boxes: each pass is one launch per compiled kernel (axis
// length and tiling), one workgroup per column tile or row group of every
// box, through the pass bodies above, each box's arguments read from a
// device table.  Float32, axis lengths that are multiples of 16.
enum { kFbZr2c = 1, kFbYfwd, kFbXfused, kFbYinv, kFbZc2r };

#define FPM_BCOL_LENGTHS(X) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)
#define FPM_BZ_LENGTHS(X) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

template <typename T, int N, int BZ, int NT, bool INV, bool KOP, bool MC, bool FWD>
__global__ __launch_bounds__(NT) void k_batch_fft_col(const ColArgs<T>* __restrict__ tab,
                                                       const int64_t* __restrict__ pre,
                                                       const int64_t* __restrict__ cnt, int nb,
                                                       int64_t total) {
  const int64_t it = blockIdx.x;
  if (it >= total) return;
  const int b = batch_box(pre, nb, it);
  const int64_t j = it - pre[b];
  if (j >= cnt[b]) return;
  fft_col_body<T, N, BZ, NT, INV, KOP, MC, FWD>(tab[b], j);
}

template <typename T, int M, int NL, int NT, bool INV>
__global__ __launch_bounds__(NT) void k_batch_fft_z(const ZArgs<T>* __restrict__ tab,
                                                     const int64_t* __restrict__ pre,
                                                     const int64_t* __restrict__ cnt, int nb,
                                                     int64_t total) {
  const int64_t it = blockIdx.x;
  if (it >= total) return;
  const int b = batch_box(pre, nb, it);
  const int64_t j = it - pre[b];
  if (j >= cnt[b]) return;
  const ZArgs<T> a = tab[b];  // a local copy: the passes' stores cannot alias it
  if constexpr (INV) fft_z_c2r_body<T, M, NL, NT>(a, j);
  else fft_z_r2c_body<T, M, NL, NT>(a, j);
}

template <typename T, int N, bool INV, bool KOP, bool FWD, int WIDE>
static void batch_col_nw(const BatchLaunch& L, hipStream_t st) {
  constexpr int BZ = ColW<T, N, WIDE>::BZ;
  constexpr int NT = threads_for<T>(N * BZ);
  hipLaunchKernelGGL((k_batch_fft_col<T, N, BZ, NT, INV, KOP, false, FWD>),
                     dim3(batch_grid(L.total)), dim3(NT), 0, st,
                     static_cast<const ColArgs<T>*>(L.tab), L.pre, L.cnt, L.nb, L.total);
}

template <typename T, int M, bool INV, int ZV>
static void batch_z_mv(const BatchLaunch& L, hipStream_t st) {
  constexpr int NL = z_lines(M, ZV);
  constexpr int NT = clampi(round64(NL * M / FPM_Z_EPT), 64, 1024);
  hipLaunchKernelGGL((k_batch_fft_z<T, M, NL, NT, INV>), dim3(batch_grid(L.total)), dim3(NT), 0,
                     st, static_cast<const ZArgs<T>*>(L.tab), L.pre, L.cnt, L.nb, L.total);
}

// column tile width (z columns) of an axis of length n with tiling `wide`
template <typename T>
static int col_bz(int n, int wide) {
  switch (n) {
#define FPM_CASE(N_) \
  case N_: return wide ? ColW<T, N_, 1>::BZ : ColW<T, N_, 0>::BZ;
    FPM_BCOL_LENGTHS(FPM_CASE)
#undef FPM_CASE
    default: return 0;
  }
}

static bool batch_col_len(int n) {
  switch (n) {
#define FPM_CASE(N_) case N_:
    FPM_BCOL_LENGTHS(FPM_CASE)
#undef FPM_CASE
    return true;
    default: return false;
  }
}

static bool batch_z_len(int m) {
  switch (m) {
#define FPM_CASE(N_) case N_:
    FPM_BZ_LENGTHS(FPM_CASE)
#undef FPM_CASE
    return true;
    default: return false;
  }
}

int fft_batch_build(int nbox, const fpm_batch_box* boxes, DevArena* arena,
                    std::vector<BatchLaunch>& out) {
  using T = float;
  std::vector<ZArgs<T>> zf(nbox), zi(nbox);
  std::vector<ColArgs<T>> yf(nbox), xf(nbox), yi(nbox);
  std::vector<BatchKey> kz(nbox), ky(nbox), kx(nbox);
  for (int i = 0; i < nbox; ++i) {
    const fpm_batch_box& bx = boxes[i];
    const fpm_mesh* m = &bx.mesh;
    int rc = check_mesh(m, bx.pitch2);
    if (rc) return rc;
    if (!bx.rho || !bx.phi || !bx.work || !bx.tw[0] || !bx.tw[1] || !bx.tw[2]) return FPM_ERR_ARG;
    const int n0 = m->nmesh[0], n1 = m->nmesh[1], n2 = m->nmesh[2];
    if (!batch_col_len(n0) || !batch_col_len(n1) || !batch_z_len(n2 / 2))
      return FPM_ERR_UNSUPPORTED;
    const int nh = n2 / 2 + 1;
    C<T>* wk = static_cast<C<T>*>(bx.work);
    const int zv = FPM_VAR_FIELD(m->variant, FPM_VAR_Z_LINES);
    const int plane = nyq_plane_ok(n1, n2, bx.pitch2);  // as in the box's own transform
    ZArgs<T> z{};
    z.rin = static_cast<const T*>(bx.rho);
    z.cout = wk;
    z.nrows = (int64_t)n0 * n1;
    z.pitch2 = bx.pitch2;
    z.tw = static_cast<const C<T>*>(bx.tw[2]);
    z.nyq_plane = plane;
    zf[i] = z;
    ZArgs<T> zc{};
    zc.cin = wk;
    zc.rout = static_cast<T*>(bx.phi);
    zc.nrows = (int64_t)n0 * n1;
    zc.pitch2 = bx.pitch2;
    zc.tw = static_cast<const C<T>*>(bx.tw[2]);
    zc.nyq_plane = plane;
    zi[i] = zc;
    kz[i] = BatchKey{n2 / 2, zv, 0, 0};
    const int wy = col_choice(m, (int)sizeof(T), n1), wx = col_choice(m, (int)sizeof(T), n0);
    ColArgs<T> y{};
    y.in = wk;
    y.out[0] = wk;
    y.ncomp = 1;
    y.estride = bx.pitch2;
    y.ostride = (int64_t)n1 * bx.pitch2;
    y.nouter = n0;
    y.ncols = nh;
    y.scale = (T)1;
    y.tw = static_cast<const C<T>*>(bx.tw[1]);
    y.nyq_split = 1;
    y.nyq_plane = plane;
    y.nyq_eblk = 1;
    col_tiles(y, col_bz<T>(n1, wy));
    yf[i] = y;
    yi[i] = y;
    ky[i] = BatchKey{n1, wy, 0, 0};
    ColArgs<T> x{};
    x.in = wk;
    x.kout = nullptr;
    x.kscale = (T)(1.0 / ((double)n0 * n1 * n2));
    x.out[0] = wk;
    x.ncomp = 1;
    x.estride = (int64_t)n1 * bx.pitch2;
    x.ostride = bx.pitch2;
    x.nouter = n1;
    x.ncols = nh;
    x.scale = (T)1;
    x.tw = static_cast<const C<T>*>(bx.tw[0]);
    x.eaxis = 0;
    x.oaxis = 1;
    if ((rc = fill_kop(x.kop, 1, &bx.common, m))) return rc;
    x.nyq_split = 1;
    x.nyq_plane = plane;
    col_tiles(x, col_bz<T>(n0, wx));
    xf[i] = x;
    kx[i] = BatchKey{n0, wx, 0, 0};
  }
  // per-box keys and item counts, found by the element's place in its vector
  auto zfkey = [&](const ZArgs<T>& a) { return kz[&a - zf.data()]; };
  auto zikey = [&](const ZArgs<T>& a) { return kz[&a - zi.data()]; };
  auto zitems = [](const BatchKey& k, const ZArgs<T>& a) {
    const int nl = z_lines(k.a, k.b);
    return (a.nrows + nl - 1) / nl;
  };
  auto zfitems = [&](const ZArgs<T>& a) { return zitems(zfkey(a), a); };
  auto ziitems = [&](const ZArgs<T>& a) { return zitems(zikey(a), a); };
  auto citems = [](const ColArgs<T>& a) { return a.nblk; };
  int rc;
  if ((rc = batch_groups(zf, kFbZr2c, zfkey, zfitems, arena, out))) return rc;
  if ((rc = batch_groups(yf, kFbYfwd, [&](const ColArgs<T>& a) { return ky[&a - yf.data()]; },
                         citems, arena, out)))
    return rc;
  if ((rc = batch_groups(xf, kFbXfused, [&](const ColArgs<T>& a) { return kx[&a - xf.data()]; },
                         citems, arena, out)))
    return rc;
  if ((rc = batch_groups(yi, kFbYinv, [&](const ColArgs<T>& a) { return ky[&a - yi.data()]; },
                         citems, arena, out)))
    return rc;
  if ((rc = batch_groups(zi, kFbZc2r, zikey, ziitems, arena, out))) return rc;
  return FPM_OK;
}

int fft_batch_run(const std::vector<BatchLaunch>& ls, hipStream_t st) {
  using T = float;
  for (const BatchLaunch& L : ls) {
    bool ok = true;
    switch (L.kind) {
      case kFbZr2c:
      case kFbZc2r: {
        const bool inv = L.kind == kFbZc2r;
        switch (L.a * 8 + L.b * 2 + (inv ? 1 : 0)) {
#define FPM_CASE(M_)                                                             \
  case M_ * 8 + 0: batch_z_mv<T, M_, false, 0>(L, st); break;                     \
  case M_ * 8 + 1: batch_z_mv<T, M_, true, 0>(L, st); break;                      \
  case M_ * 8 + 2: batch_z_mv<T, M_, false, 1>(L, st); break;                     \
  case M_ * 8 + 3: batch_z_mv<T, M_, true, 1>(L, st); break;                      \
  case M_ * 8 + 4: batch_z_mv<T, M_, false, 2>(L, st); break;                     \
  case M_ * 8 + 5: batch_z_mv<T, M_, true, 2>(L, st); break;
          FPM_BZ_LENGTHS(FPM_CASE)
#undef FPM_CASE
          default: ok = false;
        }
        break;
      }
      case kFbYfwd:
      case kFbXfused:
      case kFbYinv: {
        const int k = L.kind == kFbYfwd ? 0 : (L.kind == kFbXfused ? 1 : 2);
        switch (L.a * 8 + L.b * 4 + k) {
#define FPM_CASE(N_)                                                                     \
  case N_ * 8 + 0: batch_col_nw<T, N_, false, false, false, 0>(L, st); break;            \
  case N_ * 8 + 1: batch_col_nw<T, N_, true, true, true, 0>(L, st); break;               \
  case N_ * 8 + 2: batch_col_nw<T, N_, true, false, false, 0>(L, st); break;             \
  case N_ * 8 + 4: batch_col_nw<T, N_, false, false, false, 1>(L, st); break;            \
  case N_ * 8 + 5: batch_col_nw<T, N_, true, true, true, 1>(L, st); break;               \
  case N_ * 8 + 6: batch_col_nw<T, N_, true, false, false, 1>(L, st); break;
          FPM_BCOL_LENGTHS(FPM_CASE)
#undef FPM_CASE
          default: ok = false;
        }
        break;
      }
      default: ok = false;
    }
    if (!ok) return FPM_ERR_UNSUPPORTED;
    FPM_LAUNCH_CHECK();
  }
  return FPM_OK;
}

}  // namespace fpm

using namespace fpm;

extern "C" {

int fpm_fft_supported(int n) { return supported_len(n) ? 1 : 0; }

int fpm_fft_set_tiling(int elem_bytes, int n, int wide) {
  if ((elem_bytes != 4 && elem_bytes != 8) || !supported_len(n) || (wide != 0 && wide != 1))
    return FPM_ERR_ARG;
  g_col_wide[elem_bytes == 8 ? 1 : 0][n] = wide;
  return FPM_OK;
}

int fpm_fft_get_tiling(int elem_bytes, int n) {
  if ((elem_bytes != 4 && elem_bytes != 8) || !supported_len(n)) return FPM_ERR_ARG;
  return col_wide(elem_bytes, n);
}
int fpm_twiddles_f32(int n, float* o) { return twiddles_impl<float>(n, o); }
int fpm_twiddles_f64(int n, double* o) { return twiddles_impl<double>(n, o); }

int fpm_fft_r2c_f32(const void* in, void* out, const fpm_mesh* mesh, int64_t pitch2,
                    const void* tw0, const void* tw1, const void* tw2, fpm_stream_t s) {
  return r2c_impl<float>(in, out, mesh, pitch2, tw0, tw1, tw2, (hipStream_t)s);
}
int fpm_fft_r2c_f64(const void* in, void* out, const fpm_mesh* mesh, int64_t pitch2,
                    const void* tw0, const void* tw1, const void* tw2, fpm_stream_t s) {
  return r2c_impl<double>(in, out, mesh, pitch2, tw0, tw1, tw2, (hipStream_t)s);
}
int fpm_fft_r2c_c2r_f32(const void* in, void* kout, int ncomp, const fpm_kop* ops,
                        void* const* outs, void* const* work, const fpm_mesh* mesh, int64_t pitch2,
                        const void* tw0, const void* tw1, const void* tw2, fpm_stream_t s) {
  return r2c_c2r_impl<float>(in, kout, ncomp, ops, outs, work, mesh, pitch2, tw0, tw1, tw2,
                             (hipStream_t)s);
}
int fpm_fft_r2c_c2r_f64(const void* in, void* kout, int ncomp, const fpm_kop* ops,
                        void* const* outs, void* const* work, const fpm_mesh* mesh, int64_t pitch2,
                        const void* tw0, const void* tw1, const void* tw2, fpm_stream_t s) {
  return r2c_c2r_impl<double>(in, kout, ncomp, ops, outs, work, mesh, pitch2, tw0, tw1, tw2,
                              (hipStream_t)s);
}
int fpm_fft_c2r_f32(const void* in, int ncomp, const fpm_kop* ops, void* const* outs,
                    void* const* work, const fpm_mesh* mesh, int64_t pitch2, const void* tw0,
                    const void* tw1, const void* tw2, fpm_stream_t s) {
  return c2r_impl<float>(in, ncomp, ops, outs, work, mesh, pitch2, tw0, tw1, tw2, (hipStream_t)s);
}
int fpm_fft_c2r_f64(const void* in, int ncomp, const fpm_kop* ops, void* const* outs,
                    void* const* work, const fpm_mesh* mesh, int64_t pitch2, const void* tw0,
                    const void* tw1, const void* tw2, fpm_stream_t s) {
  return c2r_impl<double>(in, ncomp, ops, outs, work, mesh, pitch2, tw0, tw1, tw2, (hipStream_t)s);
}
int fpm_fft_c2r_grad3_f32(const void* in, const fpm_kop* common, void* const* outs,
                          void* const* work, const fpm_mesh* mesh, int64_t pitch2, const void* tw0,
                          const void* tw1, const void* tw2, fpm_stream_t s) {
  return c2r_grad3_impl<float>(in, common, outs, work, mesh, pitch2, tw0, tw1, tw2, (hipStream_t)s);
}
int fpm_fft_c2r_grad3_f64(const void* in, const fpm_kop* common, void* const* outs,
                          void* const* work, const fpm_mesh* mesh, int64_t pitch2, const void* tw0,
                          const void* tw1, const void* tw2, fpm_stream_t s) {
  return c2r_grad3_impl<double>(in, common, outs, work, mesh, pitch2, tw0, tw1, tw2,
                                (hipStream_t)s);
}
int fpm_kspace_apply_f32(void* f, const fpm_kop* op, const fpm_mesh* m, int64_t p, fpm_stream_t s) {
  return kspace_apply_impl<float>(f, op, m, p, (hipStream_t)s);
}
int fpm_kspace_apply_f64(void* f, const fpm_kop* op, const fpm_mesh* m, int64_t p, fpm_stream_t s) {
  return kspace_apply_impl<double>(f, op, m, p, (hipStream_t)s);
}

}  // extern "C"
