// Device-side FFT building blocks for gfx950 (CDNA4, wave64).
//
// Complex values are stored interleaved; every transform length is
// L = 3 * 2^k (the 3/2-padded grid M = 3N/2).  A transform is a short
// sequence of Stockham autosort passes (radix 3*2^a first in the inverse
// direction, powers of two after it; the forward direction runs the same plan
// reversed).  Each pass keeps whole radix-R butterflies in registers and uses
// LDS only for the exchange between passes.  The passes at the ends of a
// transform are split out so that kernels can feed them from / drain them to
// HBM or registers directly (see palx.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <utility>

namespace palx {

template <typename R>
struct alignas(2 * sizeof(R)) cx {
  R x, y;
};

template <typename R>
__device__ __forceinline__ cx<R> cadd(cx<R> a, cx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ cx<R> csub(cx<R> a, cx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R>
__device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__device__ __forceinline__ cx<R> cscale(cx<R> a, R s) { return {a.x * s, a.y * s}; }
template <typename R>
__device__ __forceinline__ cx<R> cconj(cx<R> a) { return {a.x, -a.y}; }
// multiply by sign*i  (sign=-1 forward, +1 inverse)
template <bool INV, typename R>
__device__ __forceinline__ cx<R> cmul_si(cx<R> a) {
  return INV ? cx<R>{-a.y, a.x} : cx<R>{a.y, -a.x};
}

// ---------------------------------------------------------------- constexpr trig
// cos/sin(2*pi*p/q) evaluated at compile time (octant reduction + Taylor).
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double ce_sin_t(double x) {
  double term = x, sum = x;
  for (int n = 1; n < 14; ++n) {
    term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double ce_cos_t(double x) {
  double term = 1.0, sum = 1.0;
  for (int n = 1; n < 14; ++n) {
    term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
    sum += term;
  }
  return sum;
}
// angle = 2*pi*p/q; returns cos (which=0) or sin (which=1)
constexpr double ce_trig(long p, long q, int which) {
  long m = ((p % q) + q) % q;  // [0, q)
  // express angle in units of 1/(8q) turns: a = 8m (octants of q)
  long a = 8 * m;
  int oct = (int)(a / q);           // 0..7
  long rem = a - (long)oct * q;     // [0, q)
  // base angle within octant, in turns: rem / (8q)
  double c = 0, s = 0;
  double t_lo = kTwoPi * (double)rem / (8.0 * (double)q);           // [0, pi/4)
  double t_hi = kTwoPi * (double)(q - rem) / (8.0 * (double)q);     // (0, pi/4]
  switch (oct) {
    case 0: c = ce_cos_t(t_lo); s = ce_sin_t(t_lo); break;
    case 1: c = ce_sin_t(t_hi); s = ce_cos_t(t_hi); break;
    case 2: c = -ce_sin_t(t_lo); s = ce_cos_t(t_lo); break;
    case 3: c = -ce_cos_t(t_hi); s = ce_sin_t(t_hi); break;
    case 4: c = -ce_cos_t(t_lo); s = -ce_sin_t(t_lo); break;
    case 5: c = -ce_sin_t(t_hi); s = -ce_cos_t(t_hi); break;
    case 6: c = ce_sin_t(t_lo); s = -ce_cos_t(t_lo); break;
    default: c = ce_cos_t(t_hi); s = -ce_sin_t(t_hi); break;
  }
  return which == 0 ? c : s;
}

// ---------------------------------------------------------------- register DFTs
// In-place DFT of R values in natural order; INV selects exp(+2 pi i nk/R).
template <int R, bool INV, typename T>
struct DFT;

template <bool INV, typename T>
struct DFT<1, INV, T> {
  static __device__ __forceinline__ void run(cx<T>*) {}
};

template <bool INV, typename T>
struct DFT<2, INV, T> {
  static __device__ __forceinline__ void run(cx<T>* v) {
    cx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

// a + k*b and a - k*b as two fused multiply-adds per component
template <typename T>
__device__ __forceinline__ void fma_pair(cx<T> a, cx<T> b, T k, cx<T>& plus, cx<T>& minus) {
  plus = cx<T>{fma(k, b.x, a.x), fma(k, b.y, a.y)};
  minus = cx<T>{fma(-k, b.x, a.x), fma(-k, b.y, a.y)};
}

template <bool INV, typename T>
struct DFT<3, INV, T> {
  static __device__ __forceinline__ void run(cx<T>* v) {
    constexpr T c = (T)-0.5;
    constexpr T d = (T)0.86602540378443864676372317075294;
    const cx<T> t1 = cadd(v[1], v[2]);
    const cx<T> t2 = cmul_si<INV>(csub(v[1], v[2]));  // +-i (v1 - v2), sign flips only
    const cx<T> m = cx<T>{fma(c, t1.x, v[0].x), fma(c, t1.y, v[0].y)};
    v[0] = cadd(v[0], t1);
    fma_pair(m, t2, d, v[1], v[2]);
  }
};

template <bool INV, typename T>
struct DFT<4, INV, T> {
  static __device__ __forceinline__ void run(cx<T>* v) {
    cx<T> a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    cx<T> c = cadd(v[1], v[3]), d = cmul_si<INV>(csub(v[1], v[3]));
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
  }
};

template <bool INV, typename T>
struct DFT<8, INV, T> {
  static __device__ __forceinline__ void run(cx<T>* v) {
    cx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cx<T> o[4] = {v[1], v[3], v[5], v[7]};
    DFT<4, INV, T>::run(e);
    DFT<4, INV, T>::run(o);
    constexpr T h = (T)0.70710678118654752440084436210485;
    // W8^1 = (1 + s i)/sqrt2, W8^2 = s i, W8^3 = (-1 + s i)/sqrt2, s = +1 (INV) / -1;
    // the 1/sqrt2 goes into fused multiply-adds with e[1], e[3]
    const cx<T> o1 = cadd(o[1], cmul_si<INV>(o[1]));
    const cx<T> o2 = cmul_si<INV>(o[2]);
    const cx<T> o3 = csub(cmul_si<INV>(o[3]), o[3]);
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    fma_pair(e[1], o1, h, v[1], v[5]);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    fma_pair(e[3], o3, h, v[3], v[7]);
  }
};

// Composite R = R1*R2 (Cooley-Tukey in registers):
//   y[k1 + R1 k2] = sum_n2 W_R2^{n2 k2} W_R^{n2 k1} sum_n1 x[n1 R2 + n2] W_R1^{n1 k1}
// Twiddle exponents are template parameters so every constant folds.
template <int R, int P, bool INV, typename T>
__device__ __forceinline__ void twiddle_const(cx<T>& a) {
  constexpr int p = P % R;
  if constexpr (p == 0) {
  } else if constexpr ((4 * p) % R == 0) {
    constexpr int q = (4 * p) / R;  // 1: W = s i, 2: -1, 3: -s i
    if constexpr (q == 1) a = cmul_si<INV>(a);
    else if constexpr (q == 2) a = cx<T>{-a.x, -a.y};
    else { cx<T> b = cmul_si<INV>(a); a = cx<T>{-b.x, -b.y}; }
  } else {
    constexpr T c = (T)ce_trig(p, R, 0);
    constexpr T s = (T)(INV ? ce_trig(p, R, 1) : -ce_trig(p, R, 1));
    a = cmul(a, cx<T>{c, s});
  }
}
template <int R, int N2, bool INV, typename T, int... K1>
__device__ __forceinline__ void twiddle_row(cx<T>* a, std::integer_sequence<int, K1...>) {
  (twiddle_const<R, N2 * K1, INV, T>(a[K1]), ...);
}
template <int R1, int R2, bool INV, typename T, int N2>
__device__ __forceinline__ void comp_col(cx<T>* v, cx<T>* a) {
#pragma unroll
  for (int n1 = 0; n1 < R1; ++n1) a[n1] = v[n1 * R2 + N2];
  DFT<R1, INV, T>::run(a);
  twiddle_row<R1 * R2, N2, INV, T>(a, std::make_integer_sequence<int, R1>{});
}
template <int R1, int R2, bool INV, typename T, int... N2>
__device__ __forceinline__ void comp_cols(cx<T>* v, cx<T> (*a)[R1], std::integer_sequence<int, N2...>) {
  (comp_col<R1, R2, INV, T, N2>(v, a[N2]), ...);
}
template <int R1, int R2, bool INV, typename T>
__device__ __forceinline__ void dft_comp(cx<T>* v) {
  cx<T> a[R2][R1];
  comp_cols<R1, R2, INV, T>(v, a, std::make_integer_sequence<int, R2>{});
#pragma unroll
  for (int k1 = 0; k1 < R1; ++k1) {
    cx<T> b[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; ++n2) b[n2] = a[n2][k1];
    DFT<R2, INV, T>::run(b);
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = b[k2];
  }
}

#define PALX_COMPOSITE(R, R1, R2)                                                 \
  template <bool INV, typename T>                                                 \
  struct DFT<R, INV, T> {                                                         \
    static __device__ __forceinline__ void run(cx<T>* v) { dft_comp<R1, R2, INV, T>(v); } \
  };
PALX_COMPOSITE(6, 3, 2)
PALX_COMPOSITE(12, 3, 4)
PALX_COMPOSITE(16, 4, 4)
PALX_COMPOSITE(24, 3, 8)
PALX_COMPOSITE(32, 4, 8)
#undef PALX_COMPOSITE

// ---------------------------------------------------------------- plans
// Per length L and variant V: radices g[0..n), T, the threads that
// cooperate on one transform, and the LDS layout (pad).  g[0] is a multiple
// of 3, so that the zero middle third of a 3/2-padded spectrum is a whole
// third of each first-pass butterfly; for L = 9*2^k (grids N = 3*2^k) a
// later radix carries the second factor 3.  Inverse transforms run g in
// order, forward transforms run it reversed.  V = 3*alt + k: alt 0 is the
// base plan, alt 1 the alternative (L = 3*2^k >= 384); k = 0: padded LDS
// (pad 1 below), k = 1: unpadded, k = 2: padded with the column pass loading
// its second field during the first field's transform (pf; more registers,
// less exposed load latency).  The autotuner picks among the variants.
struct Plan {
  int n;
  int g[4];
  int T;
  int pad;
  int pf;
  int sp = 0;  // split exchange: LDS holds real parts, then imaginary parts
};
constexpr int kVariants = 8;
constexpr Plan base_plan(int L, int alt) {
  return L == 12     ? Plan{2, {3, 4, 1, 1}, 4, 1, 0}
         : L == 24   ? Plan{2, {3, 8, 1, 1}, 8, 1, 0}
         : L == 48   ? Plan{2, {6, 8, 1, 1}, 8, 1, 0}
         : L == 96   ? Plan{2, {12, 8, 1, 1}, 8, 1, 0}
         : L == 192  ? Plan{2, {12, 16, 1, 1}, 16, 1, 0}
         : L == 384  ? (alt ? Plan{3, {6, 8, 8, 1}, 64, 1, 0} : Plan{2, {24, 16, 1, 1}, 16, 1, 0})
         : L == 768  ? (alt ? Plan{2, {24, 32, 1, 1}, 32, 1, 0} : Plan{3, {12, 8, 8, 1}, 64, 1, 0})
         : L == 1536 ? (alt ? Plan{3, {6, 16, 16, 1}, 256, 1, 0} : Plan{3, {24, 8, 8, 1}, 192, 1, 0})
         : L == 3072 ? (alt ? Plan{4, {6, 8, 8, 8}, 512, 1, 0} : Plan{3, {12, 16, 16, 1}, 256, 1, 0})
         : L == 6144 ? (alt ? Plan{4, {12, 8, 8, 8}, 512, 1, 0} : Plan{3, {24, 16, 16, 1}, 384, 1, 0})
         // FP64 only with the split exchange: a complex double transform needs 192 KB of LDS
         : L == 12288 ? (alt ? Plan{4, {24, 8, 8, 8}, 512, 1, 0} : Plan{4, {12, 16, 8, 8}, 768, 1, 0})
         // FP32 only and split exchange only (N = 16384): a complex float
         // transform needs 196 KB of LDS, its split form 99 KB
         : L == 24576 ? (alt ? Plan{4, {24, 8, 8, 16}, 1024, 1, 0} : Plan{4, {24, 16, 8, 8}, 1024, 1, 0})
         // L = 9*2^k: padded lengths of the grids N = 3*2^k (base plan only)
         : L == 72   ? Plan{2, {12, 6, 1, 1}, 6, 1, 0}
         : L == 144  ? Plan{2, {12, 12, 1, 1}, 12, 1, 0}
         : L == 288  ? Plan{2, {24, 12, 1, 1}, 12, 1, 0}
         : L == 576  ? Plan{2, {24, 24, 1, 1}, 24, 1, 0}
         : L == 1152 ? Plan{3, {12, 12, 8, 1}, 96, 1, 0}
         : L == 2304 ? Plan{3, {12, 12, 16, 1}, 192, 1, 0}
         : L == 4608 ? Plan{4, {12, 12, 8, 4}, 384, 1, 0}
         : L == 9216 ? Plan{4, {12, 12, 8, 8}, 768, 1, 0}
         // N = 12288: 144 KB of LDS per complex float transform, 148 KB per
         // double one with the split exchange (variant 6 only in FP64)
         : L == 18432 ? Plan{4, {24, 12, 8, 8}, 768, 1, 0}
                     : Plan{0, {1, 1, 1, 1}, 0, 1, 0};
}
// L = 3*2^k (the padded lengths of power-of-two grids)
constexpr bool three_pow2(int L) { return L % 3 == 0 && ((L / 3) & (L / 3 - 1)) == 0; }
// whether L has the alternative plan (variants 3-5, 7): the lengths 3*2^k >= 384
constexpr bool has_alt(int L) { return L >= 384 && three_pow2(L); }
// variant numbers in use at L: 8 where the alternative plan or the split
// exchange exists (the long 9*2^k lengths, 9216 and up, have the base plan in
// variants 0-2 and 6 only), else the base plan's three layouts
constexpr int n_variants(int L) { return has_alt(L) || L >= 9216 ? 8 : 3; }
// V 0-5: plan alt = V/3, layout k = V%3 (above); V 6, 7: plan alt = V-6 with
// the split exchange (padded, no prefetch), which halves the LDS of a
// transform at the cost of two more barriers per exchange
constexpr Plan plan_of(int L, int V = 0) {
  if (V >= 6) {
    Plan p = base_plan(L, V - 6);
    p.pad = 1;
    p.pf = 0;
    p.sp = 1;
    return p;
  }
  Plan p = base_plan(L, V >= 3 ? 1 : 0);
  p.pad = (V % 3) == 1 ? 0 : 1;
  p.pf = (V % 3) == 2 ? 1 : 0;
  return p;
}
// radix of pass i (0-based) in direction INV, and the sub-transform size NS
// before it (product of the radices already applied)
constexpr int pass_radix(Plan P, bool inv, int i) { return inv ? P.g[i] : P.g[P.n - 1 - i]; }
constexpr int pass_ns(Plan P, bool inv, int i) {
  int ns = 1;
  for (int q = 0; q < i; ++q) ns *= pass_radix(P, inv, q);
  return ns;
}
// largest power of two <= x
constexpr int pow2_floor(int x) {
  int p = 1;
  while (2 * p <= x) p *= 2;
  return p;
}

// LDS element index of logical position i in the state written by a pass of
// radix RW at sub-transform size NSW (a pass writes positions
// (j/NSW)*NSW*RW + r*NSW + j%NSW, i.e. runs of NSW consecutive lanes that
// jump by NSW*RW):
//   pad 0: i (no padding)
//   pad 1: one slot every 32 elements
//   pad 2: NSW slots after every NSW*RW elements when NSW < 16, so the runs
//          of one write start on different banks (stride-R writes of the
//          first pass become stride R+1); runs of >= 16 lanes need none.
//          (Measured no faster than pad 0/1 on gfx950; kept selectable.)
constexpr int lay_index(int pad, int RW, int NSW, int i) {
  return pad == 0 ? i : pad == 1 ? i + (i >> 5) : (NSW >= 16 ? i : i + (i / (NSW * RW)) * NSW);
}
template <int PAD, int RW, int NSW>
__device__ __forceinline__ int lay(int i) { return lay_index(PAD, RW, NSW, i); }
// index of the Hermitian-split exchange of the row pass (not a Stockham state)
template <int PAD>
__device__ __forceinline__ int split_idx(int i) { return PAD == 0 ? i : i + (i >> 5); }
// elements of one transform's LDS buffer: the largest index of every state
// of both directions, and of the split exchange
constexpr int plan_lds_len(int L, int V) {
  const Plan P = plan_of(L, V);
  int m = P.pad == 0 ? L : L + (L >> 5) + 1;
  for (int d = 0; d < 2; ++d)
    for (int p = 0; p + 1 < P.n; ++p) {
      const int e = lay_index(P.pad, pass_radix(P, d == 0, p), pass_ns(P, d == 0, p), L - 1) + 1;
      if (e > m) m = e;
    }
  return m;
}

// LDS of one transform in complex elements: plan_lds_len complex values, or
// for the split exchange plan_lds_len reals (rounded up to whole complex)
constexpr int buf_cx_len(int L, int V) {
  return plan_of(L, V).sp ? (plan_lds_len(L, V) + 1) / 2 : plan_lds_len(L, V);
}

// Twiddle exp(-/+ 2 pi i m / L) from a two-level table staged in LDS:
// tws[a] = W^a (a < 64), tws[64 + b] = W^(64 b) (b < L/64), W = exp(-2 pi i / L).
constexpr int tw_len(int L) { return 64 + (L + 63) / 64; }

// Transform groups of T threads per block for complex elements of csz bytes
// (<= 256 threads unless one group needs more; LDS kept <= 64 KB when more
// than one group), the LDS of one block, and whether variant V of length L
// is compiled: it exists for L and its block fits the 160 KB a block may use.
constexpr int groups_per_block_cs(int L, int V, int csz) {
  const int T = plan_of(L, V).T;
  int g = 256 / T;
  if (g < 1) g = 1;
  while (g > 1 && (long)g * buf_cx_len(L, V) * (long)csz > 65536) g /= 2;
  return g;
}
constexpr long block_lds_bytes(int L, int V, int csz) {
  return ((long)groups_per_block_cs(L, V, csz) * buf_cx_len(L, V) + tw_len(L)) * csz;
}
constexpr long kMaxBlockLds = 163840;
constexpr bool variant_fits(int L, int V, int csz) {
  return V >= 0 && V < n_variants(L) && (has_alt(L) || V <= 2 || V == 6) &&
         block_lds_bytes(L, V, csz) <= kMaxBlockLds;
}
// V where it fits, else the split exchange of the base plan (the only form in
// which the longest FP32 lengths fit)
constexpr int fitting_variant(int L, int V, int csz) { return variant_fits(L, V, csz) ? V : 6; }

template <bool INV, typename T>
__device__ __forceinline__ cx<T> twid(const cx<T>* tws, int m) {
  const cx<T> w = cmul(tws[m & 63], tws[64 + (m >> 6)]);
  return INV ? cconj(w) : w;
}

// Multiply v[1..R-1] by w^1..w^(R-1), w = W^m, from two-level table lookups
// (two LDS reads and a product each) and products of those.
//   float:  the powers w^(2^b) are lookups, the others products of the
//           lookups along the set bits of r, so every power is at most
//           popcount(r) products away from table values (no error growth
//           with the exponent): log2(R) + 1 lookups per butterfly.  The
//           cheaper schemes below moved the FP32 path's error against FP64
//           by up to 40% (N=1024) for 1-3% of step time.
//   double: one lookup of w, then w^r = w^(r/2) w^(r - r/2) (product depth
//           ceil(log2 r): a few ulp, far below the FP64 parity bar): one
//           lookup per butterfly, fewer LDS reads and registers (N=4096
//           RK3 step 2.40 -> 2.29 ms).
template <int R, bool INV, typename T>
__device__ __forceinline__ void twiddle_butterfly_inputs(cx<T>* v, const cx<T>* tws, int m) {
  if constexpr (R > 1 && sizeof(T) == 8) {
    cx<T> pw[R];
    pw[1] = twid<INV>(tws, m);
#pragma unroll
    for (int r = 2; r < R; ++r) pw[r] = cmul(pw[r / 2], pw[r - r / 2]);
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], pw[r]);
  } else if constexpr (R > 1) {
    constexpr int NB = R <= 2 ? 1 : R <= 4 ? 2 : R <= 8 ? 3 : R <= 16 ? 4 : 5;
    cx<T> pw[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) pw[b] = twid<INV>(tws, m << b);
#pragma unroll
    for (int r = 1; r < R; ++r) {
      int b = 0;
      while (!((r >> b) & 1)) ++b;
      cx<T> w = pw[b];
#pragma unroll
      for (int c = b + 1; c < NB; ++c)
        if ((r >> c) & 1) w = cmul(w, pw[c]);
      v[r] = cmul(v[r], w);
    }
  }
}

// fill the LDS twiddle table of one block from the FP64-accurate global table
template <int L, typename T>
__device__ __forceinline__ void load_twiddles(cx<T>* tws, const cx<T>* __restrict__ tw) {
  for (int i = threadIdx.x; i < tw_len(L); i += blockDim.x) {
    const int m = i < 64 ? (i < L ? i : 0) : 64 * (i - 64);
    tws[i] = (i < 64 && i >= L) ? cx<T>{1, 0} : tw[m < L ? m : 0];
  }
}

// Stockham pass LDS -> LDS.  Pass with radix R at sub-transform size NS; it
// reads the state written by the pass (RP, NSP) and writes its own.
template <int L, int R, int NS, int T, bool INV, int PAD, int RP, int NSP, typename Tr>
__device__ __forceinline__ void pass_lds(cx<Tr>* buf, const cx<Tr>* tw, int t) {
  constexpr int NB = L / R;
  constexpr int PER = (NB + T - 1) / T;
  cx<Tr> v[PER][R];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int j = t + i * T;
    if (NB % T == 0 || j < NB) {
      const int k = j % NS;
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[lay<PAD, RP, NSP>(j + r * NB)];
      if (NS > 1) twiddle_butterfly_inputs<R, INV>(v[i], tw, k * (L / (NS * R)));
      DFT<R, INV, Tr>::run(v[i]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int j = t + i * T;
    if (NB % T == 0 || j < NB) {
      const int k = j % NS;
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[lay<PAD, R, NS>(base + r * NS)] = v[i][r];
    }
  }
  __syncthreads();
}

// Last pass (NS = L/R), LDS -> registers; v[i][r] is output position j + r*L/R.
// Reads the state written by the pass (RP, NSP).
template <int L, int R, int T, bool INV, int PAD, int RP, int NSP, typename Tr>
__device__ __forceinline__ void pass_last_to_regs(const cx<Tr>* buf, const cx<Tr>* tw,
                                                  int t, cx<Tr> (*v)[R]) {
  constexpr int NB = L / R;
  constexpr int PER = (NB + T - 1) / T;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int j = t + i * T;
    if (NB % T == 0 || j < NB) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[lay<PAD, RP, NSP>(j + r * NB)];
      twiddle_butterfly_inputs<R, INV>(v[i], tw, j);
      DFT<R, INV, Tr>::run(v[i]);
    }
  }
}

// First pass (NS = 1) from registers; v[i][r] holds input position j + r*L/R.
// Caller guarantees nobody still reads buf, and barriers afterwards.
template <int L, int R, int T, bool INV, int PAD, typename Tr>
__device__ __forceinline__ void pass_first_from_regs(cx<Tr>* buf, int t, cx<Tr> (*v)[R]) {
  constexpr int NB = L / R;
  constexpr int PER = (NB + T - 1) / T;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int j = t + i * T;
    if (NB % T == 0 || j < NB) {
      DFT<R, INV, Tr>::run(v[i]);
#pragma unroll
      for (int r = 0; r < R; ++r) buf[lay<PAD, R, 1>(j * R + r)] = v[i][r];
    }
  }
}

// A whole transform of variant V in direction INV, in three calls: the first
// pass from registers, the LDS -> LDS passes between (each ending with a
// barrier), the last pass to registers.
template <int L, int V, bool INV, typename Tr>
__device__ __forceinline__ void first_pass(cx<Tr>* buf, int t,
                                           cx<Tr> (*v)[pass_radix(plan_of(L, V), INV, 0)]) {
  constexpr Plan P = plan_of(L, V);
  pass_first_from_regs<L, pass_radix(P, INV, 0), P.T, INV, P.pad>(buf, t, v);
}

template <int L, int V, bool INV, typename Tr>
__device__ __forceinline__ void mid_passes(cx<Tr>* buf, const cx<Tr>* tws, int t) {
  constexpr Plan P = plan_of(L, V);
  if constexpr (P.n >= 3)
    pass_lds<L, pass_radix(P, INV, 1), pass_ns(P, INV, 1), P.T, INV, P.pad, pass_radix(P, INV, 0),
             pass_ns(P, INV, 0)>(buf, tws, t);
  if constexpr (P.n >= 4)
    pass_lds<L, pass_radix(P, INV, 2), pass_ns(P, INV, 2), P.T, INV, P.pad, pass_radix(P, INV, 1),
             pass_ns(P, INV, 1)>(buf, tws, t);
}

template <int L, int V, bool INV, typename Tr>
__device__ __forceinline__ void last_pass(const cx<Tr>* buf, const cx<Tr>* tws, int t,
                                          cx<Tr> (*v)[pass_radix(plan_of(L, V), INV, plan_of(L, V).n - 1)]) {
  constexpr Plan P = plan_of(L, V);
  pass_last_to_regs<L, pass_radix(P, INV, P.n - 1), P.T, INV, P.pad, pass_radix(P, INV, P.n - 2),
                    pass_ns(P, INV, P.n - 2)>(buf, tws, t, v);
}

template <int L, int R>
constexpr int per_of(int T) {
  return (L / R + T - 1) / T;
}

// ---------------------------------------------------------------- split exchange
// A whole transform with the state in registers between passes; each
// exchange moves the real parts through the LDS buffer (Tr elements), then
// the imaginary parts, so the buffer is half the size of a complex state.
// Pass p writes its outputs at (j - k) R + k + r NS (k = j mod NS) and the
// next pass reads j' + r' L/R'; the LDS index of both is lay<PAD, R, NS>.
template <int L, int RW, int NSW, int RR, int T, int PAD, bool IM, typename Tr>
__device__ __forceinline__ void split_half(Tr* b, int t, const cx<Tr> (*w)[RW], cx<Tr> (*r)[RR]) {
  constexpr int NBW = L / RW, PERW = (NBW + T - 1) / T;
  constexpr int NBR = L / RR, PERR = (NBR + T - 1) / T;
#pragma unroll
  for (int i = 0; i < PERW; ++i) {
    const int j = t + i * T;
    if (NBW % T == 0 || j < NBW) {
      const int k = j % NSW;
      const int base = (j - k) * RW + k;
#pragma unroll
      for (int q = 0; q < RW; ++q) b[lay<PAD, RW, NSW>(base + q * NSW)] = IM ? w[i][q].y : w[i][q].x;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PERR; ++i) {
    const int j = t + i * T;
    if (NBR % T == 0 || j < NBR) {
#pragma unroll
      for (int q = 0; q < RR; ++q) {
        const Tr x = b[lay<PAD, RW, NSW>(j + q * NBR)];
        if (IM) r[i][q].y = x;
        else r[i][q].x = x;
      }
    }
  }
}

template <int L, int V, bool INV, int p, typename Tr>
struct SplitPass {
  static constexpr Plan P = plan_of(L, V);
  static constexpr int R = pass_radix(P, INV, p), NS = pass_ns(P, INV, p);
  static constexpr int NB = L / R, PER = (NB + P.T - 1) / P.T;
  static constexpr int RL = pass_radix(P, INV, P.n - 1);
  // v: inputs of pass p (position j + r NB); out: outputs of the last pass
  static __device__ __forceinline__ void run(Tr* b, const cx<Tr>* tws, int t, cx<Tr> (*v)[R],
                                             cx<Tr> (*out)[RL]) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int j = t + i * P.T;
      if (NB % P.T == 0 || j < NB) {
        if constexpr (p > 0) twiddle_butterfly_inputs<R, INV>(v[i], tws, (j % NS) * (L / (NS * R)));
        DFT<R, INV, Tr>::run(v[i]);
      }
    }
    if constexpr (p + 1 < P.n) {
      constexpr int R2 = pass_radix(P, INV, p + 1);
      constexpr int PER2 = (L / R2 + P.T - 1) / P.T;
      cx<Tr> nxt[PER2][R2];
      __syncthreads();  // earlier readers of the buffer are done
      split_half<L, R, NS, R2, P.T, P.pad, false>(b, t, v, nxt);
      __syncthreads();
      split_half<L, R, NS, R2, P.T, P.pad, true>(b, t, v, nxt);
      SplitPass<L, V, INV, p + 1, Tr>::run(b, tws, t, nxt, out);
    } else {
#pragma unroll
      for (int i = 0; i < PER; ++i)
#pragma unroll
        for (int q = 0; q < R; ++q) out[i][q] = v[i][q];
    }
  }
};

// The transform of plan V in direction INV from registers v (input position
// j + r L/R0) to registers out (output position j + r L/RL), for split plans.
template <int L, int V, bool INV, typename Tr>
__device__ __forceinline__ void transform_split(cx<Tr>* buf, const cx<Tr>* tws, int t,
                                                cx<Tr> (*v)[pass_radix(plan_of(L, V), INV, 0)],
                                                cx<Tr> (*out)[pass_radix(plan_of(L, V), INV, plan_of(L, V).n - 1)]) {
  SplitPass<L, V, INV, 0, Tr>::run(reinterpret_cast<Tr*>(buf), tws, t, v, out);
}

}  // namespace palx
