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
// lay(i + c) for a c that is a constant after unrolling.  With pad 1 and c a
// multiple of 32, (i + c) >> 5 = (i >> 5) + c / 32, so the padded index of
// every access of a butterfly is one shared term plus a constant that the
// LDS instruction takes as its offset (instead of a shift and two adds per
// access).
template <int PAD, int RW, int NSW>
__device__ __forceinline__ int lay_plus(int i, int c) {
  if (PAD == 1 && c % 32 == 0) return lay<PAD, RW, NSW>(i) + c + c / 32;
  return lay<PAD, RW, NSW>(i + c);
}
// LDS index of output q of butterfly j of a pass of radix RW at sub-transform
// size NSW: position base + q NSW with base = (j - k) RW + k, k = j mod NSW
// (the state that pass writes).  Equal to lay_plus<PAD, RW, NSW>(base, q NSW);
// with pad 1 the one-slot-per-32 term (i >> 5) is shared where that is exact,
// so that the index is one per-thread value plus a constant the LDS
// instruction takes as its offset (about three VALU instructions fewer per
// write):
//   NSW a power of two dividing 32 and NSW RW a multiple of 32: base - k is a
//     multiple of 32 and k + q NSW (k < NSW) never crosses a multiple of 32,
//     so the term is (base - k) / 32 + (q NSW) / 32;
//   NSW = 1: base = j RW is a multiple of A = gcd(RW, 32), and the positions
//     base + q with the same q - q mod A share the term.
constexpr int pow2_part_32(int x) {
  int a = 1;
  while (a < 32 && x % (2 * a) == 0) a *= 2;
  return a;
}
template <int PAD, int RW, int NSW>
__device__ __forceinline__ int lay_write(int base, int k, int q) {
  if constexpr (PAD == 1 && NSW <= 32 && 32 % NSW == 0 && (NSW * RW) % 32 == 0) {
    return base + ((base - k) >> 5) + q * NSW + ((q * NSW) >> 5);
  } else if constexpr (PAD == 1 && NSW == 1) {
    constexpr int A = pow2_part_32(RW);
    return lay_plus<PAD, RW, NSW>(base, q - q % A) + q % A;
  } else {
    return lay_plus<PAD, RW, NSW>(base, q * NSW);
  }
}
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

// Twiddles staged in LDS per block, W = exp(-2 pi i / L):
//   a two-level table, tws[a] = W^a (a < 64), tws[64 + b] = W^(64 b)
//   (b < L/64): W^m is the product of two reads;
//   per-pass tables for the first pass with twiddles of each direction (pass
//   1, sub-transform size NS = the first radix), where small: W^(S k r) for
//   r = 1..R-1, k < NS (S = L / (NS R)) at (r - 1) NS + k, so that each
//   twiddle is one read at a constant offset from k, and the NS distinct
//   values one instruction reads lie in consecutive slots (no bank
//   conflicts; the same exponents read from the second level hit a few
//   banks: k r multiples of 4 and 8).  At L = 6144 (plan 12, 8, 8, 8) they
//   hold 84 (inverse) and 56 (forward) entries.
// PT selects them per kernel: they trade product VALU work for LDS reads
// (seven distinct reads per radix-8 butterfly instead of three plus three
// broadcasts), which pays in a pass that waits little on LDS.
constexpr int kPassTableMax = 128;  // entries per direction
constexpr int tw_two_level_len(int L) { return 64 + (L + 63) / 64; }
constexpr int p1_entries(int L, int V, bool inv, bool PT = true) {
  const Plan P = plan_of(L, V);
  if (!PT || P.n < 2) return 0;
  const int e = pass_ns(P, inv, 1) * (pass_radix(P, inv, 1) - 1);
  return e <= kPassTableMax ? e : 0;
}
// offset of direction inv's pass table in the block's twiddles, -1 if none
constexpr int p1_offset(int L, int V, bool inv, bool PT = true) {
  return p1_entries(L, V, inv, PT) == 0 ? -1 : tw_two_level_len(L) + (inv ? 0 : p1_entries(L, V, true, PT));
}
constexpr int tw_len(int L, int V, bool PT = true) {
  return tw_two_level_len(L) + p1_entries(L, V, true, PT) + p1_entries(L, V, false, PT);
}

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
  return ((long)groups_per_block_cs(L, V, csz) * buf_cx_len(L, V) + tw_len(L, V)) * csz;
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
// W^(S2 k) for a compile-time S2: a multiple of 64 needs only the second
// level (its first-level factor is W^0 = 1), one read and no product
template <int S2, bool INV, typename T>
__device__ __forceinline__ cx<T> twid_k(const cx<T>* tws, int k) {
  if constexpr (S2 % 64 == 0) {
    const cx<T> w = tws[64 + k * (S2 / 64)];
    return INV ? cconj(w) : w;
  } else {
    return twid<INV>(tws, k * S2);
  }
}

// Multiply v[1..R-1] by w^1..w^(R-1), w = W^(S k), S the compile-time stride
// L / (NS R) of a pass at sub-transform size NS.  TB >= 0: the pass's own
// table at TB (one read per power, exact: the FP64 table rounded once).
// Otherwise from the two-level table (a lookup whose exponent is a multiple
// of 64 for every k is one read, twid_k):
//   float:  the powers w^(2^b) are lookups, the rest w^r = w^low(r)
//           w^(r - low(r)) (low = lowest set bit), so every power is at most
//           popcount(r) products away from table values (no error growth
//           with the exponent): log2(R) + 1 lookups per butterfly.  Chains
//           of products over the exponent moved the FP32 path's error
//           against FP64 by 40-200% (config-4 members: N=512 1.27e-6 ->
//           3.97e-6), lookups of w^(4^b) by up to 40%.
//   double: one lookup of w, then w^r = w^(r/2) w^(r - r/2) (product depth
//           ceil(log2 r): a few ulp, far below the FP64 parity bar): one
//           lookup per butterfly, fewer LDS reads and registers (N=4096
//           RK3 step 2.40 -> 2.29 ms).
template <int r, bool INV, int S, int TB, int NS, typename T>
__device__ __forceinline__ void twiddle_power(cx<T>* v, cx<T>* pw, const cx<T>* tws, int k) {
  constexpr int low = r & -r;
  if constexpr (TB >= 0) {
    const cx<T> w = tws[TB + (r - 1) * NS + k];
    pw[r] = INV ? cconj(w) : w;
  } else if constexpr (sizeof(T) == 8) {
    if constexpr (r == 1) pw[r] = twid_k<S, INV>(tws, k);
    else pw[r] = cmul(pw[r / 2], pw[r - r / 2]);
  } else {
    if constexpr (r == low) pw[r] = twid_k<S * r, INV>(tws, k);
    else pw[r] = cmul(pw[low], pw[r - low]);
  }
  v[r] = cmul(v[r], pw[r]);
}
template <bool INV, int S, int TB, int NS, typename T, int... Q>
__device__ __forceinline__ void twiddle_powers(cx<T>* v, cx<T>* pw, const cx<T>* tws, int k,
                                               std::integer_sequence<int, Q...>) {
  (twiddle_power<Q + 1, INV, S, TB, NS>(v, pw, tws, k), ...);
}
template <int R, bool INV, int S, int TB, int NS, typename T>
__device__ __forceinline__ void twiddle_butterfly_inputs(cx<T>* v, const cx<T>* tws, int k) {
  if constexpr (R > 1) {
    cx<T> pw[R];
    twiddle_powers<INV, S, TB, NS>(v, pw, tws, k, std::make_integer_sequence<int, R - 1>{});
  }
}

// exponent m of LDS twiddle slot i (variant V of length L), -1 for the unused
// first-level slots of L < 64 (they hold 1)
template <int L, int V, bool PT>
__device__ __forceinline__ int twiddle_exponent(int i) {
  constexpr Plan P = plan_of(L, V);
  constexpr int T2 = tw_two_level_len(L);
  if (i < 64) return i < L ? i : -1;
  if (i < T2) return 64 * (i - 64);
  const bool inv = i < T2 + p1_entries(L, V, true, PT);
  const int e = i - (inv ? p1_offset(L, V, true, PT) : p1_offset(L, V, false, PT));
  const int NS = inv ? pass_ns(P, true, 1) : pass_ns(P, false, 1);
  const int R = inv ? pass_radix(P, true, 1) : pass_radix(P, false, 1);
  return (L / (NS * R)) * (e % NS) * (e / NS + 1);
}

// The LDS twiddles of one block of NT threads, staged from the FP64-accurate
// global table tw[m] = W^m in two steps: fetch() issues this thread's loads
// into registers, commit() writes them to LDS.  A kernel fetches first, issues
// its HBM loads, and commits before the barrier that precedes the first pass
// with twiddles, so its HBM loads do not wait behind the table's round trip.
template <int L, int V, bool PT, int NT, typename T>
struct TwiddleStage {
  static constexpr int n = tw_len(L, V, PT);
  static constexpr int IT = (n + NT - 1) / NT;
  cx<T> r[IT];
  __device__ __forceinline__ void fetch(const cx<T>* __restrict__ tw) {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int i = (int)threadIdx.x + it * NT;
      r[it] = cx<T>{1, 0};
      if (n % NT == 0 || i < n) {
        const int m = twiddle_exponent<L, V, PT>(i);
        if (m >= 0) r[it] = tw[m];
      }
    }
  }
  __device__ __forceinline__ void commit(cx<T>* tws) const {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int i = (int)threadIdx.x + it * NT;
      if (n % NT == 0 || i < n) tws[i] = r[it];
    }
  }
};

// Stockham pass LDS -> LDS.  Pass with radix R at sub-transform size NS; it
// reads the state written by the pass (RP, NSP) and writes its own.  TB: its
// twiddle table's offset, or -1 (two-level table).
template <int L, int R, int NS, int T, bool INV, int PAD, int RP, int NSP, int TB, typename Tr>
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
      for (int r = 0; r < R; ++r) v[i][r] = buf[lay_plus<PAD, RP, NSP>(t, i * T + r * NB)];
      if (NS > 1) twiddle_butterfly_inputs<R, INV, L / (NS * R), TB, NS>(v[i], tw, k);
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
      for (int r = 0; r < R; ++r) buf[lay_write<PAD, R, NS>(base, k, r)] = v[i][r];
    }
  }
  __syncthreads();
}

// Last pass (NS = L/R), LDS -> registers; v[i][r] is output position j + r*L/R.
// Reads the state written by the pass (RP, NSP); TB as in pass_lds.
template <int L, int R, int T, bool INV, int PAD, int RP, int NSP, int TB, typename Tr>
__device__ __forceinline__ void pass_last_to_regs(const cx<Tr>* buf, const cx<Tr>* tw,
                                                  int t, cx<Tr> (*v)[R]) {
  constexpr int NB = L / R;
  constexpr int PER = (NB + T - 1) / T;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int j = t + i * T;
    if (NB % T == 0 || j < NB) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[lay_plus<PAD, RP, NSP>(t, i * T + r * NB)];
      twiddle_butterfly_inputs<R, INV, 1, TB, L / R>(v[i], tw, j);
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
      for (int r = 0; r < R; ++r) buf[lay_write<PAD, R, 1>(j * R, 0, r)] = v[i][r];
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

template <int L, int V, bool INV, bool PT, typename Tr>
__device__ __forceinline__ void mid_passes(cx<Tr>* buf, const cx<Tr>* tws, int t) {
  constexpr Plan P = plan_of(L, V);
  if constexpr (P.n >= 3)
    pass_lds<L, pass_radix(P, INV, 1), pass_ns(P, INV, 1), P.T, INV, P.pad, pass_radix(P, INV, 0),
             pass_ns(P, INV, 0), p1_offset(L, V, INV, PT)>(buf, tws, t);
  if constexpr (P.n >= 4)
    pass_lds<L, pass_radix(P, INV, 2), pass_ns(P, INV, 2), P.T, INV, P.pad, pass_radix(P, INV, 1),
             pass_ns(P, INV, 1), -1>(buf, tws, t);
}

template <int L, int V, bool INV, bool PT, typename Tr>
__device__ __forceinline__ void last_pass(const cx<Tr>* buf, const cx<Tr>* tws, int t,
                                          cx<Tr> (*v)[pass_radix(plan_of(L, V), INV, plan_of(L, V).n - 1)]) {
  constexpr Plan P = plan_of(L, V);
  pass_last_to_regs<L, pass_radix(P, INV, P.n - 1), P.T, INV, P.pad, pass_radix(P, INV, P.n - 2),
                    pass_ns(P, INV, P.n - 2), P.n == 2 ? p1_offset(L, V, INV, PT) : -1>(buf, tws, t, v);
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
      for (int q = 0; q < RW; ++q) b[lay_write<PAD, RW, NSW>(base, k, q)] = IM ? w[i][q].y : w[i][q].x;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PERR; ++i) {
    const int j = t + i * T;
    if (NBR % T == 0 || j < NBR) {
#pragma unroll
      for (int q = 0; q < RR; ++q) {
        const Tr x = b[lay_plus<PAD, RW, NSW>(t, i * T + q * NBR)];
        if (IM) r[i][q].y = x;
        else r[i][q].x = x;
      }
    }
  }
}

template <int L, int V, bool INV, bool PT, int p, typename Tr>
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
        if constexpr (p > 0)
          twiddle_butterfly_inputs<R, INV, L / (NS * R), p == 1 ? p1_offset(L, V, INV, PT) : -1, NS>(v[i], tws, j % NS);
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
      SplitPass<L, V, INV, PT, p + 1, Tr>::run(b, tws, t, nxt, out);
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
template <int L, int V, bool INV, bool PT, typename Tr>
__device__ __forceinline__ void transform_split(cx<Tr>* buf, const cx<Tr>* tws, int t,
                                                cx<Tr> (*v)[pass_radix(plan_of(L, V), INV, 0)],
                                                cx<Tr> (*out)[pass_radix(plan_of(L, V), INV, plan_of(L, V).n - 1)]) {
  SplitPass<L, V, INV, PT, 0, Tr>::run(reinterpret_cast<Tr*>(buf), tws, t, v, out);
}

}  // namespace palx
