// palx.hip -- MI355X (gfx950) kernels and C ABI for the palinstrophy DNS step.
//
// One stage of the reference pipeline (turbo_simulator.py: dns_step2b ->
// dns_step3 / RK3 stage -> dns_step2a) is two HBM passes here:
//
//   row pass (k_row):    for each padded physical row z (M rows)
//       read  V[c][z][kx<N/2]   (z-inverse-transformed u_hat, w_hat)
//       x-inverse of u + i w (one complex FFT gives both real rows),
//       square: (u + i w)^2 = (u^2 - w^2) + 2i uw,
//       x-forward, split Hermitian halves, keep kx < N/2,
//       write P[c][z][kx<N/2] = F_x[u^2 - w^2], F_x[uw]  (in place over V)
//   column pass (k_col): for each kx < N/2
//       z-forward of both P planes, keep the N retained kz (kz=-N/2 forced 0),
//       nonlinear term, implicit update of omega/fnm1, velocity
//       reconstruction, zero-padded z-inverse of u_hat, w_hat -> V.
//
// F[u^2] - F[w^2] of the reference is computed as F[u^2 - w^2] (same value up
// to rounding), which lets the row pass carry two planes instead of three.
// The padded zeros of the 3/2 rule (rows N/2..N-1 and kx >= N/2) are never
// stored: they are a known-zero middle third inside the radix-3 end passes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "palx.h"
#include "palx_fft.h"

using namespace palx;

// ============================================================================ kernels
namespace {

constexpr int kMaxGroupThreads = 512;


template <typename Tr>
constexpr int groups_per_block(int L, int T) {
  // groups of T threads per block (<= 256 threads unless one group needs more),
  // LDS kept <= 64 KB when more than one group
  int g = 256 / T;
  if (g < 1) g = 1;
  while (g > 1 && (long)g * lds_len<true>(L) * (long)sizeof(cx<Tr>) > 65536) g /= 2;
  return g;
}

enum RowMode { ROW_STEP = 0, ROW_PHYS = 1, ROW_PHYS1 = 2, ROW_ADJ = 3 };
enum ColMode {
  COL_CNAB2 = 0, COL_RK3 = 1, COL_INIT = 2, COL_OMEGA = 3, COL_STREAM = 4,
  COL_RECON = 5,  // velocity planes from omega (reconstruction + padded z-inverse)
  COL_ADJ = 6     // adjoint of one RK3 stage, column half
};
// flags of COL_ADJ, packed above the stage index
enum { ADJ_HAS_G = 16, ADJ_UPDATE = 32, ADJ_WANT_Q = 64 };

// ----------------------------------------------------------------- row pass
// x-inverse of one packed pair of mixed planes (A + iB, each the half
// spectrum of a real row): column-major planes [c][kx][z], element k of row
// `row` at k*M.  Follows pocketfft c2r semantics: the imaginary part of the
// kx=0 bin is dropped.  Result c[x] = a(x) + i b(x) at x = j + r*L/RL.
template <typename Tr, int L, bool ONE>
__device__ __forceinline__ void row_inverse(const cx<Tr>* __restrict__ Ap, const cx<Tr>* __restrict__ Bp,
                                            bool active, cx<Tr>* buf, const cx<Tr>* tws, int t,
                                            cx<Tr> (*out)[plan_of(L).g[plan_of(L).n - 1]]) {
  constexpr Plan P = plan_of(L);
  constexpr int T = P.T;
  constexpr int M = L, H = L / 3;
  constexpr int R1 = P.g[0];
  constexpr int RL = P.g[P.n - 1];
  constexpr int NB1 = L / R1, PER1 = (NB1 + T - 1) / T;
  {
    cx<Tr> v[PER1][R1];
#pragma unroll
    for (int i = 0; i < PER1; ++i) {
      const int j = t + i * T;
      if (NB1 % T == 0 || j < NB1) {
#pragma unroll
        for (int r = 0; r < R1; ++r) {
          const int pos = j + r * NB1;
          cx<Tr> c = {0, 0};
          if (active) {
            if (r < R1 / 3) {  // C[k] = A[k] + i B[k]
              const int k = pos;
              const cx<Tr> U = Ap[(size_t)k * M];
              const cx<Tr> Wv = ONE ? cx<Tr>{0, 0} : Bp[(size_t)k * M];
              c = (k == 0) ? cx<Tr>{U.x, Wv.x} : cx<Tr>{U.x - Wv.y, U.y + Wv.x};
            } else if (r >= 2 * R1 / 3) {  // C[M-k] = conj A[k] + i conj B[k]
              const int k2 = M - pos;
              if (k2 < H) {
                const cx<Tr> U = Ap[(size_t)k2 * M];
                const cx<Tr> Wv = ONE ? cx<Tr>{0, 0} : Bp[(size_t)k2 * M];
                c = cx<Tr>{U.x + Wv.y, Wv.x - U.y};
              }
            }
          }
          v[i][r] = c;
        }
      }
    }
    pass_first_from_regs<L, R1, T, true, true>(buf, t, v);
  }
  __syncthreads();
  if constexpr (P.n == 3) pass_lds<L, P.g[1], R1, T, true, true>(buf, tws, t);
  pass_last_to_regs<L, RL, T, true, true>(buf, tws, t, out);
}

// x-forward of a packed real pair g = a + i b (registers at x = j + r*L/RL,
// NS=1 first pass) and split of the Hermitian halves for kx < H:
//   Fa = (G[k] + conj G[M-k]) / 2,  Fb = (G[k] - conj G[M-k]) / (2i) * bscale.
// Written row-major [c][z][kx] into Pout.
template <typename Tr, int L>
__device__ __forceinline__ void row_forward_split(cx<Tr> (*v)[plan_of(L).g[plan_of(L).n - 1]], Tr bscale,
                                                  bool active, int row, cx<Tr>* buf, const cx<Tr>* tws,
                                                  int t, cx<Tr>* __restrict__ Pout) {
  constexpr Plan P = plan_of(L);
  constexpr int T = P.T;
  constexpr int M = L, H = L / 3;
  constexpr int R1 = P.g[0];
  constexpr int RL = P.g[P.n - 1];
  constexpr int NB1 = L / R1, PER1 = (NB1 + T - 1) / T;
  __syncthreads();  // everyone finished reading buf
  pass_first_from_regs<L, RL, T, false, true>(buf, t, v);
  __syncthreads();
  if constexpr (P.n == 3) pass_lds<L, P.g[1], RL, T, false, true>(buf, tws, t);
  cx<Tr> d[PER1][R1];
  pass_last_to_regs<L, R1, T, false, true>(buf, tws, t, d);
  __syncthreads();
  // retained outputs D[k] (k<H) -> buf[k];  D[M-k'] (k' in (0,H]) -> buf[2H-k']
#pragma unroll
  for (int i = 0; i < PER1; ++i) {
    const int j = t + i * T;
    if (NB1 % T == 0 || j < NB1) {
#pragma unroll
      for (int r = 0; r < R1; ++r) {
        const int pos = j + r * NB1;
        if (r < R1 / 3) buf[lidx<true>(pos)] = d[i][r];
        else if (r >= 2 * R1 / 3) buf[lidx<true>(pos - H)] = d[i][r];
      }
    }
  }
  __syncthreads();
  if (active) {
    const Tr hb = (Tr)0.5 * bscale;
    for (int k = t; k < H; k += T) {
      const cx<Tr> Dk = buf[lidx<true>(k)];
      const cx<Tr> Dm = buf[lidx<true>(k == 0 ? 0 : 2 * H - k)];
      const cx<Tr> A = {(Tr)0.5 * (Dk.x + Dm.x), (Tr)0.5 * (Dk.y - Dm.y)};
      const cx<Tr> B = {hb * (Dk.y + Dm.y), -hb * (Dk.x - Dm.x)};
      Pout[(size_t)row * H + k] = A;
      Pout[(size_t)M * H + (size_t)row * H + k] = B;
    }
  }
}

// W: velocity work planes, 2 x (H x M) complex column-major [c][kx][z].
// Q (ROW_ADJ): adjoint planes in the same layout.
// Pout: products / adjoint planes, 2 x (M x H) complex row-major [c][z][kx].
// ur: 3 planes of M x M real.  part: per-row max(|u|+|w|) at [z],
// sum(u^2+w^2) at [M + z] (ROW_PHYS).
template <typename Tr, int L, int MODE>
__global__ void __launch_bounds__(kMaxGroupThreads)
    k_row(const cx<Tr>* __restrict__ W, const cx<Tr>* __restrict__ Q, cx<Tr>* __restrict__ Pout,
          Tr* __restrict__ ur, double* __restrict__ part, const cx<Tr>* __restrict__ twg,
          int xcd_remap) {
  constexpr Plan P = plan_of(L);
  constexpr int T = P.T;
  constexpr int G = groups_per_block<Tr>(L, T);
  constexpr int M = L, H = L / 3;
  constexpr int RL = P.g[P.n - 1];
  constexpr int NBL = L / RL, PERL = (NBL + T - 1) / T;
  extern __shared__ unsigned char smem_raw[];
  cx<Tr>* smem = reinterpret_cast<cx<Tr>*>(smem_raw);

  int blk = blockIdx.x;
  if (xcd_remap) {  // consecutive rows (which read the same lines) on one XCD
    const int nb = gridDim.x;
    blk = (blk % 8) * (nb / 8) + blk / 8;
  }
  const int g = threadIdx.x / T, t = threadIdx.x % T;
  const int row = blk * G + g;
  const bool active = row < M;
  const int rows = active ? row : 0;
  cx<Tr>* buf = smem + g * lds_len<true>(L);
  cx<Tr>* tws = smem + G * lds_len<true>(L);
  load_twiddles<L>(tws, twg);

  // c = u + i w on the padded row (x = j + r*NBL)
  cx<Tr> v[PERL][RL];
  row_inverse<Tr, L, MODE == ROW_PHYS1>(W + rows, W + (size_t)M * H + rows, active, buf, tws, t, v);

  if constexpr (MODE == ROW_PHYS || MODE == ROW_PHYS1) {
    double mx = 0.0, sm = 0.0;
    bool nan = false;
#pragma unroll
    for (int i = 0; i < PERL; ++i) {
      const int j = t + i * T;
      if ((NBL % T == 0 || j < NBL) && active) {
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          const int x = j + r * NBL;
          const Tr u = v[i][r].x, w = v[i][r].y;
          if (MODE == ROW_PHYS) {
            ur[(size_t)row * M + x] = u;
            ur[(size_t)M * M + (size_t)row * M + x] = w;
            const Tr a = (Tr)(fabs(u) + fabs(w));
            if (a != a) nan = true;
            mx = fmax(mx, (double)a);
            sm += (double)(u * u + w * w);
          } else {
            ur[2 * (size_t)M * M + (size_t)row * M + x] = u;
          }
        }
      }
    }
    if constexpr (MODE == ROW_PHYS) {
      __syncthreads();
      double* red = reinterpret_cast<double*>(buf);  // 2*T doubles fit in L complex
      red[t] = nan ? NAN : mx;
      red[T + t] = sm;
      __syncthreads();
      constexpr int P2 = (T & (T - 1)) == 0 ? T : (T >= 256 ? 256 : (T >= 128 ? 128 : 64));
      if (P2 != T) {  // fold the threads above the largest power of two
        if (t < T - P2) {
          const double a = red[t], b = red[t + P2];
          red[t] = (a != a || b != b) ? NAN : fmax(a, b);
          red[T + t] += red[T + t + P2];
        }
        __syncthreads();
      }
      for (int s = P2 / 2; s > 0; s >>= 1) {
        if (t < s) {
          const double a = red[t], b = red[t + s];
          red[t] = (a != a || b != b) ? NAN : fmax(a, b);
          red[T + t] += red[T + t + s];
        }
        __syncthreads();
      }
      if (t == 0 && active) {
        part[row] = red[0];
        part[M + row] = red[T];
      }
    }
  } else if constexpr (MODE == ROW_STEP) {
    // (u + i w)^2 = (u^2 - w^2) + 2i uw  ->  F_x[u^2 - w^2], F_x[uw] = half of the 2i part
#pragma unroll
    for (int i = 0; i < PERL; ++i) {
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const cx<Tr> c = v[i][r];
        v[i][r] = cx<Tr>{c.x * c.x - c.y * c.y, (Tr)2 * (c.x * c.y)};
      }
    }
    row_forward_split<Tr, L>(v, (Tr)0.5, active, row, buf, tws, t, Pout);
  } else {  // ROW_ADJ: lambda_u = 2u qA + w qC, lambda_w = -2w qA + u qC
    cx<Tr> q[PERL][RL];
    __syncthreads();  // buf reused by the second inverse
    row_inverse<Tr, L, false>(Q + rows, Q + (size_t)M * H + rows, active, buf, tws, t, q);
#pragma unroll
    for (int i = 0; i < PERL; ++i) {
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const Tr u = v[i][r].x, w = v[i][r].y, qa = q[i][r].x, qc = q[i][r].y;
        v[i][r] = cx<Tr>{(Tr)2 * u * qa + w * qc, u * qc - (Tr)2 * w * qa};
      }
    }
    row_forward_split<Tr, L>(v, (Tr)1, active, row, buf, tws, t, Pout);
  }
}

// ----------------------------------------------------------------- column pass
// Layouts: Pin = products (or adjoint planes), 2 planes row-major [c][z][kx];
// om, fnm1 column-major [kx][zc] (N per column); src (COL_INIT) compact
// row-major [c][zc][kx]; out = velocity work planes (or adjoint planes)
// column-major [c][kx][z], or the one-plane getter scratch.  The column is
// written contiguously; the strided reads of Pin are served by L2 for the
// neighbouring columns that run on the same XCD at the same time.
//
// COL_ADJ (adjoint of one LS-IMEX-RK3 stage; om/fnm1 hold lambda_omega /
// lambda_fnm1, stage = st | flags):
//   ADJ_HAS_G : acc = w'(conj c_u F_z[Gu] + conj c_w F_z[Gw])   (C2R adjoint, w' = 1|2)
//   ADJ_UPDATE: lambda_om <- D1 lambda_om + acc, lambda_fnm1 <- D3 lambda_om  (stage st)
//   ADJ_WANT_Q: lambda_fn = D2 lambda_om + lambda_fnm1 for stage (st+2)%3, then
//               out = z-inverse of (GA, G2-A2) divxz lambda_fn, halved for kx>=1 (R2C adjoint)
template <typename Tr, int L, int MODE>
__global__ void __launch_bounds__(kMaxGroupThreads)
    k_col(const cx<Tr>* __restrict__ Pin, cx<Tr>* __restrict__ om, cx<Tr>* __restrict__ fnm1,
          const cx<Tr>* __restrict__ src, cx<Tr>* __restrict__ out,
          const cx<Tr>* __restrict__ twg, const double* __restrict__ scal, int stage,
          int xcd_remap) {
  constexpr Plan P = plan_of(L);
  constexpr int T = P.T;
  constexpr int G = groups_per_block<Tr>(L, T);
  constexpr int M = L, H = L / 3, N = 2 * H;
  constexpr int R1 = P.g[0];
  constexpr int RL = P.g[P.n - 1];
  constexpr int NB1 = L / R1, PER1 = (NB1 + T - 1) / T;
  constexpr int NBL = L / RL, PERL = (NBL + T - 1) / T;
  constexpr int NR = 2 * R1 / 3;  // retained outputs per radix-R1 butterfly
  constexpr bool STEP = (MODE == COL_CNAB2 || MODE == COL_RK3);
  constexpr bool ADJ = (MODE == COL_ADJ);
  extern __shared__ unsigned char smem_raw[];
  cx<Tr>* smem = reinterpret_cast<cx<Tr>*>(smem_raw);

  int blk = blockIdx.x;
  if (xcd_remap) {  // consecutive columns on one XCD at the same time share L2 lines
    const int nb = gridDim.x;
    blk = (blk % 8) * (nb / 8) + blk / 8;
  }
  const int g = threadIdx.x / T, t = threadIdx.x % T;
  const int kx = blk * G + g;
  const bool active = kx < H;
  const int kxs = active ? kx : 0;
  cx<Tr>* buf = smem + g * lds_len<false>(L);
  cx<Tr>* tws = smem + G * lds_len<false>(L);
  load_twiddles<L>(tws, twg);
  const Tr ax = (Tr)kxs;
  cx<Tr>* omc = om + (size_t)kxs * N;
  cx<Tr>* fnc = fnm1 + (size_t)kxs * N;
  const int st = stage & 15;
  const int flags = ADJ ? (stage & ~15) : 0;
  const bool want_q = !ADJ || (flags & ADJ_WANT_Q);
  const int NF_OUT = (MODE == COL_OMEGA || MODE == COL_STREAM) ? 1 : (want_q ? 2 : 0);

  // stage coefficients (turbo_simulator.py:1745-1752 CNAB2, 1927-1937 RK3)
  const Tr visc = (Tr)scal[PALX_S_VISC];
  const Tr dt = (Tr)scal[PALX_S_DT];
  const Tr cnm1 = (Tr)scal[PALX_S_CNM1];
  const Tr divxz = (Tr)1 / (((Tr)1.5 * (Tr)(N)) * ((Tr)1.5 * (Tr)(N)));
  auto alpha = [](int s) { return s == 0 ? (Tr)(8.0 / 15.0) : s == 1 ? (Tr)(5.0 / 12.0) : (Tr)(3.0 / 4.0); };
  auto beta = [](int s) { return s == 0 ? (Tr)0.0 : s == 1 ? (Tr)(-17.0 / 60.0) : (Tr)(-5.0 / 12.0); };

  // per retained mode: fn accumulator, then omega (step) / lambda_fn (adjoint)
  cx<Tr> acc[PER1][NR];
#pragma unroll
  for (int i = 0; i < PER1; ++i)
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[i][q] = cx<Tr>{0, 0};

  if (STEP || (ADJ && (flags & ADJ_HAS_G))) {
#pragma unroll 1
    for (int f = 0; f < 2; ++f) {
      const cx<Tr>* Pc = Pin + (size_t)f * M * H + (size_t)kxs;  // row-major [c][z][kx]
      {
        cx<Tr> v[PERL][RL];
#pragma unroll
        for (int i = 0; i < PERL; ++i) {
          const int j = t + i * T;
          if (NBL % T == 0 || j < NBL) {
#pragma unroll
            for (int r = 0; r < RL; ++r) v[i][r] = active ? Pc[(size_t)(j + r * NBL) * H] : cx<Tr>{0, 0};
          }
        }
        if (f == 1) __syncthreads();  // field 0 finished reading buf
        pass_first_from_regs<L, RL, T, false, false>(buf, t, v);
      }
      __syncthreads();
      if constexpr (P.n == 3) pass_lds<L, P.g[1], RL, T, false, false>(buf, tws, t);
      cx<Tr> d[PER1][R1];
      pass_last_to_regs<L, R1, T, false, false>(buf, tws, t, d);
#pragma unroll
      for (int i = 0; i < PER1; ++i) {
        const int j = t + i * T;
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          const int r = q < R1 / 3 ? q : q + R1 / 3;
          const int pos = j + r * NB1;
          const int zc = pos < H ? pos : pos - H;
          const Tr gz = (Tr)(pos < H ? pos : pos - M);
          const cx<Tr> X = d[i][r];
          if constexpr (STEP) {
            // fn = (GA F_z[A] + (G2-A2) F_z[C]) / M^2 ; row kz=-N/2 stays zero
            if (f == 0) {
              const Tr GA = gz * ax;
              acc[i][q] = cx<Tr>{X.x * GA, X.y * GA};
            } else {
              const Tr G2mA2 = gz * gz - ax * ax;
              acc[i][q] = cx<Tr>{acc[i][q].x + X.x * G2mA2, acc[i][q].y + X.y * G2mA2};
            }
          } else {
            // transpose of the reconstruction and of the x-inverse: w' conj(c) X
            const Tr wp = kxs >= 1 ? (Tr)2 : (Tr)1;
            if (f == 0) {  // conj(c_u) = i gz/K2 (kx>=1), i/gz (kx=0)
              Tr s = 0;
              if (kxs >= 1) s = gz / (ax * ax + gz * gz);
              else s = (zc >= 1 && gz != (Tr)0) ? (Tr)1 / gz : (Tr)0;
              acc[i][q] = cx<Tr>{-X.y * s * wp, X.x * s * wp};
            } else if (kxs >= 1) {  // conj(c_w) = -i ax/K2
              const Tr s = ax / (ax * ax + gz * gz);
              acc[i][q] = cx<Tr>{acc[i][q].x + X.y * s * wp, acc[i][q].y - X.x * s * wp};
            }
          }
        }
      }
    }
  }

  if constexpr (STEP) {
    // ---- spectral update (turbo_simulator.py:1731-1771 CNAB2, 1912-1943 RK3)
    Tr VT = 0, c2 = 0, c3 = 0, bdv = 0, adt = 0, bdt = 0, adv = 0;
    if constexpr (MODE == COL_CNAB2) {
      VT = (Tr)0.5 * visc * dt;
      c2 = (Tr)0.5 * dt * ((Tr)2 + cnm1);
      c3 = -(Tr)0.5 * dt * cnm1;
    } else {
      const Tr al = alpha(st), be = beta(st);
      bdv = be * dt * visc;
      adt = al * dt;
      bdt = be * dt;
      adv = al * dt * visc;
    }
#pragma unroll
    for (int i = 0; i < PER1; ++i) {
      const int j = t + i * T;
      if ((NB1 % T == 0 || j < NB1) && active) {
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          const int r = q < R1 / 3 ? q : q + R1 / 3;
          const int pos = j + r * NB1;
          const int zc = pos < H ? pos : pos - H;
          const Tr gz = (Tr)(pos < H ? pos : pos - M);
          const Tr K2 = ax * ax + gz * gz;
          cx<Tr> f1 = {0, 0};
          if (zc != H) f1 = cx<Tr>{acc[i][q].x * divxz, acc[i][q].y * divxz};
          const cx<Tr> w0 = omc[zc];
          const cx<Tr> f0 = fnc[zc];
          cx<Tr> wn;
          if constexpr (MODE == COL_CNAB2) {
            const Tr ARG = K2 * VT;
            const Tr DEN = ARG + (Tr)1;
            cx<Tr> num = {w0.x - w0.x * ARG, w0.y - w0.y * ARG};
            num = cx<Tr>{num.x + f1.x * c2, num.y + f1.y * c2};
            num = cx<Tr>{num.x + f0.x * c3, num.y + f0.y * c3};
            const Tr inv = (Tr)1 / DEN;
            wn = cx<Tr>{num.x * inv, num.y * inv};
          } else {
            const Tr tmp = (Tr)1 - K2 * bdv;
            cx<Tr> num = {w0.x * tmp, w0.y * tmp};
            num = cx<Tr>{num.x + f1.x * adt, num.y + f1.y * adt};
            num = cx<Tr>{num.x + f0.x * bdt, num.y + f0.y * bdt};
            const Tr den = K2 * adv + (Tr)1;
            const Tr inv = (Tr)1 / den;
            wn = cx<Tr>{num.x * inv, num.y * inv};
          }
          omc[zc] = wn;
          fnc[zc] = f1;
          acc[i][q] = wn;
        }
      }
    }
  } else if constexpr (ADJ || MODE == COL_RECON) {
    // transposed stage update: omega' = (tmp omega + adt fn + bdt fnm1) / den
    Tr bdv = 0, adv = 0, bdt = 0, padt = 0, padv = 0;
    if constexpr (ADJ) {
      const Tr al = alpha(st), be = beta(st);
      bdv = be * dt * visc;              // tmp = 1 - K2 bdv
      adv = al * dt * visc;              // den = K2 adv + 1
      bdt = be * dt;
      const int p = (st + 2) % 3;        // the stage before st
      padt = alpha(p) * dt;
      padv = alpha(p) * dt * visc;
    }
#pragma unroll
    for (int i = 0; i < PER1; ++i) {
      const int j = t + i * T;
      if ((NB1 % T == 0 || j < NB1) && active) {
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          const int r = q < R1 / 3 ? q : q + R1 / 3;
          const int pos = j + r * NB1;
          const int zc = pos < H ? pos : pos - H;
          if constexpr (MODE == COL_RECON) {
            acc[i][q] = omc[zc];
          } else {
            const Tr gz = (Tr)(pos < H ? pos : pos - M);
            const Tr K2 = ax * ax + gz * gz;
            cx<Tr> lo = omc[zc], lf = fnc[zc];
            if (flags & ADJ_UPDATE) {
              const Tr inv = (Tr)1 / (K2 * adv + (Tr)1);
              const Tr D1 = ((Tr)1 - K2 * bdv) * inv;
              const Tr D3 = bdt * inv;
              const cx<Tr> g = acc[i][q];
              lf = cx<Tr>{lo.x * D3, lo.y * D3};
              lo = cx<Tr>{lo.x * D1 + g.x, lo.y * D1 + g.y};
              omc[zc] = lo;
              fnc[zc] = lf;
            }
            if (flags & ADJ_WANT_Q) {  // lambda_fn of stage (st+2)%3 = D2 lambda_om + lambda_fnm1
              const Tr invp = (Tr)1 / (K2 * padv + (Tr)1);
              const Tr D2 = padt * invp;
              acc[i][q] = cx<Tr>{lo.x * D2 + lf.x, lo.y * D2 + lf.y};
            }
          }
        }
      }
    }
  }

  // ---- zero-padded z-inverse of the retained spectra -> out[f][kx][z]
#pragma unroll 1
  for (int f = 0; f < NF_OUT; ++f) {
    {
      cx<Tr> v[PER1][R1];
#pragma unroll
      for (int i = 0; i < PER1; ++i) {
        const int j = t + i * T;
#pragma unroll
        for (int r = 0; r < R1; ++r) {
          cx<Tr> a = {0, 0};
          if (r < R1 / 3 || r >= 2 * R1 / 3) {
            const int q = r < R1 / 3 ? r : r - R1 / 3;
            const int pos = j + r * NB1;
            const int zc = pos < H ? pos : pos - H;
            const Tr gz = (Tr)(pos < H ? pos : pos - M);
            if constexpr (STEP || MODE == COL_RECON) {
              // velocity reconstruction from omega (turbo_simulator.py:1555-1585)
              const cx<Tr> wn = acc[i][q];
              if (kxs >= 1) {
                const Tr K2 = ax * ax + gz * gz;
                const Tr invK2 = (Tr)1 / (K2 + (Tr)1.0e-30);
                const Tr px = wn.x * invK2, py = wn.y * invK2;
                a = f == 0 ? cx<Tr>{py * gz, -(px * gz)} : cx<Tr>{-(py * ax), px * ax};
              } else if (f == 0) {
                const Tr invg = (zc >= 1 && gz != (Tr)0) ? (Tr)1 / gz : (Tr)0;
                a = cx<Tr>{wn.y * invg, -wn.x * invg};
              }
            } else if constexpr (ADJ) {
              // R2C / truncation adjoint of the products: GA, G2-A2, divxz, half for kx>=1
              if (zc != H) {
                const Tr cf = (f == 0 ? gz * ax : gz * gz - ax * ax) * divxz * (kxs >= 1 ? (Tr)0.5 : (Tr)1);
                a = cx<Tr>{acc[i][q].x * cf, acc[i][q].y * cf};
              }
            } else if (active && (NB1 % T == 0 || j < NB1)) {
              if constexpr (MODE == COL_INIT) {
                a = src[(size_t)f * N * H + (size_t)zc * H + kxs];
              } else if (zc != H) {  // _spectral_band_to_phys_full_grid zeroes row N
                a = omc[zc];
                if constexpr (MODE == COL_STREAM) {
                  const Tr K2 = (ax * ax + gz * gz) + (Tr)1.0e-30;
                  const Tr inv = (Tr)1 / K2;  // complex / real in numpy: reciprocal multiply
                  a = cx<Tr>{a.x * inv, a.y * inv};
                }
              }
            }
          }
          v[i][r] = a;
        }
      }
      __syncthreads();  // previous readers of buf are done
      pass_first_from_regs<L, R1, T, true, false>(buf, t, v);
    }
    __syncthreads();
    if constexpr (P.n == 3) pass_lds<L, P.g[1], R1, T, true, false>(buf, tws, t);
    cx<Tr> v[PERL][RL];
    pass_last_to_regs<L, RL, T, true, false>(buf, tws, t, v);
    if (active) {
      cx<Tr>* Of = out + (size_t)f * M * H;
#pragma unroll
      for (int i = 0; i < PERL; ++i) {
        const int j = t + i * T;
        if (NBL % T == 0 || j < NBL) {
#pragma unroll
          for (int r = 0; r < RL; ++r) Of[(size_t)kxs * M + (j + r * NBL)] = v[i][r];
        }
      }
    }
  }
}

// ----------------------------------------------------------------- small kernels
// Reduce the per-row partials of the physical pass; optionally NEXTDT.
// f32: CFLM = float32(max * float32(inv_dx)) like the reference's float32 max.
__global__ void k_reduce_phys(const double* __restrict__ part, int M, double* __restrict__ scal,
                              int f32, int do_nextdt) {
  __shared__ double smx[256], ssm[256];
  const int t = threadIdx.x;
  double mx = 0.0, sm = 0.0;
  bool nan = false;
  for (int z = t; z < M; z += 256) {
    const double a = part[z];
    if (a != a) nan = true;
    else mx = fmax(mx, a);
    sm += part[M + z];
  }
  smx[t] = nan ? NAN : mx;
  ssm[t] = sm;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      const double a = smx[t], b = smx[t + s];
      smx[t] = (a != a || b != b) ? NAN : fmax(a, b);
      ssm[t] += ssm[t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    const double m = smx[0];
    double cflm;
    if (f32) cflm = (double)((float)m * (float)scal[PALX_S_INVDX]);
    else cflm = m * scal[PALX_S_INVDX];
    scal[PALX_S_CFLM] = cflm;
    scal[PALX_S_USUM] = ssm[0];
    if (m != m) scal[PALX_S_NAN] = 1.0;
    if (do_nextdt) {  // turbo_simulator.py:2161-2166; a NaN CFLM falls through, as there
      const double dt = scal[PALX_S_DT];
      if (!(cflm <= 0.0 || dt <= 0.0)) {
        const double CFL = cflm * dt * 3.141592653589793;
        const double cn = 0.8 + 0.2 * (scal[PALX_S_CFLNUM] / CFL);
        scal[PALX_S_CN] = cn;
        scal[PALX_S_DT] = dt * cn;
      }
    }
  }
}

// cnm1 = cn at the end of every CNAB2 update (turbo_simulator.py:1771, GPU 857-871)
__global__ void k_copy_cn(double* __restrict__ scal) {
  if (threadIdx.x == 0) scal[PALX_S_CNM1] = scal[PALX_S_CN];
}

// dt of the step just taken, before NEXTDT may change it
__global__ void k_save_dt(double* __restrict__ scal, double* __restrict__ dst) {
  if (threadIdx.x == 0) dst[0] = scal[PALX_S_DT];
}

// t += dt_old; it += 1 (turbo_simulator.py:2833, 2848)
__global__ void k_advance_t(double* __restrict__ scal, const double* __restrict__ dt_old) {
  if (threadIdx.x == 0) {
    scal[PALX_S_T] = scal[PALX_S_T] + dt_old[0];
    scal[PALX_S_IT] = scal[PALX_S_IT] + 1.0;
  }
}

// spectral sums over omega (N x H): per-block partials of
// {E, Lsum, Z, Pal, P00} -- flow_eddy_metrics (turbo_simulator.py:772-808) and
// pal_over_ens_kmax2 (turbo_main.py:2015-2073), element math in Tr, sums FP64.
template <typename Tr>
__global__ void k_diag(const cx<Tr>* __restrict__ om, int N, double* __restrict__ part) {
  const int H = N / 2;
  __shared__ double red[5][256];
  double acc[5] = {0, 0, 0, 0, 0};
  const size_t total = (size_t)N * H;
  for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * 256) {
    const int kx = (int)(idx / N), z = (int)(idx % N);  // column-major [kx][z]
    const Tr gz = (Tr)(z < H ? z : z - N), ax = (Tr)kx;
    const Tr k2 = ax * ax + gz * gz;
    const cx<Tr> w = om[idx];
    const Tr p = w.x * w.x + w.y * w.y;
    const Tr wt = (kx == 0 || kx == H - 1) ? (Tr)1 : (Tr)2;
    if (k2 > (Tr)0) {
      const Tr power = p / k2;
      const Tr wp = power * wt;
      const Tr invk = (Tr)1 / sqrt(k2);
      acc[0] += (double)wp;
      acc[1] += (double)(wp * invk);
    }
    acc[2] += (double)(wt * p);
    acc[3] += (double)(wt * k2 * p);
    if (idx == 0) acc[4] += (double)p;
  }
  for (int q = 0; q < 5; ++q) red[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int q = 0; q < 5; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 5; ++q) part[(size_t)blockIdx.x * 5 + q] = red[q][0];
}

__global__ void k_diag_final(const double* __restrict__ part, int nblk, double* __restrict__ out) {
  if (threadIdx.x == 0) {
    double s[5] = {0, 0, 0, 0, 0};
    for (int b = 0; b < nblk; ++b)
      for (int q = 0; q < 5; ++q) s[q] += part[(size_t)b * 5 + q];
    for (int q = 0; q < 5; ++q) out[q] = s[q];
  }
}

// radial spectrum bins (turbo_simulator.py:2232-2306), FP64 atomics
template <typename Tr>
__global__ void k_spectrum(const cx<Tr>* __restrict__ om, int N, int nbins, double r_max,
                           double k_nyq, double fft_scale2, double* __restrict__ esum,
                           int* __restrict__ cnt) {
  const int H = N / 2;
  const size_t total = (size_t)N * H;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(idx / N), z = (int)(idx % N);  // column-major [kx][z]
    const int kz = z < N / 2 ? z : z - N;
    const double k2 = (double)kx * kx + (double)kz * kz;
    if (k2 > 0.0) {
      const double r = sqrt(k2) / k_nyq;
      int bin = (int)floor((r / r_max) * (double)nbins);
      if (bin < 0) bin = 0;
      if (bin >= nbins) bin = nbins - 1;
      const cx<Tr> w = om[idx];
      const double o2 = (double)w.x * (double)w.x + (double)w.y * (double)w.y;
      double p = (o2 / k2) * fft_scale2;
      const int wt = kx > 0 ? 2 : 1;
      if (wt == 2) p *= 2.0;
      atomicAdd(&esum[bin], p);
      atomicAdd(&cnt[bin], wt);
    }
  }
}

template <typename Tr>
__global__ void k_kinetic(Tr* __restrict__ ur, size_t plane) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < plane;
       i += (size_t)gridDim.x * blockDim.x) {
    const Tr u = ur[i], w = ur[plane + i];
    ur[2 * plane + i] = sqrt(u * u + w * w);
  }
}

template <typename Tr>
__global__ void k_zero(Tr* __restrict__ p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = (Tr)0;
}

// dst[b][a] = src[a][b] for an (na x nb) complex matrix
template <typename Tr>
__global__ void k_transpose(const cx<Tr>* __restrict__ src, cx<Tr>* __restrict__ dst, int na, int nb) {
  __shared__ cx<Tr> tile[32][33];
  const int b0 = blockIdx.x * 32, a0 = blockIdx.y * 32;
  for (int y = threadIdx.y; y < 32; y += 8) {
    const int a = a0 + y, b = b0 + threadIdx.x;
    if (a < na && b < nb) tile[y][threadIdx.x] = src[(size_t)a * nb + b];
  }
  __syncthreads();
  for (int y = threadIdx.y; y < 32; y += 8) {
    const int b = b0 + y, a = a0 + threadIdx.x;
    if (a < na && b < nb) dst[(size_t)b * na + a] = tile[threadIdx.x][y];
  }
}

}  // namespace

// ============================================================================ host side
namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(PALX_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
  } while (0)

}  // namespace

struct palx_t {
  int N = 0, M = 0, H = 0, prec = 8;
  size_t csz = 16;  // complex size
  void* om = nullptr;
  void* fnm1 = nullptr;
  void* W = nullptr;     // velocity work planes, 2 x M x H complex, row-major [c][z][kx]
  void* Wp = nullptr;    // products, 2 x H x M complex, column-major [c][kx][z]
  void* ur = nullptr;    // 3 x M x M real
  void* uc0 = nullptr;   // 2 x N x H complex
  void* gs = nullptr;    // getter scratch, M x H complex
  void* tw = nullptr;    // M complex twiddles exp(-2 pi i m / M)
  double* scal = nullptr;
  double* part = nullptr;  // 2*M (row partials) + diag partials
  double* dtsave = nullptr;
  // adjoint of the fixed-dt LS-IMEX-RK3 trajectory
  void* lam_om = nullptr;   // lambda_omega, column-major like om
  void* lam_fn = nullptr;   // lambda_fnm1
  void* Q = nullptr;        // adjoint z-inverse planes, layout of W
  std::vector<void*> ckpt;  // (omega, fnm1) every ckpt_every steps
  std::vector<void*> vslot; // velocity planes of every stage of one segment
  int adj_nsteps = 0, adj_every = 0;
  hipStream_t stream = nullptr;
  int phase = 0;  // 0: W holds velocity, 1: W holds products
  int64_t it = 0;
  std::vector<void*> allocs;
};

namespace {

template <typename Tr, int L>
struct Kern {
  static constexpr Plan P = plan_of(L);
  static constexpr int G = groups_per_block<Tr>(L, P.T);
  static size_t lds() { return ((size_t)G * lds_len<true>(L) + tw_len(L)) * sizeof(cx<Tr>); }
  static int row_blocks() { return (L + G - 1) / G; }
  static int col_blocks() { return (L / 3 + G - 1) / G; }
  static dim3 block() { return dim3(G * P.T); }

  static int set_attrs() {
    static std::once_flag once;
    static hipError_t err = hipSuccess;
    std::call_once(once, [] {
      const int b = (int)lds();
      hipError_t e = hipSuccess;
      auto set = [&](const void* f) {
        hipError_t r = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, b);
        if (r != hipSuccess) e = r;
      };
      set((const void*)k_row<Tr, L, ROW_STEP>);
      set((const void*)k_row<Tr, L, ROW_PHYS>);
      set((const void*)k_row<Tr, L, ROW_PHYS1>);
      set((const void*)k_row<Tr, L, ROW_ADJ>);
      set((const void*)k_col<Tr, L, COL_RECON>);
      set((const void*)k_col<Tr, L, COL_ADJ>);
      set((const void*)k_col<Tr, L, COL_CNAB2>);
      set((const void*)k_col<Tr, L, COL_RK3>);
      set((const void*)k_col<Tr, L, COL_INIT>);
      set((const void*)k_col<Tr, L, COL_OMEGA>);
      set((const void*)k_col<Tr, L, COL_STREAM>);
      err = e;
    });
    if (err != hipSuccess) return fail(PALX_EHIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(err));
    return PALX_OK;
  }

  static int row(palx_t* h, int mode, void* target_W) {
    int rc = set_attrs();
    if (rc) return rc;
    auto* Wv = (const cx<Tr>*)target_W;
    auto* Qp = (const cx<Tr>*)h->Q;
    auto* Pout = (cx<Tr>*)h->Wp;
    auto* ur = (Tr*)h->ur;
    auto* tw = (const cx<Tr>*)h->tw;
    const int nb = row_blocks();
    const int remap = (nb % 8 == 0 && nb >= 64) ? 1 : 0;
    dim3 grid(nb);
    if (mode == ROW_STEP)
      hipLaunchKernelGGL((k_row<Tr, L, ROW_STEP>), grid, block(), lds(), h->stream, Wv, Qp, Pout, ur, h->part, tw, remap);
    else if (mode == ROW_PHYS)
      hipLaunchKernelGGL((k_row<Tr, L, ROW_PHYS>), grid, block(), lds(), h->stream, Wv, Qp, Pout, ur, h->part, tw, remap);
    else if (mode == ROW_PHYS1)
      hipLaunchKernelGGL((k_row<Tr, L, ROW_PHYS1>), grid, block(), lds(), h->stream, Wv, Qp, Pout, ur, h->part, tw, remap);
    else
      hipLaunchKernelGGL((k_row<Tr, L, ROW_ADJ>), grid, block(), lds(), h->stream, Wv, Qp, Pout, ur, h->part, tw, remap);
    HIPCHK(hipGetLastError());
    return PALX_OK;
  }

  static int col(palx_t* h, int mode, int stage) {
    int rc = set_attrs();
    if (rc) return rc;
    const int nb = col_blocks();
    const int remap = (nb % 8 == 0 && nb >= 64) ? 1 : 0;
    auto* Pin = (const cx<Tr>*)h->Wp;
    const bool adj = mode == COL_ADJ;
    auto* om = (cx<Tr>*)(adj ? h->lam_om : h->om);
    auto* fn = (cx<Tr>*)(adj ? h->lam_fn : h->fnm1);
    auto* src = (const cx<Tr>*)h->uc0;
    auto* tw = (const cx<Tr>*)h->tw;
    cx<Tr>* out = (mode == COL_OMEGA || mode == COL_STREAM) ? (cx<Tr>*)h->gs
                  : adj ? (cx<Tr>*)h->Q : (cx<Tr>*)h->W;
    dim3 grid(nb);
#define PALX_COL(MODE_)                                                                          \
  hipLaunchKernelGGL((k_col<Tr, L, MODE_>), grid, block(), lds(), h->stream, Pin, om, fn, src, out, \
                     tw, (const double*)h->scal, stage, remap)
    switch (mode) {
      case COL_CNAB2: PALX_COL(COL_CNAB2); break;
      case COL_RK3: PALX_COL(COL_RK3); break;
      case COL_INIT: PALX_COL(COL_INIT); break;
      case COL_OMEGA: PALX_COL(COL_OMEGA); break;
      case COL_STREAM: PALX_COL(COL_STREAM); break;
      case COL_RECON: PALX_COL(COL_RECON); break;
      default: PALX_COL(COL_ADJ); break;
    }
#undef PALX_COL
    HIPCHK(hipGetLastError());
    return PALX_OK;
  }
};

template <typename Tr>
int dispatch_row(palx_t* h, int mode, void* target) {
  switch (h->M) {
    case 12: return Kern<Tr, 12>::row(h, mode, target);
    case 24: return Kern<Tr, 24>::row(h, mode, target);
    case 48: return Kern<Tr, 48>::row(h, mode, target);
    case 96: return Kern<Tr, 96>::row(h, mode, target);
    case 192: return Kern<Tr, 192>::row(h, mode, target);
    case 384: return Kern<Tr, 384>::row(h, mode, target);
    case 768: return Kern<Tr, 768>::row(h, mode, target);
    case 1536: return Kern<Tr, 1536>::row(h, mode, target);
    case 3072: return Kern<Tr, 3072>::row(h, mode, target);
    case 6144: return Kern<Tr, 6144>::row(h, mode, target);
  }
  return fail(PALX_EUNSUP, "unsupported grid size");
}

template <typename Tr>
int dispatch_col(palx_t* h, int mode, int stage) {
  switch (h->M) {
    case 12: return Kern<Tr, 12>::col(h, mode, stage);
    case 24: return Kern<Tr, 24>::col(h, mode, stage);
    case 48: return Kern<Tr, 48>::col(h, mode, stage);
    case 96: return Kern<Tr, 96>::col(h, mode, stage);
    case 192: return Kern<Tr, 192>::col(h, mode, stage);
    case 384: return Kern<Tr, 384>::col(h, mode, stage);
    case 768: return Kern<Tr, 768>::col(h, mode, stage);
    case 1536: return Kern<Tr, 1536>::col(h, mode, stage);
    case 3072: return Kern<Tr, 3072>::col(h, mode, stage);
    case 6144: return Kern<Tr, 6144>::col(h, mode, stage);
  }
  return fail(PALX_EUNSUP, "unsupported grid size");
}

int run_row(palx_t* h, int mode, void* target) {
  return h->prec == 4 ? dispatch_row<float>(h, mode, target) : dispatch_row<double>(h, mode, target);
}
int run_col(palx_t* h, int mode, int stage) {
  return h->prec == 4 ? dispatch_col<float>(h, mode, stage) : dispatch_col<double>(h, mode, stage);
}

bool supported_N(int N) {
  switch (N) {
    case 8: case 16: case 32: case 64: case 128: case 256: case 512: case 1024: case 2048:
    case 4096:
      return true;
  }
  return false;
}

int dev_alloc(palx_t* h, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(PALX_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  h->allocs.push_back(*p);
  hipError_t z = hipMemset(*p, 0, bytes);
  if (z != hipSuccess) return fail(PALX_EHIP, std::string("hipMemset: ") + hipGetErrorString(z));
  return PALX_OK;
}

// omega / fnm1 live column-major [kx][z] on the device; the ABI exchanges
// them in the reference's row-major [z][kx] layout through the getter scratch.
int transpose_dev(palx_t* h, const void* src, void* dst, int na, int nb) {
  dim3 grid((nb + 31) / 32, (na + 31) / 32), block(32, 8);
  if (h->prec == 4)
    hipLaunchKernelGGL(k_transpose<float>, grid, block, 0, h->stream, (const cx<float>*)src, (cx<float>*)dst, na, nb);
  else
    hipLaunchKernelGGL(k_transpose<double>, grid, block, 0, h->stream, (const cx<double>*)src, (cx<double>*)dst, na, nb);
  HIPCHK(hipGetLastError());
  return PALX_OK;
}

bool is_spectral(int field) {
  return field == PALX_F_OM || field == PALX_F_FNM1 || field == PALX_F_ADJ_OM;
}

int field_info(palx_t* h, int field, void** ptr, size_t* bytes) {
  const size_t NH = (size_t)h->N * h->H;
  const size_t rsz = (size_t)h->prec;
  switch (field) {
    case PALX_F_OM: *ptr = h->om; *bytes = NH * h->csz; return PALX_OK;
    case PALX_F_FNM1: *ptr = h->fnm1; *bytes = NH * h->csz; return PALX_OK;
    case PALX_F_UR: *ptr = h->ur; *bytes = 3 * (size_t)h->M * h->M * rsz; return PALX_OK;
    case PALX_F_WORK: *ptr = h->W; *bytes = 2 * (size_t)h->M * h->H * h->csz; return PALX_OK;
    case PALX_F_UC0: *ptr = h->uc0; *bytes = 2 * NH * h->csz; return PALX_OK;
    case PALX_F_SCAL: *ptr = h->scal; *bytes = PALX_NSCAL * sizeof(double); return PALX_OK;
    case PALX_F_ADJ_OM: *ptr = h->lam_om; *bytes = NH * h->csz; return PALX_OK;
  }
  return fail(PALX_EINVAL, "unknown field id");
}

int physical(palx_t* h, int do_nextdt) {
  if (h->phase != 0) return fail(PALX_ESTATE, "work planes hold products; run the update stage first");
  int rc = run_row(h, ROW_PHYS, h->W);
  if (rc) return rc;
  hipLaunchKernelGGL(k_reduce_phys, dim3(1), dim3(256), 0, h->stream, (const double*)h->part, h->M,
                     h->scal, h->prec == 4 ? 1 : 0, do_nextdt);
  HIPCHK(hipGetLastError());
  return PALX_OK;
}

// column pass of one update; CNAB2 also performs cnm1 = cn like dns_step3
int update(palx_t* h, int method, int stage) {
  int rc = run_col(h, method == PALX_CNAB2 ? COL_CNAB2 : COL_RK3, stage);
  if (rc) return rc;
  if (method == PALX_CNAB2) {
    hipLaunchKernelGGL(k_copy_cn, dim3(1), dim3(64), 0, h->stream, h->scal);
    HIPCHK(hipGetLastError());
  }
  return PALX_OK;
}

}  // namespace

// ============================================================================ C ABI
extern "C" {

int palx_last_error(char* buf, size_t n) {
  if (!buf || n == 0) return PALX_EINVAL;
  std::snprintf(buf, n, "%s", g_err.c_str());
  return PALX_OK;
}

int palx_create(int32_t N, int32_t precision, palx_t** out) {
  if (!out) return fail(PALX_EINVAL, "out is null");
  *out = nullptr;
  if (precision != 4 && precision != 8) return fail(PALX_EINVAL, "precision must be 4 or 8");
  if (!supported_N(N)) return fail(PALX_EUNSUP, "N must be a power of two in [8, 4096]");
  palx_t* h = new palx_t();
  h->N = N;
  h->M = 3 * N / 2;
  h->H = N / 2;
  h->prec = precision;
  h->csz = 2 * (size_t)precision;
  const size_t NH = (size_t)N * h->H, MH = (size_t)h->M * h->H;
  int rc = PALX_OK;
  if (!rc) rc = dev_alloc(h, &h->om, NH * h->csz);
  if (!rc) rc = dev_alloc(h, &h->fnm1, NH * h->csz);
  if (!rc) rc = dev_alloc(h, &h->W, 2 * MH * h->csz);
  if (!rc) rc = dev_alloc(h, &h->Wp, 2 * MH * h->csz);
  if (!rc) rc = dev_alloc(h, &h->ur, 3 * (size_t)h->M * h->M * precision);
  if (!rc) rc = dev_alloc(h, &h->uc0, 2 * NH * h->csz);
  if (!rc) rc = dev_alloc(h, &h->gs, MH * h->csz);
  if (!rc) rc = dev_alloc(h, &h->lam_om, NH * h->csz);
  if (!rc) rc = dev_alloc(h, &h->lam_fn, NH * h->csz);
  if (!rc) rc = dev_alloc(h, &h->Q, 2 * MH * h->csz);
  if (!rc) rc = dev_alloc(h, &h->tw, (size_t)h->M * h->csz);
  if (!rc) rc = dev_alloc(h, (void**)&h->scal, PALX_NSCAL * sizeof(double));
  if (!rc) rc = dev_alloc(h, (void**)&h->part, (2 * (size_t)h->M + 5 * 1024 + 16) * sizeof(double));
  if (!rc) rc = dev_alloc(h, (void**)&h->dtsave, 16 * sizeof(double));
  if (rc) {
    palx_destroy(h);
    return rc;
  }
  // twiddles exp(-2 pi i m / M), FP64 on the host
  {
    std::vector<double> tw(2 * (size_t)h->M);
    for (int m = 0; m < h->M; ++m) {
      const double a = -2.0 * M_PI * (double)m / (double)h->M;
      tw[2 * m] = std::cos(a);
      tw[2 * m + 1] = std::sin(a);
    }
    hipError_t e;
    if (precision == 8) {
      e = hipMemcpy(h->tw, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice);
    } else {
      std::vector<float> tf(tw.begin(), tw.end());
      e = hipMemcpy(h->tw, tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
      palx_destroy(h);
      return fail(PALX_EHIP, std::string("twiddle upload: ") + hipGetErrorString(e));
    }
  }
  double sc[PALX_NSCAL] = {0};
  sc[PALX_S_CN] = 1.0;
  sc[PALX_S_INVDX] = (double)N / (2.0 * M_PI);  // turbo_simulator.py:940
  hipError_t e = hipMemcpy(h->scal, sc, sizeof(sc), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    palx_destroy(h);
    return fail(PALX_EHIP, std::string("scalar upload: ") + hipGetErrorString(e));
  }
  *out = h;
  return PALX_OK;
}

int palx_destroy(palx_t* h) {
  if (!h) return PALX_OK;
  for (void* p : h->allocs) (void)hipFree(p);
  delete h;
  return PALX_OK;
}

int palx_set_stream(palx_t* h, void* s) {
  if (!h) return fail(PALX_EINVAL, "null handle");
  h->stream = (hipStream_t)s;
  return PALX_OK;
}

int palx_info(const palx_t* h, int32_t* N, int32_t* M, int32_t* precision) {
  if (!h) return fail(PALX_EINVAL, "null handle");
  if (N) *N = h->N;
  if (M) *M = h->M;
  if (precision) *precision = h->prec;
  return PALX_OK;
}

int palx_sync(palx_t* h) {
  if (!h) return fail(PALX_EINVAL, "null handle");
  HIPCHK(hipStreamSynchronize(h->stream));
  return PALX_OK;
}

int palx_load_initial(palx_t* h, const float* uc0, const float* om0) {
  if (!h || !uc0 || !om0) return fail(PALX_EINVAL, "null argument");
  const size_t NH = (size_t)h->N * h->H;
  HIPCHK(hipStreamSynchronize(h->stream));
  if (h->prec == 4) {
    HIPCHK(hipMemcpy(h->uc0, uc0, 2 * NH * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->gs, om0, NH * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<double> a(4 * NH), b(2 * NH);
    for (size_t i = 0; i < 4 * NH; ++i) a[i] = (double)uc0[i];
    for (size_t i = 0; i < 2 * NH; ++i) b[i] = (double)om0[i];
    HIPCHK(hipMemcpy(h->uc0, a.data(), a.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->gs, b.data(), b.size() * 8, hipMemcpyHostToDevice));
  }
  {
    int rt = transpose_dev(h, h->gs, h->om, h->N, h->H);
    if (rt) return rt;
  }
  HIPCHK(hipMemset(h->fnm1, 0, NH * h->csz));
  h->phase = 0;
  int rc = run_col(h, COL_INIT, 0);
  if (rc) return rc;
  return PALX_OK;
}

int palx_load_velocity(palx_t* h) {
  if (!h) return fail(PALX_EINVAL, "null handle");
  h->phase = 0;
  return run_col(h, COL_INIT, 0);
}

int palx_field_ptr(palx_t* h, int32_t field, void** ptr, size_t* bytes) {
  if (!h || !ptr || !bytes) return fail(PALX_EINVAL, "null argument");
  return field_info(h, field, ptr, bytes);
}

int palx_download(palx_t* h, int32_t field, void* dst, size_t bytes) {
  if (!h || !dst) return fail(PALX_EINVAL, "null argument");
  void* p;
  size_t b;
  int rc = field_info(h, field, &p, &b);
  if (rc) return rc;
  if (bytes != b) return fail(PALX_EINVAL, "byte count does not match the field size");
  if (is_spectral(field)) {
    rc = transpose_dev(h, p, h->gs, h->H, h->N);  // [kx][z] -> [z][kx]
    if (rc) return rc;
    p = h->gs;
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(dst, p, b, hipMemcpyDefault));
  return PALX_OK;
}

int palx_upload(palx_t* h, int32_t field, const void* src, size_t bytes) {
  if (!h || !src) return fail(PALX_EINVAL, "null argument");
  void* p;
  size_t b;
  int rc = field_info(h, field, &p, &b);
  if (rc) return rc;
  if (bytes != b) return fail(PALX_EINVAL, "byte count does not match the field size");
  HIPCHK(hipStreamSynchronize(h->stream));
  if (is_spectral(field)) {
    HIPCHK(hipMemcpy(h->gs, src, b, hipMemcpyDefault));
    rc = transpose_dev(h, h->gs, p, h->N, h->H);  // [z][kx] -> [kx][z]
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return PALX_OK;
  }
  HIPCHK(hipMemcpy(p, src, b, hipMemcpyDefault));
  return PALX_OK;
}

int palx_set_scalar(palx_t* h, int32_t id, double value) {
  if (!h || id < 0 || id >= PALX_NSCAL) return fail(PALX_EINVAL, "bad scalar id");
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(h->scal + id, &value, sizeof(double), hipMemcpyHostToDevice));
  return PALX_OK;
}

int palx_get_scalars(palx_t* h, double* out, int32_t n) {
  if (!h || !out || n < 0 || n > PALX_NSCAL) return fail(PALX_EINVAL, "bad argument");
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(out, h->scal, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  return PALX_OK;
}

int palx_stage_products(palx_t* h) {
  if (!h) return fail(PALX_EINVAL, "null handle");
  if (h->phase != 0) return fail(PALX_ESTATE, "products already formed; run the update stage");
  int rc = run_row(h, ROW_STEP, h->W);
  if (rc) return rc;
  h->phase = 1;
  return PALX_OK;
}

int palx_stage_update(palx_t* h, int32_t method, int32_t stage) {
  if (!h) return fail(PALX_EINVAL, "null handle");
  if (h->phase != 1) return fail(PALX_ESTATE, "update needs the products of dns_step2b first");
  if (method != PALX_CNAB2 && method != PALX_LS_IMEX_RK3) return fail(PALX_EINVAL, "unknown method");
  if (stage < 0 || stage > 2) return fail(PALX_EINVAL, "stage must be 0..2");
  int rc = update(h, method, stage);
  if (rc) return rc;
  h->phase = 0;
  return PALX_OK;
}

int palx_set_iteration(palx_t* h, int64_t it) {
  if (!h) return fail(PALX_EINVAL, "null handle");
  h->it = it;
  return PALX_OK;
}

int palx_step(palx_t* h, int32_t nsteps, int32_t method, int32_t nextdt_every) {
  if (!h) return fail(PALX_EINVAL, "null handle");
  if (method != PALX_CNAB2 && method != PALX_LS_IMEX_RK3) return fail(PALX_EINVAL, "unknown method");
  if (h->phase != 0) return fail(PALX_ESTATE, "work planes hold products");
  const int cn = method == PALX_CNAB2 ? 1 : 0;
  for (int s = 0; s < nsteps; ++s) {
    const int nst = cn ? 1 : 3;
    for (int st = 0; st < nst; ++st) {
      int rc = run_row(h, ROW_STEP, h->W);
      if (rc) return rc;
      rc = update(h, method, st);
      if (rc) return rc;
    }
    h->it += 1;
    const bool ndt = nextdt_every > 0 && ((h->it % nextdt_every) == 0 || h->it == 1);
    if (ndt) {
      hipLaunchKernelGGL(k_save_dt, dim3(1), dim3(64), 0, h->stream, h->scal, h->dtsave);
      HIPCHK(hipGetLastError());
      int rc = physical(h, 1);
      if (rc) return rc;
      hipLaunchKernelGGL(k_advance_t, dim3(1), dim3(64), 0, h->stream, h->scal, (const double*)h->dtsave);
    } else {
      hipLaunchKernelGGL(k_advance_t, dim3(1), dim3(64), 0, h->stream, h->scal,
                         (const double*)(h->scal + PALX_S_DT));
    }
    HIPCHK(hipGetLastError());
  }
  return PALX_OK;
}

int palx_physical(palx_t* h) {
  if (!h) return fail(PALX_EINVAL, "null handle");
  return physical(h, 0);
}

int palx_cflm(palx_t* h, double* cflm) {
  if (!h || !cflm) return fail(PALX_EINVAL, "null argument");
  int rc = physical(h, 0);
  if (rc) return rc;
  double sc[PALX_NSCAL];
  rc = palx_get_scalars(h, sc, PALX_NSCAL);
  if (rc) return rc;
  *cflm = sc[PALX_S_CFLM];
  return PALX_OK;
}

int palx_next_dt(palx_t* h) {
  if (!h) return fail(PALX_EINVAL, "null handle");
  return physical(h, 1);
}

int palx_diag(palx_t* h, double* out) {
  if (!h || !out) return fail(PALX_EINVAL, "null argument");
  int rc = physical(h, 0);
  if (rc) return rc;
  const int nblk = 1024;
  double* dpart = h->part + 2 * (size_t)h->M;
  double* dres = h->part + 2 * (size_t)h->M + 5 * 1024;
  if (h->prec == 4)
    hipLaunchKernelGGL(k_diag<float>, dim3(nblk), dim3(256), 0, h->stream, (const cx<float>*)h->om, h->N, dpart);
  else
    hipLaunchKernelGGL(k_diag<double>, dim3(nblk), dim3(256), 0, h->stream, (const cx<double>*)h->om, h->N, dpart);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, h->stream, (const double*)dpart, nblk, dres);
  HIPCHK(hipGetLastError());
  double s[5], sc[PALX_NSCAL];
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(s, dres, sizeof(s), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(sc, h->scal, sizeof(sc), hipMemcpyDeviceToHost));
  const double M2 = (double)h->M * (double)h->M;
  const double usum = sc[PALX_S_USUM];
  const double U = std::sqrt(std::max(0.0, usum / M2));
  const double E = s[0];
  double L = NAN, tau = NAN;
  if (E > 0.0) {
    L = 2.0 * M_PI * s[1] / E;
    tau = U > 0.0 ? L / U : NAN;
  }
  const double gz = (double)(h->N / 2), ax = (double)(h->H - 1);
  out[0] = U;
  out[1] = L;
  out[2] = tau;
  out[3] = s[2] - s[4];                 // enstrophy (weighted, minus the k=0 term)
  out[4] = s[3];                        // palinstrophy
  out[5] = ax * ax + gz * gz;           // max K2 over the compact band
  out[6] = 0.5 * (2.0 * M_PI) * (2.0 * M_PI) * usum / M2;  // E in joules
  out[7] = E;
  return PALX_OK;
}

int palx_spectrum(palx_t* h, double* esum, int32_t* count, int32_t nbins) {
  if (!h || !esum || !count || nbins <= 0) return fail(PALX_EINVAL, "bad argument");
  double* de = nullptr;
  int* dc = nullptr;
  HIPCHK(hipMalloc((void**)&de, (size_t)nbins * sizeof(double)));
  hipError_t e = hipMalloc((void**)&dc, (size_t)nbins * sizeof(int));
  if (e != hipSuccess) {
    (void)hipFree(de);
    return fail(PALX_ENOMEM, "hipMalloc spectrum");
  }
  int rc = PALX_OK;
  do {
    if (hipMemsetAsync(de, 0, (size_t)nbins * sizeof(double), h->stream) != hipSuccess ||
        hipMemsetAsync(dc, 0, (size_t)nbins * sizeof(int), h->stream) != hipSuccess) {
      rc = fail(PALX_EHIP, "memset spectrum");
      break;
    }
    const double r_max = std::sqrt(2.0), k_nyq = 0.5 * (double)h->N;
    const double fs = (double)h->M * (double)h->M;
    const size_t total = (size_t)h->N * h->H;
    const int blocks = (int)std::min<size_t>(65535, (total + 255) / 256);
    if (h->prec == 4)
      hipLaunchKernelGGL(k_spectrum<float>, dim3(blocks), dim3(256), 0, h->stream, (const cx<float>*)h->om,
                         h->N, nbins, r_max, k_nyq, fs * fs, de, dc);
    else
      hipLaunchKernelGGL(k_spectrum<double>, dim3(blocks), dim3(256), 0, h->stream, (const cx<double>*)h->om,
                         h->N, nbins, r_max, k_nyq, fs * fs, de, dc);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) {
      rc = fail(PALX_EHIP, "spectrum kernel");
      break;
    }
    if (hipMemcpy(esum, de, (size_t)nbins * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(count, dc, (size_t)nbins * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
      rc = fail(PALX_EHIP, "spectrum download");
      break;
    }
  } while (0);
  (void)hipFree(de);
  (void)hipFree(dc);
  return rc;
}

int palx_recon_velocity(palx_t* h) {
  if (!h) return fail(PALX_EINVAL, "null handle");
  h->phase = 0;
  return run_col(h, COL_RECON, 0);
}

int palx_adj_forward(palx_t* h, int32_t nsteps, int32_t ckpt_every) {
  if (!h || nsteps <= 0 || ckpt_every <= 0) return fail(PALX_EINVAL, "bad adjoint forward arguments");
  const size_t NH = (size_t)h->N * h->H, MH = (size_t)h->M * h->H;
  const int nck = (nsteps + ckpt_every - 1) / ckpt_every;
  while ((int)h->ckpt.size() < nck) {
    void* p = nullptr;
    int rc = dev_alloc(h, &p, 2 * NH * h->csz);
    if (rc) return rc;
    h->ckpt.push_back(p);
  }
  while ((int)h->vslot.size() < 3 * ckpt_every) {
    void* p = nullptr;
    int rc = dev_alloc(h, &p, 2 * MH * h->csz);
    if (rc) return rc;
    h->vslot.push_back(p);
  }
  h->adj_nsteps = nsteps;
  h->adj_every = ckpt_every;
  // the trajectory starts from velocity reconstructed from omega_0 and fnm1 = 0
  HIPCHK(hipMemsetAsync(h->fnm1, 0, NH * h->csz, h->stream));
  int rc = run_col(h, COL_RECON, 0);
  if (rc) return rc;
  h->phase = 0;
  for (int k = 0; k < nsteps; ++k) {
    if (k % ckpt_every == 0) {
      void* c = h->ckpt[k / ckpt_every];
      HIPCHK(hipMemcpyAsync(c, h->om, NH * h->csz, hipMemcpyDeviceToDevice, h->stream));
      HIPCHK(hipMemcpyAsync((char*)c + NH * h->csz, h->fnm1, NH * h->csz, hipMemcpyDeviceToDevice, h->stream));
    }
    for (int st = 0; st < 3; ++st) {
      rc = run_row(h, ROW_STEP, h->W);
      if (rc) return rc;
      rc = update(h, PALX_LS_IMEX_RK3, st);
      if (rc) return rc;
    }
  }
  return PALX_OK;
}

int palx_adj_backward(palx_t* h, const void* lam_T) {
  if (!h || !lam_T) return fail(PALX_EINVAL, "null argument");
  if (h->adj_nsteps <= 0) return fail(PALX_ESTATE, "palx_adj_forward must run first");
  const size_t NH = (size_t)h->N * h->H, MH = (size_t)h->M * h->H;
  const int nsteps = h->adj_nsteps, every = h->adj_every;
  int rc = palx_upload(h, PALX_F_ADJ_OM, lam_T, NH * h->csz);
  if (rc) return rc;
  HIPCHK(hipMemsetAsync(h->lam_fn, 0, NH * h->csz, h->stream));
  rc = run_col(h, COL_ADJ, 0 | ADJ_WANT_Q);  // lambda_fn of stage (0+2)%3, the last one -> Q
  if (rc) return rc;
  void* W0 = h->W;
  const int nseg = (nsteps + every - 1) / every;
  for (int sg = nseg - 1; sg >= 0; --sg) {
    const int k0 = sg * every, k1 = std::min(nsteps, k0 + every);
    void* c = h->ckpt[sg];
    HIPCHK(hipMemcpyAsync(h->om, c, NH * h->csz, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->fnm1, (char*)c + NH * h->csz, NH * h->csz, hipMemcpyDeviceToDevice, h->stream));
    // replay the segment; the velocity entering stage s is kept in slot s
    h->W = h->vslot[0];
    rc = run_col(h, COL_RECON, 0);
    for (int k = k0; k < k1 && !rc; ++k) {
      for (int st = 0; st < 3 && !rc; ++st) {
        const int s = (k - k0) * 3 + st;
        if (k == k1 - 1 && st == 2) break;  // only the input velocity of the last stage is needed
        h->W = h->vslot[s];
        rc = run_row(h, ROW_STEP, h->W);
        if (rc) break;
        h->W = h->vslot[s + 1];
        rc = update(h, PALX_LS_IMEX_RK3, st);
      }
    }
    if (rc) {
      h->W = W0;
      return rc;
    }
    // reverse sweep through the segment
    for (int k = k1 - 1; k >= k0 && !rc; --k) {
      for (int st = 2; st >= 0 && !rc; --st) {
        const int s = (k - k0) * 3 + st;
        rc = run_row(h, ROW_ADJ, h->vslot[s]);
        if (rc) break;
        const bool first = (k == 0 && st == 0);
        rc = run_col(h, COL_ADJ, st | ADJ_HAS_G | ADJ_UPDATE | (first ? 0 : ADJ_WANT_Q));
      }
    }
    h->W = W0;
    if (rc) return rc;
  }
  (void)MH;
  h->phase = 0;
  return PALX_OK;
}

int palx_getter(palx_t* h, int32_t var) {
  if (!h) return fail(PALX_EINVAL, "null handle");
  if (h->phase != 0) return fail(PALX_ESTATE, "work planes hold products");
  int rc = physical(h, 0);
  if (rc) return rc;
  const size_t plane = (size_t)h->M * h->M;
  switch (var) {
    case PALX_VAR_U:
    case PALX_VAR_V:
      return PALX_OK;
    case PALX_VAR_ENERGY: {
      const int blocks = (int)std::min<size_t>(4096, (plane + 255) / 256);
      if (h->prec == 4)
        hipLaunchKernelGGL(k_kinetic<float>, dim3(blocks), dim3(256), 0, h->stream, (float*)h->ur, plane);
      else
        hipLaunchKernelGGL(k_kinetic<double>, dim3(blocks), dim3(256), 0, h->stream, (double*)h->ur, plane);
      HIPCHK(hipGetLastError());
      return PALX_OK;
    }
    case PALX_VAR_OMEGA:
    case PALX_VAR_STREAM:
      rc = run_col(h, var == PALX_VAR_OMEGA ? COL_OMEGA : COL_STREAM, 0);
      if (rc) return rc;
      return run_row(h, ROW_PHYS1, h->gs);
  }
  return fail(PALX_EINVAL, "unknown getter variable");
}

}  // extern "C"
