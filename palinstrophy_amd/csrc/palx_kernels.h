// palx_kernels.h -- the row and column pass kernels of one DNS stage (gfx950).
//
// One stage of the reference pipeline (turbo_simulator.py: dns_step2b ->
// dns_step3 / RK3 stage -> dns_step2a) is two HBM passes here:
//
//   row pass (k_row):    for each padded physical row z (M rows)
//       read  V[c](kx<N/2, z)   (z-inverse-transformed u_hat, w_hat; tiled by z)
//       x-inverse of u + i w (one complex FFT gives both real rows),
//       square: (u + i w)^2 = (u^2 - w^2) + 2i uw,
//       x-forward, split Hermitian halves, keep kx < N/2,
//       write P[c](z, kx<N/2) = F_x[u^2 - w^2], F_x[uw]  (tiled by kx)
//   column pass (k_col): for each kx < N/2
//       z-forward of both P planes, keep the N retained kz (kz=-N/2 forced 0),
//       nonlinear term, implicit update of omega/fnm1, velocity
//       reconstruction, zero-padded z-inverse of u_hat, w_hat -> V.
//
// F[u^2] - F[w^2] of the reference is computed as F[u^2 - w^2] (same value up
// to rounding), which lets the row pass carry two planes instead of three.
// The padded zeros of the 3/2 rule (rows N/2..N-1 and kx >= N/2) are never
// stored: they are a known-zero middle third inside the radix-3 end passes.
//
// Every kernel serves one member (`one`) or a batch of members (`tab`, one
// grid row per member: ensembles of equal padded length M in one launch).
#pragma once
#include <hip/hip_runtime.h>

#include "palx.h"
#include "palx_fft.h"
#include "palx_launch.h"

namespace palx {
namespace {


template <typename Tr>
constexpr int groups_per_block(int L, int V) { return groups_per_block_cs(L, V, sizeof(cx<Tr>)); }
// Per-pass twiddle tables (palx_fft.h) in the column pass only.  They cut the
// VALU instructions of both passes, but the row pass waits on LDS for a
// quarter of its cycles and their extra distinct reads slowed it (N=4096
// FP32 row 0.182 -> 0.192 ms), while the column pass, which waits on LDS
// for a few percent, gained (0.228 -> 0.224 ms).
constexpr bool kRowPT = false;
constexpr bool kColPT = true;
// variants of the non-step modes (init, getters, reconstruction, physical
// pass): the defaults where they fit, else the split exchange
template <typename Tr>
constexpr int nonstep_row_variant(int L) { return fitting_variant(L, kDefaultRowVariant, sizeof(cx<Tr>)); }
template <typename Tr>
constexpr int nonstep_col_variant(int L) { return fitting_variant(L, kDefaultColVariant, sizeof(cx<Tr>)); }
// variants of the adjoint passes: the length's default step variants where
// the table names them (the adjoint passes run the same transforms in the same
// alternation; N=512 FP32 adjoint 32.6 -> 31.6 ms with the column pass on
// variant 2 instead of 1), else those of the other non-step modes
template <typename Tr>
constexpr int adjoint_variant(int L, bool row) {
  for (const DefaultVariants& d : kDefaultStepVariants) {
    if (d.M != L) continue;
    const int v = sizeof(Tr) == 4 ? (row ? d.f32_row : d.f32_col) : (row ? d.f64_row : d.f64_col);
    if (v >= 0 && variant_fits(L, v, sizeof(cx<Tr>))) return v;
  }
  return row ? nonstep_row_variant<Tr>(L) : nonstep_col_variant<Tr>(L);
}

// threads per block of variant V (its launch bound: the register budget the
// compiler may use follows from it)
template <typename Tr>
constexpr int block_threads(int L, int V) { return groups_per_block<Tr>(L, V) * plan_of(L, V).T; }
// Minimum waves per SIMD the row pass is compiled for.  An FP64 row of the
// split exchange needs half the LDS of a complex one (48 KB at L = 6144), so
// two blocks fit per CU if their registers do: the bound holds the kernel to
// 128 VGPRs (from 134, a 28-byte spill) and doubles its resident waves.
// N=4096 FP64: row pass 0.52 -> 0.36 ms.  The column pass is not bounded:
// at 128 VGPRs it spills hundreds of bytes and slows down 2x.
template <typename Tr>
constexpr int row_min_waves(int L, int V) {
  return (sizeof(Tr) == 8 && plan_of(L, V).sp && block_threads<Tr>(L, V) <= 512 &&
          2 * block_lds_bytes(L, V, sizeof(cx<Tr>)) <= kMaxBlockLds)
             ? (2 * block_threads<Tr>(L, V) + 255) / 256
             : 1;
}

// ----------------------------------------------------------------- buffer access
// Strided HBM accesses go through raw buffer instructions: one 32-bit VGPR
// offset per thread plus a uniform (SGPR) offset per access, instead of a
// 64-bit address per element, and range-checked by the hardware (reads past
// `bytes` return 0, writes there are dropped).  Word 3 of the gfx9 buffer
// resource: raw 32-bit data format.
struct Buf {
  __amdgpu_buffer_rsrc_t r;
};
// A pointer every lane of the wave holds (a kernel argument, or a member-table
// entry selected by blockIdx.y), moved into scalar registers.  Without this the
// compiler cannot prove the buffer resource uniform and wraps every buffer
// access in a readfirstlane loop (about ten extra instructions per access;
// FP32 N=4096 row pass 0.202 -> 0.183 ms).  FP64 kernels keep the loop: with
// the scalar resource, the register-spilling FP64 column pass of variant 7 at
// L = 12288 returned values off by 1e-13 in a single-mode test, while the
// same source with the loop is exact; the FP64 passes gained nothing from the
// scalar form (N=4096 step 2.40 ms either way).
template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (T*)(((unsigned long long)hi << 32) | lo);
}
template <typename Tr>
__device__ __forceinline__ Buf make_buf(const void* p, size_t bytes) {
  if constexpr (sizeof(Tr) == 4) {
    p = uniform_ptr(p);
  } else {
    // FP64: the pointer is made opaque so that the resource stays a vector
    // value whatever the compiler proves about where the pointer came from (it
    // loads the member arguments into scalar registers once the kernel's early
    // loads are not behind a vector-load wait).  With the scalar resource the
    // FP64 column pass of variant 7 at L = 12288 returned stray values of
    // 1e-11 at kx = 0 that changed from run to run (variant 6 and every other
    // FP64 kernel exact), the same kernel that was off by 1e-13 in that form
    // earlier.
    asm volatile("" : "+v"(p));
  }
  return Buf{__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000)};
}
template <typename Tr>
__device__ __forceinline__ cx<Tr> bld(Buf b, unsigned voff, unsigned soff) {
  if constexpr (sizeof(Tr) == 4)
    return __builtin_bit_cast(cx<Tr>, __builtin_amdgcn_raw_buffer_load_b64(b.r, voff, soff, 0));
  else
    return __builtin_bit_cast(cx<Tr>, __builtin_amdgcn_raw_buffer_load_b128(b.r, voff, soff, 0));
}
template <typename Tr>
__device__ __forceinline__ void bst(cx<Tr> v, Buf b, unsigned voff, unsigned soff) {
  if constexpr (sizeof(Tr) == 4) {
    using V2 = __attribute__((ext_vector_type(2))) unsigned;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), b.r, voff, soff, 0);
  } else {
    using V4 = __attribute__((ext_vector_type(4))) unsigned;
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V4, v), b.r, voff, soff, 0);
  }
}

// The buffers of this block's member: `one` for a single field, tab[blockIdx.y]
// in a batched launch.  FP32: a uniform branch around the table read, so the
// arguments load into scalar registers (a select between the two addresses
// becomes one flat vector load, whose wait also holds up every load issued
// after it).  FP64 keeps the select (its code generation is left as tested).
template <typename Tr, typename A>
__device__ __forceinline__ A member_args(const A& one, const A* __restrict__ tab) {
  if constexpr (sizeof(Tr) == 4) {
    if (tab != nullptr) return tab[blockIdx.y];
    return one;
  } else {
    return tab ? tab[blockIdx.y] : one;
  }
}

// ----------------------------------------------------------------- row pass
// x-inverse of one packed pair of mixed planes (A + iB, each the half
// spectrum of a real row): planes tiled by z at `base` (A, then B at M*H),
// element k of row `row` at w_index(H, PW, k, row).  Follows pocketfft c2r
// semantics: the imaginary part of the kx=0 bin is dropped.  Result
// c[x] = a(x) + i b(x) at x = j + r*L/RL.  ONE: plane A only (B = 0).
// after_first() runs once the row's loads are issued and before the barrier
// that precedes the first pass with twiddles (the kernel's twiddle commit).
template <typename Tr, int L, int V, bool ONE, typename F>
__device__ __forceinline__ void row_inverse(const cx<Tr>* __restrict__ base, int row, F&& after_first,
                                            cx<Tr>* buf, const cx<Tr>* tws, int t,
                                            cx<Tr> (*out)[plan_of(L, V).g[plan_of(L, V).n - 1]]) {
  constexpr Plan P = plan_of(L, V);
  constexpr int T = P.T;
  constexpr int M = L, H = L / 3;
  constexpr int R1 = P.g[0];
  constexpr int NB1 = L / R1, PER1 = (NB1 + T - 1) / T;
  constexpr unsigned CS = sizeof(cx<Tr>);
  constexpr int PW = tile_width(M, M, CS);
  const Buf bw = make_buf<Tr>(base, (ONE ? 1 : 2) * work_plane(M) * CS);
  {
    cx<Tr> v[PER1][R1];
#pragma unroll
    for (int i = 0; i < PER1; ++i) {
      const int j = t + i * T;
      if (NB1 % T == 0 || j < NB1) {
        // element k*M + row: k = j + r*NB1 (low third), k = (R1-r)*NB1 - j (high third)
        // element k of the row at w_index(H, PW, k, row): linear in k with stride PW.
        // Every lane loads every element unconditionally (rows beyond M read row 0,
        // whose results are never stored), so the loads of a row issue back to back
        // and wait once; a load inside a lane-dependent branch waits before the next
        // branch.  The one element outside the half spectrum, k = H (j = 0 at
        // r = 2 R1 / 3), is read at k = H - 1 and replaced by zero.
        const unsigned vlo = w_index(H, PW, j, row) * CS;
        const unsigned vhi = w_index(H, PW, NB1 - j, row) * CS;
        const unsigned vhi0 = w_index(H, PW, NB1 - (j == 0 ? 1 : j), row) * CS;
        cx<Tr> U[R1], Wv[R1];
#pragma unroll
        for (int r = 0; r < R1; ++r) {
          if (r < R1 / 3 || r >= 2 * R1 / 3) {
            const bool lo = r < R1 / 3;
            const unsigned vo = lo ? vlo : (r == 2 * R1 / 3 ? vhi0 : vhi);
            const unsigned so = (unsigned)((lo ? r : R1 - r - 1) * NB1) * PW * CS;
            U[r] = bld<Tr>(bw, vo, so);
            Wv[r] = ONE ? cx<Tr>{0, 0} : bld<Tr>(bw, vo, so + (unsigned)(M * H) * CS);
          }
        }
#pragma unroll
        for (int r = 0; r < R1; ++r) {
          cx<Tr> c = {0, 0};
          if (r < R1 / 3) {  // C[k] = A[k] + i B[k]; k = 0 drops the imaginary parts
            c = (r == 0 && j == 0) ? cx<Tr>{U[r].x, Wv[r].x} : cx<Tr>{U[r].x - Wv[r].y, U[r].y + Wv[r].x};
          } else if (r >= 2 * R1 / 3) {  // C[M-k] = conj A[k] + i conj B[k]
            c = cx<Tr>{U[r].x + Wv[r].y, Wv[r].x - U[r].y};
            if (r == 2 * R1 / 3 && j == 0) c = cx<Tr>{0, 0};
          }
          v[i][r] = c;
        }
      }
    }
    if constexpr (P.sp) {
      after_first();
      transform_split<L, V, true, kRowPT>(buf, tws, t, v, out);
    } else {
      first_pass<L, V, true>(buf, t, v);
      after_first();
    }
  }
  if constexpr (!P.sp) {
    __syncthreads();
    mid_passes<L, V, true, kRowPT>(buf, tws, t);
    last_pass<L, V, true, kRowPT>(buf, tws, t, out);
  }
}

// x-forward of a packed real pair g = a + i b (registers at x = j + r*L/RL,
// NS=1 first pass) and split of the Hermitian halves for kx < H:
//   Fa = (G[k] + conj G[M-k]) / 2,  Fb = (G[k] - conj G[M-k]) / (2i) * bscale.
// Written tiled by kx (element (row, k) at p_index(M, PP, row, k), plane B at
// M*H) into Pout.
template <typename Tr, int L, int V>
__device__ __forceinline__ void row_forward_split(cx<Tr> (*v)[plan_of(L, V).g[plan_of(L, V).n - 1]], Tr bscale,
                                                  bool active, int row, cx<Tr>* buf, const cx<Tr>* tws,
                                                  int t, cx<Tr>* __restrict__ Pout) {
  constexpr Plan P = plan_of(L, V);
  constexpr int T = P.T;
  constexpr int M = L, H = L / 3, PP = tile_width(M, H, sizeof(cx<Tr>));
  constexpr int R1 = P.g[0];
  constexpr int NB1 = L / R1, PER1 = (NB1 + T - 1) / T;
  cx<Tr> d[PER1][R1];
  if constexpr (P.sp) {
    transform_split<L, V, false, kRowPT>(buf, tws, t, v, d);
  } else {
    __syncthreads();  // everyone finished reading buf
    first_pass<L, V, false>(buf, t, v);
    __syncthreads();
    mid_passes<L, V, false, kRowPT>(buf, tws, t);
    last_pass<L, V, false, kRowPT>(buf, tws, t, d);
  }
  __syncthreads();
  // element (row, k) of plane A at p_index(M, PP, row, k), plane B at M*H
  // after it.  FP32: buffer stores, lane offset of k = t and a scalar offset
  // per k = t + q*T (T a multiple of PP: the tile index grows by q*T/PP);
  // FP64: flat stores (its buffer resources are vector values)
  constexpr unsigned CS = sizeof(cx<Tr>);
  const Buf bp = make_buf<Tr>(Pout, 2 * work_plane(M) * CS);
  const unsigned poff = p_index(M, PP, row, t) * CS;
  auto store_p = [&](int q, int k, cx<Tr> A, cx<Tr> B) {
    if constexpr (sizeof(Tr) == 4 && T % PP == 0) {
      const unsigned so = (unsigned)(q * T) * M * CS;
      bst<Tr>(A, bp, poff, so);
      bst<Tr>(B, bp, poff, so + (unsigned)(M * H) * CS);
    } else {
      const unsigned e = p_index(M, PP, row, k);
      Pout[e] = A;
      Pout[(size_t)M * H + e] = B;
    }
  };
  if constexpr (P.sp) {  // the same exchange in two halves: real parts, then imaginary parts
    Tr* br = reinterpret_cast<Tr*>(buf);
    constexpr int KPT = (H + T - 1) / T;
    Tr dkx[KPT], dmx[KPT];
#pragma unroll
    for (int half = 0; half < 2; ++half) {
#pragma unroll
      for (int i = 0; i < PER1; ++i) {
        const int j = t + i * T;
        if (NB1 % T == 0 || j < NB1) {
#pragma unroll
          for (int r = 0; r < R1; ++r) {
            const int pos = j + r * NB1;
            const Tr x = half ? d[i][r].y : d[i][r].x;
            if (r < R1 / 3) br[split_idx<P.pad>(pos)] = x;
            else if (r >= 2 * R1 / 3) br[split_idx<P.pad>(pos - H)] = x;
          }
        }
      }
      __syncthreads();
      if (half == 0) {
#pragma unroll
        for (int q = 0; q < KPT; ++q) {
          const int k = t + q * T;
          if (k < H) {
            dkx[q] = br[split_idx<P.pad>(k)];
            dmx[q] = br[split_idx<P.pad>(k == 0 ? 0 : 2 * H - k)];
          }
        }
        __syncthreads();
      }
    }
    if (active) {
      const Tr hb = (Tr)0.5 * bscale;
#pragma unroll
      for (int q = 0; q < KPT; ++q) {
        const int k = t + q * T;
        if (k < H) {
          const cx<Tr> Dk = {dkx[q], br[split_idx<P.pad>(k)]};
          const cx<Tr> Dm = {dmx[q], br[split_idx<P.pad>(k == 0 ? 0 : 2 * H - k)]};
          const cx<Tr> A = {(Tr)0.5 * (Dk.x + Dm.x), (Tr)0.5 * (Dk.y - Dm.y)};
          const cx<Tr> B = {hb * (Dk.y + Dm.y), -hb * (Dk.x - Dm.x)};
          store_p(q, k, A, B);
        }
      }
    }
    return;
  }
  // retained outputs D[k] (k<H) -> buf[k];  D[M-k'] (k' in (0,H]) -> buf[2H-k']
#pragma unroll
  for (int i = 0; i < PER1; ++i) {
    const int j = t + i * T;
    if (NB1 % T == 0 || j < NB1) {
#pragma unroll
      for (int r = 0; r < R1; ++r) {
        const int pos = j + r * NB1;
        if (r < R1 / 3) buf[split_idx<P.pad>(pos)] = d[i][r];
        else if (r >= 2 * R1 / 3) buf[split_idx<P.pad>(pos - H)] = d[i][r];
      }
    }
  }
  __syncthreads();
  if (active) {
    const Tr hb = (Tr)0.5 * bscale;
    constexpr int KPT = (H + T - 1) / T;
#pragma unroll
    for (int q = 0; q < KPT; ++q) {
      const int k = t + q * T;
      if (H % T == 0 || k < H) {
        const cx<Tr> Dk = buf[split_idx<P.pad>(k)];
        const cx<Tr> Dm = buf[split_idx<P.pad>(k == 0 ? 0 : 2 * H - k)];
        const cx<Tr> A = {(Tr)0.5 * (Dk.x + Dm.x), (Tr)0.5 * (Dk.y - Dm.y)};
        const cx<Tr> B = {hb * (Dk.y + Dm.y), -hb * (Dk.x - Dm.x)};
        store_p(q, k, A, B);
      }
    }
  }
}

// W: velocity work planes, 2 x M*H complex tiled by z (w_index).
// Q (ROW_ADJ): adjoint planes in the same layout.
// Pout: products / adjoint planes, 2 x M*H complex tiled by kx (p_index).
// ur: 3 planes of M x M real.  part: per-row max(|u|+|w|) at [z],
// sum(u^2+w^2) at [M + z] (ROW_PHYS).
template <typename Tr, int L, int MODE, int V>
__global__ void __launch_bounds__(block_threads<Tr>(L, V), row_min_waves<Tr>(L, V))
    k_row(RowArgs one, const RowArgs* __restrict__ tab, const cx<Tr>* __restrict__ twg, int xcd_remap) {
  constexpr Plan P = plan_of(L, V);
  constexpr int T = P.T;
  constexpr int G = groups_per_block<Tr>(L, V);
  constexpr int M = L;
  constexpr int RL = P.g[P.n - 1];
  constexpr int NBL = L / RL, PERL = (NBL + T - 1) / T;
  extern __shared__ unsigned char smem_raw[];
  cx<Tr>* smem = reinterpret_cast<cx<Tr>*>(smem_raw);
  const RowArgs a = member_args<Tr>(one, tab);  // member of a batched launch
  const cx<Tr>* __restrict__ W = (const cx<Tr>*)a.W;
  const cx<Tr>* __restrict__ Q = (const cx<Tr>*)a.Q;
  cx<Tr>* __restrict__ Pout = (cx<Tr>*)a.Pout;
  Tr* __restrict__ ur = (Tr*)a.ur;
  double* __restrict__ part = a.part;

  int blk = blockIdx.x;
  if (xcd_remap) {  // consecutive rows (which read the same lines) on one XCD
    const int nb = gridDim.x;
    blk = (blk % 8) * (nb / 8) + blk / 8;
  }
  const int g = threadIdx.x / T, t = threadIdx.x % T;
  const int row = blk * G + g;
  const bool active = row < M;
  const int rows = active ? row : 0;
  cx<Tr>* buf = smem + g * buf_cx_len(L, V);
  cx<Tr>* tws = smem + G * buf_cx_len(L, V);
  TwiddleStage<L, V, kRowPT, block_threads<Tr>(L, V), Tr> tstage;
  tstage.fetch(twg);

  // c = u + i w on the padded row (x = j + r*NBL)
  cx<Tr> v[PERL][RL];
  row_inverse<Tr, L, V, MODE == ROW_PHYS1>(W, rows, [&] { tstage.commit(tws); }, buf, tws, t, v);

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
      constexpr int P2 = pow2_floor(T);
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
    row_forward_split<Tr, L, V>(v, (Tr)0.5, active, row, buf, tws, t, Pout);
  } else {  // ROW_ADJ: lambda_u = 2u qA + w qC, lambda_w = -2w qA + u qC
    cx<Tr> q[PERL][RL];
    __syncthreads();  // buf reused by the second inverse
    row_inverse<Tr, L, V, false>(Q, rows, [] {}, buf, tws, t, q);
#pragma unroll
    for (int i = 0; i < PERL; ++i) {
#pragma unroll
      for (int r = 0; r < RL; ++r) {
        const Tr u = v[i][r].x, w = v[i][r].y, qa = q[i][r].x, qc = q[i][r].y;
        v[i][r] = cx<Tr>{(Tr)2 * u * qa + w * qc, u * qc - (Tr)2 * w * qa};
      }
    }
    row_forward_split<Tr, L, V>(v, (Tr)1, active, row, buf, tws, t, Pout);
  }
}

// ----------------------------------------------------------------- column pass
// Layouts: Pin = products (or adjoint planes), 2 planes tiled by kx (p_index);
// om, fnm1 column-major [kx][zc] (N per column); src (COL_INIT) compact
// row-major [c][zc][kx]; out = velocity work planes (or adjoint planes) tiled
// by z (w_index), or the one-plane getter scratch in that layout.  A column is
// read in 64-byte steps and written in 64-byte chunks; the other elements of
// each tile belong to the neighbouring columns, which run on the same XCD at
// the same time and find the lines in L2.
//
// COL_RK3, stage = st | flags: RK3_PRE_IN reads the stage's pre-update term
// from fnm1's storage in place of omega and fnm1; RK3_PRE_OUT stores the next
// stage's term there in place of omega and fn (palx_launch.h, rk3_step_stage).
// COL_ADJ (adjoint of one LS-IMEX-RK3 stage; om/fnm1 hold lambda_omega /
// lambda_fnm1, stage = st | flags):
//   ADJ_HAS_G : acc = w'(conj c_u F_z[Gu] + conj c_w F_z[Gw])   (C2R adjoint, w' = 1|2)
//   ADJ_UPDATE: lambda_om <- D1 lambda_om + acc, lambda_fnm1 <- D3 lambda_om  (stage st)
//   ADJ_WANT_Q: lambda_fn = D2 lambda_om + lambda_fnm1 for stage (st+2)%3, then
//               out = z-inverse of (GA, G2-A2) divxz lambda_fn, halved for kx>=1 (R2C adjoint)
// Minimum waves per SIMD the column pass is compiled for: FP32 columns of
// 512-thread blocks up to L = 6144 are held to 128 VGPRs, so that two blocks
// share a CU where their LDS allows it.  The padded and the prefetching
// variants of the alternative 6144-point plan use 130 unbounded, which
// leaves one block per CU (0.30 vs 0.24 ms); bounded they take 128 (the
// prefetching one with a 2-register spill).  Longer FP32 transforms (radix-24
// first pass at L = 12288: 158-201 VGPRs) would spill 20-100 registers, and
// FP64 columns hundreds of bytes, so those are not bounded.
template <typename Tr>
constexpr int col_min_waves(int L, int V) {
  return (sizeof(Tr) == 4 && L <= 6144 && block_threads<Tr>(L, V) == 512 &&
          2 * block_lds_bytes(L, V, sizeof(cx<Tr>)) <= kMaxBlockLds)
             ? 4
             : 1;
}
template <typename Tr, int L, int MODE, int V>
__global__ void __launch_bounds__(block_threads<Tr>(L, V), col_min_waves<Tr>(L, V))
    k_col(ColArgs one, const ColArgs* __restrict__ tab, const cx<Tr>* __restrict__ twg, int stage,
          int xcd_remap) {
  constexpr Plan P = plan_of(L, V);
  constexpr int T = P.T;
  constexpr int G = groups_per_block<Tr>(L, V);
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
  const ColArgs a = member_args<Tr>(one, tab);  // member of a batched launch
  const cx<Tr>* __restrict__ Pin = (const cx<Tr>*)a.Pin;
  cx<Tr>* __restrict__ om = (cx<Tr>*)a.om;
  cx<Tr>* __restrict__ fnm1 = (cx<Tr>*)a.fnm1;
  const cx<Tr>* __restrict__ src = (const cx<Tr>*)a.src;
  cx<Tr>* __restrict__ out = (cx<Tr>*)a.out;
  const double* __restrict__ scal = a.scal;

  int blk = blockIdx.x;
  if (xcd_remap) {  // consecutive columns on one XCD at the same time share L2 lines
    const int nb = gridDim.x;
    blk = (blk % 8) * (nb / 8) + blk / 8;
  }
  const int g = threadIdx.x / T;
  int t = threadIdx.x % T;
  const int kx = blk * G + g;
  const bool active = kx < H;
  const int kxs = active ? kx : 0;
  cx<Tr>* buf = smem + g * buf_cx_len(L, V);
  cx<Tr>* tws = smem + G * buf_cx_len(L, V);
  // the twiddle table's loads are issued first and written to LDS after the
  // first field's loads and first pass (or before the inverse transforms when
  // there is no forward part)
  TwiddleStage<L, V, kColPT, block_threads<Tr>(L, V), Tr> tstage;
  tstage.fetch(twg);
  bool tw_committed = false;
  const Tr ax = (Tr)kxs;
  cx<Tr>* omc = om + (size_t)kxs * N;
  cx<Tr>* fnc = fnm1 + (size_t)kxs * N;
  const int st = stage & 15;
  const int flags = (ADJ || MODE == COL_RK3) ? (stage & ~15) : 0;
  const bool pre_in = MODE == COL_RK3 && (flags & RK3_PRE_IN);
  const bool pre_out = MODE == COL_RK3 && (flags & RK3_PRE_OUT);
  const bool want_q = !ADJ || (flags & ADJ_WANT_Q);
  const int NF_OUT = (MODE == COL_OMEGA || MODE == COL_STREAM) ? 1 : (want_q ? 2 : 0);
  constexpr unsigned CS = sizeof(cx<Tr>);
  constexpr int PP = tile_width(M, H, CS), PW = tile_width(M, M, CS);
  const Buf bin = make_buf<Tr>(Pin, 2 * work_plane(M) * CS);
  const Buf bout = make_buf<Tr>(out, (size_t)(NF_OUT > 0 ? NF_OUT : 1) * work_plane(M) * CS);

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

  // ---- spectral update (turbo_simulator.py:1731-1771 CNAB2, 1912-1943 RK3),
  // first half: pre = omega (1 - K2 bdv) + fnm1 bdt (RK3; CNAB2: omega -
  // omega K2 VT + fnm1 c3), the part of the numerator without the new
  // nonlinear term; the second half adds fn c_fn and divides by den.  FP64
  // does the first half while the first product plane's loads are in flight,
  // so the loads of omega and fnm1 are off the path between the transforms
  // and the update (N=4096 FP64 column pass 0.425 -> 0.411 ms); in FP32 the
  // registers pre holds through the transforms push the column pass over its
  // two-blocks-per-CU budget (128 VGPRs and a spill: 0.198 -> 0.208 ms), so
  // FP32 does it after the transforms.
  constexpr bool kEarlyUpdate = sizeof(Tr) == 8;
  struct {
    Tr VT = 0, c3 = 0, bdv = 0, bdt = 0, adv = 0, cfn = 0, nbdv = 0, nbdt = 0;
  } upd;
  if constexpr (MODE == COL_CNAB2) {
    upd.VT = (Tr)0.5 * visc * dt;
    upd.cfn = (Tr)0.5 * dt * ((Tr)2 + cnm1);
    upd.c3 = -(Tr)0.5 * dt * cnm1;
  } else if constexpr (MODE == COL_RK3) {
    const Tr al = alpha(st), be = beta(st);
    upd.bdv = be * dt * visc;
    upd.cfn = al * dt;
    upd.bdt = be * dt;
    upd.adv = al * dt * visc;
    const Tr bn = beta(st == 2 ? 0 : st + 1);  // stage st+1's, for RK3_PRE_OUT
    upd.nbdv = bn * dt * visc;
    upd.nbdt = bn * dt;
  }
  // RK3 pre of a mode: omega (1 - K2 bdv) + fnm1 bdt, for the stage that
  // reads omega / fnm1 and for the stage before it that forms pre
  // (RK3_PRE_OUT).  The fused multiply-adds are written out: left to
  // contraction, the two sites fused different products (FP64 omega 1 ulp
  // apart after two steps).
  auto rk3_pre = [](cx<Tr> w0, cx<Tr> f0, Tr K2, Tr bdv, Tr bdt) {
    auto fm = [](Tr a, Tr b, Tr c) -> Tr {
      if constexpr (sizeof(Tr) == 4) return __builtin_fmaf(a, b, c);
      else return __builtin_fma(a, b, c);
    };
    const Tr tmp = fm(-K2, bdv, (Tr)1);
    const Tr fx = f0.x * bdt, fy = f0.y * bdt;
    return cx<Tr>{fm(w0.x, tmp, fx), fm(w0.y, tmp, fy)};
  };
  // FP32: omega / fnm1 of mode zc through buffer instructions, lane offset
  // t and a scalar offset per mode (column kxs and the mode's constant part);
  // FP64 keeps the flat pointers (its resources are vector values, and every
  // buffer access would run a readfirstlane loop)
  const Buf bom = make_buf<Tr>(om, (size_t)N * H * CS), bfn = make_buf<Tr>(fnm1, (size_t)N * H * CS);
  auto mode_soff = [&](int i, int q) {
    const int zc0 = q < R1 / 3 ? q * NB1 : (q + R1 / 3) * NB1 - H;  // zc - j
    return ((unsigned)kxs * N + (unsigned)(zc0 + i * T)) * CS;
  };
  cx<Tr> pre[PER1][NR];
  auto prepare_update = [&]() {
    if constexpr (STEP) {
#pragma unroll
      for (int i = 0; i < PER1; ++i) {
        const int j = t + i * T;
        if ((NB1 % T == 0 || j < NB1) && active) {
          const unsigned joff = (unsigned)t * CS;
          if (pre_in) {  // pre formed by the stage before
#pragma unroll
            for (int q = 0; q < NR; ++q) {
              const int pos = j + (q < R1 / 3 ? q : q + R1 / 3) * NB1;
              const int zc = q < R1 / 3 ? pos : pos - H;
              if constexpr (sizeof(Tr) == 4) pre[i][q] = bld<Tr>(bfn, joff, mode_soff(i, q));
              else pre[i][q] = fnc[zc];
            }
            continue;
          }
          // RK3 stage 0 has beta = 0: its fnm1 term vanishes, so fnm1 is not read
          // (a tenth of that column pass's HBM reads)
          const bool need_f0 = !(MODE == COL_RK3 && st == 0);
          cx<Tr> w0s[NR], f0s[NR];
#pragma unroll
          for (int q = 0; q < NR; ++q) {  // all loads first, then the arithmetic
            const int pos = j + (q < R1 / 3 ? q : q + R1 / 3) * NB1;
            const int zc = q < R1 / 3 ? pos : pos - H;
            if constexpr (sizeof(Tr) == 4) {
              w0s[q] = bld<Tr>(bom, joff, mode_soff(i, q));
              f0s[q] = need_f0 ? bld<Tr>(bfn, joff, mode_soff(i, q)) : cx<Tr>{0, 0};
            } else {
              w0s[q] = omc[zc];
              f0s[q] = need_f0 ? fnc[zc] : cx<Tr>{0, 0};
            }
          }
#pragma unroll
          for (int q = 0; q < NR; ++q) {
            const int pos = j + (q < R1 / 3 ? q : q + R1 / 3) * NB1;
            const Tr gz = (Tr)(q < R1 / 3 ? pos : pos - M);
            const Tr K2 = ax * ax + gz * gz;
            const cx<Tr> w0 = w0s[q], f0 = f0s[q];
            if constexpr (MODE == COL_CNAB2) {
              const Tr ARG = K2 * upd.VT;
              pre[i][q] = cx<Tr>{(w0.x - w0.x * ARG) + f0.x * upd.c3, (w0.y - w0.y * ARG) + f0.y * upd.c3};
            } else {
              pre[i][q] = rk3_pre(w0, f0, K2, upd.bdv, upd.bdt);
            }
          }
        }
      }
    }
  };

  if (STEP || (ADJ && (flags & ADJ_HAS_G))) {
    // plane f tiled by kx: element (z, kx) at p_index(M, PP, z, kx), linear in z with stride PP
    cx<Tr> v[PERL][RL];
    auto load_field = [&](int f) {
#pragma unroll
      for (int i = 0; i < PERL; ++i) {
        const int j = t + i * T;
        if (NBL % T == 0 || j < NBL) {
          const unsigned vo = p_index(M, PP, j, kxs) * CS;
          // unconditional (an inactive group reads column 0 and stores nothing), so
          // the loads are not split by branches that each end in a wait
#pragma unroll
          for (int r = 0; r < RL; ++r)
            v[i][r] = bld<Tr>(bin, vo, ((unsigned)f * M * H + (unsigned)(r * NBL) * PP) * CS);
        }
      }
    };
    if constexpr (P.pf) load_field(0);
#pragma unroll 1
    for (int f = 0; f < 2; ++f) {
      asm volatile("" : "+v"(t));  // index math stays inside the loop: fewer live registers
      if constexpr (!P.pf) load_field(f);
      if (kEarlyUpdate && f == 0) prepare_update();
      cx<Tr> d[PER1][R1];
      if constexpr (P.sp) {
        if (f == 0) tstage.commit(tws);
        transform_split<L, V, false, kColPT>(buf, tws, t, v, d);  // starts with a barrier
      } else {
        if (f == 1) __syncthreads();  // field 0 finished reading buf
        first_pass<L, V, false>(buf, t, v);
        if (f == 0) tstage.commit(tws);  // before field 1's loads: its wait leaves them in flight
        if constexpr (P.pf) {
          if (f == 0) load_field(1);  // field 1's strided loads fly during field 0's passes
        }
        __syncthreads();
        mid_passes<L, V, false, kColPT>(buf, tws, t);
        last_pass<L, V, false, kColPT>(buf, tws, t, d);
      }
#pragma unroll
      for (int i = 0; i < PER1; ++i) {
        const int j = t + i * T;
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          const int r = q < R1 / 3 ? q : q + R1 / 3;
          const int pos = j + r * NB1;
          const bool lo = q < R1 / 3;  // known after unrolling: pos < H exactly when lo (j < NB1)
          const int zc = lo ? pos : pos - H;
          const Tr gz = (Tr)(lo ? pos : pos - M);
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
    tw_committed = true;
  }

  if constexpr (STEP) {
    if constexpr (!kEarlyUpdate) prepare_update();
    // ---- spectral update, second half: num = pre + fn c_fn, omega = num / den
#pragma unroll
    for (int i = 0; i < PER1; ++i) {
      const int j = t + i * T;
      if ((NB1 % T == 0 || j < NB1) && active) {
        const unsigned joff = (unsigned)t * CS;
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          const int r = q < R1 / 3 ? q : q + R1 / 3;
          const int pos = j + r * NB1;
          const bool lo = q < R1 / 3;
          const int zc = lo ? pos : pos - H;
          const Tr gz = (Tr)(lo ? pos : pos - M);
          const Tr K2 = ax * ax + gz * gz;
          cx<Tr> f1 = {0, 0};
          if (zc != H) f1 = cx<Tr>{acc[i][q].x * divxz, acc[i][q].y * divxz};
          const cx<Tr> num = {pre[i][q].x + f1.x * upd.cfn, pre[i][q].y + f1.y * upd.cfn};
          const Tr inv = (Tr)1 / (MODE == COL_CNAB2 ? K2 * upd.VT + (Tr)1 : K2 * upd.adv + (Tr)1);
          const cx<Tr> wn = {num.x * inv, num.y * inv};
          if (pre_out) {  // the next stage's pre in fnm1's storage; omega's storage untouched
            const cx<Tr> pn = rk3_pre(wn, f1, K2, upd.nbdv, upd.nbdt);
            if constexpr (sizeof(Tr) == 4) bst<Tr>(pn, bfn, joff, mode_soff(i, q));
            else fnc[zc] = pn;
          } else if constexpr (sizeof(Tr) == 4) {
            bst<Tr>(wn, bom, joff, mode_soff(i, q));
            bst<Tr>(f1, bfn, joff, mode_soff(i, q));
          } else {
            omc[zc] = wn;
            fnc[zc] = f1;
          }
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
          const bool lo = q < R1 / 3;
          const int zc = lo ? pos : pos - H;
          if constexpr (MODE == COL_RECON) {
            acc[i][q] = omc[zc];
          } else {
            const Tr gz = (Tr)(lo ? pos : pos - M);
            const Tr K2 = ax * ax + gz * gz;
            cx<Tr> lom = omc[zc], lf = fnc[zc];
            if (flags & ADJ_UPDATE) {
              const Tr inv = (Tr)1 / (K2 * adv + (Tr)1);
              const Tr D1 = ((Tr)1 - K2 * bdv) * inv;
              const Tr D3 = bdt * inv;
              const cx<Tr> g = acc[i][q];
              lf = cx<Tr>{lom.x * D3, lom.y * D3};
              lom = cx<Tr>{lom.x * D1 + g.x, lom.y * D1 + g.y};
              omc[zc] = lom;
              fnc[zc] = lf;
            }
            if (flags & ADJ_WANT_Q) {  // lambda_fn of stage (st+2)%3 = D2 lambda_om + lambda_fnm1
              const Tr invp = (Tr)1 / (K2 * padv + (Tr)1);
              const Tr D2 = padt * invp;
              acc[i][q] = cx<Tr>{lom.x * D2 + lf.x, lom.y * D2 + lf.y};
            }
          }
        }
      }
    }
  }

  if (!tw_committed) tstage.commit(tws);  // read after the barriers of the inverse's first pass

  // ---- zero-padded z-inverse of the retained spectra -> out[f][kx][z]
#pragma unroll 1
  for (int f = 0; f < NF_OUT; ++f) {
    asm volatile("" : "+v"(t));
    cx<Tr> vo[PERL][RL];
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
            const bool lo = r < R1 / 3;
            const int zc = lo ? pos : pos - H;
            const Tr gz = (Tr)(lo ? pos : pos - M);
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
      if constexpr (P.sp) {
        transform_split<L, V, true, kColPT>(buf, tws, t, v, vo);  // starts with a barrier
      } else {
        __syncthreads();  // previous readers of buf are done
        first_pass<L, V, true>(buf, t, v);
      }
    }
    if constexpr (!P.sp) {
      __syncthreads();
      mid_passes<L, V, true, kColPT>(buf, tws, t);
      last_pass<L, V, true, kColPT>(buf, tws, t, vo);
    }
    if (active) {  // plane f tiled by z: element (kx, z = j + r*NBL) at w_index(H, PW, kx, z)
#pragma unroll
      for (int i = 0; i < PERL; ++i) {
        const int j = t + i * T;
        if (NBL % T == 0 || j < NBL) {
          if constexpr (NBL % PW == 0) {  // z / PW = j / PW + r * NBL / PW, z % PW = j % PW
            const unsigned off = w_index(H, PW, kxs, j) * CS;
#pragma unroll
            for (int r = 0; r < RL; ++r)
              bst<Tr>(vo[i][r], bout, off, ((unsigned)f * M * H + (unsigned)(r * NBL) * H) * CS);
          } else {
#pragma unroll
            for (int r = 0; r < RL; ++r)
              bst<Tr>(vo[i][r], bout, w_index(H, PW, kxs, j + r * NBL) * CS, ((unsigned)f * M * H) * CS);
          }
        }
      }
    }
  }
}

// ----------------------------------------------------------------- launchers
template <typename Tr, int L, int V>
struct Pass {
  static constexpr Plan P = plan_of(L, V);
  static constexpr int G = groups_per_block<Tr>(L, V);
  static size_t lds(bool pt) { return ((size_t)G * buf_cx_len(L, V) + tw_len(L, V, pt)) * sizeof(cx<Tr>); }
  static dim3 block() { return dim3(G * P.T); }

  template <int MODE>
  static hipError_t row(const RowArgs& one, const RowArgs* tab, int nmem, const void* tw, hipStream_t s) {
    static const hipError_t attr = hipFuncSetAttribute((const void*)k_row<Tr, L, MODE, V>,
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds(kRowPT));
    if (attr != hipSuccess) return attr;
    const int nb = (L + G - 1) / G;
    const int remap = (nb % 8 == 0 && nb >= 64) ? 1 : 0;
    hipLaunchKernelGGL((k_row<Tr, L, MODE, V>), dim3(nb, nmem), block(), lds(kRowPT), s, one, tab,
                       (const cx<Tr>*)tw, remap);
    return hipGetLastError();
  }

  template <int MODE>
  static hipError_t col(int stage, const ColArgs& one, const ColArgs* tab, int nmem, const void* tw,
                        hipStream_t s) {
    static const hipError_t attr = hipFuncSetAttribute((const void*)k_col<Tr, L, MODE, V>,
                                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds(kColPT));
    if (attr != hipSuccess) return attr;
    const int nb = (L / 3 + G - 1) / G;
    const int remap = (nb % 8 == 0 && nb >= 64) ? 1 : 0;
    hipLaunchKernelGGL((k_col<Tr, L, MODE, V>), dim3(nb, nmem), block(), lds(kColPT), s, one, tab,
                       (const cx<Tr>*)tw, stage, remap);
    return hipGetLastError();
  }
};

// variant V of the step modes, only where compiled for L
#define PALX_VARIANT_CASES(CALL)                                            \
  case 0: if constexpr (variant_fits(L, 0, sizeof(cx<Tr>))) return CALL(0); break; \
  case 1: if constexpr (variant_fits(L, 1, sizeof(cx<Tr>))) return CALL(1); break; \
  case 2: if constexpr (variant_fits(L, 2, sizeof(cx<Tr>))) return CALL(2); break; \
  case 3: if constexpr (variant_fits(L, 3, sizeof(cx<Tr>))) return CALL(3); break; \
  case 4: if constexpr (variant_fits(L, 4, sizeof(cx<Tr>))) return CALL(4); break; \
  case 5: if constexpr (variant_fits(L, 5, sizeof(cx<Tr>))) return CALL(5); break; \
  case 6: if constexpr (variant_fits(L, 6, sizeof(cx<Tr>))) return CALL(6); break; \
  case 7: if constexpr (variant_fits(L, 7, sizeof(cx<Tr>))) return CALL(7); break;

template <typename Tr, int L>
hipError_t row_len(int V, int mode, const RowArgs& one, const RowArgs* tab, int nmem, const void* tw,
                   hipStream_t s) {
  constexpr int D = nonstep_row_variant<Tr>(L);
  switch (mode) {
    case ROW_STEP:
      switch (V) {
#define PALX_ROW_STEP(v_) Pass<Tr, L, v_>::template row<ROW_STEP>(one, tab, nmem, tw, s)
        PALX_VARIANT_CASES(PALX_ROW_STEP)
#undef PALX_ROW_STEP
      }
      return hipErrorInvalidValue;
    case ROW_PHYS: return Pass<Tr, L, D>::template row<ROW_PHYS>(one, tab, nmem, tw, s);
    case ROW_PHYS1: return Pass<Tr, L, D>::template row<ROW_PHYS1>(one, tab, nmem, tw, s);
    case ROW_ADJ: return Pass<Tr, L, adjoint_variant<Tr>(L, true)>::template row<ROW_ADJ>(one, tab, nmem, tw, s);
  }
  return hipErrorInvalidValue;
}

template <typename Tr, int L, int MODE>
hipError_t col_step_len(int V, int stage, const ColArgs& one, const ColArgs* tab, int nmem, const void* tw,
                        hipStream_t s) {
  switch (V) {
#define PALX_COL_STEP(v_) Pass<Tr, L, v_>::template col<MODE>(stage, one, tab, nmem, tw, s)
    PALX_VARIANT_CASES(PALX_COL_STEP)
#undef PALX_COL_STEP
  }
  return hipErrorInvalidValue;
}

template <typename Tr, int L>
hipError_t col_len(int V, int mode, int stage, const ColArgs& one, const ColArgs* tab, int nmem,
                   const void* tw, hipStream_t s) {
  constexpr int D = nonstep_col_variant<Tr>(L);
  switch (mode) {
    case COL_CNAB2: return col_step_len<Tr, L, COL_CNAB2>(V, stage, one, tab, nmem, tw, s);
    case COL_RK3: return col_step_len<Tr, L, COL_RK3>(V, stage, one, tab, nmem, tw, s);
    case COL_INIT: return Pass<Tr, L, D>::template col<COL_INIT>(stage, one, tab, nmem, tw, s);
    case COL_OMEGA: return Pass<Tr, L, D>::template col<COL_OMEGA>(stage, one, tab, nmem, tw, s);
    case COL_STREAM: return Pass<Tr, L, D>::template col<COL_STREAM>(stage, one, tab, nmem, tw, s);
    case COL_RECON: return Pass<Tr, L, D>::template col<COL_RECON>(stage, one, tab, nmem, tw, s);
    case COL_ADJ: return Pass<Tr, L, adjoint_variant<Tr>(L, false)>::template col<COL_ADJ>(stage, one, tab, nmem, tw, s);
  }
  return hipErrorInvalidValue;
}

#define PALX_FOR_EACH_LENGTH(X) X(12) X(24) X(48) X(96) X(192) X(384) X(768) X(1536) X(3072) X(6144) X(12288) \
  X(72) X(144) X(288) X(576) X(1152) X(2304) X(4608) X(9216) X(18432)

// lengths compiled for FP32 only (their FP64 transform does not fit in LDS,
// even split); 12288 and 18432 exist in FP64 with the split exchange only
#define PALX_FOR_EACH_LENGTH_F32(X) X(24576)

template <typename Tr>
hipError_t launch_row_any(int M, int V, int mode, const RowArgs& one, const RowArgs* tab, int nmem,
                          const void* tw, hipStream_t s) {
  switch (M) {
#define PALX_CASE(L_) case L_: return row_len<Tr, L_>(V, mode, one, tab, nmem, tw, s);
    PALX_FOR_EACH_LENGTH(PALX_CASE)
#undef PALX_CASE
#define PALX_CASE(L_) \
  case L_: if constexpr (sizeof(Tr) == 4) return row_len<Tr, L_>(V, mode, one, tab, nmem, tw, s); break;
    PALX_FOR_EACH_LENGTH_F32(PALX_CASE)
#undef PALX_CASE
  }
  return hipErrorInvalidValue;
}

template <typename Tr>
hipError_t launch_col_any(int M, int V, int mode, int stage, const ColArgs& one, const ColArgs* tab,
                          int nmem, const void* tw, hipStream_t s) {
  switch (M) {
#define PALX_CASE(L_) case L_: return col_len<Tr, L_>(V, mode, stage, one, tab, nmem, tw, s);
    PALX_FOR_EACH_LENGTH(PALX_CASE)
#undef PALX_CASE
#define PALX_CASE(L_) \
  case L_: if constexpr (sizeof(Tr) == 4) return col_len<Tr, L_>(V, mode, stage, one, tab, nmem, tw, s); break;
    PALX_FOR_EACH_LENGTH_F32(PALX_CASE)
#undef PALX_CASE
  }
  return hipErrorInvalidValue;
}

}  // namespace
}  // namespace palx
