// palx_launch.h -- launch interface of the row / column pass kernels.
//
// The FFT pass kernels are compiled in their own translation units, one per
// precision (palx_launch.hip built with PALX_REAL=float and =double), so the
// many template instances build in parallel.  The host code of libpalx calls
// them only through the launchers declared here.
#pragma once
#include <hip/hip_runtime.h>

namespace palx {

enum RowMode { ROW_STEP = 0, ROW_PHYS = 1, ROW_PHYS1 = 2, ROW_ADJ = 3 };
enum ColMode {
  COL_CNAB2 = 0, COL_RK3 = 1, COL_INIT = 2, COL_OMEGA = 3, COL_STREAM = 4,
  COL_RECON = 5,  // velocity planes from omega (reconstruction + padded z-inverse)
  COL_ADJ = 6     // adjoint of one RK3 stage, column half
};
// flags of COL_ADJ, packed above the stage index
enum { ADJ_HAS_G = 16, ADJ_UPDATE = 32, ADJ_WANT_Q = 64 };
// flags of COL_RK3 inside one full step, packed above the stage index.  The
// update of stage s needs only pre_s = omega (1 - K2 beta_s dt visc) + fnm1
// beta_s dt, which the column pass of stage s-1 can form from the omega and fn
// it holds.  RK3_PRE_OUT: store pre_{s+1} in fnm1's storage in place of fn and
// leave omega's storage alone; RK3_PRE_IN: read pre_s from fnm1's storage in
// place of omega and fnm1.  A step runs stage 0 with PRE_OUT, 1 with both,
// 2 with PRE_IN (which stores omega and fn as the plain stage does): four of
// the eleven omega / fnm1 plane transfers of a step fewer, the same values.
enum { RK3_PRE_IN = 128, RK3_PRE_OUT = 256 };
constexpr int rk3_step_stage(int st) {
  return st | (st > 0 ? RK3_PRE_IN : 0) | (st < 2 ? RK3_PRE_OUT : 0);
}

// Layout of the two work planes that each pass reads across the other
// pass's transform direction.  A plane is stored in tiles of P consecutive
// elements of the direction its writer walks (P complex = 64 bytes: 8 FP32,
// 4 FP64; 32 bytes, 4 FP32, on FP32 grids with M <= 3072; halved until it
// divides the extent), so the writer stores P-element chunks and the reader,
// one element per lane at a P-element stride, touches half (a quarter) as
// many 128-byte lines per instruction as lanes instead of one line per lane:
//   velocity / adjoint planes W (written by the column pass along z, read by
//   the row pass along kx): element (kx, z) at ((z/PW)*H + kx)*PW + z%PW,
//   PW = tile_width(M, M, csz);
//   product planes P (written by the row pass along kx, read by the column
//   pass along z): element (z, kx) at ((kx/PP)*M + z)*PP + kx%PP,
//   PP = tile_width(M, H, csz).
// Each plane holds M*H complex.  At N=4096 this took the FP32 RK3 step from
// 1.59 to 1.35 ms against the column-major / row-major planes (one 8- or
// 16-byte element per 128-byte line per lane in the strided reads).  32-byte
// FP32 tiles measured 3-6% faster at N=1024/2048 and 5-8% slower at
// N=4096/8192; 32-byte FP64 tiles were slower at every size.
constexpr int tile_width(int M, int extent, int csz) {
  int p = (csz == 8 && M <= 3072 ? 32 : 64) / csz;
  while (p > 1 && extent % p != 0) p /= 2;
  return p;
}
constexpr size_t work_plane(int M) { return (size_t)M * (M / 3); }
__host__ __device__ constexpr unsigned w_index(int H, int PW, unsigned kx, unsigned z) {
  return ((z / PW) * H + kx) * PW + z % PW;
}
__host__ __device__ constexpr unsigned p_index(int M, int PP, unsigned z, unsigned kx) {
  return ((kx / PP) * M + z) * PP + kx % PP;
}

// Buffers of one member for a row pass.
//   W:    velocity planes, 2 x M*H complex, tiled by z (w_index)
//   Q:    adjoint planes in the layout of W (ROW_ADJ)
//   Pout: products / adjoint planes, 2 x M*H complex, tiled by kx (p_index)
//   ur:   3 planes of M x M real (ROW_PHYS, ROW_PHYS1)
//   part: per-row max(|u|+|w|) at [z], sum(u^2+w^2) at [M + z] (ROW_PHYS)
struct RowArgs {
  const void* W;
  const void* Q;
  void* Pout;
  void* ur;
  double* part;
};

// Buffers of one member for a column pass.
//   Pin:  products (or adjoint planes), 2 x M*H complex, tiled by kx (p_index)
//   om, fnm1: column-major [kx][zc] (N per column)
//   src:  compact velocity spectra [c][zc][kx] (COL_INIT)
//   out:  velocity planes (or adjoint planes), 2 x M*H complex tiled by z
//         (w_index), or the one-plane getter scratch in the same layout
//   scal: FP64 scalar block (dt, cnm1, visc, ...)
struct ColArgs {
  const void* Pin;
  void* om;
  void* fnm1;
  const void* src;
  void* out;
  const double* scal;
};

// Number of kernel variants (plan / LDS layout) compiled for the hot step
// modes (ROW_STEP, COL_CNAB2, COL_RK3) at padded length M; other modes always
// use the default variant.
int n_pass_variants(int M);
// Variants of the non-step modes (always these): the row pass pads its LDS
// rows, the column pass does not.
constexpr int kDefaultRowVariant = 0;
constexpr int kDefaultColVariant = 1;
// Step-pass variants a handle starts with (nvar = n_pass_variants(M)): the
// autotuner's choice per padded length and precision from one MI355X sweep
// (single field, each pass timed inside the row/column alternation of RK3
// stages, 10 launches; entries whose best and second-best variants were
// within 1% keep the earlier sweep's pick).  The larger differences it found
// against one fixed rule: the base plan with the prefetching layout for the
// 9*2^k column passes (N=6144 FP32 column 0.536 vs 0.603 ms, FP64 1.48 vs
// 1.61 ms), the split exchange for FP32 rows at N=8192 (0.69 vs 1.15 ms) and
// FP64 rows at N=4096 (0.29 vs 0.46 ms, two blocks per CU: row_min_waves),
// where one row fills a CU's LDS otherwise; the alternative plan for FP32
// N=4096 (column 0.192 vs 0.270 ms) and N=1024.  Lengths not listed keep the
// base plan.  The autotuner replaces these per device and size.
struct DefaultVariants {
  int M, f32_row, f32_col, f64_row, f64_col;
};
constexpr DefaultVariants kDefaultStepVariants[] = {
    {384, 3, 5, 3, 5},   {768, 2, 2, 2, 2},   {1152, 0, 2, 2, 2}, {1536, 3, 5, 5, 5},
    {2304, 0, 0, 0, 2},  {3072, 0, 0, 2, 2},  {4608, 2, 0, 1, 0}, {6144, 5, 3, 7, 3},
    {9216, 2, 2, 2, 2},  {12288, 7, 2, 7, 7}, {18432, 2, 0, 6, 6}, {24576, 7, 6, -1, -1}};
inline void default_step_variants(int M, int nvar, int precision, int* vrow, int* vcol) {
  *vrow = kDefaultRowVariant;
  *vcol = kDefaultColVariant;
  for (const DefaultVariants& d : kDefaultStepVariants) {
    if (d.M != M) continue;
    const int r = precision == 4 ? d.f32_row : d.f64_row;
    const int c = precision == 4 ? d.f32_col : d.f64_col;
    if (r >= 0 && r < nvar && c >= 0 && c < nvar) {
      *vrow = r;
      *vcol = c;
    }
  }
}

// Launch one row / column pass of padded length M with variant V.
// tab == nullptr: a single member described by `one` (grid.y = 1);
// otherwise tab[0..nmem) in device memory, one grid row per member.
hipError_t launch_row_f32(int M, int V, int mode, const RowArgs& one, const RowArgs* tab, int nmem,
                          const void* tw, hipStream_t s);
hipError_t launch_col_f32(int M, int V, int mode, int stage, const ColArgs& one, const ColArgs* tab,
                          int nmem, const void* tw, hipStream_t s);
hipError_t launch_row_f64(int M, int V, int mode, const RowArgs& one, const RowArgs* tab, int nmem,
                          const void* tw, hipStream_t s);
hipError_t launch_col_f64(int M, int V, int mode, int stage, const ColArgs& one, const ColArgs* tab,
                          int nmem, const void* tw, hipStream_t s);

}  // namespace palx
