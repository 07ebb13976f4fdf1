// palx.hip -- host side of libpalx (the C ABI of include/palx.h) and the
// small kernels of the DNS step: CFL / NEXTDT reductions, scalar updates,
// diagnostics, spectrum, getters, ensembles (batched launches captured in
// HIP graphs) and the autotuner.  The row / column pass kernels are in
// palx_kernels.h and are reached through the launchers of palx_launch.h.
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
#include "palx_frame.h"
#include "palx_launch.h"

using namespace palx;

// ============================================================================ kernels
namespace {

// ----------------------------------------------------------------- small kernels
// Reduce the per-row partials of the physical pass; optionally NEXTDT.
// f32: CFLM = float32(max * float32(inv_dx)) like the reference's float32 max.
__device__ void reduce_phys(const double* __restrict__ part, int M, double* __restrict__ scal, int f32,
                            int do_nextdt) {
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

__global__ void __launch_bounds__(256) k_reduce_phys(const double* __restrict__ part, int M,
                                                     double* __restrict__ scal, int f32, int do_nextdt) {
  reduce_phys(part, M, scal, f32, do_nextdt);
}

// Per-member scalar work of an ensemble step; member m = blockIdx.x.
struct EnsScal {
  double* scal;
  const double* part;
  int M;
};
__global__ void __launch_bounds__(256) k_ens_reduce_phys(const EnsScal* __restrict__ mem, int f32,
                                                         int do_nextdt) {
  const EnsScal e = mem[blockIdx.x];
  reduce_phys(e.part, e.M, e.scal, f32, do_nextdt);
}

// op 0: dtsave[m] = dt (before NEXTDT);  op 1: t += dtsave[m], it += 1;
// op 2: t += dt, it += 1 (no NEXTDT this step);  op 3: cnm1 = cn (CNAB2)
__global__ void k_ens_scalars(const EnsScal* __restrict__ mem, int n, double* __restrict__ dtsave, int op) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  double* sc = mem[m].scal;
  if (op == 0) {
    dtsave[m] = sc[PALX_S_DT];
  } else if (op == 1 || op == 2) {
    sc[PALX_S_T] = sc[PALX_S_T] + (op == 1 ? dtsave[m] : sc[PALX_S_DT]);
    sc[PALX_S_IT] = sc[PALX_S_IT] + 1.0;
  } else {
    sc[PALX_S_CNM1] = sc[PALX_S_CN];
  }
}

// cnm1 = cn at the end of every CNAB2 update (turbo_simulator.py:1771, GPU 857-871)
__global__ void k_copy_cn(double* __restrict__ scal) {
  if (threadIdx.x == 0) scal[PALX_S_CNM1] = scal[PALX_S_CN];
}

// end of a CNAB2 step without NEXTDT, in one launch: cnm1 = cn (dns_step3),
// then t += dt, it += 1 (turbo_simulator.py:1771, 2833, 2848)
__global__ void k_cnab2_tail(double* __restrict__ scal) {
  if (threadIdx.x == 0) {
    scal[PALX_S_CNM1] = scal[PALX_S_CN];
    scal[PALX_S_T] = scal[PALX_S_T] + scal[PALX_S_DT];
    scal[PALX_S_IT] = scal[PALX_S_IT] + 1.0;
  }
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

// Radial spectrum bins (turbo_simulator.py:2232-2306): bin of a mode with
// k2 = kx^2 + kz^2 > 0 is floor(sqrt(k2)/k_nyq/r_max * nbins) clipped to
// [0, nbins); energy (|om|^2/k2) * M^4, twice for kx > 0; count 1 or 2.
__device__ __forceinline__ int spectrum_bin(double k2, double k_nyq, double r_max, int nbins) {
  const double r = sqrt(k2) / k_nyq;
  int b = (int)floor((r / r_max) * (double)nbins);
  return b < 0 ? 0 : (b >= nbins ? nbins - 1 : b);
}

// One wavefront per bin, so every bin is summed in one fixed order and the
// result is the same bits on every run (no atomics).  The bin index is
// nondecreasing in |kz| at fixed kx, so the modes of bin b in column kx are
// the rows |kz| in [m0, m1): lanes take the columns kx = lane, lane + 64, ...,
// find m0, m1 from the annulus radii and correct them with the exact bin
// formula, and add the rows kz = +m and kz = -m of the column (contiguous in
// the column-major omega); a fixed shuffle tree sums the lanes.  Every bin is
// written, so the outputs need no clearing.
template <typename Tr>
__global__ void __launch_bounds__(256) k_spectrum(const cx<Tr>* __restrict__ om, int N, int nbins, double r_max,
                                                  double k_nyq, double fft_scale2, double* __restrict__ esum,
                                                  int* __restrict__ cnt) {
  const int lane = threadIdx.x & 63;
  const int b = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (b >= nbins) return;  // the whole wave
  const int H = N / 2, half = N / 2;
  const double width = r_max * k_nyq / (double)nbins;  // annulus width in wavenumber units
  // first |kz| >= 0 whose bin is >= target in column kx (half + 1: none)
  auto first_at_least = [&](double kx2, int target) -> int {
    if (target <= 0) return 0;
    if (target >= nbins) return half + 1;
    const double rr = (double)target * width;
    const double d = rr * rr - kx2;
    int m = d > 0.0 ? (int)ceil(sqrt(d)) : 0;
    if (m > half + 1) m = half + 1;
    while (m > 0 && spectrum_bin(kx2 + (double)(m - 1) * (m - 1), k_nyq, r_max, nbins) >= target) --m;
    while (m <= half && spectrum_bin(kx2 + (double)m * m, k_nyq, r_max, nbins) < target) ++m;
    return m;
  };
  double e = 0.0;
  int c = 0;
  for (int kx = lane; kx < H; kx += 64) {
    const double kx2 = (double)kx * (double)kx;
    int m0 = first_at_least(kx2, b);
    const int m1 = first_at_least(kx2, b + 1);
    if (kx == 0 && m0 == 0) m0 = 1;  // k = 0 has no bin
    const int wt = kx > 0 ? 2 : 1;
    const cx<Tr>* col = om + (size_t)kx * N;
    for (int m = m0; m < m1; ++m) {
      const double k2 = kx2 + (double)m * (double)m;
      // rows kz = +m (z = m, m < N/2) and kz = -m (z = N - m, 1 <= m <= N/2)
      for (int sgn = 0; sgn < 2; ++sgn) {
        const int z = sgn == 0 ? m : N - m;
        if ((sgn == 0 && m >= half) || (sgn == 1 && (m == 0 || m > half))) continue;
        const cx<Tr> w = col[z];
        const double o2 = (double)w.x * (double)w.x + (double)w.y * (double)w.y;
        double p = (o2 / k2) * fft_scale2;
        if (wt == 2) p *= 2.0;
        e += p;
        c += wt;
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off, 64);
    c += __shfl_xor(c, off, 64);
  }
  if (lane == 0) {
    esum[b] = e;
    cnt[b] = c;
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
  int dev = 0;      // device current at palx_create
  size_t csz = 16;  // complex size
  void* om = nullptr;
  void* fnm1 = nullptr;
  void* W = nullptr;     // velocity work planes, 2 x M*H complex, tiled by z (w_index)
  void* Wp = nullptr;    // products, 2 x M*H complex, tiled by kx (p_index)
  void* ur = nullptr;    // 3 x M x M real
  void* uc0 = nullptr;   // 2 x N x H complex
  void* gs = nullptr;    // getter scratch: one work plane, or N x H spectra in transit
  void* tw = nullptr;    // M complex twiddles exp(-2 pi i m / M)
  double* scal = nullptr;
  double* part = nullptr;  // 2*M (row partials) + diag partials
  double* dtsave = nullptr;
  // display frame: min/max partials (2 x kFrameBlocks doubles), pixel sums
  // (2 x kFrameBlocks u64), controller state (kCtlState doubles)
  double* fr = nullptr;
  // adjoint of the fixed-dt LS-IMEX-RK3 trajectory
  void* lam_om = nullptr;   // lambda_omega, column-major like om
  void* lam_fn = nullptr;   // lambda_fnm1
  void* Q = nullptr;        // adjoint z-inverse planes, layout of W
  std::vector<void*> ckpt;  // (omega, fnm1) every ckpt_every steps
  void* ckpt_T = nullptr;   // (omega, fnm1) at the end of the forward trajectory
  std::vector<void*> vslot; // velocity planes of every stage of one segment
  int adj_nsteps = 0, adj_every = 0;
  hipStream_t stream = nullptr;
  int phase = 0;  // 0: W holds velocity, 1: W holds products
  // next stage of an LS-IMEX-RK3 step run with PALX_STAGE_IN_STEP, -1 when
  // none is open (omega / fnm1 hold intermediates while one is)
  int rk3_next = -1;
  int64_t it = 0;
  int vrow = kDefaultRowVariant, vcol = kDefaultColVariant;  // kernel variants of the step passes
  // graph stepping (palx_set_graph): one captured step per (method, NEXTDT
  // step), and kGraphChunk steps without NEXTDT per method in one graph
  int use_graph = 0;
  hipStream_t cap = nullptr;
  hipGraph_t graph[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  hipGraphExec_t exec[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  std::vector<void*> allocs;
};

namespace {

int run_row(palx_t* h, int mode, void* target) {
  const RowArgs a{target, h->Q, h->Wp, h->ur, h->part};
  const int V = mode == ROW_STEP ? h->vrow : kDefaultRowVariant;
  const hipError_t e = h->prec == 4 ? launch_row_f32(h->M, V, mode, a, nullptr, 1, h->tw, h->stream)
                                    : launch_row_f64(h->M, V, mode, a, nullptr, 1, h->tw, h->stream);
  if (e != hipSuccess) return fail(PALX_EHIP, std::string("row pass launch: ") + hipGetErrorString(e));
  return PALX_OK;
}

int rk3_open_error(palx_t* h) {
  return fail(PALX_ESTATE,
              "an LS-IMEX-RK3 step run with PALX_STAGE_IN_STEP is open (omega / fnm1 hold intermediates): "
              "run its remaining stages first");
}

// omega / fnm1 hold intermediates while an in-step RK3 step is open
bool holds_rk3_intermediate(const palx_t* h, int32_t field) {
  return h->rk3_next >= 0 && (field == PALX_F_OM || field == PALX_F_FNM1);
}

int run_col(palx_t* h, int mode, int stage) {
  // while an in-step RK3 step is open only its own later stages and the
  // velocity load (which touches neither omega nor fnm1) may run
  if (h->rk3_next >= 0 && mode != COL_INIT && !(mode == COL_RK3 && (stage & RK3_PRE_IN))) return rk3_open_error(h);
  const bool adj = mode == COL_ADJ;
  void* out = (mode == COL_OMEGA || mode == COL_STREAM) ? h->gs : adj ? h->Q : h->W;
  const ColArgs a{h->Wp, adj ? h->lam_om : h->om, adj ? h->lam_fn : h->fnm1, h->uc0, out, h->scal};
  const int V = (mode == COL_CNAB2 || mode == COL_RK3) ? h->vcol : kDefaultColVariant;
  const hipError_t e = h->prec == 4 ? launch_col_f32(h->M, V, mode, stage, a, nullptr, 1, h->tw, h->stream)
                                    : launch_col_f64(h->M, V, mode, stage, a, nullptr, 1, h->tw, h->stream);
  if (e != hipSuccess) return fail(PALX_EHIP, std::string("column pass launch: ") + hipGetErrorString(e));
  return PALX_OK;
}

// Grids with a compiled padded length M = 3N/2: N = 2^k (8..4096, and 8192
// in FP32) and N = 3*2^k (48..6144), the sizes of the reference GUI's N menu
// up to 8192 (turbo_main.py:589-591) except 7776.
bool supported_N(int N, int precision) {
  switch (N) {
    case 8: case 16: case 32: case 64: case 128: case 256: case 512: case 1024: case 2048:
    case 4096: case 8192:
    case 48: case 96: case 192: case 384: case 768: case 1536: case 3072: case 6144: case 12288:
      return true;
    case 16384:
      return precision == 4;
  }
  return false;
}
#define PALX_UNSUP_MSG "N must be 2^k in [8, 8192] (16384 in FP32) or 3*2^k in [48, 12288]"

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
    case PALX_F_WORK: *ptr = h->W; *bytes = 2 * work_plane(h->M) * h->csz; return PALX_OK;
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

// One full step of a handle on h->stream (stages, then NEXTDT and the time
// advance of the run_dns loop, turbo_simulator.py:2831-2848).
int enqueue_step(palx_t* h, int method, bool ndt) {
  if (method == PALX_CNAB2 && !ndt) {  // the scalar work of the step in one launch
    int rc = run_row(h, ROW_STEP, h->W);
    if (rc) return rc;
    if ((rc = run_col(h, COL_CNAB2, 0))) return rc;
    hipLaunchKernelGGL(k_cnab2_tail, dim3(1), dim3(64), 0, h->stream, h->scal);
    HIPCHK(hipGetLastError());
    return PALX_OK;
  }
  const int nst = method == PALX_CNAB2 ? 1 : 3;
  for (int st = 0; st < nst; ++st) {
    int rc = run_row(h, ROW_STEP, h->W);
    if (rc) return rc;
    rc = update(h, method, method == PALX_CNAB2 ? st : rk3_step_stage(st));
    if (rc) return rc;
  }
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
  return PALX_OK;
}

void drop_graphs(palx_t* h) {
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 3; ++b) {
      if (h->exec[a][b]) (void)hipGraphExecDestroy(h->exec[a][b]);
      if (h->graph[a][b]) (void)hipGraphDestroy(h->graph[a][b]);
      h->exec[a][b] = nullptr;
      h->graph[a][b] = nullptr;
    }
}

// Steps without NEXTDT that one graph launch runs when at least that many
// follow: a graph launch costs more host time than a small grid's step takes
// on the GPU (N=256 CNAB2 FP64: 24 us per step replaying one-step graphs
// against 19.5 us launch by launch).
constexpr int kGraphChunk = 8;

// The executable graph of one step (chunk = false) or of kGraphChunk steps
// without NEXTDT (chunk = true), captured on the handle's private stream on
// first use (the launch stream may be the null stream, which cannot be
// captured).  Kernel nodes only.
int step_graph(palx_t* h, int method, bool ndt, bool chunk, hipGraphExec_t* out) {
  const int a = method == PALX_CNAB2 ? 0 : 1, b = chunk ? 2 : ndt ? 1 : 0;
  if (h->exec[a][b]) {
    *out = h->exec[a][b];
    return PALX_OK;
  }
  if (!h->cap) HIPCHK(hipStreamCreateWithFlags(&h->cap, hipStreamNonBlocking));
  hipStream_t launch = h->stream;
  h->stream = h->cap;
  hipError_t e = hipStreamBeginCapture(h->cap, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) {
    h->stream = launch;
    return fail(PALX_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
  }
  int rc = PALX_OK;
  for (int k = 0; k < (chunk ? kGraphChunk : 1) && !rc; ++k) rc = enqueue_step(h, method, ndt);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(h->cap, &g);
  h->stream = launch;
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(PALX_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(PALX_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e));
  }
  h->graph[a][b] = g;
  h->exec[a][b] = x;
  *out = x;
  return PALX_OK;
}

// The var-specific part of a getter, after the physical pass: plane 2 of
// PALX_F_UR for ENERGY / OMEGA / STREAM (planes 0, 1 already hold u, w).
int getter_plane(palx_t* h, int var) {
  const size_t plane = (size_t)h->M * h->M;
  int rc = PALX_OK;
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


// Display size of a frame: scale > 1 keeps every scale-th row / column
// (numpy's pix[::s, ::s]), scale < -1 repeats every pixel -scale times.
int frame_dims(const palx_t* h, int scale, int* down, int* up, int* w, int* hgt) {
  if (scale > h->M || (scale < -1 && (int64_t)h->M * -(int64_t)scale > 16384))
    return fail(PALX_EINVAL, "frame scale out of range (display side at most 16384 pixels)");
  *down = scale > 1 ? scale : 1;
  *up = scale < -1 ? -scale : 1;
  if (*up > 1) {
    *w = h->M * *up;
  } else {
    *w = (h->M + *down - 1) / *down;
  }
  *hgt = *w;
  return PALX_OK;
}

}  // namespace

// ============================================================================ C ABI
// Staging bins of palx_spectrum_noalloc (esum doubles, then int counts):
// one copy per GPU, held until the process ends, so that call allocates no
// device memory.  Host code never passes `scratch` itself to a kernel or a
// copy (in host code the name is a stand-in, not a device address): it uses
// the address hipGetSymbolAddress returns for the current device, looked up
// once per device.
static __device__ __align__(256) unsigned char
    scratch[(size_t)PALX_SPECTRUM_NOALLOC_MAX_BINS * (sizeof(double) + sizeof(int))];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// device address of `scratch` on the handle's device (which must be current)
int noalloc_scratch(const palx_t* h, void** out) {
  int cur = -1;
  HIPCHK(hipGetDevice(&cur));
  if (cur != h->dev) return fail(PALX_ESTATE, "the handle's device must be the current device");
  if (cur < 0 || cur >= kMaxDevices) return fail(PALX_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[cur]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(PALX_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[cur] = p;
  }
  *out = g_scratch_dev[cur];
  return PALX_OK;
}
}  // namespace

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
  if (!supported_N(N, precision)) return fail(PALX_EUNSUP, PALX_UNSUP_MSG);
  palx_t* h = new palx_t();
  h->N = N;
  h->M = 3 * N / 2;
  h->H = N / 2;
  h->prec = precision;
  if (hipGetDevice(&h->dev) != hipSuccess) {
    delete h;
    return fail(PALX_EHIP, "hipGetDevice");
  }
  default_step_variants(h->M, n_pass_variants(h->M), precision, &h->vrow, &h->vcol);
  h->vrow = fitting_variant(h->M, h->vrow, 2 * precision);
  h->vcol = fitting_variant(h->M, h->vcol, 2 * precision);
  h->csz = 2 * (size_t)precision;
  const size_t NH = (size_t)N * h->H;
  const size_t WH = work_plane(h->M);  // one work plane
  int rc = PALX_OK;
  if (!rc) rc = dev_alloc(h, &h->om, NH * h->csz);
  if (!rc) rc = dev_alloc(h, &h->fnm1, NH * h->csz);
  if (!rc) rc = dev_alloc(h, &h->W, 2 * WH * h->csz);
  if (!rc) rc = dev_alloc(h, &h->Wp, 2 * WH * h->csz);
  if (!rc) rc = dev_alloc(h, &h->ur, 3 * (size_t)h->M * h->M * precision);
  if (!rc) rc = dev_alloc(h, &h->uc0, 2 * NH * h->csz);
  if (!rc) rc = dev_alloc(h, &h->gs, std::max(WH, NH) * h->csz);
  if (!rc) rc = dev_alloc(h, &h->lam_om, NH * h->csz);
  if (!rc) rc = dev_alloc(h, &h->lam_fn, NH * h->csz);
  if (!rc) rc = dev_alloc(h, &h->Q, 2 * WH * h->csz);
  if (!rc) rc = dev_alloc(h, &h->tw, (size_t)h->M * h->csz);
  if (!rc) rc = dev_alloc(h, (void**)&h->scal, PALX_NSCAL * sizeof(double));
  if (!rc) rc = dev_alloc(h, (void**)&h->part, (2 * (size_t)h->M + 5 * 1024 + 16) * sizeof(double));
  if (!rc) rc = dev_alloc(h, (void**)&h->dtsave, 16 * sizeof(double));
  if (!rc) rc = dev_alloc(h, (void**)&h->fr, (4 * (size_t)kFrameBlocks + kCtlState) * sizeof(double));
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
  // work still queued on the handle (or its capture stream) uses the buffers
  (void)hipStreamSynchronize(h->stream);
  if (h->cap) (void)hipStreamSynchronize(h->cap);
  drop_graphs(h);
  if (h->cap) (void)hipStreamDestroy(h->cap);
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
  h->rk3_next = -1;
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
  if (holds_rk3_intermediate(h, field)) return rk3_open_error(h);
  return field_info(h, field, ptr, bytes);
}

int palx_download(palx_t* h, int32_t field, void* dst, size_t bytes) {
  if (!h || !dst) return fail(PALX_EINVAL, "null argument");
  if (holds_rk3_intermediate(h, field)) return rk3_open_error(h);
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
  if (holds_rk3_intermediate(h, field)) return rk3_open_error(h);
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
  const bool in_step = (stage & PALX_STAGE_IN_STEP) != 0;
  const int st = stage & ~PALX_STAGE_IN_STEP;
  if (st < 0 || st > 2) return fail(PALX_EINVAL, "stage must be 0..2");
  if (in_step && method != PALX_LS_IMEX_RK3) return fail(PALX_EINVAL, "PALX_STAGE_IN_STEP is for LS-IMEX-RK3 stages");
  if (in_step && st != (h->rk3_next < 0 ? 0 : h->rk3_next))
    return fail(PALX_ESTATE, "the stages of a step run with PALX_STAGE_IN_STEP go 0, 1, 2 in order");
  int rc = update(h, method, in_step ? rk3_step_stage(st) : st);  // run_col refuses a plain stage inside one
  if (rc) return rc;
  if (in_step) h->rk3_next = st < 2 ? st + 1 : -1;
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
  if (h->rk3_next >= 0) return rk3_open_error(h);  // a replayed graph does not pass run_col
  auto is_ndt = [&](int64_t it) { return nextdt_every > 0 && ((it % nextdt_every) == 0 || it == 1); };
  for (int s = 0; s < nsteps;) {
    int rc;
    if (h->use_graph) {
      // kGraphChunk steps in one launch when none of them runs NEXTDT
      bool chunk = nsteps - s >= kGraphChunk;
      for (int k = 1; chunk && k <= kGraphChunk; ++k) chunk = !is_ndt(h->it + k);
      const bool ndt = !chunk && is_ndt(h->it + 1);
      hipGraphExec_t x = nullptr;
      if ((rc = step_graph(h, method, ndt, chunk, &x))) return rc;
      HIPCHK(hipGraphLaunch(x, h->stream));
      const int n = chunk ? kGraphChunk : 1;
      h->it += n;
      s += n;
    } else {
      h->it += 1;
      if ((rc = enqueue_step(h, method, is_ndt(h->it)))) return rc;
      s += 1;
    }
  }
  return PALX_OK;
}

int palx_set_graph(palx_t* h, int32_t enable) {
  if (!h) return fail(PALX_EINVAL, "null handle");
  h->use_graph = enable ? 1 : 0;
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
  if (h->rk3_next >= 0) return rk3_open_error(h);
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

int palx_spectrum_into(palx_t* h, double* esum_dev, int32_t* count_dev, int32_t nbins) {
  if (!h || !esum_dev || !count_dev || nbins <= 0) return fail(PALX_EINVAL, "bad argument");
  if (h->rk3_next >= 0) return rk3_open_error(h);
  // every bin is written by its own wavefront: nothing to clear
  const double r_max = std::sqrt(2.0), k_nyq = 0.5 * (double)h->N;
  const double fs = (double)h->M * (double)h->M;
  const int blocks = (nbins + 3) / 4;
  if (h->prec == 4)
    hipLaunchKernelGGL(k_spectrum<float>, dim3(blocks), dim3(256), 0, h->stream, (const cx<float>*)h->om,
                       h->N, nbins, r_max, k_nyq, fs * fs, esum_dev, (int*)count_dev);
  else
    hipLaunchKernelGGL(k_spectrum<double>, dim3(blocks), dim3(256), 0, h->stream, (const cx<double>*)h->om,
                       h->N, nbins, r_max, k_nyq, fs * fs, esum_dev, (int*)count_dev);
  HIPCHK(hipGetLastError());
  return PALX_OK;
}

int palx_spectrum(palx_t* h, double* esum, int32_t* count, int32_t nbins) {
  if (!h || !esum || !count || nbins <= 0) return fail(PALX_EINVAL, "bad argument");
  if (h->rk3_next >= 0) return rk3_open_error(h);
  double* de = nullptr;
  int* dc = nullptr;
  HIPCHK(hipMalloc((void**)&de, (size_t)nbins * sizeof(double)));
  if (hipMalloc((void**)&dc, (size_t)nbins * sizeof(int)) != hipSuccess) {
    (void)hipFree(de);
    return fail(PALX_ENOMEM, "hipMalloc spectrum");
  }
  int rc = palx_spectrum_into(h, de, dc, nbins);
  if (!rc && (hipStreamSynchronize(h->stream) != hipSuccess ||
              hipMemcpy(esum, de, (size_t)nbins * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(count, dc, (size_t)nbins * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess))
    rc = fail(PALX_EHIP, "spectrum download");
  (void)hipFree(de);
  (void)hipFree(dc);
  return rc;
}

int palx_spectrum_noalloc(palx_t* h, double* esum, int32_t* count, int32_t nbins) {
  if (!h || !esum || !count || nbins <= 0) return fail(PALX_EINVAL, "bad argument");
  if (h->rk3_next >= 0) return rk3_open_error(h);
  if (nbins > PALX_SPECTRUM_NOALLOC_MAX_BINS)
    return fail(PALX_EINVAL, "nbins exceeds PALX_SPECTRUM_NOALLOC_MAX_BINS (" +
                                 std::to_string(PALX_SPECTRUM_NOALLOC_MAX_BINS) + ")");
  void* s = nullptr;
  int rc = noalloc_scratch(h, &s);
  if (rc) return rc;
  double* de = static_cast<double*>(s);
  int* dc = reinterpret_cast<int*>(de + PALX_SPECTRUM_NOALLOC_MAX_BINS);
  if ((rc = palx_spectrum_into(h, de, dc, nbins))) return rc;  // writes every bin: nothing to clear
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(esum, de, (size_t)nbins * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(count, dc, (size_t)nbins * sizeof(int), hipMemcpyDeviceToHost));
  return PALX_OK;
}

int palx_recon_velocity(palx_t* h) {
  if (!h) return fail(PALX_EINVAL, "null handle");
  h->phase = 0;
  return run_col(h, COL_RECON, 0);
}

int palx_adj_forward(palx_t* h, int32_t nsteps, int32_t ckpt_every) {
  if (!h || nsteps <= 0 || ckpt_every <= 0) return fail(PALX_EINVAL, "bad adjoint forward arguments");
  const size_t NH = (size_t)h->N * h->H;
  const int nck = (nsteps + ckpt_every - 1) / ckpt_every;
  while ((int)h->ckpt.size() < nck) {
    void* p = nullptr;
    int rc = dev_alloc(h, &p, 2 * NH * h->csz);
    if (rc) return rc;
    h->ckpt.push_back(p);
  }
  while ((int)h->vslot.size() < 3 * ckpt_every) {
    void* p = nullptr;
    int rc = dev_alloc(h, &p, 2 * work_plane(h->M) * h->csz);
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
      rc = update(h, PALX_LS_IMEX_RK3, rk3_step_stage(st));
      if (rc) return rc;
    }
  }
  // the end state, restored by palx_adj_backward (whose replays overwrite omega)
  if (!h->ckpt_T && (rc = dev_alloc(h, &h->ckpt_T, 2 * NH * h->csz))) return rc;
  HIPCHK(hipMemcpyAsync(h->ckpt_T, h->om, NH * h->csz, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(hipMemcpyAsync((char*)h->ckpt_T + NH * h->csz, h->fnm1, NH * h->csz, hipMemcpyDeviceToDevice,
                        h->stream));
  return PALX_OK;
}

int palx_adj_backward(palx_t* h, const void* lam_T) {
  if (!h || !lam_T) return fail(PALX_EINVAL, "null argument");
  if (h->adj_nsteps <= 0) return fail(PALX_ESTATE, "palx_adj_forward must run first");
  const size_t NH = (size_t)h->N * h->H;
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
        rc = update(h, PALX_LS_IMEX_RK3, rk3_step_stage(st));
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
  // back to the end of the forward trajectory: omega(T), fnm1(T); the work
  // planes W were not written by the replays and still hold the velocity of omega(T)
  HIPCHK(hipMemcpyAsync(h->om, h->ckpt_T, NH * h->csz, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->fnm1, (char*)h->ckpt_T + NH * h->csz, NH * h->csz, hipMemcpyDeviceToDevice,
                        h->stream));
  h->phase = 0;
  return PALX_OK;
}

int palx_getter(palx_t* h, int32_t var) {
  if (!h) return fail(PALX_EINVAL, "null handle");
  if (h->phase != 0) return fail(PALX_ESTATE, "work planes hold products");
  if (var < PALX_VAR_U || var > PALX_VAR_STREAM) return fail(PALX_EINVAL, "unknown getter variable");
  int rc = physical(h, 0);
  if (rc) return rc;
  return getter_plane(h, var);
}


int palx_frame_dims(const palx_t* h, int32_t scale, int32_t* width, int32_t* height, size_t* bytes) {
  if (!h) return fail(PALX_EINVAL, "null handle");
  int down, up, w, hg;
  const int rc = frame_dims(h, scale, &down, &up, &w, &hg);
  if (rc) return rc;
  if (width) *width = w;
  if (height) *height = hg;
  if (bytes) *bytes = PALX_FRAME_NREC * sizeof(double) + (size_t)w * hg;
  return PALX_OK;
}

int palx_frame_reset(palx_t* h, double re) {
  if (!h) return fail(PALX_EINVAL, "null handle");
  hipLaunchKernelGGL(k_frame_ctl_reset, dim3(1), dim3(64), 0, h->stream, h->fr + 4 * (size_t)kFrameBlocks, re);
  HIPCHK(hipGetLastError());
  return PALX_OK;
}

int palx_frame(palx_t* h, int32_t var, int32_t scale, int32_t flags, const palx_frame_ctl* ctl, void* out,
               size_t out_bytes) {
  if (!h || !out) return fail(PALX_EINVAL, "null argument");
  if (var < PALX_VAR_U || var > PALX_VAR_STREAM) return fail(PALX_EINVAL, "unknown frame variable");
  if (flags & ~(PALX_FRAME_NEXTDT | PALX_FRAME_STRETCH | PALX_FRAME_ADAPT))
    return fail(PALX_EINVAL, "unknown frame flags");
  if ((flags & PALX_FRAME_ADAPT) && !ctl) return fail(PALX_EINVAL, "the viscosity controller needs ctl");
  if (((uintptr_t)out) % 8 != 0) return fail(PALX_EINVAL, "frame buffer must be 8-byte aligned");
  if (h->phase != 0) return fail(PALX_ESTATE, "work planes hold products");
  if (h->rk3_next >= 0) return rk3_open_error(h);  // the frame's records read omega
  int down, up, w, hg;
  int rc = frame_dims(h, scale, &down, &up, &w, &hg);
  if (rc) return rc;
  if (out_bytes < PALX_FRAME_NREC * sizeof(double) + (size_t)w * hg)
    return fail(PALX_EINVAL, "frame buffer smaller than palx_frame_dims");
  palx_frame_ctl c{2.5, 20.0, 0.25, 0.0, 25.0, 0.001, 0.01, 5.0, 0.0, 1000.0};
  if (ctl) c = *ctl;
  // 1-2: u, w (+ NEXTDT), then the getter plane
  if ((rc = physical(h, (flags & PALX_FRAME_NEXTDT) ? 1 : 0))) return rc;
  if ((rc = getter_plane(h, var))) return rc;
  const size_t plane = (size_t)h->M * h->M;
  const int pidx = var == PALX_VAR_U ? 0 : var == PALX_VAR_V ? 1 : 2;
  double* rec = (double*)out;
  uint8_t* pix = (uint8_t*)out + PALX_FRAME_NREC * sizeof(double);
  double* mm = h->fr;
  unsigned long long* ps = (unsigned long long*)(h->fr + 2 * (size_t)kFrameBlocks);
  double* cst = h->fr + 4 * (size_t)kFrameBlocks;
  double* dpart = h->part + 2 * (size_t)h->M;
  const int nb = kFrameBlocks;
  const size_t nout = (size_t)w * hg;
  const int ob = (int)std::min<size_t>(4096, (nout + 255) / 256);
  const int stretch = (flags & PALX_FRAME_STRETCH) ? 1 : 0;
  const int adapt = (flags & PALX_FRAME_ADAPT) ? 1 : 0;
  // 3-7: range, pixel sums, spectral sums, record + controller, display
  if (h->prec == 4) {
    const float* f = (const float*)h->ur + pidx * plane;
    hipLaunchKernelGGL(k_frame_minmax<float>, dim3(nb), dim3(256), 0, h->stream, f, plane, mm);
    hipLaunchKernelGGL(k_frame_range, dim3(1), dim3(256), 0, h->stream, (const double*)mm, nb, rec);
    hipLaunchKernelGGL(k_frame_pixsum<float>, dim3(nb), dim3(256), 0, h->stream, f, plane, (const double*)rec, ps);
    hipLaunchKernelGGL(k_diag<float>, dim3(nb), dim3(256), 0, h->stream, (const cx<float>*)h->om, h->N, dpart);
  } else {
    const double* f = (const double*)h->ur + pidx * plane;
    hipLaunchKernelGGL(k_frame_minmax<double>, dim3(nb), dim3(256), 0, h->stream, f, plane, mm);
    hipLaunchKernelGGL(k_frame_range, dim3(1), dim3(256), 0, h->stream, (const double*)mm, nb, rec);
    hipLaunchKernelGGL(k_frame_pixsum<double>, dim3(nb), dim3(256), 0, h->stream, f, plane, (const double*)rec, ps);
    hipLaunchKernelGGL(k_diag<double>, dim3(nb), dim3(256), 0, h->stream, (const cx<double>*)h->om, h->N, dpart);
  }
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_frame_record, dim3(1), dim3(256), 0, h->stream, (const unsigned long long*)ps,
                     (const double*)dpart, nb, h->scal, cst, c, h->N, h->M, w, hg, adapt, rec);
  HIPCHK(hipGetLastError());
  if (h->prec == 4)
    hipLaunchKernelGGL(k_frame_display<float>, dim3(ob), dim3(256), 0, h->stream,
                       (const float*)h->ur + pidx * plane, h->M, down, up, w, hg, (const double*)rec, stretch,
                       c.k_std, pix);
  else
    hipLaunchKernelGGL(k_frame_display<double>, dim3(ob), dim3(256), 0, h->stream,
                       (const double*)h->ur + pidx * plane, h->M, down, up, w, hg, (const double*)rec, stretch,
                       c.k_std, pix);
  HIPCHK(hipGetLastError());
  return PALX_OK;
}
}  // extern "C"

// ============================================================================ ensembles
// Members of equal padded length M form a class; each class has device
// tables of its members' buffers so one launch runs a pass for all of them
// (grid row = member).  A step is enqueued as: save dt, fork one stream per
// class, the class's stage passes, join, the scalar work of all members.
// palx_ens_step captures that sequence once per (method, NEXTDT-step) into a
// HIP graph and launches one graph per step.
struct EnsClass {
  int N = 0, M = 0;
  std::vector<int> idx;
  RowArgs* drow = nullptr;
  ColArgs* dcol = nullptr;
  const void* tw = nullptr;
  int vrow = kDefaultRowVariant, vcol = kDefaultColVariant;
  hipStream_t side = nullptr;
  hipEvent_t done = nullptr;
};

struct palx_ens_t {
  int prec = 4;
  std::vector<palx_t*> mem;
  std::vector<EnsClass> cls;
  EnsScal* dscal = nullptr;
  double* dtsave = nullptr;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t fork = nullptr;
  int64_t it = 0;
  hipGraph_t graph[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  hipGraphExec_t exec[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  std::vector<void*> allocs;
};

namespace {

void ens_drop_graphs(palx_ens_t* e) {
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 3; ++b) {
      if (e->exec[a][b]) (void)hipGraphExecDestroy(e->exec[a][b]);
      if (e->graph[a][b]) (void)hipGraphDestroy(e->graph[a][b]);
      e->exec[a][b] = nullptr;
      e->graph[a][b] = nullptr;
    }
}

int ens_row(palx_ens_t* e, EnsClass& c, int mode, int V, hipStream_t s) {
  const RowArgs none{};
  const int n = (int)c.idx.size();
  const hipError_t r = e->prec == 4 ? launch_row_f32(c.M, V, mode, none, c.drow, n, c.tw, s)
                                    : launch_row_f64(c.M, V, mode, none, c.drow, n, c.tw, s);
  if (r != hipSuccess) return fail(PALX_EHIP, std::string("batched row pass: ") + hipGetErrorString(r));
  return PALX_OK;
}

int ens_col(palx_ens_t* e, EnsClass& c, int mode, int V, int stage, hipStream_t s) {
  const ColArgs none{};
  const int n = (int)c.idx.size();
  const hipError_t r = e->prec == 4 ? launch_col_f32(c.M, V, mode, stage, none, c.dcol, n, c.tw, s)
                                    : launch_col_f64(c.M, V, mode, stage, none, c.dcol, n, c.tw, s);
  if (r != hipSuccess) return fail(PALX_EHIP, std::string("batched column pass: ") + hipGetErrorString(r));
  return PALX_OK;
}

int ens_scalars(palx_ens_t* e, int op) {
  const int n = (int)e->mem.size();
  hipLaunchKernelGGL(k_ens_scalars, dim3((n + 63) / 64), dim3(64), 0, e->stream, (const EnsScal*)e->dscal, n,
                     e->dtsave, op);
  HIPCHK(hipGetLastError());
  return PALX_OK;
}

// One ensemble step (the same order as palx_step for every member).
int ens_enqueue_step(palx_ens_t* e, int method, bool cadence, bool serial) {
  int rc = PALX_OK;
  if (cadence && (rc = ens_scalars(e, 0))) return rc;
  if (!serial) HIPCHK(hipEventRecord(e->fork, e->stream));
  const int nst = method == PALX_CNAB2 ? 1 : 3;
  const int cmode = method == PALX_CNAB2 ? COL_CNAB2 : COL_RK3;
  for (auto& c : e->cls) {
    hipStream_t s = serial ? e->stream : c.side;
    if (!serial) HIPCHK(hipStreamWaitEvent(s, e->fork, 0));
    for (int st = 0; st < nst; ++st) {
      if ((rc = ens_row(e, c, ROW_STEP, c.vrow, s))) return rc;
      if ((rc = ens_col(e, c, cmode, c.vcol, method == PALX_CNAB2 ? st : rk3_step_stage(st), s))) return rc;
    }
    if (cadence && (rc = ens_row(e, c, ROW_PHYS, kDefaultRowVariant, s))) return rc;
    if (!serial) {
      HIPCHK(hipEventRecord(c.done, s));
      HIPCHK(hipStreamWaitEvent(e->stream, c.done, 0));
    }
  }
  if (method == PALX_CNAB2 && (rc = ens_scalars(e, 3))) return rc;
  if (cadence) {
    hipLaunchKernelGGL(k_ens_reduce_phys, dim3((unsigned)e->mem.size()), dim3(256), 0, e->stream,
                       (const EnsScal*)e->dscal, e->prec == 4 ? 1 : 0, 1);
    HIPCHK(hipGetLastError());
    return ens_scalars(e, 1);
  }
  return ens_scalars(e, 2);
}

// The executable graph of one ensemble step (chunk = false) or of
// kGraphChunk steps without the NEXTDT cadence (chunk = true).
int ens_graph(palx_ens_t* e, int method, bool cadence, bool chunk, hipGraphExec_t* out) {
  const int a = method == PALX_CNAB2 ? 0 : 1, b = chunk ? 2 : cadence ? 1 : 0;
  if (e->exec[a][b]) {
    *out = e->exec[a][b];
    return PALX_OK;
  }
  HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed));
  int rc = PALX_OK;
  for (int k = 0; k < (chunk ? kGraphChunk : 1) && !rc; ++k) rc = ens_enqueue_step(e, method, cadence, false);
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(e->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ce != hipSuccess) return fail(PALX_EHIP, std::string("graph capture: ") + hipGetErrorString(ce));
  hipGraphExec_t x = nullptr;
  const hipError_t ie = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (ie != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(PALX_EHIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
  }
  e->graph[a][b] = g;
  e->exec[a][b] = x;
  *out = x;
  return PALX_OK;
}

}  // namespace

extern "C" {

int palx_n_variants(int32_t N) {
  if (!supported_N(N, 4)) return fail(PALX_EUNSUP, PALX_UNSUP_MSG);
  return n_pass_variants(3 * N / 2);
}

int palx_variant_ok(int32_t N, int32_t precision, int32_t V) {
  if ((precision != 4 && precision != 8) || !supported_N(N, precision)) return fail(PALX_EUNSUP, PALX_UNSUP_MSG);
  return variant_fits(3 * N / 2, V, 2 * precision) ? 1 : 0;
}

int palx_set_variants(palx_t* h, int32_t vrow, int32_t vcol) {
  if (!h) return fail(PALX_EINVAL, "null handle");
  const int csz = 2 * h->prec;
  if (!variant_fits(h->M, vrow, csz) || !variant_fits(h->M, vcol, csz))
    return fail(PALX_EINVAL, "variant out of range (or not compiled for this length and precision)");
  h->vrow = vrow;
  h->vcol = vcol;
  drop_graphs(h);  // captured launches hold the old kernels
  return PALX_OK;
}

int palx_get_variants(const palx_t* h, int32_t* vrow, int32_t* vcol) {
  if (!h || !vrow || !vcol) return fail(PALX_EINVAL, "null argument");
  *vrow = h->vrow;
  *vcol = h->vcol;
  return PALX_OK;
}

int palx_ens_destroy(palx_ens_t* e) {
  if (!e) return PALX_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  ens_drop_graphs(e);
  for (auto& c : e->cls) {
    if (c.side) (void)hipStreamDestroy(c.side);
    if (c.done) (void)hipEventDestroy(c.done);
  }
  if (e->fork) (void)hipEventDestroy(e->fork);
  for (palx_t* h : e->mem) palx_destroy(h);
  for (void* p : e->allocs) (void)hipFree(p);
  if (e->own) (void)hipStreamDestroy(e->own);
  delete e;
  return PALX_OK;
}

int palx_ens_create(int32_t n, const int32_t* Ns, int32_t precision, palx_ens_t** out) {
  if (!out || !Ns || n <= 0) return fail(PALX_EINVAL, "bad ensemble arguments");
  *out = nullptr;
  if (precision != 4 && precision != 8) return fail(PALX_EINVAL, "precision must be 4 or 8");
  for (int m = 0; m < n; ++m)
    if (!supported_N(Ns[m], precision)) return fail(PALX_EUNSUP, "member " PALX_UNSUP_MSG);
  palx_ens_t* e = new palx_ens_t();
  e->prec = precision;
  int rc = PALX_OK;
  auto bail = [&](int code) {
    palx_ens_destroy(e);
    return code;
  };
  if (hipStreamCreateWithFlags(&e->own, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(PALX_EHIP, "hipStreamCreate"));
  e->stream = e->own;
  if (hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess)
    return bail(fail(PALX_EHIP, "hipEventCreate"));
  for (int m = 0; m < n; ++m) {
    palx_t* h = nullptr;
    if ((rc = palx_create(Ns[m], precision, &h))) return bail(rc);
    h->stream = e->stream;
    e->mem.push_back(h);
  }
  // classes in order of first appearance
  for (int m = 0; m < n; ++m) {
    auto it = std::find_if(e->cls.begin(), e->cls.end(), [&](const EnsClass& c) { return c.N == Ns[m]; });
    if (it == e->cls.end()) {
      EnsClass c;
      c.N = Ns[m];
      c.M = 3 * Ns[m] / 2;
      c.tw = e->mem[m]->tw;
      c.vrow = e->mem[m]->vrow;  // the members' default step variants
      c.vcol = e->mem[m]->vcol;
      e->cls.push_back(c);
      it = e->cls.end() - 1;
    }
    it->idx.push_back(m);
  }
  for (auto& c : e->cls) {
    std::vector<RowArgs> rows;
    std::vector<ColArgs> cols;
    for (int m : c.idx) {
      palx_t* h = e->mem[m];
      rows.push_back(RowArgs{h->W, h->Q, h->Wp, h->ur, h->part});
      cols.push_back(ColArgs{h->Wp, h->om, h->fnm1, h->uc0, h->W, h->scal});
    }
    void* p = nullptr;
    if (hipMalloc(&p, rows.size() * sizeof(RowArgs)) != hipSuccess) return bail(fail(PALX_ENOMEM, "hipMalloc"));
    e->allocs.push_back(p);
    c.drow = (RowArgs*)p;
    if (hipMalloc(&p, cols.size() * sizeof(ColArgs)) != hipSuccess) return bail(fail(PALX_ENOMEM, "hipMalloc"));
    e->allocs.push_back(p);
    c.dcol = (ColArgs*)p;
    if (hipMemcpy(c.drow, rows.data(), rows.size() * sizeof(RowArgs), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c.dcol, cols.data(), cols.size() * sizeof(ColArgs), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(PALX_EHIP, "table upload"));
    if (hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c.done, hipEventDisableTiming) != hipSuccess)
      return bail(fail(PALX_EHIP, "class stream"));
  }
  {
    std::vector<EnsScal> sc;
    for (palx_t* h : e->mem) sc.push_back(EnsScal{h->scal, h->part, h->M});
    void* p = nullptr;
    if (hipMalloc(&p, sc.size() * sizeof(EnsScal)) != hipSuccess) return bail(fail(PALX_ENOMEM, "hipMalloc"));
    e->allocs.push_back(p);
    e->dscal = (EnsScal*)p;
    if (hipMemcpy(e->dscal, sc.data(), sc.size() * sizeof(EnsScal), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(PALX_EHIP, "table upload"));
    if (hipMalloc(&p, (size_t)n * sizeof(double)) != hipSuccess) return bail(fail(PALX_ENOMEM, "hipMalloc"));
    e->allocs.push_back(p);
    e->dtsave = (double*)p;
  }
  *out = e;
  return PALX_OK;
}

int palx_ens_set_stream(palx_ens_t* e, void* s) {
  if (!e) return fail(PALX_EINVAL, "null ensemble");
  if (!s) return fail(PALX_EINVAL, "ensembles need a created stream (the null stream cannot be captured)");
  HIPCHK(hipStreamSynchronize(e->stream));
  ens_drop_graphs(e);
  e->stream = (hipStream_t)s;
  for (palx_t* h : e->mem) h->stream = e->stream;
  return PALX_OK;
}

int palx_ens_member(palx_ens_t* e, int32_t m, palx_t** out) {
  if (!e || !out || m < 0 || m >= (int32_t)e->mem.size()) return fail(PALX_EINVAL, "bad member index");
  *out = e->mem[m];
  return PALX_OK;
}

int palx_ens_classes(palx_ens_t* e, int32_t* nclass, int32_t* N, int32_t* count, int32_t* vrow, int32_t* vcol) {
  if (!e || !nclass) return fail(PALX_EINVAL, "null argument");
  *nclass = (int32_t)e->cls.size();
  for (size_t c = 0; c < e->cls.size(); ++c) {
    if (N) N[c] = e->cls[c].N;
    if (count) count[c] = (int32_t)e->cls[c].idx.size();
    if (vrow) vrow[c] = e->cls[c].vrow;
    if (vcol) vcol[c] = e->cls[c].vcol;
  }
  return PALX_OK;
}

int palx_ens_set_variants(palx_ens_t* e, int32_t N, int32_t vrow, int32_t vcol) {
  if (!e) return fail(PALX_EINVAL, "null ensemble");
  for (auto& c : e->cls) {
    if (c.N != N) continue;
    if (!variant_fits(c.M, vrow, 2 * e->prec) || !variant_fits(c.M, vcol, 2 * e->prec))
      return fail(PALX_EINVAL, "variant out of range (or not compiled for this length and precision)");
    c.vrow = vrow;
    c.vcol = vcol;
    for (int m : c.idx) {
      e->mem[m]->vrow = vrow;
      e->mem[m]->vcol = vcol;
      drop_graphs(e->mem[m]);
    }
    ens_drop_graphs(e);
    return PALX_OK;
  }
  return fail(PALX_EINVAL, "no member of that size");
}

int palx_ens_tune(palx_ens_t* e, int32_t reps, float* ms) {
  if (!e || reps <= 0 || reps > 1000) return fail(PALX_EINVAL, "bad tune arguments");
  for (palx_t* h : e->mem)
    if (h->rk3_next >= 0) return rk3_open_error(h);
  HIPCHK(hipStreamSynchronize(e->stream));
  // Each pass is timed where a step runs it: between launches of the other
  // pass (at its current variant), RK3 stages cycling, events around the
  // timed launches only.  A pass repeated on the same buffers reads inputs
  // another launch of itself has just read, which favours some variants: at
  // N=4096 FP32 the split-exchange row pass tied the best one repeated and
  // took 10% longer inside the step.  Rows are tuned first, then columns
  // with the chosen row variant.
  std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
  int rc = PALX_OK;
  for (auto& x : ev)
    if (!rc && hipEventCreate(&x) != hipSuccess) rc = fail(PALX_EHIP, "hipEventCreate");
  for (size_t ci = 0; ci < e->cls.size() && !rc; ++ci) {
    EnsClass& c = e->cls[ci];
    float best[2] = {1e30f, 1e30f};
    int bv[2] = {c.vrow, c.vcol};
    for (int pass = 0; pass < 2 && !rc; ++pass) {
      for (int v = 0; v < kVariants && !rc; ++v) {
        float t = -1.0f;
        if (variant_fits(c.M, v, 2 * e->prec)) {
          const int vr = pass == 0 ? v : bv[0], vc = pass == 0 ? bv[1] : v;
          // warm-up (code object load, LDS attribute) of both passes
          if ((rc = ens_row(e, c, ROW_STEP, vr, e->stream)) ||
              (rc = ens_col(e, c, COL_RK3, vc, rk3_step_stage(0), e->stream)))
            break;
          for (int r = 0; r < reps && !rc; ++r) {
            if (pass == 1 && (rc = ens_row(e, c, ROW_STEP, vr, e->stream))) break;
            if (hipEventRecord(ev[2 * r], e->stream) != hipSuccess) { rc = fail(PALX_EHIP, "event"); break; }
            rc = pass == 0 ? ens_row(e, c, ROW_STEP, vr, e->stream)
                           : ens_col(e, c, COL_RK3, vc, rk3_step_stage(r % 3), e->stream);
            if (rc) break;
            if (hipEventRecord(ev[2 * r + 1], e->stream) != hipSuccess) { rc = fail(PALX_EHIP, "event"); break; }
            if (pass == 0 && (rc = ens_col(e, c, COL_RK3, vc, rk3_step_stage(r % 3), e->stream))) break;
          }
          if (rc) break;
          if (hipEventSynchronize(ev[2 * reps - 1]) != hipSuccess) { rc = fail(PALX_EHIP, "event timing"); break; }
          float sum = 0.0f;
          for (int r = 0; r < reps && !rc; ++r) {
            float dt = 0.0f;
            if (hipEventElapsedTime(&dt, ev[2 * r], ev[2 * r + 1]) != hipSuccess) rc = fail(PALX_EHIP, "event timing");
            sum += dt;
          }
          if (rc) break;
          t = sum / (float)reps;
          if (t < best[pass]) {
            best[pass] = t;
            bv[pass] = v;
          }
        }
        if (ms) ms[(ci * 2 + pass) * kVariants + v] = t;
      }
    }
    if (!rc) rc = palx_ens_set_variants(e, c.N, bv[0], bv[1]);
  }
  for (auto& x : ev)
    if (x) (void)hipEventDestroy(x);
  for (palx_t* h : e->mem) h->phase = 0;
  return rc;
}

int palx_ens_step(palx_ens_t* e, int32_t nsteps, int32_t method, int32_t nextdt_every, int32_t flags) {
  if (!e || nsteps < 0) return fail(PALX_EINVAL, "bad step arguments");
  if (method != PALX_CNAB2 && method != PALX_LS_IMEX_RK3) return fail(PALX_EINVAL, "unknown method");
  for (palx_t* h : e->mem) {
    if (h->phase != 0) return fail(PALX_ESTATE, "a member's work planes hold products");
    if (h->rk3_next >= 0) return rk3_open_error(h);
  }
  const bool use_graph = (flags & PALX_ENS_GRAPH) != 0, serial = (flags & PALX_ENS_SERIAL) != 0;
  auto is_cad = [&](int64_t it) { return nextdt_every > 0 && ((it % nextdt_every) == 0 || it == 1); };
  for (int s = 0; s < nsteps;) {
    int rc;
    if (use_graph) {
      // kGraphChunk steps in one launch when none of them runs the cadence
      bool chunk = nsteps - s >= kGraphChunk;
      for (int k = 1; chunk && k <= kGraphChunk; ++k) chunk = !is_cad(e->it + k);
      const bool cad = !chunk && is_cad(e->it + 1);
      hipGraphExec_t x = nullptr;
      if ((rc = ens_graph(e, method, cad, chunk, &x))) return rc;
      HIPCHK(hipGraphLaunch(x, e->stream));
      const int n = chunk ? kGraphChunk : 1;
      e->it += n;
      s += n;
    } else {
      e->it += 1;
      if ((rc = ens_enqueue_step(e, method, is_cad(e->it), serial))) return rc;
      s += 1;
    }
  }
  for (palx_t* h : e->mem) h->it = e->it;
  return PALX_OK;
}

int palx_ens_set_iteration(palx_ens_t* e, int64_t it) {
  if (!e) return fail(PALX_EINVAL, "null ensemble");
  e->it = it;
  for (palx_t* h : e->mem) h->it = it;
  return PALX_OK;
}

int palx_ens_sync(palx_ens_t* e) {
  if (!e) return fail(PALX_EINVAL, "null ensemble");
  HIPCHK(hipStreamSynchronize(e->stream));
  return PALX_OK;
}

}  // extern "C"
