// palx_frame.h -- kernels of the display frame (palx_frame in include/palx.h).
//
// The reference builds each rendered frame with several CuPy reductions and
// then finishes it on the host: mean / std of the full-resolution pixels,
// the palinstrophy ratio, the viscosity controller and the display stretch
// (turbo_wrapper.py:317-337, 480-515; turbo_main.py:2015-2214).  Here one
// stream-ordered chain of small kernels does all of it on the device, so a
// frame costs one device-to-host copy of the display-sized buffer:
//   k_frame_minmax -> k_frame_range -> k_frame_pixsum -> (k_diag) ->
//   k_frame_record -> k_frame_display.
// Every kernel reads the plane with 16-byte loads and the reductions keep
// the pixel sums as exact integers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "palx.h"

namespace palx {

constexpr int kFrameBlocks = 1024;  // partial sums per reduction
constexpr int kCtlState = 8;        // controller state: filt, filt1, e_prev, e_int, Re

// 16-byte vector of the plane's real type
template <typename Tr> struct Vec16;
template <> struct Vec16<float> { using T = float4; static constexpr int n = 4; };
template <> struct Vec16<double> { using T = double2; static constexpr int n = 2; };

template <typename Tr>
__device__ __forceinline__ Tr lane(const typename Vec16<Tr>::T& v, int i) {
  return reinterpret_cast<const Tr*>(&v)[i];
}

// Pixel of one value, _float_to_pixels_gpu (turbo_wrapper.py:317-337): a
// constant plane maps to 128, otherwise trunc(clip(1 + (v - fmin)/rng*254, 1, 255))
// with each operation rounded in Tr on its own (no fused multiply-add), like
// the reference's elementwise CuPy kernels.  NaN maps to 0.
template <typename Tr>
__device__ __forceinline__ unsigned pixel_of(Tr v, Tr fmin, Tr rng, bool cst) {
#pragma clang fp contract(off)
  if (cst) return 128u;
  const Tr norm = (v - fmin) / rng;
  Tr pf = (Tr)1 + norm * (Tr)254;
  if (!(pf == pf)) return 0u;
  pf = pf < (Tr)1 ? (Tr)1 : (pf > (Tr)255 ? (Tr)255 : pf);
  return (unsigned)pf;
}

// min / max of the plane, the range (fmax - fmin in Tr) and the
// constant-plane test |rng| <= float32(1e-12) of the reference
template <typename Tr>
struct PixMap {
  Tr fmin, rng;
  bool cst;
  __device__ explicit PixMap(const double* rec) {
#pragma clang fp contract(off)
    fmin = (Tr)rec[PALX_FR_FMIN];
    rng = (Tr)rec[PALX_FR_FMAX] - fmin;
    cst = fabs(rng) <= (Tr)1.0e-12f;
  }
  __device__ unsigned operator()(Tr v) const { return pixel_of<Tr>(v, fmin, rng, cst); }
};

// The display stretch (turbo_main.py:2109-2113) in NumPy's float32
// arithmetic: lo and inv are formed in double like the Python scalars, then
// rounded to float32 when they meet the float32 pixels; round half to even.
struct Stretch {
  float lo, inv;
  __device__ explicit Stretch(const double* rec, double k_std) {
    const double mu = rec[PALX_FR_MU], sig = rec[PALX_FR_SIGMA];
    const double l = mu - k_std * sig, hgh = mu + k_std * sig;
    lo = (float)l;
    inv = (hgh - l) != 0.0 ? (float)(255.0 / (hgh - l)) : 0.0f;
  }
  __device__ unsigned operator()(unsigned p) const {
#pragma clang fp contract(off)
    float v = ((float)p - lo) * inv;
    v = rintf(v);
    v = v < 0.0f ? 0.0f : (v > 255.0f ? 255.0f : v);
    return (unsigned)v;
  }
};

// ---- 1. per-block min / max (NaN if the block saw one)
template <typename Tr>
__global__ void __launch_bounds__(256) k_frame_minmax(const Tr* __restrict__ f, size_t n,
                                                      double* __restrict__ part) {
  using V = Vec16<Tr>;
  __shared__ double smn[256], smx[256];
  double mn = INFINITY, mx = -INFINITY;
  bool nan = false;
  const size_t nv = n / V::n;
  const typename V::T* fv = reinterpret_cast<const typename V::T*>(f);
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < nv; i += (size_t)gridDim.x * 256) {
    const typename V::T x = fv[i];
#pragma unroll
    for (int q = 0; q < V::n; ++q) {
      const double v = (double)lane<Tr>(x, q);
      if (v != v) nan = true;
      mn = fmin(mn, v);
      mx = fmax(mx, v);
    }
  }
  smn[threadIdx.x] = nan ? NAN : mn;
  smx[threadIdx.x] = nan ? NAN : mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double a = smn[threadIdx.x], b = smn[threadIdx.x + s];
      const double c = smx[threadIdx.x], d = smx[threadIdx.x + s];
      smn[threadIdx.x] = (a != a || b != b) ? NAN : fmin(a, b);
      smx[threadIdx.x] = (c != c || d != d) ? NAN : fmax(c, d);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * (size_t)blockIdx.x] = smn[0];
    part[2 * (size_t)blockIdx.x + 1] = smx[0];
  }
}

// ---- 2. the plane's min / max into the record (one block)
__global__ void __launch_bounds__(256) k_frame_range(const double* __restrict__ part, int nblk,
                                                     double* __restrict__ rec) {
  __shared__ double smn[256], smx[256];
  double mn = INFINITY, mx = -INFINITY;
  bool nan = false;
  for (int b = threadIdx.x; b < nblk; b += 256) {
    const double a = part[2 * b], c = part[2 * b + 1];
    if (a != a || c != c) nan = true;
    mn = fmin(mn, a);
    mx = fmax(mx, c);
  }
  smn[threadIdx.x] = nan ? NAN : mn;
  smx[threadIdx.x] = nan ? NAN : mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double a = smn[threadIdx.x], b = smn[threadIdx.x + s];
      const double c = smx[threadIdx.x], d = smx[threadIdx.x + s];
      smn[threadIdx.x] = (a != a || b != b) ? NAN : fmin(a, b);
      smx[threadIdx.x] = (c != c || d != d) ? NAN : fmax(c, d);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    rec[PALX_FR_FMIN] = smn[0];
    rec[PALX_FR_FMAX] = smx[0];
  }
}

// ---- 3. per-block sum and sum of squares of the [1,255] pixels
template <typename Tr>
__global__ void __launch_bounds__(256) k_frame_pixsum(const Tr* __restrict__ f, size_t n,
                                                      const double* __restrict__ rec,
                                                      unsigned long long* __restrict__ part) {
  using V = Vec16<Tr>;
  __shared__ unsigned long long s1[256], s2[256];
  const PixMap<Tr> pm(rec);
  unsigned long long a1 = 0, a2 = 0;
  const size_t nv = n / V::n;
  const typename V::T* fv = reinterpret_cast<const typename V::T*>(f);
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < nv; i += (size_t)gridDim.x * 256) {
    const typename V::T x = fv[i];
#pragma unroll
    for (int q = 0; q < V::n; ++q) {
      const unsigned p = pm(lane<Tr>(x, q));
      a1 += p;
      a2 += p * p;
    }
  }
  s1[threadIdx.x] = a1;
  s2[threadIdx.x] = a2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      s1[threadIdx.x] += s1[threadIdx.x + s];
      s2[threadIdx.x] += s2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * (size_t)blockIdx.x] = s1[0];
    part[2 * (size_t)blockIdx.x + 1] = s2[0];
  }
}

// ---- 4. the frame record and the viscosity controller (one block).
// diag: kFrameBlocks x 5 partials of k_diag {E, Lsum, Zw, Pal, P00}.
// ctl_state: {filt, filt1, e_prev, e_int, Re} persisting between frames.
__global__ void __launch_bounds__(256) k_frame_record(const unsigned long long* __restrict__ pix,
                                                      const double* __restrict__ diag, int nblk,
                                                      double* __restrict__ scal,
                                                      double* __restrict__ ctl_state,
                                                      palx_frame_ctl c, int N, int M, int width,
                                                      int height, int adapt,
                                                      double* __restrict__ rec) {
  __shared__ unsigned long long s1[256], s2[256];
  __shared__ double sd[5][256];
  const int t = threadIdx.x;
  unsigned long long a1 = 0, a2 = 0;
  double d[5] = {0, 0, 0, 0, 0};
  for (int b = t; b < nblk; b += 256) {
    a1 += pix[2 * b];
    a2 += pix[2 * b + 1];
    for (int q = 0; q < 5; ++q) d[q] += diag[(size_t)b * 5 + q];
  }
  s1[t] = a1;
  s2[t] = a2;
  for (int q = 0; q < 5; ++q) sd[q][t] = d[q];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      s1[t] += s1[t + s];
      s2[t] += s2[t + s];
      for (int q = 0; q < 5; ++q) sd[q][t] += sd[q][t + s];
    }
    __syncthreads();
  }
  if (t != 0) return;
  const double npx = (double)M * (double)M;
  // mean and population standard deviation of the pixels (numpy mean / std)
  const double S1 = (double)s1[0], S2 = (double)s2[0];
  const double mu = S1 / npx;
  const double var = fmax(0.0, (S2 - S1 * mu) / npx);
  const double sig = sqrt(var);
  const bool stop = sig < 1.0;
  // flow_eddy_metrics (turbo_simulator.py:764-810), energy_joules
  // (turbo_main.py:1947-1959), pal_over_ens_kmax2 (turbo_main.py:2015-2073)
  const double usum = scal[PALX_S_USUM];
  const double U = sqrt(fmax(0.0, usum / npx));
  const double E = sd[0][0];
  double L = NAN, tau = NAN;
  if (E > 0.0) {
    L = 2.0 * M_PI * sd[1][0] / E;
    tau = U > 0.0 ? L / U : NAN;
  }
  const double total = sd[2][0] - sd[4][0];
  const double gz = (double)(N / 2), ax = (double)(N / 2 - 1);
  const double kmax2 = ax * ax + gz * gz;
  const double ratio = (total <= 0.0 || kmax2 <= 0.0) ? 0.0 : sd[3][0] / (total * kmax2);
  double adapted = 0.0;
  if (adapt && !stop) {
    // _palin_filter_2nd (turbo_main.py:2154-2166): two cascaded EMAs
    const double p_raw = 10000.0 * ratio;
    const double alpha = 1.0 / (c.palin_tau + 1.0);
    if (ctl_state[0] == 0.0) {
      ctl_state[1] = p_raw;
      ctl_state[0] = p_raw;
    } else {
      ctl_state[1] = ctl_state[1] + alpha * (p_raw - ctl_state[1]);
      ctl_state[0] = ctl_state[0] + alpha * (ctl_state[1] - ctl_state[0]);
    }
    // adapt_visc (turbo_main.py:2168-2214): multiplicative update in log space
    const double e = (ctl_state[0] - c.target) / c.target;
    if (!(fabs(e) < c.deadband)) {
      if (c.ki != 0.0) ctl_state[3] += e;
      const double de = e - ctl_state[2];
      ctl_state[2] = e;
      double u = c.kp * e + c.ki * ctl_state[3] + c.kd * de;
      u = fmax(-c.max_frac, fmin(c.max_frac, u));
      const double nu_new = scal[PALX_S_VISC] * exp(u);
      const double re_eff = fmax(c.re0 / c.re_factor, fmin(c.re_factor * c.re0, 1.0 / nu_new));
      ctl_state[4] = re_eff;
      scal[PALX_S_VISC] = 1.0 / re_eff;
      adapted = 1.0;
    }
  }
  for (int q = 0; q < PALX_FRAME_NREC; ++q)
    if (q != PALX_FR_FMIN && q != PALX_FR_FMAX) rec[q] = 0.0;
  rec[PALX_FR_T] = scal[PALX_S_T];
  rec[PALX_FR_DT] = scal[PALX_S_DT];
  rec[PALX_FR_CN] = scal[PALX_S_CN];
  rec[PALX_FR_VISC] = scal[PALX_S_VISC];
  rec[PALX_FR_RE] = ctl_state[4];
  rec[PALX_FR_IT] = scal[PALX_S_IT];
  rec[PALX_FR_MU] = mu;
  rec[PALX_FR_SIGMA] = sig;
  rec[PALX_FR_STOP] = stop ? 1.0 : 0.0;
  rec[PALX_FR_RATIO] = ratio;
  rec[PALX_FR_FILT] = ctl_state[0];
  rec[PALX_FR_ADAPTED] = adapted;
  rec[PALX_FR_U] = U;
  rec[PALX_FR_L] = L;
  rec[PALX_FR_TAU_L] = tau;
  rec[PALX_FR_EJ] = 0.5 * (2.0 * M_PI) * (2.0 * M_PI) * usum / npx;
  rec[PALX_FR_ENSTROPHY] = total;
  rec[PALX_FR_PALINSTROPHY] = sd[3][0];
  rec[PALX_FR_KMAX2] = kmax2;
  rec[PALX_FR_ENERGY] = E;
  rec[PALX_FR_WIDTH] = (double)width;
  rec[PALX_FR_HEIGHT] = (double)height;
  rec[PALX_FR_CFLM] = scal[PALX_S_CFLM];
}

// ---- 5. display pixels: map (and stretch), then integer down / up scaling
template <typename Tr>
__global__ void __launch_bounds__(256) k_frame_display(const Tr* __restrict__ f, int M, int down, int up,
                                                       int width, int height, const double* __restrict__ rec,
                                                       int stretch, double k_std,
                                                       uint8_t* __restrict__ out) {
  const PixMap<Tr> pm(rec);
  const Stretch st(rec, k_std);
  const size_t n = (size_t)width * height;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const int oy = (int)(i / width), ox = (int)(i % width);
    const int sy = up > 1 ? oy / up : oy * down, sx = up > 1 ? ox / up : ox * down;
    unsigned p = pm(f[(size_t)sy * M + sx]);
    if (stretch) p = st(p);
    out[i] = (uint8_t)p;
  }
}

__global__ void k_frame_ctl_reset(double* __restrict__ ctl_state, double re) {
  if (threadIdx.x < kCtlState) ctl_state[threadIdx.x] = threadIdx.x == 4 ? re : 0.0;
}

}  // namespace palx
