// Host-side seeded initial field for the palinstrophy DNS (PAO / KM3 spectra).
//
// Follows palinstrophy/turbo_simulator.py:213-485 (_pao_build_ur_and_stats_impl):
// a Fortran LCG with a 97-entry shuffle table draws one phase per (z, kx<N/2)
// mode in z-outer / x-inner order; amplitudes follow exp(-(k/K0)^2)/(pi K0^2)
// split divergence-free; the kx=0 column is made Hermitian and the z=N/2 row
// zeroed; KM3 then redistributes integer k-shells toward k^-3 with a sin-hash
// jitter, renormalised to the PAO energy.
//
// The LCG is inherently sequential (each draw depends on the shuffle table),
// so one pass draws every mode's phase in loop order (integer work, ~8.4M
// cheap iterations at N=4096).  The per-mode transcendental work (phases,
// amplitudes, the KM3 hash and rescale) then runs on host threads, and every
// floating-point accumulation (KM3 shell sums, the averages) keeps the
// reference's loop order, so the result does not depend on the thread
// count.
//
// Arithmetic contract: float32 where the reference uses np.float32 scalars,
// double where it uses Python floats.  Transcendentals of float32 arguments
// are the correctly rounded float32 of the double function.  This file must
// be compiled with -ffp-contract=off so that no fused multiply-add changes a
// float32 rounding.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "palx.h"

namespace {

constexpr int IMM = 420029, IT = 2017, ID = 5011;  // turbo_simulator.py:242-244

inline float f32_exp(float x) { return (float)std::exp((double)x); }
inline float f32_cos(float x) { return (float)std::cos((double)x); }
inline float f32_sin(float x) { return (float)std::sin((double)x); }

// turbo_simulator.py:195-204
double hash01(int x, int z, int seed, int salt) {
  double a = ((double)x + 1.0) * 12.9898 + ((double)z + 1.0) * 78.233 +
             ((double)seed + 1.0) * 0.010001 + ((double)salt + 1.0) * 37.719;
  double s = std::sin(a) * 43758.5453123;
  return s - std::floor(s);
}

struct C64 {
  float re, im;
};

// Run f(lo, hi) over [0, n) split into contiguous chunks on up to
// PALX_INIT_THREADS (default: hardware threads, at most 16) threads.
template <typename F>
void parallel_for(int n, F f) {
  int nt = (int)std::thread::hardware_concurrency();
  if (const char* e = std::getenv("PALX_INIT_THREADS")) nt = std::atoi(e);
  nt = std::max(1, std::min(nt, 16));
  nt = std::min(nt, std::max(1, n / 64));
  if (nt <= 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt);
  for (int i = 0; i < nt; ++i) {
    const int lo = (int)((int64_t)n * i / nt), hi = (int)((int64_t)n * (i + 1) / nt);
    th.emplace_back([=] { f(lo, hi); });
  }
  for (auto& t : th) t.join();
}

}  // namespace

extern "C" int palx_pao_generate(int32_t N, double K0d, double Re, int32_t seed_init,
                                 int32_t km3, float* uc0, double* stats) {
  if (N < 4 || (N & 1) || uc0 == nullptr) return PALX_EINVAL;
  const int ND2 = N / 2, NE = N, NED2 = NE / 2;
  const float PI = (float)3.14159265358979;  // turbo_simulator.py:237
  const float K0 = (float)K0d;

  std::vector<float> alfa(ND2), gamma(NE, 0.0f);  // turbo_simulator.py:1079-1094
  for (int x = 0; x < ND2; ++x) alfa[x] = (float)x;
  for (int z = 1; z <= NED2; ++z) {
    gamma[z] = (float)z;
    gamma[NE - z] = -gamma[z];
  }

  // UR[x][z][c]; every entry is written by the amplitude loop below, so the
  // buffer is not cleared first (its pages are first touched by the threads)
  std::unique_ptr<C64[]> UR(new C64[(size_t)ND2 * NE * 2]);
  auto at = [&](int x, int z, int c) -> C64& { return UR[((size_t)x * NE + z) * 2 + c]; };

  int seed = seed_init;
  float RANVEC[97];
  for (int i = 0; i < 97; ++i) seed = (int)(((int64_t)seed * IMM + IT) % ID);
  for (int i = 0; i < 97; ++i) {
    seed = (int)(((int64_t)seed * IMM + IT) % ID);
    RANVEC[i] = (float)seed / (float)ID;
  }
  const float NORM = PI * K0 * K0;
  const float AMP2_FLOOR = 1.0e-40f;
  const bool use_km3 = km3 != 0;

  // the shuffled draws of every mode in loop order, z outer, x inner
  // (turbo_simulator.py:276-293)
  std::unique_ptr<float[]> draw(new float[(size_t)NE * NED2]);
  for (int z = 0; z < NE; ++z) {
    for (int x = 0; x < NED2; ++x) {
      seed = (int)(((int64_t)seed * IMM + IT) % ID);
      float r = (float)seed / (float)ID;
      int idx = (int)((double)r * 97.0);
      if (idx < 0) idx = 0;
      if (idx > 96) idx = 96;
      draw[(size_t)z * NED2 + x] = RANVEC[idx];
      RANVEC[idx] = r;
    }
  }

  parallel_for(NE, [&](int z0, int z1) {  // turbo_simulator.py:293-319
  for (int z = z0; z < z1; ++z) {
    const float gz = gamma[z];
    for (int x = 0; x < NED2; ++x) {
      float v = draw[(size_t)z * NED2 + x];
      float th = (2.0f * PI) * v;
      float ca = f32_cos(th), sa = f32_sin(th);
      float ax = alfa[x];
      float K2 = ax * ax + gz * gz;
      float K = K2 > 0.0f ? std::sqrt(K2) : 0.0f;
      float q = K / K0;
      float e = f32_exp((-q) * q);
      if (ax == 0.0f) {
        at(x, z, 1) = C64{0.0f, 0.0f};
        float absu2 = e / NORM;
        if (use_km3 && absu2 < AMP2_FLOOR) absu2 = AMP2_FLOOR;
        float amp = std::sqrt(absu2);
        at(x, z, 0) = C64{amp * ca, amp * sa};
      } else {
        float denom = 1.0f + (gz * gz) / (ax * ax);
        float absw2 = e / (denom * NORM);
        if (use_km3 && absw2 < AMP2_FLOOR) absw2 = AMP2_FLOOR;
        float ampw = std::sqrt(absw2);
        C64 w{ampw * ca, ampw * sa};
        float ratio = -(gz / ax);
        at(x, z, 1) = w;
        at(x, z, 0) = C64{ratio * w.re, ratio * w.im};
      }
    }
  }
  });
  at(0, 0, 0) = C64{0.0f, 0.0f};  // turbo_simulator.py:322-323
  at(0, 0, 1) = C64{0.0f, 0.0f};
  for (int z = 1; z < NED2; ++z) {  // turbo_simulator.py:328-330
    for (int c = 0; c < 2; ++c) {
      C64 v = at(0, z, c);
      at(0, NE - z, c) = C64{v.re, -v.im};
    }
  }
  for (int x = 0; x < ND2; ++x) {  // turbo_simulator.py:333-335
    at(x, NED2, 0) = C64{0.0f, 0.0f};
    at(x, NED2, 1) = C64{0.0f, 0.0f};
  }

  if (use_km3) {  // turbo_simulator.py:337-426
    const double k_nyq = (double)ND2;
    const int kmax = ND2;
    std::vector<double> shell_energy(kmax + 1, 0.0), shell_weight(kmax + 1, 0.0),
        target(kmax + 1, 0.0);
    // per mode (x outer, z inner): shell (-1: outside), m (u1u1 + u3u3), m jit
    const size_t nm = (size_t)ND2 * NE;
    std::unique_ptr<int[]> m_shell(new int[nm]);  // written for every mode
    std::unique_ptr<double[]> m_energy(new double[nm]), m_weight(new double[nm]),
        m_jit(new double[nm]);  // read only where m_shell >= 0, which writes them
    parallel_for(ND2, [&](int x0, int x1) {
      for (int x = x0; x < x1; ++x) {
        const double ax2 = (double)alfa[x] * (double)alfa[x];
        const double m = (x == 0) ? 1.0 : 2.0;
        for (int z = 0; z < NE; ++z) {
          const size_t i = (size_t)x * NE + z;
          C64 U1 = at(x, z, 0), U3 = at(x, z, 1);
          double u1u1 = (double)U1.re * (double)U1.re + (double)U1.im * (double)U1.im;
          double u3u3 = (double)U3.re * (double)U3.re + (double)U3.im * (double)U3.im;
          double gz2 = (double)gamma[z] * (double)gamma[z];
          double k = std::sqrt(ax2 + gz2);
          int shell = (int)(k + 0.5);
          m_shell[i] = -1;
          if (k > 0.0 && k <= k_nyq && shell <= kmax) {
            int z_hash = (x == 0 && z > NED2) ? NE - z : z;
            double jit = 0.75 + 0.50 * hash01(x, z_hash, seed_init, 17);
            m_shell[i] = shell;
            m_energy[i] = m * (u1u1 + u3u3);
            m_weight[i] = m * jit;
            m_jit[i] = jit;
          }
        }
      }
    });
    for (size_t i = 0; i < nm; ++i) {  // turbo_simulator.py:355-372, in loop order
      const int sh = m_shell[i];
      if (sh >= 0) {
        shell_energy[sh] += m_energy[i];
        shell_weight[sh] += m_weight[i];
      }
    }
    double k0f = (double)K0;
    if (k0f < 1.0) k0f = 1.0;
    for (int sh = 1; sh <= kmax; ++sh) {
      double k = (double)sh;
      if (k < k0f) {
        double low = k / k0f;
        target[sh] = (low * low * low * low) / (k0f * k0f * k0f);
      } else {
        target[sh] = 1.0 / (k * k * k);
      }
    }
    double orig = 0.0, tgt = 0.0;
    for (int sh = 1; sh <= kmax; ++sh) {
      orig += shell_energy[sh];
      tgt += target[sh];
    }
    if (orig > 0.0 && tgt > 0.0) {
      const double norm = orig / tgt;
      parallel_for(ND2, [&](int x0, int x1) {
      for (int x = x0; x < x1; ++x) {
        const double ax = (double)alfa[x];
        const double ax2 = ax * ax;
        for (int z = 0; z < NE; ++z) {
          const double gz = (double)gamma[z];
          const double gz2 = gz * gz;
          double k = std::sqrt(ax2 + gz2);
          int shell = (int)(k + 0.5);
          if (k > 0.0 && k <= k_nyq && shell <= kmax && shell_weight[shell] > 0.0) {
            int z_hash = z;
            double sign = 1.0;
            if (x == 0 && z > NED2) {
              z_hash = NE - z;
              sign = -1.0;
            }
            const double jit = m_jit[(size_t)x * NE + z];
            double me = (norm * target[shell]) * jit / shell_weight[shell];
            double th = sign * 2.0 * M_PI * hash01(x, z_hash, seed_init, 53);
            float pr = (float)std::cos(th), pi_ = (float)std::sin(th);
            if (ax == 0.0) {
              float a = (float)std::sqrt(me);
              at(x, z, 0) = C64{a * pr, a * pi_};
              at(x, z, 1) = C64{0.0f, 0.0f};
            } else {
              double denom = 1.0 + gz2 / ax2;
              float a = (float)std::sqrt(me / denom);
              C64 w{a * pr, a * pi_};
              float r = (float)(-gz / ax);
              at(x, z, 0) = C64{r * w.re, r * w.im};
              at(x, z, 1) = w;
            }
          } else {
            at(x, z, 0) = C64{0.0f, 0.0f};
            at(x, z, 1) = C64{0.0f, 0.0f};
          }
        }
      }
      });
    }
  }

  // averages (turbo_simulator.py:431-470)
  double A1 = 0, A2 = 0, A3 = 0, A4 = 0, A5 = 0, A6 = 0, A7 = 0, E110 = 0;
  for (int x = 0; x < ND2; ++x) {
    const double ax2 = (double)alfa[x] * (double)alfa[x];
    const double m = (x == 0) ? 1.0 : 2.0;
    for (int z = 0; z < NE; ++z) {
      C64 U1 = at(x, z, 0), U3 = at(x, z, 1);
      double u1u1 = (double)U1.re * (double)U1.re + (double)U1.im * (double)U1.im;
      double u3u3 = (double)U3.re * (double)U3.re + (double)U3.im * (double)U3.im;
      double gz2 = (double)gamma[z] * (double)gamma[z];
      double K2f = ax2 + gz2;
      A1 += m * u1u1;
      A2 += m * u3u3;
      A3 += m * u1u1 * ax2;
      A4 += m * u1u1 * gz2;
      A5 += m * u3u3 * ax2;
      A6 += m * u3u3 * gz2;
      A7 += m * (u1u1 + u3u3) * K2f * K2f;
      if (x == 0) E110 += u1u1;
    }
  }
  // compact velocity spectra, uc0[c][z][kx] interleaved complex64
  parallel_for(NE, [&](int z0, int z1) {
    for (int c = 0; c < 2; ++c)
      for (int z = z0; z < z1; ++z)
        for (int x = 0; x < ND2; ++x) {
          C64 v = at(x, z, c);
          size_t o = (((size_t)c * NE + z) * ND2 + x) * 2;
          uc0[o] = v.re;
          uc0[o + 1] = v.im;
        }
  });
  if (stats) {
    const float Ref = (float)Re;
    const float visc = (float)(1.0 / (double)Ref);
    stats[0] = (double)seed;  // seed out
    stats[1] = A1 + A2;       // Q2
    stats[2] = A3 + A4 + A5 + A6;  // W2
    stats[3] = E110;
    stats[4] = (double)visc;
    stats[5] = A1;
    stats[6] = A2;
    stats[7] = A3;
    stats[8] = A4;
    stats[9] = A5;
    stats[10] = A6;
    stats[11] = A7;
  }
  return PALX_OK;
}

// omega_0 = i (gz * u1 - ax * u2) in float32 (turbo_simulator.py:1314-1361)
extern "C" int palx_calcom_f32(int32_t N, const float* uc0, float* om) {
  if (N < 4 || (N & 1) || !uc0 || !om) return PALX_EINVAL;
  const int H = N / 2;
  for (int z = 0; z < N; ++z) {
    const float gz = (z < H) ? (float)z : (float)(z - N);  // gamma[H] = -H
    for (int x = 0; x < H; ++x) {
      float ax = (float)x;
      size_t o1 = ((size_t)z * H + x) * 2;
      size_t o2 = (((size_t)N + z) * H + x) * 2;
      float dr = gz * uc0[o1] - ax * uc0[o2];
      float di = gz * uc0[o1 + 1] - ax * uc0[o2 + 1];
      om[o1] = -di;
      om[o1 + 1] = dr;
    }
  }
  return PALX_OK;
}
