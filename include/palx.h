/* palx.h -- C ABI of the MI355X-native palinstrophy DNS time step (libpalx.so).
 *
 * The reference (mannetroll/palinstrophy) has no FFI: its "operator API" is a
 * set of Python functions on a mutable DnsState in
 * palinstrophy/turbo_simulator.py.  Each entry point below replaces the
 * reference function cited beside it; the Python facade
 * palinstrophy_amd/turbo_simulator.py binds them with ctypes under the
 * reference's own names.
 *
 * Conventions
 *   - Every function returns 0 (PALX_OK) or a negative PALX_E* code; the
 *     message of the last failure on the calling thread is available from
 *     palx_last_error().
 *   - A handle owns all device memory of one simulation and one HIP stream
 *     (set with palx_set_stream, default: the null stream).  Handles are not
 *     thread-safe; calls only enqueue work unless documented as blocking.
 *   - precision is 4 (float32 arithmetic, complex64 spectra) or 8 (float64).
 *   - Complex arrays are interleaved (re, im) pairs in the handle precision.
 *   - Spectral layout is compact [z][kx], z in [0,N) (kz = z for z < N/2,
 *     z - N otherwise), kx in [0,N/2).
 */
#ifndef PALX_H
#define PALX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct palx_t palx_t;

enum {
  PALX_OK = 0,
  PALX_EINVAL = -1,   /* bad argument (size, precision, field id, ...) */
  PALX_ENOMEM = -2,   /* device allocation failed */
  PALX_EHIP = -3,     /* HIP runtime error */
  PALX_ESTATE = -4,   /* call out of order (e.g. update before products) */
  PALX_EUNSUP = -5    /* grid size not supported by this build */
};

/* Time steppers (turbo_simulator.py:51 TimeStepper). */
enum { PALX_CNAB2 = 0, PALX_LS_IMEX_RK3 = 1 };
enum { PALX_STAGE_IN_STEP = 0x1000 }; /* palx_stage_update flag, below */

/* Fields for palx_download / palx_upload / palx_field_ptr. */
enum {
  PALX_F_OM = 0,    /* omega_hat, complex (N, N/2)          -- DnsState.om2      */
  PALX_F_FNM1 = 1,  /* previous nonlinear term, complex (N, N/2) -- DnsState.fnm1 */
  PALX_F_UR = 2,    /* physical planes, real (3, M, M), M=3N/2 -- DnsState.ur_full */
  PALX_F_WORK = 3,  /* velocity work planes, z physical / kx spectral, complex 2 x M*N/2 tiled by z:
                       (kx, z) at ((z/P)*N/2 + kx)*P + z%P, P = 8 (FP32; 4 when M <= 3072) / 4 (FP64),
                       halved until it divides M */
  PALX_F_UC0 = 4,   /* compact velocity spectra u_hat, w_hat, complex (2, N, N/2) */
  PALX_F_SCAL = 5,  /* FP64 scalar block, PALX_NSCAL doubles */
  PALX_F_ADJ_OM = 6 /* adjoint lambda_omega / gradient, complex (N, N/2) */
};

/* FP64 device scalars (turbo_simulator.py:690-697 t/dt/cn/cnm1, 670-671 visc/cflnum). */
enum {
  PALX_S_DT = 0, PALX_S_CN = 1, PALX_S_CNM1 = 2, PALX_S_T = 3, PALX_S_VISC = 4,
  PALX_S_CFLNUM = 5, PALX_S_CFLM = 6, PALX_S_USUM = 7, PALX_S_INVDX = 8, PALX_S_IT = 9,
  PALX_S_NAN = 10,  /* set to 1 when the CFL reduction met a NaN */
  PALX_NSCAL = 16
};

/* Getter variables (turbo_wrapper.py:27-31 VAR_*). */
enum { PALX_VAR_U = 0, PALX_VAR_V = 1, PALX_VAR_ENERGY = 2, PALX_VAR_OMEGA = 3, PALX_VAR_STREAM = 4 };

/* ---- lifetime: create_dns_state (turbo_simulator.py:878-1054) ----------------
 * N: 2^k in [8, 8192] (and 16384 with precision 4) or 3*2^k in [48, 12288]:
 * the reference GUI's N menu up to 8192 except 7776, and 12288 and 16384
 * (turbo_main.py:589-591); other N return PALX_EUNSUP. */
int palx_create(int32_t N, int32_t precision, palx_t** out);
int palx_destroy(palx_t* h);
int palx_set_stream(palx_t* h, void* hip_stream);
int palx_info(const palx_t* h, int32_t* N, int32_t* M, int32_t* precision);
int palx_last_error(char* buf, size_t n);
int palx_sync(palx_t* h); /* blocking: waits for the handle stream */

/* ---- seeded initial field: _pao_build_ur_and_stats (turbo_simulator.py:213-485)
 * Host-only, no handle.  uc0 receives complex64 (2, N, N/2) u_hat/w_hat in
 * compact order; stats[16] receives {seed_out, Q2, W2, E110, visc, A1..A7}.
 * Runs on up to 16 host threads (environment PALX_INIT_THREADS overrides);
 * the result is the same bits for any thread count. */
int palx_pao_generate(int32_t N, double K0, double Re, int32_t seed, int32_t km3, float* uc0,
                      double* stats);
/* omega_0 = i(gamma u1 - alfa u2) in float32: dns_calcom_from_uc_full (turbo_simulator.py:1314-1361). */
int palx_calcom_f32(int32_t N, const float* uc0, float* om);

/* Load a complete initial state (dns_pao_host_init scatter + calcom,
 * turbo_simulator.py:1187-1239): uc0 complex64 (2,N,N/2), om complex64 (N,N/2)
 * (both host pointers, promoted for precision 8), fnm1 cleared, then the
 * padded inverse of uc0 (dns_step2a, turbo_simulator.py:1949-2037) fills the
 * work planes. */
int palx_load_initial(palx_t* h, const float* uc0_c64, const float* om_c64);
/* Rebuild the velocity work planes from the velocity spectra currently in
 * PALX_F_UC0 (uploaded in the handle precision): the padded inverse of
 * dns_step2a (turbo_simulator.py:1949-2037) used after a restart load. */
int palx_load_velocity(palx_t* h);

/* ---- data movement ------------------------------------------------------------
 * download/upload exchange PALX_F_OM and PALX_F_FNM1 in the compact [z][kx]
 * layout above; on the device they are stored column-major [kx][z], which
 * is what palx_field_ptr exposes for those two fields. */
int palx_field_ptr(palx_t* h, int32_t field, void** ptr, size_t* bytes);
int palx_download(palx_t* h, int32_t field, void* dst, size_t bytes);      /* blocking */
int palx_upload(palx_t* h, int32_t field, const void* src, size_t bytes);  /* blocking */
int palx_set_scalar(palx_t* h, int32_t id, double value);                 /* blocking */
int palx_get_scalars(palx_t* h, double* out, int32_t n);                   /* blocking */

/* ---- the stage pipeline ---------------------------------------------------------
 * palx_stage_products: dns_step2b (turbo_simulator.py:1378-1513) -- row pass:
 *   x-inverse of the velocity work planes, u^2-w^2 and uw, x-forward truncated
 *   to kx<N/2.
 * palx_stage_update: dns_step3 CNAB2 (turbo_simulator.py:1588-1771) or one
 *   LS-IMEX-RK3 stage (turbo_simulator.py:1774-1943, stage 0..2) -- column pass:
 *   z-forward, nonlinear term, implicit update of omega/fnm1, velocity
 *   reconstruction, padded z-inverse (the z half of dns_step2a).
 *   stage | PALX_STAGE_IN_STEP (LS-IMEX-RK3 only; no reference counterpart:
 *   the reference runs its three stages inside dns_step_ls_imex_rk3,
 *   turbo_simulator.py:1900-1943, where no caller sees between them): the
 *   stages of one whole step, run 0, 1, 2 in that order.  Stages 0 and 1 then
 *   leave omega's storage alone and put the next stage's pre-update term
 *   omega (1 - K2 beta dt visc) + fnm1 beta dt in fnm1's storage, and
 *   stages 1 and 2 read it in place of omega and fnm1: four of the eleven
 *   omega / fnm1 plane transfers of a step fewer.  After stage 2 omega, fnm1
 *   and the velocity are bit for bit what the plain stages give.  While such
 *   a step is open (after its stage 0 or 1) everything that reads or writes
 *   omega / fnm1 -- palx_step, plain stages, palx_download / palx_upload /
 *   palx_field_ptr of PALX_F_OM / PALX_F_FNM1, the omega and stream-function
 *   getters and frames, palx_recon_velocity, the adjoint, ensemble steps and
 *   tuning -- fails with PALX_ESTATE; palx_stage_products, palx_physical and
 *   the velocity getters do not (the velocity planes hold every stage's
 *   velocity as usual), and palx_load_initial closes the step.  palx_step
 *   runs its LS-IMEX-RK3 stages this way.
 * palx_step: nsteps full steps; when nextdt_every > 0, NEXTDT runs after step
 *   it (1-based, counted from palx_set_iteration) when it%nextdt_every==0 or
 *   it==1 -- the cadence of run_dns (turbo_simulator.py:2842). Advances the
 *   device time by the pre-update dt like turbo_simulator.py:2848. */
int palx_stage_products(palx_t* h);
int palx_stage_update(palx_t* h, int32_t method, int32_t stage);
int palx_step(palx_t* h, int32_t nsteps, int32_t method, int32_t nextdt_every);
int palx_set_iteration(palx_t* h, int64_t it);
/* Graph stepping (no reference counterpart; the reference launches CuPy
 * kernels one by one): with enable != 0, palx_step captures one step per
 * (method, NEXTDT step or not), and 8 steps without NEXTDT per method, into
 * HIP graphs on first use, on a private stream, and launches them on the
 * handle stream: 8 steps per launch wherever the next 8 steps of the call
 * run no NEXTDT, one step per launch elsewhere.  Results are identical to
 * the launch-by-launch path.  Off by default: on MI355X the eager path is as
 * fast for a single field. */
int palx_set_graph(palx_t* h, int32_t enable);

/* ---- physical fields, CFL and NEXTDT ----------------------------------------------
 * palx_physical: materialise u, w on the padded grid into planes 0,1 of
 *   PALX_F_UR and reduce max(|u|+|w|) and sum(u^2+w^2) (compute_cflm,
 *   turbo_simulator.py:2106-2125; U of flow_eddy_metrics, 764-770).
 * palx_cflm: palx_physical + blocking read of CFLM (FP64).
 * palx_next_dt: CFL-adaptive dt on device (next_dt, turbo_simulator.py:2127-2167). */
int palx_physical(palx_t* h);
int palx_cflm(palx_t* h, double* cflm);
int palx_next_dt(palx_t* h);

/* ---- diagnostics (blocking) ---------------------------------------------------------
 * out[8] = {U, L, tau_L, enstrophy, palinstrophy, kmax2, E_joules, energy_sum}
 * (flow_eddy_metrics turbo_simulator.py:764-810; pal_over_ens_kmax2
 * turbo_main.py:2015-2073; energy_joules turbo_main.py:1947-1959).
 * Runs palx_physical first. */
int palx_diag(palx_t* h, double* out);
/* compute_energy_spectrum_uv_bins (turbo_simulator.py:2225-2383).  palx_spectrum
 * returns host arrays (blocking); palx_spectrum_into writes into device
 * arrays on the handle stream without synchronising, for accumulating time
 * averages on the device (turbo_simulator.py:2386-2497 does it on the host).
 * Each bin is summed in one fixed order (no atomics): the same omega gives
 * the same bits on every call. */
int palx_spectrum(palx_t* h, double* esum, int32_t* count, int32_t nbins);
int palx_spectrum_into(palx_t* h, double* esum_dev, int32_t* count_dev, int32_t nbins);
/* No counterpart in the reference: palx_spectrum for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, blocking, same bins; it
 * calls no hipMalloc / hipFree: the device-side bins are staged in a static
 * device array of PALX_SPECTRUM_NOALLOC_MAX_BINS bins per GPU (384 KB, held
 * until the process ends).  nbins above that limit: PALX_EINVAL.  The
 * handle's device must be the current device (else PALX_ESTATE).  Calls that
 * overlap on one GPU share that array and race; each call writes every bin
 * it returns, so nothing stale from an earlier call is read. */
#define PALX_SPECTRUM_NOALLOC_MAX_BINS 32768
int palx_spectrum_noalloc(palx_t* h, double* esum, int32_t* count, int32_t nbins);

/* ---- field getters into plane 2 of PALX_F_UR ------------------------------------------
 * dns_kinetic / dns_om2_phys / dns_stream_func (turbo_simulator.py:2629-2708);
 * PALX_VAR_U / PALX_VAR_V only materialise planes 0,1. */
int palx_getter(palx_t* h, int32_t var);

/* ---- adjoint of the fixed-dt LS-IMEX-RK3 trajectory -------------------------------
 * No counterpart in the reference (which has no adjoint); the forward map is
 * dns_step_ls_imex_rk3 (turbo_simulator.py:1900-1943) at fixed dt with the
 * velocity always reconstructed from omega (turbo_simulator.py:1555-1585).
 * palx_recon_velocity: work planes from the current omega.
 * palx_adj_forward: from the current omega (fnm1 := 0) run nsteps steps,
 *   checkpointing (omega, fnm1) every ckpt_every steps; leaves omega(T).
 * palx_adj_backward: lam_T = dL/d omega(T) (host or device, compact [z][kx],
 *   handle precision, complex gradient convention dL/dRe + i dL/dIm); replays
 *   each segment from its checkpoint and sweeps the adjoint stages backwards.
 *   The gradient dL/d omega_0 is then PALX_F_ADJ_OM, and the handle holds the
 *   end of the forward trajectory again (omega(T), fnm1(T) and the velocity
 *   of omega(T)), as right after palx_adj_forward.  A second palx_adj_forward
 *   replaces the checkpoints: a backward then differentiates the new
 *   trajectory. */
int palx_recon_velocity(palx_t* h);
int palx_adj_forward(palx_t* h, int32_t nsteps, int32_t ckpt_every);
int palx_adj_backward(palx_t* h, const void* lam_T);

/* ---- kernel variants of the step passes -------------------------------------------
 * No counterpart in the reference, whose FFT plans are cuFFT's / pocketfft's
 * (turbo_simulator.py:970-979, 1255-1261).  A variant fixes the radix
 * sequence, the threads per transform and the LDS padding of the row pass
 * (palx_stage_products) or of the column pass of CNAB2 / RK3
 * (palx_stage_update); variants differ only in rounding.  Variants 6 and 7
 * exchange through LDS in two halves (real, then imaginary parts), halving
 * the LDS of a transform.
 * palx_n_variants returns the number of variant numbers in use at N (3, or 8
 * for N = 2^k >= 256 and N = 3*2^k >= 6144, where 3-5 and 7 do not exist) or
 * an error: not every V below it is compiled, so callers that loop over the
 * range filter it with palx_variant_ok; palx_variant_ok(N, precision, V) returns 1 if
 * variant V is compiled for N in that precision (it exists and its block
 * fits the LDS: at N=16384, and at N=8192 in FP64, only the split variants
 * 6 and 7 do, at N=12288 in FP64 only 6), 0 if not,
 * or an error for an unsupported N.  palx_set_variants rejects variants
 * that are not. */
#define PALX_MAX_VARIANTS 8
int palx_n_variants(int32_t N);
int palx_variant_ok(int32_t N, int32_t precision, int32_t V);
int palx_set_variants(palx_t* h, int32_t vrow, int32_t vcol);
int palx_get_variants(const palx_t* h, int32_t* vrow, int32_t* vcol);

/* ---- ensembles of independent runs of uneven N ----------------------------------
 * No counterpart in the reference (one DnsState per process); each member is
 * a full handle (palx_ens_member, owned by the ensemble) that the calls above
 * load, read and diagnose.  Members of equal N form a class whose passes run
 * as one batched launch; palx_ens_step advances every member by nsteps steps
 * with the palx_step semantics (own dt, NEXTDT cadence per member) and, with
 * PALX_ENS_GRAPH, launches captured HIP graphs: one of 8 steps wherever the
 * next 8 steps run no NEXTDT, one of a single step elsewhere.  The ensemble
 * runs on its own stream (palx_ens_set_stream: a created stream, not the
 * null stream, which cannot be captured).
 * palx_ens_tune times every variant of the row and column pass of every class
 * (batched, reps launches each, 1 <= reps <= 1000) where a step runs it:
 * alternating with the other pass at its current variant, RK3 stages cycling,
 * events around the timed launches; rows first, then columns with the chosen
 * row variant.  It keeps the fastest, and overwrites the member fields, so it
 * runs before the initial fields are loaded.  ms (may be null) receives
 * [class][pass][PALX_MAX_VARIANTS] milliseconds per launch (-1: absent).
 * palx_ens_classes lists the classes in order of first appearance. */
typedef struct palx_ens_t palx_ens_t;
enum { PALX_ENS_GRAPH = 1, PALX_ENS_SERIAL = 2 };
int palx_ens_create(int32_t n, const int32_t* N, int32_t precision, palx_ens_t** out);
int palx_ens_destroy(palx_ens_t* e);
int palx_ens_set_stream(palx_ens_t* e, void* hip_stream);
int palx_ens_member(palx_ens_t* e, int32_t m, palx_t** out);
int palx_ens_classes(palx_ens_t* e, int32_t* nclass, int32_t* N, int32_t* count, int32_t* vrow,
                     int32_t* vcol);
int palx_ens_set_variants(palx_ens_t* e, int32_t N, int32_t vrow, int32_t vcol);
int palx_ens_tune(palx_ens_t* e, int32_t reps, float* ms);
int palx_ens_step(palx_ens_t* e, int32_t nsteps, int32_t method, int32_t nextdt_every, int32_t flags);
int palx_ens_set_iteration(palx_ens_t* e, int64_t it);
int palx_ens_sync(palx_ens_t* e); /* blocking */

/* ---- display frame: the per-frame work of the reference GUI, on the device -----
 * palx_frame enqueues on the handle stream, with no host synchronisation, what
 * the reference does on the host for each rendered frame:
 *   1. u, w on the padded grid (the physical pass) and, for var ENERGY /
 *      OMEGA / STREAM, the getter into plane 2 (turbo_wrapper.py:353-386);
 *   2. with PALX_FRAME_NEXTDT, the deferred NEXTDT of get_frame_pixels
 *      (turbo_wrapper.py:488-495; next_dt, turbo_simulator.py:2127-2167);
 *   3. the uint8 [1,255] map of the plane (_float_to_pixels_gpu,
 *      turbo_wrapper.py:317-337: constant plane -> 128, else
 *      trunc(clip(1 + (v - min)/(max - min)*254, 1, 255)) in the handle precision);
 *   4. mean and standard deviation of those pixels, exact (integer sums), and
 *      sigma < 1 -> PALX_FR_STOP (turbo_main.py:2080-2087);
 *   5. pal_over_ens_kmax2, flow_eddy_metrics, energy_joules
 *      (turbo_main.py:2015-2073, 1947-1959; turbo_simulator.py:764-810);
 *   6. with PALX_FRAME_ADAPT and not stopped, the viscosity controller:
 *      _palin_filter_2nd and adapt_visc (turbo_main.py:2154-2214), which writes
 *      the new viscosity into the device scalars read by the next step;
 *   7. the display pixels: with PALX_FRAME_STRETCH the map of step 3 stretched
 *      to [mu - k sigma, mu + k sigma] -> [0,255] (turbo_main.py:2109-2113,
 *      float32, round half to even), else the step-3 pixels; then the integer
 *      scaling of _upscale_downscale_u8 (turbo_main.py:875-890): scale > 1 keeps
 *      every scale-th row and column (ceil(M/scale) pixels a side), scale < -1
 *      repeats each pixel -scale times, otherwise 1:1.
 * out is a device buffer of out_bytes >= palx_frame_dims(...) bytes (smaller:
 * PALX_EINVAL): PALX_FRAME_NREC doubles (the record below), then the display
 * pixels, row-major (z rows, x columns).
 * ctl may be null without PALX_FRAME_ADAPT (defaults of turbo_main.py:391,
 * 474-478, 2176-2181).  Controller state (filters, last error, Re) lives on the
 * device; palx_frame_reset clears it and sets Re (on_visc_adapt_toggled,
 * turbo_main.py:1698-1705).
 * Calls on one handle run in stream order; the record and pixels of a call are
 * valid once the handle stream has reached its end. */
enum { PALX_FRAME_NEXTDT = 1, PALX_FRAME_STRETCH = 2, PALX_FRAME_ADAPT = 4 };
enum {
  PALX_FR_T = 0, PALX_FR_DT = 1, PALX_FR_CN = 2, PALX_FR_VISC = 3, PALX_FR_RE = 4, PALX_FR_IT = 5,
  PALX_FR_MU = 6, PALX_FR_SIGMA = 7, PALX_FR_STOP = 8,
  PALX_FR_RATIO = 9,    /* pal_over_ens_kmax2 */
  PALX_FR_FILT = 10,    /* low-passed 10000 * ratio of the controller */
  PALX_FR_ADAPTED = 11, /* 1 when the controller changed the viscosity */
  PALX_FR_U = 12, PALX_FR_L = 13, PALX_FR_TAU_L = 14, PALX_FR_EJ = 15,
  PALX_FR_ENSTROPHY = 16, PALX_FR_PALINSTROPHY = 17, PALX_FR_KMAX2 = 18, PALX_FR_ENERGY = 19,
  PALX_FR_FMIN = 20, PALX_FR_FMAX = 21, PALX_FR_WIDTH = 22, PALX_FR_HEIGHT = 23, PALX_FR_CFLM = 24,
  PALX_FRAME_NREC = 32
};
typedef struct palx_frame_ctl {
  double k_std;     /* display stretch half-width in sigmas, 2.5 */
  double target;    /* set point of 10000 * ratio, 20 */
  double kp, ki, kd;/* 0.25, 0, 25 */
  double deadband;  /* relative error band with no update, 0.001 */
  double max_frac;  /* clamp of the log-space update, 0.01 */
  double palin_tau; /* filter time constant in frames, 5 */
  double re0;       /* Re_from_N_K0(N, K0) (turbo_main.py:2323-2328) */
  double re_factor; /* Re stays within [re0 / f, re0 * f], 1000 */
} palx_frame_ctl;
int palx_frame_dims(const palx_t* h, int32_t scale, int32_t* width, int32_t* height, size_t* bytes);
int palx_frame_reset(palx_t* h, double re);
int palx_frame(palx_t* h, int32_t var, int32_t scale, int32_t flags, const palx_frame_ctl* ctl, void* out,
               size_t out_bytes);

#ifdef __cplusplus
}
#endif
#endif /* PALX_H */
