/* fastwigner.h -- C ABI of libfastwigner.so, the MI355X (gfx950) engine of the
 * split-operator phase-space propagators.
 *
 * The reference crosses into native code in two ways, both replaced here:
 *   1. runtime-compiled CUDA kernels with the user's V/K/mask strings spliced in
 *      (pycuda SourceModule, wigner_moyal_cuda_1d.py:244-277, templates :564-731),
 *   2. the ctypes FFI over cuFFT (cufft.py:16-343: cufftPlanMany, cufftExec*, cufftDestroy).
 * Here one handle owns the hipRTC-compiled module (same spliced strings), the
 * device buffers, the stream and the captured graphs; the FFTs are the library's
 * own LDS Stockham kernels fused with the phase multiplies.
 *
 * Conventions
 *   - Every call returns FW_OK (0) or a negative status; fw_last_error() gives the
 *     message of the last failure on the calling thread (cf. the typed cuFFT
 *     exceptions of cufft.py:31-102).
 *   - Host pointers are borrowed for the duration of the call only.
 *   - A handle is not re-entrant; distinct handles are independent.
 *   - Arrays are row-major: the Wigner function is (p_grid_dim, x_grid_dim) real
 *     with x contiguous (wigner_moyal_cuda_1d.py:302); the density matrix is
 *     (N, N) complex with rho[i][j] = <x_j|rho|x'_i> (rho_vneumann_cuda_1d.py:222).
 */
#ifndef FASTWIGNER_H
#define FASTWIGNER_H

#ifdef __cplusplus
extern "C" {
#endif

#define FW_OK 0
#define FW_ERR_INVALID_ARG (-1) /* bad sizes, null handle, unsupported grid */
#define FW_ERR_COMPILE (-2)     /* hipRTC rejected the generated source (bad V/K/mask string) */
#define FW_ERR_HIP (-3)         /* a HIP runtime call failed */
#define FW_ERR_NO_DEVICE (-4)   /* no usable GPU */
#define FW_ERR_UNSUPPORTED (-5) /* configuration not implemented */

#define FW_NSER 9 /* per-step sums: S0 Sx Sp Sxx Spp SdK SdV SH SWW */

#define FW_FP64 0
#define FW_FP32 1 /* W and FFTs in FP32; phase argument, its mod-2pi reduction and all sums in FP64 */

#define FW_BLOCH_CUDA 0 /* exp(-db/4 [V+V-Vmin]), exp(-db/2 [K+K-Kmin]) (wigner_bloch_cuda_1d.py:299-340) */
#define FW_BLOCH_FFTW 1 /* exp(-db/2 [..]) * mask, exp(-db [..]) (wigner_bloch_fftw_1d.py:74-97) */

/* ------------------------------------------------------------------------ */
/* Library                                                                   */
/* ------------------------------------------------------------------------ */

/* Library version (major*10000 + minor*100 + patch). */
int fw_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* fw_last_error(void);

/* Number of visible GPUs (0 on a machine without one; never an error). */
int fw_device_count(int* count);

/* ------------------------------------------------------------------------ */
/* Wigner-function propagators (Moyal, Bloch, random collision)             */
/* replaces WignerMoyalCUDA1D (wigner_moyal_cuda_1d.py:10-731),             */
/* WignerBlochCUDA1D (wigner_bloch_cuda_1d.py:8-445) and the random-        */
/* collision step of wigner_random_collision_model_fftw_1d.py:56-76          */
/* ------------------------------------------------------------------------ */

typedef struct fw_wigner fw_wigner;

typedef struct fw_wigner_desc {
  int x_grid_dim;          /* X_gridDIM: even, 8..4096: 2^a m, m odd with the product of its distinct
                              primes at most 15 (2^a 3^b 5^c, 2^a 7^b, 2^a 11^b, 2^a 13^b) */
  int p_grid_dim;          /* P_gridDIM: as X_gridDIM */
  double x_amplitude;      /* X_amplitude: x in [-X_amplitude, X_amplitude) */
  double p_amplitude;      /* P_amplitude */
  double dt;               /* time step */
  double phase_time_shift; /* V, K evaluated at t_{n+1} + shift*dt; 0.5 = the CUDA kernels (wigner_moyal_cuda_1d.py:602,643) */
  int precision;           /* FW_FP64 or FW_FP32 */
  int device;              /* HIP device ordinal */
  int col_e, col_p;        /* V pass: elements per thread, column pairs per block (0 = default) */
  int row_e, row_p;        /* K pass: elements per thread, row pairs per block (0 = default) */
  int phase_tables;        /* cache the V/K factors in tables (as the FFTW class caches _expV/_expK,
                              wigner_moyal_fftw_1d.py:164-213): 0 = when the V (K) string and its
                              mask do not contain the token t (the FP64 K passes then generate
                              linear-in-Lambda K factors instead of reading the table, see
                              fw_wigner_k_factors), 1 = always, -1 = never */
  int step_passes;         /* Moyal steps without the dissipator: 0 or 2 = two passes per step (the
                              closing V/2 of a step fused with the opening V/2 of the next), 3 = three */
  int v_table;             /* 0 = the V factors follow phase_tables; -1 = evaluate them in the column
                              pass even when they are time independent (no table read: faster where
                              the pass is bound by the table's traffic, e.g. 2048^2 FP64) */
} fw_wigner_desc;

typedef struct fw_wigner_exprs {
  const char* consts;                /* C declarations (constants, __device__ functions), verbatim; may be NULL */
  const char* V;                     /* V(X, t) */
  const char* K;                     /* K(P, t) */
  const char* diff_V;                /* dV/dX (X, t) for the Ehrenfest sums; NULL = "0." */
  const char* diff_K;                /* dK/dP (P, t); NULL = "0." */
  const char* abs_boundary_x_theta;  /* mask of (X, Theta, t) in the V factor; NULL = "1." */
  const char* abs_boundary_lambda_p; /* mask of (Lambda, P, t) in the K factor; NULL = "1." */
  const char* abs_boundary_x_p;      /* mask of (X, P, t) applied after each step; NULL = none */
} fw_wigner_exprs;

/* The grid constants spliced into the kernels: dX, dP, dLambda, dTheta
 * (wigner_moyal_cuda_1d.py:113-135), computed as numpy does. */
int fw_wigner_grid_constants(const fw_wigner_desc* desc, double out[4]);

/* The generated prelude of the kernel source (configuration macros, constants and the
 * spliced expressions), without compiling: *len = its size + 1; up to cap - 1 characters
 * and a terminating 0 are copied to out (out may be NULL). Replaces reading the source the
 * reference assembles for SourceModule (wigner_moyal_cuda_1d.py:244-277). */
int fw_wigner_prelude(const fw_wigner_desc* desc, const fw_wigner_exprs* exprs, char* out, size_t cap,
                      size_t* len);
/* Generate and compile the kernel source with hipRTC without touching a device.
 * On FW_ERR_COMPILE the compiler log is in fw_last_error(). */
int fw_wigner_check(const fw_wigner_desc* desc, const fw_wigner_exprs* exprs);

/* Compile (cached per source), allocate W and scratch, upload twiddles. */
int fw_wigner_create(const fw_wigner_desc* desc, const fw_wigner_exprs* exprs, fw_wigner** out);
int fw_wigner_destroy(fw_wigner* h);

/* W <- host array (p_grid_dim*x_grid_dim doubles), no normalisation. */
int fw_wigner_set_state(fw_wigner* h, const double* host_w);
/* W <- expr(X, P) evaluated on the grid (init_wigner, wigner_moyal_cuda_1d.py:666-685); i, j, X, P in scope. */
int fw_wigner_set_state_expr(fw_wigner* h, const char* expr);
/* W <- value everywhere (set_wignerfunction(float), wigner_moyal_cuda_1d.py:485-487). */
int fw_wigner_fill(fw_wigner* h, double value);
/* W /= sum(W) dXdP (wigner_moyal_cuda_1d.py:491-492); returns sum(W) dXdP before scaling. */
int fw_wigner_normalize(fw_wigner* h, double* norm_before);
/* Copy W to a host array of doubles. */
int fw_wigner_get_state(fw_wigner* h, double* host_w);
/* Device address of W and its element size (8 or 4 bytes). */
int fw_wigner_state_ptr(fw_wigner* h, void** dev_ptr, int* elem_size);

/* Advance `steps` Moyal steps (wigner_moyal_cuda_1d.py:521-540). times[s] is t
 * after the s-th increment (host-accumulated, as the reference's t += dt).
 * After each step W is renormalised and the FW_NSER sums over the step's
 * unnormalised W are written to series[s*FW_NSER..] when series != NULL.
 * No host synchronisation inside; steps are replayed from a captured hipGraph. */
int fw_wigner_propagate(fw_wigner* h, const double* times, int steps, double* series);

/* One Moyal step at time t (single_step_propagation, wigner_moyal_cuda_1d.py:496-519;
 * with the dissipator, wigner_random_collision_model_fftw_1d.py:56-75): W is left
 * unnormalised, as the reference leaves it; a later fw_wigner_propagate takes it as
 * it is. sums (optional) gets the FW_NSER sums over the new W. */
int fw_wigner_single_step(fw_wigner* h, double t, double* sums);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): fw_wigner_propagate's two-pass Moyal steps on the
 * caller's device buffer w_dev (p_grid_dim*x_grid_dim elements of the handle's
 * precision, updated in place; the handle's own W is untouched), with the
 * run's scalars, step times, step sums and per-block partial sums in one
 * file-scope static device array of 81 KiB per GPU instead of handle buffers.
 * No hipMalloc, hipFree or other allocating call: the factor tables are built
 * by fw_wigner_create, the sums come back through the handle's pinned staging.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (two host threads, two handles) race. Errors: FW_ERR_INVALID_ARG for a handle
 * with the dissipator or step_passes = 3, or tiles needing more than 1024
 * blocks per pass (the partial sums the array holds). times and series as in
 * fw_wigner_propagate. */
int fw_wigner_propagate_noalloc(fw_wigner* h, void* w_dev, const double* times, int steps, double* series);

/* The FW_NSER sums over the current W at time t (get_average building block). */
int fw_wigner_moments(fw_wigner* h, double t, double sums[FW_NSER]);

/* sum(W * expr(X, P, t)) dXdP (get_average, wigner_moyal_cuda_1d.py:378-396),
 * compiled once per expression string and cached on the handle. */
int fw_wigner_get_average(fw_wigner* h, const char* expr, double t, double* out);

/* Imaginary-time (Bloch) setup: compute V_min and K_min over the theta-x and
 * p-lambda grids at t_initial (wigner_bloch_cuda_1d.py:72-102). */
int fw_wigner_bloch_setup(fw_wigner* h, int convention, double t_initial, double mins_out[2]);

/* `steps` Bloch steps of size dbeta with renormalisation each step
 * (wigner_bloch_cuda_1d.py:229-254). If purity_guard > 0 the run stops after the
 * first step whose purity is not < purity_guard (wigner_bloch_cuda_1d.py:149-163);
 * *steps_done reports how many steps were applied. */
int fw_wigner_bloch_propagate(fw_wigner* h, double dbeta, int steps, double purity_guard, double* series,
                              int* steps_done);

/* Random-collision dissipator W <- w W + W0s around every step
 * (wigner_random_collision_model_fftw_1d.py:63-74). host_w0s = (1-w) W_Gibbs;
 * NULL disables it. */
int fw_wigner_set_dissipator(fw_wigner* h, const double* host_w0s, double w_coeff);

/* Snapshot / restore of W for the adaptive roll-back of get_ground_state
 * (wigner_bloch_cuda_1d.py:186,209,218): which = 0 save, 1 restore. */
int fw_wigner_snapshot(fw_wigner* h, int which);

/* Mean device time of the step kernels over the last propagate call, in
 * microseconds, measured with HIP events on the handle's stream when timing is
 * enabled: (v_first, k, v_last) for three-pass steps, (vv, k_norm, 0) for
 * two-pass steps (fw_wigner_step_passes tells which ran). */
int fw_wigner_set_timing(fw_wigner* h, int enable);
int fw_wigner_kernel_times(fw_wigner* h, double us_out[3]);
/* Mean device time (us) per launch of the two step kernels of two-pass steps, fw_vv and
 * fw_k_norm, in the propagator's own sequence from time t: the step time from HIP events
 * around max(1, reps/64) replays of the 64-step graph on the handle's stream, split between
 * the two kernels by their shares in one further chunk with an event around every kernel.
 * W is restored afterwards. No counterpart in the reference (measurement). */
int fw_wigner_kernel_bench(fw_wigner* h, double t, int reps, double us_out[2]);
/* Passes per step of the last Moyal propagate call (2 or 3). */
int fw_wigner_step_passes(fw_wigner* h, int* passes);
/* How the Moyal K passes obtain their factors exp(-i dt (K(P + L/2) - K(P - L/2))) (the
 * expK kernels, wigner_moyal_cuda_1d.py:592-622): *generated = 1 when they form them as
 * powers z^k of one phase per row (FP64, phase_tables 0, power-of-two rows of at least
 * four elements per thread, and K(P + L/2) - K(P - L/2) linear in L as for P^2/2: the
 * powers match the table to 1e-11 on every grid point, checked at creation), 0 when they
 * read the table or evaluate each factor; *max_dev = that check's largest
 * |generated - table| (-1 if not run). No counterpart in the reference (introspection). */
int fw_wigner_k_factors(fw_wigner* h, int* generated, double* max_dev);

/* ------------------------------------------------------------------------ */
/* Density-matrix propagators (von Neumann, Bloch)                           */
/* replaces RhoVNeumannCUDA1D (rho_vneumann_cuda_1d.py:15-858) and           */
/* RhoBlochCUDA1D (rho_bloch_cuda_1d.py:9-366)                              */
/* ------------------------------------------------------------------------ */

#define FW_RSER 20 /* per-step sums of the density-matrix propagators:
   [0,1] Re/Im Tr rho, [2,3] sum rho_ii x_i, [4,5] sum rho_ii x_i^2, [6,7] sum rho_ii V'(x_i),
   [8,9] sum rho_ii V(x_i), [10] sum |rho|^2, [11,12] Tr(rho P), [13,14] Tr(rho P^2),
   [15,16] Tr(rho K'(P)), [17,18] Tr(rho K(P)) -- over the unnormalised rho of the step */

typedef struct fw_rho fw_rho;

typedef struct fw_rho_desc {
  int x_grid_dim;          /* X_gridDIM: even, 8..4096, as fw_wigner_desc's (2^a m, distinct primes of m
                              with a product of at most 15) */
  int precision;           /* FW_FP64 (complex128) or FW_FP32 (complex64, FP64 phases and sums) */
  double x_amplitude;      /* X_amplitude */
  double dt;
  double phase_time_shift; /* V, K at t_{n+1} + shift*dt; 0 = the reference kernels (rho_vneumann_cuda_1d.py:660,681) */
  int device;
  int row_e, row_p;        /* row passes: elements per thread, rows per block (0 = default) */
  int col_e, col_p;        /* column pass: elements per thread, columns per block (0 = default) */
} fw_rho_desc;

typedef struct fw_rho_exprs {
  const char* consts;         /* C declarations, verbatim (may be NULL) */
  const char* V;              /* V(X, t) */
  const char* K;              /* K(P, t) */
  const char* diff_V;         /* NULL = "0." */
  const char* diff_K;         /* NULL = "0." */
  const char* abs_boundary_x; /* mask of (X, X_prime, t) in the V factor (rho_vneumann_cuda_1d.py:683); NULL = none */
  const char* abs_boundary_p; /* mask of (P, P_prime, t) in the K factor (:662); NULL = none */
} fw_rho_exprs;

/* dX and dP (P = fftfreq(N, dX/2pi), rho_vneumann_cuda_1d.py:94-101) */
int fw_rho_grid_constants(const fw_rho_desc* desc, double out[2]);
int fw_rho_check(const fw_rho_desc* desc, const fw_rho_exprs* exprs);
int fw_rho_create(const fw_rho_desc* desc, const fw_rho_exprs* exprs, fw_rho** out);
int fw_rho_destroy(fw_rho* h);
/* rho <- host array of N*N complex values stored as (re, im) pairs */
int fw_rho_set_state(fw_rho* h, const double* host_rho);
/* rho <- expr(X, X_prime, i, j) (init_rho, rho_vneumann_cuda_1d.py:687-708); cuda_complex is available */
int fw_rho_set_state_expr(fw_rho* h, const char* expr);
/* rho /= Tr(rho) dX (complex, rho_vneumann_cuda_1d.py:336); trace_out = Tr(rho) before */
int fw_rho_normalize(fw_rho* h, double trace_out[2]);
int fw_rho_get_state(fw_rho* h, double* host_rho);
int fw_rho_state_ptr(fw_rho* h, void** dev_ptr, int* elem_size);
/* `steps` von Neumann steps (rho_vneumann_cuda_1d.py:355-374): t = times[s], rho /= Tr(rho) dX after
 * each; series (optional) gets FW_RSER sums per step; momentum_sums != 0 adds the Tr(rho g(P)) family
 * (one extra read of rho per step). */
int fw_rho_propagate(fw_rho* h, const double* times, int steps, int momentum_sums, double* series);
/* One von Neumann step at time t (single_step_propagation, rho_vneumann_cuda_1d.py:340-353): rho is
 * left unnormalised, as the reference leaves it. series (optional) gets the trace family (the first
 * 11 of FW_RSER sums) over the new rho. */
int fw_rho_single_step(fw_rho* h, double t, double* series);
/* FW_RSER sums over the current rho at time t (the momentum family included) */
int fw_rho_moments(fw_rho* h, double t, double sums[FW_RSER]);
/* Tr[rho F1(x) g1(p) F2(x) ...] dX, real part (get_average, rho_vneumann_cuda_1d.py:431-472); obs[k] NULL
 * skips a slot; each string is compiled once */
int fw_rho_get_average(fw_rho* h, const char* const* obs, int nobs, double t, double* out);
/* Bloch: V_min = min of V(X)+V(X') at t_initial, K_min = 0 as in the reference (rho_bloch_cuda_1d.py:354-365) */
int fw_rho_bloch_setup(fw_rho* h, double t_initial, double mins_out[2]);
int fw_rho_bloch_propagate(fw_rho* h, double dbeta, int steps, double purity_guard, double* series,
                           int* steps_done);
int fw_rho_snapshot(fw_rho* h, int which);
/* Wigner function of rho on the (X/sqrt2, P/sqrt2) grid (get_wignerfunction, rho_vneumann_cuda_1d.py:502-592),
 * complex N*N (re, im) pairs, rows in fftshift(P) order; kept on the device for fw_rho_wigner_time */
int fw_rho_wigner(fw_rho* h, double* host_w);
/* W_init of the Wigner time (the reference's wigner_initial attribute, rho_vneumann_cuda_1d.py:409-412):
 * N*N doubles on the Wigner grid, uploaded once and kept on the device until replaced. */
int fw_rho_set_wigner_initial(fw_rho* h, const double* host_w_init);
/* get_wigner_time (rho_vneumann_cuda_1d.py:594-620,830-858): the integral over the Wigner grid of
 * 0.5(1 + signbit(H(t) + Ip)) (Re W - W_init) dxdp, with the reference's indicator, for the Wigner
 * function last computed (fw_rho_wigner) and the W_init set by fw_rho_set_wigner_initial. */
int fw_rho_wigner_time(fw_rho* h, double t, double Ip, double* out);
/* fw_rho_propagate with the reference's per-step Wigner time (get_Ehrenfest with wigner_initial and
 * Ip, rho_vneumann_cuda_1d.py:408-412): after each step the Wigner function (three shears and the
 * sign-flipped FFT) and its Wigner time, on the device, into wigner_time[s]; afterwards the
 * normalised Wigner function of the last step is the one fw_rho_wigner would give. */
int fw_rho_propagate_wigner_time(fw_rho* h, const double* times, int steps, int momentum_sums, double Ip,
                                 double* series, double* wigner_time);
int fw_rho_set_timing(fw_rho* h, int enable);
int fw_rho_kernel_times(fw_rho* h, double us_out[3]);


/* ------------------------------------------------------------------------ */
/* Batched propagator: many independent Wigner functions per launch          */
/* No counterpart in the reference, which runs one grid per Python object    */
/* (WignerMoyalCUDA1D, wigner_moyal_cuda_1d.py:10-731; the random-collision  */
/* step wigner_random_collision_model_fftw_1d.py:56-76; the Bloch thermal    */
/* state wigner_bloch_fftw_1d.py:27-109): each instance here is that         */
/* propagator on its own square grid with its own parameters.                */
/* ------------------------------------------------------------------------ */

typedef struct fw_batch fw_batch;

#define FW_BATCH_NPRM 8

typedef struct fw_batch_desc {
  int n_instances;
  int precision;           /* FW_FP32 (W and FFTs in FP32; phases, sums in FP64) or FW_FP64 */
  int device;
  int mode;                /* 0: one persistent launch per step; 1: one launch per pass */
  double dt;               /* common time step */
  double phase_time_shift; /* V, K at t_{n+1} + shift*dt */
  int e_for_n[5];          /* FFT elements per thread for N = 64, 128, 256, 512, 1024 (0 = default) */
  int blocks_per_cu;       /* mode 0: persistent blocks per CU (0 = occupancy limit) */
} fw_batch_desc;

typedef struct fw_batch_instance {
  int grid_dim;                  /* N: the instance's grid is N x N, N in {64, 128, 256, 512, 1024} */
  double x_amplitude, p_amplitude;
  double params[FW_BATCH_NPRM];  /* values of the parameter names in the expressions */
} fw_batch_instance;

typedef struct fw_batch_exprs {
  const char* consts;      /* C declarations, verbatim (may be NULL) */
  const char* param_names; /* comma-separated names bound to params[0..7], e.g. "a2,a3,a4" */
  const char* V;           /* V(X, t) in terms of X, t, dt, the constants and the parameters */
  const char* K;           /* K(P, t) */
  const char* diff_V;      /* NULL = "0." */
  const char* diff_K;      /* NULL = "0." */
} fw_batch_exprs;

/* Compile the batch's kernels for the target without a device. */
int fw_batch_check(const fw_batch_desc* desc, const fw_batch_instance* inst, const fw_batch_exprs* exprs);
int fw_batch_create(const fw_batch_desc* desc, const fw_batch_instance* inst, const fw_batch_exprs* exprs,
                    fw_batch** out);
int fw_batch_destroy(fw_batch* h);
/* W_i <- host array of N_i*N_i doubles (row-major, x contiguous), no normalisation. */
int fw_batch_set_state(fw_batch* h, int i, const double* host_w);
int fw_batch_get_state(fw_batch* h, int i, double* host_w);
/* Random-collision source of instance i: W <- w W + host_w0s around every step
 * (host_w0s = (1-w) W_Gibbs, wigner_random_collision_model_fftw_1d.py:63-74); NULL disables. */
int fw_batch_set_dissipator(fw_batch* h, int i, const double* host_w0s, double w_coeff);
/* Every instance with w_coeff[i] < 1 takes W0s = (1 - w_coeff[i]) W_i from its current state
 * (the Gibbs states left by fw_batch_bloch_propagate); w_coeff[i] = 1 disables it. */
int fw_batch_gibbs_dissipators(fw_batch* h, const double* w_coeff);
/* Imaginary-time steps of every instance from its current state: instance i takes steps[i]
 * steps of dbeta[i] (exponents of `convention`, FW_BLOCH_CUDA or FW_BLOCH_FFTW, potentials at
 * t_initial, wigner_bloch_fftw_1d.py:27-109) with renormalisation; W is normalised at the end. */
int fw_batch_bloch_propagate(fw_batch* h, int convention, double t_initial, const double* dbeta,
                             const int* steps);
/* `steps` real-time steps of every instance; times[s] = t after the s-th increment. After each
 * step each W_i is renormalised; series (optional) gets steps * n_instances * FW_NSER sums over
 * the unnormalised W_i of each step. One graph replay per 64 steps. */
int fw_batch_propagate(fw_batch* h, const double* times, int steps, double* series);
/* Launches per real-time step: 1 (mode 0, persistent), 3 (mode 1), or 2 (mode 1 when V does not
 * depend on t: the K pass with the previous step's deferred norm and dissipator, then one column
 * pass with the closing V/2, dissipator and sums of a step and the opening V/2 of the next). No
 * counterpart in the reference (the batch is new). */
int fw_batch_passes(fw_batch* h, int* passes_per_step);
/* total grid points, tasks per step, persistent grid size */
int fw_batch_info(fw_batch* h, long long* points, int* tasks_per_step, int* grid);
/* Device time of one step (us), averaged over the floor(steps/64)*64 steps replayed from the
 * graph between HIP events on the handle's stream (steps >= 64). The state advances by that
 * many steps, with V and K evaluated at times (s+1)*dt of each 64-step chunk. */
int fw_batch_step_time(fw_batch* h, int steps, double* us_per_step);

/* Device address of instance i's W (N*N elements of elem_size bytes, row-major), valid until
 * fw_batch_destroy; the caller may read it between calls (e.g. to copy training pairs). */
int fw_batch_state_ptr(fw_batch* h, int i, void** dev_ptr, int* elem_size);

/* Every instance's W of src copied into dst on the device (both batches must have the same
 * instance grids and precision; dst's dissipators and series are left as they are). No
 * counterpart in the reference: it lets a second batch take a step of different length from
 * the first one's states (the coarse base of the surrogate's training pairs). */
int fw_batch_copy_states(fw_batch* dst, fw_batch* src);

/* ------------------------------------------------------------------------ */
/* Fourier-neural-operator surrogate of k propagator steps                   */
/* No counterpart in the reference. Its network is trained on trajectories  */
/* of the batched propagator and checked by the plain propagator (the        */
/* Moyal / random-collision step of wigner_moyal_cuda_1d.py:496-540 and     */
/* wigner_random_collision_model_fftw_1d.py:56-76).                          */
/* ------------------------------------------------------------------------ */

typedef struct fw_fno fw_fno;

typedef struct fw_fno_desc {
  int grid_dim;           /* N: inputs are N x N Wigner functions, N = 64, 128 or 256 */
  int width;              /* channels C: 32 (the fused layer kernels are written for it) */
  int modes;              /* retained modes M per axis, 8 or 16, 4 M <= N (ky < M; kp < M or kp >= N - M) */
  int layers;             /* Fourier layers, 1 .. 16 */
  int max_batch;          /* samples per forward / training step, 1 .. 64 */
  int device;
  double x_amplitude, p_amplitude;
} fw_fno_desc;

/* Per-sample features after W: a2, a3, a4, gamma, kT (5 floats). Parameters are one FP32
 * vector: Pw[8][C], pb[C], per layer {R[2M*M][C][C][re,im], Wp[C][C], b[C]}, Q[C], qb. */
int fw_fno_create(const fw_fno_desc* desc, fw_fno** out);
int fw_fno_destroy(fw_fno* h);
int fw_fno_param_count(fw_fno* h, long long* n);
int fw_fno_set_params(fw_fno* h, const float* host_params); /* also resets the Adam state */
int fw_fno_get_params(fw_fno* h, float* host_params);
/* Output of the network: out = W + scale c with c = v_L Q + qb, or, with zero_mass != 0,
 * out = W + scale (c - mean c) per sample, so that out has the mass (norm) of W. Defaults:
 * scale 1, zero_mass 0. The scale lets a network learning a small correction keep O(1)
 * projection weights. No counterpart in the reference (the surrogate is new). */
int fw_fno_set_output(fw_fno* h, float scale, int zero_mass);
/* prediction W(t + k dt) for `batch` samples (host arrays: batch*N*N inputs, batch*5 features) */
int fw_fno_forward(fw_fno* h, const float* host_in, const float* host_feat, int batch, float* host_out);
/* Moments of the last fw_fno_forward's `batch` outputs for the physical referee, computed on the
 * device in FP64: per sample 3N doubles, [x marginal (sum over p) | p marginal (sum over x) |
 * per-row sum of W^2]. No counterpart in the reference. */
int fw_fno_moments(fw_fno* h, int batch, double* host_mom);
/* loss L = mean_b |pred_b - target_b|^2 / |target_b|^2 and, if host_grad != NULL, dL/dparams */
int fw_fno_loss_grad(fw_fno* h, const float* host_in, const float* host_feat, const float* host_target, int batch,
                     double* loss, float* host_grad);
/* device-resident training set of (input, target, features) pairs; fw_fno_dataset_put copies N*N
 * floats from device addresses (either may be NULL to leave that half of pair k unchanged) */
int fw_fno_dataset_reserve(fw_fno* h, int pairs);
int fw_fno_dataset_put(fw_fno* h, int k, const void* dev_in, const void* dev_target, const float feat[5]);
/* one Adam step (lr; betas 0.9, 0.999) on the pairs host_idx[0..batch) */
int fw_fno_train_step(fw_fno* h, const int* host_idx, int batch, float lr, double* loss);
/* relative L2 error of the prediction for each listed pair */
int fw_fno_eval_pairs(fw_fno* h, const int* host_idx, int n, double* rel_l2_out);
/* device time of one forward pass over `batch` samples (us), averaged over reps */
int fw_fno_infer_time(fw_fno* h, int batch, int reps, double* us_per_batch);

#ifdef __cplusplus
}
#endif
#endif /* FASTWIGNER_H */
