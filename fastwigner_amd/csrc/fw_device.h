// fw_device.h -- CDNA4 (gfx950) device code of the split-operator propagators.
//
// This header is embedded in libfastwigner.so and compiled by hipRTC together
// with a generated prelude that holds, in this order:
//   * configuration macros  FW_REAL (double|float), FW_NX, FW_NP, FW_COL_E,
//     FW_COL_P, FW_ROW_E, FW_ROW_P, FW_PHASE_SHIFT
//   * the grid constants the reference splices into its kernels
//     (X_gridDIM, P_gridDIM, X_amplitude, P_amplitude, dt, dX, dP, dLambda,
//     dTheta; wigner_moyal_cuda_1d.py:151-166) and the caller's own constants
//     and functions, verbatim
//   * the spliced expressions as device functions:
//       double V(double X, double t), double K(double P, double t)
//       double fw_diff_V(double X, double t), double fw_diff_K(double P, double t)
//       double fw_mask_x_theta(double X, double Theta, double t)
//       double fw_mask_lambda_p(double Lambda, double P, double t)
//       double fw_mask_x_p(double X, double P, double t)
//
// State layout (identical to the reference, wigner_moyal_cuda_1d.py:302):
// W is real, row-major, shape (P_gridDIM, X_gridDIM), x contiguous.
//
// One Strang step V/2-K-V/2 (wigner_moyal_cuda_1d.py:496-519) is three fused
// passes, each a single read and a single write of W:
//   V pass (columns, along p): r2c FFT -> x exp(-i dt/2 [V(x-th/2)-V(x+th/2)]) -> c2r
//   K pass (rows, along x):    r2c FFT -> x exp(-i dt [K(p+l/2)-K(p-l/2)])     -> c2r
//   V pass again, whose epilogue also applies the x-p mask, the dissipator and
//   reduces the nine per-step sums (norm, <x>, <p>, <x^2>, <p^2>, <K'>, <V'>,
//   <H>, sum W^2). The normalisation W /= sum(W) dXdP is deferred into the load
//   of the next step's first pass.
//
// FFT engine. Two real lines a, b are packed into one complex line z = a + i b,
// transformed by an LDS Stockham FFT of length N held E elements per thread
// (thread t of a line holds elements t + (N/E) e, e = 0..E-1, in registers),
// and separated in the spectrum: A_k = (Z_k + conj Z_{N-k})/2,
// B_k = (Z_k - conj Z_{N-k})/(2i). Each thread owns pairs (k, N-k), so every
// phase factor is evaluated once. DC and Nyquist keep only their real part times
// the real part of the factor, which is exactly the c2r rule of cuFFT and FFTW.
// The inverse transform is normalised by 1/N like pyFFTW.
#pragma once

#include "fw_common.h"

namespace fwk {

// sums of the Wigner propagators: S0 Sx Sp Sxx Spp SdK SdV SH SWW

// ---------------------------------------------------------------------------
// Per-launch arguments (one struct, passed by value)
// ---------------------------------------------------------------------------
// scal[] slots. SC_FRESH = 1: W is normalised, the partials are stale (set by the
// host before a run, cleared by the first K pass). SC_OUT.. = 9 moments.
enum { SC_NORM = 0, SC_STOP = 1, SC_FRESH = 2, SC_OUT = 8 };

// Weights of the theta-spectrum sums of the fused column pass (vvpass_theta),
// entry k <= N_p/2, from F_k = sum_n f(P_n) exp(-2 pi i n k / N_p):
//   0 < k < N_p/2: a = Im 2 F_k(P) (= dP N cot(pi k / N)), b = Re 2 F_k(P^2),
//                  kk = 2 F_k(K), dk = 2 F_k(K') (re, im);
//   k = 0, N_p/2:  a = F_k(P), b = F_k(P^2), kk = (F_k(K), 0), dk = (F_k(K'), 0).
// For P_n = dP (n - N/2) and 0 < k < N/2, Re 2 F_k(P) = -dP N and Im F_k(P^2) = 0
// exactly (F_k(P) = dP N (-1/2 + i cot(pi k/N)/2), F_k(P^2) = dP^2 N / (2 sin^2(pi k/N))),
// so these two are not stored.
struct __attribute__((aligned(16))) SumW {
  double a, b;
  double kk[2], dk[2];
};

struct Args {
  real_t* W;                   // state
  const real_t* W0s;           // (1-w) * Gibbs state for the dissipator, or null
  const tw_t* tw_col;          // exp(-2 pi i j / N_p), j < N_p
  const tw_t* tw_row;          // exp(-2 pi i j / N_x), j < N_x
  const double* times;         // t after the increment, per step
  double* scal;                // device scalars (norm factor, stop flag, moments)
  double* partials;            // per-block partial sums
  unsigned int* counter;       // last-block ticket
  double* series;              // NSER sums per step
  const cpx* vtab;             // cached V factors, (N_p/2+1) x N_x, or null
  const cpx* ktab;             // cached K factors, N_p x (N_x/2+1), or null
  const SumW* sw;              // theta-spectrum sum weights, N_p/2+1, or null
  int step;
  int flags;
  int nbp;                     // number of per-block partials of the last V pass
  int pad_;
  double dbeta, cv, ck, vmin, kmin, t_initial;
  double w_coeff;              // exp(-gamma dt/2), random-collision dissipator
  double guard;                // stop when purity >= guard (Bloch), 0 = off
};

enum {
  F_FIRST = 1,       // first pass of a step: apply deferred norm (+ dissipator) on load
  F_LAST = 2,        // last pass: x-p mask (+ dissipator), sums, deferred norm
  F_BLOCH = 4,       // imaginary-time (real) factors
  F_DISSIP = 8,      // random-collision dissipator
  F_MASK_XP = 16,    // x-p mask present
  F_BLOCH_MASK = 32, // Bloch V factor carries the x-theta mask (FFTW convention)
  F_VTAB = 64,       // V factors from the cached table
  F_KTAB = 128,      // K factors from the cached table
  F_FROM_PARTIALS = 256,  // apply_norm: norm from the last V pass's partials (end of a run)
  F_KNORM = 512,     // two-pass steps: the K pass applies the previous step's deferred norm
  F_NO_SCALE = 1024, // apply_norm: publish the sums only, W stays unnormalised (one bare step)
  F_KGEN = 2048,     // K passes: factors generated as powers (kgen_factors) instead of the table
};

// two adjacent factors (columns c, c+1) in one load
struct __attribute__((aligned(4 * sizeof(real_t)))) cpx2 {
  cpx a, b;
};

// ---------------------------------------------------------------------------
// Phase factors
// ---------------------------------------------------------------------------
// expV_bulk / expV_boundary, wigner_moyal_cuda_1d.py:631-663
template <bool EXACT = false>
FWD cpx vfactor(double X, double Theta, double tau, double tnow) {
  const double ph = -0.5 * ::dt * (::V(X - 0.5 * Theta, tau) - ::V(X + 0.5 * Theta, tau));
  const double m = ::fw_mask_x_theta(X, Theta, tnow);
  return EXACT ? unit_phase_exact(ph, m) : unit_phase(ph, m);
}
// The V of the cached factor table: V itself, or its time-free part when V
// splits as V0(X) + X F(t) (FW_V_SPLIT, set by the host), whose field factor
// exp(+i dt/2 Theta F(tau)) the passes apply per step (v_field_phase).
#ifndef FW_V_SPLIT
#define FW_V_SPLIT 0
#endif
#ifndef FW_DV_SPLIT
#define FW_DV_SPLIT 0
#endif
FWD double v_tab_fn(double X, double t) {
#if FW_V_SPLIT
  return ::fw_V_static(X, t);
#else
  return ::V(X, t);
#endif
}
template <bool EXACT = false>
FWD cpx vfactor_tab(double X, double Theta, double tau, double tnow) {
  const double ph = -0.5 * ::dt * (v_tab_fn(X - 0.5 * Theta, tau) - v_tab_fn(X + 0.5 * Theta, tau));
  const double m = ::fw_mask_x_theta(X, Theta, tnow);
  return EXACT ? unit_phase_exact(ph, m) : unit_phase(ph, m);
}
// expK_bulk / expK_boundary, wigner_moyal_cuda_1d.py:592-622
template <bool EXACT = false>
FWD cpx kfactor(double Lambda, double P, double tau, double tnow) {
  const double ph = -::dt * (::K(P + 0.5 * Lambda, tau) - ::K(P - 0.5 * Lambda, tau));
  const double m = ::fw_mask_lambda_p(Lambda, P, tnow);
  return EXACT ? unit_phase_exact(ph, m) : unit_phase(ph, m);
}
// bloch_expV_bulk / _boundary, wigner_bloch_cuda_1d.py:329-359 (cv = 1/4);
// wigner_bloch_fftw_1d.py:65-82 (cv = 1/2, absorbing mask applied).
FWD cpx bloch_vfactor(const Args& a, double X, double Theta) {
  double f = exp(-a.cv * a.dbeta *
                 (::V(X - 0.5 * Theta, a.t_initial) + ::V(X + 0.5 * Theta, a.t_initial) - a.vmin));
  if (a.flags & F_BLOCH_MASK) f *= ::fw_mask_x_theta(X, Theta, a.t_initial);
  return mk((real_t)f, (real_t)0);
}
// bloch_expK_bulk / _boundary, wigner_bloch_cuda_1d.py:290-320 (ck = 1/2);
// wigner_bloch_fftw_1d.py:86-100 (ck = 1).
FWD cpx bloch_kfactor(const Args& a, double Lambda, double P) {
  const double f = exp(-a.ck * a.dbeta *
                       (::K(P + 0.5 * Lambda, a.t_initial) + ::K(P - 0.5 * Lambda, a.t_initial) - a.kmin));
  return mk((real_t)f, (real_t)0);
}

// Separate the packed spectra of lines a, b at (k, N-k), multiply by fa, fb,
// repack, scale by s. k = 0 carries DC (z0) and Nyquist (zn) and fa0/fb0, faN/fbN.
FWD void pair_apply(cpx& zk, cpx& zm, cpx fa, cpx fb, real_t s) {
  const real_t h = (real_t)0.5 * s;
  const cpx A = mk(h * (zk.x + zm.x), h * (zk.y - zm.y));
  const cpx B = mk(h * (zk.y + zm.y), h * (zm.x - zk.x));
  const cpx Ap = cmul(A, fa), Bp = cmul(B, fb);
  zk = mk(Ap.x - Bp.y, Ap.y + Bp.x);
  zm = mk(Ap.x + Bp.y, Bp.x - Ap.y);
}
FWD cpx self_apply(cpx z, cpx fa, cpx fb, real_t s) {
  return mk(z.x * fa.x * s, z.y * fb.x * s);
}

// The x-p mask of column pair (Xa, Xb) at row momentum Pn: when the mask
// string does not reference P (FW_MASK_XP_X_ONLY, set by the host), one value
// per column, evaluated once per pass (ma, mb) instead of once per element.
#ifndef FW_MASK_XP_X_ONLY
#define FW_MASK_XP_X_ONLY 0
#endif
struct MaskXP {
  bool on;
  real_t ma, mb;
  double Xa, Xb, tnow;
  FWD MaskXP(bool on_, double Xa_, double Xb_, double tnow_) : on(on_), ma(1), mb(1), Xa(Xa_), Xb(Xb_), tnow(tnow_) {
    if (FW_MASK_XP_X_ONLY && on) {
      ma = (real_t)::fw_mask_x_p(Xa, 0.0, tnow);
      mb = (real_t)::fw_mask_x_p(Xb, 0.0, tnow);
    }
  }
  FWD void apply(cpx& w, double Pn) const {
    if (!on) return;
    if (FW_MASK_XP_X_ONLY) {
      w.x *= ma;
      w.y *= mb;
    } else {
      w.x *= (real_t)::fw_mask_x_p(Xa, Pn, tnow);
      w.y *= (real_t)::fw_mask_x_p(Xb, Pn, tnow);
    }
  }
};

// Partial sums of the column passes of Wigner steps are stored by sum (see
// write_partial_rows): the S0 row, which every block of the next pass adds for
// the deferred norm, is nbp contiguous doubles.
// The S0 row of the previous column pass, loaded with the pass's own global
// loads and added in sum_partials' order (b = tid, tid + NT, ..., compensated,
// then waves in order), so every block obtains the same bits.
template <int NT>
struct NormRow {
  static constexpr int NBP = FW_NX / 2 / FW_COL_P;  // blocks of the column pass
  static constexpr int PL = (NBP + NT - 1) / NT;
  double x[PL];
  FWD void load(const double* __restrict__ partials) {
#pragma unroll
    for (int j = 0; j < PL; ++j) {
      const int b = threadIdx.x + j * NT;
      x[j] = (b < NBP) ? partials[b] : 0.0;
    }
  }
  // block-reduction slots into red[0..BlockRed<NT>::RN); a barrier must
  // separate this from total()
  FWD void reduce_waves(double* red) const {
    double sum = 0.0, comp = 0.0;
#pragma unroll
    for (int j = 0; j < PL; ++j) {
      if (threadIdx.x + j * NT < NBP) {
        const double tt = sum + x[j];
        comp += (fabs(sum) >= fabs(x[j])) ? (sum - tt) + x[j] : (x[j] - tt) + sum;
        sum = tt;
      }
    }
    BlockRed<NT>::put(sum + comp, red, 1, 0);
  }
  FWD static double total(const double* red) { return BlockRed<NT>::total(red, 1, 0); }
};

// Block b of the nb blocks of a pass adds rows b, b + nb, ... (< NSER) of the
// previous column pass's partials and publishes them to series[step] (called
// by whole blocks, after their own work).
template <int NT>
FWD void publish_sums(const Args& a, int blk, int nb, double* red) {
  for (int i = blk; i < NSER; i += nb) {
    const double tot = sum_partials<NT, 1>(a.partials + (size_t)i * a.nbp, a.nbp, 0, red);
    if (threadIdx.x == 0) a.series[(size_t)a.step * NSER + i] = tot;
  }
}

// Deferred norm of the previous step from its S0 (preloaded in a NormRow whose
// reduction slots are in red, and a barrier passed). With the purity guard
// armed, a tripped guard raises SC_STOP, publishes the sums and returns false.
template <int NT>
FWD bool previous_norm(const Args& a, int blk, double* red, double& nrm) {
  const double dxdp = ::dX * ::dP;
  const double S0 = NormRow<NT>::total(red);
  nrm = 1.0 / (S0 * dxdp);
  if (a.guard > 0.0) {
    const double S2 = sum_partials<NT, 1>(a.partials + (size_t)8 * a.nbp, a.nbp, 0, red);
    if (!(2.0 * M_PI * S2 / (S0 * S0 * dxdp) < a.guard)) {
      if (blk == 0 && threadIdx.x == 0) a.scal[SC_STOP] = (double)(a.step + 1);
      publish_sums<NT>(a, blk, gridDim.x, red);
      return false;
    }
  }
  return true;
}

// The pass that follows the first reader of the `fresh` flag in a step (the
// three-pass K pass, the fused VV pass) clears it, so the next step takes the
// partials. At the end of the kernel: a flag load and store between the W
// loads and their first use would make every wave wait for all of its loads
// (the wait counter is not known after the branch).
FWD void clear_fresh(const Args& a, int blk, int tid) {
  if (blk == 0 && tid == 0 && a.scal[SC_FRESH] != 0.0) a.scal[SC_FRESH] = 0.0;
}

// ---------------------------------------------------------------------------
// V pass: columns (along p), lines = pairs of adjacent columns.
// ---------------------------------------------------------------------------
// Twiddles of the FP64 steps are computed, not staged (TwTab mode 2: one FP64
// sine/cosine per butterfly, the powers by products): no table in LDS (32 KiB per
// block at N = 2048), no table loads, twice the blocks per CU; measured 20-27%
// faster per pass than the staged full table (mode 0) or the two-level table
// (mode 1) on gfx950. The FP32 form keeps the full compensated table.
#ifndef FW_SW_AHEAD
// the theta-sum weights of a thread's items loaded with the pass's other loads
// (FP64, where the column tiles leave one wave per SIMD; the FP32 pass kept
// them in scratch memory)
#define FW_SW_AHEAD (!FW_FP32 && FW_NP <= 1024)
#endif
#ifndef FW_TW_PLAN
#define FW_TW_PLAN 1   // base twiddles of the spectral steps computed ahead (FP64)
#endif
#ifndef FW_TW_PLAN_COL
// the column pass too, unless its tile already fills the registers (E >= 16)
#define FW_TW_PLAN_COL (FW_TW_PLAN && FW_COL_E < 16)
#endif
#ifndef FW_TW2
#define FW_TW2 (FW_FP32 ? 0 : 2)
#endif
using VL = Line<FW_NP, FW_COL_E, FW_COL_P, false, FW_NP, FW_TW2>;
using KL = Line<FW_NX, FW_ROW_E, FW_ROW_P, true, FW_NX, FW_TW2>;
// Passes whose whole grid is resident at once, at most one wave per SIMD of
// the MI355X's 256 CUs x 4 SIMDs: extra registers held through such a pass
// cost no occupancy.
constexpr bool PUB_AHEAD_K = (FW_NP / 2 / FW_ROW_P) * ((KL::NT + 63) / 64) <= 1024;
constexpr bool PUB_AHEAD_V = (FW_NX / 2 / FW_COL_P) * ((VL::NT + 63) / 64) <= 1024;

// The V factor table is stored by column pair: entry (k, x) at
// ((x/2)(N_p/2+1) + k) 2 + x%2, so a column pass reads each pair's N_p/2+1
// factors contiguously (row-major storage would make these reads as
// fragmented as the column reads of W).
FWD size_t vtab_index(int x, int k) {
  return ((size_t)(x >> 1) * (FW_NP / 2 + 1) + k) * 2 + (x & 1);
}

// V factors of column pair (c, c+1) that a thread of a column pass applies,
// from the table: slot i < E/2 for its pair item i, slot E/2 for k = N/2
// (used by t = 0 only).
FWD void v_factors(const Args& a, int c, int t, cpx2 (&ftab)[FW_COL_E / 2 + 1]) {
  constexpr int N = FW_NP, E = FW_COL_E;
#pragma unroll
  for (int i = 0; i < E / 2; ++i)
    ftab[i] = *reinterpret_cast<const cpx2*>(a.vtab + vtab_index(c, spec_k<VL>(t, i)));
  ftab[E / 2] = *reinterpret_cast<const cpx2*>(a.vtab + vtab_index(c, N / 2));
}

// The field phases exp(+i g k) of split potentials on a pair-fused column
// line in FP64, as powers: item m of thread t sits at k = t + m S (m < RL/2),
// (RL - m) S - t (m >= RL/2), or for thread 0 at S/2 + (m - RL/2) S, so with
// u = exp(i g t), h = exp(i g S/2) and w = h^2 every item is u, conj(u) or h
// times w^j, j <= RL/2, and the Nyquist entry is w^(RL/2): two sincos per
// half-step instead of one per item.
constexpr bool FIELD_POW = FW_V_SPLIT && !FW_FP32 && VL::PAIRED && VL::BS == 1;
struct FieldPow {
  double ur = 1.0, ui = 0.0, hr = 1.0, hi = 0.0, wr = 1.0, wi = 0.0;
  FWD void init(double g, int t) {
    sincos_fp64(g * t, ui, ur);
    sincos_fp64(g * (0.5 * VL::S), hi, hr);
    wr = hr * hr - hi * hi;
    wi = 2.0 * hr * hi;
  }
  FWD cpx item(int t, int m) const {
    constexpr int RL = VL::RL;
    double pr, pi;
    int j;
    if (m < RL / 2) {
      pr = ur; pi = ui; j = m;
    } else {
      const bool z = (t == 0);
      pr = z ? hr : ur;
      pi = z ? hi : -ui;
      j = z ? m - RL / 2 : RL - m;
    }
#pragma unroll
    for (int q = 0; q < RL / 2; ++q)
      if (q < j) {
        const double x = pr * wr - pi * wi;
        pi = pr * wi + pi * wr;
        pr = x;
      }
    return mk((real_t)pr, (real_t)pi);
  }
  FWD cpx nyquist() const {
    double pr = 1.0, pi = 0.0;
#pragma unroll
    for (int q = 0; q < VL::RL / 2; ++q) {
      const double x = pr * wr - pi * wi;
      pi = pr * wi + pi * wr;
      pr = x;
    }
    return mk((real_t)pr, (real_t)pi);
  }
};

// The field factor exp(+i dt/2 Theta_k F(tau)) of split potentials on a
// thread's slots (both columns of the pair share it); nothing otherwise.
FWD void v_field_phase(int t, double F, cpx2 (&ftab)[FW_COL_E / 2 + 1]) {
#if FW_V_SPLIT
  constexpr int N = FW_NP, E = FW_COL_E;
  const double g = 0.5 * ::dt * F * ::dTheta;
  if constexpr (FIELD_POW) {
    // FP64 pair-fused lines: the phases as powers (FieldPow); the FP32 form
    // evaluates every phase (its series is cheap)
    FieldPow fp;
    fp.init(g, t);
#pragma unroll
    for (int i = 0; i < E / 2; ++i) {
      const cpx e = fp.item(t, i);
      ftab[i].a = cmul(ftab[i].a, e);
      ftab[i].b = cmul(ftab[i].b, e);
    }
    const cpx en = fp.nyquist();
    ftab[E / 2].a = cmul(ftab[E / 2].a, en);
    ftab[E / 2].b = cmul(ftab[E / 2].b, en);
  } else {
#pragma unroll
    for (int i = 0; i <= E / 2; ++i) {
      const cpx e = unit_phase(g * (i < E / 2 ? spec_k<VL>(t, i) : N / 2), 1.0);
      ftab[i].a = cmul(ftab[i].a, e);
      ftab[i].b = cmul(ftab[i].b, e);
    }
  }
#else
  (void)t;
  (void)F;
  (void)ftab;
#endif
}

// Field values of split potentials, filled per chunk next to the step times
// (fw_begin): F(t_s + shift dt) for the factors, F(t_s) for the sums.
#if FW_V_SPLIT
FWD double field_tau(const Args& a, int s) { return a.times[FW_TCHUNK + s]; }
FWD double field_now(const Args& a, int s) { return a.times[2 * FW_TCHUNK + s]; }
FWD void field_fill(const Args& a) {
  double* f = const_cast<double*>(a.times);
  for (int s = threadIdx.x; s < a.step; s += blockDim.x) {
    const double ts = a.times[s];
    f[FW_TCHUNK + s] = ::fw_V_field(ts + FW_PHASE_SHIFT * ::dt);
    f[2 * FW_TCHUNK + s] = ::fw_V_field(ts);
#if FW_DV_SPLIT
    f[3 * FW_TCHUNK + s] = ::fw_dV_field(ts);
#endif
  }
}
// V(X, t_s) in the sums: the table's V0 plus X F(t_s)
FWD double v_now(const Args& a, double X, int s) { return ::fw_V_static(X, 0.0) + X * field_now(a, s); }
#else
FWD double field_tau(const Args&, int) { return 0.0; }
FWD void field_fill(const Args&) {}
FWD double v_now(const Args& a, double X, int s) { return ::V(X, a.times[s]); }
#endif
// V'(X, t_s) in the sums: dV0(X) plus the offset G(t_s) when V' splits
#if FW_DV_SPLIT
FWD double dv_now(const Args& a, double X, int s) { return ::fw_dV_static(X, 0.0) + a.times[3 * FW_TCHUNK + s]; }
#else
FWD double dv_now(const Args& a, double X, int s) { return ::fw_diff_V(X, a.times[s]); }
#endif

// The same factors evaluated at time tau (masks at tnow), for time-dependent V.
// One straight-line block before the spectral step: the evaluations share
// their time-only subexpressions (a field E(t) in V, say), which the compiler
// then computes once per pass instead of once per factor.
FWD void v_factors_at(int c, int t, double tau, double tnow, cpx2 (&ftab)[FW_COL_E / 2 + 1]) {
  constexpr int N = FW_NP, E = FW_COL_E;
  const double Xa = ::dX * (c - 0.5 * ::X_gridDIM);
  const double Xb = ::dX * (c + 1 - 0.5 * ::X_gridDIM);
#pragma unroll
  for (int i = 0; i <= E / 2; ++i) {
    const double Th = ::dTheta * (i < E / 2 ? spec_k<VL>(t, i) : N / 2);
    ftab[i].a = vfactor(Xa, Th, tau, tnow);
    ftab[i].b = vfactor(Xb, Th, tau, tnow);
  }
}

// Pair multiply of the column passes: real-pair separation of columns a, b,
// their V factors from the register array ftab (read from the table or
// evaluated before the step), repacking, scale 1/N.
struct VMul {
  const cpx2 (&ftab)[FW_COL_E / 2 + 1];
  FWD void pair(int i, cpx& zk, cpx& zm) const {
    pair_apply(zk, zm, ftab[i].a, ftab[i].b, (real_t)1 / (real_t)FW_NP);
  }
  FWD void self(cpx& z0, cpx& zn) const {
    const real_t s = (real_t)1 / (real_t)FW_NP;
    z0 = self_apply(z0, ftab[0].a, ftab[0].b, s);
    zn = self_apply(zn, ftab[FW_COL_E / 2].a, ftab[FW_COL_E / 2].b, s);
  }
};

// Generated K factors. When K(P + L/2) - K(P - L/2) is linear in L (K
// quadratic in P, as P^2/2 without a lambda-p mask) the factor of row P at
// spectral index k is z^k, z = exp(i beta) with beta = phase(N/2) / (N/2).
// The items of thread t of a pair-fused row line are k = t + m S (m < RL/2)
// and k = (RL - m) S - t (m >= RL/2), and thread 0 holds S/2 + (m - RL/2) S
// instead, so with a = z^t, h = z^(S/2) and w = h^2 = z^S every factor is a
// or conj(a) or h times a power w^j, j <= RL/2 (two sincos per row, the rest
// products; the Nyquist factor is w^(RL/2)). Used when these values match the
// table on every grid point (kgen_check), in place of reading it.
template <class L>
FWD void kgen_factors(double P, double tau, int t, cpx (&f)[L::LE / 2 + 1]) {
  static_assert(L::PAIRED && L::BS == 1, "one slot per thread of a pair-fused line");
  constexpr int N = L::LN, RL = L::RL, S = L::S;
  const double Lh = ::dLambda * (N / 2);
  const double beta = -::dt * (::K(P + 0.5 * Lh, tau) - ::K(P - 0.5 * Lh, tau)) * (2.0 / N);
  double hs, hc, as, ac;
  sincos_fp64(beta * (0.5 * S), hs, hc);
  sincos_fp64(beta * t, as, ac);
  double wr[RL / 2 + 1], wi[RL / 2 + 1];
  wr[0] = 1.0;
  wi[0] = 0.0;
  wr[1] = hc * hc - hs * hs;
  wi[1] = 2.0 * hc * hs;
#pragma unroll
  for (int j = 2; j <= RL / 2; ++j) {
    wr[j] = wr[j - 1] * wr[1] - wi[j - 1] * wi[1];
    wi[j] = wr[j - 1] * wi[1] + wi[j - 1] * wr[1];
  }
  const bool t0 = (t == 0);
#pragma unroll
  for (int m = 0; m < RL; ++m) {
    double br, bi;   // the factor multiplying w^j
    int j;
    if (m < RL / 2) {
      br = ac; bi = as; j = m;
    } else {
      br = t0 ? hc : ac;
      bi = t0 ? hs : -as;
      j = t0 ? m - RL / 2 : RL - m;
    }
    f[m] = mk((real_t)(br * wr[j] - bi * wi[j]), (real_t)(br * wi[j] + bi * wr[j]));
  }
  f[L::LE / 2] = mk((real_t)wr[RL / 2], (real_t)wi[RL / 2]);
}

// Pair multiply of the row passes: rows a, b and their K factors.
struct KMul {
  const cpx (&fta)[FW_ROW_E / 2 + 1];
  const cpx (&ftb)[FW_ROW_E / 2 + 1];
  FWD void pair(int i, cpx& zk, cpx& zm) const {
    pair_apply(zk, zm, fta[i], ftb[i], (real_t)1 / (real_t)FW_NX);
  }
  FWD void self(cpx& z0, cpx& zn) const {
    const real_t s = (real_t)1 / (real_t)FW_NX;
    z0 = self_apply(z0, fta[0], ftb[0], s);
    zn = self_apply(zn, fta[FW_ROW_E / 2], ftb[FW_ROW_E / 2], s);
  }
};

// ---------------------------------------------------------------------------
// Sums on the theta spectrum. After the closing V/2 a column is
//   a'_n = sum_k alpha_k exp(+2 pi i n k / N),   alpha_k = A_k f_k / N
// (A_k its FFT along p, f_k the factor), so for a real function f of P
//   sum_n f(P_n) a'_n = sum_k alpha_k conj(F_k),   sum_n a'_n^2 = N sum_k |alpha_k|^2
// with F_k = sum_n f(P_n) exp(-2 pi i n k / N) (Parseval). a' is real, so k and
// N - k contribute conjugate terms and the sums run over k <= N/2 with weights
// c_k (SumW). When nothing acts on W in x-p between the closing V/2 of step s
// and the opening V/2 of step s+1 but a mask of x alone (which scales a whole
// column, so its spectrum too), and K, K' do not depend on t, the column pass
// takes the nine sums of step s on alpha_k and multiplies by the opening
// factors in the same registers: one FFT pair per pass instead of two, the
// same W and the same sums in exact arithmetic.
// ---------------------------------------------------------------------------
#ifndef FW_THETA_SUMS
#define FW_THETA_SUMS 0
#endif
#ifndef FW_V_STATIC
#define FW_V_STATIC 0
#endif
struct VVMul {
  const cpx2 (&f0)[FW_COL_E / 2 + 1];   // closing V/2 (step s)
  const cpx2 (&f1)[FW_COL_E / 2 + 1];   // opening V/2 (step s+1)
  real_t ma, mb;                        // x-only x-p mask of columns a, b (1 without a mask)
  // weights: FW_SW_AHEAD, this thread's items (then the Nyquist entry), loaded
  // with the pass's other loads; otherwise the table, read at the multiply
  const SumW* __restrict__ sw;
  int t;                                // lane of the line
  double acc[4];                        // sum c_k Re(sigma_k conj F_k), sigma = alpha_a + alpha_b,
                                        // for P, P^2, K', K
  double pur;                           // sum c_k (|alpha_a|^2 + |alpha_b|^2)
  double a0, b0;                        // alpha_0 of columns a, b (thread 0 of a line)
  double g0, g1;                        // split V: field phase per unit k at the two half-steps
  FieldPow p0{}, p1{};                  // FIELD_POW: the phases of g0, g1 as powers
  // the factors of item i (k = spec_k): split potentials take the table of V0
  // (f0 = f1) times exp(+i g k), evaluated here so that only one factor set
  // stays in registers
  FWD cpx2 fac(const cpx2 (&f)[FW_COL_E / 2 + 1], int i, double g, const FieldPow& fp) const {
#if FW_V_SPLIT
    cpx e;
    if constexpr (FIELD_POW) e = fp.item(t, i);
    else e = unit_phase(g * (double)spec_k<VL>(t, i), 1.0);
    cpx2 r;
    r.a = cmul(f[i].a, e);
    r.b = cmul(f[i].b, e);
    return r;
#else
    (void)g;
    (void)fp;
    return f[i];
#endif
  }
  FWD void pair(int i, cpx& zk, cpx& zm) {
    const real_t h = (real_t)0.5 / (real_t)FW_NP;
    const cpx A = mk(h * (zk.x + zm.x), h * (zk.y - zm.y));
    const cpx B = mk(h * (zk.y + zm.y), h * (zm.x - zk.x));
    const cpx2 F0 = fac(f0, i, g0, p0);
    const cpx A1 = cscale(cmul(A, F0.a), ma), B1 = cscale(cmul(B, F0.b), mb);
    const double sx = (double)A1.x + (double)B1.x, sy = (double)A1.y + (double)B1.y;
#if FW_SW_AHEAD
    const SumW w = sw[i];
#else
    const SumW w = sw[spec_k<VL>(t, i)];
#endif
    acc[0] += sy * w.a - sx * (::dP * FW_NP);
    acc[1] += sx * w.b;
    acc[2] += sx * w.dk[0] + sy * w.dk[1];
    acc[3] += sx * w.kk[0] + sy * w.kk[1];
    const double ax = A1.x, ay = A1.y, bx = B1.x, by = B1.y;
    pur += 2.0 * (ax * ax + ay * ay + bx * bx + by * by);
    const cpx2 F1 = fac(f1, i, g1, p1);
    const cpx A2 = cmul(A1, F1.a), B2 = cmul(B1, F1.b);
    zk = mk(A2.x - B2.y, A2.y + B2.x);
    zm = mk(A2.x + B2.y, B2.x - A2.y);
  }
  // DC and Nyquist: real part times the real part of the factor (c2r rule)
  FWD void self(cpx& z0, cpx& zn) {
    constexpr int E2 = FW_COL_E / 2;
    const real_t s = (real_t)1 / (real_t)FW_NP;
#if FW_V_SPLIT
    // Nyquist: the real part of the phased factor (the DC phase is 1)
    const cpx en0 = FIELD_POW ? p0.nyquist() : unit_phase(g0 * (FW_NP / 2), 1.0);
    const cpx en1 = FIELD_POW ? p1.nyquist() : unit_phase(g1 * (FW_NP / 2), 1.0);
    const cpx2 n0 = {cmul(f0[E2].a, en0), cmul(f0[E2].b, en0)}, n1 = {cmul(f1[E2].a, en1), cmul(f1[E2].b, en1)};
#else
    const cpx2 n0 = f0[E2], n1 = f1[E2];
#endif
    const real_t ra = z0.x * s * f0[0].a.x * ma, rb = z0.y * s * f0[0].b.x * mb;
    const real_t na = zn.x * s * n0.a.x * ma, nb = zn.y * s * n0.b.x * mb;
    a0 = ra;
    b0 = rb;
    const double s0 = (double)ra + (double)rb, sn = (double)na + (double)nb;
#if FW_SW_AHEAD
    const SumW w0 = sw[0], wn = sw[FW_COL_E / 2];   // thread 0: item 0 is k = 0
#else
    const SumW w0 = sw[0], wn = sw[FW_NP / 2];
#endif
    acc[0] += s0 * w0.a + sn * wn.a;
    acc[1] += s0 * w0.b + sn * wn.b;
    acc[2] += s0 * w0.dk[0] + sn * wn.dk[0];
    acc[3] += s0 * w0.kk[0] + sn * wn.kk[0];
    pur += (double)ra * ra + (double)rb * rb + (double)na * na + (double)nb * nb;
    z0 = mk(ra * f1[0].a.x, rb * f1[0].b.x);
    zn = mk(na * n1.a.x, nb * n1.b.x);
  }
};

template <int FL>
FWD void vpass(const Args& a) {
  using L = VL;
  constexpr int N = FW_NP, E = FW_COL_E, P = FW_COL_P, TPL = L::TPL, NT = L::NT;
  __shared__ cpx lds[L::LDS];
  __shared__ tw_t twl[L::TW::SIZE];
  __shared__ double red[NSER * BlockRed<NT>::RN];
  if (a.scal[SC_STOP] != 0.0) return;
  const int tid = threadIdx.x;
  const int q = tid % P, t = tid / P;
  const int nb = gridDim.x;
  const int blk = xcd_remap(blockIdx.x, nb);
  const int c = (blk * P + q) * 2;
  const double tnow = a.times[a.step];
  const double tau = tnow + FW_PHASE_SHIFT * ::dt;
  // the factors come from the cached table for Bloch steps and for
  // time-independent V (FW_VTAB); otherwise they are evaluated here
  constexpr bool use_tab = (FL & F_BLOCH) || FW_VTAB;
  real_t* __restrict__ W = a.W;

  // issue every global load of the pass up front: twiddles, W, factor table
  TwStage<typename L::TW, NT> tws;
  tws.load(a.tw_col, tid);
  cpx v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = *reinterpret_cast<const cpx*>(W + (size_t)(t + TPL * e) * FW_NX + c);
  // the previous step's S0 row (deferred norm), issued before the table loads
  const bool deferred = (FL & F_FIRST) && a.scal[SC_FRESH] == 0.0;
  NormRow<NT> norm_row;
  if (deferred) norm_row.load(a.partials);
  // the row of the previous step's sums this block publishes at its end,
  // loaded with the rest when the grid is resident at once (PUB_AHEAD_V)
  const bool pub_pre = PUB_AHEAD_V && deferred && gridDim.x >= NSER && blk < NSER;
  NormRow<NT> pub_row;
  if (pub_pre && blk > 0) pub_row.load(a.partials + (size_t)blk * NormRow<NT>::NBP);
  cpx2 ftab[E / 2 + 1];
  if constexpr (use_tab) {
    v_factors(a, c, t, ftab);
    if constexpr (!(FL & F_BLOCH)) v_field_phase(t, field_tau(a, a.step), ftab);
  } else {
    v_factors_at(c, t, tau, tnow, ftab);
  }
  tws.store(twl, tid);
  if (deferred) norm_row.reduce_waves(red);
  __syncthreads();

  if constexpr (FL & F_FIRST) {
    double nrm = 1.0;
    if (deferred && !previous_norm<NT>(a, blk, red, nrm)) return;
    const real_t rn = (real_t)nrm;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      cpx w = v[e];
      w.x *= rn; w.y *= rn;
      if constexpr (FL & F_DISSIP) {
        const cpx g = *reinterpret_cast<const cpx*>(a.W0s + (size_t)(t + TPL * e) * FW_NX + c);
        w.x = w.x * (real_t)a.w_coeff + g.x;
        w.y = w.y * (real_t)a.w_coeff + g.y;
      }
      v[e] = w;
    }
  }
  // spectral multiply on the theta x representation
  const double Xa = ::dX * (c - 0.5 * ::X_gridDIM);
  const double Xb = ::dX * (c + 1 - 0.5 * ::X_gridDIM);
  VMul mul{ftab};
  spectral_step<L>(v, lds, q, t, twl, mul);

  double sums[NSER];
#pragma unroll
  for (int i = 0; i < NSER; ++i) sums[i] = 0.0;
  double ca = 0.0, cb = 0.0;
  const MaskXP mask((FL & F_LAST) && (a.flags & F_MASK_XP), Xa, Xb, tnow);
  // the dissipator's source terms, loaded ahead of the stores of W (a load
  // after a store through asm cannot be moved above it)
  cpx g0s[(FL & F_LAST) && (FL & F_DISSIP) ? E : 1];
  if constexpr ((FL & F_LAST) && (FL & F_DISSIP)) {
#pragma unroll
    for (int e = 0; e < E; ++e) g0s[e] = *reinterpret_cast<const cpx*>(a.W0s + (size_t)(t + TPL * e) * FW_NX + c);
  }
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = t + TPL * e;
    const size_t idx = (size_t)n * FW_NX + c;
    cpx w = v[e];
    if constexpr (FL & F_LAST) {
      const double Pn = ::dP * (n - 0.5 * ::P_gridDIM);
      mask.apply(w, Pn);
      if constexpr (FL & F_DISSIP) {
        const cpx g = g0s[e];
        w.x = w.x * (real_t)a.w_coeff + g.x;
        w.y = w.y * (real_t)a.w_coeff + g.y;
      }
      const double wa = (double)w.x, wb = (double)w.y, ws = wa + wb;
      ca += wa; cb += wb;
      sums[2] += Pn * ws;
      sums[4] += Pn * Pn * ws;
      sums[5] += ::fw_diff_K(Pn, tnow) * ws;
      sums[7] += ::K(Pn, tnow) * ws;
      sums[8] += wa * wa + wb * wb;
    }
    st_w(reinterpret_cast<cpx*>(W + idx), w);
  }
  if constexpr (FL & F_LAST) {
    sums[0] = ca + cb;
    sums[1] = Xa * ca + Xb * cb;
    sums[3] = Xa * Xa * ca + Xb * Xb * cb;
    if constexpr (FL & F_BLOCH) {
      // imaginary time: the whole V at t_initial (no field values uploaded)
      sums[6] = ::fw_diff_V(Xa, tnow) * ca + ::fw_diff_V(Xb, tnow) * cb;
      sums[7] += ::V(Xa, tnow) * ca + ::V(Xb, tnow) * cb;
    } else {
      sums[6] = dv_now(a, Xa, a.step) * ca + dv_now(a, Xb, a.step) * cb;
      sums[7] += v_now(a, Xa, a.step) * ca + v_now(a, Xb, a.step) * cb;
    }
    // the previous step's sums were published by the K pass (three-pass steps
    // open with this pass only after a fresh start or a K pass)
    write_partial_rows<NT>(sums, red, a.partials, blk, nb);
  }
  if constexpr (FL & F_FIRST) {
    if (pub_pre) {
      // the preloaded row added in sum_partials' order: the same bits (the
      // barrier: red may still be read by write_partial_rows above)
      __syncthreads();
      (blk == 0 ? norm_row : pub_row).reduce_waves(red);
      __syncthreads();
      if (tid == 0) a.series[(size_t)a.step * NSER + blk] = NormRow<NT>::total(red);
    } else if (deferred) {
      publish_sums<NT>(a, blk, gridDim.x, red);
    }
  }
}


// ---------------------------------------------------------------------------
// Fused column pass of two-pass steps: the closing V/2 of step s and the
// opening V/2 of step s+1 in one read and one write of W. Between them, in
// registers: the x-p mask and the nine sums of step s (partials; the norm is
// applied by the next K pass, which is linear like everything before it).
// ---------------------------------------------------------------------------
FWD void vvpass(const Args& a) {
  using L = VL;
  constexpr int N = FW_NP, E = FW_COL_E, P = FW_COL_P, TPL = L::TPL, NT = L::NT;
  __shared__ cpx lds[L::LDS];
  __shared__ tw_t twl[L::TW::SIZE];
  __shared__ double red[NSER * BlockRed<NT>::RN];
  // two-pass steps are Moyal steps: the stop flag (purity guard of Bloch runs)
  // is never raised, and not reading it lets the W loads issue at once
  const int tid = threadIdx.x;
  const int q = tid % P, t = tid / P;
  const int nb = gridDim.x;
  const int blk = xcd_remap(blockIdx.x, nb);
  const int c = (blk * P + q) * 2;
  const double t0 = a.times[a.step], t1 = a.times[a.step + 1];
  constexpr bool use_tab = FW_VTAB;   // one table read serves both half-steps
  real_t* __restrict__ W = a.W;

  TwStage<typename L::TW, NT> tws;
  tws.load(a.tw_col, tid);
  cpx v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = *reinterpret_cast<const cpx*>(W + (size_t)(t + TPL * e) * FW_NX + c);
  cpx2 ftab[E / 2 + 1];
  if constexpr (use_tab) {
    v_factors(a, c, t, ftab);
    v_field_phase(t, field_tau(a, a.step), ftab);
  } else {
    v_factors_at(c, t, t0 + FW_PHASE_SHIFT * ::dt, t0, ftab);
  }
  tws.store(twl, tid);
  __syncthreads();

  const double Xa = ::dX * (c - 0.5 * ::X_gridDIM);
  const double Xb = ::dX * (c + 1 - 0.5 * ::X_gridDIM);
  // closing V/2 of step s
  VMul mul0{ftab};
  spectral_step<L>(v, lds, q, t, twl, mul0);
  // end of step s: x-p mask, sums
  double sums[NSER];
#pragma unroll
  for (int i = 0; i < NSER; ++i) sums[i] = 0.0;
  double ca = 0.0, cb = 0.0;
  const MaskXP mask((a.flags & F_MASK_XP) != 0, Xa, Xb, t0);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = t + TPL * e;
    cpx w = v[e];
    const double Pn = ::dP * (n - 0.5 * ::P_gridDIM);
    mask.apply(w, Pn);
    const double wa = (double)w.x, wb = (double)w.y, ws = wa + wb;
    ca += wa; cb += wb;
    sums[2] += Pn * ws;
    sums[4] += Pn * Pn * ws;
    sums[5] += ::fw_diff_K(Pn, t0) * ws;
    sums[7] += ::K(Pn, t0) * ws;
    sums[8] += wa * wa + wb * wb;
    v[e] = w;
  }
  sums[0] = ca + cb;
  sums[1] = Xa * ca + Xb * cb;
  sums[3] = Xa * Xa * ca + Xb * Xb * cb;
  sums[6] = dv_now(a, Xa, a.step) * ca + dv_now(a, Xb, a.step) * cb;
  sums[7] += v_now(a, Xa, a.step) * ca + v_now(a, Xb, a.step) * cb;
  write_partial_rows<NT>(sums, red, a.partials, blk, nb);
  // opening V/2 of step s+1. The thread index passes through an empty asm so
  // that the compiler recomputes this step's twiddles and LDS indices instead
  // of keeping those of the closing step live across it (register pressure).
  int t2 = t;
  asm volatile("" : "+v"(t2));
  if constexpr (!use_tab) {
    v_factors_at(c, t2, t1 + FW_PHASE_SHIFT * ::dt, t1, ftab);
  } else if constexpr (FW_V_SPLIT) {
    // the table again (L2) with the opening step's field factor
    v_factors(a, c, t2, ftab);
    v_field_phase(t2, field_tau(a, a.step + 1), ftab);
  }
  VMul mul1{ftab};
  spectral_step<L>(v, lds, q, t2, twl, mul1);
#pragma unroll
  for (int e = 0; e < E; ++e) st_w(reinterpret_cast<cpx*>(W + (size_t)(t + TPL * e) * FW_NX + c), v[e]);
  clear_fresh(a, blk, tid);
}

// The fused column pass with the sums on the theta spectrum (VVMul): no x-p
// mask or one of x alone, K and K' independent of t (FW_THETA_SUMS, set by
// the host).
FWD void vvpass_theta(const Args& a) {
  using L = VL;
  constexpr int N = FW_NP, E = FW_COL_E, P = FW_COL_P, TPL = L::TPL, NT = L::NT;
  __shared__ cpx lds[L::LDS];
  __shared__ tw_t twl[L::TW::SIZE];
  __shared__ double red[NSER * BlockRed<NT>::RN];
  // two-pass steps are Moyal steps: the stop flag (purity guard of Bloch runs)
  // is never raised, and not reading it lets the W loads issue at once
  const int tid = threadIdx.x;
  const int q = tid % P, t = tid / P;
  const int nb = gridDim.x;
  const int blk = xcd_remap(blockIdx.x, nb);
  const int c = (blk * P + q) * 2;
  const double t0 = a.times[a.step], t1 = a.times[a.step + 1];
  constexpr bool use_tab = FW_VTAB;   // one table read serves both half-steps
  real_t* __restrict__ W = a.W;

  TwStage<typename L::TW, NT> tws;
  tws.load(a.tw_col, tid);
  cpx v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = *reinterpret_cast<const cpx*>(W + (size_t)(t + TPL * e) * FW_NX + c);
  static_assert(!FW_V_SPLIT || FW_VTAB, "split potentials use the table of V0");
  cpx2 f0[E / 2 + 1], f1[E / 2 + 1];
  if constexpr (use_tab) {
    // split potentials: the multiply phases the one table per half-step
    v_factors(a, c, t, f0);
#pragma unroll
    for (int i = 0; i <= E / 2; ++i) f1[i] = f0[i];
  } else if constexpr (FW_V_STATIC) {
    v_factors_at(c, t, t0 + FW_PHASE_SHIFT * ::dt, t0, f0);
#pragma unroll
    for (int i = 0; i <= E / 2; ++i) f1[i] = f0[i];
  } else {
    v_factors_at(c, t, t0 + FW_PHASE_SHIFT * ::dt, t0, f0);
    v_factors_at(c, t, t1 + FW_PHASE_SHIFT * ::dt, t1, f1);
  }
#if FW_SW_AHEAD
  SumW swa[E / 2 + 1];
#pragma unroll
  for (int i = 0; i < E / 2; ++i) swa[i] = a.sw[spec_k<L>(t, i)];
  swa[E / 2] = a.sw[N / 2];
  const SumW* swp = swa;
#else
  const SumW* swp = a.sw;
#endif
  // FP64 pair-fused steps: the base twiddles while the loads are in flight
  typename L::Plan plan;
  if constexpr (L::PLANNED && FW_TW_PLAN_COL) L::plan(plan, t, twl);
  tws.store(twl, tid);
  __syncthreads();

  const double Xa = ::dX * (c - 0.5 * ::X_gridDIM);
  const double Xb = ::dX * (c + 1 - 0.5 * ::X_gridDIM);
  // the x-only mask at t0, once per column (MaskXP); 1 without a mask
  const MaskXP mask((a.flags & F_MASK_XP) != 0, Xa, Xb, t0);
  static_assert(!FW_THETA_SUMS || FW_MASK_XP_X_ONLY, "the theta-spectrum pass takes x-only masks");
  double g0 = 0.0, g1 = 0.0;
#if FW_V_SPLIT
  g0 = 0.5 * ::dt * field_tau(a, a.step) * ::dTheta;
  g1 = 0.5 * ::dt * field_tau(a, a.step + 1) * ::dTheta;
#endif
  VVMul mul{f0, f1, mask.ma, mask.mb, swp, t, {0.0, 0.0, 0.0, 0.0}, 0.0, 0.0, 0.0, g0, g1};
  if constexpr (FIELD_POW) {
    mul.p0.init(g0, t);
    mul.p1.init(g1, t);
  }
  if constexpr (L::PLANNED && FW_TW_PLAN_COL) spectral_step_planned<L>(v, lds, q, t, twl, mul, plan);
  else spectral_step<L>(v, lds, q, t, twl, mul);
#pragma unroll
  for (int e = 0; e < E; ++e) st_w(reinterpret_cast<cpx*>(W + (size_t)(t + TPL * e) * FW_NX + c), v[e]);

  // the nine sums of step s (column sums N alpha_0 on thread 0 of each line)
  const double ca = (double)N * mul.a0, cb = (double)N * mul.b0;
  double sums[NSER];
  sums[0] = ca + cb;
  sums[1] = Xa * ca + Xb * cb;
  sums[2] = mul.acc[0];
  sums[3] = Xa * Xa * ca + Xb * Xb * cb;
  sums[4] = mul.acc[1];
  sums[5] = mul.acc[2];
  sums[6] = 0.0;
  sums[7] = mul.acc[3];
  if (t == 0) {
    sums[6] = dv_now(a, Xa, a.step) * ca + dv_now(a, Xb, a.step) * cb;
    sums[7] += v_now(a, Xa, a.step) * ca + v_now(a, Xb, a.step) * cb;
  }
  sums[8] = (double)N * mul.pur;
  write_partial_rows<NT>(sums, red, a.partials, blk, nb);
  clear_fresh(a, blk, tid);
}

// Theta-spectrum sum weights (SumW), one entry per k <= N_p/2, by direct
// summation in FP64 (the angle index (n k) mod N_p is exact). K and K' do not
// depend on t when these weights are used. acc[o] = (re, im) of F_k for
// f = P, P^2, K', K.
FWD void build_sumw(const Args& a, SumW* sw) {
  constexpr int N = FW_NP;
  const double tk = a.t_initial;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k <= N / 2; k += gridDim.x * blockDim.x) {
    double acc[4][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (int n = 0; n < N; ++n) {
      const int m = (int)(((long long)n * k) % N);
      double sn, cs;
      sincospi(-2.0 * (double)m / (double)N, &sn, &cs);
      const double P = ::dP * (n - 0.5 * ::P_gridDIM);
      const double f[4] = {P, P * P, ::fw_diff_K(P, tk), ::K(P, tk)};
#pragma unroll
      for (int o = 0; o < 4; ++o) {
        acc[o][0] += f[o] * cs;
        acc[o][1] += f[o] * sn;
      }
    }
    SumW w;
    if (k == 0 || k == N / 2) {
      w.a = acc[0][0];
      w.b = acc[1][0];
      w.dk[0] = acc[2][0];
      w.dk[1] = 0.0;
      w.kk[0] = acc[3][0];
      w.kk[1] = 0.0;
    } else {
      w.a = 2.0 * acc[0][1];
      w.b = 2.0 * acc[1][0];
      w.dk[0] = 2.0 * acc[2][0];
      w.dk[1] = 2.0 * acc[2][1];
      w.kk[0] = 2.0 * acc[3][0];
      w.kk[1] = 2.0 * acc[3][1];
    }
    sw[k] = w;
  }
}

// ---------------------------------------------------------------------------
// K pass: rows (along x), lines = pairs of adjacent rows.
// ---------------------------------------------------------------------------
template <int FL>
FWD void kpass(const Args& a) {
  using L = KL;
  constexpr int N = FW_NX, E = FW_ROW_E, P = FW_ROW_P, TPL = L::TPL, NT = L::NT;
  constexpr int KW = FW_NX / 2 + 1;
  __shared__ cpx lds[L::LDS];
  __shared__ tw_t twl[L::TW::SIZE];
  __shared__ double red[(FL & F_KNORM) ? BlockRed<NT>::RN : 1];
  // F_KNORM: the row pass of two-pass Moyal steps, which never stop early
  if (!(FL & F_KNORM) && a.scal[SC_STOP] != 0.0) return;
  const int tid = threadIdx.x;
  const int t = tid % TPL, q = tid / TPL;
  const int nb = gridDim.x;
  const int blk = xcd_remap(blockIdx.x, nb);
  const int r = (blk * P + q) * 2;
  const double tnow = a.times[a.step];
  const double tau = tnow + FW_PHASE_SHIFT * ::dt;
  constexpr bool use_tab = (FL & F_BLOCH) || FW_KTAB;
  real_t* __restrict__ Wa = a.W + (size_t)r * FW_NX;
  real_t* __restrict__ Wb = Wa + FW_NX;

  TwStage<typename L::TW, NT> tws;
  tws.load(a.tw_row, tid);
  cpx v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = mk(Wa[t + TPL * e], Wb[t + TPL * e]);
  // two-pass steps: the previous step's S0 row (deferred norm), issued before the table loads
  const bool deferred = (FL & F_KNORM) && a.scal[SC_FRESH] == 0.0;
  NormRow<NT> norm_row;
  if (deferred) norm_row.load(a.partials);
  // the row of the previous step's sums this block publishes at its end (row
  // blk < NSER; row 0 is S0, already in norm_row), loaded with the rest when
  // the whole grid is resident at once (PUB_AHEAD_K: one wave per SIMD, where
  // the registers it holds cost no occupancy; at 2048^2 they cost a wave)
  const bool pub_pre = PUB_AHEAD_K && deferred && gridDim.x >= NSER && blk < NSER;
  NormRow<NT> pub_row;
  if (pub_pre && blk > 0) pub_row.load(a.partials + (size_t)blk * NormRow<NT>::NBP);
  cpx fta[E / 2 + 1], ftb[E / 2 + 1];
  constexpr bool can_gen = !(FL & F_BLOCH) && L::PAIRED && L::BS == 1;
  if (can_gen && (a.flags & F_KGEN)) {
    if constexpr (can_gen) {
      const double Pa = ::dP * (r - 0.5 * ::P_gridDIM);
      kgen_factors<L>(Pa, tau, t, fta);
      kgen_factors<L>(::dP * (r + 1 - 0.5 * ::P_gridDIM), tau, t, ftb);
    }
  } else if constexpr (use_tab) {
#pragma unroll
    for (int i = 0; i < E / 2; ++i) {
      const int k = spec_k<L>(t, i);
      fta[i] = a.ktab[(size_t)r * KW + k];
      ftb[i] = a.ktab[(size_t)(r + 1) * KW + k];
    }
    fta[E / 2] = a.ktab[(size_t)r * KW + N / 2];
    ftb[E / 2] = a.ktab[(size_t)(r + 1) * KW + N / 2];
  } else {
    // time-dependent K: evaluated here in one block (see v_factors_at)
    const double Pa = ::dP * (r - 0.5 * ::P_gridDIM);
    const double Pb = ::dP * (r + 1 - 0.5 * ::P_gridDIM);
#pragma unroll
    for (int i = 0; i <= E / 2; ++i) {
      const double Lm = ::dLambda * (i < E / 2 ? spec_k<L>(t, i) : N / 2);
      fta[i] = kfactor(Lm, Pa, tau, tnow);
      ftb[i] = kfactor(Lm, Pb, tau, tnow);
    }
  }
  // FP64 pair-fused steps: the base twiddles while the loads are in flight
  typename L::Plan plan;
  if constexpr (L::PLANNED && FW_TW_PLAN) L::plan(plan, t, twl);
  tws.store(twl, tid);
  if (deferred) norm_row.reduce_waves(red);
  __syncthreads();
  if constexpr (FL & F_KNORM) {
    double nrm = 1.0;
    if (deferred && !previous_norm<NT>(a, blk, red, nrm)) return;
    const real_t rn = (real_t)nrm;
#pragma unroll
    for (int e = 0; e < E; ++e) { v[e].x *= rn; v[e].y *= rn; }
  }
  KMul mul{fta, ftb};
  if constexpr (L::PLANNED && FW_TW_PLAN) spectral_step_planned<L>(v, lds, q, t, twl, mul, plan);
  else spectral_step<L>(v, lds, q, t, twl, mul);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    st_w(&Wa[t + TPL * e], v[e].x);
    st_w(&Wb[t + TPL * e], v[e].y);
  }
  // the previous step's sums to the series, after this block's own work (the
  // preloaded row added in sum_partials' order: the same bits)
  if constexpr (FL & F_KNORM) {
    if (pub_pre) {
      (blk == 0 ? norm_row : pub_row).reduce_waves(red);
      __syncthreads();
      if (tid == 0) a.series[(size_t)a.step * NSER + blk] = NormRow<NT>::total(red);
    } else if (deferred) {
      publish_sums<NT>(a, blk, gridDim.x, red);
    }
  }
  if constexpr (!(FL & F_KNORM)) clear_fresh(a, blk, tid);
}

// ---------------------------------------------------------------------------
// Whole-grid helpers
// ---------------------------------------------------------------------------
constexpr int HT = 256;   // threads per block of the helper kernels

// The nine sums over the current W at time times[step]; result in scal[SC_OUT..],
// and scal[SC_NORM] = 1/(S0 dXdP). One-off use (normalisation, getters): the
// last-block pattern with one atomic per block is acceptable here.
FWD void moments(const Args& a) {
  __shared__ double red[NSER * 4];
  const int tid = threadIdx.x;
  const double tnow = a.times[a.step];
  double s[NSER];
#pragma unroll
  for (int i = 0; i < NSER; ++i) s[i] = 0.0;
  for (int row = blockIdx.x; row < FW_NP; row += gridDim.x) {
    const double Pn = ::dP * (row - 0.5 * ::P_gridDIM);
    const double dK = ::fw_diff_K(Pn, tnow), Kp = ::K(Pn, tnow);
    double rs = 0.0;
    for (int j = tid; j < FW_NX; j += HT) {
      const double w = (double)a.W[(size_t)row * FW_NX + j];
      const double X = ::dX * (j - 0.5 * ::X_gridDIM);
      rs += w;
      s[1] += X * w;
      s[3] += X * X * w;
      s[6] += ::fw_diff_V(X, tnow) * w;
      s[7] += ::V(X, tnow) * w;
      s[8] += w * w;
    }
    s[0] += rs;
    s[2] += Pn * rs;
    s[4] += Pn * Pn * rs;
    s[5] += dK * rs;
    s[7] += Kp * rs;
  }
  if (block_reduce_publish<HT>(s, red, a.partials, a.counter, blockIdx.x, gridDim.x, a.scal + SC_OUT)) {
    if (tid == 0) a.scal[SC_NORM] = 1.0 / (a.scal[SC_OUT] * (::dX * ::dP));
  }
}

// Start of a chunk of steps (one small block): the step times
// times[s] = times[s-1] + inc from times[0] = t0 when n > 0, the same
// sequence of IEEE additions the caller made (t += dt), so the same bits; with
// reset, the run flags (no stop, W fresh); then the field values of split
// potentials for steps s < nf. One launch replaces a host-to-device copy of
// the times and the flag writes.
FWD void begin_chunk(Args a, double t0, double inc, int n, int reset, int nf) {
  if (threadIdx.x == 0) {
    double* tm = const_cast<double*>(a.times);
    double tt = t0;
    for (int s = 0; s < n; ++s) {
      if (s > 0) tt = tt + inc;
      tm[s] = tt;
    }
    if (reset) {
      a.scal[SC_STOP] = 0.0;
      a.scal[SC_FRESH] = 1.0;
    }
  }
  __syncthreads();
  a.step = nf;
  field_fill(a);
}

// W *= norm: the norm is scal[SC_NORM], or, at the end of a run
// (F_FROM_PARTIALS), that of the last step, whose sums blocks 0..NSER-1
// publish to series[step] (the grid has at least NSER blocks). F_NO_SCALE
// publishes and leaves W as it is.
FWD void apply_norm(const Args& a) {
  __shared__ double red[16];
  const bool scale = !(a.flags & F_NO_SCALE);
  const size_t n = (size_t)FW_NX * FW_NP, stride = (size_t)gridDim.x * HT;
  size_t i0 = (size_t)blockIdx.x * HT + threadIdx.x;
  // the first eight elements' loads ahead of the norm's reduction
  real_t w[8];
  if (scale) {
#pragma unroll
    for (int j = 0; j < 8; ++j) w[j] = (i0 + j * stride < n) ? a.W[i0 + j * stride] : (real_t)0;
  }
  double c;
  if (a.flags & F_FROM_PARTIALS) {
    const double S0 = sum_partials<HT, 1>(a.partials, a.nbp, 0, red);
    c = 1.0 / (S0 * (::dX * ::dP));
    if (blockIdx.x < NSER) {
      const double tot =
          blockIdx.x == 0 ? S0 : sum_partials<HT, 1>(a.partials + (size_t)blockIdx.x * a.nbp, a.nbp, 0, red);
      if (threadIdx.x == 0) a.series[(size_t)a.step * NSER + blockIdx.x] = tot;
    }
  } else {
    c = a.scal[SC_NORM];
  }
  if (!scale) return;
  const real_t cr = (real_t)c;
  // eight loads, then eight stores: a load after a store through asm cannot
  // be moved above it, so interleaved they would wait one latency each
  for (bool first = true; i0 < n; i0 += 8 * stride, first = false) {
    if (!first) {
#pragma unroll
      for (int j = 0; j < 8; ++j) w[j] = (i0 + j * stride < n) ? a.W[i0 + j * stride] : (real_t)0;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (i0 + j * stride < n) st_w(&a.W[i0 + j * stride], w[j] * cr);
  }
}

// Cached factor tables for time-independent V/K/masks (the reference caches
// _expV/_expK the same way, wigner_moyal_fftw_1d.py:164-213). Entries are the
// exact values the on-the-fly path computes. F_BLOCH selects the Bloch factors.
FWD void build_vtab(const Args& a, cpx* tab) {
  const size_t n = (size_t)(FW_NP / 2 + 1) * FW_NX;
  const double tnow = a.times[0], tau = tnow + FW_PHASE_SHIFT * ::dt;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(idx / FW_NX), j = (int)(idx % FW_NX);
    const double X = ::dX * (j - 0.5 * ::X_gridDIM), Th = ::dTheta * k;
    tab[vtab_index(j, k)] = (a.flags & F_BLOCH) ? bloch_vfactor(a, X, Th) : vfactor_tab<true>(X, Th, tau, tnow);
  }
}
FWD void build_ktab(const Args& a, cpx* tab) {
  constexpr int KW = FW_NX / 2 + 1;
  const size_t n = (size_t)FW_NP * KW;
  const double tnow = a.times[0], tau = tnow + FW_PHASE_SHIFT * ::dt;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx / KW), k = (int)(idx % KW);
    const double P = ::dP * (i - 0.5 * ::P_gridDIM), Lm = ::dLambda * k;
    tab[idx] = (a.flags & F_BLOCH) ? bloch_kfactor(a, Lm, P) : kfactor<true>(Lm, P, tau, tnow);
  }
}

// Largest |generated - table| K factor over the grid (the rows of the K
// pass's blocks, each thread its items), as the bits of a non-negative double
// maxed into scal[SC_OUT] (cleared by the host); NaN counts as infinite. Lines
// that are not pair-fused report infinity: their passes read the table.
FWD void kgen_check(const Args& a) {
  using L = KL;
  constexpr int N = FW_NX, E = FW_ROW_E, P = FW_ROW_P, TPL = L::TPL, KW = FW_NX / 2 + 1;
  double d = __builtin_inf();
  if constexpr (L::PAIRED && L::BS == 1) {
    const int t = threadIdx.x % TPL, q = threadIdx.x / TPL;
    const int r = (blockIdx.x * P + q) * 2;
    const double tnow = a.times[0], tau = tnow + FW_PHASE_SHIFT * ::dt;
    d = 0.0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      cpx f[E / 2 + 1];
      kgen_factors<L>(::dP * (r + h - 0.5 * ::P_gridDIM), tau, t, f);
#pragma unroll
      for (int i = 0; i <= E / 2; ++i) {
        const int k = i < E / 2 ? spec_k<L>(t, i) : N / 2;
        const cpx g = a.ktab[(size_t)(r + h) * KW + k];
        const double e = fabs((double)f[i].x - (double)g.x) + fabs((double)f[i].y - (double)g.y);
        d = (e <= d) ? d : (e == e ? e : __builtin_inf());
      }
    }
  }
  unsigned long long* out = reinterpret_cast<unsigned long long*>(a.scal + SC_OUT);
  if constexpr (L::NT % 64 == 0) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) d = fmax(d, __shfl_xor(d, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(d));
  } else {
    atomicMax(out, (unsigned long long)__double_as_longlong(d));
  }
}

// Minima of V(X-Th/2)+V(X+Th/2) over the theta-x grid and of K(P+L/2)+K(P-L/2)
// over the p-lambda grid at t_initial (fill_V_K + gpuarray.min,
// wigner_bloch_cuda_1d.py:72-102,362-445). One block; result in scal[SC_OUT..+1].
FWD void vk_min(const Args& a) {
  __shared__ double rv[16], rk[16];
  const int tid = threadIdx.x, nt = blockDim.x;
  const double ti = a.t_initial;
  double mv = __builtin_inf(), mkk = __builtin_inf();
  constexpr int NTH = FW_NP / 2 + 1, NLA = FW_NX / 2 + 1;
  for (int idx = tid; idx < NTH * FW_NX; idx += nt) {
    const int i = idx / FW_NX, j = idx % FW_NX;
    const double X = ::dX * (j - 0.5 * ::X_gridDIM), Th = ::dTheta * i;
    mv = fmin(mv, ::V(X - 0.5 * Th, ti) + ::V(X + 0.5 * Th, ti));
  }
  for (int idx = tid; idx < FW_NP * NLA; idx += nt) {
    const int i = idx / NLA, j = idx % NLA;
    const double Lm = ::dLambda * j, P = ::dP * (i - 0.5 * ::P_gridDIM);
    mkk = fmin(mkk, ::K(P + 0.5 * Lm, ti) + ::K(P - 0.5 * Lm, ti));
  }
  mv = wave_min(mv);
  mkk = wave_min(mkk);
  if ((tid & 63) == 0) { rv[tid >> 6] = mv; rk[tid >> 6] = mkk; }
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < (nt + 63) / 64; ++w) { mv = fmin(mv, rv[w]); mkk = fmin(mkk, rk[w]); }
    a.scal[SC_OUT] = mv;
    a.scal[SC_OUT + 1] = mkk;
  }
}

}  // namespace fwk

// ---------------------------------------------------------------------------
// Entry points (looked up by name by the host runtime), in sets compiled on
// demand: FW_KSET 1 = Moyal step + helpers, 2 = Bloch step, 3 = Moyal step with
// the random-collision dissipator, 0 = none (expression modules only use the
// helpers above).
// ---------------------------------------------------------------------------
#ifndef FW_KSET
#define FW_KSET 1
#endif
#define FW_VK(name, FL) \
  extern "C" __global__ void __launch_bounds__(fwk::VL::NT) name(fwk::Args a) { fwk::vpass<FL>(a); }
#define FW_KK(name, FL) \
  extern "C" __global__ void __launch_bounds__(fwk::KL::NT) name(fwk::Args a) { fwk::kpass<FL>(a); }

#if FW_KSET == 1
FW_VK(fw_v_first, fwk::F_FIRST)
FW_VK(fw_v_last, fwk::F_LAST)
FW_KK(fw_k, 0)
FW_KK(fw_k_norm, fwk::F_KNORM)
extern "C" __global__ void __launch_bounds__(fwk::VL::NT) fw_vv(fwk::Args a) {
  if constexpr (FW_THETA_SUMS) fwk::vvpass_theta(a);
  else fwk::vvpass(a);
}
extern "C" __global__ void __launch_bounds__(256) fw_build_sumw(fwk::Args a, fwk::SumW* sw) { fwk::build_sumw(a, sw); }
extern "C" __global__ void __launch_bounds__(128) fw_begin(fwk::Args a, double t0, double inc, int n, int reset,
                                                          int nf) {
  fwk::begin_chunk(a, t0, inc, n, reset, nf);
}
extern "C" __global__ void __launch_bounds__(fwk::HT) fw_moments(fwk::Args a) { fwk::moments(a); }
extern "C" __global__ void __launch_bounds__(fwk::HT) fw_apply_norm(fwk::Args a) { fwk::apply_norm(a); }
extern "C" __global__ void fw_set_scalar(double* p, double v) { *p = v; }
extern "C" __global__ void __launch_bounds__(1024) fw_vk_min(fwk::Args a) { fwk::vk_min(a); }
extern "C" __global__ void __launch_bounds__(256) fw_build_vtab(fwk::Args a, fwk::cpx* tab) { fwk::build_vtab(a, tab); }
extern "C" __global__ void __launch_bounds__(256) fw_build_ktab(fwk::Args a, fwk::cpx* tab) { fwk::build_ktab(a, tab); }
extern "C" __global__ void __launch_bounds__(fwk::KL::NT) fw_kgen_check(fwk::Args a) { fwk::kgen_check(a); }
#elif FW_KSET == 2
FW_VK(fw_bv_first, fwk::F_FIRST | fwk::F_BLOCH)
FW_VK(fw_bv_last, fwk::F_LAST | fwk::F_BLOCH)
FW_KK(fw_bk, fwk::F_BLOCH)
#elif FW_KSET == 3
FW_VK(fw_v_first_dissip, fwk::F_FIRST | fwk::F_DISSIP)
FW_VK(fw_v_last_dissip, fwk::F_LAST | fwk::F_DISSIP)
#endif
