// fw_rho.h -- gfx950 device code of the density-matrix propagators
// (von Neumann and Bloch), RhoVNeumannCUDA1D / RhoBlochCUDA1D of the reference.
//
// Generated prelude (before this header): FW_REAL, FW_FP32, FW_N, FW_ROW_E,
// FW_ROW_P, FW_COL_E, FW_COL_P, FW_PHASE_SHIFT, FW_HAS_MASK_X, FW_HAS_MASK_P,
// #include "fw_common.h", the constants (X_gridDIM, X_amplitude, dt, dX, dP) and
// the spliced device functions V(X,t), K(P,t), fw_diff_V, fw_diff_K,
// fw_mask_x(X, X_prime, t), fw_mask_p(P, P_prime, t).
//
// State: rho[i][j] = <X_j| rho |X'_i>, complex N x N, row-major
// (rho_vneumann_cuda_1d.py:222,672-684). One step
//   rho *= expV; FFT2; rho *= expK; IFFT2; rho *= expV; rho /= Tr(rho) dX
// (rho_vneumann_cuda_1d.py:340-374) is three fused passes, each one read and one
// write of rho:
//   A (rows):    deferred norm, expV, forward FFT along j
//   B (columns): forward FFT along i, expK, inverse FFT along i
//   C (rows):    inverse FFT along j, expV, per-block partial sums (trace,
//                <x>, <x^2>, <V'>, <V>, sum |rho|^2)
// Both factors are separable: expV[i][j] = u_j conj(u_i) with
// u_j = exp(-i dt/2 V(X_j, tau)) and expK[k'][k] = w_k conj(w_k') with
// w_k = exp(-i dt K(P_k, tau)) (FFT-ordered momenta, :657-658), so only 2N phases
// per step are evaluated, into per-step rows computed once per chunk of steps.
// The Bloch factors exp(-db/4 [V(X)+V(X')-Vmin]) and exp(-db/2 [K(P)+K(P')-Kmin])
// (rho_bloch_cuda_1d.py:271-309) factor the same way with real rows.
#pragma once

#include "fw_common.h"

namespace fwk {

// published sums (slots of the series): Re/Im of Tr rho, sum rho_ii X_i,
// sum rho_ii X_i^2, sum rho_ii V'(X_i), sum rho_ii V(X_i); sum |rho|^2;
// then (diagonal pass) Re/Im of Tr(rho P), Tr(rho P^2), Tr(rho K'(P)), Tr(rho K(P))
enum { RSER = 20, R_TR = 0, R_X = 2, R_X2 = 4, R_DV = 6, R_V = 8, R_PUR = 10, R_P = 11, R_P2 = 13, R_DK = 15,
       R_K = 17 };
enum { RC_FRESH = 2, RC_STOP = 1 };
enum { RF_FIRST = 1, RF_BLOCH = 4 };

struct RArgs {
  cpx* rho;            // state (or the work buffer of the Wigner transform)
  cpx* aux;            // second buffer (Wigner transform target, diagonal partials)
  const tw_t* tw;      // exp(-2 pi i j / N)
  const double* times; // t after the increment, per step
  double* scal;        // device scalars
  double* partials;    // per-block partial sums (RSER per block)
  double* series;      // RSER sums per step
  const cpx* uv;       // per-step rows u_j (unitary) or real Bloch rows
  const cpx* uk;       // per-step rows w_k
  const double* dw;    // momentum weights G_o[d] (re, im), o = P, P^2, K', K, or null
  int step;
  int flags;
  int nbp;             // number of partial rows written by pass C
  int nrows;           // phase rows available (chunk)
  int pubnb;           // > 0: block 0 of pass A publishes step - 1's momentum shares (pubnb blocks)
  double dbeta, vmin, kmin, t_initial, guard;
};

FWD double xcoord(int j) { return ::dX * (j - 0.5 * ::X_gridDIM); }
FWD double pcoord(int k) { return ::dP * ((k + FW_N / 2) % FW_N - 0.5 * FW_N); }

// The partial sums of pass C are stored by sum (write_partial_rows): sum r of
// block b at partials[r * nbp + b], so each row the next pass adds is nbp
// contiguous doubles.
// The trace rows of the previous step's partials (nbp = N / FW_ROW_P blocks of
// pass C), loaded with the pass's other loads and added in sum_partials'
// order, so that the norm is the same bits without a dependent load in the
// middle of the pass.
template <int NT>
struct TraceRows {
  static constexpr int NBP = FW_N / FW_ROW_P;
  static constexpr int PL = (NBP + NT - 1) / NT;
  double re[PL], im[PL];
  FWD void load(const double* __restrict__ partials) {
#pragma unroll
    for (int j = 0; j < PL; ++j) {
      const int b = threadIdx.x + j * NT;
      re[j] = b < NBP ? partials[(size_t)R_TR * NBP + b] : 0.0;
      im[j] = b < NBP ? partials[(size_t)(R_TR + 1) * NBP + b] : 0.0;
    }
  }
  FWD static double comp_sum(const double (&x)[PL]) {
    double sum = 0.0, comp = 0.0;
#pragma unroll
    for (int j = 0; j < PL; ++j) {
      if (threadIdx.x + j * NT < NBP) {
        const double tt = sum + x[j];
        comp += (fabs(sum) >= fabs(x[j])) ? (sum - tt) + x[j] : (x[j] - tt) + sum;
        sum = tt;
      }
    }
    return sum + comp;
  }
  // the two traces (every block the same bits); red: >= 2 BlockRed<NT>::RN
  FWD void totals(double* red, double& tr, double& ti) const {
    __syncthreads();
    BlockRed<NT>::put(comp_sum(re), red, 2, 0);
    BlockRed<NT>::put(comp_sum(im), red, 2, 1);
    __syncthreads();
    tr = BlockRed<NT>::total(red, 2, 0);
    ti = BlockRed<NT>::total(red, 2, 1);
  }
};

// Complex norm of the previous step from the trace rows (the same bits in
// every block). With the purity guard armed, `stop` tells whether the step
// must not run; the caller then publishes the sums (rho_publish) and returns.
template <int NT>
FWD void rho_previous(const RArgs& a, double* red, cpx& nrm, bool& stop, const TraceRows<NT>* pre = nullptr) {
  double tr_re, tr_im;
  if (pre) {
    pre->totals(red, tr_re, tr_im);
  } else {
    tr_re = sum_partials<NT, 1>(a.partials + (size_t)R_TR * a.nbp, a.nbp, 0, red);
    tr_im = sum_partials<NT, 1>(a.partials + (size_t)(R_TR + 1) * a.nbp, a.nbp, 0, red);
  }
  // c = 1 / (Tr rho * dX), complex, as rho /= trace(rho) * dX (rho_vneumann_cuda_1d.py:369)
  const double d = (tr_re * tr_re + tr_im * tr_im) * ::dX;
  nrm = mk((real_t)(tr_re / d), (real_t)(-tr_im / d));
  stop = false;
  if (a.guard > 0.0) {
    const double pur = sum_partials<NT, 1>(a.partials + (size_t)R_PUR * a.nbp, a.nbp, 0, red);
    const double purity = pur * ::dX * ::dX / ((tr_re * tr_re + tr_im * tr_im) * ::dX * ::dX);
    stop = !(purity < a.guard);
  }
}

// The previous step's sums 0..R_PUR to series[step]: block b of nb adds rows
// b, b + nb, ...; block 0 raises the purity guard. Called by whole blocks,
// after their own work.
template <int NT>
FWD void rho_publish(const RArgs& a, int blk, int nb, double* red, bool stop) {
  for (int r = blk; r <= R_PUR; r += nb) {
    const double tot = sum_partials<NT, 1>(a.partials + (size_t)r * a.nbp, a.nbp, 0, red);
    if (threadIdx.x == 0) a.series[(size_t)a.step * RSER + r] = tot;
  }
  if (stop && blk == 0 && threadIdx.x == 0) a.scal[RC_STOP] = (double)(a.step + 1);
}

FWD cpx rowfac(const cpx* __restrict__ u, int s, int j) { return u[(size_t)s * FW_N + j]; }

// ---------------------------------------------------------------------------
// Pass A / C: rows (along j). PASS_A: [norm], expV, forward FFT.
// PASS_C: inverse FFT, expV, sums.
// ---------------------------------------------------------------------------
// FP64 steps compute their twiddles in the kernel (one sine/cosine per
// butterfly, powers by products: no table staged in LDS, as the Wigner passes
// do); the FP32 form keeps the staged compensated table.
#ifndef FW_RHO_TW
#define FW_RHO_TW (FW_FP32 ? 0 : 2)
#endif
using RL = Line<FW_N, FW_ROW_E, FW_ROW_P, true, FW_N, FW_RHO_TW>;
using CL = Line<FW_N, FW_COL_E, FW_COL_P, false, FW_N, FW_RHO_TW>;

// The 8 momentum sums of step s from the diagonal pass's shares (8 rows of
// nshare in aux), the work of fw_rho_diag_publish, done by the last blocks of
// the next step's pass A after their own rows (sum o by block nb - 1 - o, so
// that these blocks are not the ones that publish the x-family sums), so that
// the step needs no launch for it and no block delays its own loads.
template <int NT>
FWD void publish_shares(const RArgs& a, double* red, int nshare, int s, int nb) {
  const double* part = reinterpret_cast<const double*>(a.aux);
  for (int o = nb - 1 - (int)blockIdx.x; o >= 0 && o < 8; o += nb) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nshare; b += NT) v += part[(size_t)o * nshare + b];
    BlockRed<NT>::put(v, red, 1, 0);
    __syncthreads();
    if (threadIdx.x == 0) a.series[(size_t)s * RSER + R_P + o] = BlockRed<NT>::total(red, 1, 0);
    __syncthreads();   // red is reused
  }
}

template <bool PASS_A, int FL>
FWD void rho_row(const RArgs& a) {
  using L = RL;
  constexpr int N = FW_N, E = FW_ROW_E, P = FW_ROW_P, TPL = L::TPL, NT = L::NT;
  constexpr int TWN = (N + NT - 1) / NT;
  constexpr bool STAGE = L::TW::COUNT > 0;   // a twiddle table in LDS (FP32)
  __shared__ cpx lds[L::LDS];
  __shared__ tw_t twl[STAGE ? TwLds<N>::SIZE : 1];
  __shared__ double red[RSER * 32];
  static_assert(8 * BlockRed<NT>::RN <= RSER * 32, "reduction buffer");
  if (a.scal[RC_STOP] != 0.0) return;
  const int tid = threadIdx.x;
  const int t = tid % TPL, q = tid / TPL;
  const int nb = gridDim.x;
  const int blk = xcd_remap(blockIdx.x, nb);
  const int i = blk * P + q;                 // row = x'
  cpx* __restrict__ row = a.rho + (size_t)i * N;
  const double tnow = a.times[a.step];

  tw_t twr[TWN];
  if constexpr (STAGE) {
#pragma unroll
    for (int r = 0; r < TWN; ++r) twr[r] = a.tw[(tid + r * NT) % N];
  }
  cpx v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = row[t + TPL * e];
  cpx fj[E];
#pragma unroll
  for (int e = 0; e < E; ++e) fj[e] = rowfac(a.uv, a.step, t + TPL * e);
  const cpx fi = cconj(rowfac(a.uv, a.step, i));
  // pass A after a step: the previous step's trace rows with the row's loads
  TraceRows<NT> trows;
  if constexpr (PASS_A && (FL & RF_FIRST)) trows.load(a.partials);
  if constexpr (STAGE) {
#pragma unroll
    for (int r = 0; r < TWN; ++r)
      if ((N % NT) == 0 || tid + r * NT < N) twl[twpad(tid + r * NT)] = twr[r];
  }
  const double Xi = xcoord(i);

  bool deferred = false;   // pass A of a step after the first: the previous step's norm and sums
  if constexpr (PASS_A) {
    cpx nrm = mk((real_t)1, (real_t)0);
    if constexpr (FL & RF_FIRST) {
      if (a.scal[RC_FRESH] == 0.0) {
        deferred = true;
        bool stop;
        rho_previous<NT>(a, red, nrm, stop, &trows);
        if (stop) {
          rho_publish<NT>(a, blk, nb, red, true);
          if (a.pubnb > 0) publish_shares<NT>(a, red, a.pubnb, a.step - 1, nb);
          return;
        }
      } else {
        __syncthreads();
      }
    } else {
      __syncthreads();
    }
    const cpx g = cmul(nrm, fi);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      cpx f = cmul(fj[e], g);
#if FW_HAS_MASK_X
      if (!(FL & RF_BLOCH)) f = cscale(f, (real_t)::fw_mask_x(xcoord(t + TPL * e), Xi, tnow));
#endif
      v[e] = cmul(v[e], f);
    }
    L::template fwd<false>(v, lds, q, t, twl);
#pragma unroll
    for (int e = 0; e < E; ++e) st_w(&row[t + TPL * e], v[e]);
    // the previous step's sums, after this block's own work
    if (deferred) rho_publish<NT>(a, blk, nb, red, false);
    if (a.pubnb > 0) publish_shares<NT>(a, red, a.pubnb, a.step - 1, nb);
  } else {
    __syncthreads();
    L::inv(v, lds, q, t, twl);
    const real_t sN = (real_t)1 / (real_t)N;
    double s[RSER];
#pragma unroll
    for (int r = 0; r < RSER; ++r) s[r] = 0.0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int j = t + TPL * e;
      cpx f = cmul(fj[e], fi);
#if FW_HAS_MASK_X
      if (!(FL & RF_BLOCH)) f = cscale(f, (real_t)::fw_mask_x(xcoord(j), Xi, tnow));
#endif
      cpx w = cmul(cscale(v[e], sN), f);
      st_w(&row[j], w);
      const double wr = (double)w.x, wi = (double)w.y;
      s[R_PUR] += wr * wr + wi * wi;
      if (j == i) {
        const double dv = ::fw_diff_V(Xi, tnow), vv = ::V(Xi, tnow);
        s[R_TR] += wr; s[R_TR + 1] += wi;
        s[R_X] += wr * Xi; s[R_X + 1] += wi * Xi;
        s[R_X2] += wr * Xi * Xi; s[R_X2 + 1] += wi * Xi * Xi;
        s[R_DV] += wr * dv; s[R_DV + 1] += wi * dv;
        s[R_V] += wr * vv; s[R_V + 1] += wi * vv;
      }
    }
    write_partial_rows<NT, RSER>(s, red, a.partials, blk, nb);
  }
}

// ---------------------------------------------------------------------------
// Pass B: columns (along i): forward FFT, expK, inverse FFT.
// ---------------------------------------------------------------------------
template <int FL>
FWD void rho_col(const RArgs& a) {
  using L = CL;
  constexpr int N = FW_N, E = FW_COL_E, P = FW_COL_P, TPL = L::TPL, NT = L::NT;
  constexpr int TWN = (N + NT - 1) / NT;
  constexpr bool STAGE = L::TW::COUNT > 0;   // a twiddle table in LDS (FP32)
  __shared__ cpx lds[L::LDS];
  __shared__ tw_t twl[STAGE ? TwLds<N>::SIZE : 1];
  if (a.scal[RC_STOP] != 0.0) return;
  const int tid = threadIdx.x;
  const int q = tid % P, t = tid / P;
  const int nb = gridDim.x;
  const int blk = xcd_remap(blockIdx.x, nb);
  const int j = blk * P + q;                 // column = momentum index of x
  if (blk == 0 && tid == 0 && a.scal[RC_FRESH] != 0.0) a.scal[RC_FRESH] = 0.0;
  const double tnow = a.times[a.step];

  tw_t twr[TWN];
  if constexpr (STAGE) {
#pragma unroll
    for (int r = 0; r < TWN; ++r) twr[r] = a.tw[(tid + r * NT) % N];
  }
  cpx v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = a.rho[(size_t)(t + TPL * e) * N + j];
  cpx wk[E];
#pragma unroll
  for (int e = 0; e < E; ++e) wk[e] = cconj(rowfac(a.uk, a.step, t + TPL * e));
  const cpx wj = rowfac(a.uk, a.step, j);
  if constexpr (STAGE) {
#pragma unroll
    for (int r = 0; r < TWN; ++r)
      if ((N % NT) == 0 || tid + r * NT < N) twl[twpad(tid + r * NT)] = twr[r];
  }
  __syncthreads();
  L::template fwd<false>(v, lds, q, t, twl);
  const real_t sN = (real_t)1 / (real_t)N;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    cpx f = cscale(cmul(wj, wk[e]), sN);
#if FW_HAS_MASK_P
    if (!(FL & RF_BLOCH)) f = cscale(f, (real_t)::fw_mask_p(pcoord(j), pcoord(t + TPL * e), tnow));
#endif
    v[e] = cmul(v[e], f);
  }
  L::sync();
  // the lane index passes through an empty asm, so that the compiler
  // recomputes the inverse's twiddles and LDS indices instead of keeping the
  // forward transform's live across it (register pressure)
  int t2 = t;
  asm volatile("" : "+v"(t2));
  L::inv(v, lds, q, t2, twl);
#pragma unroll
  for (int e = 0; e < E; ++e) st_w(&a.rho[(size_t)(t + TPL * e) * N + j], v[e]);
}

// ---------------------------------------------------------------------------
// Per-step phase rows for `nrows` steps: times[s], s < nrows.
// Unitary: uv = exp(-i dt/2 V(X_j, tau)), uk = exp(-i dt K(P_k, tau)).
// Bloch:   uv = exp(-db/4 (V(X_j) - Vmin/2)), uk = exp(-db/2 (K(P_k) - Kmin/2)) at t_initial.
// ---------------------------------------------------------------------------
FWD void rho_rows(const RArgs& a, cpx* uv, cpx* uk) {
  const size_t n = (size_t)a.nrows * FW_N;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    const int s = (int)(idx / FW_N), j = (int)(idx % FW_N);
    const double X = xcoord(j), P = pcoord(j);
    if (a.flags & RF_BLOCH) {
      const double ti = a.t_initial;
      uv[idx] = mk((real_t)exp(-0.25 * a.dbeta * (::V(X, ti) - 0.5 * a.vmin)), (real_t)0);
      uk[idx] = mk((real_t)exp(-0.5 * a.dbeta * (::K(P, ti) - 0.5 * a.kmin)), (real_t)0);
    } else {
      const double tau = a.times[s] + FW_PHASE_SHIFT * ::dt;
      uv[idx] = unit_phase_exact(-0.5 * ::dt * ::V(X, tau), 1.0);
      uk[idx] = unit_phase_exact(-::dt * ::K(P, tau), 1.0);
    }
  }
}

// ---------------------------------------------------------------------------
// Momentum observables: c_d = sum_i rho[i][(i - d) mod N] over row tiles, then
// n_k = (1/N) sum_d c_d exp(2 pi i k d / N) and Tr(rho g(P)) = sum_k g(P_k) n_k,
// which is the reference's fft(axis 1)/N -> g(P) -> ifft -> trace
// (rho_vneumann_cuda_1d.py:452-472).
// Partial kernel: block (g, c) sums diagonals c*DC .. c*DC + DC - 1 over row
// group g (N/G rows); lane l of a DC-wide thread row takes diagonal c*DC + l, so
// a thread row reads DC consecutive (wrapped) elements of one row of rho.
// The row subgroups of a block are combined in LDS, in a fixed order, and the
// G group sums land in aux[g * N + d] for the one-block finish.
// ---------------------------------------------------------------------------
#ifndef FW_DIAG_W
#define FW_DIAG_W 0   // 1: momentum sums by the weights G (K and K' independent of t)
#endif
constexpr int DT = 256;   // threads of the diagonal partial kernel
constexpr int DDC = FW_DIAG_DC, DW = DT / DDC, DG = FW_DIAG_G, DRG = FW_N / DG;
static_assert(DT % DDC == 0 && FW_N % DDC == 0 && FW_N % DG == 0, "diagonal tiling");

FWD void rho_diag_partial(const RArgs& a) {
  __shared__ double red[2 * DT];
  constexpr int NC = FW_N / DDC;
  const int l = threadIdx.x % DDC, w = threadIdx.x / DDC;
  const int g = blockIdx.x / NC, d = (blockIdx.x % NC) * DDC + l;
  double cr = 0.0, ci = 0.0;
  if (w < DRG) {
    constexpr int RT = (DRG + DW - 1) / DW;   // rows per thread (the last ones may be absent)
#pragma unroll 8
    for (int r = 0; r < RT; ++r) {
      const int i = g * DRG + w + r * DW;
      if ((DRG % DW) == 0 || w + r * DW < DRG) {
        const cpx z = a.rho[(size_t)i * FW_N + (i - d + FW_N) % FW_N];
        cr += (double)z.x;
        ci += (double)z.y;
      }
    }
  }
  red[threadIdx.x] = cr;
  red[DT + threadIdx.x] = ci;
  __syncthreads();
  if (w == 0) {
    for (int k = 1; k < DW; ++k) {
      cr += red[k * DDC + l];
      ci += red[DT + k * DDC + l];
    }
#if !FW_DIAG_W
    a.aux[(size_t)g * FW_N + d] = mk((real_t)cr, (real_t)ci);
#endif
  }
#if FW_DIAG_W
  // Tr(rho g(P)) = sum_d c_d G_d with G_d = (1/N) sum_k g(P_k) exp(2 pi i k d / N)
  // (K and K' independent of t): this block's share over its diagonals; the
  // next kernel (rho_diag_publish) adds the shares and publishes them.
  __syncthreads();
  double* dr = red;   // 8 rows of DDC lanes (8 DDC <= 2 DT)
  if (w == 0) {
#pragma unroll
    for (int o = 0; o < 4; ++o) {
      const double gr = a.dw[((size_t)o * FW_N + d) * 2], gi = a.dw[((size_t)o * FW_N + d) * 2 + 1];
      dr[(2 * o) * DDC + l] = cr * gr - ci * gi;
      dr[(2 * o + 1) * DDC + l] = cr * gi + ci * gr;
    }
  }
  __syncthreads();
  double* part = reinterpret_cast<double*>(a.aux);   // 8 rows of gridDim.x shares
  if (threadIdx.x < 8) {
    double t = 0.0;
    for (int k = 0; k < DDC; ++k) t += dr[threadIdx.x * DDC + k];
    part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = t;
  }
#endif
}

// The momentum sums of the weights form: one block adds the nb shares of each
// of the 8 rows (every thread a strided subset, then the block; fixed order)
// and publishes them to series[step].
FWD void rho_diag_publish(const RArgs& a, int nb) {
  __shared__ double red[8 * BlockRed<DT>::RN];
  const double* part = reinterpret_cast<const double*>(a.aux);
  double v[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nb; b += DT) {
#pragma unroll
    for (int o = 0; o < 8; ++o) v[o] += part[(size_t)o * nb + b];
  }
#pragma unroll
  for (int o = 0; o < 8; ++o) BlockRed<DT>::put(v[o], red, 8, o);
  __syncthreads();
  if (threadIdx.x < 8) a.series[(size_t)a.step * RSER + R_P + threadIdx.x] = BlockRed<DT>::total(red, 8, threadIdx.x);
}

// Momentum weights G_o[d] = (1/N) sum_k g_o(P_k) exp(2 pi i k d / N) for
// g = P, P^2, K'(P), K(P) (K independent of t), by direct summation in FP64
// with the exact angle index (k d) mod N.
FWD void rho_diag_weights(const RArgs& a, double* dw) {
  const double tk = a.t_initial;
  for (int d = blockIdx.x * blockDim.x + threadIdx.x; d < FW_N; d += gridDim.x * blockDim.x) {
    double acc[4][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (int k = 0; k < FW_N; ++k) {
      const int m = (int)(((long long)k * d) % FW_N);
      double sn, cs;
      sincospi(2.0 * (double)m / (double)FW_N, &sn, &cs);
      const double P = pcoord(k);
      const double g[4] = {P, P * P, ::fw_diff_K(P, tk), ::K(P, tk)};
#pragma unroll
      for (int o = 0; o < 4; ++o) {
        acc[o][0] += g[o] * cs;
        acc[o][1] += g[o] * sn;
      }
    }
#pragma unroll
    for (int o = 0; o < 4; ++o) {
      dw[((size_t)o * FW_N + d) * 2] = acc[o][0] / FW_N;
      dw[((size_t)o * FW_N + d) * 2 + 1] = acc[o][1] / FW_N;
    }
  }
}

template <int E>
FWD void rho_diag_finish(const RArgs& a) {
  // one block of N/E threads (a single line of the FFT engine)
  using L = Line<FW_N, E, 1, true>;
  constexpr int TPL = L::TPL;
  __shared__ cpx lds[L::LDS];
  __shared__ tw_t twl[TwLds<FW_N>::SIZE];
  __shared__ double red[RSER * 32];
  const int t = threadIdx.x;
  for (int r = t; r < FW_N; r += TPL) twl[twpad(r)] = a.tw[r];
  cpx v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int d = t + TPL * e;
    double cr = 0.0, ci = 0.0;
#pragma unroll
    for (int b = 0; b < DG; ++b) {
      const cpx z = a.aux[(size_t)b * FW_N + d];
      cr += (double)z.x;
      ci += (double)z.y;
    }
    v[e] = mk((real_t)cr, (real_t)ci);
  }
  __syncthreads();
  L::inv(v, lds, 0, t, twl);   // sum_d c_d exp(+2 pi i k d / N)
  const double tnow = a.times[a.step];
  double s[RSER];
#pragma unroll
  for (int r = 0; r < RSER; ++r) s[r] = 0.0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t + TPL * e;
    const double P = pcoord(k);
    const double nr = (double)v[e].x / FW_N, ni = (double)v[e].y / FW_N;
    const double g[4] = {P, P * P, ::fw_diff_K(P, tnow), ::K(P, tnow)};
#pragma unroll
    for (int o = 0; o < 4; ++o) {
      s[R_P + 2 * o] += g[o] * nr;
      s[R_P + 2 * o + 1] += g[o] * ni;
    }
  }
  write_partials<TPL, RSER>(s, red, a.partials + (size_t)a.nbp * RSER, 0);
  __syncthreads();
  if (t < 8) a.series[(size_t)a.step * RSER + R_P + t] = a.partials[(size_t)a.nbp * RSER + R_P + t];
}

// ---------------------------------------------------------------------------
// rho *= c (complex): c from the partials of the last pass C (end of a run;
// block 0 also publishes that step's sums), or c = scal[0] + i scal[3].
// from_partials = 2 publishes the sums and leaves rho as it is (one bare step).
// ---------------------------------------------------------------------------
FWD void rho_apply_norm(const RArgs& a, int from_partials) {
  __shared__ double red[RSER * 4];
  cpx c;
  if (from_partials) {
    bool stop;   // the guard is off (guard = 0) at the end of a run
    rho_previous<256>(a, red, c, stop);
    rho_publish<256>(a, blockIdx.x, gridDim.x, red, false);
  } else {
    c = mk((real_t)a.scal[0], (real_t)a.scal[3]);
  }
  if (from_partials == 2) return;
  const size_t n = (size_t)FW_N * FW_N, stride = (size_t)gridDim.x * 256;
  // four loads, then four stores (a load after a store through asm cannot be
  // moved above it)
  for (size_t i0 = (size_t)blockIdx.x * 256 + threadIdx.x; i0 < n; i0 += 4 * stride) {
    cpx w[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) w[j] = (i0 + j * stride < n) ? a.rho[i0 + j * stride] : mk((real_t)0, (real_t)0);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (i0 + j * stride < n) st_w(&a.rho[i0 + j * stride], cmul(w[j], c));
  }
}

// Sums over the current rho (trace family + purity) into scal[8..8+R_PUR];
// scal[0] + i scal[3] = 1/(Tr rho dX). One block.
FWD void rho_moments(const RArgs& a) {
  __shared__ double red[RSER * 32];
  const int tid = threadIdx.x;
  const double tnow = a.times[a.step];
  double s[RSER];
#pragma unroll
  for (int r = 0; r < RSER; ++r) s[r] = 0.0;
  for (int i = tid; i < FW_N; i += 1024) {
    const cpx z = a.rho[(size_t)i * FW_N + i];
    const double Xi = xcoord(i), dv = ::fw_diff_V(Xi, tnow), vv = ::V(Xi, tnow);
    const double wr = (double)z.x, wi = (double)z.y;
    s[R_TR] += wr; s[R_TR + 1] += wi;
    s[R_X] += wr * Xi; s[R_X + 1] += wi * Xi;
    s[R_X2] += wr * Xi * Xi; s[R_X2 + 1] += wi * Xi * Xi;
    s[R_DV] += wr * dv; s[R_DV + 1] += wi * dv;
    s[R_V] += wr * vv; s[R_V + 1] += wi * vv;
  }
  const size_t n = (size_t)FW_N * FW_N;
  for (size_t idx = tid; idx < n; idx += 1024) {
    const cpx z = a.rho[idx];
    s[R_PUR] += (double)z.x * z.x + (double)z.y * z.y;
  }
  write_partials<1024, RSER>(s, red, a.scal + 8, 0);
  __syncthreads();
  if (tid == 0) {
    const double tr = a.scal[8 + R_TR], ti = a.scal[8 + R_TR + 1];
    const double d = (tr * tr + ti * ti) * ::dX;
    a.scal[0] = tr / d;
    a.scal[3] = -ti / d;
  }
}

// ---------------------------------------------------------------------------
// Wigner function of rho by the Larkin three-shear 45 degree rotation and the
// sign-flipped FFT (rho_vneumann_cuda_1d.py:502-592, kernels :753-800).
// MODE 0: shear along x (rows):    FFT_j, * exp(-i tan(pi/8) P_k X'_i), IFFT_j / N
// MODE 1: shear along x' (columns): FFT_i, * exp(+i sin(pi/4) P'_k X_j), IFFT_i / N
// MODE 2: columns: * (-1)^i, IFFT_i (unnormalised), * (-1)^k, block sums of the result
// ---------------------------------------------------------------------------
template <int MODE>
FWD void rho_wigner_pass(const RArgs& a, const cpx* __restrict__ src, cpx* __restrict__ dst) {
  using L = typename pick<MODE == 0, RL, CL>::type;
  constexpr int N = FW_N, E = (MODE == 0) ? FW_ROW_E : FW_COL_E, P = (MODE == 0) ? FW_ROW_P : FW_COL_P;
  constexpr int TPL = L::TPL, NT = L::NT;
  constexpr bool STAGE = L::TW::COUNT > 0;
  __shared__ cpx lds[L::LDS];
  __shared__ tw_t twl[STAGE ? TwLds<N>::SIZE : 1];
  __shared__ double red[RSER * 32];
  const int tid = threadIdx.x;
  const int t = (MODE == 0) ? tid % TPL : tid / P;
  const int q = (MODE == 0) ? tid / TPL : tid % P;
  const int line = blockIdx.x * P + q;
  if constexpr (STAGE)
    for (int r = tid; r < N; r += NT) twl[twpad(r)] = a.tw[r];
  auto at = [&](int e) -> size_t {
    const int n = t + TPL * e;
    return (MODE == 0) ? (size_t)line * N + n : (size_t)n * N + line;
  };
  cpx v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = src[at(e)];
  __syncthreads();
  if constexpr (MODE == 2) {
#pragma unroll
    for (int e = 0; e < E; ++e)
      if ((t + TPL * e) & 1) v[e] = mk(-v[e].x, -v[e].y);
    L::inv(v, lds, q, t, twl);
    double s[RSER];
#pragma unroll
    for (int r = 0; r < RSER; ++r) s[r] = 0.0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      if ((t + TPL * e) & 1) v[e] = mk(-v[e].x, -v[e].y);
      s[0] += (double)v[e].x;
      s[1] += (double)v[e].y;
      dst[at(e)] = v[e];
    }
    write_partials<NT, RSER>(s, red, a.partials, blockIdx.x);
  } else {
    L::template fwd<false>(v, lds, q, t, twl);
    const real_t sN = (real_t)1 / (real_t)N;
    const double coord = xcoord(line);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const double pk = pcoord(t + TPL * e);
      const double ph = (MODE == 0) ? -0.41421356237309504880 * pk * coord   // -tan(pi/8) P X'
                                    : 0.70710678118654752440 * pk * coord;   // +sin(pi/4) P' X
      v[e] = cmul(v[e], unit_phase_exact(ph, (double)sN));
    }
    L::sync();
    L::inv(v, lds, q, t, twl);
#pragma unroll
    for (int e = 0; e < E; ++e) dst[at(e)] = v[e];
  }
}

// W /= sum(W) * wigner_dxdp (complex), sums from the MODE 2 partials
FWD void rho_wigner_scale(const RArgs& a, cpx* w, double wigner_dxdp) {
  __shared__ double red[RSER * 4];
  const double sr = sum_partials<256, RSER>(a.partials, a.nbp, 0, red);
  const double si = sum_partials<256, RSER>(a.partials, a.nbp, 1, red);
  const double d = (sr * sr + si * si) * wigner_dxdp;
  const cpx c = mk((real_t)(sr / d), (real_t)(-si / d));
  const size_t n = (size_t)FW_N * FW_N;
  for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < n; idx += (size_t)gridDim.x * 256)
    w[idx] = cmul(w[idx], c);
}

}  // namespace fwk

#ifndef FW_KSET
#define FW_KSET 1
#endif
#if FW_KSET == 1
extern "C" __global__ void __launch_bounds__(fwk::RL::NT) fw_rho_a(fwk::RArgs a) { fwk::rho_row<true, fwk::RF_FIRST>(a); }
extern "C" __global__ void __launch_bounds__(fwk::CL::NT) fw_rho_b(fwk::RArgs a) { fwk::rho_col<0>(a); }
extern "C" __global__ void __launch_bounds__(fwk::RL::NT) fw_rho_c(fwk::RArgs a) { fwk::rho_row<false, 0>(a); }
extern "C" __global__ void __launch_bounds__(fwk::RL::NT) fw_rho_ba(fwk::RArgs a) {
  fwk::rho_row<true, fwk::RF_FIRST | fwk::RF_BLOCH>(a);
}
extern "C" __global__ void __launch_bounds__(fwk::CL::NT) fw_rho_bb(fwk::RArgs a) { fwk::rho_col<fwk::RF_BLOCH>(a); }
extern "C" __global__ void __launch_bounds__(fwk::RL::NT) fw_rho_bc(fwk::RArgs a) {
  fwk::rho_row<false, fwk::RF_BLOCH>(a);
}
extern "C" __global__ void __launch_bounds__(256) fw_rho_rows(fwk::RArgs a, fwk::cpx* uv, fwk::cpx* uk) {
  fwk::rho_rows(a, uv, uk);
}
extern "C" __global__ void __launch_bounds__(fwk::DT) fw_rho_diag_partial(fwk::RArgs a) {
  fwk::rho_diag_partial(a);
}
extern "C" __global__ void __launch_bounds__(fwk::DT) fw_rho_diag_publish(fwk::RArgs a, int nb) {
  fwk::rho_diag_publish(a, nb);
}
extern "C" __global__ void __launch_bounds__(256) fw_rho_diag_weights(fwk::RArgs a, double* dw) {
  fwk::rho_diag_weights(a, dw);
}
extern "C" __global__ void __launch_bounds__(FW_N / FW_DIAG_E) fw_rho_diag_finish(fwk::RArgs a) {
  fwk::rho_diag_finish<FW_DIAG_E>(a);
}
extern "C" __global__ void __launch_bounds__(256) fw_rho_apply_norm(fwk::RArgs a, int from_partials) {
  fwk::rho_apply_norm(a, from_partials);
}
extern "C" __global__ void __launch_bounds__(1024) fw_rho_moments(fwk::RArgs a) { fwk::rho_moments(a); }
extern "C" __global__ void __launch_bounds__(fwk::RL::NT) fw_rho_shear_x(fwk::RArgs a, const fwk::cpx* s, fwk::cpx* d) {
  fwk::rho_wigner_pass<0>(a, s, d);
}
extern "C" __global__ void __launch_bounds__(fwk::CL::NT) fw_rho_shear_y(fwk::RArgs a, const fwk::cpx* s, fwk::cpx* d) {
  fwk::rho_wigner_pass<1>(a, s, d);
}
extern "C" __global__ void __launch_bounds__(fwk::CL::NT) fw_rho_flip_ifft(fwk::RArgs a, const fwk::cpx* s, fwk::cpx* d) {
  fwk::rho_wigner_pass<2>(a, s, d);
}
extern "C" __global__ void __launch_bounds__(256) fw_rho_wigner_scale(fwk::RArgs a, fwk::cpx* w, double dxdp) {
  fwk::rho_wigner_scale(a, w, dxdp);
}
extern "C" __global__ void fw_rho_set_scalar(double* p, double v) { *p = v; }
// min over j of V(X_j, t_initial): V_min of the Bloch factor is twice it
// (fill_V + gpuarray.min over the X, X' grid, rho_bloch_cuda_1d.py:69-82,336-346)
extern "C" __global__ void __launch_bounds__(256) fw_rho_vmin(fwk::RArgs a) {
  __shared__ double r[4];
  double m = __builtin_inf();
  for (int j = threadIdx.x; j < FW_N; j += 256) m = fmin(m, ::V(fwk::xcoord(j), a.t_initial));
  m = fwk::wave_min(m);
  if ((threadIdx.x & 63) == 0) r[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) a.scal[8] = 2.0 * fmin(fmin(r[0], r[1]), fmin(r[2], r[3]));
}
#endif
