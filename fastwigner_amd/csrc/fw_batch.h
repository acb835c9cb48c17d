// fw_batch.h -- batched split-operator propagation of many independent Wigner
// functions on one gfx950 device: each instance has its own square grid
// (N x N, N a power of two in [64, 1024]), amplitudes, potential parameters,
// collision rate and temperature. Embedded in libfastwigner.so and compiled by
// hipRTC with a prelude holding:
//   FW_REAL, FW_FP32, FW_PHASE_SHIFT, FW_B_NMAX (largest N of the batch),
//   FW_B_HAS_<N> (1 for every N present), FW_B_E<N> (elements per thread of
//   the FFT of length N), FW_B_EMAX, the time step `dt`, the caller's constants,
//   and the spliced expressions
//     double V(double X, double t, const double* prm), K(P, t, prm),
//     fw_diff_V(X, t, prm), fw_diff_K(P, t, prm)
//   in which the caller's parameter names are bound to prm[0..7].
//
// Work decomposition. A step of one instance is the three passes of the single
// grid propagator (V columns, K rows, V columns; see fw_device.h), cut into
// tiles of P line pairs, one tile per 256-thread block. Tasks (pass, instance,
// tile) sit in eight queues, one per XCD: each instance's tasks go to one
// queue, so the column tiles that share 128-byte lines, and the passes that
// re-read the instance, stay on one XCD's L2. Queues are ordered pass-major.
//
// One launch per step (mode 0): a persistent grid takes tasks from its XCD's
// queue (then from the others), and a task of pass p of instance i waits until
// every tile of pass p-1 of instance i has been released (per-instance
// counters; agent-scope release/acquire). Tasks earlier in a queue never wait
// on later ones and a task is only taken by a running block, so the waits
// always end; each wait is also bounded (spin_limit), after which the block
// raises the error flag and every block drains out.
// One launch per pass (mode 1): grid = tasks of that pass, one task per block,
// no waits. In mode 1, when V does not depend on t, a real-time run takes two
// passes per step (BArgs.vv = 1): the K pass, which first applies the previous
// step's deferred norm and dissipator as W = alpha Y + G, and one fused column
// pass (b_vvpass) holding the closing V/2 of a step, its dissipator and sums,
// and the opening V/2 of the next step. With U the state after the closing V/2
// and dissipator and S its sum, the next step starts from w U / (S dXdP) + W0s,
// and the opening V/2 is linear, so the pass stores Y = Vopen(U) and the K pass
// forms (w / (S dXdP)) Y + G with G = Vopen(W0s) (b_gpass, once per source).
// A run opens with the first V pass and closes with the last V pass.
#pragma once

#include "fw_common.h"

namespace fwk {

constexpr int BT = 256;   // threads per block of the batched kernels

// One instance (host mirror in fw_batch.cpp; field order and types must match).
struct BInst {
  long long off;          // element offset of W and W0s
  int n;                  // grid dimension N (N x N)
  int ntile;              // tiles per pass (N / (2 P))
  int part;               // first partial row of the instance
  int dissip;             // 1: random-collision dissipator on
  int nbloch;             // Bloch runs: steps of this instance
  int swoff;              // first entry of this instance's theta-sum weights (BSumW)
  double dX, dP, dLambda, dTheta;
  double w_coeff;         // exp(-gamma dt / 2)
  double dbeta;           // Bloch step of this instance
  double prm[8];          // parameters of the spliced expressions
};

struct BTask {
  int inst;
  int code;               // pass << 16 | tile; inst < 0: no work (padding)
};

// Weights of the theta-spectrum sums of one instance, entry k <= N/2 (the
// layout of the single-grid SumW): 0 < k < N/2: a = Im 2F_k(P), b = Re 2F_k(P^2),
// kk = 2F_k(K), dk = 2F_k(K'); k = 0, N/2: a = F_k(P), b = F_k(P^2), kk, dk = (F_k, 0).
struct __attribute__((aligned(16))) BSumW {
  double a, b;
  double kk[2], dk[2];
};

struct BArgs {
  real_t* W;
  const real_t* W0s;
  real_t* G;              // Vopen(W0s) of the two-pass steps (dissipative instances), or null
  real_t* W0t;            // W0s on the theta spectrum (two-pass theta steps), or null
  const BSumW* sw;        // theta-sum weights of every instance (two-pass theta steps), or null
  const BInst* inst;
  const BTask* tasks;     // mode 0: queues concatenated; mode 1: passes concatenated
  const int* qoff;        // [9] queue offsets (mode 0)
  const tw_t* tw;         // exp(-2 pi i j / FW_B_NMAX)
  const double* times;    // t after the increment, per chunk-local step
  double* partials;       // per instance NSER rows of ntile sums (row k: sum k of every tile)
  double* series;         // [slot][instance][NSER]
  unsigned int* qhead;    // [chunk step][8]
  unsigned int* done;     // [chunk step][instance][3]
  int* fresh;             // [instance]: W normalised, partials stale
  int* err;               // dependency wait exceeded spin_limit
  const double* vkmin;    // Bloch: [instance][2] V_min, K_min
  int step;               // chunk-local step (series slot, counters)
  int gstep;              // global step (Bloch: instances stop after nbloch)
  int ninst;
  int mode;
  int task_base;          // mode 1: first task of this launch
  int spin_limit;
  int vv;                 // mode 1 Moyal: 0 three-pass step; 1 two-pass step (pass 0 tasks run
                          // the fused column pass); 2: pass 0 tasks compute G (and W0t) from W0s
  int pad_;
  double cv, ck;          // Bloch exponent factors (1/4, 1/2 or 1/2, 1)
};

enum { BK_MOYAL = 0, BK_BLOCH = 1 };

// ---------------------------------------------------------------------------
// Phase factors evaluated on the fly (FP64 argument; the FP32 form reduces it
// mod 2 pi in FP64 before its FP32 sincos)
// ---------------------------------------------------------------------------
FWD cpx b_vfactor(double X, double Th, double tau, const double* prm) {
  return unit_phase(-0.5 * ::dt * (::V(X - 0.5 * Th, tau, prm) - ::V(X + 0.5 * Th, tau, prm)), 1.0);
}
FWD cpx b_kfactor(double Lm, double P, double tau, const double* prm) {
  return unit_phase(-::dt * (::K(P + 0.5 * Lm, tau, prm) - ::K(P - 0.5 * Lm, tau, prm)), 1.0);
}
FWD cpx b_bloch_v(const BArgs& a, int i, double X, double Th, const double* prm, double db) {
  const double t0 = a.times[0];
  const double f = exp(-a.cv * db * (::V(X - 0.5 * Th, t0, prm) + ::V(X + 0.5 * Th, t0, prm) - a.vkmin[2 * i]));
  return mk((real_t)f, (real_t)0);
}
FWD cpx b_bloch_k(const BArgs& a, int i, double Lm, double P, const double* prm, double db) {
  const double t0 = a.times[0];
  const double f = exp(-a.ck * db * (::K(P + 0.5 * Lm, t0, prm) + ::K(P - 0.5 * Lm, t0, prm) - a.vkmin[2 * i + 1]));
  return mk((real_t)f, (real_t)0);
}

FWD void b_pair_apply(cpx& zk, cpx& zm, cpx fa, cpx fb, real_t s) {
  const real_t h = (real_t)0.5 * s;
  const cpx A = mk(h * (zk.x + zm.x), h * (zk.y - zm.y));
  const cpx B = mk(h * (zk.y + zm.y), h * (zm.x - zk.x));
  const cpx Ap = cmul(A, fa), Bp = cmul(B, fb);
  zk = mk(Ap.x - Bp.y, Ap.y + Bp.x);
  zm = mk(Ap.x + Bp.y, Bp.x - Ap.y);
}
FWD cpx b_self_apply(cpx z, cpx fa, cpx fb, real_t s) { return mk(z.x * fa.x * s, z.y * fb.x * s); }

// Pair multiply of the batched passes (the callback of spectral_step) with the
// factors of item i in registers: fa[i], fb[i] for lines a, b; slot E/2 holds
// the factors at the Nyquist index. The factors are evaluated before the step
// in one straight-line block (b_factors), where the evaluations share their
// instance- and time-only subexpressions.
template <int N, int E>
struct BMul {
  const cpx (&fa)[E / 2 + 1];
  const cpx (&fb)[E / 2 + 1];
  FWD void pair(int item, cpx& zk, cpx& zm) const {
    b_pair_apply(zk, zm, fa[item], fb[item], (real_t)1 / (real_t)N);
  }
  FWD void self(cpx& z0, cpx& zn) const {
    const real_t s = (real_t)1 / (real_t)N;
    z0 = b_self_apply(z0, fa[0], fb[0], s);
    zn = b_self_apply(zn, fa[E / 2], fb[E / 2], s);
  }
};

// ---------------------------------------------------------------------------
// V pass of one column tile: P column pairs c = 2 (tile P + q), q < P.
// FIRST: deferred normalisation (+ dissipator) on load. LAST: dissipator + the
// nine sums into partials.
// ---------------------------------------------------------------------------
template <int N, int E, int KIND, bool FIRST, bool LAST>
FWD void b_vpass(const BArgs& a, int i, int tile, int tid, cpx* lds, const tw_t* twl, double* red) {
  constexpr int TPL = N / E, P = BT / TPL;
  static_assert(P >= 1 && P * TPL == BT, "a tile must fill the block");
  using L = Line<N, E, P, false, FW_B_NMAX>;
  const BInst* in = a.inst + i;
  const int q = tid % P, t = tid / P;
  const int c = (tile * P + q) * 2;
  real_t* __restrict__ W = a.W + in->off;
  const double* prm = in->prm;
  const double tnow = a.times[a.step];
  const double tau = tnow + FW_PHASE_SHIFT * ::dt;
  const double dX = in->dX, dP = in->dP, dTh = in->dTheta;
  const bool dis = KIND == BK_MOYAL && in->dissip;
  const real_t wc = (real_t)in->w_coeff;

  cpx v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = *reinterpret_cast<const cpx*>(W + (size_t)(t + TPL * e) * N + c);

  if constexpr (FIRST) {
    // the previous step's norm (its sums are published by the K pass)
    double nrm = 1.0;
    if (a.fresh[i] == 0)
      nrm = 1.0 / (sum_partials<BT, 1>(a.partials + (size_t)in->part * NSER, in->ntile, 0, red) * (dX * dP));
    const real_t rn = (real_t)nrm;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      cpx w = v[e];
      w.x *= rn; w.y *= rn;
      if (dis) {
        const cpx g = *reinterpret_cast<const cpx*>(a.W0s + in->off + (size_t)(t + TPL * e) * N + c);
        w.x = w.x * wc + g.x;
        w.y = w.y * wc + g.y;
      }
      v[e] = w;
    }
  }
  __syncthreads();
  const double Xa = dX * (c - 0.5 * N), Xb = dX * (c + 1 - 0.5 * N);
  cpx fa[E / 2 + 1], fb[E / 2 + 1];
#pragma unroll
  for (int k = 0; k <= E / 2; ++k) {
    const double Th = dTh * (k < E / 2 ? spec_k<L>(t, k) : N / 2);
    if constexpr (KIND == BK_BLOCH) {
      fa[k] = b_bloch_v(a, i, Xa, Th, prm, in->dbeta); fb[k] = b_bloch_v(a, i, Xb, Th, prm, in->dbeta);
    } else {
      fa[k] = b_vfactor(Xa, Th, tau, prm); fb[k] = b_vfactor(Xb, Th, tau, prm);
    }
  }
  BMul<N, E> mul{fa, fb};
  spectral_step<L>(v, lds, q, t, twl, mul);

  double sums[NSER];
#pragma unroll
  for (int k = 0; k < NSER; ++k) sums[k] = 0.0;
  double ca = 0.0, cb = 0.0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = t + TPL * e;
    const size_t idx = (size_t)n * N + c;
    cpx w = v[e];
    if constexpr (LAST) {
      const double Pn = dP * (n - 0.5 * N);
      if (dis) {
        const cpx g = *reinterpret_cast<const cpx*>(a.W0s + in->off + idx);
        w.x = w.x * wc + g.x;
        w.y = w.y * wc + g.y;
      }
      const double wa = (double)w.x, wb = (double)w.y, ws = wa + wb;
      ca += wa; cb += wb;
      sums[2] += Pn * ws;
      sums[4] += Pn * Pn * ws;
      sums[5] += ::fw_diff_K(Pn, tnow, prm) * ws;
      sums[7] += ::K(Pn, tnow, prm) * ws;
      sums[8] += wa * wa + wb * wb;
    }
    st_w(reinterpret_cast<cpx*>(W + idx), w);
  }
  if constexpr (LAST) {
    sums[0] = ca + cb;
    sums[1] = Xa * ca + Xb * cb;
    sums[3] = Xa * Xa * ca + Xb * Xb * cb;
    sums[6] = ::fw_diff_V(Xa, tnow, prm) * ca + ::fw_diff_V(Xb, tnow, prm) * cb;
    sums[7] += ::V(Xa, tnow, prm) * ca + ::V(Xb, tnow, prm) * cb;
    write_partial_rows<BT>(sums, red, a.partials + (size_t)in->part * NSER, tile, in->ntile);
    // every K tile of this step has read `fresh`: a run's first step ends here
    if (tile == 0 && tid == 0 && a.fresh[i] != 0) a.fresh[i] = 0;
  }
}

// K pass of one row tile: P row pairs r = 2 (tile P + q).
template <int N, int E, int KIND>
FWD void b_kpass(const BArgs& a, int i, int tile, int tid, cpx* lds, const tw_t* twl, double* red) {
  constexpr int TPL = N / E, P = BT / TPL;
  using L = Line<N, E, P, true, FW_B_NMAX>;
  const BInst* in = a.inst + i;
  const int t = tid % TPL, q = tid / TPL;
  const int r = (tile * P + q) * 2;
  const double* prm = in->prm;
  const double tnow = a.times[a.step];
  const double tau = tnow + FW_PHASE_SHIFT * ::dt;
  real_t* __restrict__ Wa = a.W + in->off + (size_t)r * N;
  real_t* __restrict__ Wb = Wa + N;
  // `fresh` is cleared by the last V pass of the step, after every K tile has read it
  const int fr = a.fresh[i];
  cpx v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = mk(Wa[t + TPL * e], Wb[t + TPL * e]);
  if (KIND == BK_MOYAL && a.vv == 1 && fr == 0) {
    // two-pass steps: W = alpha Y + G, the previous step's norm and dissipator
    const bool dis = in->dissip;
    cpx g[E];
    if (dis) {
      const real_t* __restrict__ Ga = a.G + in->off + (size_t)r * N;
#pragma unroll
      for (int e = 0; e < E; ++e) g[e] = mk(Ga[t + TPL * e], Ga[N + t + TPL * e]);
    }
    const double S0 = sum_partials<BT, 1>(a.partials + (size_t)in->part * NSER, in->ntile, 0, red);
    const real_t al = (real_t)((dis ? in->w_coeff : 1.0) / (S0 * (in->dX * in->dP)));
#pragma unroll
    for (int e = 0; e < E; ++e) {
      v[e].x *= al;
      v[e].y *= al;
      if (dis) {
        v[e].x += g[e].x;
        v[e].y += g[e].y;
      }
    }
  }
  __syncthreads();
  const double dP = in->dP, dL = in->dLambda;
  const double Pa = dP * (r - 0.5 * N), Pb = dP * (r + 1 - 0.5 * N);
  cpx fa[E / 2 + 1], fb[E / 2 + 1];
#pragma unroll
  for (int k = 0; k <= E / 2; ++k) {
    const double Lm = dL * (k < E / 2 ? spec_k<L>(t, k) : N / 2);
    if constexpr (KIND == BK_BLOCH) {
      fa[k] = b_bloch_k(a, i, Lm, Pa, prm, in->dbeta); fb[k] = b_bloch_k(a, i, Lm, Pb, prm, in->dbeta);
    } else {
      fa[k] = b_kfactor(Lm, Pa, tau, prm); fb[k] = b_kfactor(Lm, Pb, tau, prm);
    }
  }
  BMul<N, E> mul{fa, fb};
  spectral_step<L>(v, lds, q, t, twl, mul);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    st_w(&Wa[t + TPL * e], v[e].x);
    st_w(&Wb[t + TPL * e], v[e].y);
  }
  // the previous step's sums to series[step]: tile k adds rows k, k + ntile, ...
  // (the last V pass, which overwrites the partials, runs after this pass)
  if (fr == 0) {
    const double* pp = a.partials + (size_t)in->part * NSER;
    for (int k = tile; k < NSER; k += in->ntile) {
      const double v2 = sum_partials<BT, 1>(pp + (size_t)k * in->ntile, in->ntile, 0, red);
      if (tid == 0) a.series[((size_t)a.step * a.ninst + i) * NSER + k] = v2;
    }
  }
}

// V factors of a column tile's items (V independent of t in the two-pass steps:
// one set serves the closing and the opening half-steps)
template <int N, int E, class L>
FWD void b_vfactors(const BInst* in, int c, int t, double tau, cpx (&fa)[E / 2 + 1], cpx (&fb)[E / 2 + 1]) {
  const double Xa = in->dX * (c - 0.5 * N), Xb = in->dX * (c + 1 - 0.5 * N);
#pragma unroll
  for (int k = 0; k <= E / 2; ++k) {
    const double Th = in->dTheta * (k < E / 2 ? spec_k<L>(t, k) : N / 2);
    fa[k] = b_vfactor(Xa, Th, tau, in->prm);
    fb[k] = b_vfactor(Xb, Th, tau, in->prm);
  }
}

// Fused column pass of the two-pass steps: closing V/2 of step s, dissipator,
// the nine sums of step s, opening V/2 of step s+1; stores Y (unnormalised).
template <int N, int E>
FWD void b_vvpass(const BArgs& a, int i, int tile, int tid, cpx* lds, const tw_t* twl, double* red) {
  constexpr int TPL = N / E, P = BT / TPL;
  using L = Line<N, E, P, false, FW_B_NMAX>;
  const BInst* in = a.inst + i;
  const int q = tid % P, t = tid / P;
  const int c = (tile * P + q) * 2;
  real_t* __restrict__ W = a.W + in->off;
  const double* prm = in->prm;
  const double tnow = a.times[a.step];
  const double tau = tnow + FW_PHASE_SHIFT * ::dt;
  const double dX = in->dX, dP = in->dP;
  const bool dis = in->dissip;
  const real_t wc = (real_t)in->w_coeff;
  cpx v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = *reinterpret_cast<const cpx*>(W + (size_t)(t + TPL * e) * N + c);
  cpx fa[E / 2 + 1], fb[E / 2 + 1];
  b_vfactors<N, E, L>(in, c, t, tau, fa, fb);
  __syncthreads();
  BMul<N, E> mul{fa, fb};
  spectral_step<L>(v, lds, q, t, twl, mul);
  // end of step s: dissipator, sums
  const double Xa = dX * (c - 0.5 * N), Xb = dX * (c + 1 - 0.5 * N);
  double sums[NSER];
#pragma unroll
  for (int k = 0; k < NSER; ++k) sums[k] = 0.0;
  double ca = 0.0, cb = 0.0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = t + TPL * e;
    const double Pn = dP * (n - 0.5 * N);
    cpx w = v[e];
    if (dis) {
      const cpx g = *reinterpret_cast<const cpx*>(a.W0s + in->off + (size_t)n * N + c);
      w.x = w.x * wc + g.x;
      w.y = w.y * wc + g.y;
    }
    const double wa = (double)w.x, wb = (double)w.y, ws = wa + wb;
    ca += wa; cb += wb;
    sums[2] += Pn * ws;
    sums[4] += Pn * Pn * ws;
    sums[5] += ::fw_diff_K(Pn, tnow, prm) * ws;
    sums[7] += ::K(Pn, tnow, prm) * ws;
    sums[8] += wa * wa + wb * wb;
    v[e] = w;
  }
  sums[0] = ca + cb;
  sums[1] = Xa * ca + Xb * cb;
  sums[3] = Xa * Xa * ca + Xb * Xb * cb;
  sums[6] = ::fw_diff_V(Xa, tnow, prm) * ca + ::fw_diff_V(Xb, tnow, prm) * cb;
  sums[7] += ::V(Xa, tnow, prm) * ca + ::V(Xb, tnow, prm) * cb;
  write_partial_rows<BT>(sums, red, a.partials + (size_t)in->part * NSER, tile, in->ntile);
  // every K tile of this step has read `fresh`
  if (tile == 0 && tid == 0 && a.fresh[i] != 0) a.fresh[i] = 0;
  // opening V/2 of step s+1 (the lane index through an empty asm: the
  // compiler recomputes this step's twiddles instead of keeping the first's)
  int t2 = t;
  asm volatile("" : "+v"(t2));
  __syncthreads();
  BMul<N, E> mul1{fa, fb};
  spectral_step<L>(v, lds, q, t2, twl, mul1);
#pragma unroll
  for (int e = 0; e < E; ++e) st_w(reinterpret_cast<cpx*>(W + (size_t)(t + TPL * e) * N + c), v[e]);
}

// ---------------------------------------------------------------------------
// Two-pass steps on the theta spectrum (FW_B_THETA: K and K' independent of
// t). The fused column pass never returns to x-p space between its halves:
// with alpha_k = A_k f_k / N after the closing factor (A_k the FFT along p),
// the dissipator is alpha <- w alpha + alpha0s (alpha0s the same transform of
// W0s, stored in W0t), the nine sums are taken on alpha (Parseval, weights
// BSumW; the single-grid vvpass_theta explains them), and the opening factor
// multiplies alpha in registers: one FFT pair per pass instead of two.
// W0t of column pair p of an instance: entries k < N/2 at off + 2 N p + 4 k as
// (Re A0s_k, Im A0s_k, Re B0s_k, Im B0s_k); entry 0 carries the real DC and
// Nyquist values as (A0s_0, A0s_N/2, B0s_0, B0s_N/2). Same size as W0s.
// ---------------------------------------------------------------------------
#ifndef FW_B_THETA
#define FW_B_THETA 0
#endif
template <int N, int E, class L>
struct BVVMul {
  const cpx (&fa)[E / 2 + 1];
  const cpx (&fb)[E / 2 + 1];
  const real_t* __restrict__ w0t;   // this pair's W0t entries, or null (no dissipator)
  const BSumW* __restrict__ sw;     // this instance's weights
  real_t wc;
  double dPN;                       // dP N
  int t;
  double acc[4];                    // P, P^2, K', K
  double pur;
  double a0, b0;                    // alpha_0 of columns a, b after the dissipator (thread 0)
  FWD void pair(int i, cpx& zk, cpx& zm) {
    const real_t h = (real_t)0.5 / (real_t)N;
    const int k = spec_k<L>(t, i);
    const cpx A = mk(h * (zk.x + zm.x), h * (zk.y - zm.y));
    const cpx B = mk(h * (zk.y + zm.y), h * (zm.x - zk.x));
    cpx A1 = cmul(A, fa[i]), B1 = cmul(B, fb[i]);
    if (w0t) {
      const real_t* e = w0t + 4 * k;
      A1 = mk(wc * A1.x + e[0], wc * A1.y + e[1]);
      B1 = mk(wc * B1.x + e[2], wc * B1.y + e[3]);
    }
    const double sx = (double)A1.x + (double)B1.x, sy = (double)A1.y + (double)B1.y;
    const BSumW w = sw[k];
    acc[0] += sy * w.a - sx * dPN;
    acc[1] += sx * w.b;
    acc[2] += sx * w.dk[0] + sy * w.dk[1];
    acc[3] += sx * w.kk[0] + sy * w.kk[1];
    const double ax = A1.x, ay = A1.y, bx = B1.x, by = B1.y;
    pur += 2.0 * (ax * ax + ay * ay + bx * bx + by * by);
    const cpx A2 = cmul(A1, fa[i]), B2 = cmul(B1, fb[i]);
    zk = mk(A2.x - B2.y, A2.y + B2.x);
    zm = mk(A2.x + B2.y, B2.x - A2.y);
  }
  FWD void self(cpx& z0, cpx& zn) {
    const real_t s = (real_t)1 / (real_t)N;
    real_t ra = z0.x * s * fa[0].x, rb = z0.y * s * fb[0].x;
    real_t na = zn.x * s * fa[E / 2].x, nb = zn.y * s * fb[E / 2].x;
    if (w0t) {
      ra = wc * ra + w0t[0];
      na = wc * na + w0t[1];
      rb = wc * rb + w0t[2];
      nb = wc * nb + w0t[3];
    }
    a0 = ra;
    b0 = rb;
    const double s0 = (double)ra + (double)rb, sn = (double)na + (double)nb;
    const BSumW w0 = sw[0], wn = sw[N / 2];
    acc[0] += s0 * w0.a + sn * wn.a;
    acc[1] += s0 * w0.b + sn * wn.b;
    acc[2] += s0 * w0.dk[0] + sn * wn.dk[0];
    acc[3] += s0 * w0.kk[0] + sn * wn.kk[0];
    pur += (double)ra * ra + (double)rb * rb + (double)na * na + (double)nb * nb;
    z0 = mk(ra * fa[0].x, rb * fb[0].x);
    zn = mk(na * fa[E / 2].x, nb * fb[E / 2].x);
  }
};

template <int N, int E>
FWD void b_vvpass_theta(const BArgs& a, int i, int tile, int tid, cpx* lds, const tw_t* twl, double* red) {
  constexpr int TPL = N / E, P = BT / TPL;
  using L = Line<N, E, P, false, FW_B_NMAX>;
  const BInst* in = a.inst + i;
  const int q = tid % P, t = tid / P;
  const int c = (tile * P + q) * 2;
  real_t* __restrict__ W = a.W + in->off;
  const double* prm = in->prm;
  const double tnow = a.times[a.step];
  const double tau = tnow + FW_PHASE_SHIFT * ::dt;
  const double dX = in->dX, dP = in->dP;
  cpx v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = *reinterpret_cast<const cpx*>(W + (size_t)(t + TPL * e) * N + c);
  cpx fa[E / 2 + 1], fb[E / 2 + 1];
  b_vfactors<N, E, L>(in, c, t, tau, fa, fb);
  __syncthreads();
  const real_t* w0t = in->dissip ? a.W0t + in->off + (size_t)N * c : nullptr;   // pair c/2: 2 N entries
  BVVMul<N, E, L> mul{fa, fb, w0t, a.sw + in->swoff, (real_t)in->w_coeff, dP * N, t, {0.0, 0.0, 0.0, 0.0},
                      0.0, 0.0, 0.0};
  spectral_step<L>(v, lds, q, t, twl, mul);
#pragma unroll
  for (int e = 0; e < E; ++e) st_w(reinterpret_cast<cpx*>(W + (size_t)(t + TPL * e) * N + c), v[e]);
  // the nine sums of step s (column sums N alpha_0 on thread 0 of each line)
  const double Xa = dX * (c - 0.5 * N), Xb = dX * (c + 1 - 0.5 * N);
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
    sums[6] = ::fw_diff_V(Xa, tnow, prm) * ca + ::fw_diff_V(Xb, tnow, prm) * cb;
    sums[7] += ::V(Xa, tnow, prm) * ca + ::V(Xb, tnow, prm) * cb;
  }
  sums[8] = (double)N * mul.pur;
  write_partial_rows<BT>(sums, red, a.partials + (size_t)in->part * NSER, tile, in->ntile);
  // every K tile of this step has read `fresh`
  if (tile == 0 && tid == 0 && a.fresh[i] != 0) a.fresh[i] = 0;
}

// W0t of a dissipative instance: the transform of W0s on the theta spectrum in
// the convention of alpha (A_k / N), recorded by a spectral step whose multiply
// stores the separated spectra and leaves them unchanged.
template <int N, int E, class L>
struct BStoreMul {
  real_t* __restrict__ out;   // this pair's W0t entries
  int t;
  FWD void pair(int i, cpx& zk, cpx& zm) const {
    const real_t h = (real_t)0.5 / (real_t)N;
    const int k = spec_k<L>(t, i);
    real_t* e = out + 4 * k;
    e[0] = h * (zk.x + zm.x);
    e[1] = h * (zk.y - zm.y);
    e[2] = h * (zk.y + zm.y);
    e[3] = h * (zm.x - zk.x);
  }
  FWD void self(cpx& z0, cpx& zn) const {
    const real_t s = (real_t)1 / (real_t)N;
    out[0] = z0.x * s;
    out[1] = zn.x * s;
    out[2] = z0.y * s;
    out[3] = zn.y * s;
  }
};

template <int N, int E>
FWD void b_w0tpass(const BArgs& a, int i, int tile, int tid, cpx* lds, const tw_t* twl) {
  constexpr int TPL = N / E, P = BT / TPL;
  using L = Line<N, E, P, false, FW_B_NMAX>;
  const BInst* in = a.inst + i;
  if (!in->dissip) return;   // uniform: one instance per block
  const int q = tid % P, t = tid / P;
  const int c = (tile * P + q) * 2;
  const real_t* __restrict__ S = a.W0s + in->off;
  cpx v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = *reinterpret_cast<const cpx*>(S + (size_t)(t + TPL * e) * N + c);
  __syncthreads();
  BStoreMul<N, E, L> mul{a.W0t + in->off + (size_t)N * c, t};
  spectral_step<L>(v, lds, q, t, twl, mul);
}

// Theta-sum weights of every instance: block i computes instance i's entries
// k <= N/2 by direct summation in FP64 (exact angle index (n k) mod N).
FWD void b_sumw(const BArgs& a, BSumW* sw) {
  const int i = blockIdx.x;
  const BInst* in = a.inst + i;
  const int N = in->n;
  const double tk = 0.0;   // K and K' do not depend on t here
  for (int k = threadIdx.x; k <= N / 2; k += blockDim.x) {
    double acc[4][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (int n = 0; n < N; ++n) {
      const int m = (int)(((long long)n * k) & (N - 1));
      double sn, cs;
      sincospi(-2.0 * (double)m / (double)N, &sn, &cs);
      const double P = in->dP * (n - 0.5 * N);
      const double f[4] = {P, P * P, ::fw_diff_K(P, tk, in->prm), ::K(P, tk, in->prm)};
#pragma unroll
      for (int o = 0; o < 4; ++o) {
        acc[o][0] += f[o] * cs;
        acc[o][1] += f[o] * sn;
      }
    }
    BSumW w;
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
    sw[in->swoff + k] = w;
  }
}

// G = Vopen(W0s) for a dissipative instance (the opening V/2 alone).
template <int N, int E>
FWD void b_gpass(const BArgs& a, int i, int tile, int tid, cpx* lds, const tw_t* twl) {
  constexpr int TPL = N / E, P = BT / TPL;
  using L = Line<N, E, P, false, FW_B_NMAX>;
  const BInst* in = a.inst + i;
  if (!in->dissip) return;   // uniform: one instance per block
  const int q = tid % P, t = tid / P;
  const int c = (tile * P + q) * 2;
  const real_t* __restrict__ S = a.W0s + in->off;
  real_t* __restrict__ G = a.G + in->off;
  const double tau = a.times[a.step] + FW_PHASE_SHIFT * ::dt;
  cpx v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = *reinterpret_cast<const cpx*>(S + (size_t)(t + TPL * e) * N + c);
  cpx fa[E / 2 + 1], fb[E / 2 + 1];
  b_vfactors<N, E, L>(in, c, t, tau, fa, fb);
  __syncthreads();
  BMul<N, E> mul{fa, fb};
  spectral_step<L>(v, lds, q, t, twl, mul);
#pragma unroll
  for (int e = 0; e < E; ++e) *reinterpret_cast<cpx*>(G + (size_t)(t + TPL * e) * N + c) = v[e];
}

template <int N, int E, int KIND>
FWD void b_run_n(const BArgs& a, int i, int pass, int tile, int tid, cpx* lds, const tw_t* twl, double* red) {
  if (pass == 0) {
    if constexpr (KIND == BK_MOYAL) {
      if (a.vv == 1) {
        if constexpr (FW_B_THETA) b_vvpass_theta<N, E>(a, i, tile, tid, lds, twl, red);
        else b_vvpass<N, E>(a, i, tile, tid, lds, twl, red);
        return;
      }
      if (a.vv == 2) {
        b_gpass<N, E>(a, i, tile, tid, lds, twl);
        return;
      }
      if (a.vv == 3) {
        if constexpr (FW_B_THETA) b_w0tpass<N, E>(a, i, tile, tid, lds, twl);
        return;
      }
    }
    b_vpass<N, E, KIND, true, false>(a, i, tile, tid, lds, twl, red);
  } else if (pass == 1) {
    b_kpass<N, E, KIND>(a, i, tile, tid, lds, twl, red);
  } else {
    b_vpass<N, E, KIND, false, true>(a, i, tile, tid, lds, twl, red);
  }
}

template <int KIND>
FWD void b_run(const BArgs& a, int i, int pass, int tile, int tid, cpx* lds, const tw_t* twl, double* red) {
  switch (a.inst[i].n) {
#if FW_B_HAS_64
    case 64: b_run_n<64, FW_B_E64, KIND>(a, i, pass, tile, tid, lds, twl, red); break;
#endif
#if FW_B_HAS_128
    case 128: b_run_n<128, FW_B_E128, KIND>(a, i, pass, tile, tid, lds, twl, red); break;
#endif
#if FW_B_HAS_256
    case 256: b_run_n<256, FW_B_E256, KIND>(a, i, pass, tile, tid, lds, twl, red); break;
#endif
#if FW_B_HAS_512
    case 512: b_run_n<512, FW_B_E512, KIND>(a, i, pass, tile, tid, lds, twl, red); break;
#endif
#if FW_B_HAS_1024
    case 1024: b_run_n<1024, FW_B_E1024, KIND>(a, i, pass, tile, tid, lds, twl, red); break;
#endif
    default: break;
  }
}

constexpr int B_LDS = BT * FW_B_EMAX + (BT * FW_B_EMAX) / 16 + 1;

// Wait until every tile of pass `pass` of instance i is released (thread 0
// spins, bounded), then acquire. Returns false if the wait gave up.
FWD bool b_wait(const BArgs& a, int i, int pass, unsigned int target) {
  __shared__ int ok;
  if (threadIdx.x == 0) {
    unsigned int* c = a.done + ((size_t)a.step * a.ninst + i) * 3 + pass;
    int it = 0, good = 1;
    while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(2);
      if (++it > a.spin_limit) {
        __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        good = 0;
        break;
      }
    }
    ok = good;
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return ok != 0;
}

FWD void b_release(const BArgs& a, int i, int pass) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_fetch_add(a.done + ((size_t)a.step * a.ninst + i) * 3 + pass, 1u, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
}

template <int KIND>
FWD void b_step(const BArgs& a) {
  __shared__ cpx lds[B_LDS];
  __shared__ tw_t twl[TwLds<FW_B_NMAX>::SIZE];
  __shared__ double red[NSER * 16];
  __shared__ int s_task;
  if (a.mode == 1) {
    const BTask tk = a.tasks[a.task_base + blockIdx.x];
    const int i = tk.inst, pass = tk.code >> 16, tile = tk.code & 0xffff;   // uniform (scalar loads)
    if (i < 0 || (KIND == BK_BLOCH && a.gstep >= a.inst[i].nbloch)) return;
    // the FFT of length n reads the table entries at stride FW_B_NMAX / n only
    const int n = a.inst[i].n, stride = FW_B_NMAX / n;
    for (int j = threadIdx.x; j < n; j += BT) twl[twpad(j * stride)] = a.tw[j * stride];
    __syncthreads();
    b_run<KIND>(a, i, pass, tile, threadIdx.x, lds, twl, red);
    return;
  }
  for (int j = threadIdx.x; j < FW_B_NMAX; j += BT) twl[twpad(j)] = a.tw[j];
  __syncthreads();
  const int home = blockIdx.x & 7;
  for (;;) {
    if (threadIdx.x == 0) {
      int got = -1;
      if (__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        for (int k = 0; k < 8 && got < 0; ++k) {
          const int qn = (home + k) & 7;
          const int len = a.qoff[qn + 1] - a.qoff[qn];
          if (len == 0) continue;
          unsigned int* h = a.qhead + (size_t)a.step * 8 + qn;
          if (__hip_atomic_load(h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= (unsigned int)len) continue;
          const unsigned int slot = __hip_atomic_fetch_add(h, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (slot < (unsigned int)len) got = a.qoff[qn] + (int)slot;
        }
      }
      s_task = got;
    }
    __syncthreads();
    // the task is the same for the whole block: make that visible to the
    // compiler, so the instance fields are scalar loads and the size switch a
    // uniform branch
    const int task = __builtin_amdgcn_readfirstlane(s_task);
    __syncthreads();
    if (task < 0) break;
    const BTask tk = a.tasks[task];
    const int i = tk.inst, pass = tk.code >> 16, tile = tk.code & 0xffff;
    bool ok = true;
    if (pass > 0) ok = b_wait(a, i, pass - 1, (unsigned int)a.inst[i].ntile);
    const bool active = !(KIND == BK_BLOCH && a.gstep >= a.inst[i].nbloch);
    // the lane index is re-materialised per task: otherwise the compiler keeps
    // every lane-dependent LDS/global offset of every size and pass variant
    // live across the task loop (loop-invariant hoisting), > 200 VGPRs
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    if (ok && active) b_run<KIND>(a, i, pass, tile, tid, lds, twl, red);
    if (pass < 2) b_release(a, i, pass);
    __syncthreads();
  }
}

// End of a run: W /= sum(W) dXdP from the last V pass's partials; block 0 of
// each instance publishes the sums to series[slot]. Grid: ninst x 8 blocks.
FWD void b_norm(const BArgs& a, int slot) {
  __shared__ double red[16];
  const int i = blockIdx.x >> 3, part = blockIdx.x & 7;
  const BInst* in = a.inst + i;
  if (a.fresh[i]) return;
  const double* pp = a.partials + (size_t)in->part * NSER;
  const double S0 = sum_partials<BT, 1>(pp, in->ntile, 0, red);
  if (part == 0) {
    double tot[NSER];
#pragma unroll
    for (int k = 0; k < NSER; ++k) tot[k] = (k == 0) ? S0 : sum_partials<BT, 1>(pp + (size_t)k * in->ntile, in->ntile, 0, red);
    if (threadIdx.x == 0)
      for (int k = 0; k < NSER; ++k) a.series[((size_t)slot * a.ninst + i) * NSER + k] = tot[k];
  }
  const real_t c = (real_t)(1.0 / (S0 * (in->dX * in->dP)));
  const size_t n = (size_t)in->n * in->n;
  real_t* W = a.W + in->off;
  for (size_t j = (size_t)part * BT + threadIdx.x; j < n; j += 8 * BT) W[j] *= c;
}

// Minima of V(X-Th/2)+V(X+Th/2) and K(P+L/2)+K(P-L/2) at times[0] per instance
// (as vk_min of the single grid). One block per instance.
FWD void b_vkmin(const BArgs& a, double* out) {
  __shared__ double rv[4], rk[4];
  const int i = blockIdx.x, tid = threadIdx.x;
  const BInst* in = a.inst + i;
  const double* prm = in->prm;
  const int N = in->n;
  const double t0 = a.times[0];
  double mv = __builtin_inf(), mkk = __builtin_inf();
  const int nh = N / 2 + 1;
  for (int idx = tid; idx < nh * N; idx += BT) {
    const int k = idx / N, j = idx % N;
    const double X = in->dX * (j - 0.5 * N), Th = in->dTheta * k;
    mv = fmin(mv, ::V(X - 0.5 * Th, t0, prm) + ::V(X + 0.5 * Th, t0, prm));
    const int r = idx / nh, l = idx % nh;
    const double P = in->dP * (r - 0.5 * N), Lm = in->dLambda * l;
    mkk = fmin(mkk, ::K(P + 0.5 * Lm, t0, prm) + ::K(P - 0.5 * Lm, t0, prm));
  }
  mv = wave_min(mv);
  mkk = wave_min(mkk);
  if ((tid & 63) == 0) { rv[tid >> 6] = mv; rk[tid >> 6] = mkk; }
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < BT / 64; ++w) { mv = fmin(mv, rv[w]); mkk = fmin(mkk, rk[w]); }
    out[2 * i] = mv;
    out[2 * i + 1] = mkk;
  }
}

}  // namespace fwk

// FW_B_WAVES: resident waves per SIMD the step kernel's registers must allow
// (0: the compiler's choice)
#ifndef FW_B_WAVES
#define FW_B_WAVES 4
#endif
#if FW_B_WAVES
#define FW_B_OCC __attribute__((amdgpu_waves_per_eu(FW_B_WAVES)))
#else
#define FW_B_OCC
#endif
extern "C" __global__ void __launch_bounds__(fwk::BT) FW_B_OCC fw_batch_step(fwk::BArgs a) {
  fwk::b_step<fwk::BK_MOYAL>(a);
}
extern "C" __global__ void __launch_bounds__(fwk::BT) fw_batch_bloch_step(fwk::BArgs a) {
  fwk::b_step<fwk::BK_BLOCH>(a);
}
extern "C" __global__ void __launch_bounds__(fwk::BT) fw_batch_norm(fwk::BArgs a, int slot) { fwk::b_norm(a, slot); }
extern "C" __global__ void __launch_bounds__(256) fw_batch_sumw(fwk::BArgs a, fwk::BSumW* sw) { fwk::b_sumw(a, sw); }
extern "C" __global__ void __launch_bounds__(fwk::BT) fw_batch_vkmin(fwk::BArgs a, double* out) {
  fwk::b_vkmin(a, out);
}
// W0s <- (1 - w_i) W for every instance with w_i < 1 (the random-collision
// source from the Gibbs states held in W). Grid: ninst x 64 blocks.
extern "C" __global__ void __launch_bounds__(fwk::BT) fw_batch_gibbs(fwk::BArgs a, fwk::real_t* w0s) {
  const int i = blockIdx.x >> 6, part = blockIdx.x & 63;
  const fwk::BInst* in = a.inst + i;
  const fwk::real_t c = (fwk::real_t)(1.0 - in->w_coeff);
  const size_t n = (size_t)in->n * in->n;
  for (size_t j = (size_t)part * fwk::BT + threadIdx.x; j < n; j += 64 * fwk::BT)
    w0s[in->off + j] = c * a.W[in->off + j];
}
// zero n words (chunk counters); fill an int array
extern "C" __global__ void __launch_bounds__(256) fw_batch_zero(unsigned int* p, int n) {
  for (int j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) p[j] = 0u;
}
extern "C" __global__ void __launch_bounds__(256) fw_batch_fill_int(int* p, int n, int v) {
  for (int j = blockIdx.x * 256 + threadIdx.x; j < n; j += gridDim.x * 256) p[j] = v;
}
