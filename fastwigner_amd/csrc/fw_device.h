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

#ifndef FW_REAL
#error "FW_REAL must be defined"
#endif

namespace fwk {

typedef FW_REAL real_t;

struct __attribute__((aligned(2 * sizeof(real_t)))) cpx {
  real_t x, y;
};

#define FWD __device__ __forceinline__

FWD cpx mk(real_t a, real_t b) { cpx r; r.x = a; r.y = b; return r; }
FWD cpx cadd(cpx a, cpx b) { return mk(a.x + b.x, a.y + b.y); }
FWD cpx csub(cpx a, cpx b) { return mk(a.x - b.x, a.y - b.y); }
FWD cpx cmul(cpx a, cpx b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
FWD cpx cmul_mi(cpx a) { return mk(a.y, -a.x); }   // a * (-i)
FWD cpx cconj(cpx a) { return mk(a.x, -a.y); }
FWD cpx cscale(cpx a, real_t s) { return mk(a.x * s, a.y * s); }

// Multiplication by a constant rotation wr + i wi given in double precision.
// In the FP32 form the constant is split into float hi + lo parts and the
// product compensated: the fixed rounding of an FP32 constant would otherwise
// perturb the step operator identically every step and the error would grow
// linearly with the number of steps.
FWD cpx cmulk(cpx a, double wr, double wi) {
#if FW_FP32
  const float hr = (float)wr, hi = (float)wi;
  const float lr = (float)(wr - (double)hr), li = (float)(wi - (double)hi);
  return cadd(cmul(a, mk(hr, hi)), cmul(a, mk(lr, li)));
#else
  return cmul(a, mk(wr, wi));
#endif
}

// FFT twiddle table entry exp(-2 pi i j / N): FP64 value, or FP32 hi + lo pair.
#if FW_FP32
struct tw_t {
  cpx hi, lo;
};
FWD cpx twmul(cpx a, const tw_t& w) { return cadd(cmul(a, w.hi), cmul(a, w.lo)); }
#else
typedef cpx tw_t;
FWD cpx twmul(cpx a, const tw_t& w) { return cmul(a, w); }
#endif

// ---------------------------------------------------------------------------
// Forward DFT kernels of radix 2, 4, 8, 16 on registers, natural order in/out.
// X_q = sum_m a_m exp(-2 pi i m q / R).
// ---------------------------------------------------------------------------
FWD void dft4(cpx& a0, cpx& a1, cpx& a2, cpx& a3) {
  const cpx t0 = cadd(a0, a2), t1 = csub(a0, a2);
  const cpx t2 = cadd(a1, a3), t3 = cmul_mi(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <int R> struct Dft;

template <> struct Dft<1> {
  FWD static void run(cpx*) {}
};

template <> struct Dft<2> {
  FWD static void run(cpx* a) {
    const cpx t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};

template <> struct Dft<4> {
  FWD static void run(cpx* a) { dft4(a[0], a[1], a[2], a[3]); }
};

template <> struct Dft<8> {
  FWD static void run(cpx* a) {
    const real_t c = (real_t)0.70710678118654752440;
    cpx e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
    cpx o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
    dft4(e0, e1, e2, e3);
    dft4(o0, o1, o2, o3);
#if FW_FP32
    o1 = cmulk(o1, 0.70710678118654752440, -0.70710678118654752440);
    o2 = cmul_mi(o2);
    o3 = cmulk(o3, -0.70710678118654752440, -0.70710678118654752440);
#else
    o1 = mk(c * (o1.x + o1.y), c * (o1.y - o1.x));     // * exp(-i pi/4)
    o2 = cmul_mi(o2);                                  // * exp(-i pi/2)
    o3 = mk(c * (o3.y - o3.x), -c * (o3.x + o3.y));    // * exp(-3 i pi/4)
#endif
    a[0] = cadd(e0, o0); a[4] = csub(e0, o0);
    a[1] = cadd(e1, o1); a[5] = csub(e1, o1);
    a[2] = cadd(e2, o2); a[6] = csub(e2, o2);
    a[3] = cadd(e3, o3); a[7] = csub(e3, o3);
  }
};

template <> struct Dft<16> {
  // 4 x 4 decomposition: X_{k1 + 4 k2} = sum_n2 W16^{n2 k1} W4^{n2 k2} sum_n1 a_{4 n1 + n2} W4^{n1 k1}
  FWD static void run(cpx* a) {
    const double c1 = 0.92387953251128675613;   // cos(pi/8)
    const double s1 = 0.38268343236508977173;   // sin(pi/8)
    const double h = 0.70710678118654752440;    // cos(pi/4)
    cpx b[16];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      cpx x0 = a[n2], x1 = a[4 + n2], x2 = a[8 + n2], x3 = a[12 + n2];
      dft4(x0, x1, x2, x3);
      b[4 * n2 + 0] = x0; b[4 * n2 + 1] = x1; b[4 * n2 + 2] = x2; b[4 * n2 + 3] = x3;
    }
    // twiddles W16^{n2 k1}
    b[5] = cmulk(b[5], c1, -s1);        // n2=1,k1=1 : W^1
    b[6] = cmulk(b[6], h, -h);          // n2=1,k1=2 : W^2
    b[7] = cmulk(b[7], s1, -c1);        // n2=1,k1=3 : W^3
    b[9] = cmulk(b[9], h, -h);          // n2=2,k1=1 : W^2
    b[10] = cmul_mi(b[10]);             // n2=2,k1=2 : W^4
    b[11] = cmulk(b[11], -h, -h);       // n2=2,k1=3 : W^6
    b[13] = cmulk(b[13], s1, -c1);      // n2=3,k1=1 : W^3
    b[14] = cmulk(b[14], -h, -h);       // n2=3,k1=2 : W^6
    b[15] = cmulk(b[15], -c1, s1);      // n2=3,k1=3 : W^9
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      cpx y0 = b[k1], y1 = b[4 + k1], y2 = b[8 + k1], y3 = b[12 + k1];
      dft4(y0, y1, y2, y3);
      a[k1] = y0; a[k1 + 4] = y1; a[k1 + 8] = y2; a[k1 + 12] = y3;
    }
  }
};

// ---------------------------------------------------------------------------
// Synchronisation
// ---------------------------------------------------------------------------
FWD void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).
// Remap so each XCD processes a contiguous range of tiles: neighbouring column
// tiles share 128-byte lines and then share the XCD's L2.
FWD int xcd_remap(int b, int nb) {
  if (nb & 7) return b;
  return (b & 7) * (nb >> 3) + (b >> 3);
}

// Twiddles are staged in LDS with one pad entry per 16: a stage reads entries
// m*k*N/(NS*R) for consecutive k, whose power-of-two stride would otherwise put
// the lanes of a wave on the same banks.
FWD int twpad(int i) { return i + (i >> 4); }
template <int N> struct TwLds { static constexpr int SIZE = N + N / 16 + 1; };

// ---------------------------------------------------------------------------
// Line FFT engine: P lines of length N per block, E complex elements per thread.
// LM (line-major) lane mapping: t = tid % TPL, q = tid / TPL (rows, contiguous);
// interleaved mapping: q = tid % P, t = tid / P (columns: a wave reads 2P
// adjacent columns of several rows per instruction).
// ---------------------------------------------------------------------------
template <int N, int E, int P, bool LM>
struct Line {
  static_assert((N & (N - 1)) == 0 && N >= 2, "line length must be a power of two");
  static_assert(E <= N && (E & (E - 1)) == 0, "elements per thread must be a power of two <= N");
  static constexpr int TPL = N / E;
  static constexpr int NT = P * TPL;
  static constexpr int RMAX = E < 16 ? E : 16;
  static constexpr int RAW = P * N;
  static constexpr int LDS = RAW + (RAW >> 4) + 1;
  static constexpr bool WAVE_LOCAL = (NT <= 64) || (LM && TPL <= 64 && (64 % TPL) == 0);

  FWD static int lidx(int q, int i) {
    const int r = LM ? q * N + i : i * P + q;
    return r + (r >> 4);
  }
  FWD static void sync() {
    if constexpr (WAVE_LOCAL) wave_sync();
    else __syncthreads();
  }
  FWD static void load_natural(cpx (&v)[E], const cpx* lds, int q, int t) {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = lds[lidx(q, t + TPL * e)];
  }

  // One Stockham radix-R stage; NS = product of the radices already applied.
  // Butterfly j = t + TPL b takes x[j + m N/R] = v[b + m B] (B = E/R), applies
  // the twiddles w^{m k}, k = j mod NS, w = exp(-2 pi i/(NS R)), and writes
  // output q to index (j/NS) NS R + k + q NS. TOREG: the last stage, whose
  // outputs land on the thread's own natural positions, stays in registers.
  template <int R, int NS, bool TOREG>
  FWD static void stage(cpx (&v)[E], cpx* lds, int q, int t, const tw_t* __restrict__ tw) {
    constexpr int B = E / R;
    constexpr int TWS = N / (NS * R);
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int j = t + TPL * b;
      cpx a[R];
#pragma unroll
      for (int m = 0; m < R; ++m) a[m] = v[b + m * B];
      if constexpr (NS > 1) {
        const int k = j & (NS - 1);
#pragma unroll
        for (int m = 1; m < R; ++m) a[m] = twmul(a[m], tw[twpad(m * k * TWS)]);
      }
      Dft<R>::run(a);
      if constexpr (TOREG) {
#pragma unroll
        for (int m = 0; m < R; ++m) v[b + m * B] = a[m];
      } else {
        const int base = (j & ~(NS - 1)) * R + (j & (NS - 1));
#pragma unroll
        for (int m = 0; m < R; ++m) lds[lidx(q, base + m * NS)] = a[m];
      }
    }
  }

  // Forward FFT. TOLDS: leave the spectrum in LDS in natural order (synced);
  // otherwise leave it in registers at the natural positions t + TPL e.
  template <bool TOLDS, int NS = 1>
  FWD static void fwd(cpx (&v)[E], cpx* lds, int q, int t, const tw_t* __restrict__ tw) {
    constexpr int REM = N / NS;
    constexpr int R = REM < RMAX ? REM : RMAX;
    constexpr bool LAST = (NS * R == N);
    if constexpr (LAST && !TOLDS) {
      stage<R, NS, true>(v, lds, q, t, tw);
    } else {
      stage<R, NS, false>(v, lds, q, t, tw);
      sync();
      if constexpr (!LAST) {
        load_natural(v, lds, q, t);
        sync();
        fwd<TOLDS, NS * R>(v, lds, q, t, tw);
      }
    }
  }

  // Unnormalised inverse FFT, registers to registers (conjugation identity).
  FWD static void inv(cpx (&v)[E], cpx* lds, int q, int t, const tw_t* __restrict__ tw) {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e].y = -v[e].y;
    fwd<false>(v, lds, q, t, tw);
#pragma unroll
    for (int e = 0; e < E; ++e) v[e].y = -v[e].y;
  }
};

// ---------------------------------------------------------------------------
// Per-launch arguments (one struct, passed by value)
// ---------------------------------------------------------------------------
enum { NSER = 9 };               // S0 Sx Sp Sxx Spp SdK SdV SH SWW
// scal[] slots. SC_FRESH = 1: W is normalised, the partials are stale (set by the
// host before a run, cleared by the first K pass). SC_OUT.. = 9 moments.
enum { SC_NORM = 0, SC_STOP = 1, SC_FRESH = 2, SC_OUT = 8 };

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
};

// two adjacent factors (columns c, c+1) in one load
struct __attribute__((aligned(4 * sizeof(real_t)))) cpx2 {
  cpx a, b;
};

// ---------------------------------------------------------------------------
// Phase factors
// ---------------------------------------------------------------------------
FWD cpx unit_phase(double ph, double m) {
#if FW_FP32
  // FP32 form: argument and its mod-2pi reduction in FP64, sincos in FP32.
  const double two_pi = 6.283185307179586476925;
  const double r = ph - two_pi * rint(ph * 0.15915494309189533577);
  const float rf = (float)r;
  const float rl = (float)(r - (double)rf);   // argument rounding, corrected to first order
  float s, c;
  sincosf(rf, &s, &c);
  const float cc = c - s * rl, ss = s + c * rl;
  const float mf = (float)m;
  return mk(cc * mf, ss * mf);
#else
  double s, c;
  sincos(ph, &s, &c);
  return mk(c * m, s * m);
#endif
}

// Table entries are built once, so they take the FP64 sincos in both forms and
// are rounded once to the storage precision.
FWD cpx unit_phase_exact(double ph, double m) {
  double s, c;
  sincos(ph, &s, &c);
  return mk((real_t)(c * m), (real_t)(s * m));
}

// expV_bulk / expV_boundary, wigner_moyal_cuda_1d.py:631-663
template <bool EXACT = false>
FWD cpx vfactor(double X, double Theta, double tau, double tnow) {
  const double ph = -0.5 * ::dt * (::V(X - 0.5 * Theta, tau) - ::V(X + 0.5 * Theta, tau));
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

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------
FWD double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-reduce NSER sums into partials[blk], then the last block to finish adds
// the partials of all blocks (every thread takes a strided share, Neumaier-
// compensated FP64, fixed order: deterministic), publishes the totals to out[]
// and resets the ticket. red: >= NSER * 16 doubles of LDS.
template <int NT>
FWD double block_sum(double v, double* red, int i) {
  constexpr int NW = (NT + 63) / 64;
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[(threadIdx.x >> 6) * NSER + i] = v;
  (void)NW;
  return v;
}

template <int NT>
FWD bool block_reduce_publish(double (&s)[NSER], double* red, double* partials, unsigned int* counter,
                              int blk, int nb, double* out) {
  const int tid = threadIdx.x;
  constexpr int NW = (NT + 63) / 64;
  __shared__ int is_last;
#pragma unroll
  for (int i = 0; i < NSER; ++i) block_sum<NT>(s[i], red, i);
  __syncthreads();
  if (tid < NSER) {
    double acc = 0.0;
    for (int w = 0; w < NW; ++w) acc += red[w * NSER + tid];
    __hip_atomic_store(partials + (size_t)blk * NSER + tid, acc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __threadfence();
  __syncthreads();
  if (tid == 0) {
    const unsigned int tk = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    is_last = (tk == (unsigned int)(nb - 1));
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
  double sum[NSER], comp[NSER];
#pragma unroll
  for (int i = 0; i < NSER; ++i) { sum[i] = 0.0; comp[i] = 0.0; }
  for (int b = tid; b < nb; b += NT) {
#pragma unroll
    for (int i = 0; i < NSER; ++i) {
      const double x = __hip_atomic_load(partials + (size_t)b * NSER + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const double tt = sum[i] + x;
      comp[i] += (fabs(sum[i]) >= fabs(x)) ? (sum[i] - tt) + x : (x - tt) + sum[i];
      sum[i] = tt;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NSER; ++i) block_sum<NT>(sum[i] + comp[i], red, i);
  __syncthreads();
  if (tid < NSER) {
    double acc = 0.0;
    for (int w = 0; w < NW; ++w) acc += red[w * NSER + tid];
    out[tid] = acc;
  }
  __syncthreads();
  if (tid == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// ---------------------------------------------------------------------------
// Deferred normalisation. The last V pass of a step writes one row of NSER
// partial sums per block (plain stores, no atomics: an agent-scope fence or
// atomic per block serialises the grid). The next kernel that needs the sums
// adds the partials itself: every block takes the same strided order and the
// same tree, so every block obtains the same bits.
// ---------------------------------------------------------------------------
template <int NT>
FWD double sum_partials(const double* __restrict__ partials, int nb, int i, double* red) {
  constexpr int NW = (NT + 63) / 64;
  double sum = 0.0, comp = 0.0;
  for (int b = threadIdx.x; b < nb; b += NT) {
    const double x = partials[(size_t)b * NSER + i];
    const double tt = sum + x;
    comp += (fabs(sum) >= fabs(x)) ? (sum - tt) + x : (x - tt) + sum;
    sum = tt;
  }
  double v = wave_sum(sum + comp);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double tot = 0.0;
#pragma unroll
  for (int w = 0; w < NW; ++w) tot += red[w];
  return tot;
}

// Norm factor of the previous step; block 0 also publishes its NSER sums to
// series[slot] and raises the purity guard. Returns false when the step must
// not run (guard tripped).
template <int NT>
FWD bool previous_step(const Args& a, int blk, double* red, double& nrm) {
  const double dxdp = ::dX * ::dP;
  const double S0 = sum_partials<NT>(a.partials, a.nbp, 0, red);
  nrm = 1.0 / (S0 * dxdp);
  bool stop = false;
  if (a.guard > 0.0) {
    const double S2 = sum_partials<NT>(a.partials, a.nbp, 8, red);
    stop = !(2.0 * M_PI * S2 / (S0 * S0 * dxdp) < a.guard);
  }
  if (blk == 0) {
    double tot[NSER];
#pragma unroll
    for (int i = 0; i < NSER; ++i) tot[i] = (i == 0) ? S0 : sum_partials<NT>(a.partials, a.nbp, i, red);
    if (threadIdx.x == 0) {
#pragma unroll
      for (int i = 0; i < NSER; ++i) a.series[(size_t)a.step * NSER + i] = tot[i];
      if (stop) a.scal[SC_STOP] = (double)(a.step + 1);
    }
  }
  return !stop;
}

// Block sums of the last V pass -> partials[blk] (plain stores).
template <int NT>
FWD void write_partials(double (&s)[NSER], double* red, double* partials, int blk) {
  constexpr int NW = (NT + 63) / 64;
#pragma unroll
  for (int i = 0; i < NSER; ++i) {
    const double v = wave_sum(s[i]);
    if ((threadIdx.x & 63) == 0) red[(threadIdx.x >> 6) * NSER + i] = v;
  }
  __syncthreads();
  if (threadIdx.x < NSER) {
    double acc = 0.0;
    for (int w = 0; w < NW; ++w) acc += red[w * NSER + threadIdx.x];
    partials[(size_t)blk * NSER + threadIdx.x] = acc;
  }
}

// ---------------------------------------------------------------------------
// V pass: columns (along p), lines = pairs of adjacent columns.
// ---------------------------------------------------------------------------
using VL = Line<FW_NP, FW_COL_E, FW_COL_P, false>;
using KL = Line<FW_NX, FW_ROW_E, FW_ROW_P, true>;

template <int FL>
FWD void vpass(const Args& a) {
  using L = VL;
  constexpr int N = FW_NP, E = FW_COL_E, P = FW_COL_P, TPL = L::TPL, NT = L::NT;
  __shared__ cpx lds[L::LDS];
  __shared__ tw_t twl[TwLds<N>::SIZE];
  __shared__ double red[NSER * 16];
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
  constexpr int TWN = (N + NT - 1) / NT;
  tw_t twr[TWN];
#pragma unroll
  for (int i = 0; i < TWN; ++i) twr[i] = a.tw_col[(tid + i * NT) & (N - 1)];
  cpx v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = *reinterpret_cast<const cpx*>(W + (size_t)(t + TPL * e) * FW_NX + c);
  cpx2 ftab[E / 2 + 1];
  if constexpr (use_tab) {
#pragma unroll
    for (int e = 0; e < E / 2; ++e)
      ftab[e] = *reinterpret_cast<const cpx2*>(a.vtab + (size_t)(t + TPL * e) * FW_NX + c);
    ftab[E / 2] = *reinterpret_cast<const cpx2*>(a.vtab + (size_t)(N / 2) * FW_NX + c);
  }
#pragma unroll
  for (int i = 0; i < TWN; ++i)
    if ((N % NT) == 0 || tid + i * NT < N) twl[twpad(tid + i * NT)] = twr[i];

  if constexpr (FL & F_FIRST) {
    double nrm = 1.0;
    if (a.scal[SC_FRESH] == 0.0) {
      if (!previous_step<NT>(a, blk, red, nrm)) return;
    } else {
      __syncthreads();
    }
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
  } else {
    __syncthreads();
  }
  L::template fwd<true>(v, lds, q, t, twl);

  // spectral multiply on the theta x representation
  const double Xa = ::dX * (c - 0.5 * ::X_gridDIM);
  const double Xb = ::dX * (c + 1 - 0.5 * ::X_gridDIM);
  {
    const real_t s = (real_t)1 / (real_t)N;
#pragma unroll
    for (int e = 0; e < E / 2; ++e) {
      const int k = t + TPL * e;
      if (k == 0) {
        const double ThN = ::dTheta * (N / 2);
        cpx fa0, fb0, faN, fbN;
        if constexpr (use_tab) {
          fa0 = ftab[0].a; fb0 = ftab[0].b; faN = ftab[E / 2].a; fbN = ftab[E / 2].b;
        } else {
          fa0 = vfactor(Xa, 0.0, tau, tnow); fb0 = vfactor(Xb, 0.0, tau, tnow);
          faN = vfactor(Xa, ThN, tau, tnow); fbN = vfactor(Xb, ThN, tau, tnow);
        }
        const int i0 = L::lidx(q, 0), iN = L::lidx(q, N / 2);
        lds[i0] = self_apply(lds[i0], fa0, fb0, s);
        lds[iN] = self_apply(lds[iN], faN, fbN, s);
      } else {
        const double Th = ::dTheta * k;
        cpx fa, fb;
        if constexpr (use_tab) {
          fa = ftab[e].a; fb = ftab[e].b;
        } else {
          fa = vfactor(Xa, Th, tau, tnow); fb = vfactor(Xb, Th, tau, tnow);
        }
        const int ik = L::lidx(q, k), im = L::lidx(q, N - k);
        cpx zk = lds[ik], zm = lds[im];
        pair_apply(zk, zm, fa, fb, s);
        lds[ik] = zk;
        lds[im] = zm;
      }
    }
  }
  L::sync();
  L::load_natural(v, lds, q, t);
  L::sync();
  L::inv(v, lds, q, t, twl);

  double sums[NSER];
#pragma unroll
  for (int i = 0; i < NSER; ++i) sums[i] = 0.0;
  double ca = 0.0, cb = 0.0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = t + TPL * e;
    const size_t idx = (size_t)n * FW_NX + c;
    cpx w = v[e];
    if constexpr (FL & F_LAST) {
      const double Pn = ::dP * (n - 0.5 * ::P_gridDIM);
      if (a.flags & F_MASK_XP) {
        w.x *= (real_t)::fw_mask_x_p(Xa, Pn, tnow);
        w.y *= (real_t)::fw_mask_x_p(Xb, Pn, tnow);
      }
      if constexpr (FL & F_DISSIP) {
        const cpx g = *reinterpret_cast<const cpx*>(a.W0s + idx);
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
    *reinterpret_cast<cpx*>(W + idx) = w;
  }
  if constexpr (FL & F_LAST) {
    sums[0] = ca + cb;
    sums[1] = Xa * ca + Xb * cb;
    sums[3] = Xa * Xa * ca + Xb * Xb * cb;
    sums[6] = ::fw_diff_V(Xa, tnow) * ca + ::fw_diff_V(Xb, tnow) * cb;
    sums[7] += ::V(Xa, tnow) * ca + ::V(Xb, tnow) * cb;
    write_partials<NT>(sums, red, a.partials, blk);
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
  __shared__ tw_t twl[TwLds<N>::SIZE];
  if (a.scal[SC_STOP] != 0.0) return;
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
  if (blk == 0 && tid == 0 && a.scal[SC_FRESH] != 0.0) a.scal[SC_FRESH] = 0.0;

  constexpr int TWN = (N + NT - 1) / NT;
  tw_t twr[TWN];
#pragma unroll
  for (int i = 0; i < TWN; ++i) twr[i] = a.tw_row[(tid + i * NT) & (N - 1)];
  cpx v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = mk(Wa[t + TPL * e], Wb[t + TPL * e]);
  cpx fta[E / 2 + 1], ftb[E / 2 + 1];
  if constexpr (use_tab) {
#pragma unroll
    for (int e = 0; e < E / 2; ++e) {
      fta[e] = a.ktab[(size_t)r * KW + t + TPL * e];
      ftb[e] = a.ktab[(size_t)(r + 1) * KW + t + TPL * e];
    }
    fta[E / 2] = a.ktab[(size_t)r * KW + N / 2];
    ftb[E / 2] = a.ktab[(size_t)(r + 1) * KW + N / 2];
  }
#pragma unroll
  for (int i = 0; i < TWN; ++i)
    if ((N % NT) == 0 || tid + i * NT < N) twl[twpad(tid + i * NT)] = twr[i];
  __syncthreads();
  L::template fwd<true>(v, lds, q, t, twl);
  {
    const double Pa = ::dP * (r - 0.5 * ::P_gridDIM);
    const double Pb = ::dP * (r + 1 - 0.5 * ::P_gridDIM);
    const real_t s = (real_t)1 / (real_t)N;
#pragma unroll
    for (int e = 0; e < E / 2; ++e) {
      const int k = t + TPL * e;
      if (k == 0) {
        const double LN = ::dLambda * (N / 2);
        cpx fa0, fb0, faN, fbN;
        if constexpr (use_tab) {
          fa0 = fta[0]; fb0 = ftb[0]; faN = fta[E / 2]; fbN = ftb[E / 2];
        } else {
          fa0 = kfactor(0.0, Pa, tau, tnow); fb0 = kfactor(0.0, Pb, tau, tnow);
          faN = kfactor(LN, Pa, tau, tnow); fbN = kfactor(LN, Pb, tau, tnow);
        }
        const int i0 = L::lidx(q, 0), iN = L::lidx(q, N / 2);
        lds[i0] = self_apply(lds[i0], fa0, fb0, s);
        lds[iN] = self_apply(lds[iN], faN, fbN, s);
      } else {
        const double Lm = ::dLambda * k;
        cpx fa, fb;
        if constexpr (use_tab) {
          fa = fta[e]; fb = ftb[e];
        } else {
          fa = kfactor(Lm, Pa, tau, tnow); fb = kfactor(Lm, Pb, tau, tnow);
        }
        const int ik = L::lidx(q, k), im = L::lidx(q, N - k);
        cpx zk = lds[ik], zm = lds[im];
        pair_apply(zk, zm, fa, fb, s);
        lds[ik] = zk;
        lds[im] = zm;
      }
    }
  }
  L::sync();
  L::load_natural(v, lds, q, t);
  L::sync();
  L::inv(v, lds, q, t, twl);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    Wa[t + TPL * e] = v[e].x;
    Wb[t + TPL * e] = v[e].y;
  }
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

// W *= norm: the norm is scal[SC_NORM], or, at the end of a run
// (F_FROM_PARTIALS), that of the last step, whose sums block 0 publishes to
// series[step].
FWD void apply_norm(const Args& a) {
  __shared__ double red[16];
  double c;
  if (a.flags & F_FROM_PARTIALS) {
    const double S0 = sum_partials<HT>(a.partials, a.nbp, 0, red);
    c = 1.0 / (S0 * (::dX * ::dP));
    if (blockIdx.x == 0) {
      double tot[NSER];
#pragma unroll
      for (int i = 0; i < NSER; ++i) tot[i] = (i == 0) ? S0 : sum_partials<HT>(a.partials, a.nbp, i, red);
      if (threadIdx.x == 0)
        for (int i = 0; i < NSER; ++i) a.series[(size_t)a.step * NSER + i] = tot[i];
    }
  } else {
    c = a.scal[SC_NORM];
  }
  const real_t cr = (real_t)c;
  const size_t n = (size_t)FW_NX * FW_NP;
  for (size_t i = (size_t)blockIdx.x * HT + threadIdx.x; i < n; i += (size_t)gridDim.x * HT) a.W[i] *= cr;
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
    tab[idx] = (a.flags & F_BLOCH) ? bloch_vfactor(a, X, Th) : vfactor<true>(X, Th, tau, tnow);
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

FWD double wave_min(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
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
extern "C" __global__ void __launch_bounds__(fwk::HT) fw_moments(fwk::Args a) { fwk::moments(a); }
extern "C" __global__ void __launch_bounds__(fwk::HT) fw_apply_norm(fwk::Args a) { fwk::apply_norm(a); }
extern "C" __global__ void fw_set_scalar(double* p, double v) { *p = v; }
extern "C" __global__ void __launch_bounds__(1024) fw_vk_min(fwk::Args a) { fwk::vk_min(a); }
extern "C" __global__ void __launch_bounds__(256) fw_build_vtab(fwk::Args a, fwk::cpx* tab) { fwk::build_vtab(a, tab); }
extern "C" __global__ void __launch_bounds__(256) fw_build_ktab(fwk::Args a, fwk::cpx* tab) { fwk::build_ktab(a, tab); }
#elif FW_KSET == 2
FW_VK(fw_bv_first, fwk::F_FIRST | fwk::F_BLOCH)
FW_VK(fw_bv_last, fwk::F_LAST | fwk::F_BLOCH)
FW_KK(fw_bk, fwk::F_BLOCH)
#elif FW_KSET == 3
FW_VK(fw_v_first_dissip, fwk::F_FIRST | fwk::F_DISSIP)
FW_VK(fw_v_last_dissip, fwk::F_LAST | fwk::F_DISSIP)
#endif
