// fw_common.h -- shared CDNA4 (gfx950) device code of the propagators: complex
// arithmetic, radix-2/4/8/16 DFT kernels, the LDS Stockham line-FFT engine,
// phase factors and the block reductions. Embedded in libfastwigner.so and
// compiled by hipRTC; included by fw_device.h (Wigner function) and fw_rho.h
// (density matrix). Needs FW_REAL (double|float) and FW_FP32 (0|1).
#pragma once

#include "fw_math.h"

#ifndef FW_REAL
#error "FW_REAL must be defined"
#endif
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

// Complex scalar for the caller's expressions (the reference's density-matrix
// strings use pycuda's complex type under the name cuda_complex, e.g.
// "exp(-sigma*(...) + cuda_complex(0., p0*(X - X_prime)))",
// rho_vneumann_cuda_1d.py:962-966).
struct fw_complex {
  double re, im;
  __device__ fw_complex(double r = 0., double i = 0.) : re(r), im(i) {}
};
typedef fw_complex cuda_complex;
__device__ inline fw_complex operator+(fw_complex a, fw_complex b) { return fw_complex(a.re + b.re, a.im + b.im); }
__device__ inline fw_complex operator-(fw_complex a, fw_complex b) { return fw_complex(a.re - b.re, a.im - b.im); }
__device__ inline fw_complex operator-(fw_complex a) { return fw_complex(-a.re, -a.im); }
__device__ inline fw_complex operator*(fw_complex a, fw_complex b) {
  return fw_complex(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__device__ inline fw_complex operator/(fw_complex a, fw_complex b) {
  const double d = b.re * b.re + b.im * b.im;
  return fw_complex((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}
__device__ inline fw_complex operator+(fw_complex a, double b) { return fw_complex(a.re + b, a.im); }
__device__ inline fw_complex operator+(double a, fw_complex b) { return fw_complex(a + b.re, b.im); }
__device__ inline fw_complex operator-(fw_complex a, double b) { return fw_complex(a.re - b, a.im); }
__device__ inline fw_complex operator-(double a, fw_complex b) { return fw_complex(a - b.re, -b.im); }
__device__ inline fw_complex operator*(fw_complex a, double b) { return fw_complex(a.re * b, a.im * b); }
__device__ inline fw_complex operator*(double a, fw_complex b) { return fw_complex(a * b.re, a * b.im); }
__device__ inline fw_complex operator/(fw_complex a, double b) { return fw_complex(a.re / b, a.im / b); }
__device__ inline fw_complex operator/(double a, fw_complex b) { return fw_complex(a, 0.) / b; }
__device__ inline fw_complex exp(fw_complex z) {
  double s, c;
  sincos(z.im, &s, &c);
  const double e = exp(z.re);
  return fw_complex(e * c, e * s);
}
__device__ inline fw_complex conj(fw_complex z) { return fw_complex(z.re, -z.im); }
__device__ inline double real(fw_complex z) { return z.re; }
__device__ inline double imag(fw_complex z) { return z.im; }
__device__ inline double real(double x) { return x; }
__device__ inline double imag(double) { return 0.; }
__device__ inline double abs(fw_complex z) { return hypot(z.re, z.im); }
__device__ inline double norm(fw_complex z) { return z.re * z.re + z.im * z.im; }

namespace fwk {

template <bool B, class T, class F> struct pick { using type = T; };
template <class T, class F> struct pick<false, T, F> { using type = F; };

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

// Stores of the state by the passes. A pass writes a whole grid that the next
// kernel reads, on every XCD; the XCDs' L2 caches are not coherent with one
// another, so the end of a kernel writes the dirty lines of its L2 back before
// a dependent kernel starts. Stores of device scope (sc1) write through the L2
// while the pass runs, which leaves that boundary little to write back:
// measured on MI355X, the 1024^2 FP64 Moyal step 24.0 -> 19.7 us, the
// density-matrix step 57.2 -> 46.3 us, the same bits. Policy: 1 = sc1 (the
// default), 0 = plain stores (the batched propagator's passes, whose grids
// exceed the caches, are faster with them).
#ifndef FW_W_STORE
#define FW_W_STORE 1
#endif
template <class T>
FWD void st_w(T* p, T v) {
  if constexpr (FW_W_STORE == 0) {
    *p = v;
  } else {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16, "4, 8 or 16 bytes");
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    if constexpr (sizeof(T) == 16) {
      // the s_nop: a store of more than 64 bits of data needs a wait state
      // before a VALU may overwrite its data registers (the compiler inserts
      // it after its own stores but cannot see a store inside asm)
      const u4 d = __builtin_bit_cast(u4, v);
      asm volatile("global_store_dwordx4 %0, %1, off sc1\n\ts_nop 1" : : "v"(p), "v"(d) : "memory");
    } else if constexpr (sizeof(T) == 8) {
      const u2 d = __builtin_bit_cast(u2, v);
      asm volatile("global_store_dwordx2 %0, %1, off sc1" : : "v"(p), "v"(d) : "memory");
    } else {
      const unsigned int d = __builtin_bit_cast(unsigned int, v);
      asm volatile("global_store_dword %0, %1, off sc1" : : "v"(p), "v"(d) : "memory");
    }
  }
}

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

// Real constant times a complex value. In the FP32 form the constant is split
// into float hi + lo parts (the compensation of cmulk).
FWD cpx kscale(cpx a, double c) {
#if FW_FP32
  const float h = (float)c, l = (float)(c - (double)h);
  return mk(fmaf(a.x, l, a.x * h), fmaf(a.y, l, a.y * h));
#else
  return cscale(a, c);
#endif
}

// cos and sin of 2 pi j / n for compile-time constants: the angle reduced
// exactly in integers to the nearest quarter turn q and a remainder
// |phi| <= pi/4, the Taylor series of sin to phi^17 and of cos to phi^18
// (truncation below 1e-19), then the quarter-turn rotation; about 1 ulp.
struct ct_cs {
  double c, s;
};
constexpr ct_cs ct_unit(long j, long n) {
  const long u = 4 * (((j % n) + n) % n);
  const long q = (u + n / 2) / n;
  const long r = u - q * n;
  const double phi = (double)r * (3.14159265358979323846 / (2.0 * (double)n));
  const double p2 = phi * phi;
  double ps = 1.0 / 355687428096000.0;
  ps = ps * p2 - 1.0 / 1307674368000.0;
  ps = ps * p2 + 1.0 / 6227020800.0;
  ps = ps * p2 - 1.0 / 39916800.0;
  ps = ps * p2 + 1.0 / 362880.0;
  ps = ps * p2 - 1.0 / 5040.0;
  ps = ps * p2 + 1.0 / 120.0;
  ps = ps * p2 - 1.0 / 6.0;
  const double sn = phi + phi * p2 * ps;
  double pc = 1.0 / 6402373705728000.0;
  pc = pc * p2 - 1.0 / 20922789888000.0;
  pc = pc * p2 + 1.0 / 87178291200.0;
  pc = pc * p2 - 1.0 / 479001600.0;
  pc = pc * p2 + 1.0 / 3628800.0;
  pc = pc * p2 - 1.0 / 40320.0;
  pc = pc * p2 + 1.0 / 720.0;
  pc = pc * p2 - 1.0 / 24.0;
  pc = pc * p2 + 0.5;
  const double cs = 1.0 - p2 * pc;
  const long qq = q & 3;
  ct_cs w{0.0, 0.0};
  w.c = qq == 0 ? cs : qq == 1 ? -sn : qq == 2 ? -cs : sn;
  w.s = qq == 0 ? sn : qq == 1 ? cs : qq == 2 ? -sn : -cs;
  return w;
}
// cos and sin of 2 pi j / R, j < R
template <int R>
struct UnitTab {
  double c[R], s[R];
  constexpr UnitTab() : c(), s() {
    for (int j = 0; j < R; ++j) {
      const ct_cs w = ct_unit(j, R);
      c[j] = w.c;
      s[j] = w.s;
    }
  }
};

// ---------------------------------------------------------------------------
// Forward DFT kernels on registers, natural order in/out:
// X_q = sum_m a_m exp(-2 pi i m q / R). Radix 2, 4, 8, 16 by the kernels
// below; an odd radix by its conjugate-symmetric pairs (DftOdd); a radix
// 2^k * odd (k >= 1) by the prime-factor map over the two (DftPfa).
// ---------------------------------------------------------------------------
FWD void dft4(cpx& a0, cpx& a1, cpx& a2, cpx& a3) {
  const cpx t0 = cadd(a0, a2), t1 = csub(a0, a2);
  const cpx t2 = cadd(a1, a3), t3 = cmul_mi(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

// Odd R >= 3. With s_m = a_m + a_{R-m}, d_m = a_m - a_{R-m} (m = 1..H,
// H = (R-1)/2) and phi = 2 pi m q / R:
//   X_0 = a_0 + sum_m s_m,
//   X_q = c_q - i e_q,  X_{R-q} = c_q + i e_q,  q = 1..H,
//   c_q = a_0 + sum_m s_m cos(phi),  e_q = sum_m d_m sin(phi).
template <int R> struct Dft;

template <int R>
struct DftOdd {
  static_assert(R >= 3 && R % 2 == 1, "odd radix");
  FWD static void run(cpx* a) {
    constexpr int H = (R - 1) / 2;
    constexpr UnitTab<R> U;
    cpx s[H + 1], d[H + 1];
    cpx x0 = a[0];
#pragma unroll
    for (int m = 1; m <= H; ++m) {
      s[m] = cadd(a[m], a[R - m]);
      d[m] = csub(a[m], a[R - m]);
      x0 = cadd(x0, s[m]);
    }
    // the outputs q, R - q overwrite inputs already folded into s and d; a[0]
    // is read by every q and written last
#pragma unroll
    for (int q = 1; q <= H; ++q) {
      cpx c = a[0], e = mk(0, 0);
#pragma unroll
      for (int m = 1; m <= H; ++m) {
        c = cadd(c, kscale(s[m], U.c[(m * q) % R]));
        e = cadd(e, kscale(d[m], U.s[(m * q) % R]));
      }
      a[q] = mk(c.x + e.y, c.y - e.x);
      a[R - q] = mk(c.x - e.y, c.y + e.x);
    }
    a[0] = x0;
  }
};

constexpr int inv_mod(int x, int m) {
  for (int y = 0; y < m; ++y)
    if ((x * y) % m == 1 % m) return y;
  return 0;
}

// R = N1 N2 with N1, N2 coprime (Good-Thomas): input n = (N2 n1 + N1 n2) mod R
// and output k = (N2 s k1 + N1 u k2) mod R, s = N2^-1 mod N1, u = N1^-1 mod N2,
// turn the DFT of length R into N2 DFTs of length N1 followed by N1 of length
// N2, with no twiddles (every index is a compile-time constant).
template <int N1, int N2>
struct DftPfa {
  FWD static void run(cpx* a) {
    constexpr int R = N1 * N2;
    constexpr int S = inv_mod(N2 % N1, N1), U = inv_mod(N1 % N2, N2);
    cpx y[N2][N1];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2)
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) y[n2][n1] = a[(N2 * n1 + N1 * n2) % R];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) Dft<N1>::run(y[n2]);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      cpx z[N2];
#pragma unroll
      for (int n2 = 0; n2 < N2; ++n2) z[n2] = y[n2][k1];
      Dft<N2>::run(z);
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) a[(N2 * S * k1 + N1 * U * k2) % R] = z[k2];
    }
  }
};

// radices other than the powers of two specialised below
template <int R>
struct Dft {
  static constexpr int A = R & -R, B = R / A;   // power-of-two and odd parts
  static_assert(B > 1 && A <= 16, "radix: 2^k * odd, 2^k <= 16");
  FWD static void run(cpx* a) {
    if constexpr (A == 1) DftOdd<R>::run(a);
    else DftPfa<A, B>::run(a);
  }
};

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

// Twiddle table of length TWN as staged in LDS. Full (TWO = false): all TWN
// entries, padded (twpad). Two-level (TWO = true, FP64 only): w_j = A[j / S]
// B[j % S] with A[h] = w_{h S}, B[l] = w_l, S = min(64, TWN): TWN/S + S
// entries (1.5 KiB at TWN = 2048 instead of 32 KiB) for one more complex
// product per twiddle. Entry e of the staged table is entry src(e) of the
// global full table, stored at dst(e).
template <int TWN, int MODE>
struct TwTab {
  static constexpr int COUNT = TWN;
  static constexpr int SIZE = TWN + TWN / 16 + 1;
  FWD static int src(int e) { return e; }
  FWD static int dst(int e) { return twpad(e); }
  FWD static tw_t get(const tw_t* tw, int idx) { return tw[twpad(idx)]; }
};
#if !FW_FP32
// MODE 2 (FP64): no table; w_idx = exp(-2 pi i idx / TWN), 0 <= idx < TWN.
// The angle is reduced exactly in integers to the nearest quarter turn q and a
// remainder r in [-TWN/8, TWN/8], so phi = 2 pi r / TWN lies in [-pi/4, pi/4];
// sin and cos of phi by their Taylor series to phi^17 and phi^18 (truncation
// below 1e-19 there), then the quarter-turn rotation. Branch-free; about one
// ulp (checked on the CPU against 120-bit references for TWN = 8 .. 4096).
template <int TWN>
struct TwTab<TWN, 2> {
  static constexpr int COUNT = 0;
  static constexpr int SIZE = 1;
  FWD static int src(int e) { return e; }
  FWD static int dst(int e) { return e; }
  FWD static tw_t get(const tw_t*, int idx) {
    static_assert(TWN >= 4 && TWN % 2 == 0, "even twiddle table length");
    int q, r;
    double phi;
    if constexpr ((TWN & (TWN - 1)) == 0) {
      constexpr int QN = TWN / 4;
      q = (idx + QN / 2) / QN;
      r = idx - q * QN;
      phi = (double)r * (2.0 * 3.14159265358979323846 / TWN);
    } else {
      // quarter turns of 2 pi idx / TWN in units of 1/(4 TWN): r in [-TWN/2, TWN/2]
      q = (4 * idx + TWN / 2) / TWN;
      r = 4 * idx - q * TWN;
      phi = (double)r * (3.14159265358979323846 / (2.0 * TWN));
    }
    const double p2 = phi * phi;
    double sp = 1.0 / 355687428096000.0;                 // 1/17!
    sp = sp * p2 - 1.0 / 1307674368000.0;                 // 1/15!
    sp = sp * p2 + 1.0 / 6227020800.0;                    // 1/13!
    sp = sp * p2 - 1.0 / 39916800.0;                      // 1/11!
    sp = sp * p2 + 1.0 / 362880.0;                        // 1/9!
    sp = sp * p2 - 1.0 / 5040.0;
    sp = sp * p2 + 1.0 / 120.0;
    sp = sp * p2 - 1.0 / 6.0;
    const double sn = phi + phi * p2 * sp;
    double cp = 1.0 / 6402373705728000.0;                 // 1/18!
    cp = cp * p2 - 1.0 / 20922789888000.0;                // 1/16!
    cp = cp * p2 + 1.0 / 87178291200.0;                   // 1/14!
    cp = cp * p2 - 1.0 / 479001600.0;                     // 1/12!
    cp = cp * p2 + 1.0 / 3628800.0;                       // 1/10!
    cp = cp * p2 - 1.0 / 40320.0;
    cp = cp * p2 + 1.0 / 720.0;
    cp = cp * p2 - 1.0 / 24.0;
    cp = cp * p2 + 0.5;
    const double cs = 1.0 - p2 * cp;
    // angle = q pi/2 + phi; exp(-i angle) = (cos angle, -sin angle)
    const int qq = q & 3;
    const double ca = (qq == 0) ? cs : (qq == 1) ? -sn : (qq == 2) ? -cs : sn;
    const double sa = (qq == 0) ? sn : (qq == 1) ? cs : (qq == 2) ? -sn : -cs;
    return mk(ca, -sa);
  }
};
template <int TWN>
struct TwTab<TWN, 1> {
  static constexpr int S = TWN < 64 ? TWN : 64;
  static constexpr int NA = TWN / S;
  static constexpr int COUNT = NA + S;
  static constexpr int SIZE = NA + S;
  FWD static int src(int e) { return e < NA ? e * S : e - NA; }
  FWD static int dst(int e) { return e; }
  FWD static tw_t get(const tw_t* tw, int idx) { return cmul(tw[idx / S], tw[NA + (idx % S)]); }
};
#endif

// Loads the staged entries of a TwTab from global memory into registers (issue
// early) and stores them into LDS (before the first barrier of the FFT).
template <class T, int NT>
struct TwStage {
  static constexpr int R = (T::COUNT + NT - 1) / NT;   // 0: nothing staged
  tw_t r[R > 0 ? R : 1];
  // COUNT a multiple of NT (the usual case): every e = tid + i NT (i < R) is
  // below COUNT, so no bounds test, which would make the compiler keep r[] in
  // scratch memory
  static constexpr bool EVEN = T::COUNT % NT == 0;
  FWD void load(const tw_t* __restrict__ g, int tid) {
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int e = tid + i * NT;
      if (EVEN || e < T::COUNT) r[i] = g[T::src(e)];
    }
  }
  FWD void store(tw_t* l, int tid) const {
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int e = tid + i * NT;
      if (EVEN || e < T::COUNT) l[T::dst(e)] = r[i];
    }
  }
};

// ---------------------------------------------------------------------------
// Line FFT engine: P lines of length N per block, E complex elements per thread.
// LM (line-major) lane mapping: t = tid % TPL, q = tid / TPL (rows, contiguous);
// interleaved mapping: q = tid % P, t = tid / P (columns: a wave reads 2P
// adjacent columns of several rows per instruction).
// ---------------------------------------------------------------------------
// TWN: length of the twiddle table passed in (a table for a multiple of N
// serves N with stride TWN/N; the batched kernels share one table).
// Radix of the next Stockham stage over the remaining length rem with E
// elements per thread: for powers of two the largest power of two up to 16
// dividing both; otherwise the largest radix up to min(E, 32) dividing both
// (its power-of-two part at most 16). A schedule is valid when such a radix
// exists at every stage; the host checks this with the same rule.
constexpr int next_radix(int rem, int e) {
  if ((rem & (rem - 1)) == 0 && (e & (e - 1)) == 0) {
    for (int r = 16; r >= 2; r /= 2)
      if (rem % r == 0 && e % r == 0) return r;
  }
  for (int r = e < 32 ? e : 32; r >= 2; --r)
    if (rem % r == 0 && e % r == 0 && (r & -r) <= 16) return r;
  return rem;
}

// A value materialised in registers at this point of the program (the
// compiler may neither sink its computation below nor recompute it later).
FWD cpx pin(cpx c) {
  asm volatile("" : "+v"(c.x), "+v"(c.y));
  return c;
}

// Lengths: any even N. Powers of two take radix-2/4/8/16 stages and the
// pair-fused spectral step; other lengths add odd-radix stages (every radix
// dividing E) and take the spectral step through LDS.
template <int N, int E, int P, bool LM, int TWN = N, int TWO = 0>
struct Line {
  using TW = TwTab<TWN, TWO>;
  static constexpr bool POW2 = (N & (N - 1)) == 0;
  static_assert(N >= 2 && N % 2 == 0, "line length must be even");
  static_assert(E <= N && N % E == 0 && (POW2 || E % 2 == 0), "elements per thread must divide N (and be even)");
  static constexpr int LN = N, LE = E;   // line length, elements per thread
  static constexpr int TPL = N / E;
  static constexpr int NT = P * TPL;
  static constexpr int RMAX = E < 16 ? E : 16;
  struct Plan;   // base twiddles of a spectral step, computed ahead (below)
  static constexpr int RAW = P * N;
#ifndef FW_LDS_SWZ
#define FW_LDS_SWZ 0
#endif
#if FW_LDS_SWZ == 1
  static constexpr int LDS = RAW + (RAW >> 3) + 1;
#elif FW_LDS_SWZ == 2
  static constexpr int LDS = RAW;
#else
  static constexpr int LDS = RAW + (RAW >> 4) + 1;
#endif
  static constexpr bool WAVE_LOCAL = (NT <= 64) || (LM && TPL <= 64 && (64 % TPL) == 0);

  FWD static int lidx(int q, int i) {
    const int r = LM ? q * N + i : i * P + q;
#if FW_LDS_SWZ == 1
    return r + (r >> 3);
#elif FW_LDS_SWZ == 2
    return r ^ ((r >> 3) & 7);
#else
    return r + (r >> 4);
#endif
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
  // PRE: the base twiddle of butterfly b is wb[b] (computed ahead by plan)
  template <int R, int NS, bool TOREG, bool PRE = false>
  FWD static void stage(cpx (&v)[E], cpx* lds, int q, int t, const tw_t* __restrict__ tw,
                        const cpx* wb = nullptr) {
    constexpr int B = E / R;
    constexpr int TWS = (N / (NS * R)) * (TWN / N);
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int j = t + TPL * b;
      cpx a[R];
#pragma unroll
      for (int m = 0; m < R; ++m) a[m] = v[b + m * B];
      if constexpr (NS > 1) {
        const int k = ((NS & (NS - 1)) == 0) ? (j & (NS - 1)) : j % NS;
        if constexpr (TWO == 2 && R > 16) {
          // radices above 16 (lines with odd factors): powers by a running
          // product, two twiddles live instead of R
          cpx w1;
          if constexpr (PRE) w1 = wb[b];
          else w1 = TW::get(tw, k * TWS);
          cpx wm = w1;
#pragma unroll
          for (int m = 1; m < R; ++m) {
            a[m] = cmul(a[m], wm);
            if (m + 1 < R) wm = cmul(wm, w1);
          }
        } else if constexpr (TWO == 2) {
          cpx w[R];
          if constexpr (PRE) w[1] = wb[b];
          else w[1] = TW::get(tw, k * TWS);
#pragma unroll
          for (int m = 2; m < R; ++m) w[m] = cmul(w[m / 2], w[m - m / 2]);
#pragma unroll
          for (int m = 1; m < R; ++m) a[m] = cmul(a[m], w[m]);
        } else {
#pragma unroll
          for (int m = 1; m < R; ++m) a[m] = twmul(a[m], TW::get(tw, m * k * TWS));
        }
      }
      Dft<R>::run(a);
      if constexpr (TOREG) {
#pragma unroll
        for (int m = 0; m < R; ++m) v[b + m * B] = a[m];
      } else {
        // (j / NS) NS R + j mod NS; a mask for a power-of-two NS
        const int base = ((NS & (NS - 1)) == 0) ? (j & ~(NS - 1)) * R + (j & (NS - 1)) : (j / NS) * NS * R + j % NS;
#pragma unroll
        for (int m = 0; m < R; ++m) lds[lidx(q, base + m * NS)] = a[m];
      }
    }
  }

  // Forward FFT. TOLDS: leave the spectrum in LDS in natural order (synced);
  // otherwise leave it in registers at the natural positions t + TPL e.
  // PL: the base twiddles of stage I (counting from NS) come from pl->inv
  // (the inverse of a planned spectral step, which starts at NS = RL > 1)
  template <bool TOLDS, int NS = 1, bool PL = false, int I = 0>
  FWD static void fwd(cpx (&v)[E], cpx* lds, int q, int t, const tw_t* __restrict__ tw,
                      const Plan* pl = nullptr) {
    constexpr int REM = N / NS;
    constexpr int R = POW2 ? (REM < RMAX ? REM : RMAX) : next_radix(REM, E);
    static_assert(E % R == 0, "every stage radix must divide E");
    static_assert(!PL || NS > 1, "planned stages are twiddled");
    constexpr bool LAST = (NS * R == N);
    const cpx* wb = nullptr;
    if constexpr (PL) wb = pl->inv[I];
    if constexpr (LAST && !TOLDS) {
      stage<R, NS, true, PL>(v, lds, q, t, tw, wb);
    } else {
      stage<R, NS, false, PL>(v, lds, q, t, tw, wb);
      sync();
      if constexpr (!LAST) {
        load_natural(v, lds, q, t);
        sync();
        fwd<TOLDS, NS * R, PL, I + 1>(v, lds, q, t, tw, pl);
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

  // ---- Pair-fused spectral step: forward FFT, multiply, inverse FFT ----
  // The separation of two packed real lines pairs spectral positions k and
  // N - k. The forward FFT's last stage (radix RL, butterflies j < S = N/RL,
  // butterfly j holding positions j + m S) runs with a map that gives each
  // thread the butterflies j and S - j of a slot, i.e. conjugate partners, so
  // the pair multiply happens in registers; and the inverse FFT's first stage,
  // of radix RL too, takes exactly the positions j + m S of butterfly j as its
  // inputs. Against transforms with the multiply between them through LDS
  // (spectrum out, pair gather and write back, natural reload) this saves two
  // LDS rounds and three barriers per spectral step.
  static constexpr bool PAIRED = POW2 && E >= 4;
  static constexpr int RL = PAIRED ? (E / 2 < N / 2 ? E / 2 : N / 2) : 1;
  static constexpr int S = N / RL;
  static constexpr int BS = PAIRED ? E / (2 * RL) : 0;   // slots per thread
  // A thread has E/2 items (pairs; item 0 of thread 0 is the DC/Nyquist pair
  // of self-partners). item_k: the spectral index k <= N/2 whose factors item i
  // of thread t applies; thread 0 also applies k = N/2.
  FWD static int item_k(int t, int i) {
    const int sl = t + TPL * (i / RL), m = i % RL;
    if (sl == 0) return m < RL / 2 ? m * S : S / 2 + (m - RL / 2) * S;
    const int p = sl + m * S;
    return m < RL / 2 ? p : N - p;
  }
  // forward stages before the last one; result in LDS (natural order, synced)
  // PL: the base twiddles of the I-th twiddled stage come from pl->pre
  template <int NS, bool PL = false, int I = 0>
  FWD static void fwd_pre(cpx (&v)[E], cpx* lds, int q, int t, const tw_t* __restrict__ tw,
                          const Plan* pl = nullptr) {
    constexpr int REM = S / NS;
    constexpr int R = REM < RMAX ? REM : RMAX;
    constexpr bool TWD = PL && NS > 1;
    const cpx* wb = nullptr;
    if constexpr (TWD) wb = pl->pre[I];
    stage<R, NS, false, TWD>(v, lds, q, t, tw, wb);
    sync();
    if constexpr (NS * R < S) {
      load_natural(v, lds, q, t);
      sync();
      fwd_pre<NS * R, PL, (NS > 1 ? I + 1 : I)>(v, lds, q, t, tw, pl);
    }
  }
  // butterfly h (0, 1) of slot sl
  FWD static int slot_bfly(int sl, int h) { return sl == 0 ? (h == 0 ? 0 : S / 2) : (h == 0 ? sl : S - sl); }
  // MUL: pair(i, zk, zm) for item i (zk at k = item_k, zm at N - k) and
  // self(z0, zn) for DC and Nyquist (thread 0, with item 0).
  template <class MUL, bool PL = false>
  FWD static void spectral(cpx (&v)[E], cpx* lds, int q, int t, const tw_t* __restrict__ tw, MUL& mul,
                           const Plan* pl = nullptr) {
    static_assert(PAIRED, "the pair-fused step needs E >= 4");
    constexpr int TWS = TWN / N;
    fwd_pre<1, PL>(v, lds, q, t, tw, pl);
    // last forward stage, slot map; v[(2 b + h) RL + m] = output m of butterfly h of slot b
#pragma unroll
    for (int b = 0; b < BS; ++b) {
      const int sl = t + TPL * b;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int j = slot_bfly(sl, h);
        cpx a[RL];
#pragma unroll
        for (int m = 0; m < RL; ++m) a[m] = lds[lidx(q, j + m * S)];
        if constexpr (TWO == 2) {
          cpx w[RL];
          if constexpr (PL) w[1] = pl->pr[2 * b + h];
          else w[1] = TW::get(tw, j * TWS);
#pragma unroll
          for (int m = 2; m < RL; ++m) w[m] = cmul(w[m / 2], w[m - m / 2]);
#pragma unroll
          for (int m = 1; m < RL; ++m) a[m] = cmul(a[m], w[m]);
        } else {
#pragma unroll
          for (int m = 1; m < RL; ++m) a[m] = twmul(a[m], TW::get(tw, m * j * TWS));
        }
        Dft<RL>::run(a);
#pragma unroll
        for (int m = 0; m < RL; ++m) v[(2 * b + h) * RL + m] = a[m];
      }
    }
    sync();   // every read of this exchange precedes the inverse's writes
    // pair multiply in registers
#pragma unroll
    for (int b = 0; b < BS; ++b) {
      const int sl = t + TPL * b;
      cpx* y1 = v + 2 * b * RL;
      cpx* y2 = y1 + RL;
      if (sl == 0) {
        mul.self(y1[0], y1[RL / 2]);
#pragma unroll
        for (int m = 1; m < RL / 2; ++m) mul.pair(b * RL + m, y1[m], y1[RL - m]);
#pragma unroll
        for (int m = 0; m < RL / 2; ++m) mul.pair(b * RL + RL / 2 + m, y2[m], y2[RL - 1 - m]);
      } else {
#pragma unroll
        for (int m = 0; m < RL / 2; ++m) mul.pair(b * RL + m, y1[m], y2[RL - 1 - m]);
#pragma unroll
        for (int m = RL / 2; m < RL; ++m) mul.pair(b * RL + m, y2[RL - 1 - m], y1[m]);
      }
    }
    // inverse (conjugation identity): first stage, radix RL, NS = 1, on the slot butterflies
#pragma unroll
    for (int b = 0; b < BS; ++b) {
      const int sl = t + TPL * b;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int j = slot_bfly(sl, h);
        cpx a[RL];
#pragma unroll
        for (int m = 0; m < RL; ++m) a[m] = mk(v[(2 * b + h) * RL + m].x, -v[(2 * b + h) * RL + m].y);
        Dft<RL>::run(a);
#pragma unroll
        for (int m = 0; m < RL; ++m) lds[lidx(q, j * RL + m)] = a[m];
      }
    }
    sync();
    load_natural(v, lds, q, t);
    sync();
    fwd<false, RL, PL>(v, lds, q, t, tw, pl);
#pragma unroll
    for (int e = 0; e < E; ++e) v[e].y = -v[e].y;
  }

  // ---- Base twiddles of a pair-fused spectral step, computed ahead (FP64) ----
  // Every twiddled stage needs, per butterfly, one base twiddle that depends
  // on the thread's position in the line only. A pass computes them while its
  // loads of W are in flight (before the barrier that waits for them), so
  // their sine/cosine series overlap the load latency instead of following it.
  static constexpr int PSTG = 6;   // twiddled stages per direction, at most (N <= 4096, E >= 4)
  static constexpr bool PLANNED = PAIRED && TWO == 2;
  struct Plan {
    cpx pre[PSTG][E];                 // forward stages before the pair stage
    cpx pr[2 * (BS > 0 ? BS : 1)];    // the pair stage's butterflies
    cpx inv[PSTG][E];                 // the inverse's stages after its first
  };
  template <int NS, int I>
  FWD static void plan_pre(Plan& p, int t, const tw_t* __restrict__ tw) {
    constexpr int REM = S / NS;
    constexpr int R = REM < RMAX ? REM : RMAX;
    constexpr int B = E / R;
    constexpr int TWS = (N / (NS * R)) * (TWN / N);
    if constexpr (NS > 1) {
      static_assert(I < PSTG, "twiddled stages");
#pragma unroll
      for (int b = 0; b < B; ++b) p.pre[I][b] = pin(TW::get(tw, ((t + TPL * b) & (NS - 1)) * TWS));
    }
    if constexpr (NS * R < S) plan_pre<NS * R, (NS > 1 ? I + 1 : I)>(p, t, tw);
  }
  template <int NS, int I>
  FWD static void plan_inv(Plan& p, int t, const tw_t* __restrict__ tw) {
    constexpr int REM = N / NS;
    constexpr int R = REM < RMAX ? REM : RMAX;
    constexpr int B = E / R;
    constexpr int TWS = (N / (NS * R)) * (TWN / N);
    static_assert(I < PSTG, "twiddled stages");
#pragma unroll
    for (int b = 0; b < B; ++b) p.inv[I][b] = pin(TW::get(tw, ((t + TPL * b) & (NS - 1)) * TWS));
    if constexpr (NS * R < N) plan_inv<NS * R, I + 1>(p, t, tw);
  }
  FWD static void plan(Plan& p, int t, const tw_t* __restrict__ tw) {
    static_assert(PAIRED && TWO == 2, "planned twiddles: FP64 series, pair-fused step");
    plan_pre<1, 0>(p, t, tw);
#pragma unroll
    for (int b = 0; b < BS; ++b)
#pragma unroll
      for (int h = 0; h < 2; ++h) p.pr[2 * b + h] = pin(TW::get(tw, slot_bfly(t + TPL * b, h) * (TWN / N)));
    plan_inv<RL, 0>(p, t, tw);
  }
};

enum { NSER = 9 };   // sums per step published by the propagators

// Spectral index k <= N/2 of pair item i of thread t in a pass's spectral
// step: the pair-fused map of Line::spectral, or k = t + TPL i (E = 2).
template <class L>
FWD int spec_k(int t, int i) {
  if constexpr (L::PAIRED) return L::item_k(t, i);
  else return t + L::TPL * i;
}

// Spectral step of a pass: forward FFT, mul.pair on each conjugate pair
// (k, N - k) and mul.self on DC and Nyquist, inverse FFT; registers to
// registers at the natural positions. E >= 4: the pair-fused step; E = 2:
// separate transforms with the multiply through LDS.
// The pair-fused spectral step with base twiddles computed ahead (L::plan).
template <class L, int E, class MUL>
FWD void spectral_step_planned(cpx (&v)[E], cpx* lds, int q, int t, const tw_t* __restrict__ twl, MUL& mul,
                               const typename L::Plan& pl) {
  L::template spectral<MUL, true>(v, lds, q, t, twl, mul, &pl);
}

template <class L, int E, class MUL>
FWD void spectral_step(cpx (&v)[E], cpx* lds, int q, int t, const tw_t* __restrict__ twl, MUL& mul) {
  if constexpr (L::PAIRED) {
    L::spectral(v, lds, q, t, twl, mul);
  } else {
    constexpr int N = L::TPL * E, TPL = L::TPL;
    L::template fwd<true>(v, lds, q, t, twl);
#pragma unroll
    for (int e = 0; e < E / 2; ++e) {
      const int k = t + TPL * e;
      const int ik = L::lidx(q, k), im = L::lidx(q, k == 0 ? N / 2 : N - k);
      cpx zk = lds[ik], zm = lds[im];
      if (k == 0) mul.self(zk, zm);
      else mul.pair(e, zk, zm);
      lds[ik] = zk;
      lds[im] = zm;
    }
    L::sync();
    L::load_natural(v, lds, q, t);
    L::sync();
    L::inv(v, lds, q, t, twl);
  }
}


#if FW_FP32
// sin and cos of |x| <= pi/4 + 1e-6 by their Taylor series through x^9 and
// x^10 (truncation below 2.5e-9 relative, under half an FP32 ulp), Horner in
// FP32. The argument comes already reduced, so no large-argument path is
// needed.
FWD void sincos_quarter(float x, float& s, float& c) {
  const float x2 = x * x;
  s = x * (1.0f + x2 * (-1.0f / 6.0f + x2 * (1.0f / 120.0f + x2 * (-1.0f / 5040.0f + x2 * (1.0f / 362880.0f)))));
  c = 1.0f + x2 * (-0.5f + x2 * (1.0f / 24.0f + x2 * (-1.0f / 720.0f + x2 * (1.0f / 40320.0f +
                                                                            x2 * (-1.0f / 3628800.0f)))));
}
#endif

FWD cpx unit_phase(double ph, double m) {
#if FW_FP32
  // FP32 form: the argument and its reduction to [-pi/4, pi/4] (quadrant q) in
  // FP64, the series in FP32 with a first-order correction for the rounding of
  // the reduced argument, then the quadrant rotation.
  const double pio2 = 1.57079632679489661923;
  const double k = rint(ph * 0.63661977236758134308);
  const double r = fma(-k, pio2, ph);
  const float rf = (float)r;
  const float rl = (float)(r - (double)rf);
  float s, c;
  sincos_quarter(rf, s, c);
  const float cc = c - s * rl, ss = s + c * rl;
  const int q = (int)((long long)k & 3);
  const float cq = (q == 0) ? cc : (q == 1) ? -ss : (q == 2) ? -cc : ss;
  const float sq = (q == 0) ? ss : (q == 1) ? cc : (q == 2) ? -ss : -cc;
  const float mf = (float)m;
  return mk(cq * mf, sq * mf);
#else
  double s, c;
  sincos_fp64(ph, s, c);
  return mk(c * m, s * m);
#endif
}

// Table entries are built once, so they take the FP64 sincos in both forms and
// are rounded once to the storage precision (in FP64 they equal unit_phase).
FWD cpx unit_phase_exact(double ph, double m) {
  double s, c;
  sincos_fp64(ph, s, c);
  return mk((real_t)(c * m), (real_t)(s * m));
}

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------
// Lane-to-lane move of a double through DPP (two 32-bit moves, same control).
template <int CTRL>
FWD double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
FWD double read_lane(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// Sum over the 64 lanes of a full wave, the same bits in every lane. Within
// each row of 16 lanes by DPP (quad_perm [1,0,3,2], quad_perm [2,3,0,1],
// row_half_mirror, row_mirror: VALU moves, no LDS traffic), then the four row
// sums read from lanes 0, 16, 32, 48. A cross-lane shuffle through
// ds_bpermute costs two LDS instructions per step and double.
FWD double wave_sum(double v) {
  v += dpp_move<0xB1>(v);
  v += dpp_move<0x4E>(v);
  v += dpp_move<0x141>(v);
  v += dpp_move<0x140>(v);
  return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

// Block reductions over the NT threads of a block. NT >= 64: each wave adds
// its lanes with cross-lane reads and lane 0 stores one slot per wave. NT < 64:
// the block is one partial wave whose missing lanes hold undefined data that a
// cross-lane read would pick up, so every thread stores its own slot. After a
// barrier, total() adds the RN slots of quantity i in order (the same bits in
// every block). red: >= ns * RN doubles.
// Sum over the first L < 64 lanes of a wave whose other lanes do not exist
// (the last wave of a block of NT threads, NT % 64 = L): the DPP row moves
// read lanes outside the wave's execution mask as 0 (bound_ctrl), and only the
// rows that hold lanes of the wave are read.
template <int L>
FWD double wave_sum_prefix(double v) {
  v += dpp_move<0xB1>(v);
  v += dpp_move<0x4E>(v);
  v += dpp_move<0x141>(v);
  v += dpp_move<0x140>(v);
  double s = read_lane(v, 0);
  if constexpr (L > 16) s += read_lane(v, 16);
  if constexpr (L > 32) s += read_lane(v, 32);
  if constexpr (L > 48) s += read_lane(v, 48);
  return s;
}

template <int NT>
struct BlockRed {
  static constexpr int RN = NT >= 64 ? (NT + 63) / 64 : NT;
  FWD static void put(double v, double* red, int ns, int i) {
    if constexpr (NT >= 64) {
      if constexpr (NT % 64 == 0) {
        v = wave_sum(v);
      } else {
        // blocks of lines whose length has an odd factor: a partial last wave
        if ((int)(threadIdx.x >> 6) < NT / 64) v = wave_sum(v);
        else v = wave_sum_prefix<NT % 64>(v);
      }
      if ((threadIdx.x & 63) == 0) red[(threadIdx.x >> 6) * ns + i] = v;
    } else {
      red[threadIdx.x * ns + i] = v;
    }
  }
  FWD static double total(const double* red, int ns, int i) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < RN; ++w) t += red[w * ns + i];
    return t;
  }
};

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
// red: >= BlockRed<NT>::RN doubles.
template <int NT, int NS = NSER>
FWD double sum_partials(const double* __restrict__ partials, int nb, int i, double* red) {
  double sum = 0.0, comp = 0.0;
  for (int b = threadIdx.x; b < nb; b += NT) {
    const double x = partials[(size_t)b * NS + i];
    const double tt = sum + x;
    comp += (fabs(sum) >= fabs(x)) ? (sum - tt) + x : (x - tt) + sum;
    sum = tt;
  }
  __syncthreads();
  BlockRed<NT>::put(sum + comp, red, 1, 0);
  __syncthreads();
  return BlockRed<NT>::total(red, 1, 0);
}

// Block sums of the last V pass -> partials[blk] (plain stores; blocks with
// fewer threads than sums store several). red: >= NS * BlockRed<NT>::RN doubles.
template <int NT, int NS = NSER>
FWD void write_partials(double (&s)[NS], double* red, double* partials, int blk) {
#pragma unroll
  for (int i = 0; i < NS; ++i) BlockRed<NT>::put(s[i], red, NS, i);
  __syncthreads();
  for (int i = threadIdx.x; i < NS; i += NT) partials[(size_t)blk * NS + i] = BlockRed<NT>::total(red, NS, i);
}

// Block sums stored by sum: sum i of block blk at partials[i * nbp + blk], so
// that a row (one sum over all nbp blocks) is contiguous for the readers that
// add it. red: >= NS * BlockRed<NT>::RN doubles.
template <int NT, int NS = NSER>
FWD void write_partial_rows(double (&s)[NS], double* red, double* partials, int blk, int nbp) {
#pragma unroll
  for (int i = 0; i < NS; ++i) BlockRed<NT>::put(s[i], red, NS, i);
  __syncthreads();
  for (int i = threadIdx.x; i < NS; i += NT) partials[(size_t)i * nbp + blk] = BlockRed<NT>::total(red, NS, i);
}

FWD double wave_min(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}

}  // namespace fwk
