// fw_fno.cpp -- Fourier-neural-operator surrogate of the Wigner propagator:
// forward, backward and Adam on gfx950, compiled ahead of time into
// libfastwigner.so. Contract: include/fastwigner.h (fw_fno_*).
//
// Network (N x N grid, width C, retained modes M x M, L layers):
//   features f(p, x) = [W, x/X_amplitude, p/P_amplitude, a2, a3, a4, gamma, kT]
//   v0 = f Pw + pb                                   (lift, 8 -> C)
//   v_{l+1} = act_l(S_l(v_l) + v_l Wp_l + b_l)       (act = GELU for l < L-1, identity for the last)
//   c = v_L Q + qb                                   (projection, C -> 1)
//   out = W + s c, or W + s (c - mean c) per sample  (residual on W; output scale s, 1 by
//                                                     default; zero-mass correction optional)
// S_l is the truncated spectral convolution: the 2-D DFT of each channel kept
// on the modes ky in [0, M) (the real-input half axis, along x) and
// kp in [0, M) U [N-M, N) (along p), mixed per mode by complex C x C weights
// R_l[m], and transformed back as numpy's irfft2 of that truncated spectrum.
//
// With only M of N modes kept, each truncated transform is a dense product
// with a (2M x N) or (N x 2M) cosine/sine matrix: every stage is a GEMM. DFT
// stages run on FP32 MFMA (v_mfma_f32_16x16x4f32); the channel mixing (per-mode
// complex weights, pointwise C x C) runs on BF16 MFMA
// (v_mfma_f32_16x16x16bf16_1k) with BF16 weights and BF16-rounded activations,
// FP32 accumulation. Activations are channels-last, [b][p][x][c], FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "fastwigner.h"
#include "fw_internal.h"

typedef unsigned short bf16_t;
typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

namespace fno {

// --------------------------------------------------------------------------
// bf16 (round to nearest even; inputs are finite)
// --------------------------------------------------------------------------
__device__ __forceinline__ bf16_t to_bf16(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(bf16_t h) { return __uint_as_float((unsigned int)h << 16); }

// --------------------------------------------------------------------------
// Strided, batched GEMM  C[z] = alpha A[z] B[z] + beta C[z]
//   A(m, k) at A + z1 sA1 + z2 sA2 + m sAm + k sAk, likewise B(k, n), C(m, n);
//   z = z1 * nb2 + z2 = blockIdx.z.
// 64 x 64 tile per 256-thread block, each wave a 32 x 32 quadrant of 2 x 2
// MFMA 16 x 16 blocks; K in steps of 32 staged through LDS. Out-of-range
// elements load as zero; stores are masked.
// MFMA operand layout (lane l, 16 x 16 blocks): A rows i = l % 16 with
// k = 4 (l / 16) + 0..3 (bf16) or k = l / 16 (f32); B columns j = l % 16 with
// the same k; the accumulator holds rows 4 (l / 16) + 0..3 of column l % 16.
// --------------------------------------------------------------------------
struct GemmArgs {
  const void* A;
  const void* B;
  float* C;
  int M, N, K, nb2;
  long long sAm, sAk, sBk, sBn, sCm, sCn;
  long long sA1, sA2, sB1, sB2, sC1, sC2;
  float alpha, beta;
};

constexpr int GT = 64, GK = 32, GPAD = 4;

template <typename T> __device__ __forceinline__ float ld(const T* p, long long i);
template <> __device__ __forceinline__ float ld<float>(const float* p, long long i) { return p[i]; }
template <> __device__ __forceinline__ float ld<bf16_t>(const bf16_t* p, long long i) { return from_bf16(p[i]); }

// BF16: operands rounded to bf16 in LDS; F32: full FP32 operands.
template <typename TA, typename TB, bool BF>
__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs g) {
  typedef typename std::conditional<BF, bf16_t, float>::type ST;
  __shared__ ST As[GT][GK + GPAD];
  __shared__ ST Bs[GT][GK + GPAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int z = blockIdx.z, z1 = z / g.nb2, z2 = z % g.nb2;
  const TA* A = (const TA*)g.A + z1 * g.sA1 + z2 * g.sA2;
  const TB* B = (const TB*)g.B + z1 * g.sB1 + z2 * g.sB2;
  float* C = g.C + z1 * g.sC1 + z2 * g.sC2;
  const bool a_kfast = g.sAk == 1, b_nfast = g.sBn == 1;

  float4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (float4_t){0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < g.K; k0 += GK) {
    // stage A (64 x 32) and B (32 x 64, stored n-major): 8 elements per thread each
#pragma unroll
    for (int r = 0; r < (GT * GK) / 256; ++r) {
      const int e = tid + 256 * r;
      int mm, kk;
      if (a_kfast) { mm = e / GK; kk = e % GK; } else { mm = e % GT; kk = e / GT; }
      const int gm = m0 + mm, gk = k0 + kk;
      float v = 0.f;
      if (gm < g.M && gk < g.K) v = ld<TA>(A, gm * g.sAm + gk * g.sAk);
      if constexpr (BF) As[mm][kk] = to_bf16(v); else As[mm][kk] = v;
      int nn;
      if (b_nfast) { nn = e % GT; kk = e / GT; } else { nn = e / GK; kk = e % GK; }
      const int gn = n0 + nn, gk2 = k0 + kk;
      float w = 0.f;
      if (gn < g.N && gk2 < g.K) w = ld<TB>(B, gk2 * g.sBk + gn * g.sBn);
      if constexpr (BF) Bs[nn][kk] = to_bf16(w); else Bs[nn][kk] = w;
    }
    __syncthreads();
    if constexpr (BF) {
#pragma unroll
      for (int kk = 0; kk < GK; kk += 16) {
        short4_t af[2], bfr[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const ST* pa = &As[wm * 32 + i * 16 + (lane & 15)][kk + 4 * (lane >> 4)];
          af[i] = (short4_t){(short)pa[0], (short)pa[1], (short)pa[2], (short)pa[3]};
          const ST* pb = &Bs[wn * 32 + i * 16 + (lane & 15)][kk + 4 * (lane >> 4)];
          bfr[i] = (short4_t){(short)pb[0], (short)pb[1], (short)pb[2], (short)pb[3]};
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i], bfr[j], acc[i][j], 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int kk = 0; kk < GK; kk += 4) {
        float af[2], bfr[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          af[i] = As[wm * 32 + i * 16 + (lane & 15)][kk + (lane >> 4)];
          bfr[i] = Bs[wn * 32 + i * 16 + (lane & 15)][kk + (lane >> 4)];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i], bfr[j], acc[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + r;
        const int gn = n0 + wn * 32 + j * 16 + (lane & 15);
        if (gm < g.M && gn < g.N) {
          float* pc = C + gm * g.sCm + gn * g.sCn;
          const float v = g.alpha * acc[i][j][r];
          *pc = (g.beta != 0.f) ? v + g.beta * *pc : v;
        }
      }
}

// --------------------------------------------------------------------------
// Elementwise and reduction kernels
// --------------------------------------------------------------------------
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float x) {
  return 0.5f * (1.f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * __expf(-0.5f * x * x);
}

// v0[n][c] = sum_j Pw[j][c] f_j(n) + pb[c]; f = [W, x/Xa, p/Pa, feat[b][0..4]].
// One thread per output (n, c): consecutive threads write consecutive channels.
__global__ void __launch_bounds__(256) lift_fwd(const float* W, const float* feat, const float* Pw, const float* pb,
                                                float* v0, int B, int N, int C, float xs, float ps) {
  const long long total = (long long)B * N * N * C;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const long long n = i / C;
    const int c = (int)(i % C);
    const int b = (int)(n / ((long long)N * N));
    const int rem = (int)(n % ((long long)N * N));
    const int p = rem / N, x = rem % N;
    float f[8];
    f[0] = W[n];
    f[1] = xs * (x - 0.5f * N);
    f[2] = ps * (p - 0.5f * N);
#pragma unroll
    for (int j = 0; j < 5; ++j) f[3 + j] = feat[b * 5 + j];
    float s = pb[c];
#pragma unroll
    for (int j = 0; j < 8; ++j) s += Pw[j * C + c] * f[j];
    v0[i] = s;
  }
}

// per-block partial sums of f_j(n) dv0[n][c] (j < 8) and dv0[n][c]: 9 C values per block
__global__ void __launch_bounds__(256) lift_bwd_partial(const float* W, const float* feat, const float* dv0, float* part,
                                                        int B, int N, int C, float xs, float ps) {
  __shared__ float red[256];
  const long long total = (long long)B * N * N;
  // thread t accumulates for channel c = t % C over points n = t / C + k (256 / C)
  const int c = threadIdx.x % C, lanes = 256 / C, li = threadIdx.x / C;
  float acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (long long n = (long long)blockIdx.x * lanes + li; n < total; n += (long long)gridDim.x * lanes) {
    const int b = (int)(n / ((long long)N * N));
    const int rem = (int)(n % ((long long)N * N));
    const int p = rem / N, x = rem % N;
    const float d = dv0[n * C + c];
    acc[0] += W[n] * d;
    acc[1] += xs * (x - 0.5f * N) * d;
    acc[2] += ps * (p - 0.5f * N) * d;
#pragma unroll
    for (int j = 0; j < 5; ++j) acc[3 + j] += feat[b * 5 + j] * d;
    acc[8] += d;
  }
  for (int j = 0; j < 9; ++j) {
    red[threadIdx.x] = acc[j];
    __syncthreads();
    if (threadIdx.x < C) {
      float s = 0.f;
      for (int q = 0; q < lanes; ++q) s += red[q * C + threadIdx.x];
      part[((long long)blockIdx.x * 9 + j) * C + threadIdx.x] = s;
    }
    __syncthreads();
  }
}

// out[r] = sum over z < nz of part[z][r]  (fixed order)
__global__ void __launch_bounds__(256) reduce_rows(const float* part, float* out, int nz, int nr, float scale,
                                                   int accumulate) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= nr) return;
  float s = 0.f;
  for (int z = 0; z < nz; ++z) s += part[(long long)z * nr + r];
  out[r] = accumulate ? out[r] + scale * s : scale * s;
}

// out[n] = W[n] + scale (sum_c Q[c] v[n][c] + qb)
__global__ void __launch_bounds__(256) proj_fwd(const float* W, const float* v, const float* Q, const float* qb,
                                                float* out, long long total, int C, float scale) {
  for (long long n = (long long)blockIdx.x * 256 + threadIdx.x; n < total; n += (long long)gridDim.x * 256) {
    float s = qb[0];
    for (int c = 0; c < C; ++c) s += Q[c] * v[n * C + c];
    out[n] = W[n] + scale * s;
  }
}

// with d = scale dout[n]: dv[n][c] = Q[c] d; partials of sum_n v[n][c] d (C) and sum d (1) per block
__global__ void __launch_bounds__(256) proj_bwd(const float* v, const float* dout, const float* Q, float* dv,
                                                float* part, long long total, int C, float scale) {
  __shared__ float red[256];
  const int c = threadIdx.x % C, lanes = 256 / C, li = threadIdx.x / C;
  float aq = 0.f, ab = 0.f;
  for (long long n = (long long)blockIdx.x * lanes + li; n < total; n += (long long)gridDim.x * lanes) {
    const float d = scale * dout[n];
    dv[n * C + c] = Q[c] * d;
    aq += v[n * C + c] * d;
    if (c == 0) ab += d;
  }
  red[threadIdx.x] = aq;
  __syncthreads();
  if (threadIdx.x < C) {
    float s = 0.f;
    for (int q = 0; q < lanes; ++q) s += red[q * C + threadIdx.x];
    part[(long long)blockIdx.x * (C + 1) + threadIdx.x] = s;
  }
  __syncthreads();
  red[threadIdx.x] = ab;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int q = 0; q < lanes; ++q) s += red[q * C];
    part[(long long)blockIdx.x * (C + 1) + C] = s;
  }
}

// pre = a + b[c] (kept), v = gelu(pre) or pre (last layer)
__global__ void __launch_bounds__(256) bias_act(float* a, const float* bias, float* v, long long total, int C, int act) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const float pre = a[i] + bias[i % C];
    a[i] = pre;
    v[i] = act ? gelu(pre) : pre;
  }
}

// dpre = dv * gelu'(pre) in place (act) ; partial bias sums per block
__global__ void __launch_bounds__(256) act_bwd(const float* pre, float* dv, float* part, long long total, int C, int act) {
  __shared__ float red[256];
  const int c = threadIdx.x % C, lanes = 256 / C, li = threadIdx.x / C;
  const long long npts = total / C;
  float s = 0.f;
  for (long long n = (long long)blockIdx.x * lanes + li; n < npts; n += (long long)gridDim.x * lanes) {
    const long long i = n * C + c;
    float d = dv[i];
    if (act) d *= gelu_d(pre[i]);
    dv[i] = d;
    s += d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < C) {
    float t = 0.f;
    for (int q = 0; q < lanes; ++q) t += red[q * C + threadIdx.x];
    part[(long long)blockIdx.x * C + threadIdx.x] = t;
  }
}

// Z[b][m][ri][c] from G[b][r(64)][(r1 = ri_in*M + ky)*C + c], m = kpi*M + ky
// (Zr = G[cos kp, Tr] + G[sin kp, Ti], Zi = G[cos kp, Ti] - G[sin kp, Tr]); inverse=1: the transpose.
__global__ void __launch_bounds__(256) assemble(float* G, float* Z, int B, int Mm, int C, int inverse) {
  const int KP = 2 * Mm;                       // kp indices
  const long long total = (long long)B * KP * Mm * C;
  const long long gq = (long long)2 * Mm * C;  // row length of G
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const int c = (int)(i % C);
    const int ky = (int)((i / C) % Mm);
    const int kpi = (int)((i / ((long long)C * Mm)) % KP);
    const int b = (int)(i / ((long long)C * Mm * KP));
    float* Gb = G + (long long)b * (2 * KP) * gq;
    const long long cos_tr = (long long)kpi * gq + (long long)ky * C + c;
    const long long cos_ti = (long long)kpi * gq + (long long)(Mm + ky) * C + c;
    const long long sin_tr = (long long)(KP + kpi) * gq + (long long)ky * C + c;
    const long long sin_ti = (long long)(KP + kpi) * gq + (long long)(Mm + ky) * C + c;
    const int m = kpi * Mm + ky;
    float* zr = Z + ((long long)b * KP * Mm + m) * 2 * C + c;
    float* zi = zr + C;
    if (!inverse) {
      *zr = Gb[cos_tr] + Gb[sin_ti];
      *zi = Gb[cos_ti] - Gb[sin_tr];
    } else {
      Gb[cos_tr] = *zr;
      Gb[sin_ti] = *zr;
      Gb[cos_ti] = *zi;
      Gb[sin_tr] = -*zi;
    }
  }
}

// Yst[b][ri][kpi][ky][o] <-> Y[b][m = kpi*M + ky][ri*O + o]
__global__ void __launch_bounds__(256) disassemble(float* Y, float* Yst, int B, int Mm, int O, int inverse) {
  const int KP = 2 * Mm;
  const long long total = (long long)B * 2 * KP * Mm * O;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const int o = (int)(i % O);
    const int ky = (int)((i / O) % Mm);
    const int kpi = (int)((i / ((long long)O * Mm)) % KP);
    const int ri = (int)((i / ((long long)O * Mm * KP)) % 2);
    const int b = (int)(i / ((long long)O * Mm * KP * 2));
    const long long yi = ((long long)b * KP * Mm + (long long)kpi * Mm + ky) * 2 * O + (long long)ri * O + o;
    if (!inverse) Yst[i] = Y[yi];
    else Y[yi] = Yst[i];
  }
}

// expanded real weights per mode: Wm[m][2C][2O] = [[Rr, Ri], [-Ri, Rr]] as bf16
__global__ void __launch_bounds__(256) expand_modes(const float* R, bf16_t* Wm, int nm, int C, int O) {
  const long long total = (long long)nm * C * O;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const int o = (int)(i % O), c = (int)((i / O) % C), m = (int)(i / ((long long)O * C));
    const float rr = R[i * 2], ri = R[i * 2 + 1];
    bf16_t* w = Wm + (long long)m * 4 * C * O;
    w[(long long)c * 2 * O + o] = to_bf16(rr);
    w[(long long)c * 2 * O + O + o] = to_bf16(ri);
    w[(long long)(C + c) * 2 * O + o] = to_bf16(-ri);
    w[(long long)(C + c) * 2 * O + O + o] = to_bf16(rr);
  }
}

// dR[m][c][o][re, im] from dWm[m][2C][2O]: dRr = dW11 + dW22, dRi = dW12 - dW21
__global__ void __launch_bounds__(256) fold_modes(const float* dWm, float* dR, int nm, int C, int O) {
  const long long total = (long long)nm * C * O;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const int o = (int)(i % O), c = (int)((i / O) % C), m = (int)(i / ((long long)O * C));
    const float* w = dWm + (long long)m * 4 * C * O;
    dR[i * 2] = w[(long long)c * 2 * O + o] + w[(long long)(C + c) * 2 * O + O + o];
    dR[i * 2 + 1] = w[(long long)c * 2 * O + O + o] - w[(long long)(C + c) * 2 * O + o];
  }
}

// Moments of predicted Wigner functions for the physical referee, in FP64:
// per sample b, mom[b][0 .. N) = sum over p of W (the x marginal),
// mom[b][N .. 2N) = sum over x (the p marginal), mom[b][2N .. 3N) = sum over x
// of W^2 per row. Rows: block (p, b); columns: block (x chunk, b). Fixed order.
__global__ void __launch_bounds__(256) moments_rows(const float* w, int N, double* mom) {
  __shared__ double rs[256], rq[256];
  const int p = blockIdx.x, b = blockIdx.y;
  const float* row = w + ((size_t)b * N + p) * N;
  double sv = 0.0, sq = 0.0;
  for (int x = threadIdx.x; x < N; x += 256) {
    const double v = row[x];
    sv += v;
    sq += v * v;
  }
  rs[threadIdx.x] = sv;
  rq[threadIdx.x] = sq;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      rs[threadIdx.x] += rs[threadIdx.x + o];
      rq[threadIdx.x] += rq[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    mom[(size_t)b * 3 * N + N + p] = rs[0];
    mom[(size_t)b * 3 * N + 2 * N + p] = rq[0];
  }
}
__global__ void __launch_bounds__(256) moments_cols(const float* w, int N, double* mom) {
  const int x = blockIdx.x * 256 + threadIdx.x, b = blockIdx.y;
  if (x >= N) return;
  const float* col = w + (size_t)b * N * N + x;
  double sv = 0.0;
  for (int p = 0; p < N; ++p) sv += col[(size_t)p * N];
  mom[(size_t)b * 3 * N + x] = sv;
}

__global__ void __launch_bounds__(256) to_bf16_kernel(const float* s, bf16_t* d, long long n) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) d[i] = to_bf16(s[i]);
}

// Per sample b (one block of 1024 threads each): x[b][n] -= mean_n (x[b][n] - y[b][n])
// (y = NULL: 0). Forward of the zero-mass correction (x = out, y = W) and its
// backward (x = dout).
__global__ void __launch_bounds__(1024) center_kernel(float* x, const float* y, long long nn) {
  __shared__ double r[1024];
  const int b = blockIdx.x;
  float* xb = x + (long long)b * nn;
  const float* yb = y ? y + (long long)b * nn : nullptr;
  double s = 0;
#pragma unroll 8
  for (long long i = threadIdx.x; i < nn; i += 1024) s += (double)xb[i] - (yb ? (double)yb[i] : 0.0);
  r[threadIdx.x] = s;
  __syncthreads();
  for (int k = 512; k > 0; k >>= 1) {
    if (threadIdx.x < k) r[threadIdx.x] += r[threadIdx.x + k];
    __syncthreads();
  }
  const float m = (float)(r[0] / (double)nn);
#pragma unroll 8
  for (long long i = threadIdx.x; i < nn; i += 1024) xb[i] -= m;
}

// loss L = (1/B) sum_b sum_n (out - t)^2 / sum_n t^2 ; dout = dL/dout. One block per sample.
__global__ void __launch_bounds__(256) loss_kernel(const float* out, const float* tgt, float* dout, double* per_sample,
                                                   long long nn, float inv_b) {
  __shared__ double r1[256], r2[256];
  const int b = blockIdx.x;
  const float* o = out + (long long)b * nn;
  const float* t = tgt + (long long)b * nn;
  double se = 0, st = 0;
  for (long long i = threadIdx.x; i < nn; i += 256) {
    const double d = (double)o[i] - (double)t[i];
    se += d * d;
    st += (double)t[i] * t[i];
  }
  r1[threadIdx.x] = se;
  r2[threadIdx.x] = st;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) { r1[threadIdx.x] += r1[threadIdx.x + s]; r2[threadIdx.x] += r2[threadIdx.x + s]; }
    __syncthreads();
  }
  const double SE = r1[0], ST = r2[0];
  if (threadIdx.x == 0) per_sample[b] = SE / ST;
  if (dout) {
    const float sc = (float)(2.0 / ST) * inv_b;
    float* d = dout + (long long)b * nn;
    for (long long i = threadIdx.x; i < nn; i += 256) d[i] = sc * (o[i] - t[i]);
  }
}

// Adam (decoupled weight decay off) on the flat parameter vector
__global__ void __launch_bounds__(256) adam_kernel(float* p, const float* g, float* m, float* v, long long n, float lr,
                                                   float b1, float b2, float eps, float c1, float c2) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  }
}

__global__ void __launch_bounds__(256) gather_pairs(const float* din, const float* dtg, const float* dft, const int* idx,
                                                    float* in, float* tg, float* ft, int B, long long nn) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < (long long)B * nn; i += (long long)gridDim.x * 256) {
    const int b = (int)(i / nn);
    const long long j = i % nn, k = idx[b];
    in[i] = din[k * nn + j];
    tg[i] = dtg[k * nn + j];
    if (j < 5) ft[b * 5 + j] = dft[k * 5 + j];
  }
}

}  // namespace fno


using fno::GemmArgs;

struct fw_fno {
  fw_fno_desc d{};
  int N = 256, C = 32, Mm = 16, L = 4, Bmax = 8;
  int nm = 0;              // retained modes 2 Mm * Mm
  hipStream_t stream = nullptr;
  // parameters: flat FP32 vector and offsets
  long long nparam = 0;
  long long o_Pw = 0, o_pb = 0, o_Q = 0, o_qb = 0;
  float out_scale = 1.f;   // scale of the projection's output (small corrections with O(1) weights)
  int zero_mass = 0;       // subtract each sample's mean correction (out keeps the mass of W)
  std::vector<long long> o_R, o_Wp, o_b;
  float* P = nullptr;      // master weights
  float* G = nullptr;      // gradients
  float* Ma = nullptr;     // Adam moments
  float* Va = nullptr;
  long long t_adam = 0;
  bf16_t* Wm = nullptr;    // expanded spectral weights, L * nm * 4 C C
  bf16_t* Wpb = nullptr;   // pointwise weights bf16, L * C * C
  // DFT matrices (FP32)
  float *Fx = nullptr, *Fp = nullptr, *Gp = nullptr, *Gxr = nullptr, *Gxi = nullptr;
  // activations for batch Bmax
  std::vector<float*> v;   // L+1 tensors [B][N][N][C]
  std::vector<float*> pre; // L tensors
  std::vector<float*> Zs;  // L tensors [B][nm][2C] (kept for the weight gradient)
  float *T1 = nullptr, *Gb = nullptr, *Y = nullptr, *Yst = nullptr, *H = nullptr, *dv = nullptr, *dv2 = nullptr;
  float *dWm = nullptr, *part = nullptr, *out = nullptr, *dout = nullptr;
  float *bin = nullptr, *btg = nullptr, *bft = nullptr;
  int* bidx = nullptr;
  double* persample = nullptr;
  double* mom = nullptr;   // moments of the last forward's outputs (max_batch x 3N)
  // dataset
  float *din = nullptr, *dtg = nullptr, *dft = nullptr;
  int npairs = 0;
  int nblk_red = 512;      // blocks of the reduction kernels
};

namespace {

int gemm(fw_fno* h, bool bf, bool b_is_bf16, const void* A, const void* Bm, float* Cm, int M, int N, int K, int nb1,
         int nb2, long long sAm, long long sAk, long long sBk, long long sBn, long long sCm, long long sCn,
         long long sA1, long long sA2, long long sB1, long long sB2, long long sC1, long long sC2, float alpha,
         float beta) {
  GemmArgs g;
  g.A = A; g.B = Bm; g.C = Cm; g.M = M; g.N = N; g.K = K; g.nb2 = nb2;
  g.sAm = sAm; g.sAk = sAk; g.sBk = sBk; g.sBn = sBn; g.sCm = sCm; g.sCn = sCn;
  g.sA1 = sA1; g.sA2 = sA2; g.sB1 = sB1; g.sB2 = sB2; g.sC1 = sC1; g.sC2 = sC2;
  g.alpha = alpha; g.beta = beta;
  dim3 grid((N + fno::GT - 1) / fno::GT, (M + fno::GT - 1) / fno::GT, nb1 * nb2);
  if (bf) {
    if (b_is_bf16) hipLaunchKernelGGL((fno::gemm_kernel<float, bf16_t, true>), grid, dim3(256), 0, h->stream, g);
    else hipLaunchKernelGGL((fno::gemm_kernel<float, float, true>), grid, dim3(256), 0, h->stream, g);
  } else {
    hipLaunchKernelGGL((fno::gemm_kernel<float, float, false>), grid, dim3(256), 0, h->stream, g);
  }
  HIPCHK(hipGetLastError());
  return FW_OK;
}

int grid_for(long long n) { return (int)std::max(1LL, std::min(4096LL, (n + 255) / 256)); }

int refresh_weights(fw_fno* h) {
  const int C = h->C;
  for (int l = 0; l < h->L; ++l) {
    hipLaunchKernelGGL(fno::expand_modes, dim3(grid_for((long long)h->nm * C * C)), dim3(256), 0, h->stream,
                       h->P + h->o_R[l], h->Wm + (long long)l * h->nm * 4 * C * C, h->nm, C, C);
    hipLaunchKernelGGL(fno::to_bf16_kernel, dim3(grid_for((long long)C * C)), dim3(256), 0, h->stream,
                       h->P + h->o_Wp[l], h->Wpb + (long long)l * C * C, (long long)C * C);
  }
  HIPCHK(hipGetLastError());
  return FW_OK;
}

// spectral convolution of layer l: out (+)= S_l(vin) ; Z kept in h->Zs[l]
int spectral_fwd(fw_fno* h, int l, const float* vin, float* outp, int B, float beta) {
  const int N = h->N, C = h->C, Mm = h->Mm, KP = 2 * Mm, nm = h->nm;
  const long long NN = (long long)N * N;
  // S1: T1[b][p][r(2M)][c] = sum_x Fx[r][x] v[b][p][x][c]
  FWCHK(gemm(h, false, false, h->Fx, vin, h->T1, 2 * Mm, C, N, B, N, N, 1, C, 1, C, 1, 0, 0, NN * C, (long long)N * C,
             (long long)N * 2 * Mm * C, (long long)2 * Mm * C, 1.f, 0.f));
  // S2: G[b][r(2KP)][q] = sum_p Fp[r][p] T1[b][p][q], q < 2M C
  const long long Q = (long long)2 * Mm * C;
  FWCHK(gemm(h, false, false, h->Fp, h->T1, h->Gb, 2 * KP, (int)Q, N, B, 1, N, 1, Q, 1, Q, 1, 0, 0, (long long)N * Q, 0,
             (long long)2 * KP * Q, 0, 1.f, 0.f));
  float* Z = h->Zs[l];
  hipLaunchKernelGGL(fno::assemble, dim3(grid_for((long long)B * KP * Mm * C)), dim3(256), 0, h->stream, h->Gb, Z, B,
                     Mm, C, 0);
  // S3: Y[b][m][2O] = Z[b][m][2C] Wm[m]
  const bf16_t* Wm = h->Wm + (long long)l * nm * 4 * C * C;
  FWCHK(gemm(h, true, true, Z, Wm, h->Y, B, 2 * C, 2 * C, 1, nm, (long long)nm * 2 * C, 1, 2 * C, 1,
             (long long)nm * 2 * C, 1, 0, 2 * C, 0, (long long)4 * C * C, 0, 2 * C, 1.f, 0.f));
  hipLaunchKernelGGL(fno::disassemble, dim3(grid_for((long long)B * 2 * nm * C)), dim3(256), 0, h->stream, h->Y,
                     h->Yst, B, Mm, C, 0);
  // S4: H[b][r(2N)][q(M O)] = sum_s Gp[r][s] Yst[b][s][q], s < 2 KP
  const long long QO = (long long)Mm * C;
  FWCHK(gemm(h, false, false, h->Gp, h->Yst, h->H, 2 * N, (int)QO, 2 * KP, B, 1, 2 * KP, 1, QO, 1, QO, 1, 0, 0,
             (long long)2 * KP * QO, 0, (long long)2 * N * QO, 0, 1.f, 0.f));
  // S5: out[b][p][x][o] = sum_ky Gxr[x][ky] Hr[b][p][ky][o] + Gxi[x][ky] Hi[b][p][ky][o]
  FWCHK(gemm(h, false, false, h->Gxr, h->H, outp, N, C, Mm, B, N, Mm, 1, C, 1, C, 1, 0, 0, (long long)2 * N * QO, QO,
             NN * C, (long long)N * C, 1.f, beta));
  FWCHK(gemm(h, false, false, h->Gxi, h->H + (long long)N * QO, outp, N, C, Mm, B, N, Mm, 1, C, 1, C, 1, 0, 0,
             (long long)2 * N * QO, QO, NN * C, (long long)N * C, 1.f, 1.f));
  return FW_OK;
}

// forward over a batch in h->bin / h->bft; prediction in h->out
int forward(fw_fno* h, int B) {
  const int N = h->N, C = h->C;
  const long long npts = (long long)B * N * N, tot = npts * C;
  const float xs = (float)(2.0 * h->d.x_amplitude / N / h->d.x_amplitude);
  const float ps = (float)(2.0 * h->d.p_amplitude / N / h->d.p_amplitude);
  hipLaunchKernelGGL(fno::lift_fwd, dim3(grid_for(tot)), dim3(256), 0, h->stream, h->bin, h->bft, h->P + h->o_Pw,
                     h->P + h->o_pb, h->v[0], B, N, C, xs, ps);
  for (int l = 0; l < h->L; ++l) {
    float* acc = h->pre[l];
    FWCHK(spectral_fwd(h, l, h->v[l], acc, B, 0.f));
    // pointwise: acc += v Wp
    FWCHK(gemm(h, true, true, h->v[l], h->Wpb + (long long)l * C * C, acc, (int)npts, C, C, 1, 1, C, 1, C, 1, C, 1, 0,
               0, 0, 0, 0, 0, 1.f, 1.f));
    hipLaunchKernelGGL(fno::bias_act, dim3(grid_for(tot)), dim3(256), 0, h->stream, acc, h->P + h->o_b[l], h->v[l + 1],
                       tot, C, l < h->L - 1 ? 1 : 0);
  }
  hipLaunchKernelGGL(fno::proj_fwd, dim3(grid_for(npts)), dim3(256), 0, h->stream, h->bin, h->v[h->L], h->P + h->o_Q,
                     h->P + h->o_qb, h->out, npts, C, h->out_scale);
  if (h->zero_mass)
    hipLaunchKernelGGL(fno::center_kernel, dim3(B), dim3(1024), 0, h->stream, h->out, h->bin, (long long)N * N);
  HIPCHK(hipGetLastError());
  return FW_OK;
}

// backward from h->dout; gradients into h->G
int backward(fw_fno* h, int B) {
  const int N = h->N, C = h->C, Mm = h->Mm, KP = 2 * Mm, nm = h->nm;
  const long long NN = (long long)N * N, npts = (long long)B * NN, tot = npts * C;
  const int nr = h->nblk_red;
  if (npts % nr) return fw::fail(FW_ERR_INVALID_ARG, "batch * grid points must be a multiple of 512");
  HIPCHK(hipMemsetAsync(h->G, 0, sizeof(float) * h->nparam, h->stream));
  // projection
  hipLaunchKernelGGL(fno::proj_bwd, dim3(nr), dim3(256), 0, h->stream, h->v[h->L], h->dout, h->P + h->o_Q, h->dv, h->part,
                     npts, C, h->out_scale);
  hipLaunchKernelGGL(fno::reduce_rows, dim3((C + 1 + 255) / 256), dim3(256), 0, h->stream, h->part, h->G + h->o_Q, nr,
                     C + 1, 1.f, 0);   // Q then qb (adjacent)
  float* dcur = h->dv;    // gradient w.r.t. v[l+1]
  float* dnext = h->dv2;  // gradient w.r.t. v[l]
  const long long Q = (long long)2 * Mm * C, QO = (long long)Mm * C;
  for (int l = h->L - 1; l >= 0; --l) {
    // activation and bias: dcur <- dpre
    hipLaunchKernelGGL(fno::act_bwd, dim3(nr), dim3(256), 0, h->stream, h->pre[l], dcur, h->part, tot, C,
                       l < h->L - 1 ? 1 : 0);
    hipLaunchKernelGGL(fno::reduce_rows, dim3(1), dim3(256), 0, h->stream, h->part, h->G + h->o_b[l], nr, C, 1.f, 0);
    // pointwise: dnext = dpre Wp^T ; dWp = v^T dpre (split over nr chunks of points)
    FWCHK(gemm(h, true, true, dcur, h->Wpb + (long long)l * C * C, dnext, (int)npts, C, C, 1, 1, C, 1, 1, C, C, 1, 0, 0,
               0, 0, 0, 0, 1.f, 0.f));
    const long long chunk = npts / nr;   // nr divides npts (checked on entry)
    {
      FWCHK(gemm(h, true, false, h->v[l], dcur, h->part, C, C, (int)chunk, nr, 1, 1, C, C, 1, C, 1, chunk * C, 0,
                 chunk * C, 0, (long long)C * C, 0, 1.f, 0.f));
      hipLaunchKernelGGL(fno::reduce_rows, dim3((C * C + 255) / 256), dim3(256), 0, h->stream, h->part, h->G + h->o_Wp[l],
                         nr, C * C, 1.f, 0);
    }
    // spectral backward, gradient w.r.t. the spectral output is dcur
    // S5^T: dH[b][p][ky][o] = sum_x Gxr[x][ky] d[b][p][x][o] ; dHi likewise with Gxi
    FWCHK(gemm(h, false, false, h->Gxr, dcur, h->H, Mm, C, N, B, N, 1, Mm, C, 1, C, 1, 0, 0, NN * C, (long long)N * C,
               (long long)2 * N * QO, QO, 1.f, 0.f));
    FWCHK(gemm(h, false, false, h->Gxi, dcur, h->H + (long long)N * QO, Mm, C, N, B, N, 1, Mm, C, 1, C, 1, 0, 0, NN * C,
               (long long)N * C, (long long)2 * N * QO, QO, 1.f, 0.f));
    // S4^T: dYst[b][s][q] = sum_r Gp[r][s] dH[b][r][q]
    FWCHK(gemm(h, false, false, h->Gp, h->H, h->Yst, 2 * KP, (int)QO, 2 * N, B, 1, 1, 2 * KP, QO, 1, QO, 1, 0, 0,
               (long long)2 * N * QO, 0, (long long)2 * KP * QO, 0, 1.f, 0.f));
    hipLaunchKernelGGL(fno::disassemble, dim3(grid_for((long long)B * 2 * nm * C)), dim3(256), 0, h->stream, h->Y,
                       h->Yst, B, Mm, C, 1);
    // S3^T: dZ[b][m] = dY[b][m] Wm[m]^T ; dWm[m] = Z[:, m]^T dY[:, m]
    const bf16_t* Wm = h->Wm + (long long)l * nm * 4 * C * C;
    float* dZ = h->T1;   // scratch (B nm 2C <= B N 2M C)
    FWCHK(gemm(h, true, true, h->Y, Wm, dZ, B, 2 * C, 2 * C, 1, nm, (long long)nm * 2 * C, 1, 1, 2 * C,
               (long long)nm * 2 * C, 1, 0, 2 * C, 0, (long long)4 * C * C, 0, 2 * C, 1.f, 0.f));
    FWCHK(gemm(h, true, false, h->Zs[l], h->Y, h->dWm, 2 * C, 2 * C, B, 1, nm, 1, (long long)nm * 2 * C,
               (long long)nm * 2 * C, 1, 2 * C, 1, 0, 2 * C, 0, 2 * C, 0, (long long)4 * C * C, 1.f, 0.f));
    hipLaunchKernelGGL(fno::fold_modes, dim3(grid_for((long long)nm * C * C)), dim3(256), 0, h->stream, h->dWm,
                       h->G + h->o_R[l], nm, C, C);
    // S3a^T: dG from dZ
    hipLaunchKernelGGL(fno::assemble, dim3(grid_for((long long)B * KP * Mm * C)), dim3(256), 0, h->stream, h->Gb, dZ, B,
                       Mm, C, 1);
    // S2^T: dT1[b][p][q] = sum_r Fp[r][p] dG[b][r][q]
    float* dT1 = h->T1 + (long long)B * nm * 2 * C;   // after dZ
    FWCHK(gemm(h, false, false, h->Fp, h->Gb, dT1, N, (int)Q, 2 * KP, B, 1, 1, N, Q, 1, Q, 1, 0, 0,
               (long long)2 * KP * Q, 0, (long long)N * Q, 0, 1.f, 0.f));
    // S1^T: dnext[b][p][x][c] += sum_r Fx[r][x] dT1[b][p][r][c]
    FWCHK(gemm(h, false, false, h->Fx, dT1, dnext, N, C, 2 * Mm, B, N, 1, N, C, 1, C, 1, 0, 0, (long long)N * Q, Q,
               NN * C, (long long)N * C, 1.f, 1.f));
    std::swap(dcur, dnext);
  }
  // lift: dcur is the gradient w.r.t. v0
  const float xs = (float)(2.0 * h->d.x_amplitude / N / h->d.x_amplitude);
  const float ps = (float)(2.0 * h->d.p_amplitude / N / h->d.p_amplitude);
  hipLaunchKernelGGL(fno::lift_bwd_partial, dim3(nr), dim3(256), 0, h->stream, h->bin, h->bft, dcur, h->part, B, N, C,
                     xs, ps);
  hipLaunchKernelGGL(fno::reduce_rows, dim3((9 * C + 255) / 256), dim3(256), 0, h->stream, h->part, h->G + h->o_Pw, nr,
                     9 * C, 1.f, 0);   // Pw (8 C) then pb (C): adjacent in the parameter vector
  HIPCHK(hipGetLastError());
  return FW_OK;
}

int loss(fw_fno* h, int B, bool want_grad, double* out) {
  const long long nn = (long long)h->N * h->N;
  hipLaunchKernelGGL(fno::loss_kernel, dim3(B), dim3(256), 0, h->stream, h->out, h->btg, want_grad ? h->dout : nullptr,
                     h->persample, nn, 1.f / B);
  // zero-mass correction: dL/dc = s (dout - mean dout) per sample; proj_bwd applies s
  if (want_grad && h->zero_mass)
    hipLaunchKernelGGL(fno::center_kernel, dim3(B), dim3(1024), 0, h->stream, h->dout, nullptr, nn);
  HIPCHK(hipGetLastError());
  if (out) {
    std::vector<double> ps(B);
    HIPCHK(hipMemcpyAsync(ps.data(), h->persample, sizeof(double) * B, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    double s = 0;
    for (double v : ps) s += v;
    *out = s / B;
  }
  return FW_OK;
}

void dft_matrices(int N, int Mm, std::vector<float>& Fx, std::vector<float>& Fp, std::vector<float>& Gp,
                  std::vector<float>& Gxr, std::vector<float>& Gxi) {
  const int KP = 2 * Mm;
  auto kp_of = [&](int i) { return i < Mm ? i : N - KP + i; };
  const long double tp = 2.0L * 3.141592653589793238462643383279502884L / N;
  Fx.assign((size_t)2 * Mm * N, 0.f);
  for (int r = 0; r < Mm; ++r)
    for (int x = 0; x < N; ++x) {
      const long double a = tp * ((long long)r * x % N);
      Fx[(size_t)r * N + x] = (float)cosl(a);
      Fx[(size_t)(Mm + r) * N + x] = (float)-sinl(a);
    }
  Fp.assign((size_t)2 * KP * N, 0.f);
  for (int r = 0; r < KP; ++r)
    for (int p = 0; p < N; ++p) {
      const long double a = tp * ((long long)kp_of(r) * p % N);
      Fp[(size_t)r * N + p] = (float)cosl(a);
      Fp[(size_t)(KP + r) * N + p] = (float)sinl(a);
    }
  Gp.assign((size_t)2 * N * 2 * KP, 0.f);
  for (int p = 0; p < N; ++p)
    for (int s = 0; s < KP; ++s) {
      const long double a = tp * ((long long)kp_of(s) * p % N);
      const float c = (float)cosl(a), sn = (float)sinl(a);
      Gp[(size_t)p * 2 * KP + s] = c;             // Hr from Yr
      Gp[(size_t)p * 2 * KP + KP + s] = -sn;      // Hr from Yi
      Gp[(size_t)(N + p) * 2 * KP + s] = sn;      // Hi from Yr
      Gp[(size_t)(N + p) * 2 * KP + KP + s] = c;  // Hi from Yi
    }
  Gxr.assign((size_t)N * Mm, 0.f);
  Gxi.assign((size_t)N * Mm, 0.f);
  const long double inv = 1.0L / ((long double)N * N);
  for (int x = 0; x < N; ++x)
    for (int ky = 0; ky < Mm; ++ky) {
      const long double a = tp * ((long long)ky * x % N);
      const long double cw = (ky == 0 ? 1.0L : 2.0L) * inv;
      Gxr[(size_t)x * Mm + ky] = (float)(cw * cosl(a));
      Gxi[(size_t)x * Mm + ky] = (float)(-cw * sinl(a));
    }
}

template <typename T>
int dalloc(T** p, size_t n) {
  HIPCHK(hipMalloc((void**)p, sizeof(T) * std::max<size_t>(n, 1)));
  return FW_OK;
}

int ensure_batch(fw_fno* h, int B) {
  if (B < 1 || B > h->Bmax) return fw::fail(FW_ERR_INVALID_ARG, "batch must be in [1, max_batch]");
  return FW_OK;
}

}  // namespace

extern "C" int fw_fno_create(const fw_fno_desc* d, fw_fno** out) {
  fw::clear_error();
  if (!d || !out) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  *out = nullptr;
  if (d->grid_dim < 64 || d->grid_dim > 1024 || !fw::is_pow2(d->grid_dim))
    return fw::fail(FW_ERR_INVALID_ARG, "grid_dim must be a power of two in [64, 1024]");
  if (d->width < 8 || d->width > 128 || d->width % 8) return fw::fail(FW_ERR_INVALID_ARG, "width must be 8..128, a multiple of 8");
  if (d->modes < 1 || 2 * d->modes > d->grid_dim / 2) return fw::fail(FW_ERR_INVALID_ARG, "modes must be in [1, N/4]");
  if (d->layers < 1 || d->layers > 16) return fw::fail(FW_ERR_INVALID_ARG, "layers must be in [1, 16]");
  if (d->max_batch < 1 || d->max_batch > 64) return fw::fail(FW_ERR_INVALID_ARG, "max_batch must be in [1, 64]");
  if (256 % d->width != 0 && d->width > 256) return fw::fail(FW_ERR_INVALID_ARG, "width must divide 256");
  if (256 % d->width != 0) return fw::fail(FW_ERR_INVALID_ARG, "width must divide 256");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fw::fail(FW_ERR_NO_DEVICE, "no GPU visible");
  if (d->device < 0 || d->device >= ndev) return fw::fail(FW_ERR_INVALID_ARG, "device ordinal out of range");
  HIPCHK(hipSetDevice(d->device));
  auto h = new fw_fno();
  std::unique_ptr<fw_fno, int (*)(fw_fno*)> guard(h, fw_fno_destroy);
  h->d = *d;
  h->N = d->grid_dim; h->C = d->width; h->Mm = d->modes; h->L = d->layers; h->Bmax = d->max_batch;
  h->nm = 2 * h->Mm * h->Mm;
  const int N = h->N, C = h->C, Mm = h->Mm, L = h->L, B = h->Bmax, nm = h->nm;
  // parameter layout
  long long o = 0;
  h->o_Pw = o; o += 8LL * C;
  h->o_pb = o; o += C;
  for (int l = 0; l < L; ++l) {
    h->o_R.push_back(o); o += (long long)nm * C * C * 2;
    h->o_Wp.push_back(o); o += (long long)C * C;
    h->o_b.push_back(o); o += C;
  }
  h->o_Q = o; o += C;
  h->o_qb = o; o += 1;
  h->nparam = o;
  HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  FWCHK(dalloc(&h->P, o));
  FWCHK(dalloc(&h->G, o));
  FWCHK(dalloc(&h->Ma, o));
  FWCHK(dalloc(&h->Va, o));
  HIPCHK(hipMemsetAsync(h->P, 0, sizeof(float) * o, h->stream));
  HIPCHK(hipMemsetAsync(h->Ma, 0, sizeof(float) * o, h->stream));
  HIPCHK(hipMemsetAsync(h->Va, 0, sizeof(float) * o, h->stream));
  FWCHK(dalloc(&h->Wm, (size_t)L * nm * 4 * C * C));
  FWCHK(dalloc(&h->Wpb, (size_t)L * C * C));
  std::vector<float> Fx, Fp, Gp, Gxr, Gxi;
  dft_matrices(N, Mm, Fx, Fp, Gp, Gxr, Gxi);
  FWCHK(dalloc(&h->Fx, Fx.size()));
  FWCHK(dalloc(&h->Fp, Fp.size()));
  FWCHK(dalloc(&h->Gp, Gp.size()));
  FWCHK(dalloc(&h->Gxr, Gxr.size()));
  FWCHK(dalloc(&h->Gxi, Gxi.size()));
  HIPCHK(hipMemcpyAsync(h->Fx, Fx.data(), sizeof(float) * Fx.size(), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->Fp, Fp.data(), sizeof(float) * Fp.size(), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->Gp, Gp.data(), sizeof(float) * Gp.size(), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->Gxr, Gxr.data(), sizeof(float) * Gxr.size(), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->Gxi, Gxi.data(), sizeof(float) * Gxi.size(), hipMemcpyHostToDevice, h->stream));
  const size_t act = (size_t)B * N * N * C;
  for (int l = 0; l <= L; ++l) {
    float* p;
    FWCHK(dalloc(&p, act));
    h->v.push_back(p);
  }
  for (int l = 0; l < L; ++l) {
    float* p;
    FWCHK(dalloc(&p, act));
    h->pre.push_back(p);
    FWCHK(dalloc(&p, (size_t)B * nm * 2 * C));
    h->Zs.push_back(p);
  }
  // T1 also holds dZ and dT1 in the backward pass: B N 2M C + B nm 2C
  FWCHK(dalloc(&h->T1, (size_t)B * N * 2 * Mm * C + (size_t)B * nm * 2 * C));
  FWCHK(dalloc(&h->Gb, (size_t)B * 4 * Mm * 2 * Mm * C));
  FWCHK(dalloc(&h->Y, (size_t)B * nm * 2 * C));
  FWCHK(dalloc(&h->Yst, (size_t)B * nm * 2 * C));
  FWCHK(dalloc(&h->H, (size_t)B * 2 * N * Mm * C));
  FWCHK(dalloc(&h->dv, act));
  FWCHK(dalloc(&h->dv2, act));
  FWCHK(dalloc(&h->dWm, (size_t)nm * 4 * C * C));
  FWCHK(dalloc(&h->part, (size_t)h->nblk_red * std::max(9 * C, C * C + 1)));
  FWCHK(dalloc(&h->out, (size_t)B * N * N));
  FWCHK(dalloc(&h->dout, (size_t)B * N * N));
  FWCHK(dalloc(&h->bin, (size_t)B * N * N));
  FWCHK(dalloc(&h->btg, (size_t)B * N * N));
  FWCHK(dalloc(&h->bft, (size_t)B * 5));
  FWCHK(dalloc(&h->bidx, (size_t)B));
  FWCHK(dalloc(&h->persample, (size_t)B));
  FWCHK(dalloc(&h->mom, (size_t)B * 3 * N));
  FWCHK(refresh_weights(h));
  HIPCHK(hipStreamSynchronize(h->stream));
  *out = guard.release();
  return FW_OK;
}

extern "C" int fw_fno_destroy(fw_fno* h) {
  if (!h) return FW_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  std::vector<void*> bufs = {h->P, h->G, h->Ma, h->Va, h->Wm, h->Wpb, h->Fx, h->Fp, h->Gp, h->Gxr, h->Gxi, h->T1,
                             h->Gb, h->Y, h->Yst, h->H, h->dv, h->dv2, h->dWm, h->part, h->out, h->dout, h->bin,
                             h->btg, h->bft, h->bidx, h->persample, h->mom, h->din, h->dtg, h->dft};
  for (auto p : h->v) bufs.push_back(p);
  for (auto p : h->pre) bufs.push_back(p);
  for (auto p : h->Zs) bufs.push_back(p);
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return FW_OK;
}

extern "C" int fw_fno_param_count(fw_fno* h, long long* n) {
  if (!h || !n) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  *n = h->nparam;
  return FW_OK;
}

extern "C" int fw_fno_set_params(fw_fno* h, const float* host) {
  fw::clear_error();
  if (!h || !host) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  HIPCHK(hipMemcpyAsync(h->P, host, sizeof(float) * h->nparam, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemsetAsync(h->Ma, 0, sizeof(float) * h->nparam, h->stream));
  HIPCHK(hipMemsetAsync(h->Va, 0, sizeof(float) * h->nparam, h->stream));
  h->t_adam = 0;
  FWCHK(refresh_weights(h));
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

extern "C" int fw_fno_set_output(fw_fno* h, float scale, int zero_mass) {
  fw::clear_error();
  if (!h) return fw::fail(FW_ERR_INVALID_ARG, "NULL handle");
  if (!(scale > 0.f) || !(scale < 1e30f)) return fw::fail(FW_ERR_INVALID_ARG, "the output scale must be positive");
  h->out_scale = scale;
  h->zero_mass = zero_mass ? 1 : 0;
  return FW_OK;
}

extern "C" int fw_fno_get_params(fw_fno* h, float* host) {
  fw::clear_error();
  if (!h || !host) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  HIPCHK(hipMemcpyAsync(host, h->P, sizeof(float) * h->nparam, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

static int upload_batch(fw_fno* h, const float* in, const float* feat, const float* tgt, int B) {
  const size_t nn = (size_t)h->N * h->N;
  HIPCHK(hipMemcpyAsync(h->bin, in, sizeof(float) * B * nn, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->bft, feat, sizeof(float) * B * 5, hipMemcpyHostToDevice, h->stream));
  if (tgt) HIPCHK(hipMemcpyAsync(h->btg, tgt, sizeof(float) * B * nn, hipMemcpyHostToDevice, h->stream));
  return FW_OK;
}

extern "C" int fw_fno_forward(fw_fno* h, const float* host_in, const float* host_feat, int batch, float* host_out) {
  fw::clear_error();
  if (!h || !host_in || !host_feat || !host_out) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  FWCHK(ensure_batch(h, batch));
  FWCHK(upload_batch(h, host_in, host_feat, nullptr, batch));
  FWCHK(forward(h, batch));
  HIPCHK(hipMemcpyAsync(host_out, h->out, sizeof(float) * batch * h->N * h->N, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

extern "C" int fw_fno_moments(fw_fno* h, int batch, double* host_mom) {
  fw::clear_error();
  if (!h || !host_mom) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  FWCHK(ensure_batch(h, batch));
  const int N = h->N;
  hipLaunchKernelGGL(fno::moments_rows, dim3(N, batch), dim3(256), 0, h->stream, h->out, N, h->mom);
  hipLaunchKernelGGL(fno::moments_cols, dim3((N + 255) / 256, batch), dim3(256), 0, h->stream, h->out, N, h->mom);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(host_mom, h->mom, sizeof(double) * 3 * N * batch, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

extern "C" int fw_fno_loss_grad(fw_fno* h, const float* host_in, const float* host_feat, const float* host_target,
                                int batch, double* loss_out, float* host_grad) {
  fw::clear_error();
  if (!h || !host_in || !host_feat || !host_target || !loss_out) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  FWCHK(ensure_batch(h, batch));
  FWCHK(upload_batch(h, host_in, host_feat, host_target, batch));
  FWCHK(forward(h, batch));
  FWCHK(loss(h, batch, true, loss_out));
  if (host_grad) {
    FWCHK(backward(h, batch));
    HIPCHK(hipMemcpyAsync(host_grad, h->G, sizeof(float) * h->nparam, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  return FW_OK;
}

extern "C" int fw_fno_dataset_reserve(fw_fno* h, int pairs) {
  fw::clear_error();
  if (!h || pairs < 1) return fw::fail(FW_ERR_INVALID_ARG, "bad arguments");
  const size_t nn = (size_t)h->N * h->N;
  if (h->din) { (void)hipFree(h->din); (void)hipFree(h->dtg); (void)hipFree(h->dft); h->din = h->dtg = h->dft = nullptr; }
  FWCHK(dalloc(&h->din, nn * pairs));
  FWCHK(dalloc(&h->dtg, nn * pairs));
  FWCHK(dalloc(&h->dft, (size_t)5 * pairs));
  h->npairs = pairs;
  return FW_OK;
}

extern "C" int fw_fno_dataset_put(fw_fno* h, int k, const void* dev_in, const void* dev_target, const float feat[5]) {
  fw::clear_error();
  if (!h || (!dev_in && !dev_target) || !feat) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  if (k < 0 || k >= h->npairs) return fw::fail(FW_ERR_INVALID_ARG, "pair index out of range");
  const size_t nn = (size_t)h->N * h->N;
  if (dev_in) HIPCHK(hipMemcpyAsync(h->din + nn * k, dev_in, sizeof(float) * nn, hipMemcpyDeviceToDevice, h->stream));
  if (dev_target)
    HIPCHK(hipMemcpyAsync(h->dtg + nn * k, dev_target, sizeof(float) * nn, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->dft + 5 * k, feat, sizeof(float) * 5, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

static int gather(fw_fno* h, const int* host_idx, int B) {
  for (int b = 0; b < B; ++b)
    if (host_idx[b] < 0 || host_idx[b] >= h->npairs) return fw::fail(FW_ERR_INVALID_ARG, "pair index out of range");
  HIPCHK(hipMemcpyAsync(h->bidx, host_idx, sizeof(int) * B, hipMemcpyHostToDevice, h->stream));
  const long long nn = (long long)h->N * h->N;
  hipLaunchKernelGGL(fno::gather_pairs, dim3(grid_for(B * nn)), dim3(256), 0, h->stream, h->din, h->dtg, h->dft, h->bidx,
                     h->bin, h->btg, h->bft, B, nn);
  HIPCHK(hipGetLastError());
  return FW_OK;
}

extern "C" int fw_fno_train_step(fw_fno* h, const int* host_idx, int batch, float lr, double* loss_out) {
  fw::clear_error();
  if (!h || !host_idx) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  FWCHK(ensure_batch(h, batch));
  if (!h->din) return fw::fail(FW_ERR_INVALID_ARG, "no dataset");
  FWCHK(gather(h, host_idx, batch));
  FWCHK(forward(h, batch));
  FWCHK(loss(h, batch, true, loss_out));
  FWCHK(backward(h, batch));
  h->t_adam += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, (double)h->t_adam)));
  const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, (double)h->t_adam)));
  hipLaunchKernelGGL(fno::adam_kernel, dim3(grid_for(h->nparam)), dim3(256), 0, h->stream, h->P, h->G, h->Ma, h->Va,
                     h->nparam, lr, b1, b2, 1e-8f, c1, c2);
  HIPCHK(hipGetLastError());
  FWCHK(refresh_weights(h));
  if (!loss_out) HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

extern "C" int fw_fno_eval_pairs(fw_fno* h, const int* host_idx, int n, double* rel_l2_out) {
  fw::clear_error();
  if (!h || !host_idx || !rel_l2_out) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  if (!h->din) return fw::fail(FW_ERR_INVALID_ARG, "no dataset");
  for (int s = 0; s < n; s += h->Bmax) {
    const int B = std::min(h->Bmax, n - s);
    FWCHK(gather(h, host_idx + s, B));
    FWCHK(forward(h, B));
    FWCHK(loss(h, B, false, nullptr));
    std::vector<double> ps(B);
    HIPCHK(hipMemcpyAsync(ps.data(), h->persample, sizeof(double) * B, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int b = 0; b < B; ++b) rel_l2_out[s + b] = std::sqrt(ps[b]);
  }
  return FW_OK;
}

extern "C" int fw_fno_infer_time(fw_fno* h, int batch, int reps, double* us_per_batch) {
  fw::clear_error();
  if (!h || !us_per_batch || reps < 1) return fw::fail(FW_ERR_INVALID_ARG, "bad arguments");
  FWCHK(ensure_batch(h, batch));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  FWCHK(forward(h, batch));
  HIPCHK(hipEventRecord(e0, h->stream));
  for (int r = 0; r < reps; ++r) FWCHK(forward(h, batch));
  HIPCHK(hipEventRecord(e1, h->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *us_per_batch = ms * 1000.0 / reps;
  return FW_OK;
}

// Strided batched GEMM on MFMA, exported for the parity tests of the kernel itself.
extern "C" int fw_fno_gemm_test(int bf16, const float* hA, const float* hB, float* hC, int M, int N, int K,
                                int transA, int transB) {
  fw::clear_error();
  if (!hA || !hB || !hC || M < 1 || N < 1 || K < 1) return fw::fail(FW_ERR_INVALID_ARG, "bad arguments");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fw::fail(FW_ERR_NO_DEVICE, "no GPU visible");
  fw_fno tmp;
  HIPCHK(hipStreamCreateWithFlags(&tmp.stream, hipStreamNonBlocking));
  float *dA, *dB, *dC;
  HIPCHK(hipMalloc(&dA, sizeof(float) * M * K));
  HIPCHK(hipMalloc(&dB, sizeof(float) * K * N));
  HIPCHK(hipMalloc(&dC, sizeof(float) * M * N));
  HIPCHK(hipMemcpy(dA, hA, sizeof(float) * M * K, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dB, hB, sizeof(float) * K * N, hipMemcpyHostToDevice));
  // A is stored M x K row-major (transA = 0) or K x M (transA = 1); B is K x N or N x K
  const long long sAm = transA ? 1 : K, sAk = transA ? M : 1;
  const long long sBk = transB ? 1 : N, sBn = transB ? K : 1;
  int rc = gemm(&tmp, bf16 != 0, false, dA, dB, dC, M, N, K, 1, 1, sAm, sAk, sBk, sBn, N, 1, 0, 0, 0, 0, 0, 0, 1.f, 0.f);
  if (rc == FW_OK) {
    HIPCHK(hipStreamSynchronize(tmp.stream));
    HIPCHK(hipMemcpy(hC, dC, sizeof(float) * M * N, hipMemcpyDeviceToHost));
  }
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  (void)hipStreamDestroy(tmp.stream);
  return rc;
}
