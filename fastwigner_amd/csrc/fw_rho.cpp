// fw_rho.cpp -- density-matrix propagator handles (von Neumann, Bloch) of
// libfastwigner.so. See include/fastwigner.h for the contract.
#include <hip/hip_runtime.h>

#include <cctype>
#include <cmath>
#include <cstring>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "fastwigner.h"
#include "fw_internal.h"

namespace {

const int kChunk = 64;
const int RSER = FW_RSER;
enum { RC_NORM_RE = 0, RC_STOP = 1, RC_FRESH = 2, RC_NORM_IM = 3, RC_OUT = 8, RC_IP = 48 };
enum { RF_FIRST = 1, RF_BLOCH = 4 };

// Host mirror of fwk::RArgs (fw_rho.h)
struct RArgs {
  void* rho;
  void* aux;
  const void* tw;
  const double* times;
  double* scal;
  double* partials;
  double* series;
  const void* uv;
  const void* uk;
  const double* dw;
  int step;
  int flags;
  int nbp;
  int nrows;
  int pubnb;
  double dbeta, vmin, kmin, t_initial, guard;
};

struct Mod {
  hipModule_t mod = nullptr;
  hipFunction_t f1 = nullptr, f2 = nullptr;
};

// Diagonal partial kernel tiling: DC diagonals per block (lanes; a power of
// two dividing N, at most 64), 256 / DC row subgroups, G row groups (the rows
// the one-block finish adds; a divisor of N, at most 16 and N / (256 / DC)).
void diag_geom(int n, int* dc, int* g) {
  int c = 1;
  while (c < 64 && n % (2 * c) == 0) c *= 2;
  *dc = c;
  const int w = 256 / c;
  int cap = n / w;
  if (cap > 16) cap = 16;
  if (cap < 1) cap = 1;
  int gg = cap;
  while (gg > 1 && n % gg) --gg;
  *g = gg;
}

int diag_blocks(int n) {
  int dc, g;
  diag_geom(n, &dc, &g);
  return (n / dc) * g;
}

// E of the one-block momentum finish (a single line of N / E threads)
int diag_e(int n) {
  int e = n / 256;
  if (e < 2) e = 2;
  if (e > 32) e = 32;
  if (e > n) e = n;
  if (fw::is_pow2(n)) return e;
  for (int f = e; f <= 32; ++f)
    if (fw::line_schedule_ok(n, f) && n / f <= 1024) return f;
  return 0;
}

}  // namespace

struct fw_rho {
  fw_rho_desc d{};
  int N = 0;
  bool fp32 = false;
  size_t csz = 16;   // bytes per complex element
  double dX = 0, dP = 0;
  int row_e = 0, row_p = 0, col_e = 0, col_p = 0;
  int row_nb = 0, row_nt = 0, col_nb = 0, col_nt = 0;
  std::string prelude, arch;

  hipModule_t mod = nullptr;
  hipFunction_t f_a = nullptr, f_b = nullptr, f_c = nullptr, f_ba = nullptr, f_bb = nullptr, f_bc = nullptr,
                f_rows = nullptr, f_diag_p = nullptr, f_diag_f = nullptr, f_norm = nullptr, f_moments = nullptr,
                f_shx = nullptr, f_shy = nullptr, f_flip = nullptr, f_wscale = nullptr, f_set = nullptr,
                f_vmin = nullptr, f_dweights = nullptr, f_diag_pub = nullptr;
  double* dw = nullptr;             // momentum weights (FW_DIAG_W), or null
  hipStream_t stream = nullptr;

  void* rho = nullptr;
  void* aux = nullptr;
  void* wig = nullptr;
  void* winit = nullptr;
  void* snap = nullptr;
  void* tw = nullptr;
  void* uv = nullptr;
  void* uk = nullptr;
  double* times = nullptr;
  double* scal = nullptr;
  double* partials = nullptr;
  double* series = nullptr;

  double t_initial = 0, vmin = 0, kmin = 0;
  bool bloch_ready = false;

  // Wigner time (rho_vneumann_cuda_1d.py:594-620): W_init (uploaded once), the Wigner
  // transform's own block sums, the time kernel's block sums, the per-step values
  double* wpart = nullptr;
  double* wtpart = nullptr;
  double* wts = nullptr;
  Mod wtm;
  bool winit_set = false;

  hipGraphExec_t gexec = nullptr;
  int gkey = -1;
  std::map<std::string, Mod> init_mods, obs_mods;
  bool timing = false;
  double ktimes[3] = {0, 0, 0};
};

namespace {

std::string or_default(const char* s, const char* d) { return (s && *s) ? std::string(s) : std::string(d); }

int validate(const fw_rho_desc* d, const fw_rho_exprs* x) {
  if (!d || !x) return fw::fail(FW_ERR_INVALID_ARG, "descriptor or expressions are NULL");
  if (d->x_grid_dim < 8 || d->x_grid_dim > 4096 || d->x_grid_dim % 2)
    return fw::fail(FW_ERR_INVALID_ARG, "X_gridDIM must be even, in [8, 4096]");
  if (!fw::line_default_e(d->x_grid_dim, 8) || !diag_e(d->x_grid_dim))
    return fw::fail(FW_ERR_INVALID_ARG,
                    "X_gridDIM must be 2^a m with m odd and the product of the distinct primes of m at most 15 "
                    "(e.g. 2^a 3^b 5^c, 2^a 7^b, 2^a 11^b, 2^a 13^b)");
  if (!(d->x_amplitude > 0)) return fw::fail(FW_ERR_INVALID_ARG, "X_amplitude must be positive");
  if (!std::isfinite(d->dt)) return fw::fail(FW_ERR_INVALID_ARG, "dt must be finite");
  if (d->precision != FW_FP64 && d->precision != FW_FP32)
    return fw::fail(FW_ERR_INVALID_ARG, "precision must be FW_FP64 or FW_FP32");
  if (!x->V || !x->K) return fw::fail(FW_ERR_INVALID_ARG, "V and K expressions are required");
  return FW_OK;
}

void grid(const fw_rho_desc* d, double* dX, double* dP) {
  // dX = 2 X_amplitude / N; P = fftfreq(N, dX/(2 pi)), dP = 1/(N dX/(2 pi)) (rho_vneumann_cuda_1d.py:94-101)
  *dX = 2.0 * d->x_amplitude / d->x_grid_dim;
  *dP = 1.0 / ((double)d->x_grid_dim * (*dX / (2.0 * 3.141592653589793)));
}

int tiles(const fw_rho_desc* d, int& re, int& rp, int& ce, int& cp) {
  const int n = d->x_grid_dim;
  const size_t esz = d->precision == FW_FP32 ? 8 : 16;
  // defaults: 256-thread row tiles of 8 elements per thread, column tiles of
  // 16 elements per thread and 4 columns (fastest at 1024^2 FP64 on gfx950)
  re = d->row_e > 0 ? d->row_e : fw::line_default_e(n, n >= 8 ? 8 : n);
  ce = d->col_e > 0 ? d->col_e : fw::line_default_e(n, n >= 256 ? 16 : (n >= 8 ? 8 : n));
  const int tpl = n / re;
  rp = d->row_p;
  if (rp <= 0) {
    rp = 1;
    while (2 * rp * tpl <= 256 && n % (2 * rp) == 0) rp *= 2;
  }
  // FP32 stages its twiddle table in LDS; FP64 computes its twiddles (no table)
  const size_t tws = d->precision == FW_FP32 ? ((size_t)n * 17 / 16 + 1) * 16 : 16;
  auto lds = [&](int lines) {
    const size_t l = (size_t)lines * n;
    return (l + l / 16 + 1) * esz + tws + 20 * 32 * 8;
  };
  cp = d->col_p > 0 ? d->col_p : 4;
  while (cp > 1 && n % cp) cp /= 2;
  // the default column tile shrinks to what fits the 160 KiB of LDS of a CU
  while (d->col_p <= 0 && cp > 1 && lds(cp) > 160 * 1024) cp /= 2;
  if (!fw::line_schedule_ok(n, re) || !fw::line_schedule_ok(n, ce))
    return fw::fail(FW_ERR_INVALID_ARG,
                    "elements per thread must divide N, be at most 32 and hold every stage radix "
                    "(a power of two for power-of-two N)");
  if (!fw::is_pow2(rp) || !fw::is_pow2(cp) || n % rp || n % cp)
    return fw::fail(FW_ERR_INVALID_ARG, "lines per block must be a power of two dividing N");
  if (rp * (n / re) > 1024 || cp * (n / ce) > 1024) return fw::fail(FW_ERR_INVALID_ARG, "tile needs more than 1024 threads");
  if (lds(rp) > 160 * 1024 || lds(cp) > 160 * 1024)
    return fw::fail(FW_ERR_INVALID_ARG, "tile needs more than 160 KiB of LDS");
  return FW_OK;
}

// Does the expression reference the identifier t (the time argument)?
bool uses_t_tok(const char* s) {
  if (!s) return false;
  auto ident = [](char c) { return isalnum((unsigned char)c) || c == '_'; };
  const size_t n = strlen(s);
  for (size_t i = 0; i < n; ++i)
    if (s[i] == 't' && (i == 0 || !ident(s[i - 1])) && (i + 1 == n || !ident(s[i + 1]))) return true;
  return false;
}

// Momentum sums by precomputed weights (one pass, no FFT) when K and K' do not
// depend on t; otherwise the group sums go through the one-block inverse FFT.
bool diag_weights(const fw_rho_exprs* x) { return !uses_t_tok(x->K) && !uses_t_tok(x->diff_K); }

std::string make_prelude(const fw_rho_desc* d, const fw_rho_exprs* x, int re, int rp, int ce, int cp) {
  double dX, dP;
  grid(d, &dX, &dP);
  const bool fp32 = d->precision == FW_FP32;
  std::ostringstream s;
  s << "#define FW_REAL " << (fp32 ? "float" : "double") << "\n#define FW_FP32 " << (fp32 ? 1 : 0) << "\n";
  s << "#define FW_N " << d->x_grid_dim << "\n";
  s << "#define FW_ROW_E " << re << "\n#define FW_ROW_P " << rp << "\n#define FW_COL_E " << ce << "\n#define FW_COL_P "
    << cp << "\n";
  s << "#define FW_DIAG_E " << diag_e(d->x_grid_dim) << "\n";
  s << "#define FW_DIAG_W " << (diag_weights(x) ? 1 : 0) << "\n";
  {
    int dc, g;
    diag_geom(d->x_grid_dim, &dc, &g);
    s << "#define FW_DIAG_DC " << dc << "\n#define FW_DIAG_G " << g << "\n";
  }
  s << "#define FW_PHASE_SHIFT (" << fw::fmt_double(d->phase_time_shift) << ")\n";
  s << "#define FW_HAS_MASK_X " << ((x->abs_boundary_x && *x->abs_boundary_x) ? 1 : 0) << "\n";
  s << "#define FW_HAS_MASK_P " << ((x->abs_boundary_p && *x->abs_boundary_p) ? 1 : 0) << "\n";
  s << "#include \"fw_common.h\"\n";
  s << "const int X_gridDIM = " << d->x_grid_dim << ";\n";
  s << "const double X_amplitude = " << fw::fmt_double(d->x_amplitude) << ";\n";
  s << "const double dt = " << fw::fmt_double(d->dt) << ";\n";
  s << "const double dX = " << fw::fmt_double(dX) << ";\n";
  s << "const double dP = " << fw::fmt_double(dP) << ";\n";
  // the caller's sin and cos: the branch-free FP64 ones of fw_math.h (see there)
  s << "#include \"fw_math.h\"\n#define sin fw_user_sin\n#define cos fw_user_cos\n";
  s << "// ---- caller constants and functions ----\n";
  if (x->consts) s << x->consts << "\n";
  s << "// ---- spliced expressions (rho_vneumann_cuda_1d.py:631-684) ----\n";
  s << "__device__ double V(double X, double t) { return (" << x->V << "); }\n";
  s << "__device__ double K(double P, double t) { return (" << x->K << "); }\n";
  s << "__device__ double fw_diff_V(double X, double t) { return (" << or_default(x->diff_V, "0.") << "); }\n";
  s << "__device__ double fw_diff_K(double P, double t) { return (" << or_default(x->diff_K, "0.") << "); }\n";
  s << "__device__ double fw_mask_x(double X, double X_prime, double t) { return ("
    << or_default(x->abs_boundary_x, "1.") << "); }\n";
  s << "__device__ double fw_mask_p(double P, double P_prime, double t) { return ("
    << or_default(x->abs_boundary_p, "1.") << "); }\n";
  s << "#undef sin\n#undef cos\n";
  return s.str();
}

std::string kset(const std::string& pre, int k) {
  return pre + "#define FW_KSET " + std::to_string(k) + "\n#include \"fw_rho.h\"\n";
}

int launch(fw_rho* h, hipFunction_t f, int grid, int block, void** params) {
  HIPCHK(hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, 0, h->stream, params, nullptr));
  return FW_OK;
}

RArgs base(fw_rho* h) {
  RArgs a;
  std::memset(&a, 0, sizeof a);
  a.rho = h->rho;
  a.aux = h->aux;
  a.tw = h->tw;
  a.times = h->times;
  a.scal = h->scal;
  a.partials = h->partials;
  a.series = h->series;
  a.dw = h->dw;
  a.uv = h->uv;
  a.uk = h->uk;
  a.nbp = h->row_nb;
  a.nrows = kChunk;
  a.vmin = h->vmin;
  a.kmin = h->kmin;
  a.t_initial = h->t_initial;
  return a;
}

int set_scalar(fw_rho* h, int slot, double v) {
  double* p = h->scal + slot;
  void* params[] = {&p, &v};
  return launch(h, h->f_set, 1, 1, params);
}

int args1(fw_rho* h, hipFunction_t f, int grid, int block, RArgs& a) {
  void* params[] = {&a};
  return launch(h, f, grid, block, params);
}

// Momentum family of the current rho into series[a.step]: the diagonal pass,
// then the publish kernel (weights form) or the one-block inverse FFT.
int momentum_sums(fw_rho* h, RArgs& a) {
  const int nb = diag_blocks(h->N);
  FWCHK(args1(h, h->f_diag_p, nb, 256, a));
  if (h->dw) {
    int n = nb;
    void* params[] = {&a, &n};
    return launch(h, h->f_diag_pub, 1, 256, params);
  }
  return args1(h, h->f_diag_f, 1, h->N / diag_e(h->N), a);
}

// rows + steps of one chunk; bloch selects the kernel set
constexpr int kWtBlocks = 256;   // blocks of the Wigner-time reduction (<= N rows each)

// After step s: the Wigner function of rho (Larkin three shears and the
// sign-flipped inverse FFT, rho_vneumann_cuda_1d.py:502-592) in the handle's
// Wigner buffer, unnormalised, then the Wigner time of step s into wts[s].
// The transform works in the Wigner buffer (aux holds the momentum shares that
// the next step's first pass publishes) and keeps its block sums apart from the
// step's partial sums, which the next step's first pass reads for the norm.
int enqueue_wigner_time(fw_rho* h, const RArgs& a, int s) {
  RArgs w = a;
  w.partials = h->wpart;
  w.step = s;
  void* rho = h->rho;
  void* wig = h->wig;
  {
    void* p[] = {&w, &rho, &wig};
    FWCHK(launch(h, h->f_shx, h->row_nb, h->row_nt, p));
  }
  {
    void* p[] = {&w, &wig, &wig};
    FWCHK(launch(h, h->f_shy, h->col_nb, h->col_nt, p));
  }
  {
    void* p[] = {&w, &wig, &wig};
    FWCHK(launch(h, h->f_shx, h->row_nb, h->row_nt, p));
  }
  {
    void* p[] = {&w, &wig, &wig};
    FWCHK(launch(h, h->f_flip, h->col_nb, h->col_nt, p));
  }
  const void* w0 = h->winit;
  const double* tm = h->times;
  const double* sc = h->scal;
  double* part = h->wtpart;
  int step = s;
  const int nbw = std::min(h->N, kWtBlocks);
  {
    void* p[] = {&wig, &w0, &tm, &step, &sc, &part};
    HIPCHK(hipModuleLaunchKernel(h->wtm.f1, nbw, 1, 1, 256, 1, 1, 0, h->stream, p, nullptr));
  }
  double* out = h->wts;
  double dxdp = (h->dX / std::sqrt(2.0)) * (h->dP / std::sqrt(2.0));
  int nb = nbw;
  void* p[] = {&part, &nb, &out, &step, &dxdp};
  HIPCHK(hipModuleLaunchKernel(h->wtm.f2, 1, 1, 1, 256, 1, 1, 0, h->stream, p, nullptr));
  return FW_OK;
}

int enqueue_chunk(fw_rho* h, bool bloch, int n, bool momentum, const RArgs& b0, std::vector<hipEvent_t>* ev,
                  bool wtime = false) {
  RArgs a = b0;
  a.nrows = n;
  {
    void* uv = h->uv;
    void* uk = h->uk;
    void* params[] = {&a, &uv, &uk};
    FWCHK(launch(h, h->f_rows, 256, 256, params));
  }
  // momentum weights without the purity guard: pass A of step s + 1 publishes
  // step s's momentum sums (block 0); the chunk's last step publishes its own
  const bool fold = momentum && !bloch && h->dw && a.guard <= 0;
  const int nbd = diag_blocks(h->N);
  for (int s = 0; s < n; ++s) {
    a.step = s;
    a.pubnb = fold && s > 0 ? nbd : 0;
    if (ev) HIPCHK(hipEventRecord((*ev)[4 * s], h->stream));
    FWCHK(args1(h, bloch ? h->f_ba : h->f_a, h->row_nb, h->row_nt, a));
    a.pubnb = 0;
    if (ev) HIPCHK(hipEventRecord((*ev)[4 * s + 1], h->stream));
    FWCHK(args1(h, bloch ? h->f_bb : h->f_b, h->col_nb, h->col_nt, a));
    if (ev) HIPCHK(hipEventRecord((*ev)[4 * s + 2], h->stream));
    FWCHK(args1(h, bloch ? h->f_bc : h->f_c, h->row_nb, h->row_nt, a));
    if (ev) HIPCHK(hipEventRecord((*ev)[4 * s + 3], h->stream));
    if (momentum) {
      if (fold && s + 1 < n) FWCHK(args1(h, h->f_diag_p, nbd, 256, a));
      else FWCHK(momentum_sums(h, a));
    }
    if (wtime) FWCHK(enqueue_wigner_time(h, a, s));
  }
  return FW_OK;
}

// wtime (real-time steps only): the Wigner time of every step into wtime_out[s]
// normalise = false: the closing pass publishes the last step's sums and leaves rho
// unnormalised (the next run takes rho as it is: it starts with RC_FRESH = 1).
int run(fw_rho* h, bool bloch, const RArgs& b0, const double* times, int steps, bool momentum, double* series,
        int* done, double* wtime_out = nullptr, bool normalise = true) {
  FWCHK(set_scalar(h, RC_STOP, 0.0));
  FWCHK(set_scalar(h, RC_FRESH, 1.0));
  const bool wtime = wtime_out != nullptr && !bloch;
  const bool use_graph = !h->timing && !bloch && b0.guard <= 0;
  const int key = (momentum ? 1 : 0) + (wtime ? 2 : 0);
  if (use_graph && steps >= kChunk && (!h->gexec || h->gkey != key)) {
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_chunk(h, false, kChunk, momentum, b0, nullptr, wtime);
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc != FW_OK) return rc;
    if (e != hipSuccess) return fw::fail(FW_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fw::fail(FW_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    h->gkey = key;
  }
  std::vector<double> slots((size_t)RSER * (kChunk + 1));
  std::vector<double> pending(RSER, 0.0);   // momentum family of the last step of the previous chunk
  double acc[3] = {0, 0, 0};
  int nacc = 0;
  int applied = steps;
  for (int off = 0; off < steps; off += kChunk) {
    const int n = std::min(kChunk, steps - off);
    HIPCHK(hipMemcpyAsync(h->times, times + off, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
    if (use_graph && n == kChunk) {
      HIPCHK(hipGraphLaunch(h->gexec, h->stream));
    } else if (h->timing) {
      std::vector<hipEvent_t> ev(4 * n);
      for (auto& e : ev) HIPCHK(hipEventCreate(&e));
      FWCHK(enqueue_chunk(h, bloch, n, momentum, b0, &ev, wtime));
      HIPCHK(hipStreamSynchronize(h->stream));
      for (int s = 0; s < n; ++s, ++nacc)
        for (int k = 0; k < 3; ++k) {
          float ms = 0;
          HIPCHK(hipEventElapsedTime(&ms, ev[4 * s + k], ev[4 * s + k + 1]));
          acc[k] += ms * 1000.0;
        }
      for (auto& e : ev) (void)hipEventDestroy(e);
    } else {
      FWCHK(enqueue_chunk(h, bloch, n, momentum, b0, nullptr, wtime));
    }
    if (wtime) {
      HIPCHK(hipMemcpyAsync(wtime_out + off, h->wts, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
      if (!series && b0.guard <= 0) HIPCHK(hipStreamSynchronize(h->stream));   // wts is reused by the next chunk
    }
    double stop = 0;
    if (series || b0.guard > 0) {
      HIPCHK(hipMemcpyAsync(slots.data(), h->series, sizeof(double) * RSER * n, hipMemcpyDeviceToHost, h->stream));
      HIPCHK(hipMemcpyAsync(&stop, h->scal + RC_STOP, sizeof(double), hipMemcpyDeviceToHost, h->stream));
      HIPCHK(hipStreamSynchronize(h->stream));
      const int ns = stop != 0 ? (int)stop : n;
      if (series) {
        // trace family of step g is published in slot g+1 (chunk-local), the momentum
        // family of step g in slot g
        for (int s = 0; s < ns; ++s) {
          const int g = off + s - 1;
          if (g >= 0) std::memcpy(series + (size_t)g * RSER, &slots[(size_t)s * RSER], sizeof(double) * 11);
        }
        if (momentum)
          for (int s = 0; s < n && off + s < steps; ++s)
            std::memcpy(series + (size_t)(off + s) * RSER + 11, &slots[(size_t)s * RSER + 11], sizeof(double) * 9);
      }
    }
    if (stop != 0) {
      applied = off + (int)stop - 1;
      break;
    }
  }
  if (nacc)
    for (int k = 0; k < 3; ++k) h->ktimes[k] = acc[k] / nacc;
  RArgs a = b0;
  a.step = kChunk;
  int fp = normalise ? 1 : 2;
  void* params[] = {&a, &fp};
  FWCHK(launch(h, h->f_norm, 1024, 256, params));
  if (series && applied > 0)
    HIPCHK(hipMemcpyAsync(series + (size_t)(applied - 1) * RSER, h->series + (size_t)kChunk * RSER,
                          sizeof(double) * 11, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (done) *done = applied;
  return FW_OK;
}

int build_mod(fw_rho* h, const std::string& body, const char* f1, const char* f2, Mod& m) {
  std::vector<char> code;
  FWCHK(fw::compile(kset(h->prelude, 0) + body, h->arch, code));
  HIPCHK(hipModuleLoadData(&m.mod, code.data()));
  HIPCHK(hipModuleGetFunction(&m.f1, m.mod, f1));
  if (f2) HIPCHK(hipModuleGetFunction(&m.f2, m.mod, f2));
  return FW_OK;
}

int moments_now(fw_rho* h, double t) {
  HIPCHK(hipMemcpyAsync(h->times, &t, sizeof(double), hipMemcpyHostToDevice, h->stream));
  RArgs a = base(h);
  a.step = 0;
  return args1(h, h->f_moments, 1, 1024, a);
}

}  // namespace

extern "C" int fw_rho_grid_constants(const fw_rho_desc* desc, double out[2]) {
  if (!desc || !out) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  grid(desc, &out[0], &out[1]);
  return FW_OK;
}

extern "C" int fw_rho_check(const fw_rho_desc* desc, const fw_rho_exprs* exprs) {
  fw::clear_error();
  FWCHK(validate(desc, exprs));
  int re, rp, ce, cp;
  FWCHK(tiles(desc, re, rp, ce, cp));
  std::vector<char> code;
  return fw::compile(kset(make_prelude(desc, exprs, re, rp, ce, cp), 1), "gfx950", code);
}

extern "C" int fw_rho_create(const fw_rho_desc* desc, const fw_rho_exprs* exprs, fw_rho** out) {
  fw::clear_error();
  if (!out) return fw::fail(FW_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  FWCHK(validate(desc, exprs));
  int re, rp, ce, cp;
  FWCHK(tiles(desc, re, rp, ce, cp));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fw::fail(FW_ERR_NO_DEVICE, "no GPU visible");
  if (desc->device < 0 || desc->device >= ndev) return fw::fail(FW_ERR_INVALID_ARG, "device ordinal out of range");
  HIPCHK(hipSetDevice(desc->device));
  fw_rho* h = new fw_rho();
  h->d = *desc;
  h->N = desc->x_grid_dim;
  h->fp32 = desc->precision == FW_FP32;
  h->csz = h->fp32 ? 8 : 16;
  grid(desc, &h->dX, &h->dP);
  h->row_e = re; h->row_p = rp; h->col_e = ce; h->col_p = cp;
  h->row_nt = rp * (h->N / re);
  h->col_nt = cp * (h->N / ce);
  h->row_nb = h->N / rp;
  h->col_nb = h->N / cp;
  h->prelude = make_prelude(desc, exprs, re, rp, ce, cp);
  h->arch = fw::target_arch(desc->device);
  auto bail = [&](int rc) {
    fw_rho_destroy(h);
    return rc;
  };
  std::vector<char> code;
  int rc = fw::compile(kset(h->prelude, 1), h->arch, code);
  if (rc != FW_OK) return bail(rc);
  if (hipModuleLoadData(&h->mod, code.data()) != hipSuccess) return bail(fw::fail(FW_ERR_HIP, "hipModuleLoadData failed"));
  struct {
    const char* n;
    hipFunction_t* f;
  } fns[] = {{"fw_rho_a", &h->f_a},        {"fw_rho_b", &h->f_b},           {"fw_rho_c", &h->f_c},
             {"fw_rho_ba", &h->f_ba},      {"fw_rho_bb", &h->f_bb},         {"fw_rho_bc", &h->f_bc},
             {"fw_rho_rows", &h->f_rows},  {"fw_rho_diag_partial", &h->f_diag_p},
             {"fw_rho_diag_finish", &h->f_diag_f}, {"fw_rho_apply_norm", &h->f_norm},
             {"fw_rho_moments", &h->f_moments}, {"fw_rho_shear_x", &h->f_shx},
             {"fw_rho_shear_y", &h->f_shy}, {"fw_rho_flip_ifft", &h->f_flip},
             {"fw_rho_wigner_scale", &h->f_wscale}, {"fw_rho_set_scalar", &h->f_set},
             {"fw_rho_vmin", &h->f_vmin}, {"fw_rho_diag_weights", &h->f_dweights},
             {"fw_rho_diag_publish", &h->f_diag_pub}};
  for (auto& f : fns)
    if (hipModuleGetFunction(f.f, h->mod, f.n) != hipSuccess)
      return bail(fw::fail(FW_ERR_HIP, std::string("missing kernel ") + f.n));
  const size_t nn = (size_t)h->N * h->N;
  auto alloc = [&](void** p, size_t bytes) -> int {
    HIPCHK(hipMalloc(p, bytes));
    HIPCHK(hipMemset(*p, 0, bytes));
    return FW_OK;
  };
  // partial rows: pass C (row_nb), the momentum pass (+1), the Wigner flip (col_nb),
  // the Wigner-time kernel (up to 256 blocks of FW_NSER)
  const int maxnb = std::max(std::max(h->row_nb, h->col_nb) + 2, 256);
  if ((rc = alloc(&h->rho, nn * h->csz)) != FW_OK) return bail(rc);
  if ((rc = alloc(&h->aux, nn * h->csz)) != FW_OK) return bail(rc);
  if ((rc = alloc(&h->uv, (size_t)kChunk * h->N * h->csz)) != FW_OK) return bail(rc);
  if ((rc = alloc(&h->uk, (size_t)kChunk * h->N * h->csz)) != FW_OK) return bail(rc);
  if ((rc = alloc((void**)&h->times, sizeof(double) * kChunk)) != FW_OK) return bail(rc);
  if ((rc = alloc((void**)&h->scal, sizeof(double) * 64)) != FW_OK) return bail(rc);
  if ((rc = alloc((void**)&h->partials, sizeof(double) * RSER * maxnb)) != FW_OK) return bail(rc);
  if ((rc = alloc((void**)&h->series, sizeof(double) * RSER * (kChunk + 1))) != FW_OK) return bail(rc);
  if ((rc = fw::upload_twiddles(h->N, h->fp32, &h->tw)) != FW_OK) return bail(rc);
  if (hipStreamCreate(&h->stream) != hipSuccess) return bail(fw::fail(FW_ERR_HIP, "hipStreamCreate failed"));
  if (diag_weights(exprs)) {
    if ((rc = alloc((void**)&h->dw, sizeof(double) * 8 * h->N)) != FW_OK) return bail(rc);
    RArgs a = base(h);
    double* dw = h->dw;
    void* params[] = {&a, &dw};
    if ((rc = launch(h, h->f_dweights, (h->N + 255) / 256, 256, params)) != FW_OK) return bail(rc);
    if (hipStreamSynchronize(h->stream) != hipSuccess) return bail(fw::fail(FW_ERR_HIP, "weights sync failed"));
  }
  *out = h;
  return FW_OK;
}

extern "C" int fw_rho_destroy(fw_rho* h) {
  if (!h) return FW_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  for (auto* m : {&h->init_mods, &h->obs_mods})
    for (auto& kv : *m)
      if (kv.second.mod) (void)hipModuleUnload(kv.second.mod);
  if (h->wtm.mod) (void)hipModuleUnload(h->wtm.mod);
  if (h->mod) (void)hipModuleUnload(h->mod);
  for (void* p : {h->rho, h->aux, h->wig, h->winit, h->snap, h->tw, h->uv, h->uk, (void*)h->times, (void*)h->scal,
                  (void*)h->partials, (void*)h->series, (void*)h->dw, (void*)h->wpart, (void*)h->wtpart,
                  (void*)h->wts})
    if (p) (void)hipFree(p);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return FW_OK;
}

static int upload_complex(fw_rho* h, void* dst, const double* src) {
  const size_t n = (size_t)h->N * h->N;
  if (h->fp32) {
    std::vector<float> tmp(2 * n);
    for (size_t i = 0; i < 2 * n; ++i) tmp[i] = (float)src[i];
    HIPCHK(hipMemcpy(dst, tmp.data(), 2 * n * 4, hipMemcpyHostToDevice));
  } else {
    HIPCHK(hipMemcpy(dst, src, 2 * n * 8, hipMemcpyHostToDevice));
  }
  return FW_OK;
}

static int download_complex(fw_rho* h, const void* src, double* dst) {
  const size_t n = (size_t)h->N * h->N;
  HIPCHK(hipStreamSynchronize(h->stream));
  if (h->fp32) {
    std::vector<float> tmp(2 * n);
    HIPCHK(hipMemcpy(tmp.data(), src, 2 * n * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < 2 * n; ++i) dst[i] = tmp[i];
  } else {
    HIPCHK(hipMemcpy(dst, src, 2 * n * 8, hipMemcpyDeviceToHost));
  }
  return FW_OK;
}

extern "C" int fw_rho_set_state(fw_rho* h, const double* host_rho) {
  if (!h || !host_rho) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  HIPCHK(hipStreamSynchronize(h->stream));
  return upload_complex(h, h->rho, host_rho);
}

extern "C" int fw_rho_set_state_expr(fw_rho* h, const char* expr) {
  if (!h || !expr) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  auto it = h->init_mods.find(expr);
  if (it == h->init_mods.end()) {
    // init_rho_cuda_source, rho_vneumann_cuda_1d.py:687-708
    std::string body =
        "extern \"C\" __global__ void fw_rho_init(fwk::cpx* rho) {\n"
        "  const int i = blockIdx.y;\n"
        "  const int j = threadIdx.x + blockDim.x * blockIdx.x;\n"
        "  if (j >= X_gridDIM) return;\n"
        "  const size_t indexTotal = (size_t)j + (size_t)i * X_gridDIM;\n"
        "  const double X = dX * (j - 0.5 * X_gridDIM);\n"
        "  const double X_prime = dX * (i - 0.5 * X_gridDIM);\n"
        "  const fw_complex v = fw_complex(0.) + (" +
        std::string(expr) +
        ");\n"
        "  rho[indexTotal] = fwk::mk((fwk::real_t)v.re, (fwk::real_t)v.im);\n}\n";
    Mod m;
    FWCHK(build_mod(h, body, "fw_rho_init", nullptr, m));
    it = h->init_mods.emplace(expr, m).first;
  }
  void* r = h->rho;
  void* params[] = {&r};
  const int bx = std::min(h->N, 256);
  HIPCHK(hipModuleLaunchKernel(it->second.f1, (h->N + bx - 1) / bx, h->N, 1, bx, 1, 1, 0, h->stream, params, nullptr));
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

extern "C" int fw_rho_normalize(fw_rho* h, double trace_out[2]) {
  if (!h) return fw::fail(FW_ERR_INVALID_ARG, "NULL handle");
  FWCHK(moments_now(h, 0.0));
  RArgs a = base(h);
  int fp = 0;
  void* params[] = {&a, &fp};
  FWCHK(launch(h, h->f_norm, 1024, 256, params));
  double tr[2];
  HIPCHK(hipMemcpyAsync(tr, h->scal + RC_OUT, sizeof tr, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (trace_out) {
    trace_out[0] = tr[0];
    trace_out[1] = tr[1];
  }
  return FW_OK;
}

extern "C" int fw_rho_get_state(fw_rho* h, double* host_rho) {
  if (!h || !host_rho) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  return download_complex(h, h->rho, host_rho);
}

extern "C" int fw_rho_state_ptr(fw_rho* h, void** dev_ptr, int* elem_size) {
  if (!h || !dev_ptr || !elem_size) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  *dev_ptr = h->rho;
  *elem_size = (int)h->csz;
  return FW_OK;
}

extern "C" int fw_rho_propagate(fw_rho* h, const double* times, int steps, int momentum_sums, double* series) {
  if (!h || (!times && steps > 0)) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  if (steps < 0) return fw::fail(FW_ERR_INVALID_ARG, "steps must be >= 0");
  if (steps == 0) return FW_OK;
  RArgs a = base(h);
  return run(h, false, a, times, steps, momentum_sums != 0, series, nullptr);
}

extern "C" int fw_rho_single_step(fw_rho* h, double t, double* series) {
  if (!h) return fw::fail(FW_ERR_INVALID_ARG, "NULL handle");
  const double times[1] = {t};
  RArgs a = base(h);
  return run(h, false, a, times, 1, false, series, nullptr, nullptr, false);
}

extern "C" int fw_rho_moments(fw_rho* h, double t, double sums[FW_RSER]) {
  if (!h || !sums) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  FWCHK(moments_now(h, t));
  // momentum family of the current rho through the diagonal-sum kernels, into series slot 0
  RArgs a = base(h);
  a.step = 0;
  FWCHK(momentum_sums(h, a));
  HIPCHK(hipMemcpyAsync(sums, h->scal + RC_OUT, sizeof(double) * 11, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipMemcpyAsync(sums + 11, h->series + 11, sizeof(double) * 9, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

extern "C" int fw_rho_get_average(fw_rho* h, const char* const* obs, int nobs, double t, double* out) {
  if (!h || !out || (nobs > 0 && !obs)) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  // apply_observable_cuda_source (rho_vneumann_cuda_1d.py:710-736): multiply by obs(X_j, P_j, t);
  // the p slots go through fft(axis 1)/N ... ifft(axis 1) (:461-470), fused in one row pass
  const size_t bytes = (size_t)h->N * h->N * h->csz;
  HIPCHK(hipMemcpyAsync(h->aux, h->rho, bytes, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->times, &t, sizeof(double), hipMemcpyHostToDevice, h->stream));
  bool is_x = false;
  for (int k = 0; k < nobs; ++k) {
    is_x = !is_x;
    if (!obs[k] || !*obs[k]) continue;
    auto it = h->obs_mods.find(obs[k]);
    if (it == h->obs_mods.end()) {
      std::string e = obs[k];
      std::string body =
          "__device__ fw_complex fw_obs(double X, double P, double t) { return fw_complex(0.) + (" + e +
          "); }\n"
          "extern \"C\" __global__ void fw_obs_x(fwk::cpx* rho, const double* times) {\n"
          "  const int i = blockIdx.y;\n"
          "  const int j = threadIdx.x + blockDim.x * blockIdx.x;\n"
          "  if (j >= X_gridDIM) return;\n"
          "  const fw_complex o = fw_obs(fwk::xcoord(j), fwk::pcoord(j), times[0]);\n"
          "  fwk::cpx& z = rho[(size_t)i * X_gridDIM + j];\n"
          "  z = fwk::cmul(z, fwk::mk((fwk::real_t)o.re, (fwk::real_t)o.im));\n}\n"
          "extern \"C\" __global__ void __launch_bounds__(fwk::RL::NT) fw_obs_p(fwk::cpx* rho, const fwk::tw_t* tw, "
          "const double* times) {\n"
          "  using L = fwk::RL;\n"
          "  constexpr int N = FW_N, E = FW_ROW_E, P = FW_ROW_P, TPL = L::TPL, NT = L::NT;\n"
          "  __shared__ fwk::cpx lds[L::LDS];\n"
          "  __shared__ fwk::tw_t twl[fwk::TwLds<N>::SIZE];\n"
          "  const int t = threadIdx.x % TPL, q = threadIdx.x / TPL;\n"
          "  fwk::cpx* row = rho + (size_t)(blockIdx.x * P + q) * N;\n"
          "  for (int r = threadIdx.x; r < N; r += NT) twl[fwk::twpad(r)] = tw[r];\n"
          "  fwk::cpx v[E];\n"
          "  for (int e = 0; e < E; ++e) v[e] = row[t + TPL * e];\n"
          "  __syncthreads();\n"
          "  L::fwd<false>(v, lds, q, t, twl);\n"
          "  for (int e = 0; e < E; ++e) {\n"
          "    const int k = t + TPL * e;\n"
          "    const fw_complex o = fw_obs(fwk::xcoord(k), fwk::pcoord(k), times[0]) / (double)N;\n"
          "    v[e] = fwk::cmul(v[e], fwk::mk((fwk::real_t)o.re, (fwk::real_t)o.im));\n"
          "  }\n"
          "  L::sync();\n"
          "  L::inv(v, lds, q, t, twl);\n"
          "  for (int e = 0; e < E; ++e) row[t + TPL * e] = v[e];\n}\n";
      Mod m;
      FWCHK(build_mod(h, body, "fw_obs_x", "fw_obs_p", m));
      it = h->obs_mods.emplace(obs[k], m).first;
    }
    void* r = h->aux;
    void* tw = h->tw;
    double* tm = h->times;
    if (is_x) {
      void* params[] = {&r, &tm};
      const int bx = std::min(h->N, 256);
      HIPCHK(hipModuleLaunchKernel(it->second.f1, (h->N + bx - 1) / bx, h->N, 1, bx, 1, 1, 0, h->stream, params,
                                   nullptr));
    } else {
      void* params[] = {&r, &tw, &tm};
      HIPCHK(hipModuleLaunchKernel(it->second.f2, h->row_nb, 1, 1, h->row_nt, 1, 1, 0, h->stream, params, nullptr));
    }
  }
  // trace of the work buffer: moments over aux
  RArgs a = base(h);
  a.rho = h->aux;
  a.step = 0;
  FWCHK(args1(h, h->f_moments, 1, 1024, a));
  double tr[2];
  HIPCHK(hipMemcpyAsync(tr, h->scal + RC_OUT, sizeof tr, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  *out = tr[0] * h->dX;
  return FW_OK;
}

extern "C" int fw_rho_bloch_setup(fw_rho* h, double t_initial, double mins_out[2]) {
  if (!h) return fw::fail(FW_ERR_INVALID_ARG, "NULL handle");
  h->t_initial = t_initial;
  RArgs a = base(h);
  a.t_initial = t_initial;
  FWCHK(args1(h, h->f_vmin, 1, 256, a));
  double vmin = 0;
  HIPCHK(hipMemcpyAsync(&vmin, h->scal + RC_OUT, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->vmin = vmin;
  h->kmin = 0.0;
  h->bloch_ready = true;
  if (mins_out) {
    mins_out[0] = h->vmin;
    mins_out[1] = h->kmin;
  }
  return FW_OK;
}

extern "C" int fw_rho_bloch_propagate(fw_rho* h, double dbeta, int steps, double purity_guard, double* series,
                                      int* steps_done) {
  if (!h) return fw::fail(FW_ERR_INVALID_ARG, "NULL handle");
  if (!h->bloch_ready) return fw::fail(FW_ERR_INVALID_ARG, "call fw_rho_bloch_setup first");
  if (steps_done) *steps_done = 0;
  if (steps <= 0) return steps < 0 ? fw::fail(FW_ERR_INVALID_ARG, "steps must be >= 0") : FW_OK;
  RArgs a = base(h);
  a.flags = RF_BLOCH;
  a.dbeta = dbeta;
  a.guard = purity_guard;
  std::vector<double> times(steps, h->t_initial);
  return run(h, true, a, times.data(), steps, false, series, steps_done);
}

extern "C" int fw_rho_snapshot(fw_rho* h, int which) {
  if (!h) return fw::fail(FW_ERR_INVALID_ARG, "NULL handle");
  const size_t bytes = (size_t)h->N * h->N * h->csz;
  if (which == 0) {
    if (!h->snap) HIPCHK(hipMalloc(&h->snap, bytes));
    HIPCHK(hipMemcpyAsync(h->snap, h->rho, bytes, hipMemcpyDeviceToDevice, h->stream));
  } else if (which == 1) {
    if (!h->snap) return fw::fail(FW_ERR_INVALID_ARG, "no snapshot saved");
    HIPCHK(hipMemcpyAsync(h->rho, h->snap, bytes, hipMemcpyDeviceToDevice, h->stream));
  } else {
    return fw::fail(FW_ERR_INVALID_ARG, "which must be 0 (save) or 1 (restore)");
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

extern "C" int fw_rho_wigner(fw_rho* h, double* host_w) {
  if (!h) return fw::fail(FW_ERR_INVALID_ARG, "NULL handle");
  if (!h->wig) HIPCHK(hipMalloc(&h->wig, (size_t)h->N * h->N * h->csz));
  RArgs a = base(h);
  void* rho = h->rho;
  void* aux = h->aux;
  void* wig = h->wig;
  {
    void* p[] = {&a, &rho, &aux};
    FWCHK(launch(h, h->f_shx, h->row_nb, h->row_nt, p));
  }
  {
    void* p[] = {&a, &aux, &aux};
    FWCHK(launch(h, h->f_shy, h->col_nb, h->col_nt, p));
  }
  {
    void* p[] = {&a, &aux, &aux};
    FWCHK(launch(h, h->f_shx, h->row_nb, h->row_nt, p));
  }
  {
    void* p[] = {&a, &aux, &wig};
    FWCHK(launch(h, h->f_flip, h->col_nb, h->col_nt, p));
  }
  {
    RArgs b = a;
    b.nbp = h->col_nb;
    // wigner_dxdp = (X[1]-X[0])/sqrt2 * (P[1]-P[0])/sqrt2 (rho_vneumann_cuda_1d.py:103-107)
    double dxdp = (h->dX / std::sqrt(2.0)) * (h->dP / std::sqrt(2.0));
    void* p[] = {&b, &wig, &dxdp};
    FWCHK(launch(h, h->f_wscale, 1024, 256, p));
  }
  if (host_w) return download_complex(h, h->wig, host_w);
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

// The Wigner-time kernels (wigner_time_mapper_cuda_code, rho_vneumann_cuda_1d.py:830-858):
// the integral of ind (W - W_init) over the Wigner grid (X/sqrt2, P/sqrt2), with the
// reference's indicator ind = 0.5 (1 + signbit(H + Ip)) (:854). W is the flip pass's
// unnormalised result or the normalised one: with T1 = sum ind W, Sw = sum W (complex)
// and T0 = sum ind W_init, the normalised W = W/(Sw dxdp) gives Re(T1/Sw) - T0 dxdp.
// Ip comes from the scalar slot RC_IP, so it may be set after construction.
static int ensure_wt_module(fw_rho* h) {
  if (h->wtm.mod) return FW_OK;
  const std::string body =
      "extern \"C\" __global__ void __launch_bounds__(256) fw_wt_partial(const fwk::cpx* w, const double* w0, "
      "const double* times, int step, const double* scal, double* part) {\n"
      "  __shared__ double red[fwk::NSER * 4];\n"
      "  const double t = times[step];\n"
      "  const double Ip = scal[" + std::to_string((int)RC_IP) + "];\n"
      "  double s[fwk::NSER];\n"
      "  for (int k = 0; k < fwk::NSER; ++k) s[k] = 0.0;\n"
      "  for (int i = blockIdx.x; i < FW_N; i += gridDim.x) {\n"
      "    const double P = dP * (i - 0.5 * X_gridDIM) / sqrt(2.);\n"
      "    for (int j = threadIdx.x; j < FW_N; j += 256) {\n"
      "      const double X = dX * (j - 0.5 * X_gridDIM) / sqrt(2.);\n"
      "      const double hamiltonian = K(P, t) + V(X, t);\n"
      "      const double indicator_func = 0.5 * (1 + int(signbit(hamiltonian + Ip)));\n"
      "      const size_t idx = (size_t)i * FW_N + j;\n"
      "      const double wr = (double)w[idx].x, wi = (double)w[idx].y;\n"
      "      s[0] += indicator_func * wr;\n"
      "      s[1] += indicator_func * wi;\n"
      "      s[2] += wr;\n"
      "      s[3] += wi;\n"
      "      s[4] += indicator_func * w0[idx];\n"
      "    }\n"
      "  }\n"
      "  fwk::write_partials<256>(s, red, part, blockIdx.x);\n}\n"
      "extern \"C\" __global__ void __launch_bounds__(256) fw_wt_finish(const double* part, int nb, double* out, "
      "int step, double dxdp) {\n"
      "  __shared__ double red[8];\n"
      "  double v[5];\n"
      "  for (int k = 0; k < 5; ++k) v[k] = fwk::sum_partials<256, fwk::NSER>(part, nb, k, red);\n"
      "  if (threadIdx.x == 0) out[step] = (v[0] * v[2] + v[1] * v[3]) / (v[2] * v[2] + v[3] * v[3]) - v[4] * dxdp;\n"
      "}\n";
  return build_mod(h, body, "fw_wt_partial", "fw_wt_finish", h->wtm);
}

static int ensure_wigner_buffers(fw_rho* h) {
  const size_t nn = (size_t)h->N * h->N;
  if (!h->wig) HIPCHK(hipMalloc(&h->wig, nn * h->csz));
  if (!h->wpart) HIPCHK(hipMalloc(&h->wpart, sizeof(double) * RSER * std::max(h->col_nb, h->row_nb)));
  if (!h->wtpart) HIPCHK(hipMalloc(&h->wtpart, sizeof(double) * FW_NSER * kWtBlocks));
  if (!h->wts) HIPCHK(hipMalloc(&h->wts, sizeof(double) * (kChunk + 1)));
  return FW_OK;
}

extern "C" int fw_rho_set_wigner_initial(fw_rho* h, const double* host_w_init) {
  if (!h || !host_w_init) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  const size_t nn = (size_t)h->N * h->N;
  FWCHK(ensure_wigner_buffers(h));
  if (!h->winit) HIPCHK(hipMalloc(&h->winit, nn * 8));
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(h->winit, host_w_init, nn * 8, hipMemcpyHostToDevice));
  FWCHK(ensure_wt_module(h));
  h->winit_set = true;
  return FW_OK;
}

static int set_ip(fw_rho* h, double Ip) {
  HIPCHK(hipMemcpyAsync(h->scal + RC_IP, &Ip, sizeof(double), hipMemcpyHostToDevice, h->stream));
  return FW_OK;
}

extern "C" int fw_rho_wigner_time(fw_rho* h, double t, double Ip, double* out) {
  if (!h || !out) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  if (!h->wig) return fw::fail(FW_ERR_INVALID_ARG, "compute the Wigner function first (fw_rho_wigner)");
  if (!h->winit_set) return fw::fail(FW_ERR_INVALID_ARG, "set W_init first (fw_rho_set_wigner_initial)");
  FWCHK(ensure_wigner_buffers(h));
  FWCHK(set_ip(h, Ip));
  HIPCHK(hipMemcpyAsync(h->times, &t, sizeof(double), hipMemcpyHostToDevice, h->stream));
  // the time kernel reads times[step]: step 0 here
  RArgs a = base(h);
  a.step = 0;
  {
    const void* w0 = h->winit;
    const double* tm = h->times;
    const double* sc = h->scal;
    double* part = h->wtpart;
    void* wig = h->wig;
    int step = 0;
    const int nbw = std::min(h->N, kWtBlocks);
    void* p[] = {&wig, &w0, &tm, &step, &sc, &part};
    HIPCHK(hipModuleLaunchKernel(h->wtm.f1, nbw, 1, 1, 256, 1, 1, 0, h->stream, p, nullptr));
    double* o = h->wts;
    double dxdp = (h->dX / std::sqrt(2.0)) * (h->dP / std::sqrt(2.0));
    int nb = nbw;
    void* q[] = {&part, &nb, &o, &step, &dxdp};
    HIPCHK(hipModuleLaunchKernel(h->wtm.f2, 1, 1, 1, 256, 1, 1, 0, h->stream, q, nullptr));
  }
  HIPCHK(hipMemcpyAsync(out, h->wts, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

extern "C" int fw_rho_propagate_wigner_time(fw_rho* h, const double* times, int steps, int momentum_sums, double Ip,
                                            double* series, double* wigner_time) {
  if (!h || !wigner_time || (!times && steps > 0)) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  if (steps < 0) return fw::fail(FW_ERR_INVALID_ARG, "steps must be >= 0");
  if (!h->winit_set) return fw::fail(FW_ERR_INVALID_ARG, "set W_init first (fw_rho_set_wigner_initial)");
  if (steps == 0) return FW_OK;
  FWCHK(ensure_wigner_buffers(h));
  FWCHK(set_ip(h, Ip));
  RArgs a = base(h);
  FWCHK(run(h, false, a, times, steps, momentum_sums != 0, series, nullptr, wigner_time));
  // the normalised Wigner function of the last step, as get_wignerfunction leaves it
  RArgs b = base(h);
  b.partials = h->wpart;
  b.nbp = h->col_nb;
  double dxdp = (h->dX / std::sqrt(2.0)) * (h->dP / std::sqrt(2.0));
  void* wig = h->wig;
  void* p[] = {&b, &wig, &dxdp};
  FWCHK(launch(h, h->f_wscale, 1024, 256, p));
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

extern "C" int fw_rho_set_timing(fw_rho* h, int enable) {
  if (!h) return fw::fail(FW_ERR_INVALID_ARG, "NULL handle");
  h->timing = enable != 0;
  return FW_OK;
}

extern "C" int fw_rho_kernel_times(fw_rho* h, double us_out[3]) {
  if (!h || !us_out) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  for (int k = 0; k < 3; ++k) us_out[k] = h->ktimes[k];
  return FW_OK;
}
