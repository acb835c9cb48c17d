// fw_wigner.cpp -- Wigner-function propagator handles (Moyal, Bloch, random
// collision) of libfastwigner.so. See include/fastwigner.h for the contract.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cctype>
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "fastwigner.h"
#include "fw_internal.h"

using fw::KArgs;

// ---------------------------------------------------------------------------
// Scratch of fw_wigner_propagate_noalloc: the run's device scalars, step
// times, per-step sums and per-block partial sums in one file-scope array
// instead of the handle's buffers. Each GPU holds its own copy until the
// process ends. In host code `scratch` names a host-side stand-in; kernels get
// the device address from hipGetSymbolAddress (g_scratch_dev, per device).
// Layout in doubles (each part a multiple of 32 doubles, 256 bytes): scalars,
// times (4 x 65: times and the field values of split potentials), series
// slots (NSER x 65), the ticket counter, partial sums (NSER x 1024 blocks).
// ---------------------------------------------------------------------------
namespace {
constexpr int kNoAllocMaxBlocks = 1024;   // blocks per pass whose partial sums the scratch holds
constexpr size_t kScScal = 0, kScTimes = 32, kScSeries = kScTimes + 288, kScCounter = kScSeries + 608,
                 kScPartials = kScCounter + 32, kScTotal = kScPartials + (size_t)FW_NSER * kNoAllocMaxBlocks;
static_assert(kScTimes - kScScal >= fw::NSCAL, "scalar slots");
static_assert(kScSeries - kScTimes >= 4 * 65 && kScCounter - kScSeries >= FW_NSER * 65, "times and series slots");
}  // namespace

static __device__ __align__(256) double scratch[kScTotal];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// The device address of `scratch` on `device` (current), looked up once.
int scratch_on(int device, double** out) {
  if (device < 0 || device >= kMaxDevices) return fw::fail(FW_ERR_INVALID_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fw::fail(FW_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[device] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[device];
  return FW_OK;
}

const int kChunk = 64;        // steps per captured graph
const int kStageChunks = 32;  // chunks of per-step sums staged on the host before a wait

struct Tiles {
  int col_e, col_p, row_e, row_p;
};

int default_e(int n) {
  return fw::line_default_e(n, n >= 256 ? 16 : n >= 8 ? 8 : n);
}

// Default tiles: E elements per thread; column pairs per block so a wave reads
// >= 32 contiguous bytes per row while the grid keeps >= 256 blocks where the
// grid allows; rows fill one wave.
Tiles default_tiles(int nx, int np, bool fp32) {
  Tiles t;
  t.col_e = default_e(np);
  t.row_e = default_e(nx);
  const int pairs_col = nx / 2, pairs_row = np / 2;
  const int esz = fp32 ? 8 : 16;
  int p = 1;
  while (p < 4 && pairs_col % (2 * p) == 0 && pairs_col / (2 * p) >= 256 &&
         (size_t)(2 * p) * np * esz * 17 / 16 <= 48 * 1024 && (2 * p) * (np / t.col_e) <= 512)
    p *= 2;
  t.col_p = p;
  const int tpl = nx / t.row_e;
  int pr = 1;
  while (2 * pr * tpl <= 64 && pairs_row % (2 * pr) == 0) pr *= 2;
  t.row_p = pr;
  return t;
}

struct ExprMod {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
};

}  // namespace

struct fw_wigner {
  fw_wigner_desc d{};
  Tiles tl{};
  int NX = 0, NP = 0;
  bool fp32 = false;
  size_t esz = 8;
  double dX = 0, dP = 0, dLambda = 0, dTheta = 0;
  std::string prelude;  // constants + spliced expressions (shared with expression modules)
  std::string arch;
  bool has_mask_xp = false;

  hipModule_t mod = nullptr, mod_bloch = nullptr, mod_dissip = nullptr;
  hipFunction_t f_v_first = nullptr, f_v_last = nullptr, f_v_first_dissip = nullptr, f_v_last_dissip = nullptr,
                f_k = nullptr, f_bv_first = nullptr, f_bv_last = nullptr, f_bk = nullptr, f_moments = nullptr,
                f_apply_norm = nullptr, f_set_scalar = nullptr, f_vk_min = nullptr, f_vv = nullptr,
                f_k_norm = nullptr;
  hipStream_t stream = nullptr;

  void* W = nullptr;
  void* W0s = nullptr;
  void* snap = nullptr;
  void* tw_col = nullptr;
  void* tw_row = nullptr;
  double* times = nullptr;
  double* scal = nullptr;
  double* partials = nullptr;
  unsigned int* counter = nullptr;
  double* series = nullptr;
  int col_nb = 0, row_nb = 0, col_nt = 0, row_nt = 0;

  bool dissip = false;
  double w_coeff = 1.0;
  int bloch_conv = FW_BLOCH_CUDA;
  double t_initial = 0.0, vmin = 0.0, kmin = 0.0;
  bool bloch_ready = false;

  bool vtab_ok = false, ktab_ok = false;   // V (K) factors are time independent
  bool vsplit = false;                      // V = V0(X) + X F(t): field values beside the step times
  void* vtab = nullptr;
  void* sw = nullptr;   // theta-spectrum sum weights (fused column pass), or null
  void* ktab = nullptr;
  int tab_kind = -1;          // what the tables hold: 0 Moyal, 1 Bloch
  double tab_dbeta = 0.0;
  int kgen = -1;              // Moyal K factors generated in the K passes: -1 not decided, 0 no, 1 yes
  double kgen_err = -1.0;     // largest |generated - table| factor over the grid (-1: not checked)
  hipFunction_t f_build_vtab = nullptr, f_build_ktab = nullptr, f_build_sumw = nullptr, f_begin = nullptr,
                f_kgen_check = nullptr;
  double* h_stage = nullptr;   // pinned: staged per-step sums (kStageChunks chunks + the closing slot)
  double* h_stop = nullptr;    // pinned: the purity guard's stop flag

  hipGraphExec_t gexec = nullptr;
  int gkind = -1;
  bool gdissip = false;   // the three-pass graph was captured with the dissipator kernels
  hipGraphExec_t gexec2 = nullptr;   // two-pass Moyal steps
  // two-pass runs' partial chunks (key: steps | first << 8 | last << 9 | unnormalised << 10),
  // captured on a key's second use
  struct ShortGraph {
    int uses = 0;
    hipGraphExec_t exec = nullptr;
  };
  std::map<int, ShortGraph> sgraphs;
  int last_passes = 3;

  std::map<std::string, ExprMod> init_mods, avg_mods;

  bool timing = false;
  double ktimes[3] = {0, 0, 0};
};

namespace {

int validate(const fw_wigner_desc* d, const fw_wigner_exprs* x) {
  if (!d || !x) return fw::fail(FW_ERR_INVALID_ARG, "descriptor or expressions are NULL");
  if (d->x_grid_dim % 2 || d->p_grid_dim % 2)
    return fw::fail(FW_ERR_INVALID_ARG, "X_gridDIM and P_gridDIM must be even");
  if (d->x_grid_dim < 8 || d->p_grid_dim < 8 || d->x_grid_dim > 4096 || d->p_grid_dim > 4096)
    return fw::fail(FW_ERR_INVALID_ARG, "grid dimensions must be in [8, 4096]");
  if (!default_e(d->x_grid_dim) || !default_e(d->p_grid_dim))
    return fw::fail(FW_ERR_INVALID_ARG,
                    "grid dimensions must be 2^a m with m odd and the product of the distinct primes of m "
                    "at most 15 (e.g. 2^a 3^b 5^c, 2^a 7^b, 2^a 11^b, 2^a 13^b)");
  if (!(d->x_amplitude > 0) || !(d->p_amplitude > 0))
    return fw::fail(FW_ERR_INVALID_ARG, "X_amplitude and P_amplitude must be positive");
  if (!std::isfinite(d->dt)) return fw::fail(FW_ERR_INVALID_ARG, "dt must be finite");
  if (d->precision != FW_FP64 && d->precision != FW_FP32)
    return fw::fail(FW_ERR_INVALID_ARG, "precision must be FW_FP64 or FW_FP32");
  if (d->step_passes != 0 && d->step_passes != 2 && d->step_passes != 3)
    return fw::fail(FW_ERR_INVALID_ARG, "step_passes must be 0, 2 or 3");
  if (d->phase_tables < -1 || d->phase_tables > 1)
    return fw::fail(FW_ERR_INVALID_ARG, "phase_tables must be -1, 0 or 1");
  if (d->v_table != 0 && d->v_table != -1) return fw::fail(FW_ERR_INVALID_ARG, "v_table must be 0 or -1");
  if (!x->V || !x->K) return fw::fail(FW_ERR_INVALID_ARG, "V and K expressions are required");
  return FW_OK;
}

void grid_constants(const fw_wigner_desc* d, double* dX, double* dP, double* dL, double* dT) {
  // as the reference: dX = 2 X_amplitude / X_gridDIM; Lambda = fftfreq(N, dX/(2 pi))
  // whose spacing numpy computes as 1.0/(N * (dX/(2 pi))) (wigner_moyal_cuda_1d.py:113-135)
  const double two_pi = 2.0 * 3.141592653589793;
  *dX = 2.0 * d->x_amplitude / d->x_grid_dim;
  *dP = 2.0 * d->p_amplitude / d->p_grid_dim;
  *dL = 1.0 / ((double)d->x_grid_dim * (*dX / two_pi));
  *dT = 1.0 / ((double)d->p_grid_dim * (*dP / two_pi));
}

int resolve_tiles(const fw_wigner_desc* d, Tiles& t) {
  const bool fp32 = d->precision == FW_FP32;
  t = default_tiles(d->x_grid_dim, d->p_grid_dim, fp32);
  if (d->col_e > 0) t.col_e = d->col_e;
  if (d->col_p > 0) t.col_p = d->col_p;
  if (d->row_e > 0) t.row_e = d->row_e;
  if (d->row_p > 0) t.row_p = d->row_p;
  if (!fw::line_schedule_ok(d->p_grid_dim, t.col_e) || !fw::line_schedule_ok(d->x_grid_dim, t.row_e))
    return fw::fail(FW_ERR_INVALID_ARG,
                    "elements per thread must divide the line length, be at most 32 and hold every stage radix "
                    "(a power of two for power-of-two lines)");
  if (!fw::is_pow2(t.col_p) || (d->x_grid_dim / 2) % t.col_p || !fw::is_pow2(t.row_p) ||
      (d->p_grid_dim / 2) % t.row_p)
    return fw::fail(FW_ERR_INVALID_ARG, "lines per block must be a power of two dividing the line count");
  const int col_nt = t.col_p * (d->p_grid_dim / t.col_e);
  const int row_nt = t.row_p * (d->x_grid_dim / t.row_e);
  if (col_nt > 1024 || row_nt > 1024) return fw::fail(FW_ERR_INVALID_ARG, "tile needs more than 1024 threads");
  const size_t esz = fp32 ? 8 : 16;
  const size_t lds_c = (size_t)t.col_p * d->p_grid_dim, lds_r = (size_t)t.row_p * d->x_grid_dim;
  // gfx950: 160 KiB of LDS per workgroup (exchange buffer + twiddle table + reduction scratch)
  const size_t tws = 16;
  // FP64: twiddles computed in the kernel (no LDS table); FP32: the full padded table
  auto tw_entries = [&](int n) -> size_t { return fp32 ? (size_t)n * 17 / 16 + 1 : (size_t)1; };
  const size_t twc = tw_entries(d->p_grid_dim), twr = tw_entries(d->x_grid_dim);
  if ((lds_c + lds_c / 16 + 1) * esz + twc * tws + 1152 > 160 * 1024 ||
      (lds_r + lds_r / 16 + 1) * esz + twr * tws > 160 * 1024)
    return fw::fail(FW_ERR_INVALID_ARG, "tile needs more than 160 KiB of LDS");
  return FW_OK;
}

// Does the expression reference the identifier token `tok` (e.g. "t", "P")?
bool uses_token(const char* s, const char* tok) {
  if (!s) return false;
  const size_t n = strlen(s), m = strlen(tok);
  auto ident = [](char c) { return isalnum((unsigned char)c) || c == '_'; };
  for (size_t i = 0; i + m <= n; ++i)
    if (strncmp(s + i, tok, m) == 0 && (i == 0 || !ident(s[i - 1])) && (i + m == n || !ident(s[i + m])))
      return true;
  return false;
}
// Does the expression reference the time argument (identifier token "t")?
bool uses_t(const char* s) { return uses_token(s, "t"); }

std::string or_default(const char* s, const char* d) { return (s && *s) ? std::string(s) : std::string(d); }
std::string trim_ws(const std::string& t) {
  const size_t a = t.find_first_not_of(" \t\n"), b = t.find_last_not_of(" \t\n");
  return a == std::string::npos ? std::string() : t.substr(a, b - a + 1);
}

// The additive terms at the top level of an expression as (sign, text); false
// when the scan cannot read it (unbalanced parentheses or brackets, an empty
// term) or when the top level holds an operator that binds more loosely than
// + and - (?:, comparisons, shifts, bitwise and logical operators, assignment,
// the comma): in "c ? a : b + X*F(t)" the last + belongs to one branch only,
// so the top-level terms would not add up to the expression.
bool top_terms(const char* expr, std::vector<std::pair<char, std::string>>& terms) {
  const std::string v(expr);
  auto ident = [](char c) { return isalnum((unsigned char)c) || c == '_' || c == '.'; };
  int depth = 0;
  char sign = '+';
  std::string cur;
  terms.clear();
  for (size_t i = 0; i < v.size(); ++i) {
    const char ch = v[i];
    if (ch == '(' || ch == '[') ++depth;
    if (ch == ')' || ch == ']') --depth;
    if (depth < 0) return false;
    if (depth == 0 && std::strchr("?:<>=!&|^,;", ch)) return false;
    if (depth == 0 && (ch == '+' || ch == '-')) {
      // binary only after an operand; not the sign of a numeric exponent (1e-5)
      size_t j = i;
      while (j > 0 && (v[j - 1] == ' ' || v[j - 1] == '\t')) --j;
      const bool after_operand = j > 0 && (ident(v[j - 1]) || v[j - 1] == ')');
      const bool exponent = j == i && i >= 2 && (v[i - 1] == 'e' || v[i - 1] == 'E') &&
                            (isdigit((unsigned char)v[i - 2]) || v[i - 2] == '.');
      if (after_operand && !exponent) {
        terms.push_back({sign, trim_ws(cur)});
        sign = ch;
        cur.clear();
        continue;
      }
    }
    cur += ch;
  }
  if (depth != 0) return false;
  terms.push_back({sign, trim_ws(cur)});
  for (auto& tm : terms)
    if (tm.second.empty()) return false;
  return true;
}

// Splits V at the top level of its expression into time-free terms and terms
// X * F(t) (a product with the single factor X, F free of X), as in the
// strong-field potentials V0(X) + X * E(t): then V(x - th/2, t) - V(x + th/2, t)
// = [V0(x - th/2) - V0(x + th/2)] - th F(t), and the V factor is a cached table
// of V0 times exp(+i dt/2 th F(tau)) per theta. Returns false (no split) for
// any other form; a term the scan cannot read is never split.
bool split_field(const char* vexpr, std::string& vs, std::string& vf) {
  if (!vexpr || !uses_t(vexpr)) return false;
  std::vector<std::pair<char, std::string>> terms;
  if (!top_terms(vexpr, terms)) return false;
  std::string st, fl;
  for (auto& tm : terms) {
    const std::string& t = tm.second;
    if (!uses_t(t.c_str())) {
      st += std::string(" ") + tm.first + " (" + t + ")";
      continue;
    }
    // a product of factors at depth 0, exactly one of them the token X
    std::vector<std::string> fac;
    std::string f;
    int dd = 0;
    for (char ch : t) {
      if (ch == '(') ++dd;
      if (ch == ')') --dd;
      if (dd == 0 && ch == '*') {
        fac.push_back(trim_ws(f));
        f.clear();
        continue;
      }
      if (dd == 0 && std::strchr("/%+-<>=!&|?:,^", ch)) return false;
      f += ch;
    }
    fac.push_back(trim_ws(f));
    int nx = 0;
    std::string rest;
    for (auto& g : fac) {
      if (g == "X" || g == "(X)") {
        ++nx;
        continue;
      }
      if (g.empty() || uses_token(g.c_str(), "X")) return false;
      rest += (rest.empty() ? "(" : " * (") + g + ")";
    }
    if (nx != 1 || rest.empty()) return false;
    fl += std::string(" ") + tm.first + " " + rest;
  }
  vs = st.empty() ? std::string("0.") : st;
  vf = fl;
  return !vf.empty();
}

// Splits V' = dV0(X) + G(t) (every time-dependent top-level term free of X,
// as E(t) in the strong-field V' = X / (X^2 + 1)^1.5 + E(t)), so the sums of a
// split potential read G per step instead of evaluating it per column.
bool split_offset(const char* expr, std::string& ds, std::string& dg) {
  if (!expr || !uses_t(expr)) return false;
  std::vector<std::pair<char, std::string>> terms;
  if (!top_terms(expr, terms)) return false;
  std::string st, gt;
  for (auto& tm : terms) {
    const std::string& t = tm.second;
    if (!uses_t(t.c_str())) {
      st += std::string(" ") + tm.first + " (" + t + ")";
    } else {
      if (uses_token(t.c_str(), "X")) return false;
      gt += std::string(" ") + tm.first + " (" + t + ")";
    }
  }
  ds = st.empty() ? std::string("0.") : st;
  dg = gt;
  return !dg.empty();
}

// Constants and spliced expressions, wigner_moyal_cuda_1d.py:151-166,244-277.
bool tables_v(const fw_wigner_desc* d, const fw_wigner_exprs* x) {
  if (d->v_table < 0) return false;
  std::string vs, vf;
  const bool t_free = !uses_t(x->V) || split_field(x->V, vs, vf);
  return d->phase_tables > 0 || (d->phase_tables == 0 && t_free && !uses_t(x->abs_boundary_x_theta));
}
// The V factor table holds the time-free part and the field phase is applied
// per step: V splits (split_field), the table is used, and the x-theta mask is
// time-free.
bool v_split(const fw_wigner_desc* d, const fw_wigner_exprs* x, std::string& vs, std::string& vf) {
  return d->phase_tables >= 0 && tables_v(d, x) && !uses_t(x->abs_boundary_x_theta) && split_field(x->V, vs, vf);
}
bool tables_k(const fw_wigner_desc* d, const fw_wigner_exprs* x) {
  return d->phase_tables > 0 || (d->phase_tables == 0 && !uses_t(x->K) && !uses_t(x->abs_boundary_lambda_p));
}

// The fused column pass takes the step's sums on the theta spectrum when
// nothing acts on W in x-p between its two half-steps but a mask of x alone
// (no P token: it scales whole columns) and the weights of the p sums are
// fixed (K and K' independent of t).
bool theta_sums(const fw_wigner_exprs* x) {
  return !uses_token(x->abs_boundary_x_p, "P") && !uses_t(x->K) && !uses_t(x->diff_K);
}

std::string make_prelude(const fw_wigner_desc* d, const fw_wigner_exprs* x, const Tiles& t) {
  double dX, dP, dL, dT;
  grid_constants(d, &dX, &dP, &dL, &dT);
  const bool fp32 = d->precision == FW_FP32;
  std::ostringstream s;
  s << "#define FW_REAL " << (fp32 ? "float" : "double") << "\n";
  s << "#define FW_FP32 " << (fp32 ? 1 : 0) << "\n";
  s << "#define FW_NX " << d->x_grid_dim << "\n#define FW_NP " << d->p_grid_dim << "\n";
  s << "#define FW_COL_E " << t.col_e << "\n#define FW_COL_P " << t.col_p << "\n";
  s << "#define FW_ROW_E " << t.row_e << "\n#define FW_ROW_P " << t.row_p << "\n";
  s << "#define FW_PHASE_SHIFT (" << fw::fmt_double(d->phase_time_shift) << ")\n";
  s << "#define FW_VTAB " << (tables_v(d, x) ? 1 : 0) << "\n#define FW_KTAB " << (tables_k(d, x) ? 1 : 0) << "\n";
  s << "#ifndef M_PI\n#define M_PI 3.14159265358979323846\n#endif\n";
  s << "const int X_gridDIM = " << d->x_grid_dim << ";\n";
  s << "const int P_gridDIM = " << d->p_grid_dim << ";\n";
  s << "const double X_amplitude = " << fw::fmt_double(d->x_amplitude) << ";\n";
  s << "const double P_amplitude = " << fw::fmt_double(d->p_amplitude) << ";\n";
  s << "const double dt = " << fw::fmt_double(d->dt) << ";\n";
  s << "const double dX = " << fw::fmt_double(dX) << ";\n";
  s << "const double dP = " << fw::fmt_double(dP) << ";\n";
  s << "const double dLambda = " << fw::fmt_double(dL) << ";\n";
  s << "const double dTheta = " << fw::fmt_double(dT) << ";\n";
  s << "#define FW_MASK_XP_X_ONLY " << (uses_token(x->abs_boundary_x_p, "P") ? 0 : 1) << "\n";
  s << "#define FW_THETA_SUMS " << (theta_sums(x) ? 1 : 0) << "\n";
  // V and its x-theta mask free of t: the factors of the two half-steps of the
  // fused column pass are the same, evaluated once when there is no table
  s << "#define FW_V_STATIC " << ((!uses_t(x->V) && !uses_t(x->abs_boundary_x_theta)) ? 1 : 0) << "\n";
  // the caller's sin and cos: the branch-free FP64 ones of fw_math.h (see there)
  s << "#include \"fw_math.h\"\n#define sin fw_user_sin\n#define cos fw_user_cos\n";
  s << "// ---- caller constants and functions ----\n";
  if (x->consts) s << x->consts << "\n";
  s << "// ---- spliced expressions ----\n";
  s << "__device__ double V(double X, double t) { return (" << x->V << "); }\n";
  {
    std::string vs, vf;
    if (v_split(d, x, vs, vf)) {
      s << "#define FW_V_SPLIT 1\n#define FW_TCHUNK " << (kChunk + 1) << "\n";
      s << "__device__ double fw_V_static(double X, double t) { (void)t; return (" << vs << "); }\n";
      s << "__device__ double fw_V_field(double t) { return (" << vf << "); }\n";
      std::string ds, dg;
      if (split_offset(x->diff_V, ds, dg)) {
        s << "#define FW_DV_SPLIT 1\n";
        s << "__device__ double fw_dV_static(double X, double t) { (void)t; return (" << ds << "); }\n";
        s << "__device__ double fw_dV_field(double t) { return (" << dg << "); }\n";
      }
    }
  }
  s << "__device__ double K(double P, double t) { return (" << x->K << "); }\n";
  s << "__device__ double fw_diff_V(double X, double t) { return (" << or_default(x->diff_V, "0.") << "); }\n";
  s << "__device__ double fw_diff_K(double P, double t) { return (" << or_default(x->diff_K, "0.") << "); }\n";
  s << "__device__ double fw_mask_x_theta(double X, double Theta, double t) {\n"
       "  const double X_minus = X - 0.5 * Theta; const double X_plus = X + 0.5 * Theta;\n"
       "  (void)X_minus; (void)X_plus;\n  return ("
    << or_default(x->abs_boundary_x_theta, "1.") << ");\n}\n";
  s << "__device__ double fw_mask_lambda_p(double Lambda, double P, double t) { return ("
    << or_default(x->abs_boundary_lambda_p, "1.") << "); }\n";
  s << "__device__ double fw_mask_x_p(double X, double P, double t) { return ("
    << or_default(x->abs_boundary_x_p, "1.") << "); }\n";
  s << "#undef sin\n#undef cos\n";
  return s.str();
}

std::string kset_source(const std::string& prelude, int kset) {
  return prelude + "#define FW_KSET " + std::to_string(kset) + "\n#include \"fw_device.h\"\n";
}

int load_module(const std::vector<char>& code, hipModule_t* mod) {
  HIPCHK(hipModuleLoadData(mod, code.data()));
  return FW_OK;
}

int getfn(hipModule_t m, const char* name, hipFunction_t* f) {
  HIPCHK(hipModuleGetFunction(f, m, name));
  return FW_OK;
}

int launch(fw_wigner* h, hipFunction_t f, int grid, int block, KArgs& a) {
  void* params[] = {&a};
  HIPCHK(hipModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, 0, h->stream, params, nullptr));
  return FW_OK;
}

KArgs base_args(fw_wigner* h) {
  KArgs a;
  std::memset(&a, 0, sizeof a);
  a.W = h->W;
  a.W0s = h->W0s;
  a.tw_col = h->tw_col;
  a.tw_row = h->tw_row;
  a.times = h->times;
  a.scal = h->scal;
  a.partials = h->partials;
  a.counter = h->counter;
  a.series = h->series;
  a.vtab = h->vtab;
  a.ktab = h->ktab;
  a.sw = h->sw;
  a.w_coeff = h->w_coeff;
  a.flags = h->has_mask_xp ? fw::F_MASK_XP : 0;
  a.nbp = h->col_nb;
  a.dbeta = 0;
  a.cv = a.ck = 0;
  a.vmin = h->vmin;
  a.kmin = h->kmin;
  a.t_initial = h->t_initial;
  a.guard = 0;
  return a;
}

int set_scalar(fw_wigner* h, int slot, double v) {
  double* p = h->scal + slot;
  void* params[] = {&p, &v};
  HIPCHK(hipModuleLaunchKernel(h->f_set_scalar, 1, 1, 1, 1, 1, 1, 0, h->stream, params, nullptr));
  return FW_OK;
}

int helper_blocks(fw_wigner* h) { return h->NP < 128 ? h->NP : 128; }

// Start of a chunk of n steps with times t[0..n). When the times are the
// progression t[s] = t[s-1] + inc for inc = dt (the propagators' t += dt) or
// inc = 0 (Bloch runs at t_initial), checked here bit for bit, one small kernel
// (fw_begin) regenerates them on the device by the same IEEE additions;
// other times are copied. The same launch resets the run flags (reset) and,
// for real-time steps of a split potential (field), writes the field values
// F(t_s + shift dt), F(t_s) beside the times, so the passes read one number
// per half-step instead of evaluating the field in every thread. Bloch steps
// use the whole V at t_initial and need none.
int begin_chunk(fw_wigner* h, const KArgs& run, const double* t, int n, bool reset, bool field = true) {
  double inc = 0.0;
  bool prog = false;
  for (double c : {h->d.dt, 0.0}) {
    prog = true;
    for (int s = 1; s < n && prog; ++s) prog = (t[s] == t[s - 1] + c);
    if (prog) {
      inc = c;
      break;
    }
  }
  int ngen = n;
  if (!prog) {
    // a pageable source: the copy is staged before hipMemcpyAsync returns
    HIPCHK(hipMemcpyAsync((void*)run.times, t, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
    ngen = 0;
  }
  KArgs a = run;
  double t0 = n > 0 ? t[0] : 0.0;
  int rs = reset ? 1 : 0, nf = (h->vsplit && field) ? n : 0;
  void* params[] = {&a, &t0, &inc, &ngen, &rs, &nf};
  HIPCHK(hipModuleLaunchKernel(h->f_begin, 1, 1, 1, 128, 1, 1, 0, h->stream, params, nullptr));
  return FW_OK;
}

// Collects the per-step sums of a run on the host. The slots of each chunk go
// device -> pinned staging memory asynchronously, in stream order after the
// chunk's kernels; the host waits once at the end of the run (or when the
// staging area is full) and scatters them: slot s of the chunk starting at
// step off holds the sums of step off + s - 1, and the closing normalisation
// publishes those of the last step into slot kChunk.
struct SeriesSink {
  fw_wigner* h;
  const double* dser;   // the run's device series slots
  double* out;
  struct Rec {
    int off, n, k;
  };
  Rec recs[kStageChunks];
  int nrec = 0;
  int last = -1;
  int last_rec = -1;   // the record whose slot kChunk holds the last step's sums (final_with_chunk)
  SeriesSink(fw_wigner* h_, const double* dser_, double* out_) : h(h_), dser(dser_), out(out_) {}
  int chunk(int off, int n) {
    if (!out) return FW_OK;
    if (nrec == kStageChunks) FWCHK(flush());
    const int k = nrec;
    HIPCHK(hipMemcpyAsync(h->h_stage + (size_t)k * (kChunk + 1) * FW_NSER, dser, sizeof(double) * FW_NSER * n,
                          hipMemcpyDeviceToHost, h->stream));
    recs[nrec++] = {off, n, k};
    return FW_OK;
  }
  // the run's last chunk and the closing slot kChunk in one copy, after the
  // closing normalisation has published the last step's sums
  int final_with_chunk(int off, int n, int step) {
    if (!out) return FW_OK;
    if (nrec == kStageChunks) FWCHK(flush());
    const int k = nrec;
    HIPCHK(hipMemcpyAsync(h->h_stage + (size_t)k * (kChunk + 1) * FW_NSER, dser, sizeof(double) * FW_NSER * (kChunk + 1),
                          hipMemcpyDeviceToHost, h->stream));
    recs[nrec++] = {off, n, k};
    last = step;
    last_rec = k;
    return FW_OK;
  }
  int final_slot(int step) {
    if (!out || step < 0) return FW_OK;
    HIPCHK(hipMemcpyAsync(h->h_stage + (size_t)kStageChunks * (kChunk + 1) * FW_NSER, dser + (size_t)kChunk * FW_NSER,
                          sizeof(double) * FW_NSER, hipMemcpyDeviceToHost, h->stream));
    last = step;
    return FW_OK;
  }
  void scatter() {
    for (int i = 0; i < nrec; ++i) {
      const Rec& r = recs[i];
      for (int s = 0; s < r.n; ++s) {
        const int g = r.off + s - 1;
        if (g >= 0)
          std::memcpy(out + (size_t)g * FW_NSER, h->h_stage + ((size_t)r.k * (kChunk + 1) + s) * FW_NSER,
                      sizeof(double) * FW_NSER);
      }
    }
    nrec = 0;
  }
  int flush() {
    if (!out) return FW_OK;
    HIPCHK(hipStreamSynchronize(h->stream));
    scatter();
    return FW_OK;
  }
  // after the stream is synchronised
  void finish() {
    if (!out) return;
    // the closing slot: in the last chunk's record (final_with_chunk) or in the area after the records
    const double* fin = last_rec >= 0 ? h->h_stage + ((size_t)last_rec * (kChunk + 1) + kChunk) * FW_NSER
                                      : h->h_stage + (size_t)kStageChunks * (kChunk + 1) * FW_NSER;
    scatter();
    if (last >= 0) std::memcpy(out + (size_t)last * FW_NSER, fin, sizeof(double) * FW_NSER);
  }
};

// With phase_tables 0 the FP64 K passes generate the Moyal K factors as
// powers (kgen_factors) when those agree with the freshly built table to 1e-11
// on every grid point (K(P + L/2) - K(P - L/2) linear in L, no lambda-p mask;
// for P^2/2 the two differ by the rounding of the phases, below 1e-13 up to
// 2048^2). The pass then reads W alone. Otherwise, with phase_tables 1, and in
// FP32 (whose row pass the two FP64 sincos per row slow down: 2048^2 17.1 ->
// 21.9 us) they read the table.
int decide_kgen(fw_wigner* h, KArgs b) {
  const double tol = 1e-11;
  h->kgen = 0;
  if (h->d.phase_tables != 0 || h->fp32) return FW_OK;
  HIPCHK(hipMemsetAsync(h->scal + fw::SC_OUT, 0, sizeof(double), h->stream));
  FWCHK(launch(h, h->f_kgen_check, h->row_nb, h->row_nt, b));
  double err = -1.0;
  HIPCHK(hipMemcpyAsync(&err, h->scal + fw::SC_OUT, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->kgen_err = err;
  h->kgen = (err >= 0.0 && err <= tol) ? 1 : 0;
  return FW_OK;
}

// Fill the factor tables for a run of `kind` (0 Moyal, 1 Bloch with a.dbeta)
// and set the table flags in a. Tables are allocated once; contents follow kind.
int ensure_tables(fw_wigner* h, int kind, KArgs& a) {
  const bool need_v = h->vtab_ok || kind == 1, need_k = h->ktab_ok || kind == 1;
  if (!need_v && !need_k) return FW_OK;
  const size_t cs = h->fp32 ? 8 : 16;
  if (need_v && !h->vtab) HIPCHK(hipMalloc(&h->vtab, (size_t)(h->NP / 2 + 1) * h->NX * cs));
  if (need_k && !h->ktab) HIPCHK(hipMalloc(&h->ktab, (size_t)h->NP * (h->NX / 2 + 1) * cs));
  a.vtab = h->vtab;
  a.ktab = h->ktab;
  if (h->tab_kind != kind || (kind == 1 && h->tab_dbeta != a.dbeta)) {
    KArgs b = a;
    if (kind == 1) b.flags |= fw::F_BLOCH;
    const double t0 = kind == 1 ? h->t_initial : 0.0;
    HIPCHK(hipMemcpyAsync(h->times, &t0, sizeof(double), hipMemcpyHostToDevice, h->stream));
    if (need_v) {
      void* tab = h->vtab;
      void* params[] = {&b, &tab};
      HIPCHK(hipModuleLaunchKernel(h->f_build_vtab, 1024, 1, 1, 256, 1, 1, 0, h->stream, params, nullptr));
    }
    if (need_k) {
      void* tab = h->ktab;
      void* params[] = {&b, &tab};
      HIPCHK(hipModuleLaunchKernel(h->f_build_ktab, 1024, 1, 1, 256, 1, 1, 0, h->stream, params, nullptr));
    }
    h->tab_kind = kind;
    h->tab_dbeta = a.dbeta;
    if (kind == 0 && need_k && h->kgen < 0) FWCHK(decide_kgen(h, b));
  }
  if (need_v) a.flags |= fw::F_VTAB;
  if (need_k) a.flags |= fw::F_KTAB;
  if (kind == 0 && need_k && h->kgen == 1) a.flags |= fw::F_KGEN;
  return FW_OK;
}

// Step kernels of one step s (chunk-local). kind 0 = Moyal, 1 = Bloch.
int enqueue_step(fw_wigner* h, int kind, int s, const KArgs& base, hipEvent_t* ev) {
  KArgs a = base;
  a.step = s;
  hipFunction_t f1, f2, f3;
  if (kind == 1) {
    f1 = h->f_bv_first; f2 = h->f_bk; f3 = h->f_bv_last;
  } else {
    f1 = h->dissip ? h->f_v_first_dissip : h->f_v_first;
    f2 = h->f_k;
    f3 = h->dissip ? h->f_v_last_dissip : h->f_v_last;
  }
  if (ev) HIPCHK(hipEventRecord(ev[0], h->stream));
  FWCHK(launch(h, f1, h->col_nb, h->col_nt, a));
  if (ev) HIPCHK(hipEventRecord(ev[1], h->stream));
  FWCHK(launch(h, f2, h->row_nb, h->row_nt, a));
  if (ev) HIPCHK(hipEventRecord(ev[2], h->stream));
  FWCHK(launch(h, f3, h->col_nb, h->col_nt, a));
  if (ev) HIPCHK(hipEventRecord(ev[3], h->stream));
  return FW_OK;
}

int ensure_graph(fw_wigner* h, int kind, const KArgs& base) {
  if (h->gexec && h->gkind == kind && h->gdissip == h->dissip) return FW_OK;
  if (h->gexec) {
    (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
  }
  hipGraph_t g;
  HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = FW_OK;
  for (int s = 0; s < kChunk && rc == FW_OK; ++s) rc = enqueue_step(h, kind, s, base, nullptr);
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (rc != FW_OK) return rc;
  if (e != hipSuccess) return fw::fail(FW_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fw::fail(FW_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  h->gkind = kind;
  h->gdissip = h->dissip;
  return FW_OK;
}

// Run `steps` steps of `kind`; times[s] per step; series (optional) gets NSER
// sums per step. The sums of step s are published by the first pass of step
// s+1 into slot (s+1) mod chunk of the device series buffer, and those of the
// last step by the closing normalisation into slot kChunk. *done = steps
// applied before a purity-guard stop.
// normalise = false: the closing pass publishes the last step's sums and leaves W
// unnormalised (the next run takes W as it is: it starts with SC_FRESH = 1).
int run_steps(fw_wigner* h, int kind, const KArgs& base0, const double* times, int steps, double* series,
              int* done, bool normalise = true) {
  KArgs base = base0;
  FWCHK(ensure_tables(h, kind, base));
  const bool use_graph = !h->timing && kind == 0;
  if (use_graph && steps >= kChunk) FWCHK(ensure_graph(h, kind, base));
  std::vector<hipEvent_t> evs;
  SeriesSink sink(h, base.series, series);
  double acc[3] = {0, 0, 0};
  int nacc = 0;
  int applied = steps;
  for (int off = 0; off < steps; off += kChunk) {
    const int n = std::min(kChunk, steps - off);
    FWCHK(begin_chunk(h, base, times + off, n, off == 0, kind == 0));
    if (use_graph && n == kChunk) {
      HIPCHK(hipGraphLaunch(h->gexec, h->stream));
    } else if (h->timing) {
      evs.resize(4 * n);
      for (auto& e : evs) HIPCHK(hipEventCreate(&e));
      for (int s = 0; s < n; ++s) FWCHK(enqueue_step(h, kind, s, base, &evs[4 * s]));
      HIPCHK(hipStreamSynchronize(h->stream));
      for (int s = 0; s < n; ++s) {
        for (int k = 0; k < 3; ++k) {
          float ms = 0;
          HIPCHK(hipEventElapsedTime(&ms, evs[4 * s + k], evs[4 * s + k + 1]));
          acc[k] += ms * 1000.0;
        }
        ++nacc;
      }
      for (auto& e : evs) (void)hipEventDestroy(e);
    } else {
      for (int s = 0; s < n; ++s) FWCHK(enqueue_step(h, kind, s, base, nullptr));
    }
    FWCHK(sink.chunk(off, n));
    if (base.guard > 0) {
      // the purity guard: whether a step of this chunk tripped it (stop = its step + 1)
      HIPCHK(hipMemcpyAsync(h->h_stop, h->scal + fw::SC_STOP, sizeof(double), hipMemcpyDeviceToHost, h->stream));
      HIPCHK(hipStreamSynchronize(h->stream));
      const double stop = *h->h_stop;
      if (stop != 0 && sink.nrec > 0) sink.recs[sink.nrec - 1].n = (int)stop;   // slots written in this chunk
      sink.scatter();
      if (stop != 0) {
        applied = off + (int)stop - 1;
        break;
      }
    }
  }
  if (nacc)
    for (int k = 0; k < 3; ++k) h->ktimes[k] = acc[k] / nacc;
  // the deferred normalisation of the last applied step, which also publishes its sums
  KArgs a = base;
  a.flags |= fw::F_FROM_PARTIALS | (normalise ? 0 : fw::F_NO_SCALE);
  a.step = kChunk;
  FWCHK(launch(h, h->f_apply_norm, 1024, 256, a));
  FWCHK(sink.final_slot(applied - 1));
  HIPCHK(hipStreamSynchronize(h->stream));
  sink.finish();
  if (done) *done = applied;
  return FW_OK;
}

// Two-pass Moyal steps: V/2 (opening, step 0), then per step s a normalising K
// pass and a fused VV pass (closing V/2 of s + opening V/2 of s+1), the last
// step closing with the plain V/2 pass. The K pass of step s applies the norm
// of step s-1 and publishes its sums to chunk-local slot s, as the first pass
// of three-pass steps does, so the series bookkeeping is the same.
int enqueue_fused(fw_wigner* h, int s, bool last, const KArgs& base, hipEvent_t* ev) {
  KArgs a = base;
  a.step = s;
  if (ev) HIPCHK(hipEventRecord(ev[0], h->stream));
  FWCHK(launch(h, h->f_k_norm, h->row_nb, h->row_nt, a));
  if (ev) HIPCHK(hipEventRecord(ev[1], h->stream));
  FWCHK(launch(h, last ? h->f_v_last : h->f_vv, h->col_nb, h->col_nt, a));
  if (ev) HIPCHK(hipEventRecord(ev[2], h->stream));
  return FW_OK;
}

int ensure_graph_fused(fw_wigner* h, const KArgs& base) {
  if (h->gexec2) return FW_OK;
  hipGraph_t g;
  HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = FW_OK;
  for (int s = 0; s < kChunk && rc == FW_OK; ++s) rc = enqueue_fused(h, s, false, base, nullptr);
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (rc != FW_OK) return rc;
  if (e != hipSuccess) return fw::fail(FW_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&h->gexec2, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fw::fail(FW_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  return FW_OK;
}

// The closing normalisation of a two-pass run: the norm and the last step's
// sums from the partials of the closing V/2 (into series slot kChunk).
int enqueue_close(fw_wigner* h, const KArgs& base, bool normalise) {
  KArgs a = base;
  a.flags |= fw::F_FROM_PARTIALS | (normalise ? 0 : fw::F_NO_SCALE);
  a.step = kChunk;
  return launch(h, h->f_apply_norm, 1024, 256, a);
}

// The kernels of a partial chunk of n < kChunk steps, or of a run's last
// chunk: [opening V/2 when first] n x (K pass, VV pass or the closing V/2 at
// the run's end) [closing normalisation when last], captured as one graph on
// the second use of the same shape, so that repeated short runs
// (propagate(50) per frame, wigner_moyal_cuda_1d.py:854) issue one launch
// instead of 2n + 2. *out stays null until then.
int short_graph(fw_wigner* h, const KArgs& base, int n, bool first, bool last, bool normalise,
                hipGraphExec_t* out) {
  *out = nullptr;
  const int key = n | (first ? 1 << 8 : 0) | (last ? 1 << 9 : 0) | (normalise ? 0 : 1 << 10);
  fw_wigner::ShortGraph& sg = h->sgraphs[key];
  if (!sg.exec && ++sg.uses >= 2) {
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = FW_OK;
    if (first) {
      KArgs a = base;
      a.step = 0;
      a.flags |= fw::F_FIRST;
      rc = launch(h, h->f_v_first, h->col_nb, h->col_nt, a);
    }
    for (int s = 0; s < n && rc == FW_OK; ++s) rc = enqueue_fused(h, s, last && s == n - 1, base, nullptr);
    if (last && rc == FW_OK) rc = enqueue_close(h, base, normalise);
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc != FW_OK) return rc;
    if (e != hipSuccess) return fw::fail(FW_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&sg.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      sg.exec = nullptr;
      return fw::fail(FW_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
  }
  *out = sg.exec;
  return FW_OK;
}

// graphs: replay the handle's captured 64-step graph for whole chunks and the
// short-chunk graphs for the rest (their kernel arguments name the handle's
// own buffers, so a run on other buffers launches every kernel itself)
int run_steps_fused(fw_wigner* h, const KArgs& base0, const double* times, int steps, double* series,
                    bool graphs = true, bool normalise = true) {
  KArgs base = base0;
  FWCHK(ensure_tables(h, 0, base));
  const bool timing = h->timing;
  graphs = graphs && !timing;
  if (graphs && steps > kChunk) FWCHK(ensure_graph_fused(h, base));
  SeriesSink sink(h, base.series, series);
  std::vector<hipEvent_t> evs;
  double acc[2] = {0, 0};
  int nacc = 0;
  bool closed = false;   // the closing normalisation ran inside a short-chunk graph
  for (int off = 0; off < steps; off += kChunk) {
    const int n = std::min(kChunk, steps - off);
    const int nt = std::min(kChunk + 1, steps - off);   // t of steps off .. off+n (the next chunk's first)
    const bool has_last = off + n == steps;
    FWCHK(begin_chunk(h, base, times + off, nt, off == 0));
    hipGraphExec_t sg = nullptr;
    if (graphs && (has_last || n < kChunk)) FWCHK(short_graph(h, base, n, off == 0, has_last, normalise, &sg));
    if (off == 0 && !sg) {
      KArgs a = base;
      a.step = 0;
      a.flags |= fw::F_FIRST;
      FWCHK(launch(h, h->f_v_first, h->col_nb, h->col_nt, a));
    }
    if (sg) {
      HIPCHK(hipGraphLaunch(sg, h->stream));
      closed = has_last;
    } else if (graphs && !has_last && n == kChunk) {
      HIPCHK(hipGraphLaunch(h->gexec2, h->stream));
    } else if (timing) {
      evs.resize(3 * n);
      for (auto& e : evs) HIPCHK(hipEventCreate(&e));
      for (int s = 0; s < n; ++s) FWCHK(enqueue_fused(h, s, has_last && s == n - 1, base, &evs[3 * s]));
      HIPCHK(hipStreamSynchronize(h->stream));
      for (int s = 0; s < n; ++s) {
        float ms0 = 0, ms1 = 0;
        HIPCHK(hipEventElapsedTime(&ms0, evs[3 * s + 1], evs[3 * s + 2]));
        HIPCHK(hipEventElapsedTime(&ms1, evs[3 * s], evs[3 * s + 1]));
        if (!(has_last && s == n - 1)) {
          acc[0] += ms0 * 1000.0;
          acc[1] += ms1 * 1000.0;
          ++nacc;
        }
      }
      for (auto& e : evs) (void)hipEventDestroy(e);
    } else {
      for (int s = 0; s < n; ++s) FWCHK(enqueue_fused(h, s, has_last && s == n - 1, base, nullptr));
    }
    if (has_last) {
      if (!closed) FWCHK(enqueue_close(h, base, normalise));
      FWCHK(sink.final_with_chunk(off, n, steps - 1));
    } else {
      FWCHK(sink.chunk(off, n));
    }
  }
  if (nacc) {
    h->ktimes[0] = acc[0] / nacc;
    h->ktimes[1] = acc[1] / nacc;
    h->ktimes[2] = 0.0;
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  sink.finish();
  h->last_passes = 2;
  return FW_OK;
}

int ensure_kset(fw_wigner* h, int kset) {
  hipModule_t* m = kset == 2 ? &h->mod_bloch : &h->mod_dissip;
  if (*m) return FW_OK;
  std::vector<char> code;
  FWCHK(fw::compile(kset_source(h->prelude, kset), h->arch, code));
  FWCHK(load_module(code, m));
  if (kset == 2) {
    FWCHK(getfn(*m, "fw_bv_first", &h->f_bv_first));
    FWCHK(getfn(*m, "fw_bv_last", &h->f_bv_last));
    FWCHK(getfn(*m, "fw_bk", &h->f_bk));
  } else {
    FWCHK(getfn(*m, "fw_v_first_dissip", &h->f_v_first_dissip));
    FWCHK(getfn(*m, "fw_v_last_dissip", &h->f_v_last_dissip));
  }
  return FW_OK;
}

int build_expr_module(fw_wigner* h, const std::string& body, const char* fname, ExprMod& em) {
  std::string src = kset_source(h->prelude, 0) + body;
  std::vector<char> code;
  FWCHK(fw::compile(src, h->arch, code));
  FWCHK(load_module(code, &em.mod));
  FWCHK(getfn(em.mod, fname, &em.fn));
  return FW_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int fw_wigner_grid_constants(const fw_wigner_desc* desc, double out[4]) {
  if (!desc || !out) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  grid_constants(desc, &out[0], &out[1], &out[2], &out[3]);
  return FW_OK;
}

extern "C" int fw_wigner_prelude(const fw_wigner_desc* desc, const fw_wigner_exprs* exprs, char* out, size_t cap,
                                 size_t* len) {
  fw::clear_error();
  FWCHK(validate(desc, exprs));
  Tiles t;
  FWCHK(resolve_tiles(desc, t));
  const std::string pre = make_prelude(desc, exprs, t);
  if (len) *len = pre.size() + 1;
  if (out && cap > 0) {
    const size_t n = std::min(cap - 1, pre.size());
    std::memcpy(out, pre.data(), n);
    out[n] = '\0';
  }
  return FW_OK;
}

extern "C" int fw_wigner_check(const fw_wigner_desc* desc, const fw_wigner_exprs* exprs) {
  fw::clear_error();
  FWCHK(validate(desc, exprs));
  Tiles t;
  FWCHK(resolve_tiles(desc, t));
  std::vector<char> code;
  const std::string pre = make_prelude(desc, exprs, t);
  for (int kset = 1; kset <= 3; ++kset) FWCHK(fw::compile(kset_source(pre, kset), "gfx950", code));
  return FW_OK;
}

extern "C" int fw_wigner_create(const fw_wigner_desc* desc, const fw_wigner_exprs* exprs, fw_wigner** out) {
  fw::clear_error();
  if (!out) return fw::fail(FW_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  FWCHK(validate(desc, exprs));
  Tiles t;
  FWCHK(resolve_tiles(desc, t));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fw::fail(FW_ERR_NO_DEVICE, "no GPU visible");
  if (desc->device < 0 || desc->device >= ndev) return fw::fail(FW_ERR_INVALID_ARG, "device ordinal out of range");
  HIPCHK(hipSetDevice(desc->device));

  fw_wigner* h = new fw_wigner();
  h->d = *desc;
  h->tl = t;
  h->NX = desc->x_grid_dim;
  h->NP = desc->p_grid_dim;
  h->fp32 = desc->precision == FW_FP32;
  h->esz = h->fp32 ? 4 : 8;
  grid_constants(desc, &h->dX, &h->dP, &h->dLambda, &h->dTheta);
  h->has_mask_xp = exprs->abs_boundary_x_p && *exprs->abs_boundary_x_p;
  h->vtab_ok = tables_v(desc, exprs);
  {
    std::string vs, vf;
    h->vsplit = v_split(desc, exprs, vs, vf);
  }
  h->ktab_ok = tables_k(desc, exprs);
  h->prelude = make_prelude(desc, exprs, t);
  h->arch = fw::target_arch(desc->device);
  h->col_nt = t.col_p * (h->NP / t.col_e);
  h->row_nt = t.row_p * (h->NX / t.row_e);
  h->col_nb = (h->NX / 2) / t.col_p;
  h->row_nb = (h->NP / 2) / t.row_p;

  auto bail = [&](int rc) {
    fw_wigner_destroy(h);
    return rc;
  };
  std::vector<char> code;
  int rc = fw::compile(kset_source(h->prelude, 1), h->arch, code);
  if (rc != FW_OK) return bail(rc);
  if ((rc = load_module(code, &h->mod)) != FW_OK) return bail(rc);
  struct {
    const char* n;
    hipFunction_t* f;
  } fns[] = {{"fw_v_first", &h->f_v_first},   {"fw_v_last", &h->f_v_last},
             {"fw_k", &h->f_k},               {"fw_moments", &h->f_moments},
             {"fw_apply_norm", &h->f_apply_norm}, {"fw_set_scalar", &h->f_set_scalar},
             {"fw_vk_min", &h->f_vk_min},   {"fw_build_vtab", &h->f_build_vtab},
             {"fw_build_ktab", &h->f_build_ktab}, {"fw_vv", &h->f_vv},
             {"fw_k_norm", &h->f_k_norm},     {"fw_build_sumw", &h->f_build_sumw},
             {"fw_begin", &h->f_begin},       {"fw_kgen_check", &h->f_kgen_check}};
  for (auto& f : fns)
    if ((rc = getfn(h->mod, f.n, f.f)) != FW_OK) return bail(rc);

  auto alloc = [&](void** p, size_t bytes) -> int {
    HIPCHK(hipMalloc(p, bytes));
    HIPCHK(hipMemset(*p, 0, bytes));
    return FW_OK;
  };
  const size_t npts = (size_t)h->NX * h->NP;
  const int maxnb = std::max(std::max(h->col_nb, h->row_nb), 1024);
  if ((rc = alloc(&h->W, npts * h->esz)) != FW_OK) return bail(rc);
  // step times, then (split potentials) the field at t + shift dt and at t,
  // and the time-dependent offset of V' at t
  if ((rc = alloc((void**)&h->times, sizeof(double) * 4 * (kChunk + 1))) != FW_OK) return bail(rc);
  if ((rc = alloc((void**)&h->scal, sizeof(double) * fw::NSCAL)) != FW_OK) return bail(rc);
  if ((rc = alloc((void**)&h->partials, sizeof(double) * FW_NSER * maxnb)) != FW_OK) return bail(rc);
  if ((rc = alloc((void**)&h->counter, sizeof(unsigned int) * 4)) != FW_OK) return bail(rc);
  if ((rc = alloc((void**)&h->series, sizeof(double) * FW_NSER * (kChunk + 1))) != FW_OK) return bail(rc);
  if ((rc = fw::upload_twiddles(h->NP, h->fp32, &h->tw_col)) != FW_OK) return bail(rc);
  if ((rc = fw::upload_twiddles(h->NX, h->fp32, &h->tw_row)) != FW_OK) return bail(rc);
  if (hipStreamCreate(&h->stream) != hipSuccess) return bail(fw::fail(FW_ERR_HIP, "hipStreamCreate failed"));
  if (hipHostMalloc((void**)&h->h_stage, sizeof(double) * FW_NSER * ((size_t)kStageChunks * (kChunk + 1) + 1)) !=
          hipSuccess ||
      hipHostMalloc((void**)&h->h_stop, sizeof(double)) != hipSuccess)
    return bail(fw::fail(FW_ERR_HIP, "hipHostMalloc of the staging buffers failed"));
  if ((rc = set_scalar(h, fw::SC_NORM, 1.0)) != FW_OK) return bail(rc);
  if (theta_sums(exprs)) {
    if ((rc = alloc(&h->sw, (size_t)(h->NP / 2 + 1) * 48)) != FW_OK) return bail(rc);
    KArgs a = base_args(h);
    void* sw = h->sw;
    void* params[] = {&a, &sw};
    if (hipModuleLaunchKernel(h->f_build_sumw, (h->NP / 2 + 256) / 256, 1, 1, 256, 1, 1, 0, h->stream, params, nullptr) !=
        hipSuccess)
      return bail(fw::fail(FW_ERR_HIP, "fw_build_sumw launch failed"));
  }
  {
    // the Moyal factor tables now, so that no run allocates (fw_wigner_propagate_noalloc)
    KArgs a = base_args(h);
    if ((rc = ensure_tables(h, 0, a)) != FW_OK) return bail(rc);
    double* sp = nullptr;
    if ((rc = scratch_on(desc->device, &sp)) != FW_OK) return bail(rc);
  }
  if (hipStreamSynchronize(h->stream) != hipSuccess) return bail(fw::fail(FW_ERR_HIP, "initial sync failed"));
  *out = h;
  return FW_OK;
}

extern "C" int fw_wigner_destroy(fw_wigner* h) {
  if (!h) return FW_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->gexec2) (void)hipGraphExecDestroy(h->gexec2);
  for (auto& kv : h->sgraphs)
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
  for (auto* m : {&h->init_mods, &h->avg_mods})
    for (auto& kv : *m)
      if (kv.second.mod) (void)hipModuleUnload(kv.second.mod);
  for (hipModule_t m : {h->mod, h->mod_bloch, h->mod_dissip})
    if (m) (void)hipModuleUnload(m);
  for (void* p : {h->W, h->W0s, h->snap, h->tw_col, h->tw_row, h->vtab, h->ktab, h->sw, (void*)h->times, (void*)h->scal,
                  (void*)h->partials, (void*)h->counter, (void*)h->series})
    if (p) (void)hipFree(p);
  for (double* p : {h->h_stage, h->h_stop})
    if (p) (void)hipHostFree(p);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return FW_OK;
}

static int upload_real(fw_wigner* h, void* dst, const double* src) {
  const size_t n = (size_t)h->NX * h->NP;
  if (h->fp32) {
    std::vector<float> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = (float)src[i];
    HIPCHK(hipMemcpy(dst, tmp.data(), n * 4, hipMemcpyHostToDevice));
  } else {
    HIPCHK(hipMemcpy(dst, src, n * 8, hipMemcpyHostToDevice));
  }
  return FW_OK;
}

extern "C" int fw_wigner_set_state(fw_wigner* h, const double* host_w) {
  if (!h || !host_w) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  HIPCHK(hipStreamSynchronize(h->stream));
  return upload_real(h, h->W, host_w);
}

extern "C" int fw_wigner_set_state_expr(fw_wigner* h, const char* expr) {
  if (!h || !expr) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  auto it = h->init_mods.find(expr);
  if (it == h->init_mods.end()) {
    // init_wigner, wigner_moyal_cuda_1d.py:666-685
    std::string body =
        "extern \"C\" __global__ void fw_init_expr(FW_REAL* W) {\n"
        "  const int i = blockIdx.y;\n"
        "  const int j = threadIdx.x + blockDim.x * blockIdx.x;\n"
        "  if (j >= X_gridDIM) return;\n"
        "  const size_t indexTotal = (size_t)j + (size_t)i * X_gridDIM; (void)indexTotal;\n"
        "  const double X = dX * (j - 0.5 * X_gridDIM);\n"
        "  const double P = dP * (i - 0.5 * P_gridDIM);\n"
        "  W[indexTotal] = (FW_REAL)(" +
        std::string(expr) + ");\n}\n";
    ExprMod em;
    FWCHK(build_expr_module(h, body, "fw_init_expr", em));
    it = h->init_mods.emplace(expr, em).first;
  }
  void* W = h->W;
  void* params[] = {&W};
  const int bx = std::min(h->NX, 256);
  HIPCHK(hipModuleLaunchKernel(it->second.fn, (h->NX + bx - 1) / bx, h->NP, 1, bx, 1, 1, 0, h->stream, params,
                               nullptr));
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

extern "C" int fw_wigner_fill(fw_wigner* h, double value) {
  if (!h) return fw::fail(FW_ERR_INVALID_ARG, "NULL handle");
  std::vector<double> v((size_t)h->NX * h->NP, value);
  return fw_wigner_set_state(h, v.data());
}

static int moments_at(fw_wigner* h, double t) {
  HIPCHK(hipMemcpyAsync(h->times, &t, sizeof(double), hipMemcpyHostToDevice, h->stream));
  KArgs a = base_args(h);
  a.step = 0;
  FWCHK(launch(h, h->f_moments, helper_blocks(h), 256, a));
  return FW_OK;
}

extern "C" int fw_wigner_normalize(fw_wigner* h, double* norm_before) {
  if (!h) return fw::fail(FW_ERR_INVALID_ARG, "NULL handle");
  FWCHK(moments_at(h, 0.0));
  KArgs a = base_args(h);
  FWCHK(launch(h, h->f_apply_norm, 1024, 256, a));
  FWCHK(set_scalar(h, fw::SC_NORM, 1.0));
  double s0 = 0;
  HIPCHK(hipMemcpyAsync(&s0, h->scal + fw::SC_OUT, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (norm_before) *norm_before = s0 * h->dX * h->dP;
  return FW_OK;
}

extern "C" int fw_wigner_get_state(fw_wigner* h, double* host_w) {
  if (!h || !host_w) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  HIPCHK(hipStreamSynchronize(h->stream));
  const size_t n = (size_t)h->NX * h->NP;
  if (h->fp32) {
    std::vector<float> tmp(n);
    HIPCHK(hipMemcpy(tmp.data(), h->W, n * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) host_w[i] = tmp[i];
  } else {
    HIPCHK(hipMemcpy(host_w, h->W, n * 8, hipMemcpyDeviceToHost));
  }
  return FW_OK;
}

extern "C" int fw_wigner_state_ptr(fw_wigner* h, void** dev_ptr, int* elem_size) {
  if (!h || !dev_ptr || !elem_size) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  *dev_ptr = h->W;
  *elem_size = (int)h->esz;
  return FW_OK;
}

extern "C" int fw_wigner_propagate(fw_wigner* h, const double* times, int steps, double* series) {
  if (!h || (!times && steps > 0)) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  if (steps < 0) return fw::fail(FW_ERR_INVALID_ARG, "steps must be >= 0");
  if (steps == 0) return FW_OK;
  KArgs a = base_args(h);
  if (!h->dissip && h->d.step_passes != 3) return run_steps_fused(h, a, times, steps, series);
  h->last_passes = 3;
  return run_steps(h, 0, a, times, steps, series, nullptr);
}

extern "C" int fw_wigner_single_step(fw_wigner* h, double t, double* sums) {
  if (!h) return fw::fail(FW_ERR_INVALID_ARG, "NULL handle");
  const double times[1] = {t};
  KArgs a = base_args(h);
  if (!h->dissip && h->d.step_passes != 3) return run_steps_fused(h, a, times, 1, sums, false, false);
  h->last_passes = 3;
  return run_steps(h, 0, a, times, 1, sums, nullptr, false);
}

extern "C" int fw_wigner_propagate_noalloc(fw_wigner* h, void* w_dev, const double* times, int steps,
                                           double* series) {
  if (!h || !w_dev || (!times && steps > 0)) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  if (steps < 0) return fw::fail(FW_ERR_INVALID_ARG, "steps must be >= 0");
  if (h->dissip || h->d.step_passes == 3)
    return fw::fail(FW_ERR_INVALID_ARG, "fw_wigner_propagate_noalloc runs two-pass Moyal steps only (no dissipator, "
                                        "step_passes 0 or 2)");
  const int nb = std::max(h->col_nb, h->row_nb);
  if (nb > kNoAllocMaxBlocks)
    return fw::fail(FW_ERR_INVALID_ARG, "fw_wigner_propagate_noalloc: the tiles need " + std::to_string(nb) +
                                            " blocks per pass; its static scratch holds the partial sums of at most " +
                                            std::to_string(kNoAllocMaxBlocks) + " (use larger col_p / row_p)");
  double* sp = nullptr;
  FWCHK(scratch_on(h->d.device, &sp));
  if (steps == 0) return FW_OK;
  HIPCHK(hipSetDevice(h->d.device));
  KArgs a = base_args(h);
  a.W = w_dev;
  a.scal = sp + kScScal;
  a.times = sp + kScTimes;
  a.series = sp + kScSeries;
  a.counter = reinterpret_cast<unsigned int*>(sp + kScCounter);
  a.partials = sp + kScPartials;
  a.nbp = h->col_nb;
  // the tables were built at creation (the allocation happened there); a Bloch
  // run may have refilled them, which ensure_tables undoes without allocating
  return run_steps_fused(h, a, times, steps, series, false);
}

extern "C" int fw_wigner_moments(fw_wigner* h, double t, double sums[FW_NSER]) {
  if (!h || !sums) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  FWCHK(moments_at(h, t));
  HIPCHK(hipMemcpyAsync(sums, h->scal + fw::SC_OUT, sizeof(double) * FW_NSER, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

extern "C" int fw_wigner_get_average(fw_wigner* h, const char* expr, double t, double* out) {
  if (!h || !expr || !out) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  auto it = h->avg_mods.find(expr);
  if (it == h->avg_mods.end()) {
    // weighted_func_cuda_code + gpuarray.sum, wigner_moyal_cuda_1d.py:378-396,687-709
    std::string body =
        "extern \"C\" __global__ void __launch_bounds__(256) fw_avg_expr(fwk::Args a) {\n"
        "  __shared__ double red[fwk::NSER * 4];\n"
        "  const double t = a.times[a.step];\n"
        "  double s[fwk::NSER];\n"
        "  for (int k = 0; k < fwk::NSER; ++k) s[k] = 0.0;\n"
        "  for (int i = blockIdx.x; i < P_gridDIM; i += gridDim.x) {\n"
        "    for (int j = threadIdx.x; j < X_gridDIM; j += 256) {\n"
        "      const size_t indexTotal = (size_t)j + (size_t)i * X_gridDIM;\n"
        "      const double X = dX * (j - 0.5 * X_gridDIM);\n"
        "      const double P = dP * (i - 0.5 * P_gridDIM);\n"
        "      s[0] += (double)a.W[indexTotal] * (" +
        std::string(expr) +
        ");\n"
        "    }\n"
        "  }\n"
        "  fwk::block_reduce_publish<256>(s, red, a.partials, a.counter, blockIdx.x, gridDim.x, a.scal + "
        "fwk::SC_OUT);\n"
        "}\n";
    ExprMod em;
    FWCHK(build_expr_module(h, body, "fw_avg_expr", em));
    it = h->avg_mods.emplace(expr, em).first;
  }
  HIPCHK(hipMemcpyAsync(h->times, &t, sizeof(double), hipMemcpyHostToDevice, h->stream));
  KArgs a = base_args(h);
  FWCHK(launch(h, it->second.fn, helper_blocks(h), 256, a));
  double s = 0;
  HIPCHK(hipMemcpyAsync(&s, h->scal + fw::SC_OUT, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  *out = s * h->dX * h->dP;
  return FW_OK;
}

extern "C" int fw_wigner_bloch_setup(fw_wigner* h, int convention, double t_initial, double mins_out[2]) {
  if (!h) return fw::fail(FW_ERR_INVALID_ARG, "NULL handle");
  if (convention != FW_BLOCH_CUDA && convention != FW_BLOCH_FFTW)
    return fw::fail(FW_ERR_INVALID_ARG, "unknown Bloch convention");
  FWCHK(ensure_kset(h, 2));
  h->bloch_conv = convention;
  h->t_initial = t_initial;
  KArgs a = base_args(h);
  a.t_initial = t_initial;
  FWCHK(launch(h, h->f_vk_min, 1, 1024, a));
  double m[2];
  HIPCHK(hipMemcpyAsync(m, h->scal + fw::SC_OUT, sizeof m, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->vmin = m[0];
  h->kmin = m[1];
  h->bloch_ready = true;
  if (mins_out) {
    mins_out[0] = m[0];
    mins_out[1] = m[1];
  }
  return FW_OK;
}

extern "C" int fw_wigner_bloch_propagate(fw_wigner* h, double dbeta, int steps, double purity_guard,
                                         double* series, int* steps_done) {
  if (!h) return fw::fail(FW_ERR_INVALID_ARG, "NULL handle");
  if (!h->bloch_ready) return fw::fail(FW_ERR_INVALID_ARG, "call fw_wigner_bloch_setup first");
  if (steps < 0) return fw::fail(FW_ERR_INVALID_ARG, "steps must be >= 0");
  if (steps_done) *steps_done = 0;
  if (steps == 0) return FW_OK;
  KArgs a = base_args(h);
  a.dbeta = dbeta;
  const bool fftw = h->bloch_conv == FW_BLOCH_FFTW;
  a.cv = fftw ? 0.5 : 0.25;
  a.ck = fftw ? 1.0 : 0.5;
  if (fftw) a.flags |= fw::F_BLOCH_MASK;
  a.flags &= ~fw::F_MASK_XP;
  a.t_initial = h->t_initial;
  a.vmin = h->vmin;
  a.kmin = h->kmin;
  a.guard = purity_guard;
  std::vector<double> times(steps, h->t_initial);
  return run_steps(h, 1, a, times.data(), steps, series, steps_done);
}

extern "C" int fw_wigner_set_dissipator(fw_wigner* h, const double* host_w0s, double w_coeff) {
  if (!h) return fw::fail(FW_ERR_INVALID_ARG, "NULL handle");
  if (!host_w0s) {
    // the cached three-pass graph holds the dissipative kernels and the old coefficient
    if (h->gexec) {
      (void)hipGraphExecDestroy(h->gexec);
      h->gexec = nullptr;
    }
    h->dissip = false;
    h->w_coeff = 1.0;
    return FW_OK;
  }
  FWCHK(ensure_kset(h, 3));
  if (!h->W0s) HIPCHK(hipMalloc(&h->W0s, (size_t)h->NX * h->NP * h->esz));
  FWCHK(upload_real(h, h->W0s, host_w0s));
  h->dissip = true;
  h->w_coeff = w_coeff;
  if (h->gexec) {
    (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
  }
  return FW_OK;
}

extern "C" int fw_wigner_snapshot(fw_wigner* h, int which) {
  if (!h) return fw::fail(FW_ERR_INVALID_ARG, "NULL handle");
  const size_t bytes = (size_t)h->NX * h->NP * h->esz;
  if (which == 0) {
    if (!h->snap) HIPCHK(hipMalloc(&h->snap, bytes));
    HIPCHK(hipMemcpyAsync(h->snap, h->W, bytes, hipMemcpyDeviceToDevice, h->stream));
  } else if (which == 1) {
    if (!h->snap) return fw::fail(FW_ERR_INVALID_ARG, "no snapshot saved");
    HIPCHK(hipMemcpyAsync(h->W, h->snap, bytes, hipMemcpyDeviceToDevice, h->stream));
  } else {
    return fw::fail(FW_ERR_INVALID_ARG, "which must be 0 (save) or 1 (restore)");
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

extern "C" int fw_wigner_set_timing(fw_wigner* h, int enable) {
  if (!h) return fw::fail(FW_ERR_INVALID_ARG, "NULL handle");
  h->timing = enable != 0;
  return FW_OK;
}

extern "C" int fw_wigner_step_passes(fw_wigner* h, int* passes) {
  if (!h || !passes) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  *passes = h->last_passes;
  return FW_OK;
}

extern "C" int fw_wigner_k_factors(fw_wigner* h, int* generated, double* max_dev) {
  if (!h || !generated || !max_dev) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  *generated = h->kgen == 1 ? 1 : 0;
  *max_dev = h->kgen_err;
  return FW_OK;
}

extern "C" int fw_wigner_kernel_bench(fw_wigner* h, double t, int reps, double us_out[2]) {
  if (!h || !us_out || reps < 1) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument or reps < 1");
  if (h->dissip || h->d.step_passes == 3)
    return fw::fail(FW_ERR_INVALID_ARG, "the kernel bench times the two-pass step kernels");
  KArgs base = base_args(h);
  FWCHK(ensure_tables(h, 0, base));
  FWCHK(ensure_graph_fused(h, base));
  const size_t bytes = (size_t)h->NX * h->NP * h->esz;
  void* save = nullptr;
  HIPCHK(hipMalloc(&save, bytes));
  HIPCHK(hipMemcpyAsync(save, h->W, bytes, hipMemcpyDeviceToDevice, h->stream));
  std::vector<double> tt(kChunk + 1);
  for (int s = 0; s <= kChunk; ++s) tt[s] = t + (s + 1) * h->d.dt;
  FWCHK(begin_chunk(h, base, tt.data(), kChunk + 1, true));
  KArgs a = base;
  a.step = 0;
  a.flags |= fw::F_FIRST;
  const int chunks = std::max(1, reps / kChunk);
  std::vector<hipEvent_t> ev(2 + 3 * kChunk);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  int rc = launch(h, h->f_v_first, h->col_nb, h->col_nt, a);
  // the step time: graph replays (the propagator's own launch path) between two events
  if (rc == FW_OK) rc = hipEventRecord(ev[0], h->stream) == hipSuccess ? FW_OK : FW_ERR_HIP;
  for (int c = 0; c < chunks && rc == FW_OK; ++c) rc = hipGraphLaunch(h->gexec2, h->stream) == hipSuccess ? FW_OK : FW_ERR_HIP;
  if (rc == FW_OK) rc = hipEventRecord(ev[1], h->stream) == hipSuccess ? FW_OK : FW_ERR_HIP;
  // each kernel's share: one chunk with an event around every kernel
  for (int s = 0; s < kChunk && rc == FW_OK; ++s) rc = enqueue_fused(h, s, false, base, &ev[2 + 3 * s]);
  double tk = 0.0, tv = 0.0;
  float step_ms = 0.f;
  if (rc == FW_OK && hipStreamSynchronize(h->stream) == hipSuccess) {
    (void)hipEventElapsedTime(&step_ms, ev[0], ev[1]);
    for (int s = 0; s < kChunk; ++s) {
      float m0 = 0, m1 = 0;
      (void)hipEventElapsedTime(&m0, ev[2 + 3 * s], ev[3 + 3 * s]);
      (void)hipEventElapsedTime(&m1, ev[3 + 3 * s], ev[4 + 3 * s]);
      tk += m0;
      tv += m1;
    }
    const double step_us = step_ms * 1000.0 / (chunks * kChunk);
    us_out[0] = step_us * tv / (tv + tk);
    us_out[1] = step_us * tk / (tv + tk);
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  // the state as it was before the call
  HIPCHK(hipMemcpyAsync(h->W, save, bytes, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  (void)hipFree(save);
  if (rc != FW_OK) return fw::fail(FW_ERR_HIP, "kernel bench launch failed");
  return FW_OK;
}

extern "C" int fw_wigner_kernel_times(fw_wigner* h, double us_out[3]) {
  if (!h || !us_out) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  for (int k = 0; k < 3; ++k) us_out[k] = h->ktimes[k];
  return FW_OK;
}
