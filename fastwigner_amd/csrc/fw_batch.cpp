// fw_batch.cpp -- batched propagator handles of libfastwigner.so: many
// independent Wigner functions advanced by one persistent launch per step (or
// one launch per pass), replayed from a captured hipGraph per 64 steps.
// Device code: fw_batch.h. Contract: include/fastwigner.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "fastwigner.h"
#include "fw_internal.h"

namespace {

const int kChunk = 64;
const int kNSizes = 5;   // N = 64 << s, s < kNSizes
const int kBT = 256;

// host mirrors of fwk::BInst, fwk::BTask, fwk::BArgs (fw_batch.h)
struct BInstH {
  long long off;
  int n, ntile, part, dissip, nbloch, swoff;
  double dX, dP, dLambda, dTheta;
  double w_coeff, dbeta;
  double prm[FW_BATCH_NPRM];
};
static_assert(sizeof(BInstH) == 144, "BInst layout");

struct BTaskH {
  int inst, code;
};

struct BArgsH {
  void* W;
  const void* W0s;
  void* G;
  void* W0t;
  const void* sw;
  const void* inst;
  const void* tasks;
  const int* qoff;
  const void* tw;
  const double* times;
  double* partials;
  double* series;
  unsigned int* qhead;
  unsigned int* done;
  int* fresh;
  int* err;
  const double* vkmin;
  int step, gstep, ninst, mode, task_base, spin_limit;
  int vv, pad_;
  double cv, ck;
};

// Does the expression reference the identifier t (the time argument)?
bool uses_t_token(const char* s) {
  if (!s) return false;
  auto ident = [](char c) { return isalnum((unsigned char)c) || c == '_'; };
  const size_t n = strlen(s);
  for (size_t i = 0; i < n; ++i)
    if (s[i] == 't' && (i == 0 || !ident(s[i - 1])) && (i + 1 == n || !ident(s[i + 1]))) return true;
  return false;
}

int size_index(int n) {
  for (int s = 0; s < kNSizes; ++s)
    if (n == (64 << s)) return s;
  return -1;
}

int default_e(int n) {
  switch (n) {
    case 64: return 4;
    case 128: return 4;
    case 256: return 8;
    case 512: return 8;
    default: return 8;
  }
}

// elements per thread E for the FFT of length N in a 256-thread tile: the tile
// is P = 256 E / N line pairs, which must divide the N/2 pairs.
bool e_ok(int n, int e) {
  if (!fw::is_pow2(e) || e < 2 || e > 16 || e > n) return false;
  const int tpl = n / e;
  if (tpl > kBT) return false;
  const int p = kBT / tpl;
  return (n / 2) % p == 0;
}

std::string or_default(const char* s, const char* d) { return (s && *s) ? std::string(s) : std::string(d); }

std::vector<std::string> split_names(const char* s) {
  std::vector<std::string> out;
  if (!s) return out;
  std::string cur;
  for (const char* p = s;; ++p) {
    if (*p == ',' || *p == 0) {
      std::string t;
      for (char c : cur)
        if (!isspace((unsigned char)c)) t += c;
      if (!t.empty()) out.push_back(t);
      cur.clear();
      if (*p == 0) break;
    } else {
      cur += *p;
    }
  }
  return out;
}

}  // namespace

struct fw_batch {
  fw_batch_desc d{};
  std::vector<fw_batch_instance> src;
  std::vector<BInstH> inst;
  int e_for[kNSizes] = {0, 0, 0, 0, 0};
  bool fp32 = true;
  size_t esz = 4;
  int nmax = 64;
  long long points = 0;
  std::string arch, prelude;
  hipModule_t mod = nullptr;
  hipFunction_t f_step = nullptr, f_bloch = nullptr, f_norm = nullptr, f_vkmin = nullptr, f_zero = nullptr,
                f_fill = nullptr, f_gibbs = nullptr, f_sumw = nullptr;
  hipStream_t stream = nullptr;

  void* W = nullptr;
  void* W0s = nullptr;
  void* G = nullptr;          // Vopen(W0s) of the two-pass steps
  void* W0t = nullptr;        // W0s on the theta spectrum (two-pass theta steps)
  void* sw = nullptr;         // theta-sum weights of every instance (two-pass theta steps)
  bool two_pass = false;      // mode 1 real-time runs in two passes per step (V independent of t)
  bool theta = false;         // two-pass steps with the sums on the theta spectrum (K, K' independent of t)
  void* inst_d = nullptr;
  void* qtasks_d = nullptr;   // mode 0 queues
  void* ptasks_d = nullptr;   // mode 1, pass-major, XCD-interleaved
  int* qoff_d = nullptr;
  void* tw = nullptr;
  double* times = nullptr;
  double* partials = nullptr;
  double* series = nullptr;
  unsigned int* counters = nullptr;   // qhead [kChunk*8] then done [kChunk*ninst*3]
  int ncounters = 0;
  int* fresh = nullptr;
  int* err = nullptr;
  double* vkmin = nullptr;

  int ntasks = 0;                 // tasks per step (mode 0)
  int ptask_off[4] = {0, 0, 0, 0};  // mode 1 pass offsets into ptasks
  int grid = 0;                   // persistent grid (mode 0)
  hipGraphExec_t gexec = nullptr;
};

namespace {

int validate(const fw_batch_desc* d, const fw_batch_instance* in, const fw_batch_exprs* x) {
  if (!d || !in || !x) return fw::fail(FW_ERR_INVALID_ARG, "descriptor, instances or expressions are NULL");
  if (d->n_instances < 1 || d->n_instances > 65535)
    return fw::fail(FW_ERR_INVALID_ARG, "n_instances must be in [1, 65535]");
  if (d->precision != FW_FP64 && d->precision != FW_FP32)
    return fw::fail(FW_ERR_INVALID_ARG, "precision must be FW_FP64 or FW_FP32");
  if (d->mode != 0 && d->mode != 1) return fw::fail(FW_ERR_INVALID_ARG, "mode must be 0 or 1");
  if (!std::isfinite(d->dt)) return fw::fail(FW_ERR_INVALID_ARG, "dt must be finite");
  if (!x->V || !x->K) return fw::fail(FW_ERR_INVALID_ARG, "V and K expressions are required");
  if (split_names(x->param_names).size() > FW_BATCH_NPRM)
    return fw::fail(FW_ERR_INVALID_ARG, "at most 8 parameter names");
  for (int i = 0; i < d->n_instances; ++i) {
    if (size_index(in[i].grid_dim) < 0)
      return fw::fail(FW_ERR_INVALID_ARG, "instance " + std::to_string(i) +
                                              ": grid_dim must be one of 64, 128, 256, 512, 1024");
    if (!(in[i].x_amplitude > 0) || !(in[i].p_amplitude > 0))
      return fw::fail(FW_ERR_INVALID_ARG, "instance " + std::to_string(i) + ": amplitudes must be positive");
  }
  for (int s = 0; s < kNSizes; ++s)
    if (d->e_for_n[s] != 0 && !e_ok(64 << s, d->e_for_n[s]))
      return fw::fail(FW_ERR_INVALID_ARG, "e_for_n[" + std::to_string(s) + "] is not a valid tile for N = " +
                                              std::to_string(64 << s));
  return FW_OK;
}

std::string make_prelude(const fw_batch_desc* d, const fw_batch_instance* in, const fw_batch_exprs* x,
                         const int* e_for, int nmax, const bool* has) {
  const bool fp32 = d->precision == FW_FP32;
  std::ostringstream s;
  s << "#define FW_REAL " << (fp32 ? "float" : "double") << "\n";
  s << "#define FW_FP32 " << (fp32 ? 1 : 0) << "\n";
  s << "#define FW_PHASE_SHIFT (" << fw::fmt_double(d->phase_time_shift) << ")\n";
  // plain (write-back) stores of W: the batch's grids exceed the caches, its
  // passes are long against their boundaries, and device-scope stores made the
  // 256-instance FP32 step slower (842 -> 1105 us)
  s << "#define FW_W_STORE 0\n";
  s << "#define FW_B_NMAX " << nmax << "\n";
  int emax = 2;
  for (int k = 0; k < kNSizes; ++k) {
    const int n = 64 << k;
    s << "#define FW_B_HAS_" << n << " " << (has[k] ? 1 : 0) << "\n";
    s << "#define FW_B_E" << n << " " << e_for[k] << "\n";
    if (has[k]) emax = std::max(emax, e_for[k]);
  }
  s << "#define FW_B_EMAX " << emax << "\n";
  s << "#ifndef M_PI\n#define M_PI 3.14159265358979323846\n#endif\n";
  s << "const double dt = " << fw::fmt_double(d->dt) << ";\n";
  // the caller's sin and cos: the branch-free FP64 ones of fw_math.h (see there)
  s << "#include \"fw_math.h\"\n#define sin fw_user_sin\n#define cos fw_user_cos\n";
  s << "// ---- caller constants and functions ----\n";
  if (x->consts) s << x->consts << "\n";
  std::string bind;
  const auto names = split_names(x->param_names);
  for (size_t k = 0; k < names.size(); ++k)
    bind += "  const double " + names[k] + " = prm[" + std::to_string(k) + "]; (void)" + names[k] + ";\n";
  s << "// ---- spliced expressions (parameters bound per instance) ----\n";
  s << "__device__ double V(double X, double t, const double* prm) {\n" << bind << "  return (" << x->V << ");\n}\n";
  s << "__device__ double K(double P, double t, const double* prm) {\n" << bind << "  return (" << x->K << ");\n}\n";
  s << "__device__ double fw_diff_V(double X, double t, const double* prm) {\n"
    << bind << "  return (" << or_default(x->diff_V, "0.") << ");\n}\n";
  s << "__device__ double fw_diff_K(double P, double t, const double* prm) {\n"
    << bind << "  return (" << or_default(x->diff_K, "0.") << ");\n}\n";
  s << "#undef sin\n#undef cos\n";
  s << "#define FW_B_THETA "
    << ((d->mode == 1 && !uses_t_token(x->V) && !uses_t_token(x->K) && !uses_t_token(x->diff_K)) ? 1 : 0) << "\n";
  s << "#include \"fw_batch.h\"\n";
  return s.str();
}

// resolve tiles and which sizes are present; returns the prelude
int plan(const fw_batch_desc* d, const fw_batch_instance* in, const fw_batch_exprs* x, int* e_for, int* nmax,
         std::string& src) {
  FWCHK(validate(d, in, x));
  bool has[kNSizes] = {false, false, false, false, false};
  *nmax = 64;
  for (int i = 0; i < d->n_instances; ++i) {
    has[size_index(in[i].grid_dim)] = true;
    *nmax = std::max(*nmax, in[i].grid_dim);
  }
  for (int k = 0; k < kNSizes; ++k) e_for[k] = d->e_for_n[k] ? d->e_for_n[k] : default_e(64 << k);
  src = make_prelude(d, in, x, e_for, *nmax, has);
  return FW_OK;
}

int launch(fw_batch* h, hipFunction_t f, int grid, BArgsH& a) {
  void* params[] = {&a};
  HIPCHK(hipModuleLaunchKernel(f, grid, 1, 1, kBT, 1, 1, 0, h->stream, params, nullptr));
  return FW_OK;
}

BArgsH base_args(fw_batch* h) {
  BArgsH a;
  std::memset(&a, 0, sizeof a);
  a.W = h->W;
  a.W0s = h->W0s;
  a.G = h->G;
  a.W0t = h->W0t;
  a.sw = h->sw;
  a.inst = h->inst_d;
  a.tasks = h->d.mode == 0 ? h->qtasks_d : h->ptasks_d;
  a.qoff = h->qoff_d;
  a.tw = h->tw;
  a.times = h->times;
  a.partials = h->partials;
  a.series = h->series;
  a.qhead = h->counters;
  a.done = h->counters + kChunk * 8;
  a.fresh = h->fresh;
  a.err = h->err;
  a.vkmin = h->vkmin;
  a.ninst = h->d.n_instances;
  a.mode = h->d.mode;
  a.spin_limit = 1 << 22;
  return a;
}

int fill_int(fw_batch* h, int* p, int n, int v) {
  void* params[] = {&p, &n, &v};
  HIPCHK(hipModuleLaunchKernel(h->f_fill, std::max(1, std::min(1024, (n + 255) / 256)), 1, 1, 256, 1, 1, 0,
                               h->stream, params, nullptr));
  return FW_OK;
}

int zero_counters(fw_batch* h) {
  unsigned int* p = h->counters;
  int n = h->ncounters;
  void* params[] = {&p, &n};
  HIPCHK(hipModuleLaunchKernel(h->f_zero, std::max(1, std::min(1024, (n + 255) / 256)), 1, 1, 256, 1, 1, 0,
                               h->stream, params, nullptr));
  return FW_OK;
}

// one step's launches (chunk-local step s)
int enqueue_step(fw_batch* h, int kind, int s, int gstep) {
  BArgsH a = base_args(h);
  a.step = s;
  a.gstep = gstep;
  hipFunction_t f = kind == 1 ? h->f_bloch : h->f_step;
  if (h->d.mode == 0) return launch(h, f, h->grid, a);
  for (int p = 0; p < 3; ++p) {
    a.task_base = h->ptask_off[p];
    FWCHK(launch(h, f, h->ptask_off[p + 1] - h->ptask_off[p], a));
  }
  return FW_OK;
}

// one pass of the mode-1 task lists: pass 0 (column tiles), 1 (row tiles) or
// 2 (column tiles), with the two-pass selector vv
int enqueue_pass(fw_batch* h, int pass, int vv, int s, int gstep) {
  BArgsH a = base_args(h);
  a.step = s;
  a.gstep = gstep;
  a.vv = vv;
  a.task_base = h->ptask_off[pass];
  return launch(h, h->f_step, h->ptask_off[pass + 1] - h->ptask_off[pass], a);
}

// two-pass steps: the K pass of step s, then the fused column pass (closing
// V/2 of s, opening V/2 of s+1)
int enqueue_step2(fw_batch* h, int s, int gstep) {
  FWCHK(enqueue_pass(h, 1, 1, s, gstep));
  return enqueue_pass(h, 0, 1, s, gstep);
}

int ensure_graph(fw_batch* h) {
  if (h->gexec) return FW_OK;
  hipGraph_t g;
  HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = zero_counters(h);
  for (int s = 0; s < kChunk && rc == FW_OK; ++s) rc = h->two_pass ? enqueue_step2(h, s, s) : enqueue_step(h, 0, s, s);
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (rc != FW_OK) return rc;
  if (e != hipSuccess) return fw::fail(FW_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fw::fail(FW_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  return FW_OK;
}

int check_err(fw_batch* h) {
  int e = 0;
  HIPCHK(hipMemcpyAsync(&e, h->err, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (e) return fw::fail(FW_ERR_HIP, "batched step: a dependency wait exceeded its spin limit");
  return FW_OK;
}

// closing normalisation; publishes the last step's sums to slot kChunk
int final_norm(fw_batch* h) {
  BArgsH a = base_args(h);
  int slot = kChunk;
  void* params[] = {&a, &slot};
  HIPCHK(hipModuleLaunchKernel(h->f_norm, h->d.n_instances * 8, 1, 1, kBT, 1, 1, 0, h->stream, params, nullptr));
  return FW_OK;
}

int instance_check(fw_batch* h, int i) {
  if (!h) return fw::fail(FW_ERR_INVALID_ARG, "NULL handle");
  if (i < 0 || i >= h->d.n_instances) return fw::fail(FW_ERR_INVALID_ARG, "instance index out of range");
  return FW_OK;
}

int upload_instances(fw_batch* h) {
  HIPCHK(hipMemcpyAsync(h->inst_d, h->inst.data(), sizeof(BInstH) * h->inst.size(), hipMemcpyHostToDevice,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

}  // namespace

extern "C" int fw_batch_check(const fw_batch_desc* d, const fw_batch_instance* in, const fw_batch_exprs* x) {
  fw::clear_error();
  int e_for[kNSizes], nmax;
  std::string src;
  FWCHK(plan(d, in, x, e_for, &nmax, src));
  std::vector<char> code;
  return fw::compile(src, fw::target_arch(-1), code);
}

extern "C" int fw_batch_create(const fw_batch_desc* d, const fw_batch_instance* in, const fw_batch_exprs* x,
                               fw_batch** out) {
  fw::clear_error();
  if (!out) return fw::fail(FW_ERR_INVALID_ARG, "out is NULL");
  *out = nullptr;
  auto h = new fw_batch();
  std::unique_ptr<fw_batch, int (*)(fw_batch*)> guard(h, fw_batch_destroy);
  FWCHK(plan(d, in, x, h->e_for, &h->nmax, h->prelude));
  h->d = *d;
  h->src.assign(in, in + d->n_instances);
  h->two_pass = d->mode == 1 && !uses_t_token(x->V);
  h->theta = h->two_pass && !uses_t_token(x->K) && !uses_t_token(x->diff_K);
  h->fp32 = d->precision == FW_FP32;
  h->esz = h->fp32 ? 4 : 8;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fw::fail(FW_ERR_NO_DEVICE, "no GPU visible");
  if (d->device < 0 || d->device >= ndev) return fw::fail(FW_ERR_INVALID_ARG, "device ordinal out of range");
  HIPCHK(hipSetDevice(d->device));
  h->arch = fw::target_arch(d->device);
  std::vector<char> code;
  FWCHK(fw::compile(h->prelude, h->arch, code));
  HIPCHK(hipModuleLoadData(&h->mod, code.data()));
  HIPCHK(hipModuleGetFunction(&h->f_step, h->mod, "fw_batch_step"));
  HIPCHK(hipModuleGetFunction(&h->f_bloch, h->mod, "fw_batch_bloch_step"));
  HIPCHK(hipModuleGetFunction(&h->f_norm, h->mod, "fw_batch_norm"));
  HIPCHK(hipModuleGetFunction(&h->f_vkmin, h->mod, "fw_batch_vkmin"));
  HIPCHK(hipModuleGetFunction(&h->f_zero, h->mod, "fw_batch_zero"));
  HIPCHK(hipModuleGetFunction(&h->f_fill, h->mod, "fw_batch_fill_int"));
  HIPCHK(hipModuleGetFunction(&h->f_gibbs, h->mod, "fw_batch_gibbs"));
  HIPCHK(hipModuleGetFunction(&h->f_sumw, h->mod, "fw_batch_sumw"));
  HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));

  // instances: offsets, tiles, partial rows
  const int ni = d->n_instances;
  h->inst.resize(ni);
  const double two_pi = 2.0 * 3.141592653589793;
  long long off = 0;
  int part = 0, swn = 0;
  for (int i = 0; i < ni; ++i) {
    BInstH& b = h->inst[i];
    std::memset(&b, 0, sizeof b);
    const int n = in[i].grid_dim, e = h->e_for[size_index(n)];
    b.off = off;
    b.n = n;
    b.ntile = (n / 2) / (kBT * e / n);
    b.part = part;
    b.dX = 2.0 * in[i].x_amplitude / n;
    b.dP = 2.0 * in[i].p_amplitude / n;
    b.dLambda = 1.0 / ((double)n * (b.dX / two_pi));
    b.dTheta = 1.0 / ((double)n * (b.dP / two_pi));
    b.w_coeff = 1.0;
    for (int k = 0; k < FW_BATCH_NPRM; ++k) b.prm[k] = in[i].params[k];
    b.swoff = swn;
    swn += n / 2 + 1;
    off += (long long)n * n;
    part += b.ntile;
  }
  h->points = off;

  // XCD queues: instances by decreasing size, each to the least-loaded queue;
  // within a queue, pass-major, instances by decreasing size, tiles ascending
  std::vector<int> order(ni);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return in[a].grid_dim > in[b].grid_dim; });
  std::vector<std::vector<int>> qinst(8);
  std::vector<long long> load(8, 0);
  for (int i : order) {
    const int q = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    qinst[q].push_back(i);
    load[q] += (long long)in[i].grid_dim * in[i].grid_dim;
  }
  std::vector<BTaskH> qt;
  std::vector<int> qoff(9, 0);
  std::vector<std::vector<BTaskH>> perq_pass[3];
  for (int p = 0; p < 3; ++p) perq_pass[p].resize(8);
  for (int q = 0; q < 8; ++q) {
    qoff[q] = (int)qt.size();
    for (int p = 0; p < 3; ++p)
      for (int i : qinst[q])
        for (int t = 0; t < h->inst[i].ntile; ++t) {
          qt.push_back({i, (p << 16) | t});
          perq_pass[p][q].push_back({i, (p << 16) | t});
        }
  }
  qoff[8] = (int)qt.size();
  h->ntasks = (int)qt.size();
  // mode 1: per pass, block b takes queue b % 8's (b / 8)-th task (padding where a queue is short)
  std::vector<BTaskH> pt;
  for (int p = 0; p < 3; ++p) {
    h->ptask_off[p] = (int)pt.size();
    size_t longest = 0;
    for (int q = 0; q < 8; ++q) longest = std::max(longest, perq_pass[p][q].size());
    for (size_t r = 0; r < longest; ++r)
      for (int q = 0; q < 8; ++q) pt.push_back(r < perq_pass[p][q].size() ? perq_pass[p][q][r] : BTaskH{-1, 0});
  }
  h->ptask_off[3] = (int)pt.size();

  // buffers
  HIPCHK(hipMalloc(&h->W, (size_t)h->points * h->esz));
  HIPCHK(hipMemsetAsync(h->W, 0, (size_t)h->points * h->esz, h->stream));
  HIPCHK(hipMalloc(&h->inst_d, sizeof(BInstH) * ni));
  HIPCHK(hipMalloc(&h->qtasks_d, sizeof(BTaskH) * qt.size()));
  HIPCHK(hipMalloc(&h->ptasks_d, sizeof(BTaskH) * pt.size()));
  HIPCHK(hipMalloc(&h->qoff_d, sizeof(int) * 9));
  HIPCHK(hipMemcpyAsync(h->qtasks_d, qt.data(), sizeof(BTaskH) * qt.size(), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->ptasks_d, pt.data(), sizeof(BTaskH) * pt.size(), hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->qoff_d, qoff.data(), sizeof(int) * 9, hipMemcpyHostToDevice, h->stream));
  FWCHK(fw::upload_twiddles(h->nmax, h->fp32, &h->tw));
  HIPCHK(hipMalloc(&h->times, sizeof(double) * kChunk));
  HIPCHK(hipMalloc(&h->partials, sizeof(double) * FW_NSER * std::max(part, 1)));
  HIPCHK(hipMemsetAsync(h->partials, 0, sizeof(double) * FW_NSER * std::max(part, 1), h->stream));
  HIPCHK(hipMalloc(&h->series, sizeof(double) * FW_NSER * ni * (kChunk + 1)));
  h->ncounters = kChunk * 8 + kChunk * ni * 3;
  HIPCHK(hipMalloc(&h->counters, sizeof(unsigned int) * h->ncounters));
  HIPCHK(hipMalloc(&h->fresh, sizeof(int) * ni));
  HIPCHK(hipMalloc(&h->err, sizeof(int)));
  HIPCHK(hipMalloc(&h->vkmin, sizeof(double) * 2 * ni));
  FWCHK(upload_instances(h));
  if (h->theta) {   // theta-sum weights of every instance (K, K' independent of t)
    HIPCHK(hipMalloc(&h->sw, (size_t)std::max(swn, 1) * 48));
    BArgsH a = base_args(h);
    void* sw = h->sw;
    void* params[] = {&a, &sw};
    HIPCHK(hipModuleLaunchKernel(h->f_sumw, ni, 1, 1, 256, 1, 1, 0, h->stream, params, nullptr));
  }
  FWCHK(zero_counters(h));
  FWCHK(fill_int(h, h->fresh, ni, 1));
  FWCHK(fill_int(h, h->err, 1, 0));

  // persistent grid: the occupancy limit (or blocks_per_cu) on every CU
  int ncu = 0, per_cu = 0;
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, d->device));
  HIPCHK(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, h->f_step, kBT, 0));
  if (d->blocks_per_cu > 0) per_cu = std::min(per_cu, d->blocks_per_cu);
  h->grid = std::max(8, ncu * std::max(per_cu, 1));
  h->grid -= h->grid % 8;
  HIPCHK(hipStreamSynchronize(h->stream));
  *out = guard.release();
  return FW_OK;
}

extern "C" int fw_batch_destroy(fw_batch* h) {
  if (!h) return FW_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  void* bufs[] = {h->W, h->W0s, h->G, h->W0t, h->sw, h->inst_d, h->qtasks_d, h->ptasks_d, h->qoff_d, h->tw, h->times, h->partials,
                  h->series, h->counters, h->fresh, h->err, h->vkmin};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (h->mod) (void)hipModuleUnload(h->mod);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return FW_OK;
}

extern "C" int fw_batch_set_state(fw_batch* h, int i, const double* host_w) {
  fw::clear_error();
  FWCHK(instance_check(h, i));
  if (!host_w) return fw::fail(FW_ERR_INVALID_ARG, "host_w is NULL");
  const BInstH& b = h->inst[i];
  const size_t n = (size_t)b.n * b.n;
  char* dst = (char*)h->W + (size_t)b.off * h->esz;
  if (h->fp32) {
    std::vector<float> tmp(host_w, host_w + n);
    HIPCHK(hipMemcpyAsync(dst, tmp.data(), n * 4, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  } else {
    HIPCHK(hipMemcpyAsync(dst, host_w, n * 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  return FW_OK;
}

extern "C" int fw_batch_get_state(fw_batch* h, int i, double* host_w) {
  fw::clear_error();
  FWCHK(instance_check(h, i));
  if (!host_w) return fw::fail(FW_ERR_INVALID_ARG, "host_w is NULL");
  const BInstH& b = h->inst[i];
  const size_t n = (size_t)b.n * b.n;
  const char* src = (const char*)h->W + (size_t)b.off * h->esz;
  if (h->fp32) {
    std::vector<float> tmp(n);
    HIPCHK(hipMemcpyAsync(tmp.data(), src, n * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (size_t j = 0; j < n; ++j) host_w[j] = tmp[j];
  } else {
    HIPCHK(hipMemcpyAsync(host_w, src, n * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  }
  return FW_OK;
}

static int ensure_w0s(fw_batch* h) {
  if (h->W0s) return FW_OK;
  HIPCHK(hipMalloc(&h->W0s, (size_t)h->points * h->esz));
  HIPCHK(hipMemsetAsync(h->W0s, 0, (size_t)h->points * h->esz, h->stream));
  if (h->two_pass) {
    HIPCHK(hipMalloc(&h->G, (size_t)h->points * h->esz));
    HIPCHK(hipMemsetAsync(h->G, 0, (size_t)h->points * h->esz, h->stream));
  }
  if (h->theta) {
    HIPCHK(hipMalloc(&h->W0t, (size_t)h->points * h->esz));
    HIPCHK(hipMemsetAsync(h->W0t, 0, (size_t)h->points * h->esz, h->stream));
  }
  if (h->gexec) {
    (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
  }
  return FW_OK;
}

// G = Vopen(W0s) of every dissipative instance, for the two-pass steps (V does
// not depend on t there, so G is computed once per source).
static int update_g(fw_batch* h) {
  if (!h->two_pass || !h->G) return FW_OK;
  const double t0 = 0.0;
  HIPCHK(hipMemcpyAsync(h->times, &t0, sizeof(double), hipMemcpyHostToDevice, h->stream));
  FWCHK(enqueue_pass(h, 0, 2, 0, 0));
  if (h->theta) FWCHK(enqueue_pass(h, 0, 3, 0, 0));   // W0t
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

extern "C" int fw_batch_set_dissipator(fw_batch* h, int i, const double* host_w0s, double w_coeff) {
  fw::clear_error();
  FWCHK(instance_check(h, i));
  BInstH& b = h->inst[i];
  if (!host_w0s) {
    b.dissip = 0;
    b.w_coeff = 1.0;
    return upload_instances(h);
  }
  if (!(w_coeff > 0.0) || w_coeff > 1.0) return fw::fail(FW_ERR_INVALID_ARG, "w_coeff must be in (0, 1]");
  FWCHK(ensure_w0s(h));
  const size_t n = (size_t)b.n * b.n;
  char* dst = (char*)h->W0s + (size_t)b.off * h->esz;
  if (h->fp32) {
    std::vector<float> tmp(host_w0s, host_w0s + n);
    HIPCHK(hipMemcpyAsync(dst, tmp.data(), n * 4, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
  } else {
    HIPCHK(hipMemcpyAsync(dst, host_w0s, n * 8, hipMemcpyHostToDevice, h->stream));
  }
  b.dissip = 1;
  b.w_coeff = w_coeff;
  FWCHK(upload_instances(h));
  return update_g(h);
}

extern "C" int fw_batch_gibbs_dissipators(fw_batch* h, const double* w_coeff) {
  fw::clear_error();
  if (!h || !w_coeff) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  for (int i = 0; i < h->d.n_instances; ++i)
    if (!(w_coeff[i] > 0.0) || w_coeff[i] > 1.0) return fw::fail(FW_ERR_INVALID_ARG, "w_coeff must be in (0, 1]");
  FWCHK(ensure_w0s(h));
  for (int i = 0; i < h->d.n_instances; ++i) {
    h->inst[i].w_coeff = w_coeff[i];
    h->inst[i].dissip = w_coeff[i] < 1.0 ? 1 : 0;
  }
  FWCHK(upload_instances(h));
  BArgsH a = base_args(h);
  void* w0s = h->W0s;
  void* params[] = {&a, &w0s};
  HIPCHK(hipModuleLaunchKernel(h->f_gibbs, h->d.n_instances * 64, 1, 1, kBT, 1, 1, 0, h->stream, params, nullptr));
  FWCHK(update_g(h));
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

extern "C" int fw_batch_bloch_propagate(fw_batch* h, int convention, double t_initial, const double* dbeta,
                                        const int* steps) {
  fw::clear_error();
  if (!h || !dbeta || !steps) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  if (convention != FW_BLOCH_CUDA && convention != FW_BLOCH_FFTW)
    return fw::fail(FW_ERR_INVALID_ARG, "convention must be FW_BLOCH_CUDA or FW_BLOCH_FFTW");
  int nmax = 0;
  for (int i = 0; i < h->d.n_instances; ++i) {
    if (steps[i] < 0 || !std::isfinite(dbeta[i])) return fw::fail(FW_ERR_INVALID_ARG, "bad dbeta or steps");
    h->inst[i].dbeta = dbeta[i];
    h->inst[i].nbloch = steps[i];
    nmax = std::max(nmax, steps[i]);
  }
  FWCHK(upload_instances(h));
  std::vector<double> t0(kChunk, t_initial);
  HIPCHK(hipMemcpyAsync(h->times, t0.data(), sizeof(double) * kChunk, hipMemcpyHostToDevice, h->stream));
  {
    BArgsH a = base_args(h);
    double* out = h->vkmin;
    void* params[] = {&a, &out};
    HIPCHK(hipModuleLaunchKernel(h->f_vkmin, h->d.n_instances, 1, 1, kBT, 1, 1, 0, h->stream, params, nullptr));
  }
  FWCHK(fill_int(h, h->fresh, h->d.n_instances, 1));
  FWCHK(fill_int(h, h->err, 1, 0));
  for (int g = 0; g < nmax; ++g) {
    const int s = g % kChunk;
    if (s == 0) FWCHK(zero_counters(h));
    BArgsH a = base_args(h);
    a.step = s;
    a.gstep = g;
    a.cv = convention == FW_BLOCH_CUDA ? 0.25 : 0.5;
    a.ck = convention == FW_BLOCH_CUDA ? 0.5 : 1.0;
    if (h->d.mode == 0) {
      FWCHK(launch(h, h->f_bloch, h->grid, a));
    } else {
      for (int p = 0; p < 3; ++p) {
        a.task_base = h->ptask_off[p];
        FWCHK(launch(h, h->f_bloch, h->ptask_off[p + 1] - h->ptask_off[p], a));
      }
    }
    if (s == kChunk - 1) FWCHK(check_err(h));
  }
  if (nmax > 0) FWCHK(final_norm(h));
  FWCHK(fill_int(h, h->fresh, h->d.n_instances, 1));
  return check_err(h);
}

extern "C" int fw_batch_propagate(fw_batch* h, const double* times, int steps, double* series) {
  fw::clear_error();
  if (!h || (!times && steps > 0)) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  if (steps <= 0) return FW_OK;
  const int ni = h->d.n_instances;
  FWCHK(fill_int(h, h->fresh, ni, 1));
  FWCHK(fill_int(h, h->err, 1, 0));
  if (steps >= kChunk) FWCHK(ensure_graph(h));
  std::vector<double> slots;
  if (series) slots.resize((size_t)kChunk * ni * FW_NSER);
  for (int off = 0; off < steps; off += kChunk) {
    const int n = std::min(kChunk, steps - off);
    HIPCHK(hipMemcpyAsync(h->times, times + off, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
    if (h->two_pass) {
      // VF, then (K, VV) per step; the run's last step is K, VL
      if (off == 0) FWCHK(enqueue_pass(h, 0, 0, 0, 0));
      const bool last = off + n == steps;
      if (n == kChunk && !last) {
        HIPCHK(hipGraphLaunch(h->gexec, h->stream));
      } else {
        for (int s = 0; s < n - (last ? 1 : 0); ++s) FWCHK(enqueue_step2(h, s, off + s));
        if (last) {
          FWCHK(enqueue_pass(h, 1, 1, n - 1, off + n - 1));
          FWCHK(enqueue_pass(h, 2, 0, n - 1, off + n - 1));
        }
      }
    } else if (n == kChunk) {
      HIPCHK(hipGraphLaunch(h->gexec, h->stream));
    } else {
      FWCHK(zero_counters(h));
      for (int s = 0; s < n; ++s) FWCHK(enqueue_step(h, 0, s, off + s));
    }
    if (series) {
      HIPCHK(hipMemcpyAsync(slots.data(), h->series, sizeof(double) * (size_t)n * ni * FW_NSER,
                            hipMemcpyDeviceToHost, h->stream));
    }
    FWCHK(check_err(h));
    if (series)
      for (int s = 0; s < n; ++s) {
        const int g = off + s - 1;
        if (g >= 0)
          std::memcpy(series + (size_t)g * ni * FW_NSER, &slots[(size_t)s * ni * FW_NSER],
                      sizeof(double) * ni * FW_NSER);
      }
  }
  FWCHK(final_norm(h));
  if (series)
    HIPCHK(hipMemcpyAsync(series + (size_t)(steps - 1) * ni * FW_NSER, h->series + (size_t)kChunk * ni * FW_NSER,
                          sizeof(double) * ni * FW_NSER, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return FW_OK;
}

extern "C" int fw_batch_state_ptr(fw_batch* h, int i, void** dev_ptr, int* elem_size) {
  fw::clear_error();
  FWCHK(instance_check(h, i));
  if (!dev_ptr) return fw::fail(FW_ERR_INVALID_ARG, "dev_ptr is NULL");
  *dev_ptr = (char*)h->W + (size_t)h->inst[i].off * h->esz;
  if (elem_size) *elem_size = (int)h->esz;
  return FW_OK;
}

extern "C" int fw_batch_copy_states(fw_batch* dst, fw_batch* src) {
  fw::clear_error();
  if (!dst || !src) return fw::fail(FW_ERR_INVALID_ARG, "NULL handle");
  if (dst->inst.size() != src->inst.size() || dst->points != src->points || dst->esz != src->esz)
    return fw::fail(FW_ERR_INVALID_ARG, "the batches differ in instances or precision");
  for (size_t i = 0; i < dst->inst.size(); ++i)
    if (dst->inst[i].n != src->inst[i].n || dst->inst[i].off != src->inst[i].off)
      return fw::fail(FW_ERR_INVALID_ARG, "the batches differ in instance grids");
  if (dst == src) return FW_OK;
  HIPCHK(hipStreamSynchronize(src->stream));
  HIPCHK(hipMemcpyAsync(dst->W, src->W, (size_t)dst->points * dst->esz, hipMemcpyDeviceToDevice, dst->stream));
  HIPCHK(hipStreamSynchronize(dst->stream));
  return FW_OK;
}

extern "C" int fw_batch_passes(fw_batch* h, int* passes_per_step) {
  fw::clear_error();
  if (!h || !passes_per_step) return fw::fail(FW_ERR_INVALID_ARG, "NULL argument");
  *passes_per_step = h->d.mode == 0 ? 1 : (h->two_pass ? 2 : 3);
  return FW_OK;
}

extern "C" int fw_batch_info(fw_batch* h, long long* points, int* tasks_per_step, int* grid) {
  if (!h) return fw::fail(FW_ERR_INVALID_ARG, "NULL handle");
  if (points) *points = h->points;
  if (tasks_per_step) *tasks_per_step = h->ntasks;
  if (grid) *grid = h->d.mode == 0 ? h->grid : h->ptask_off[3];
  return FW_OK;
}

extern "C" int fw_batch_step_time(fw_batch* h, int steps, double* us_per_step) {
  fw::clear_error();
  if (!h || !us_per_step || steps < kChunk)
    return fw::fail(FW_ERR_INVALID_ARG, "NULL argument or fewer than 64 steps");
  const int chunks = steps / kChunk;
  FWCHK(ensure_graph(h));
  std::vector<double> t(kChunk);
  for (int s = 0; s < kChunk; ++s) t[s] = (s + 1) * h->d.dt;
  HIPCHK(hipMemcpyAsync(h->times, t.data(), sizeof(double) * kChunk, hipMemcpyHostToDevice, h->stream));
  FWCHK(fill_int(h, h->fresh, h->d.n_instances, 1));
  FWCHK(fill_int(h, h->err, 1, 0));
  if (h->two_pass) FWCHK(enqueue_pass(h, 0, 0, 0, 0));   // VF (not timed)
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, h->stream));
  for (int c = 0; c < chunks; ++c) HIPCHK(hipGraphLaunch(h->gexec, h->stream));
  HIPCHK(hipEventRecord(e1, h->stream));
  HIPCHK(hipEventSynchronize(e1));
  if (h->two_pass) {   // close the run: K, VL (not timed)
    FWCHK(enqueue_pass(h, 1, 1, kChunk - 1, kChunk - 1));
    FWCHK(enqueue_pass(h, 2, 0, kChunk - 1, kChunk - 1));
  }
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *us_per_step = ms * 1000.0 / (chunks * kChunk);
  FWCHK(final_norm(h));
  return check_err(h);
}
