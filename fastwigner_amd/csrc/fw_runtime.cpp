// fw_runtime.cpp -- host side of libfastwigner.so (C ABI in include/fastwigner.h).
//
// Replaces the two native crossings of the reference: runtime compilation of
// kernels with spliced V/K/mask strings (pycuda SourceModule,
// wigner_moyal_cuda_1d.py:244-277) and the cuFFT ctypes FFI (cufft.py). Each
// propagator handle owns its hipRTC-compiled module, device buffers, one HIP
// stream (default priority) and a captured hipGraph of a chunk of steps.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>
#include <sys/stat.h>
#include <unistd.h>

#include "fastwigner.h"
#include "fw_internal.h"

extern const char fw_common_h_src[];
extern const char fw_device_h_src[];
extern const char fw_rho_h_src[] __attribute__((weak));
extern const char fw_batch_h_src[] __attribute__((weak));
extern const char fw_fno_h_src[] __attribute__((weak));
extern const char fw_math_h_src[];

namespace fw {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
void clear_error() { g_err.clear(); }
const char* last_error() { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// hipRTC compilation with an in-process and an on-disk code-object cache
// ---------------------------------------------------------------------------
static std::mutex g_cache_mu;
static std::unordered_map<std::string, std::vector<char>> g_code_cache;

static std::string hex128(const std::string& s) {
  // two independent 64-bit FNV-1a hashes
  uint64_t h1 = 1469598103934665603ull, h2 = 1099511628211ull ^ 0x9e3779b97f4a7c15ull;
  for (unsigned char c : s) {
    h1 = (h1 ^ c) * 1099511628211ull;
    h2 = (h2 ^ (c + 0x5bu)) * 0x100000001b3ull + 0x7f4a7c15ull;
  }
  char buf[40];
  snprintf(buf, sizeof buf, "%016llx%016llx", (unsigned long long)h1, (unsigned long long)h2);
  return buf;
}

static std::string cache_dir() {
  const char* e = getenv("FASTWIGNER_CACHE_DIR");
  if (e && *e) return e;
  const char* h = getenv("HOME");
  if (h && *h) return std::string(h) + "/.cache/fastwigner_amd";
  return "";
}

static void mkdirs(const std::string& p) {
  std::string cur;
  for (size_t i = 0; i < p.size(); ++i) {
    cur += p[i];
    if (p[i] == '/' || i + 1 == p.size()) mkdir(cur.c_str(), 0755);
  }
}

std::string target_arch(int device) {
  if (device < 0) return "gfx950";
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return "gfx950";
  std::string a = prop.gcnArchName;
  size_t c = a.find(':');
  if (c != std::string::npos) a = a.substr(0, c);
  return a.empty() ? "gfx950" : a;
}

int compile(const std::string& src, const std::string& arch, std::vector<char>& code) {
  std::vector<std::string> opts = {"--offload-arch=" + arch, "-O3", "-std=c++17", "-Xclang", "-target-feature",
                                   "-Xclang", "-packed-fp32-ops", "-Wno-unused-variable",
                                   "-Wno-unused-but-set-variable"};
  // the embedded headers hipRTC may include
  std::vector<const char*> hdr_names, hdrs;
  const char* names_all[] = {"fw_common.h", "fw_device.h", "fw_rho.h", "fw_batch.h", "fw_fno.h", "fw_math.h"};
  const char* srcs_all[] = {fw_common_h_src, fw_device_h_src, &fw_rho_h_src ? fw_rho_h_src : nullptr,
                            &fw_batch_h_src ? fw_batch_h_src : nullptr, &fw_fno_h_src ? fw_fno_h_src : nullptr,
                            fw_math_h_src};
  std::string key_src = src + "\n//" + arch;
  for (int i = 0; i < 6; ++i)
    if (srcs_all[i]) {
      hdr_names.push_back(names_all[i]);
      hdrs.push_back(srcs_all[i]);
      key_src += std::string("\n//") + srcs_all[i];
    }
  for (auto& o : opts) key_src += "\n//" + o;
  const std::string key = hex128(key_src);
  {
    std::lock_guard<std::mutex> g(g_cache_mu);
    auto it = g_code_cache.find(key);
    if (it != g_code_cache.end()) {
      code = it->second;
      return FW_OK;
    }
  }
  const std::string dir = cache_dir();
  const std::string path = dir.empty() ? "" : dir + "/" + key + ".co";
  if (!path.empty()) {
    std::ifstream f(path, std::ios::binary);
    if (f) {
      std::vector<char> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
      if (!buf.empty()) {
        code = buf;
        std::lock_guard<std::mutex> g(g_cache_mu);
        g_code_cache[key] = buf;
        return FW_OK;
      }
    }
  }
  if (const char* dd = getenv("FASTWIGNER_DUMP_DIR")) {
    // debugging aid: the generated source, compilable with hipcc -I csrc
    mkdirs(dd);
    std::ofstream f(std::string(dd) + "/" + key + ".hip");
    f << src;
  }
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "fastwigner_kernels.hip", (int)hdrs.size(), hdrs.data(),
                          hdr_names.data()) != HIPRTC_SUCCESS)
    return fail(FW_ERR_COMPILE, "hiprtcCreateProgram failed");
  std::vector<const char*> copts;
  for (auto& o : opts) copts.push_back(o.c_str());
  hiprtcResult r = hiprtcCompileProgram(prog, (int)copts.size(), copts.data());
  if (r != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls + 1, '\0');
    hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    return fail(FW_ERR_COMPILE, "hipRTC compilation failed:\n" + log);
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  code.resize(cs);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  {
    std::lock_guard<std::mutex> g(g_cache_mu);
    g_code_cache[key] = code;
  }
  if (!path.empty()) {
    mkdirs(dir);
    const std::string tmp = path + ".tmp" + std::to_string(getpid());
    std::ofstream f(tmp, std::ios::binary);
    if (f) {
      f.write(code.data(), (std::streamsize)code.size());
      f.close();
      rename(tmp.c_str(), path.c_str());
    }
  }
  return FW_OK;
}

std::string fmt_double(double v) {
  char buf[64];
  snprintf(buf, sizeof buf, "%.17g", v);
  std::string s = buf;
  if (s.find_first_of(".eEn") == std::string::npos) s += ".0";  // keep it a double literal
  return s;
}

// exp(-2 pi i j / n), j < n, evaluated in long double and rounded once
void twiddles(int n, std::vector<double>& re, std::vector<double>& im) {
  re.resize(n);
  im.resize(n);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int j = 0; j < n; ++j) {
    const long double a = -2.0L * pi * (long double)j / (long double)n;
    re[j] = (double)cosl(a);
    im[j] = (double)sinl(a);
  }
  // exact values on the axes
  for (int j = 0; j < n; ++j) {
    if ((4 * j) % n == 0) {
      const int q = (4 * j) / n;
      re[j] = (q == 0) ? 1.0 : (q == 2 ? -1.0 : 0.0);
      im[j] = (q == 1) ? -1.0 : (q == 3 ? 1.0 : 0.0);
    }
  }
}

int upload_twiddles(int n, bool fp32, void** dptr) {
  std::vector<double> re, im;
  twiddles(n, re, im);
  if (fp32) {
    // hi + lo float pairs: (hi.x, hi.y, lo.x, lo.y) per entry
    std::vector<float> t(4 * n);
    for (int j = 0; j < n; ++j) {
      const float hr = (float)re[j], hi = (float)im[j];
      t[4 * j] = hr;
      t[4 * j + 1] = hi;
      t[4 * j + 2] = (float)(re[j] - (double)hr);
      t[4 * j + 3] = (float)(im[j] - (double)hi);
    }
    HIPCHK(hipMalloc(dptr, t.size() * sizeof(float)));
    HIPCHK(hipMemcpy(*dptr, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
  } else {
    std::vector<double> t(2 * n);
    for (int j = 0; j < n; ++j) { t[2 * j] = re[j]; t[2 * j + 1] = im[j]; }
    HIPCHK(hipMalloc(dptr, t.size() * sizeof(double)));
    HIPCHK(hipMemcpy(*dptr, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  return FW_OK;
}

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// the device's radix rule (fw_common.h next_radix)
static int next_radix(int rem, int e) {
  if (is_pow2(rem) && is_pow2(e)) {
    for (int r = 16; r >= 2; r /= 2)
      if (rem % r == 0 && e % r == 0) return r;
  }
  for (int r = e < 32 ? e : 32; r >= 2; --r)
    if (rem % r == 0 && e % r == 0 && (r & -r) <= 16) return r;
  return 0;
}

bool line_schedule_ok(int n, int e) {
  if (n < 2 || e < 2 || e > 32 || e > n || n % e) return false;
  if (is_pow2(n)) return is_pow2(e);
  if (e % 2) return false;
  for (int rem = n; rem > 1;) {
    const int r = next_radix(rem, e);
    if (r < 2) return false;
    rem /= r;
  }
  return true;
}

int line_default_e(int n, int pow2_default) {
  if (is_pow2(n)) return pow2_default;
  for (int e = 8; e <= 32; ++e)
    if (line_schedule_ok(n, e)) return e;
  for (int e = 6; e >= 2; --e)
    if (line_schedule_ok(n, e)) return e;
  return 0;
}

}  // namespace fw

// ---------------------------------------------------------------------------
// C ABI: library
// ---------------------------------------------------------------------------
extern "C" int fw_version(void) { return 100; }
extern "C" const char* fw_last_error(void) { return fw::last_error(); }
extern "C" int fw_device_count(int* count) {
  if (!count) return fw::fail(FW_ERR_INVALID_ARG, "count is NULL");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return FW_OK;
}
