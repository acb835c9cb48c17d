// fw_internal.h -- shared host helpers of libfastwigner.so (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace fw {

int fail(int code, const std::string& msg);
void clear_error();
const char* last_error();

// hipRTC-compile `src` (which may #include "fw_device.h") for `arch`; cached in
// process and on disk (FASTWIGNER_CACHE_DIR, default ~/.cache/fastwigner_amd).
int compile(const std::string& src, const std::string& arch, std::vector<char>& code);
std::string target_arch(int device);  // device < 0: gfx950
std::string fmt_double(double v);     // round-trip exact C double literal
void twiddles(int n, std::vector<double>& re, std::vector<double>& im);
int upload_twiddles(int n, bool fp32, void** dptr);
bool is_pow2(int n);
// Line FFT schedules (fw_common.h Line): E elements per thread over a line of
// n points is valid when E divides n, E <= 32 (even unless n is a power of two,
// then a power of two) and at every stage some radix up to min(E, 32) divides
// both E and the remaining length (the device's next_radix rule).
bool line_schedule_ok(int n, int e);
// The default E of a line of n points: the power-of-two rule of the propagator
// for powers of two; otherwise the smallest valid E >= 8, else the largest
// valid one; 0 when none exists (n = 2^a m needs E >= 2 rad(m) <= 32: odd
// primes of m whose product is at most 15).
int line_default_e(int n, int pow2_default);

// Host mirror of fwk::Args (fw_device.h); field order and types must match.
struct KArgs {
  void* W;
  const void* W0s;
  const void* tw_col;
  const void* tw_row;
  const double* times;
  double* scal;
  double* partials;
  unsigned int* counter;
  double* series;
  const void* vtab;
  const void* ktab;
  const void* sw;
  int step;
  int flags;
  int nbp;
  int pad_;
  double dbeta, cv, ck, vmin, kmin, t_initial;
  double w_coeff;
  double guard;
};

enum { SC_NORM = 0, SC_STOP = 1, SC_FRESH = 2, SC_OUT = 8, NSCAL = 32 };
enum { F_FIRST = 1, F_LAST = 2, F_BLOCH = 4, F_DISSIP = 8, F_MASK_XP = 16, F_BLOCH_MASK = 32, F_VTAB = 64,
       F_KTAB = 128, F_FROM_PARTIALS = 256, F_KNORM = 512, F_NO_SCALE = 1024, F_KGEN = 2048 };

}  // namespace fw

#define HIPCHK(x)                                                                         \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess)                                                                 \
      return ::fw::fail(FW_ERR_HIP, std::string(#x) + " failed: " + hipGetErrorString(e_)); \
  } while (0)

#define FWCHK(x)               \
  do {                         \
    int r_ = (x);              \
    if (r_ != FW_OK) return r_; \
  } while (0)
