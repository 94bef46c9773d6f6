// Error reporting, device facts and target-offset tables shared by every launcher.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "lsim_internal.h"

namespace lsim {

static thread_local char g_error[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof g_error, fmt, ap);
  va_end(ap);
}

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof g_error, fmt, ap);
  va_end(ap);
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(LSIM_EHIP, "%s: %s", what, hipGetErrorString(e));
  return LSIM_OK;
}

int make_targets(int nbits, int k, const int* targets, Targets* out) {
  LSIM_REQUIRE(k >= 1 && k <= kMaxK, "k must be between 1 and %d, not %d", kMaxK, k);
  LSIM_REQUIRE(k <= nbits, "a %d-bit operator does not fit a %d-bit state", k, nbits);
  LSIM_REQUIRE(targets != nullptr, "targets is NULL");
  uint32_t mask = 0;
  for (int i = 0; i < k; ++i) {
    LSIM_REQUIRE(targets[i] >= 0 && targets[i] < nbits, "target bit %d is outside 0..%d",
                 targets[i], nbits - 1);
    LSIM_REQUIRE(!(mask & (1u << targets[i])), "target bit %d is given twice", targets[i]);
    mask |= 1u << targets[i];
  }
  int j = 0;
  for (int b = 0; b < nbits; ++b)
    if (mask & (1u << b)) out->sorted[j++] = (uint32_t)b;
  for (; j < kMaxK; ++j) out->sorted[j] = 0;
  for (int i = 0; i < kMaxK; ++i) out->listed[i] = i < k ? (uint32_t)targets[i] : 0u;
  for (int c = 0; c < (1 << kMaxK); ++c) {
    uint32_t off = 0;
    if (c < (1 << k))
      for (int i = 0; i < k; ++i)
        if ((c >> (k - 1 - i)) & 1) off |= 1u << targets[i];
    out->offs[c] = off;
  }
  return LSIM_OK;
}

bool make_tiling(int nbits, int k, const int* targets, Tiling* t) {
  uint32_t sorted[kMaxK];
  for (int i = 0; i < k; ++i) sorted[i] = (uint32_t)targets[i];
  for (int i = 1; i < k; ++i)
    for (int j = i; j > 0 && sorted[j - 1] > sorted[j]; --j) {
      const uint32_t x = sorted[j];
      sorted[j] = sorted[j - 1];
      sorted[j - 1] = x;
    }
  // low targets: the longest prefix with sorted[j] < 6 + (j + 1); then b = 6 + (#low)
  int low = 0;
  while (low < k && sorted[low] < (uint32_t)(6 + low + 1)) ++low;
  // With targets listed from the highest bit down (the planner's order), the direct kernel
  // fetches the amplitudes that share a 128 B line with consecutive instructions and stays near
  // copy speed unless three or more targets crowd the lowest 32 amplitudes (512 B), where every
  // lane of an instruction hits its own line: tile those 3- and 4-qubit operators (measured on
  // gfx950; a 5-qubit tile does not fit the register budget).
  int below5 = 0;
  for (int i = 0; i < k; ++i) below5 += sorted[i] < 5;
  if (k < 3 || k > 4 || below5 < 3) return false;
  if ((int)(k + 6) > nbits) return false;
  t->b = 6 + low;
  t->h = k - low;
  for (int j = 0; j < kMaxK; ++j) t->high[j] = j < (int)t->h ? sorted[low + j] : 0;
  for (int s = 0; s < (1 << kMaxK); ++s) {
    uint32_t off = 0;
    if (s < (1 << t->h))
      for (int j = 0; j < (int)t->h; ++j)
        if ((s >> j) & 1) off |= 1u << t->high[j];
    t->seg[s] = off;
  }
  auto local = [&](uint32_t g) -> uint32_t {
    if (g < t->b) return g;
    for (int j = 0; j < (int)t->h; ++j)
      if (t->high[j] == g) return t->b + j;
    return 0;
  };
  for (int i = 0; i < kMaxK; ++i) t->lsorted[i] = i < k ? local(sorted[i]) : 0;
  for (int c = 0; c < (1 << kMaxK); ++c) {
    uint32_t off = 0;
    if (c < (1 << k))
      for (int i = 0; i < k; ++i)
        if ((c >> (k - 1 - i)) & 1) off |= 1u << local((uint32_t)targets[i]);
    t->loffs[c] = off;
  }
  return true;
}

}  // namespace lsim

extern "C" {

const char* lsim_last_error(void) { return lsim::g_error; }

int lsim_abi_version(void) { return 3; }

// A plain sequential loop: the same float64 additions in the same order as NumPy's accumulate
// (out[0] = in[0], then one addition per element), which is all bit-exact sampling needs, at
// about twice NumPy's speed (its accumulate loop carries per-element overhead that the
// dependent-add chain doesn't need). No reassociation: the build has no fast-math flags.
int lsim_host_cumsum(const double* in, double* out, size_t n) {
  if (n == 0) return LSIM_OK;
  LSIM_REQUIRE(in != nullptr && out != nullptr, "the input or output pointer is NULL");
  double c = in[0];
  out[0] = c;
  for (size_t i = 1; i < n; ++i) {
    c += in[i];
    out[i] = c;
  }
  return LSIM_OK;
}

int lsim_device_info(int* cu_count, long long* hbm_bytes, char* arch, int arch_len) {
  hipDeviceProp_t p;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipGetDeviceProperties(&p, dev);
  if (e != hipSuccess) return lsim::fail(LSIM_EHIP, "device info: %s", hipGetErrorString(e));
  if (cu_count) *cu_count = p.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = (long long)p.totalGlobalMem;
  if (arch && arch_len > 0) {
    strncpy(arch, p.gcnArchName, (size_t)arch_len - 1);
    arch[arch_len - 1] = 0;
  }
  return LSIM_OK;
}

}  // extern "C"
