// Shared definitions for the lucivant-sim MI355X (gfx950) kernels and their C-ABI launchers.
//
// Layout of every state buffer: `batch` states of 2^nbits amplitudes each, back to back, amplitude
// index bit q = qubit q (little-endian, as the reference engines flatten their tensors). A density
// matrix of n qubits is a "state" of 2n bits: flat index = row * 2^n + col, so vector bit n+q is
// the row bit of qubit q and vector bit q its column bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lsim.h"

namespace lsim {

// ---------------------------------------------------------------------------- error reporting

void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);
int check_launch(const char* what);

#define LSIM_REQUIRE(cond, ...)                             \
  do {                                                      \
    if (!(cond)) return ::lsim::fail(LSIM_EINVAL, __VA_ARGS__); \
  } while (0)

// ------------------------------------------------------------------------------ value types

template <typename R>
struct Cx;
template <>
struct Cx<double> {
  using V = double2;
};
template <>
struct Cx<float> {
  using V = float2;
};

constexpr int kMaxK = 5;              // widest fused dense operator (32 x 32)
constexpr int kMaxBits = 30;          // widest vector handled (2^30 amplitudes per state)
constexpr int kBlock = 256;           // threads per workgroup for the streaming kernels
constexpr int kMaxGridY = 65535;      // batch entries per launch (grid y)

// Targets of a k-qubit operator: `offs[c]` is the amplitude offset of matrix index c (the first
// listed target is the matrix's most significant bit), `sorted` the target bits ascending, used
// to spread a group number over the non-target bits.
struct Targets {
  uint32_t sorted[kMaxK];
  uint32_t listed[kMaxK];
  uint32_t offs[1 << kMaxK];
};

int make_targets(int nbits, int k, const int* targets, Targets* out);

// LDS-tiled placement of a k-qubit operator: a tile is 2^(k+6) amplitudes, the `b` lowest
// index bits (contiguous, coalesced) times every combination of the `h` targets at or above b.
// Local tile index bits: [0, b) are the global bits [0, b); b + j is high[j].
struct Tiling {
  uint32_t b, h;
  uint32_t high[kMaxK];          // targets >= b, ascending
  uint32_t seg[1 << kMaxK];      // global offset of tile segment s
  uint32_t lsorted[kMaxK];       // targets in local bits, ascending
  uint32_t loffs[1 << kMaxK];    // local offset of matrix index c
};

// True when the operator's low targets make lanes stride (the tiled kernel pays off); fills t.
bool make_tiling(int nbits, int k, const int* targets, Tiling* t);

__device__ __forceinline__ uint32_t spread(uint32_t g, const uint32_t* sorted, int k) {
#pragma unroll
  for (int j = 0; j < kMaxK; ++j) {
    if (j < k) {
      const uint32_t s = sorted[j];
      g = ((g >> s) << (s + 1)) | (g & ((1u << s) - 1u));
    }
  }
  return g;
}

// A 16-byte load or store, with or without the streaming (non-temporal) hint.
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));

template <bool NT>
__device__ __forceinline__ double2 load_amp(const double2* p) {
  d2v v = NT ? __builtin_nontemporal_load(reinterpret_cast<const d2v*>(p)) : *reinterpret_cast<const d2v*>(p);
  return make_double2(v.x, v.y);
}
template <bool NT>
__device__ __forceinline__ float2 load_amp(const float2* p) {
  f2v v = NT ? __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p)) : *reinterpret_cast<const f2v*>(p);
  return make_float2(v.x, v.y);
}
template <bool NT>
__device__ __forceinline__ void store_amp(double2* p, double2 a) {
  d2v v = {a.x, a.y};
  if (NT)
    __builtin_nontemporal_store(v, reinterpret_cast<d2v*>(p));
  else
    *reinterpret_cast<d2v*>(p) = v;
}
template <bool NT>
__device__ __forceinline__ void store_amp(float2* p, float2 a) {
  f2v v = {a.x, a.y};
  if (NT)
    __builtin_nontemporal_store(v, reinterpret_cast<f2v*>(p));
  else
    *reinterpret_cast<f2v*>(p) = v;
}

template <typename V>
__device__ __forceinline__ V cmul(V a, V b) {
  V r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

// Wavefront (64 lanes) then workgroup sum of a double; every thread returns the total.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Whether a launch over `bytes` of states uses streaming (non-temporal) accesses, by the
// process-wide switch lsim_set_dense_streaming.
bool dense_streams(uint64_t bytes);

inline size_t dtype_bytes(int dtype) { return dtype == LSIM_C64 ? 8 : 16; }

}  // namespace lsim
