// Permutations with phases: a k-qubit operator with one non-zero entry per row and column (x, y,
// cx, swap, ccx, and products of them with diagonal phases), the reference's slice moves
// (engines.py:91-104). Thread g owns the 2^k amplitudes of group g, as in the dense kernel, loads
// them and stores each one at its destination row times that row's phase: one read and one write
// of every amplitude, with no arithmetic beyond a complex multiply per amplitude (none for the
// rows whose phase is 1). Registers are indexed only by compile-time sources; the destination of
// each source is a wavefront-uniform kernel argument, so the stores take runtime addresses
// rather than runtime register indices.
#include "lsim_internal.h"

namespace lsim {

struct PermMap {
  uint32_t dst[1 << kMaxK];  // destination row of each source column
  uint32_t unit;             // bit r set: phase[r] is exactly 1 (the row is a plain move)
};

template <int K, typename R>
__global__ void __launch_bounds__(kBlock)
    k_perm(typename Cx<R>::V* __restrict__ states, uint32_t nbits, uint32_t groups,
           const typename Cx<R>::V* __restrict__ phases, const int* __restrict__ sel,
           uint32_t batch_base, Targets t, PermMap p) {
  using V = typename Cx<R>::V;
  constexpr int D = 1 << K;
  const uint32_t b = batch_base + blockIdx.y;
  const V* __restrict__ ph = phases;
  if (sel != nullptr) {
    const int m = sel[b];
    if (m < 0) return;
    ph += (size_t)m * D;
  }
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= groups) return;
  V* __restrict__ st = states + ((size_t)b << nbits);
  const uint32_t base = spread(g, t.sorted, K);
  V in[D];
#pragma unroll
  for (int c = 0; c < D; ++c) in[c] = st[base + t.offs[c]];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const uint32_t r = p.dst[c];
    st[base + t.offs[r]] = ((p.unit >> r) & 1u) ? in[c] : cmul(ph[r], in[c]);
  }
}

template <typename R>
static void launch_perm(void* states, int nbits, long long batch, int k, const Targets& t,
                        const PermMap& p, const void* phases, const int* sel, hipStream_t s) {
  using V = typename Cx<R>::V;
  const uint32_t groups = (uint32_t)(1ull << (nbits - k));
  for (long long b0 = 0; b0 < batch; b0 += kMaxGridY) {
    const uint32_t gy = (uint32_t)(batch - b0 < kMaxGridY ? batch - b0 : kMaxGridY);
    const dim3 grid((groups + kBlock - 1) / kBlock, gy);
    V* st = (V*)states;
    const V* ph = (const V*)phases;
    const uint32_t bb = (uint32_t)b0;
    switch (k) {
      case 1: k_perm<1, R><<<grid, kBlock, 0, s>>>(st, nbits, groups, ph, sel, bb, t, p); break;
      case 2: k_perm<2, R><<<grid, kBlock, 0, s>>>(st, nbits, groups, ph, sel, bb, t, p); break;
      case 3: k_perm<3, R><<<grid, kBlock, 0, s>>>(st, nbits, groups, ph, sel, bb, t, p); break;
      case 4: k_perm<4, R><<<grid, kBlock, 0, s>>>(st, nbits, groups, ph, sel, bb, t, p); break;
      default: k_perm<5, R><<<grid, kBlock, 0, s>>>(st, nbits, groups, ph, sel, bb, t, p); break;
    }
  }
}

}  // namespace lsim

using namespace lsim;

extern "C" {

int lsim_sv_apply_perm(void* states, int dtype, int nbits, long long batch, int k,
                       const int* targets, const int* perm, unsigned int unit_rows,
                       const void* phases, const int* sel, void* stream) {
  LSIM_REQUIRE(states != nullptr, "the state pointer is NULL");
  LSIM_REQUIRE(dtype == LSIM_C128 || dtype == LSIM_C64, "dtype must be LSIM_C128 or LSIM_C64");
  LSIM_REQUIRE(nbits >= 0 && nbits <= kMaxBits, "nbits must be between 0 and %d, not %d", kMaxBits, nbits);
  LSIM_REQUIRE(batch >= 1 && batch <= (1ll << 31), "batch must be between 1 and 2^31");
  LSIM_REQUIRE(phases != nullptr, "the phase pointer is NULL");
  LSIM_REQUIRE(perm != nullptr, "the permutation is NULL");
  Targets t;
  if (int rc = make_targets(nbits, k, targets, &t)) return rc;
  const int D = 1 << k;
  PermMap p = {};
  uint32_t seen = 0;
  for (int r = 0; r < D; ++r) {
    LSIM_REQUIRE(perm[r] >= 0 && perm[r] < D, "perm[%d] = %d is outside 0..%d", r, perm[r], D - 1);
    LSIM_REQUIRE(!((seen >> perm[r]) & 1u), "perm repeats source %d", perm[r]);
    seen |= 1u << perm[r];
    p.dst[perm[r]] = (uint32_t)r;
  }
  p.unit = unit_rows & (D == 32 ? 0xffffffffu : (1u << D) - 1u);
  if (dtype == LSIM_C128)
    launch_perm<double>(states, nbits, batch, k, t, p, phases, sel, as_stream(stream));
  else
    launch_perm<float>(states, nbits, batch, k, t, p, phases, sel, as_stream(stream));
  return check_launch("permutation");
}

}  // extern "C"
