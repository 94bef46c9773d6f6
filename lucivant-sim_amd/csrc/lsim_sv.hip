// State-update kernels for gfx950: fused dense k-qubit operators, diagonal operators, global
// scaling, basis-state initialisation and |psi|^2. All are HBM-streaming kernels: one read and
// one write of every amplitude they touch.
//
// Dense operator (the replacement of engines.py:38-119): thread g owns the 2^k amplitudes whose
// index has the non-target bits of g; it loads them into registers, forms out[r] = sum_c
// U[r][c] in[c] and stores in place. U is addressed by a wavefront-uniform index, so the
// compiler reads it with scalar loads into SGPRs and every FMA takes its matrix operand from
// there: no LDS or vector-register traffic for the operator. Lanes map to consecutive
// non-target indices, so for targets at bit 6 and above each load instruction of a wavefront
// reads 64 consecutive amplitudes; lower targets make the lanes stride, and the partner loads of
// the same wavefront complete the cache lines.
#include "lsim_internal.h"

namespace lsim {

// rows of a 5-qubit operator the dense kernel reads through the scalar cache (the rest from LDS)
static int g_k5_scalar_rows = 0;

// SR: for K = 5, how many leading rows of U are read through the scalar cache; the other rows
// are staged in LDS (see below). Ignored for K < 5 (every row through the scalar cache).
template <int K, typename R, int SR = 1 << K>
__global__ void __launch_bounds__(kBlock)
    k_dense(typename Cx<R>::V* __restrict__ states, uint32_t nbits, uint32_t groups,
            const typename Cx<R>::V* __restrict__ mats, const int* __restrict__ sel,
            uint32_t batch_base, Targets t) {
  using V = typename Cx<R>::V;
  constexpr int D = 1 << K;
  // rows of a 5-qubit operator read from LDS
  constexpr int LR = K == 5 ? D - SR : 0;
  const uint32_t b = batch_base + blockIdx.y;
  const V* __restrict__ U = mats;
  if (sel != nullptr) {
    const int m = sel[b];
    if (m < 0) return;
    U += (size_t)m * (D * D);
  }
  // A 5-qubit operator is 16 KiB, as large as the scalar cache. Rows SR.. are staged once per
  // workgroup in LDS and read as wavefront-uniform broadcasts; rows ..SR-1 stream through the
  // scalar cache. The split divides the operator reads between the two paths: an LDS broadcast
  // of 16 B still returns 1 KiB per wavefront, so all rows from LDS make LDS the bottleneck,
  // and all rows through the scalar cache thrash it.
  __shared__ V ulds[LR > 0 ? LR * D : 1];
  if constexpr (LR > 0) {
    for (int e = threadIdx.x; e < LR * D; e += kBlock) ulds[e] = U[SR * D + e];
    __syncthreads();
  }
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= groups) return;
  V* __restrict__ st = states + ((size_t)b << nbits);
  const uint32_t base = spread(g, t.sorted, K);
  V in[D];
#pragma unroll
  for (int c = 0; c < D; ++c) in[c] = st[base + t.offs[c]];
#pragma unroll
  for (int r = 0; r < D; ++r) {
    R re = 0, im = 0;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      const V u = (LR > 0 && r >= SR) ? ulds[(r - SR) * D + c] : U[r * D + c];
      re = fma(u.x, in[c].x, re);
      re = fma(-u.y, in[c].y, re);
      im = fma(u.x, in[c].y, im);
      im = fma(u.y, in[c].x, im);
    }
    V o;
    o.x = re;
    o.y = im;
    st[base + t.offs[r]] = o;
  }
}

// LDS-tiled dense operator for placements with low target bits. One wavefront owns a tile of
// 2^(K+6) amplitudes: 2^h segments of 2^b contiguous amplitudes (b low bits, h high targets).
// Segments are loaded with consecutive lanes on consecutive amplitudes (full 128 B lines), kept
// in LDS under an XOR swizzle of the 16-byte bank group, each lane then applies U to one group
// of 2^K amplitudes, and the tile is stored back the same coalesced way. HBM sees each line once.
__device__ __forceinline__ uint32_t swz(uint32_t i) { return i ^ ((i >> 4) & 15u); }

template <int K, typename R>
__global__ void __launch_bounds__(64)
    k_dense_tiled(typename Cx<R>::V* __restrict__ states, uint32_t nbits, uint32_t tiles,
                  const typename Cx<R>::V* __restrict__ mats, const int* __restrict__ sel,
                  uint32_t batch_base, Tiling t) {
  using V = typename Cx<R>::V;
  constexpr int D = 1 << K;
  constexpr uint32_t M = 1u << (K + 6);
  __shared__ V tile[M];
  const uint32_t b = batch_base + blockIdx.y;
  const V* __restrict__ U = mats;
  if (sel != nullptr) {
    const int m = sel[b];
    if (m < 0) return;
    U += (size_t)m * (D * D);
  }
  const uint32_t lane = threadIdx.x;
  V* __restrict__ st = states + ((size_t)b << nbits);
  const uint32_t seglen = 1u << t.b, nseg = 1u << t.h;
  for (uint32_t tid = blockIdx.x; tid < tiles; tid += gridDim.x) {
    // global base of the tile: tile number above bit b, zeros spread in at the high targets
    uint32_t base = tid << t.b;
#pragma unroll
    for (int j = 0; j < K; ++j)
      if (j < (int)t.h) {
        const uint32_t s = t.high[j];
        base = ((base >> s) << (s + 1)) | (base & ((1u << s) - 1u));
      }
    // every lane loads M / 64 = D amplitudes, 64 consecutive per instruction. Row r of 64
    // lies in segment r >> (b - 6) at offset (r & (2^(b-6) - 1)) * 64, a wavefront-uniform
    // address: each load is a scalar base plus the lane number.
    const uint32_t rsh = t.b - 6u, rmask = (1u << rsh) - 1u;
#pragma unroll
    for (int r = 0; r < D; ++r) {
      const uint32_t row = base + t.seg[(uint32_t)r >> rsh] + (((uint32_t)r & rmask) << 6);
      tile[swz((uint32_t)r * 64u + lane)] = st[row + lane];
    }
    (void)nseg;
    (void)seglen;
    __syncthreads();
    uint32_t g = lane;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const uint32_t s = t.lsorted[j];
      g = ((g >> s) << (s + 1)) | (g & ((1u << s) - 1u));
    }
    V in[D];
#pragma unroll
    for (int c = 0; c < D; ++c) in[c] = tile[swz(g + t.loffs[c])];
#pragma unroll
    for (int r = 0; r < D; ++r) {
      R re = 0, im = 0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        const V u = U[r * D + c];
        re = fma(u.x, in[c].x, re);
        re = fma(-u.y, in[c].y, re);
        im = fma(u.x, in[c].y, im);
        im = fma(u.y, in[c].x, im);
      }
      V o;
      o.x = re;
      o.y = im;
      tile[swz(g + t.loffs[r])] = o;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < D; ++r) {
      const uint32_t row = base + t.seg[(uint32_t)r >> rsh] + (((uint32_t)r & rmask) << 6);
      st[row + lane] = tile[swz((uint32_t)r * 64u + lane)];
    }
    __syncthreads();
  }
}

// Diagonal operator: every amplitude is scaled by the diagonal entry its target bits select.
// The (at most 32-entry) diagonal is staged in LDS; four amplitudes per thread.
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_diag(typename Cx<R>::V* __restrict__ states, uint32_t nbits, int k,
           const typename Cx<R>::V* __restrict__ diags, const int* __restrict__ sel,
           uint32_t batch_base, Targets t) {
  using V = typename Cx<R>::V;
  __shared__ V d[1 << kMaxK];
  const uint32_t b = batch_base + blockIdx.y;
  const int D = 1 << k;
  const V* __restrict__ dg = diags;
  if (sel != nullptr) {
    const int m = sel[b];
    if (m < 0) return;
    dg += (size_t)m * D;
  }
  if ((int)threadIdx.x < D) d[threadIdx.x] = dg[threadIdx.x];
  __syncthreads();
  const uint64_t size = 1ull << nbits;
  V* __restrict__ st = states + ((size_t)b << nbits);
  const uint64_t first = (uint64_t)blockIdx.x * (kBlock * 4) + threadIdx.x;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint64_t i = first + (uint64_t)j * kBlock;
    if (i < size) {
      uint32_t m = 0;
      for (int q = 0; q < k; ++q) m |= (uint32_t)((i >> t.listed[q]) & 1u) << (k - 1 - q);
      st[i] = cmul(d[m], st[i]);
    }
  }
}

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_scale(typename Cx<R>::V* __restrict__ states, uint64_t total, double re, double im) {
  using V = typename Cx<R>::V;
  V c;
  c.x = (R)re;
  c.y = (R)im;
  const uint64_t first = (uint64_t)blockIdx.x * (kBlock * 4) + threadIdx.x;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint64_t i = first + (uint64_t)j * kBlock;
    if (i < total) states[i] = cmul(c, states[i]);
  }
}

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_set_basis(typename Cx<R>::V* __restrict__ states, uint32_t nbits, uint64_t total,
                uint64_t index) {
  using V = typename Cx<R>::V;
  const uint64_t first = (uint64_t)blockIdx.x * (kBlock * 4) + threadIdx.x;
  const uint64_t mask = (1ull << nbits) - 1;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint64_t i = first + (uint64_t)j * kBlock;
    if (i < total) {
      V v;
      v.x = (i & mask) == index ? (R)1 : (R)0;
      v.y = 0;
      states[i] = v;
    }
  }
}

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_probs(const typename Cx<R>::V* __restrict__ states, uint64_t total,
            double* __restrict__ probs) {
  const uint64_t first = (uint64_t)blockIdx.x * (kBlock * 4) + threadIdx.x;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint64_t i = first + (uint64_t)j * kBlock;
    if (i < total) {
      const double x = (double)states[i].x, y = (double)states[i].y;
      probs[i] = x * x + y * y;
    }
  }
}

// y <- y + a x over whole buffers (the accumulation of an out-of-place Kraus sum).
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_axpy(typename Cx<R>::V* __restrict__ y, const typename Cx<R>::V* __restrict__ x,
           uint64_t total, double re, double im) {
  using V = typename Cx<R>::V;
  V a;
  a.x = (R)re;
  a.y = (R)im;
  const uint64_t first = (uint64_t)blockIdx.x * (kBlock * 4) + threadIdx.x;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint64_t i = first + (uint64_t)j * kBlock;
    if (i < total) {
      const V p = cmul(a, x[i]);
      V o = y[i];
      o.x += p.x;
      o.y += p.y;
      y[i] = o;
    }
  }
}

// ----------------------------------------------------------------------------- launchers

template <typename R>
static int launch_dense_tiled(void* states, int nbits, long long batch, int k, const Tiling& t,
                              const void* mats, const int* sel, hipStream_t s) {
  using V = typename Cx<R>::V;
  const uint32_t tiles = 1u << (nbits - k - 6);
  // one tile per wavefront up to ~8 waves' worth per CU, then grid-stride
  const uint32_t gx = tiles < 32768u ? tiles : 32768u;
  for (long long b0 = 0; b0 < batch; b0 += kMaxGridY) {
    const uint32_t gy = (uint32_t)(batch - b0 < kMaxGridY ? batch - b0 : kMaxGridY);
    dim3 grid(gx, gy), block(64);
    V* st = (V*)states;
    const V* m = (const V*)mats;
    switch (k) {
      case 2: k_dense_tiled<2, R><<<grid, block, 0, s>>>(st, nbits, tiles, m, sel, (uint32_t)b0, t); break;
      case 3: k_dense_tiled<3, R><<<grid, block, 0, s>>>(st, nbits, tiles, m, sel, (uint32_t)b0, t); break;
      case 4: k_dense_tiled<4, R><<<grid, block, 0, s>>>(st, nbits, tiles, m, sel, (uint32_t)b0, t); break;
      default: return fail(LSIM_EINVAL, "tiled dense operator needs k between 2 and 4");
    }
  }
  return check_launch("tiled dense operator");
}

template <typename R>
static int launch_dense(void* states, int nbits, long long batch, int k, const Targets& t,
                        const void* mats, const int* sel, hipStream_t s) {
  using V = typename Cx<R>::V;
  const uint32_t groups = 1u << (nbits - k);
  const uint32_t gx = (groups + kBlock - 1) / kBlock;
  for (long long b0 = 0; b0 < batch; b0 += kMaxGridY) {
    const uint32_t gy = (uint32_t)(batch - b0 < kMaxGridY ? batch - b0 : kMaxGridY);
    dim3 grid(gx, gy), block(kBlock);
    V* st = (V*)states;
    const V* m = (const V*)mats;
    switch (k) {
      case 1: k_dense<1, R><<<grid, block, 0, s>>>(st, nbits, groups, m, sel, (uint32_t)b0, t); break;
      case 2: k_dense<2, R><<<grid, block, 0, s>>>(st, nbits, groups, m, sel, (uint32_t)b0, t); break;
      case 3: k_dense<3, R><<<grid, block, 0, s>>>(st, nbits, groups, m, sel, (uint32_t)b0, t); break;
      case 4: k_dense<4, R><<<grid, block, 0, s>>>(st, nbits, groups, m, sel, (uint32_t)b0, t); break;
      case 5:
        switch (g_k5_scalar_rows) {
          case 0: k_dense<5, R, 0><<<grid, block, 0, s>>>(st, nbits, groups, m, sel, (uint32_t)b0, t); break;
          case 8: k_dense<5, R, 8><<<grid, block, 0, s>>>(st, nbits, groups, m, sel, (uint32_t)b0, t); break;
          case 12: k_dense<5, R, 12><<<grid, block, 0, s>>>(st, nbits, groups, m, sel, (uint32_t)b0, t); break;
          case 16: k_dense<5, R, 16><<<grid, block, 0, s>>>(st, nbits, groups, m, sel, (uint32_t)b0, t); break;
          case 20: k_dense<5, R, 20><<<grid, block, 0, s>>>(st, nbits, groups, m, sel, (uint32_t)b0, t); break;
          case 24: k_dense<5, R, 24><<<grid, block, 0, s>>>(st, nbits, groups, m, sel, (uint32_t)b0, t); break;
          default: k_dense<5, R, 32><<<grid, block, 0, s>>>(st, nbits, groups, m, sel, (uint32_t)b0, t); break;
        }
        break;
    }
  }
  return check_launch("dense operator");
}

static int check_state(const void* states, int dtype, int nbits, long long batch) {
  LSIM_REQUIRE(states != nullptr, "the state pointer is NULL");
  LSIM_REQUIRE(dtype == LSIM_C128 || dtype == LSIM_C64, "dtype must be LSIM_C128 or LSIM_C64");
  LSIM_REQUIRE(nbits >= 0 && nbits <= kMaxBits, "nbits must be between 0 and %d, not %d", kMaxBits,
               nbits);
  LSIM_REQUIRE(batch >= 1 && batch <= (1ll << 31), "batch must be between 1 and 2^31");
  return LSIM_OK;
}

static unsigned blocks_for(uint64_t total) {
  return (unsigned)((total + kBlock * 4 - 1) / (kBlock * 4));
}

static int g_tiling_enabled = 1;

}  // namespace lsim

using namespace lsim;

extern "C" {

int lsim_set_dense_tiling(int enabled) {
  const int old = g_tiling_enabled;
  g_tiling_enabled = enabled != 0;
  return old;
}

int lsim_set_dense_k5_split(int scalar_rows) {
  if (scalar_rows < 0) return g_k5_scalar_rows;
  LSIM_REQUIRE(scalar_rows == 0 || scalar_rows == 8 || scalar_rows == 12 || scalar_rows == 16 ||
                   scalar_rows == 20 || scalar_rows == 24 || scalar_rows == 32,
               "the 5-qubit split must be 0, 8, 12, 16, 20, 24 or 32 rows, not %d", scalar_rows);
  g_k5_scalar_rows = scalar_rows;
  return LSIM_OK;
}

int lsim_sv_set_basis(void* states, int dtype, int nbits, long long batch, long long index,
                      void* stream) {
  if (int rc = check_state(states, dtype, nbits, batch)) return rc;
  LSIM_REQUIRE(index >= 0 && index < (1ll << nbits), "basis index %lld is outside the state",
               index);
  const uint64_t total = (uint64_t)batch << nbits;
  if (dtype == LSIM_C128)
    k_set_basis<double><<<blocks_for(total), kBlock, 0, as_stream(stream)>>>(
        (double2*)states, nbits, total, (uint64_t)index);
  else
    k_set_basis<float><<<blocks_for(total), kBlock, 0, as_stream(stream)>>>(
        (float2*)states, nbits, total, (uint64_t)index);
  return check_launch("set basis");
}

int lsim_sv_apply_dense(void* states, int dtype, int nbits, long long batch, int k,
                        const int* targets, const void* mats, const int* sel, void* stream) {
  if (int rc = check_state(states, dtype, nbits, batch)) return rc;
  LSIM_REQUIRE(mats != nullptr, "the matrix pointer is NULL");
  Targets t;
  if (int rc = make_targets(nbits, k, targets, &t)) return rc;
  Tiling tl;
  if (k >= 2 && g_tiling_enabled && make_tiling(nbits, k, targets, &tl)) {
    if (dtype == LSIM_C128)
      return launch_dense_tiled<double>(states, nbits, batch, k, tl, mats, sel, as_stream(stream));
    return launch_dense_tiled<float>(states, nbits, batch, k, tl, mats, sel, as_stream(stream));
  }
  if (dtype == LSIM_C128)
    return launch_dense<double>(states, nbits, batch, k, t, mats, sel, as_stream(stream));
  return launch_dense<float>(states, nbits, batch, k, t, mats, sel, as_stream(stream));
}

int lsim_sv_apply_diag(void* states, int dtype, int nbits, long long batch, int k,
                       const int* targets, const void* diags, const int* sel, void* stream) {
  if (int rc = check_state(states, dtype, nbits, batch)) return rc;
  LSIM_REQUIRE(diags != nullptr, "the diagonal pointer is NULL");
  Targets t;
  if (int rc = make_targets(nbits, k, targets, &t)) return rc;
  const uint64_t size = 1ull << nbits;
  for (long long b0 = 0; b0 < batch; b0 += kMaxGridY) {
    const uint32_t gy = (uint32_t)(batch - b0 < kMaxGridY ? batch - b0 : kMaxGridY);
    dim3 grid(blocks_for(size), gy);
    if (dtype == LSIM_C128)
      k_diag<double><<<grid, kBlock, 0, as_stream(stream)>>>(
          (double2*)states, nbits, k, (const double2*)diags, sel, (uint32_t)b0, t);
    else
      k_diag<float><<<grid, kBlock, 0, as_stream(stream)>>>(
          (float2*)states, nbits, k, (const float2*)diags, sel, (uint32_t)b0, t);
  }
  return check_launch("diagonal operator");
}

int lsim_sv_scale(void* states, int dtype, int nbits, long long batch, double re, double im,
                  void* stream) {
  if (int rc = check_state(states, dtype, nbits, batch)) return rc;
  const uint64_t total = (uint64_t)batch << nbits;
  if (dtype == LSIM_C128)
    k_scale<double><<<blocks_for(total), kBlock, 0, as_stream(stream)>>>((double2*)states, total,
                                                                        re, im);
  else
    k_scale<float><<<blocks_for(total), kBlock, 0, as_stream(stream)>>>((float2*)states, total, re,
                                                                       im);
  return check_launch("scale");
}

int lsim_sv_axpy(void* y, const void* x, int dtype, int nbits, long long batch, double re,
                 double im, void* stream) {
  if (int rc = check_state(y, dtype, nbits, batch)) return rc;
  LSIM_REQUIRE(x != nullptr, "the source pointer is NULL");
  const uint64_t total = (uint64_t)batch << nbits;
  if (dtype == LSIM_C128)
    k_axpy<double><<<blocks_for(total), kBlock, 0, as_stream(stream)>>>(
        (double2*)y, (const double2*)x, total, re, im);
  else
    k_axpy<float><<<blocks_for(total), kBlock, 0, as_stream(stream)>>>(
        (float2*)y, (const float2*)x, total, re, im);
  return check_launch("axpy");
}

int lsim_sv_probabilities(const void* states, int dtype, int nbits, long long batch,
                          double* probs, void* stream) {
  if (int rc = check_state(states, dtype, nbits, batch)) return rc;
  LSIM_REQUIRE(probs != nullptr, "the output pointer is NULL");
  const uint64_t total = (uint64_t)batch << nbits;
  if (dtype == LSIM_C128)
    k_probs<double><<<blocks_for(total), kBlock, 0, as_stream(stream)>>>((const double2*)states,
                                                                        total, probs);
  else
    k_probs<float><<<blocks_for(total), kBlock, 0, as_stream(stream)>>>((const float2*)states,
                                                                       total, probs);
  return check_launch("probabilities");
}

}  // extern "C"
