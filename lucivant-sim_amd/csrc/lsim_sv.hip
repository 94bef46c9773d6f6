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
#include <atomic>

#include "lsim_internal.h"

namespace lsim {

// bit k set: complex128 k-qubit operators (k = 3..5) run on the matrix cores (k_dense_mfma), on
// every placement (26 qubits: 51.0 -> 50.7 ms per step against the vector kernel for 3-4 qubit
// operators whose targets are all at bit 3 or above)
static std::atomic<int> g_mfma_widths{(1 << 3) | (1 << 4) | (1 << 5)};
// streaming (non-temporal) loads and stores: 0 never; 1 (the default) unless the launch's states
// fit the last-level cache but not the L2s; 2 at every size. The dense kernels also need a
// placement whose load instructions cover whole lines.
static std::atomic<int> g_dense_streaming{1};
// MI355X: 8 XCDs with 4 MiB of L2 each, and a 256 MiB Infinity Cache (MALL) in front of HBM.
// Complex128 passes with 40 random 1-5 qubit operators, per pass:
//   - 20 qubits (16 MiB): 10.6 us streaming, 14.4 us cached. The state sits in the L2s, and a
//     cached kernel ends by writing its dirty lines back, which streaming stores avoid;
//   - 22-24 qubits (64-256 MiB): cached wins, 24.2 / 41.3 / 79.1 us against 25.4 / 48.0 / 93.7,
//     because the MALL keeps the state from one pass to the next;
//   - 25-26 qubits (512 MiB, 1 GiB): every pass streams from HBM and the hint wins, 185 / 380 us
//     against 197 / 408.
constexpr uint64_t kL2Bytes = 32ull << 20;
constexpr uint64_t kLastLevelCacheBytes = 256ull << 20;

static bool streams(uint64_t bytes) {
  const int mode = g_dense_streaming.load(std::memory_order_relaxed);
  if (mode == 2) return true;
  return mode == 1 && (bytes <= kL2Bytes || bytes > kLastLevelCacheBytes);
}

bool dense_streams(uint64_t bytes) { return streams(bytes); }

// loads and stores with the streaming (non-temporal) hint, or plain ones
template <bool NT, typename T>
__device__ __forceinline__ T ld(const T* p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void st_(T* p, T v) {
  if constexpr (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// a complex value read or written as its two components, with or without the streaming hint
template <bool NT, typename V>
__device__ __forceinline__ V ldc(const V* p) {
  const auto* q = reinterpret_cast<const decltype(p->x)*>(p);
  V v;
  v.x = ld<NT>(q);
  v.y = ld<NT>(q + 1);
  return v;
}
template <bool NT, typename V>
__device__ __forceinline__ void stc(V* p, V v) {
  auto* q = reinterpret_cast<decltype(p->x)*>(p);
  st_<NT>(q, v.x);
  st_<NT>(q + 1, v.y);
}

// NT: streaming (non-temporal) loads and stores, chosen (for the sizes streams() picks) when
// every target is at bit 3 or above: each load instruction then covers whole 128-byte lines, and
// keeping them in cache gains nothing (lower targets rely on the cache for the partner loads of
// the same line).
template <int K, typename R, bool NT>
__global__ void __launch_bounds__(kBlock)
    k_dense(typename Cx<R>::V* __restrict__ states, uint32_t nbits, uint32_t groups,
            const typename Cx<R>::V* __restrict__ mats, const int* __restrict__ sel,
            uint32_t batch_base, Targets t) {
  using V = typename Cx<R>::V;
  constexpr int D = 1 << K;
  const uint32_t b = batch_base + blockIdx.y;
  const V* __restrict__ U = mats;
  if (sel != nullptr) {
    const int m = sel[b];
    if (m < 0) return;
    U += (size_t)m * (D * D);
  }
  // A 5-qubit operator is 16 KiB, as large as the scalar cache: it is staged once per workgroup
  // in LDS and read as wavefront-uniform broadcasts. (complex128 5-qubit operators normally run
  // on the matrix cores instead, see k_dense_mfma.)
  __shared__ V ulds[K == 5 ? D * D : 1];
  if constexpr (K == 5) {
    for (int e = threadIdx.x; e < D * D; e += kBlock) ulds[e] = U[e];
    __syncthreads();
  }
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= groups) return;
  R* __restrict__ st = reinterpret_cast<R*>(states + ((size_t)b << nbits));
  const uint32_t base = spread(g, t.sorted, K);
  V in[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    const R* p = st + 2 * (size_t)(base + t.offs[c]);
    in[c].x = ld<NT>(p);
    in[c].y = ld<NT>(p + 1);
  }
#pragma unroll
  for (int r = 0; r < D; ++r) {
    R re = 0, im = 0;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      const V u = K == 5 ? ulds[r * D + c] : U[r * D + c];
      re = fma(u.x, in[c].x, re);
      re = fma(-u.y, in[c].y, re);
      im = fma(u.x, in[c].y, im);
      im = fma(u.y, in[c].x, im);
    }
    R* p = st + 2 * (size_t)(base + t.offs[r]);
    st_<NT>(p, re);
    st_<NT>(p + 1, im);
  }
}

// K-qubit operator (K = 3..5) on the FP64 matrix cores. The complex 2^K x 2^K operator U acts as
// the real 2^(K+1) x 2^(K+1) M = [[Re U, -Im U], [Im U, Re U]] on x = [Re psi_g; Im psi_g], and
// one wavefront applies it to 16 groups at a time: Y = M X with X, Y of 2^(K+1) x 16, i.e.
// 2^(K+1)/16 row tiles x 2^(K+1)/4 k-steps of v_mfma_f64_16x16x4f64. The lane layout of that
// instruction on gfx950 (A: lane l holds A[l%16][l/16]; B: B[l/16][l%16]; C/D: register v of
// tile tt holds D[16 tt + 4v + l/16][l%16], checked with a one-instruction probe) puts, on lane
// (i = l%16, k = l/16):
//   - operator fragment (tile tt, step kk): M[16 tt + i][4 kk + k], loaded once per wavefront and
//     kept in registers, so no operand is broadcast from LDS or the scalar cache;
//   - input steps j and j + 2^(K-2): Re and Im of amplitude 4 j + k of group i (one 16-byte load);
//   - output register R = 4 tt + v holds row 4R + k: Re of amplitude 4R + k for R < 2^(K-2), and
//     its Im in register R + 2^(K-2).
// For one amplitude index the 16 lanes of a row read 16 consecutive groups. The flop count equals
// the vector kernel's (4 real products per complex one); what changes is operand delivery, which
// limits the vector kernel at K = 5 (an LDS broadcast per four FMAs). NT: streaming (non-temporal)
// loads and stores, chosen when each load instruction covers whole 128-byte lines (see
// launch_dense_mfma), so nothing is gained by keeping them in cache.
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

template <int K, bool NT>
__global__ void __launch_bounds__(kBlock)
    k_dense_mfma(double2* __restrict__ states, uint32_t nbits, uint32_t batches,
                 const double2* __restrict__ mats, const int* __restrict__ sel,
                 uint32_t batch_base, Targets t) {
  constexpr int D = 1 << K;            // complex dimension
  constexpr int TILES = 2 * D / 16;    // row tiles of M
  constexpr int STEPS = 2 * D / 4;     // k-steps
  constexpr int LOADS = D / 4;         // amplitudes per lane per 16 groups
  const uint32_t b = batch_base + blockIdx.y;
  const double2* __restrict__ U = mats;
  if (sel != nullptr) {
    const int m = sel[b];
    if (m < 0) return;
    U += (size_t)m * (D * D);
  }
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t i = lane & 15u, k = lane >> 4;
  double a[TILES][STEPS];
#pragma unroll
  for (int tt = 0; tt < TILES; ++tt)
#pragma unroll
    for (int kk = 0; kk < STEPS; ++kk) {
      const uint32_t row = 16u * tt + i, col = 4u * kk + k;
      const double2 u = U[(row % D) * D + (col % D)];
      a[tt][kk] = ((row < (uint32_t)D) == (col < (uint32_t)D)) ? u.x : (row < (uint32_t)D ? -u.y : u.y);
    }
  // amplitude offsets split into a wavefront-uniform part and this lane's part (index bit
  // (K-1-j) of the matrix index selects target listed[j]; the lane's k is bits 1..0 of both its
  // input and its output amplitude indices)
  const uint32_t in_lane = (((k >> 1) & 1u) << t.listed[K - 2]) | ((k & 1u) << t.listed[K - 1]);
  double* __restrict__ st = reinterpret_cast<double*>(states + ((size_t)b << nbits));
  const uint32_t wave = (blockIdx.x * kBlock + threadIdx.x) >> 6;
  const uint32_t nwaves = gridDim.x * (kBlock / 64);
  // two adjacent steps (32 consecutive groups) per iteration: their loads for one amplitude index
  // are neighbouring 256-byte runs issued back to back, which the memory system serves better
  // than lone runs (26 qubits: 5-qubit passes 0.414 -> 0.402 ms)
  for (uint32_t pair = wave; 2 * pair < batches; pair += nwaves) {
    uint32_t base[2];
    f64x2 x[2][LOADS];
#pragma unroll
    for (int h = 0; h < 2; ++h) base[h] = spread((2 * pair + h) * 16u + i, t.sorted, K);
#pragma unroll
    for (int j = 0; j < LOADS; ++j)
#pragma unroll
      for (int h = 0; h < 2; ++h)
        x[h][j] = ld<NT>(reinterpret_cast<const f64x2*>(st + 2 * (size_t)(base[h] + t.offs[4 * j] + in_lane)));
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      f64x4 acc[TILES];
#pragma unroll
      for (int tt = 0; tt < TILES; ++tt) {
        acc[tt] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int j = 0; j < LOADS; ++j) acc[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[tt][j], x[h][j].x, acc[tt], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < LOADS; ++j) acc[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[tt][j + LOADS], x[h][j].y, acc[tt], 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < LOADS; ++r) {
        const f64x2 o = {acc[r / 4][r % 4], acc[(r + LOADS) / 4][(r + LOADS) % 4]};
        st_<NT>(reinterpret_cast<f64x2*>(st + 2 * (size_t)(base[h] + t.offs[4 * r] + in_lane)), o);
      }
    }
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
template <typename R, bool NT>
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
      stc<NT>(st + i, cmul(d[m], ldc<NT>(st + i)));
    }
  }
}

template <typename R, bool NT>
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
    if (i < total) stc<NT>(states + i, cmul(c, ldc<NT>(states + i)));
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
// state dst[j] <- state src[j] of a batch, for j = first + blockIdx.y; indices outside the batch
// are skipped
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_copy_states(typename Cx<R>::V* states, uint32_t nbits, uint32_t batch,
                  const int* __restrict__ src, const int* __restrict__ dst, uint32_t first) {
  using V = typename Cx<R>::V;
  const uint32_t j = first + blockIdx.y;
  const int from = src[j], to = dst[j];
  if (from < 0 || to < 0 || (uint32_t)from >= batch || (uint32_t)to >= batch) return;
  const uint64_t size = 1ull << nbits;
  const V* __restrict__ in = states + ((size_t)from << nbits);
  V* __restrict__ out = states + ((size_t)to << nbits);
  const uint64_t i0 = (uint64_t)blockIdx.x * (kBlock * 4) + threadIdx.x;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const uint64_t i = i0 + (uint64_t)r * kBlock;
    if (i < size) out[i] = in[i];
  }
}

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

static int launch_dense_mfma(void* states, int nbits, long long batch, int k, const Targets& t,
                             const void* mats, const int* sel, hipStream_t s) {
  const uint32_t batches = 1u << (nbits - k - 4);  // 16 groups of 2^k amplitudes per wavefront step
  // one iteration (two steps) per wavefront: 2^(nbits-k-5) wavefronts per state. Measured at 26
  // qubits and k = 5 against fewer wavefronts that stride over the state: 8192 -> 394 us per pass,
  // 16384 -> 389, 32768 -> 384, one iteration each (65536) -> 374, and 416 -> 390 us where two
  // or more targets are at bit 20 and above. Each wavefront loads its operator fragments (16 KiB
  // at k = 5) from cache, against 32 KiB of state it reads and writes.
  const uint32_t waves = batches / 2 > 0 ? batches / 2 : 1u;
  const uint32_t gx = (waves + 3) / 4;
  // Streaming accesses (sizes as for streams()) when each load instruction covers whole 128-byte
  // lines (index bits 0-2):
  // an instruction spans the 4 lowest non-target bits (16 consecutive groups) and the two targets
  // the lanes' k selects (listed[k-2], listed[k-1]), so every target below bit 3 must be one of
  // those two.
  uint32_t low_targets = 0;
  for (int j = 0; j < k; ++j)
    if (t.listed[j] < 3) low_targets |= 1u << t.listed[j];
  const uint32_t lane_targets = (1u << t.listed[k - 2]) | (1u << t.listed[k - 1]);
  const bool nt = streams(((uint64_t)batch << nbits) * 16) && (low_targets & ~lane_targets) == 0;
  for (long long b0 = 0; b0 < batch; b0 += kMaxGridY) {
    const uint32_t gy = (uint32_t)(batch - b0 < kMaxGridY ? batch - b0 : kMaxGridY);
    const dim3 grid(gx, gy), block(kBlock);
    double2* st = (double2*)states;
    const double2* m = (const double2*)mats;
    const uint32_t bb = (uint32_t)b0;
    switch (k * 2 + (nt ? 1 : 0)) {
      case 6: k_dense_mfma<3, false><<<grid, block, 0, s>>>(st, nbits, batches, m, sel, bb, t); break;
      case 7: k_dense_mfma<3, true><<<grid, block, 0, s>>>(st, nbits, batches, m, sel, bb, t); break;
      case 8: k_dense_mfma<4, false><<<grid, block, 0, s>>>(st, nbits, batches, m, sel, bb, t); break;
      case 9: k_dense_mfma<4, true><<<grid, block, 0, s>>>(st, nbits, batches, m, sel, bb, t); break;
      case 10: k_dense_mfma<5, false><<<grid, block, 0, s>>>(st, nbits, batches, m, sel, bb, t); break;
      case 11: k_dense_mfma<5, true><<<grid, block, 0, s>>>(st, nbits, batches, m, sel, bb, t); break;
      default: return fail(LSIM_EINVAL, "the matrix-core dense operator needs k between 3 and 5");
    }
  }
  return check_launch("dense operator (matrix cores)");
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
    const bool nt = streams(((uint64_t)batch << nbits) * sizeof(V)) && t.sorted[0] >= 3;
#define LSIM_DENSE(KK)                                                                     \
  do {                                                                                     \
    if (nt)                                                                                \
      k_dense<KK, R, true><<<grid, block, 0, s>>>(st, nbits, groups, m, sel, (uint32_t)b0, t);  \
    else                                                                                   \
      k_dense<KK, R, false><<<grid, block, 0, s>>>(st, nbits, groups, m, sel, (uint32_t)b0, t); \
  } while (0)
    switch (k) {
      case 1: LSIM_DENSE(1); break;
      case 2: LSIM_DENSE(2); break;
      case 3: LSIM_DENSE(3); break;
      case 4: LSIM_DENSE(4); break;
      case 5: LSIM_DENSE(5); break;
    }
#undef LSIM_DENSE
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

static std::atomic<int> g_tiling_enabled{1};

}  // namespace lsim

using namespace lsim;

extern "C" {

int lsim_set_dense_tiling(int enabled) {
  if (enabled >= 0) return g_tiling_enabled.exchange(enabled != 0 ? 1 : 0);
  return g_tiling_enabled.load();
}

int lsim_set_dense_mfma(int widths) {
  if (widths >= 0) return g_mfma_widths.exchange(widths & ((1 << 3) | (1 << 4) | (1 << 5)));
  return g_mfma_widths.load();
}

int lsim_set_dense_streaming(int enabled) {
  if (enabled >= 0) return g_dense_streaming.exchange(enabled > 2 ? 2 : enabled);
  return g_dense_streaming.load();
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
  if (k >= 2 && g_tiling_enabled.load(std::memory_order_relaxed) && make_tiling(nbits, k, targets, &tl)) {
    if (dtype == LSIM_C128)
      return launch_dense_tiled<double>(states, nbits, batch, k, tl, mats, sel, as_stream(stream));
    return launch_dense_tiled<float>(states, nbits, batch, k, tl, mats, sel, as_stream(stream));
  }
  if (dtype == LSIM_C128) {
    if (k >= 3 && ((g_mfma_widths.load(std::memory_order_relaxed) >> k) & 1) && nbits >= k + 5)
      return launch_dense_mfma(states, nbits, batch, k, t, mats, sel, as_stream(stream));
    return launch_dense<double>(states, nbits, batch, k, t, mats, sel, as_stream(stream));
  }
  return launch_dense<float>(states, nbits, batch, k, t, mats, sel, as_stream(stream));
}

int lsim_sv_apply_diag(void* states, int dtype, int nbits, long long batch, int k,
                       const int* targets, const void* diags, const int* sel, void* stream) {
  if (int rc = check_state(states, dtype, nbits, batch)) return rc;
  LSIM_REQUIRE(diags != nullptr, "the diagonal pointer is NULL");
  Targets t;
  if (int rc = make_targets(nbits, k, targets, &t)) return rc;
  const uint64_t size = 1ull << nbits;
  const bool nt = streams(((uint64_t)batch << nbits) * dtype_bytes(dtype));
  for (long long b0 = 0; b0 < batch; b0 += kMaxGridY) {
    const uint32_t gy = (uint32_t)(batch - b0 < kMaxGridY ? batch - b0 : kMaxGridY);
    dim3 grid(blocks_for(size), gy);
    hipStream_t s = as_stream(stream);
    const uint32_t bb = (uint32_t)b0;
    if (dtype == LSIM_C128) {
      double2* st = (double2*)states;
      const double2* dg = (const double2*)diags;
      if (nt)
        k_diag<double, true><<<grid, kBlock, 0, s>>>(st, nbits, k, dg, sel, bb, t);
      else
        k_diag<double, false><<<grid, kBlock, 0, s>>>(st, nbits, k, dg, sel, bb, t);
    } else {
      float2* st = (float2*)states;
      const float2* dg = (const float2*)diags;
      if (nt)
        k_diag<float, true><<<grid, kBlock, 0, s>>>(st, nbits, k, dg, sel, bb, t);
      else
        k_diag<float, false><<<grid, kBlock, 0, s>>>(st, nbits, k, dg, sel, bb, t);
    }
  }
  return check_launch("diagonal operator");
}

int lsim_sv_scale(void* states, int dtype, int nbits, long long batch, double re, double im,
                  void* stream) {
  if (int rc = check_state(states, dtype, nbits, batch)) return rc;
  const uint64_t total = (uint64_t)batch << nbits;
  const bool nt = streams(total * dtype_bytes(dtype));
  const dim3 grid(blocks_for(total));
  hipStream_t s = as_stream(stream);
  if (dtype == LSIM_C128) {
    if (nt)
      k_scale<double, true><<<grid, kBlock, 0, s>>>((double2*)states, total, re, im);
    else
      k_scale<double, false><<<grid, kBlock, 0, s>>>((double2*)states, total, re, im);
  } else {
    if (nt)
      k_scale<float, true><<<grid, kBlock, 0, s>>>((float2*)states, total, re, im);
    else
      k_scale<float, false><<<grid, kBlock, 0, s>>>((float2*)states, total, re, im);
  }
  return check_launch("scale");
}

int lsim_sv_copy_states(void* states, int dtype, int nbits, long long batch, long long count,
                        const int* src, const int* dst, void* stream) {
  if (int rc = check_state(states, dtype, nbits, batch)) return rc;
  LSIM_REQUIRE(count >= 0 && count <= batch, "count must be between 0 and the batch size");
  if (count == 0) return LSIM_OK;
  LSIM_REQUIRE(src != nullptr && dst != nullptr, "the index pointer is NULL");
  const uint64_t size = 1ull << nbits;
  hipStream_t s = as_stream(stream);
  for (long long c0 = 0; c0 < count; c0 += kMaxGridY) {
    const uint32_t gy = (uint32_t)(count - c0 < kMaxGridY ? count - c0 : kMaxGridY);
    const dim3 grid(blocks_for(size), gy);
    if (dtype == LSIM_C128)
      k_copy_states<double><<<grid, kBlock, 0, s>>>((double2*)states, nbits, (uint32_t)batch, src,
                                                    dst, (uint32_t)c0);
    else
      k_copy_states<float><<<grid, kBlock, 0, s>>>((float2*)states, nbits, (uint32_t)batch, src,
                                                   dst, (uint32_t)c0);
  }
  return check_launch("copy states");
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
