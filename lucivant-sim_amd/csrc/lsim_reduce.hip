// Reductions over states: reduced density matrices on 1-2 target bits (measurement
// probabilities and Kraus branch weights), Pauli-string expectations for statevectors and
// density matrices, and density-matrix diagonals.
//
// Every reduction is deterministic: workgroup b of a state sums a fixed, strided subset of the
// state with wavefront shuffles and LDS, writes one partial per value, and a second kernel adds
// the partials of each state in a fixed order. Repeating a call gives the same bits, which keeps
// seeded trajectory decisions reproducible run to run. Accumulation is float64 for both dtypes.
#include <string.h>

#include <mutex>

#include "lsim_internal.h"

namespace lsim {

// Workgroups per state for a reduction over `items` work items: about 8 items per thread, and
// up to 4096 workgroups (16 waves per CU on 256 CUs) so one large state keeps HBM busy.
static uint32_t reduce_blocks(uint64_t items) {
  uint64_t nb = items / (kBlock * 8);
  if (nb < 1) nb = 1;
  if (nb > 4096) nb = 4096;
  return (uint32_t)nb;
}

// Sum NV per-thread values across the workgroup; thread 0 writes them to out[0..NV).
template <int NV>
__device__ __forceinline__ void block_sum_store(double (&acc)[NV], double* __restrict__ out) {
  __shared__ double part[kBlock / 64][NV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const double s = wave_sum(acc[v]);
    if (lane == 0) part[wave][v] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) s += part[w][threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// Reduced density matrix on K target bits. Values per state: D diagonal reals, then for each
// pair r < c (row-major) the real and imaginary parts of sum a_r conj(a_c).
template <int K, typename R>
__global__ void __launch_bounds__(kBlock)
    k_rdm(const typename Cx<R>::V* __restrict__ states, uint32_t nbits, uint32_t groups,
          uint32_t nb, Targets t, double* __restrict__ partial, uint32_t batch_base) {
  using V = typename Cx<R>::V;
  constexpr int D = 1 << K;
  constexpr int NV = D * D;
  const uint32_t b = batch_base + blockIdx.y;
  const V* __restrict__ st = states + ((size_t)b << nbits);
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0;
  for (uint32_t g = blockIdx.x * kBlock + threadIdx.x; g < groups; g += nb * kBlock) {
    const uint32_t base = spread(g, t.sorted, K);
    double x[D], y[D];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      const V a = st[base + t.offs[c]];
      x[c] = (double)a.x;
      y[c] = (double)a.y;
    }
    int v = D;
#pragma unroll
    for (int r = 0; r < D; ++r) {
      acc[r] += x[r] * x[r] + y[r] * y[r];
#pragma unroll
      for (int c = r + 1; c < D; ++c) {
        acc[v++] += x[r] * x[c] + y[r] * y[c];
        acc[v++] += y[r] * x[c] - x[r] * y[c];
      }
    }
  }
  block_sum_store<NV>(acc, partial + ((size_t)b * nb + blockIdx.x) * NV);
}

// Sum, in a fixed order, the nb partials of every value: one workgroup per (state, value), a
// strided per-thread sum then a wavefront/LDS tree. out_sum[b * nv + v].
__global__ void __launch_bounds__(kBlock)
    k_sum_values(const double* __restrict__ partial, uint32_t nb, uint32_t nv,
                 double* __restrict__ out_sum) {
  const uint32_t v = blockIdx.x, b = blockIdx.y;
  const double* p = partial + ((size_t)b * nb) * nv + v;
  double acc[1] = {0};
  for (uint32_t j = threadIdx.x; j < nb; j += kBlock) acc[0] += p[(size_t)j * nv];
  block_sum_store<1>(acc, out_sum + (size_t)b * nv + v);
}

// Expand the summed values of each state to a full Hermitian 2^K x 2^K complex matrix.
template <int K>
__global__ void k_rdm_finish(const double* __restrict__ sums, uint64_t batch,
                             double2* __restrict__ out) {
  constexpr int D = 1 << K;
  constexpr int NV = D * D;
  const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= batch * NV) return;
  const uint64_t b = id / NV;
  const int rc = (int)(id % NV), r = rc / D, c = rc % D;
  const double* p = sums + b * NV;
  double2 v;
  if (r == c) {
    v.x = p[r];
    v.y = 0;
  } else {
    const int lo = r < c ? r : c, hi = r < c ? c : r;
    // position of pair (lo, hi) among the pairs r < c in row-major order
    const int pair = lo * D - lo * (lo + 1) / 2 + (hi - lo - 1);
    v.x = p[D + 2 * pair];
    v.y = r < c ? p[D + 2 * pair + 1] : -p[D + 2 * pair + 1];
  }
  out[b * NV + rc] = v;
}

struct PauliTerms {
  uint64_t x[LSIM_MAX_TERMS];
  uint64_t z[LSIM_MAX_TERMS];
};

__device__ __forceinline__ double re_times_ipow(int w, double zx, double zy) {
  // Re(i^w * z)
  switch (w & 3) {
    case 0: return zx;
    case 1: return -zy;
    case 2: return -zx;
    default: return zy;
  }
}

// <psi|P|psi> = sum_j conj(psi[j^x]) i^nY (-1)^popcount(j&z) psi[j]. With x != 0 the terms of j
// and j^x are complex conjugates, so the sum runs over j with the top bit of x clear and adds
// twice the real part: each amplitude is read once per string.
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_pauli_sv(const typename Cx<R>::V* __restrict__ states, uint32_t nbits, uint32_t nb,
               int nterms, PauliTerms terms, double* __restrict__ partial, uint32_t batch_base) {
  using V = typename Cx<R>::V;
  const uint32_t t = blockIdx.y % nterms;
  const uint32_t b = batch_base + blockIdx.y / nterms;
  const uint64_t X = terms.x[t], Z = terms.z[t];
  const int w = __popcll(X & Z) & 3;
  const V* __restrict__ st = states + ((size_t)b << nbits);
  double acc[1] = {0};
  if (X == 0) {
    const uint64_t size = 1ull << nbits;
    for (uint64_t j = blockIdx.x * kBlock + threadIdx.x; j < size; j += (uint64_t)nb * kBlock) {
      const V a = st[j];
      const double p = (double)a.x * a.x + (double)a.y * a.y;
      acc[0] += (__popcll(j & Z) & 1) ? -p : p;
    }
  } else {
    const uint32_t h = 63 - __clzll(X);
    const uint64_t half = 1ull << (nbits - 1);
    for (uint64_t g = blockIdx.x * kBlock + threadIdx.x; g < half; g += (uint64_t)nb * kBlock) {
      const uint64_t j = ((g >> h) << (h + 1)) | (g & ((1ull << h) - 1));
      const V a = st[j], p = st[j ^ X];
      // z = conj(p) * a
      const double zx = (double)p.x * a.x + (double)p.y * a.y;
      const double zy = (double)p.x * a.y - (double)p.y * a.x;
      const double v = 2.0 * re_times_ipow(w, zx, zy);
      acc[0] += (__popcll(j & Z) & 1) ? -v : v;
    }
  }
  block_sum_store<1>(acc, partial + ((size_t)b * nterms + t) * nb + blockIdx.x);
}

// ---- Pauli strings in groups: one read of the state for every string of a group ----------
//
// <psi|P|psi> = sum_j Re[i^w conj(psi[j ^ X]) psi[j]] (-1)^popcount(j & Z), w = popcount(X & Z).
// The strings of a group have X masks whose span V has at most kGroupDims dimensions above the
// low `low` bits. Take W = V + span(bits 0..low-1), with a basis b_i of V reduced to zero low
// bits and distinct pivot bits. Every address is j = r ^ B(c) ^ l: r a coset representative (zero
// at the low bits and the pivots), B(c) = xor of the b_i selected by c, l < 2^low. One workgroup
// stages a coset, 2^(low + dims) = 4096 amplitudes read as runs of 2^low contiguous ones, in LDS;
// for every string the partner j ^ X is r ^ B(c ^ c_t) ^ (l ^ l_t), inside the same tile, at
// tile index e ^ tx with e = (c << low) | l. The sign factors as well: parity(j & Z) =
// parity(r & Z) ^ parity(e & tz), where tz holds Z's low bits and parity(b_i & Z) at bit low + i.
constexpr int kTileBits = 12;   // amplitudes per coset tile: 4096 (64 KiB complex128)
constexpr int kGroupDims = 8;   // span dimensions above the low bits, at most
constexpr int kGroupTerms = 16; // strings per group (accumulators per thread)

struct PauliGroup {
  uint64_t basis[kGroupDims];  // b_i, ordered by pivot bit
  uint64_t z[kGroupTerms];     // each string's Z mask (the coset sign)
  uint32_t tx[kGroupTerms];    // tile-index X mask (c_t << low | l_t)
  uint32_t tz[kGroupTerms];    // tile-index Z mask
  uint32_t kpar[kGroupTerms];  // bit k: parity of tz & (k * kBlock), the sign of entry k of a thread
  int32_t pairbit[kGroupTerms];  // a bit b of tx >> 8 (the k bits), or -1 when tx has none there
  int32_t slot[kGroupTerms];   // the string's position in the call's output (-1: padding)
  uint32_t wodd, wneg;         // bit t: w_t odd (Re[i^w z] is +-Im z), Re[i^w z] negated (w = 1, 2)
  uint8_t zeros[kTileBits];    // ascending bit positions a coset representative leaves zero
  int nterms, low, dims;
  int diagonal;                // every X mask is 0: the strings need no partner amplitudes
};

// NT = the group's string count rounded up to a power of two (padding strings read a partner and
// are never written out), so every loop below unrolls and every accumulator is a register.
// SL: streaming (non-temporal) loads, for states the caches cannot keep.
template <typename R, int NT, bool SL>
__global__ void __launch_bounds__(kBlock)
    k_pauli_group(const typename Cx<R>::V* __restrict__ states, uint32_t nbits, uint32_t nb,
                  int call_terms, PauliGroup g, double* __restrict__ partial, uint32_t batch_base) {
  using V = typename Cx<R>::V;
  constexpr int PER = (1 << kTileBits) / kBlock;
  constexpr uint32_t kTileMask = (1u << kTileBits) - 1u;
  __shared__ V tile[1 << kTileBits];
  const uint32_t b = batch_base + blockIdx.y;
  const V* __restrict__ st = states + ((size_t)b << nbits);
  const uint64_t cosets = 1ull << (nbits - kTileBits);
  const uint32_t lmask = (1u << g.low) - 1u;
  // offset B(c) ^ l of each of this thread's tile entries e = tid + k * kBlock (the same in
  // every coset; addresses stay below 2^kMaxBits)
  uint32_t off[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint32_t e = threadIdx.x + k * kBlock;
    const uint32_t c = e >> g.low;
    uint64_t o = e & lmask;
    for (int i = 0; i < g.dims; ++i)
      if ((c >> i) & 1u) o ^= g.basis[i];
    off[k] = (uint32_t)o;
  }
  // the sign (-1)^parity(e & tz) splits into this thread's part, parity(tid & tz), applied at the
  // end, and the entry's part kpar (e = tid + k * kBlock: the two use disjoint bits)
  uint32_t tpar = 0;
#pragma unroll
  for (int t = 0; t < NT; ++t) tpar |= (uint32_t)(__popc(threadIdx.x & g.tz[t]) & 1) << t;
  double acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = 0;
  // coset q's representative: q with zeros spread in at g.zeros
  auto representative = [&](uint64_t q) {
    for (int i = 0; i < kTileBits; ++i) {
      const uint32_t p = g.zeros[i];
      q = ((q >> p) << (p + 1)) | (q & ((1ull << p) - 1));
    }
    return (uint32_t)q;
  };
  // Software pipeline: the amplitudes of the workgroup's next coset are loaded into `a` while
  // the strings of the current one are evaluated from LDS. (The barriers wait only for LDS, so
  // the loads stay in flight until the next coset's LDS stores need them.) After its last coset
  // a workgroup loads coset 0 once more, which every workgroup shares and the cache holds: the
  // loads stay unconditional, so the registers they fill are the ones the next iteration reads.
  R are[PER], aim[PER];  // (as two arrays of scalars: an array of V is kept in scratch memory)
  uint32_t r = blockIdx.x < cosets ? representative(blockIdx.x) : 0u;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const V v = load_amp<SL>(st + (r ^ off[k]));
    are[k] = v.x;
    aim[k] = v.y;
  }
  for (uint64_t q = blockIdx.x; q < cosets; q += nb) {
    // a wave staging its amplitudes holds up its workgroup at the barrier, while the other
    // workgroup on the compute unit has strings to evaluate: it issues first until the barrier
    // (2% per 8-string call at 26 qubits)
    __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      V v;
      v.x = are[k];
      v.y = aim[k];
      tile[threadIdx.x + k * kBlock] = v;
    }
    __syncthreads();
    __builtin_amdgcn_s_setprio(0);
    uint32_t cpar = 0;  // coset sign of each string, parity(r & Z) (wavefront-uniform)
#pragma unroll
    for (int t = 0; t < NT; ++t) cpar |= (uint32_t)(__popcll(r & g.z[t]) & 1) << t;
    r = q + nb < cosets ? representative(q + nb) : 0u;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const V v = load_amp<SL>(st + (r ^ off[k]));
      are[k] = v.x;
      aim[k] = v.y;
    }
    // String by string. Entries e and e ^ tx have equal terms (the term of j ^ X is the complex
    // conjugate of the term of j inside Re[i^w ...], times (-1)^popcount(X & Z) = (-1)^w), so a
    // string whose tx has a bit b among the k bits (8..11 of e = tid + k * kBlock) visits only the
    // k with bit b clear and doubles the sum; the others visit every k. Own amplitude and partner
    // come from LDS; the form of the term (Re or Im of conj(p) a, by the parity of w) is a
    // wavefront-uniform branch per string.
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (t >= g.nterms) break;
      const uint32_t tx = g.tx[t];
      const V* __restrict__ pt = tile + (threadIdx.x ^ (tx & (kBlock - 1u)));  // partner's thread part
      const V* __restrict__ ot = tile + threadIdx.x;
      const uint32_t xh = tx >> 8;  // the partner of k is k ^ xh
      const int pb = g.pairbit[t];
      const uint32_t neg = g.kpar[t] ^ (((cpar >> t) & 1u) ? 0xFFFFu : 0u);  // bit k: term sign
      const int count = pb >= 0 ? PER / 2 : PER;
      double sum = 0.0;
      if ((g.wodd >> t) & 1u) {
        for (int kk = 0; kk < count; ++kk) {
          const uint32_t k = pb >= 0 ? (((uint32_t)kk >> pb) << (pb + 1)) | ((uint32_t)kk & ((1u << pb) - 1u)) : (uint32_t)kk;
          const V a = ot[k * kBlock], p = pt[(k ^ xh) * kBlock];
          sum = fma(((neg >> k) & 1u) ? -1.0 : 1.0, (double)p.x * a.y - (double)p.y * a.x, sum);
        }
      } else {
        for (int kk = 0; kk < count; ++kk) {
          const uint32_t k = pb >= 0 ? (((uint32_t)kk >> pb) << (pb + 1)) | ((uint32_t)kk & ((1u << pb) - 1u)) : (uint32_t)kk;
          const V a = ot[k * kBlock], p = pt[(k ^ xh) * kBlock];
          sum = fma(((neg >> k) & 1u) ? -1.0 : 1.0, (double)p.x * a.x + (double)p.y * a.y, sum);
        }
      }
      acc[t] += pb >= 0 ? 2.0 * sum : sum;
    }
    __syncthreads();  // the tile is refilled by the next coset
  }
  // workgroup sums in a fixed order; string t's partial goes to its own row
  __shared__ double part[kBlock / 64][NT];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const double v = ((tpar ^ g.wneg) >> t) & 1u ? -acc[t] : acc[t];
    const double sum = wave_sum(v);
    if (lane == 0) part[wave][t] = sum;
  }
  __syncthreads();
  if ((int)threadIdx.x < g.nterms) {
    double sum = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) sum += part[w][threadIdx.x];
    partial[((size_t)b * call_terms + g.slot[threadIdx.x]) * nb + blockIdx.x] = sum;
  }
}

// Groups of diagonal strings (X = 0, only I and Z): <psi|P|psi> = sum_j (-1)^popcount(j & Z)
// |psi[j]|^2 needs no partner amplitude, so no LDS tile. Workgroups of kDiagBlock threads (no
// LDS, so the compute units hold more of them and more loads stay in flight) step through
// chunks of kDiagBlock x kDiagLoads amplitudes: amplitude j = chunk base + u * kDiagBlock + tid.
// The sign splits over those disjoint bit fields: parity(tid & Z) per thread, applied at the
// end; parity(u * kDiagBlock & Z) and parity(base & Z), wavefront-uniform.
constexpr int kDiagBlock = 1024;
constexpr int kDiagLoads = 4;
static_assert(kDiagBlock * kDiagLoads == 1 << kTileBits, "a chunk is one coset tile");

template <typename R, int NT, bool SL>
__global__ void __launch_bounds__(kDiagBlock)
    k_pauli_diag(const typename Cx<R>::V* __restrict__ states, uint32_t nbits, uint32_t nb,
                 int call_terms, PauliGroup g, double* __restrict__ partial, uint32_t batch_base) {
  using V = typename Cx<R>::V;
  const uint32_t b = batch_base + blockIdx.y;
  const V* __restrict__ st = states + ((size_t)b << nbits);
  const uint64_t chunks = 1ull << (nbits - kTileBits);
  uint32_t tpar = 0, upar[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    tpar |= (uint32_t)(__popcll(threadIdx.x & g.z[t]) & 1) << t;
    upar[t] = 0;
#pragma unroll
    for (int u = 0; u < kDiagLoads; ++u)
      upar[t] |= (uint32_t)(__popcll((uint64_t)(u * kDiagBlock) & g.z[t]) & 1) << u;
  }
  double acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = 0;
  for (uint64_t c = blockIdx.x; c < chunks; c += nb) {
    const uint64_t base = c << kTileBits;
    V a[kDiagLoads];
#pragma unroll
    for (int u = 0; u < kDiagLoads; ++u) a[u] = load_amp<SL>(st + base + u * kDiagBlock + threadIdx.x);
    uint32_t cpar = 0;  // parity(base & Z) of each string
#pragma unroll
    for (int t = 0; t < NT; ++t) cpar |= (uint32_t)(__popcll(base & g.z[t]) & 1) << t;
#pragma unroll
    for (int u = 0; u < kDiagLoads; ++u) {
      const double p = (double)a[u].x * a[u].x + (double)a[u].y * a[u].y;
#pragma unroll
      for (int t = 0; t < NT; ++t)
        acc[t] = fma((((upar[t] >> u) ^ (cpar >> t)) & 1u) ? -1.0 : 1.0, p, acc[t]);
    }
  }
  // workgroup sums in a fixed order; string t's partial goes to its own row
  __shared__ double part[kDiagBlock / 64][NT];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const double sum = wave_sum((tpar >> t) & 1u ? -acc[t] : acc[t]);
    if (lane == 0) part[wave][t] = sum;
  }
  __syncthreads();
  if ((int)threadIdx.x < g.nterms) {
    double sum = 0;
#pragma unroll
    for (int w = 0; w < kDiagBlock / 64; ++w) sum += part[w][threadIdx.x];
    partial[((size_t)b * call_terms + g.slot[threadIdx.x]) * nb + blockIdx.x] = sum;
  }
}

// Tr(P rho) = sum_m i^nY (-1)^popcount(m&z) rho[m][m^x] (real part) for an n-qubit rho.
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_pauli_dm(const typename Cx<R>::V* __restrict__ rho, uint32_t n, uint32_t nb, int nterms,
               PauliTerms terms, double* __restrict__ partial, uint32_t batch_base) {
  using V = typename Cx<R>::V;
  const uint32_t t = blockIdx.y % nterms;
  const uint32_t b = batch_base + blockIdx.y / nterms;
  const uint64_t X = terms.x[t], Z = terms.z[t];
  const int w = __popcll(X & Z) & 3;
  const V* __restrict__ m0 = rho + ((size_t)b << (2 * n));
  const uint64_t dim = 1ull << n;
  double acc[1] = {0};
  for (uint64_t m = blockIdx.x * kBlock + threadIdx.x; m < dim; m += (uint64_t)nb * kBlock) {
    const V e = m0[m * dim + (m ^ X)];
    const double v = re_times_ipow(w, (double)e.x, (double)e.y);
    acc[0] += (__popcll(m & Z) & 1) ? -v : v;
  }
  block_sum_store<1>(acc, partial + ((size_t)b * nterms + t) * nb + blockIdx.x);
}


template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_dm_probs(const typename Cx<R>::V* __restrict__ rho, uint32_t n, uint64_t batch,
               double* __restrict__ probs) {
  const uint64_t dim = 1ull << n;
  const uint64_t id = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (id >= batch * dim) return;
  const uint64_t b = id / dim, i = id % dim;
  const double v = (double)rho[(b << (2 * n)) + i * dim + i].x;
  probs[id] = v > 0.0 ? v : 0.0;
}

// out[r] = sum of the nb consecutive partials of row r, one workgroup per row (fixed order).
static void sum_rows(const double* partial, uint32_t nb, uint64_t rows, double* out,
                     hipStream_t s) {
  for (uint64_t r0 = 0; r0 < rows; r0 += kMaxGridY) {
    const uint32_t gy = (uint32_t)(rows - r0 < kMaxGridY ? rows - r0 : kMaxGridY);
    k_sum_values<<<dim3(1, gy), kBlock, 0, s>>>(partial + r0 * nb, nb, 1, out + r0);
  }
}

static int check_common(const void* states, int dtype, int nbits, long long batch) {
  LSIM_REQUIRE(states != nullptr, "the state pointer is NULL");
  LSIM_REQUIRE(dtype == LSIM_C128 || dtype == LSIM_C64, "dtype must be LSIM_C128 or LSIM_C64");
  LSIM_REQUIRE(nbits >= 1 && nbits <= kMaxBits, "nbits must be between 1 and %d, not %d", kMaxBits,
               nbits);
  LSIM_REQUIRE(batch >= 1 && batch <= (1ll << 31), "batch must be between 1 and 2^31");
  return LSIM_OK;
}

static int load_terms(int nterms, const unsigned long long* xmask, const unsigned long long* zmask,
                      int nbits, PauliTerms* terms) {
  LSIM_REQUIRE(nterms >= 1 && nterms <= LSIM_MAX_TERMS, "nterms must be between 1 and %d",
               LSIM_MAX_TERMS);
  LSIM_REQUIRE(xmask != nullptr && zmask != nullptr, "the Pauli masks are NULL");
  const unsigned long long valid = nbits >= 64 ? ~0ull : ((1ull << nbits) - 1);
  for (int t = 0; t < LSIM_MAX_TERMS; ++t) {
    terms->x[t] = t < nterms ? xmask[t] : 0;
    terms->z[t] = t < nterms ? zmask[t] : 0;
    LSIM_REQUIRE((terms->x[t] | terms->z[t]) <= valid, "Pauli term %d acts on a qubit beyond %d",
                 t, nbits);
  }
  return LSIM_OK;
}

template <typename R, bool SL>
static void launch_group(int m, dim3 grid, hipStream_t s, const typename Cx<R>::V* states,
                         uint32_t nbits, uint32_t nb, int call_terms, const PauliGroup& g,
                         double* partial, uint32_t b0) {
  const int nt = m <= 1 ? 1 : m <= 2 ? 2 : m <= 4 ? 4 : m <= 8 ? 8 : 16;
#define LSIM_PAULI_LAUNCH(KERNEL, NT, BLOCK) \
  KERNEL<R, NT, SL><<<grid, BLOCK, 0, s>>>(states, nbits, nb, call_terms, g, partial, b0)
  if (g.diagonal) {
    switch (nt) {
      case 1: LSIM_PAULI_LAUNCH(k_pauli_diag, 1, kDiagBlock); break;
      case 2: LSIM_PAULI_LAUNCH(k_pauli_diag, 2, kDiagBlock); break;
      case 4: LSIM_PAULI_LAUNCH(k_pauli_diag, 4, kDiagBlock); break;
      case 8: LSIM_PAULI_LAUNCH(k_pauli_diag, 8, kDiagBlock); break;
      default: LSIM_PAULI_LAUNCH(k_pauli_diag, 16, kDiagBlock); break;
    }
  } else {
    switch (nt) {
      case 1: LSIM_PAULI_LAUNCH(k_pauli_group, 1, kBlock); break;
      case 2: LSIM_PAULI_LAUNCH(k_pauli_group, 2, kBlock); break;
      case 4: LSIM_PAULI_LAUNCH(k_pauli_group, 4, kBlock); break;
      case 8: LSIM_PAULI_LAUNCH(k_pauli_group, 8, kBlock); break;
      default: LSIM_PAULI_LAUNCH(k_pauli_group, 16, kBlock); break;
    }
  }
#undef LSIM_PAULI_LAUNCH
}

// Workgroups per state, i.e. partials per (state, string), of the expectation path for nbits.
static uint32_t pauli_sv_blocks(int nbits) {
  if (nbits < kTileBits) return reduce_blocks(1ull << (nbits - 1));
  // two 64 KiB tiles fit a compute unit's LDS: 512 workgroups are one round on 256 compute units,
  // and each steps through its cosets with the next one's loads in flight (26 qubits, 8 strings
  // in one group: 334 us with 4096 workgroups of 4 cosets, 269-288 us with 512)
  const uint64_t cosets = 1ull << (nbits - kTileBits);
  return (uint32_t)(cosets < 512 ? cosets : 512);
}

static size_t pauli_sv_workspace_bytes(int nbits, long long batch, int nterms) {
  return (size_t)batch * nterms * pauli_sv_blocks(nbits) * sizeof(double);
}

// Rank over GF(2) of the masks v >> low.
static int rank_above(const uint64_t* xs, int count, int low) {
  uint64_t basis[64] = {};
  int rank = 0;
  for (int i = 0; i < count; ++i) {
    uint64_t v = xs[i] >> low;
    for (int bit = 63; bit >= 0 && v; --bit)
      if ((v >> bit) & 1) {
        if (basis[bit]) {
          v ^= basis[bit];
        } else {
          basis[bit] = v;
          ++rank;
          break;
        }
      }
  }
  return rank;
}

// Splits the call's strings into groups: the diagonal strings (X = 0) first, in call order, in
// groups of up to kGroupTerms; then the others, in call order: a string joins the current group
// while the group has fewer than kGroupTerms strings and its X masks keep a span of at most
// kGroupDims dimensions above bit 4 (runs of at least 16 contiguous amplitudes). Each string's
// value lands in its call position (slot). Returns the number of groups, or -1 if a group breaks
// an invariant the kernel's indexing relies on (low + dims == kTileBits, every tile mask below
// 2^kTileBits, every address bit below nbits).
static int make_groups(const PauliTerms& terms, int nterms, int nbits, PauliGroup* groups) {
  uint64_t xo[LSIM_MAX_TERMS], zo[LSIM_MAX_TERMS];
  int io[LSIM_MAX_TERMS], n = 0;
  for (int pass = 0; pass < 2; ++pass)
    for (int i = 0; i < nterms; ++i)
      if ((terms.x[i] == 0) == (pass == 0)) {
        xo[n] = terms.x[i];
        zo[n] = terms.z[i];
        io[n++] = i;
      }
  int count = 0, first = 0;
  while (first < nterms) {
    const bool diagonal = xo[first] == 0;
    int last = first + 1;  // [first, last) is the group
    while (last < nterms && last - first < kGroupTerms && (xo[last] == 0) == diagonal &&
           rank_above(xo + first, last + 1 - first, 4) <= kGroupDims)
      ++last;
    const int m = last - first;
    const uint64_t* xs = xo + first;
    // the widest low part that keeps the tile at kTileBits: low + rank_above(low) == kTileBits
    int low = 4;
    for (int l = 4; l <= kTileBits; ++l)
      if (l + rank_above(xs, m, l) <= kTileBits) low = l;
    PauliGroup& g = groups[count++];
    memset(&g, 0, sizeof g);
    g.nterms = m;
    g.low = low;
    g.diagonal = diagonal ? 1 : 0;
    // reduced basis of the masks with their low bits cleared: distinct pivots, each pivot bit in
    // its own vector only
    uint64_t basis[64] = {};
    for (int i = 0; i < m; ++i) {
      uint64_t v = xs[i] & ~((1ull << low) - 1);
      for (int bit = 63; bit >= 0 && v; --bit)
        if (((v >> bit) & 1) && basis[bit]) v ^= basis[bit];
      if (!v) continue;
      const int pivot = 63 - __builtin_clzll(v);
      for (int bit = 0; bit < 64; ++bit)
        if (basis[bit] && ((basis[bit] >> pivot) & 1)) basis[bit] ^= v;
      basis[pivot] = v;
    }
    int pivots[kGroupDims];
    for (int bit = 0; bit < 64; ++bit)
      if (basis[bit]) {
        pivots[g.dims] = bit;
        g.basis[g.dims++] = basis[bit];
      }
    // zero positions of a coset representative: the low bits, then the pivots (ascending)
    int nz = 0;
    for (int bit = 0; bit < low; ++bit) g.zeros[nz++] = (uint8_t)bit;
    for (int i = 0; i < g.dims; ++i) g.zeros[nz++] = (uint8_t)pivots[i];
    for (int i = 0; i < m; ++i) {
      const uint64_t x = xs[i], z = zo[first + i];
      uint32_t c = 0;
      uint64_t rest = x;
      for (int d = 0; d < g.dims; ++d)
        if ((x >> pivots[d]) & 1) {
          c |= 1u << d;
          rest ^= g.basis[d];
        }
      uint32_t tz = (uint32_t)(z & ((1ull << low) - 1));
      for (int d = 0; d < g.dims; ++d) tz |= (uint32_t)(__builtin_popcountll(g.basis[d] & z) & 1) << (low + d);
      if (rest >> low) return -1;  // the mask lies outside the group's span
      const int w = __builtin_popcountll(x & z) & 3;
      g.tx[i] = (c << low) | (uint32_t)rest;
      g.tz[i] = tz;
      g.z[i] = z;
      for (int k = 0; k < (1 << kTileBits) / kBlock; ++k)
        g.kpar[i] |= (uint32_t)(__builtin_popcount(tz & (uint32_t)(k * kBlock)) & 1) << k;
      g.pairbit[i] = (g.tx[i] >> 8) ? 31 - __builtin_clz(g.tx[i] >> 8) : -1;
      g.wodd |= (uint32_t)(w & 1) << i;
      g.wneg |= (uint32_t)(((w + 1) >> 1) & 1) << i;
      g.slot[i] = io[first + i];
    }
    if (g.low + g.dims != kTileBits) return -1;
    for (int d = 0; d < g.dims; ++d)
      if (nbits < 64 && (g.basis[d] >> nbits)) return -1;
    for (int i = 0; i < kTileBits; ++i)
      if (g.zeros[i] >= nbits) return -1;
    first = last;
  }
  return count;
}

int launch_pauli_sv(const void* states, int dtype, int nbits, long long batch, int nterms,
                    const PauliTerms& terms, double* out, double* workspace, void* stream);

// Scratch of lsim_sv_pauli_expect_noalloc: a file-scope device array, so that entry point calls
// no allocator. Each GPU the library runs on holds its own copy until the process ends. In host
// code the array's name is not a device address: the address comes from hipGetSymbolAddress,
// looked up once per device with that device current, and only that address is passed on.
constexpr size_t kScratchBytes = 4u << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
static std::mutex scratch_mutex;
static void* scratch_on[kMaxDevices] = {};

static int scratch_address(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(LSIM_EHIP, "no current device: %s", hipGetErrorString(e));
  LSIM_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d is beyond the %d this library serves", dev,
               kMaxDevices);
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (scratch_on[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return fail(LSIM_EHIP, "scratch address on device %d: %s", dev, hipGetErrorString(e));
    scratch_on[dev] = p;
  }
  *out = static_cast<double*>(scratch_on[dev]);
  return LSIM_OK;
}

}  // namespace lsim

using namespace lsim;

extern "C" {

size_t lsim_reduce_workspace_bytes(int nbits, long long batch, int width) {
  if (nbits < 1 || nbits > kMaxBits || batch < 1 || width < 1) return 0;
  // the widest user: a reduced density matrix on 2 bits (16 values per workgroup), or one value
  // per Pauli term
  const uint64_t nb = reduce_blocks(1ull << (nbits - 1));
  const uint64_t per = width <= 2 ? (uint64_t)(1 << (2 * width)) : (uint64_t)width;
  const uint64_t per_rdm = 16;
  const uint64_t per_state = per > per_rdm ? per : per_rdm;
  // partials, then one summed value per (state, value) for the reduced density matrices
  return (size_t)((batch * nb * per_state + batch * per_rdm) * sizeof(double));
}

int lsim_sv_reduced_dm(const void* states, int dtype, int nbits, long long batch, int k,
                       const int* targets, double* out, double* workspace, void* stream) {
  if (int rc = check_common(states, dtype, nbits, batch)) return rc;
  LSIM_REQUIRE(k == 1 || k == 2, "reduced density matrices are on 1 or 2 bits, not %d", k);
  LSIM_REQUIRE(out != nullptr && workspace != nullptr, "output or workspace pointer is NULL");
  Targets t;
  if (int rc = make_targets(nbits, k, targets, &t)) return rc;
  const uint32_t groups = 1u << (nbits - k);
  const uint32_t nb = reduce_blocks(groups);
  hipStream_t s = as_stream(stream);
  for (long long b0 = 0; b0 < batch; b0 += kMaxGridY) {
    const uint32_t gy = (uint32_t)(batch - b0 < kMaxGridY ? batch - b0 : kMaxGridY);
    dim3 grid(nb, gy);
    if (dtype == LSIM_C128) {
      if (k == 1)
        k_rdm<1, double><<<grid, kBlock, 0, s>>>((const double2*)states, nbits, groups, nb, t,
                                                 workspace, (uint32_t)b0);
      else
        k_rdm<2, double><<<grid, kBlock, 0, s>>>((const double2*)states, nbits, groups, nb, t,
                                                 workspace, (uint32_t)b0);
    } else {
      if (k == 1)
        k_rdm<1, float><<<grid, kBlock, 0, s>>>((const float2*)states, nbits, groups, nb, t,
                                                workspace, (uint32_t)b0);
      else
        k_rdm<2, float><<<grid, kBlock, 0, s>>>((const float2*)states, nbits, groups, nb, t,
                                                workspace, (uint32_t)b0);
    }
  }
  const uint32_t nv = 1u << (2 * k);
  double* sums = workspace + (size_t)batch * nb * nv;
  for (long long b0 = 0; b0 < batch; b0 += kMaxGridY) {
    const uint32_t gy = (uint32_t)(batch - b0 < kMaxGridY ? batch - b0 : kMaxGridY);
    k_sum_values<<<dim3(nv, gy), kBlock, 0, s>>>(workspace + (size_t)b0 * nb * nv, nb, nv,
                                               sums + (size_t)b0 * nv);
  }
  const uint64_t count = (uint64_t)batch << (2 * k);
  const unsigned fb = (unsigned)((count + kBlock - 1) / kBlock);
  if (k == 1)
    k_rdm_finish<1><<<fb, kBlock, 0, s>>>(sums, (uint64_t)batch, (double2*)out);
  else
    k_rdm_finish<2><<<fb, kBlock, 0, s>>>(sums, (uint64_t)batch, (double2*)out);
  return check_launch("reduced density matrix");
}

int lsim_sv_pauli_expect(const void* states, int dtype, int nbits, long long batch, int nterms,
                         const unsigned long long* xmask, const unsigned long long* zmask,
                         double* out, double* workspace, void* stream) {
  if (int rc = check_common(states, dtype, nbits, batch)) return rc;
  LSIM_REQUIRE(out != nullptr && workspace != nullptr, "output or workspace pointer is NULL");
  PauliTerms terms;
  if (int rc = load_terms(nterms, xmask, zmask, nbits, &terms)) return rc;
  return launch_pauli_sv(states, dtype, nbits, batch, nterms, terms, out, workspace, stream);
}

int lsim_sv_pauli_expect_noalloc(const void* states, int dtype, int nbits, long long batch,
                                 int nterms, const unsigned long long* xmask,
                                 const unsigned long long* zmask, double* out, void* stream) {
  if (int rc = check_common(states, dtype, nbits, batch)) return rc;
  LSIM_REQUIRE(out != nullptr, "the output pointer is NULL");
  PauliTerms terms;
  if (int rc = load_terms(nterms, xmask, zmask, nbits, &terms)) return rc;
  const size_t need = pauli_sv_workspace_bytes(nbits, batch, nterms);
  LSIM_REQUIRE(need <= kScratchBytes,
               "%lld states of %d bits with %d strings need %zu bytes of scratch; this entry point "
               "has %zu (use lsim_sv_pauli_expect with a workspace)",
               batch, nbits, nterms, need, kScratchBytes);
  double* ws = nullptr;
  if (int rc = scratch_address(&ws)) return rc;
  return launch_pauli_sv(states, dtype, nbits, batch, nterms, terms, out, ws, stream);
}

}  // extern "C"

namespace lsim {

// The expectation launches shared by both entry points: per-string partials into `workspace`
// (every slot of batch x nterms x nb is written before it is read), then their fixed-order sums.
int launch_pauli_sv(const void* states, int dtype, int nbits, long long batch, int nterms,
                    const PauliTerms& terms, double* out, double* workspace, void* stream) {
  const uint32_t nb = pauli_sv_blocks(nbits);
  hipStream_t s = as_stream(stream);
  if (nbits >= kTileBits) {
    PauliGroup groups[LSIM_MAX_TERMS];
    const int ng = make_groups(terms, nterms, nbits, groups);
    if (ng < 0) return fail(LSIM_EINVAL, "Pauli strings: grouping failed its own checks");
    // streaming loads where the dense passes stream (measured on a 1 GiB state: one diagonal
    // string 176 -> 158 us with them)
    const bool sl = dense_streams(((uint64_t)batch << nbits) * (dtype == LSIM_C128 ? 16 : 8));
    for (long long b0 = 0; b0 < batch; b0 += kMaxGridY) {
      const uint32_t gy = (uint32_t)(batch - b0 < kMaxGridY ? batch - b0 : kMaxGridY);
      for (int i = 0; i < ng; ++i) {
        const dim3 grid(nb, gy);
        const int m = groups[i].nterms;
        if (dtype == LSIM_C128 && sl)
          launch_group<double, true>(m, grid, s, (const double2*)states, nbits, nb, nterms,
                                     groups[i], workspace, (uint32_t)b0);
        else if (dtype == LSIM_C128)
          launch_group<double, false>(m, grid, s, (const double2*)states, nbits, nb, nterms,
                                      groups[i], workspace, (uint32_t)b0);
        else if (sl)
          launch_group<float, true>(m, grid, s, (const float2*)states, nbits, nb, nterms,
                                    groups[i], workspace, (uint32_t)b0);
        else
          launch_group<float, false>(m, grid, s, (const float2*)states, nbits, nb, nterms,
                                     groups[i], workspace, (uint32_t)b0);
      }
    }
    sum_rows(workspace, nb, (uint64_t)batch * nterms, out, s);
    return check_launch("Pauli expectation");
  }
  const long long per_launch = kMaxGridY / nterms;
  for (long long b0 = 0; b0 < batch; b0 += per_launch) {
    const long long nbatch = batch - b0 < per_launch ? batch - b0 : per_launch;
    dim3 grid(nb, (uint32_t)(nbatch * nterms));
    if (dtype == LSIM_C128)
      k_pauli_sv<double><<<grid, kBlock, 0, s>>>((const double2*)states, nbits, nb, nterms, terms,
                                                 workspace, (uint32_t)b0);
    else
      k_pauli_sv<float><<<grid, kBlock, 0, s>>>((const float2*)states, nbits, nb, nterms, terms,
                                                workspace, (uint32_t)b0);
  }
  sum_rows(workspace, nb, (uint64_t)batch * nterms, out, s);
  return check_launch("Pauli expectation");
}

}  // namespace lsim

extern "C" {

int lsim_dm_probabilities(const void* rho, int dtype, int n, long long batch, double* probs,
                          void* stream) {
  if (int rc = check_common(rho, dtype, 2 * n, batch)) return rc;
  LSIM_REQUIRE(probs != nullptr, "the output pointer is NULL");
  const uint64_t count = (uint64_t)batch << n;
  const unsigned blocks = (unsigned)((count + kBlock - 1) / kBlock);
  if (dtype == LSIM_C128)
    k_dm_probs<double><<<blocks, kBlock, 0, as_stream(stream)>>>((const double2*)rho, n, batch,
                                                                 probs);
  else
    k_dm_probs<float><<<blocks, kBlock, 0, as_stream(stream)>>>((const float2*)rho, n, batch,
                                                                probs);
  return check_launch("density-matrix probabilities");
}

int lsim_dm_pauli_expect(const void* rho, int dtype, int n, long long batch, int nterms,
                         const unsigned long long* xmask, const unsigned long long* zmask,
                         double* out, double* workspace, void* stream) {
  if (int rc = check_common(rho, dtype, 2 * n, batch)) return rc;
  LSIM_REQUIRE(out != nullptr && workspace != nullptr, "output or workspace pointer is NULL");
  PauliTerms terms;
  if (int rc = load_terms(nterms, xmask, zmask, n, &terms)) return rc;
  const uint32_t nb = reduce_blocks(1ull << n);
  hipStream_t s = as_stream(stream);
  const long long per_launch = kMaxGridY / nterms;
  for (long long b0 = 0; b0 < batch; b0 += per_launch) {
    const long long nbatch = batch - b0 < per_launch ? batch - b0 : per_launch;
    dim3 grid(nb, (uint32_t)(nbatch * nterms));
    if (dtype == LSIM_C128)
      k_pauli_dm<double><<<grid, kBlock, 0, s>>>((const double2*)rho, n, nb, nterms, terms,
                                                 workspace, (uint32_t)b0);
    else
      k_pauli_dm<float><<<grid, kBlock, 0, s>>>((const float2*)rho, n, nb, nterms, terms,
                                                workspace, (uint32_t)b0);
  }
  sum_rows(workspace, nb, (uint64_t)batch * nterms, out, s);
  return check_launch("density-matrix Pauli expectation");
}

}  // extern "C"
