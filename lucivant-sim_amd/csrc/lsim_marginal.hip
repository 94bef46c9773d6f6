// Marginal distributions on the device: the probabilities of a subset of qubits, summed over the
// others in exactly the order NumPy sums them, so that sampling from the marginal gives the same
// outcomes bit for bit as the reference's host marginal (engines.py:267-275).
//
// The reference reduces the [2]*n tensor of probabilities over the dropped axes with
// `tensor.sum(axis=drop)`. NumPy 2's reduction of a C-contiguous float64 array walks the flat
// index in increasing order. When the lowest r qubits are all dropped (r >= 1), each run of
// 2^r contiguous entries is fed to the inner loop in pieces of at most 8192 entries (NumPy's
// iterator buffer size, np.getbufsize()), each piece is summed by NumPy's pairwise summation,
// and the piece sums are added one by one to the output entry, which starts at 0. When qubit 0
// is kept, every entry is added on its own, in flat order. So with C = 2^min(r, 13):
//
//   out[j] = 0 + S(c_0) + S(c_1) + ...   over the chunks c of C entries whose kept bits spell j,
//                                         in increasing order of c,
//   S(c)   = NumPy's pairwise sum of the chunk (for C = 1, the entry itself).
//
// NumPy's pairwise sum of a power-of-two count m: m < 8: 0 + a0 + a1 + ... in order; 8 <= m <=
// 128: eight running sums r_j = a_j + a_{j+8} + a_{j+16} + ..., then ((r0+r1)+(r2+r3)) +
// ((r4+r5)+(r6+r7)); m > 128: the sum of the two halves, recursively. Sampling code that
// replaces the host marginal with this one is pinned by a test comparing both bit for bit.
//
// Stage 1 writes S(c) for every chunk (skipped for C = 1); stage 2 forms each output entry by
// the ordered additions, one thread per entry.
#include "lsim_internal.h"

namespace lsim {

constexpr int kChunkBits = 13;  // log2 of NumPy's iterator buffer (8192 entries)

// Operands of an ordered chain loaded ahead of their additions (per is a power of two, so a
// chain of at least kBatch entries is a whole number of batches).
constexpr int kBatch = 32;

// Chunk-index bit of each kept qubit, in the order of the output index bits.
struct MarginalBits {
  uint32_t pos[kMaxBits];
};

// NumPy's pairwise sum of m = 2^mb <= 128 consecutive entries starting at a.
__device__ __forceinline__ double np_pairwise_small(const double* __restrict__ a, int m) {
  if (m < 8) {
    double res = 0.0;
    for (int i = 0; i < m; ++i) res += a[i];
    return res;
  }
  double r[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = a[j];
  for (int i = 8; i < m; i += 8) {
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] += a[i + j];
  }
  return ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
}

// Chunks of at most 128 entries: one thread per chunk.
__global__ void __launch_bounds__(kBlock)
    k_chunk_sums_small(const double* __restrict__ probs, uint64_t chunks, int m,
                       double* __restrict__ sums) {
  const uint64_t c = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c < chunks) sums[c] = np_pairwise_small(probs + c * (uint64_t)m, m);
}

// Chunks of 256 .. 8192 entries: one 64-thread workgroup per chunk. Thread i sums block i of 128
// entries; the blocks are then combined as the recursive halving does, left half + right half,
// which for a power-of-two block count is the balanced tree over the blocks in order.
__global__ void __launch_bounds__(64)
    k_chunk_sums_tree(const double* __restrict__ probs, int blocks, double* __restrict__ sums) {
  __shared__ double part[64];
  const uint64_t c = blockIdx.x;
  const int i = threadIdx.x;
  if (i < blocks) part[i] = np_pairwise_small(probs + (c * (uint64_t)blocks + i) * 128, 128);
  __syncthreads();
  for (int s = 1; s < blocks; s <<= 1) {
    if ((i & (2 * s - 1)) == 0 && i + s < blocks) part[i] = part[i] + part[i + s];
    __syncthreads();
  }
  if (i == 0) sums[c] = part[0];
}

// out[b][j] = 0 + sums over the chunks whose kept bits spell j, in increasing chunk order. In
// chunk-index bits, `kept` holds the kept qubits' positions and `drop` the mask of the others;
// the next chunk of the same j is ((d | ~drop) + 1) & drop, the usual walk through a mask's
// subsets in increasing order.
__global__ void __launch_bounds__(kBlock)
    k_ordered_accumulate(const double* __restrict__ sums, uint32_t cbits, int m, MarginalBits kb,
                         uint64_t drop, uint64_t per, uint64_t outputs, double* __restrict__ out) {
  const uint64_t id = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (id >= outputs) return;
  const uint64_t b = id >> m, j = id & ((1ull << m) - 1);
  uint64_t base = 0;
  for (int i = 0; i < m; ++i) base |= ((j >> i) & 1ull) << kb.pos[i];
  const double* __restrict__ row = sums + (b << cbits);
  double acc = 0.0;
  uint64_t d = 0;
  if (per >= kBatch) {
    // the additions are a chain, but their operands are not: load a batch ahead, then add in
    // order (one wait per batch instead of one memory latency per addition)
    for (uint64_t t = 0; t < per; t += kBatch) {
      double v[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        v[u] = row[base | d];
        d = ((d | ~drop) + 1) & drop;
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u) acc += v[u];
    }
  } else {
    for (uint64_t t = 0; t < per; ++t) {
      acc += row[base | d];
      d = ((d | ~drop) + 1) & drop;
    }
  }
  out[id] = acc;
}

}  // namespace lsim

using namespace lsim;

namespace {

// The chunk size (as log2) and the lowest run of dropped qubits for this keep set.
int chunk_bits(int n, int nkeep, const int* keep) {
  int r = 0;
  while (r < n) {
    bool kept = false;
    for (int i = 0; i < nkeep; ++i) kept |= keep[i] == r;
    if (kept) break;
    ++r;
  }
  return r < kChunkBits ? r : kChunkBits;
}

int check_keep(int n, int nkeep, const int* keep) {
  LSIM_REQUIRE(n >= 1 && n <= kMaxBits, "n must be between 1 and %d, not %d", kMaxBits, n);
  LSIM_REQUIRE(nkeep >= 1 && nkeep <= n, "between 1 and %d kept qubits, not %d", n, nkeep);
  LSIM_REQUIRE(keep != nullptr, "the kept qubits are NULL");
  for (int i = 0; i < nkeep; ++i) {
    LSIM_REQUIRE(keep[i] >= 0 && keep[i] < n, "kept qubit %d is outside 0..%d", keep[i], n - 1);
    LSIM_REQUIRE(i == 0 || keep[i] > keep[i - 1], "the kept qubits must ascend");
  }
  return LSIM_OK;
}

}  // namespace

extern "C" {

size_t lsim_marginal_workspace_bytes(int n, long long batch, int nkeep, const int* keep) {
  if (check_keep(n, nkeep, keep) != LSIM_OK || batch < 1) return 0;
  const int c = chunk_bits(n, nkeep, keep);
  return c == 0 ? 0 : (size_t)batch * ((size_t)1 << (n - c)) * sizeof(double);
}

int lsim_marginal(const double* probs, int n, long long batch, int nkeep, const int* keep,
                  double* out, double* workspace, void* stream) {
  if (int rc = check_keep(n, nkeep, keep); rc != LSIM_OK) return rc;
  LSIM_REQUIRE(probs != nullptr && out != nullptr, "the probability or output pointer is NULL");
  LSIM_REQUIRE(batch >= 1 && batch <= (1ll << 31), "batch must be between 1 and 2^31");
  LSIM_REQUIRE((uint64_t)batch << n <= (1ull << 38), "batch x 2^n entries exceed 2^38");
  const int c = chunk_bits(n, nkeep, keep);
  LSIM_REQUIRE(c == 0 || workspace != nullptr,
               "this keep set needs a workspace of lsim_marginal_workspace_bytes bytes");
  hipStream_t s = as_stream(stream);
  const uint64_t total = (uint64_t)batch << n;
  const double* sums = probs;
  if (c > 0) {
    const uint64_t chunks = total >> c;
    if (c <= 7) {
      k_chunk_sums_small<<<(uint32_t)((chunks + kBlock - 1) / kBlock), kBlock, 0, s>>>(probs, chunks, 1 << c,
                                                                                   workspace);
    } else {
      k_chunk_sums_tree<<<(uint32_t)chunks, 64, 0, s>>>(probs, 1 << (c - 7), workspace);
    }
    sums = workspace;
  }
  // chunk-index bits: kept qubit q sits at q - c (q >= c, since qubits below c are dropped)
  MarginalBits kb = {};
  uint64_t keepmask = 0;
  for (int i = 0; i < nkeep; ++i) {
    kb.pos[i] = (uint32_t)(keep[i] - c);
    keepmask |= 1ull << (keep[i] - c);
  }
  const uint32_t cbits = (uint32_t)(n - c);
  const uint64_t drop = ((cbits == 64 ? ~0ull : (1ull << cbits) - 1)) & ~keepmask;
  const uint64_t per = 1ull << (cbits - nkeep);
  const uint64_t outputs = (uint64_t)batch << nkeep;
  k_ordered_accumulate<<<(uint32_t)((outputs + kBlock - 1) / kBlock), kBlock, 0, s>>>(sums, cbits, nkeep, kb, drop,
                                                                                    per, outputs, out);
  return check_launch("marginal");
}

}  // extern "C"
