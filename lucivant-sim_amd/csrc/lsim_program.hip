// Whole-circuit programs on LDS-resident states: one workgroup runs one circuit instance (a
// circuit of a batch, or one trajectory shot) from |0...0> to its outputs without touching HBM
// in between. A 13-qubit complex128 state is 128 KiB and fits the 160 KiB of LDS of a gfx950
// CU; a 14-qubit complex64 state likewise.
//
// The program is a list of ops interpreted by every thread of the workgroup in lockstep:
//  - DENSE / DIAG: a k <= 5 qubit operator from the matrix pool (engines.py:38-119);
//  - MIX: a Pauli-mixture channel, branch chosen from one draw against cumulative probabilities
//    formed in the reference's order (engines.py:162-170);
//  - KRAUS: a Kraus channel on <= 2 qubits: branch weights from the reduced density matrix,
//    first branch whose running sum passes the draw, applied as K / sqrt(w) (engines.py:171-183);
//  - MEASURE / RESET: p1 reduced in float64, outcome = draw < p1, collapse and divide by
//    sqrt(p), optional readout flip from the next draw (engines.py:185-205,
//    simulator.py:300-306); RESET also flips a 1 back to 0;
//  - COND: evaluates an if/else condition on the classical bits into a flag bit; every op carries
//    a guard (flag mask and value) and runs only where it holds (simulator.py:311-312);
//  - PHASE: a global phase (engines.py:156-157).
// Decisions are taken by thread 0 in float64 and broadcast through LDS; draws come from a
// per-instance slice of a host-drawn buffer, so shot k reads exactly the reference's stream
// positions. Outputs: the classical bits, and optionally |psi|^2, the state, or Pauli
// expectations.
#include "lsim_internal.h"

namespace lsim {

enum : uint8_t {
  OP_DENSE = 0,
  OP_DIAG = 1,
  OP_MIX = 2,
  OP_KRAUS = 3,
  OP_MEASURE = 4,
  OP_RESET = 5,
  OP_COND = 6,
  OP_PHASE = 7,
};

constexpr int kProgThreads = 256;

__device__ __forceinline__ bool guard_ok(const lsim_prog_op& op, uint32_t flags) {
  return (flags & op.guard_mask) == op.guard_value;
}

// Apply a dense K-qubit matrix U (global or LDS) to the LDS state of 2^n amplitudes.
template <int K, typename R>
__device__ __forceinline__ void lds_dense_k(typename Cx<R>::V* s, int n, const uint8_t* t,
                                            const double2* __restrict__ U) {
  using V = typename Cx<R>::V;
  constexpr int D = 1 << K;
  uint32_t sorted[K];
#pragma unroll
  for (int i = 0; i < K; ++i) sorted[i] = t[i];
#pragma unroll
  for (int i = 1; i < K; ++i)
#pragma unroll
    for (int j = i; j > 0; --j)
      if (sorted[j - 1] > sorted[j]) {
        const uint32_t x = sorted[j];
        sorted[j] = sorted[j - 1];
        sorted[j - 1] = x;
      }
  uint32_t offs[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < K; ++i)
      if ((c >> (K - 1 - i)) & 1) o |= 1u << t[i];
    offs[c] = o;
  }
  const uint32_t groups = 1u << (n - K);
  for (uint32_t g = threadIdx.x; g < groups; g += blockDim.x) {
    uint32_t base = g;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const uint32_t b = sorted[i];
      base = ((base >> b) << (b + 1)) | (base & ((1u << b) - 1u));
    }
    V in[D];
#pragma unroll
    for (int c = 0; c < D; ++c) in[c] = s[base + offs[c]];
    // one output row at a time for wide operators: fully unrolled, the 2^K x 2^K operator reads
    // would be scheduled ahead and spill (this kernel holds every op kind at once)
    constexpr int kRowUnroll = K >= 3 ? 1 : D;
#pragma unroll kRowUnroll
    for (int r = 0; r < D; ++r) {
      R re = 0, im = 0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        const double2 u = U[r * D + c];
        re = fma((R)u.x, in[c].x, re);
        re = fma((R)-u.y, in[c].y, re);
        im = fma((R)u.x, in[c].y, im);
        im = fma((R)u.y, in[c].x, im);
      }
      V o;
      o.x = re;
      o.y = im;
      s[base + offs[r]] = o;
    }
  }
}

// Dense operators in programs have at most 4 qubits (the host refuses wider ones): a 5-qubit
// operator's 32 inputs per group would hold 128 VGPRs and halve the whole interpreter's
// occupancy (4 -> 2 waves per SIMD in complex128), which every program pays.
template <typename R>
__device__ __forceinline__ void lds_dense(typename Cx<R>::V* s, int n, int k, const uint8_t* t,
                                          const double2* __restrict__ U) {
  switch (k) {
    case 1: lds_dense_k<1, R>(s, n, t, U); break;
    case 2: lds_dense_k<2, R>(s, n, t, U); break;
    case 3: lds_dense_k<3, R>(s, n, t, U); break;
    default: lds_dense_k<4, R>(s, n, t, U); break;
  }
}

template <typename R>
__device__ __forceinline__ void lds_diag(typename Cx<R>::V* s, int n, int k, const uint8_t* t,
                                         const double2* __restrict__ d) {
  using V = typename Cx<R>::V;
  const uint32_t size = 1u << n;
  for (uint32_t i = threadIdx.x; i < size; i += blockDim.x) {
    uint32_t m = 0;
#pragma unroll
    for (int q = 0; q < 5; ++q)
      if (q < k) m |= ((i >> t[q]) & 1u) << (k - 1 - q);
    const double2 f = d[m];
    const V a = s[i];
    V o;
    o.x = (R)f.x * a.x - (R)f.y * a.y;
    o.y = (R)f.x * a.y + (R)f.y * a.x;
    s[i] = o;
  }
}

// Workgroup sum of NV doubles per thread; every thread gets the totals in `red`.
template <int NV>
__device__ __forceinline__ void wg_sum(double (&v)[NV], double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const double w = wave_sum(v[i]);
    if (lane == 0) red[wave * NV + i] = w;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0;
    for (int w = 0; w < waves; ++w) s += red[w * NV + threadIdx.x];
    red[kProgThreads / 64 * NV + threadIdx.x] = s;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = red[kProgThreads / 64 * NV + i];
  __syncthreads();
}

// Reduced density matrix on K <= 2 target qubits (row-major D x D complex, Hermitian). K is a
// template parameter so that the accumulator indices are constants and stay in registers.
template <int K, typename R>
__device__ void lds_rdm_k(const typename Cx<R>::V* s, int n, const uint8_t* t, double* red,
                          double2* rho) {
  using V = typename Cx<R>::V;
  constexpr int D = 1 << K;
  constexpr int NV = D * (D + 1);  // (re, im) of sum a_r conj(a_c) for r <= c
  double acc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) acc[i] = 0;
  const uint32_t groups = 1u << (n - K);
  uint32_t lo = t[0], hi = t[0];
  if constexpr (K == 2) {
    lo = t[0] < t[1] ? t[0] : t[1];
    hi = t[0] < t[1] ? t[1] : t[0];
  }
  uint32_t offs[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < K; ++i)
      if ((c >> (K - 1 - i)) & 1) o |= 1u << t[i];
    offs[c] = o;
  }
  for (uint32_t g = threadIdx.x; g < groups; g += blockDim.x) {
    uint32_t base = ((g >> lo) << (lo + 1)) | (g & ((1u << lo) - 1u));
    if constexpr (K == 2) base = ((base >> hi) << (hi + 1)) | (base & ((1u << hi) - 1u));
    double x[D], y[D];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      const V a = s[base | offs[c]];
      x[c] = (double)a.x;
      y[c] = (double)a.y;
    }
    int v = 0;
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int c = r; c < D; ++c) {
        acc[v++] += x[r] * x[c] + y[r] * y[c];
        acc[v++] += y[r] * x[c] - x[r] * y[c];
      }
  }
  wg_sum<NV>(acc, red);
  if (threadIdx.x == 0) {
    int v = 0;
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int c = r; c < D; ++c) {
        rho[r * D + c] = make_double2(acc[v], r == c ? 0.0 : acc[v + 1]);
        rho[c * D + r] = make_double2(acc[v], r == c ? 0.0 : -acc[v + 1]);
        v += 2;
      }
  }
}

template <typename R>
__device__ void lds_rdm(const typename Cx<R>::V* s, int n, int k, const uint8_t* t, double* red,
                        double2* rho) {
  if (k == 1)
    lds_rdm_k<1, R>(s, n, t, red, rho);
  else
    lds_rdm_k<2, R>(s, n, t, red, rho);
}

struct ProgShared {
  double red[(kProgThreads / 64 + 1) * 20];
  double2 rho[16];
  double2 mat[16];  // a normalised Kraus operator chosen by thread 0
  int choice;       // -1: nothing to apply
  int outcome;
  double scale;
};

template <typename R>
__global__ void __launch_bounds__(kProgThreads)
    k_program(int n, const lsim_prog_op* __restrict__ ops, const int* __restrict__ op_start,
              const int* __restrict__ op_count, const double2* __restrict__ mats,
              const double* __restrict__ reals, const double* __restrict__ draws,
              const int* __restrict__ draw_start, unsigned long long* __restrict__ clbits_out,
              double* __restrict__ probs_out, typename Cx<R>::V* __restrict__ state_out,
              int nterms, const unsigned long long* __restrict__ pauli_x,
              const unsigned long long* __restrict__ pauli_z, double* __restrict__ expect_out,
              double* __restrict__ decisions) {
  using V = typename Cx<R>::V;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  ProgShared& sh = *reinterpret_cast<ProgShared*>(lds_raw);
  V* s = reinterpret_cast<V*>(lds_raw + ((sizeof(ProgShared) + 15) & ~size_t(15)));
  const int b = blockIdx.x;
  const uint32_t size = 1u << n;
  for (uint32_t i = threadIdx.x; i < size; i += blockDim.x) {
    V z;
    z.x = i == 0 ? (R)1 : (R)0;
    z.y = 0;
    s[i] = z;
  }
  __syncthreads();
  const lsim_prog_op* prog = ops + op_start[b];
  const int count = op_count[b];
  const double* dr = draws != nullptr ? draws + draw_start[b] : nullptr;
  // optional record of every state-dependent decision, laid out as the draws (two doubles per
  // draw column: the probability the draw was compared with, and the branch taken)
  double* rec = decisions != nullptr ? decisions + 2 * (size_t)draw_start[b] : nullptr;
  unsigned long long clbits = 0;
  uint32_t flags = 0;
  for (int p = 0; p < count; ++p) {
    // the op is the same for the whole workgroup: read it into scalar registers, so that every
    // field, and every address built from one (the operator's matrix above all), is uniform
    lsim_prog_op op;
    {
      static_assert(sizeof(lsim_prog_op) % 4 == 0, "ops are whole 32-bit words");
      const uint32_t* src = reinterpret_cast<const uint32_t*>(prog + p);
      uint32_t* dst = reinterpret_cast<uint32_t*>(&op);
#pragma unroll
      for (int w = 0; w < (int)(sizeof(lsim_prog_op) / 4); ++w) dst[w] = __builtin_amdgcn_readfirstlane(src[w]);
    }
    if (!guard_ok(op, flags)) continue;
    switch (op.kind) {
      case OP_DENSE:
        lds_dense<R>(s, n, op.k, op.t, mats + op.mat);
        break;
      case OP_DIAG:
        lds_diag<R>(s, n, op.k, op.t, mats + op.mat);
        break;
      case OP_PHASE:
        for (uint32_t i = threadIdx.x; i < size; i += blockDim.x) {
          const double2 f = mats[op.mat];
          const V a = s[i];
          V o;
          o.x = (R)f.x * a.x - (R)f.y * a.y;
          o.y = (R)f.x * a.y + (R)f.y * a.x;
          s[i] = o;
        }
        break;
      case OP_MIX: {
        // reals[aux .. aux+nb): cumulative probabilities; reals[aux+nb ..): matrix offset of
        // each branch as a double (-1: the identity)
        const double r = dr[op.draw];
        int pick = -1;
        for (int j = 0; j < op.nb; ++j)
          if (r < reals[op.aux + j]) {
            pick = j;
            break;
          }
        const int off = pick < 0 ? -1 : (int)reals[op.aux + op.nb + pick];
        if (off >= 0) lds_dense<R>(s, n, op.k, op.t, mats + off);
        break;
      }
      case OP_KRAUS: {
        lds_rdm<R>(s, n, op.k, op.t, sh.red, sh.rho);
        __syncthreads();
        if (threadIdx.x == 0) {
          const int D = 1 << op.k;
          const double r = dr[op.draw];
          double acc = 0, w = 0;
          int pick = -1;
          for (int j = 0; j < op.nb; ++j) {
            const double2* K = mats + op.mat + j * D * D;
            w = 0;
            for (int row = 0; row < D; ++row)
              for (int a = 0; a < D; ++a) {
                double2 ka = K[row * D + a];
                for (int c = 0; c < D; ++c) {
                  const double2 kc = K[row * D + c];
                  const double2 rr = sh.rho[a * D + c];
                  // Re(ka * rr * conj(kc))
                  const double tx = ka.x * rr.x - ka.y * rr.y, ty = ka.x * rr.y + ka.y * rr.x;
                  w += tx * kc.x + ty * kc.y;
                }
              }
            acc += w;
            if (r < acc && w > 0) {
              pick = j;
              break;
            }
          }
          if (pick < 0 && w > 0) pick = op.nb - 1;
          sh.choice = pick;
          if (rec != nullptr) {
            rec[2 * op.draw] = acc;
            rec[2 * op.draw + 1] = pick;
          }
          if (pick >= 0) {
            const double norm = sqrt(w);
            const double2* K = mats + op.mat + pick * D * D;
            for (int e = 0; e < D * D; ++e) sh.mat[e] = make_double2(K[e].x / norm, K[e].y / norm);
          }
        }
        __syncthreads();
        if (sh.choice >= 0) lds_dense<R>(s, n, op.k, op.t, sh.mat);
        break;
      }
      case OP_MEASURE:
      case OP_RESET: {
        const uint32_t q = op.t[0];
        double v[1] = {0};
        for (uint32_t i = threadIdx.x; i < size; i += blockDim.x)
          if ((i >> q) & 1u) {
            const V a = s[i];
            v[0] += (double)a.x * a.x + (double)a.y * a.y;
          }
        wg_sum<1>(v, sh.red);
        const double p1 = fmin(fmax(v[0], 0.0), 1.0);
        const int outcome = dr[op.draw] < p1 ? 1 : 0;
        if (rec != nullptr && threadIdx.x == 0) {
          rec[2 * op.draw] = p1;
          rec[2 * op.draw + 1] = outcome;
        }
        const double p = outcome ? p1 : 1.0 - p1;
        const double root = sqrt(p);
        for (uint32_t i = threadIdx.x; i < size; i += blockDim.x) {
          V a = s[i];
          if ((int)((i >> q) & 1u) != outcome) {
            a.x = 0;
            a.y = 0;
          } else if (p > 0) {
            a.x = (R)((double)a.x / root);
            a.y = (R)((double)a.y / root);
          }
          s[i] = a;
        }
        __syncthreads();
        if (op.kind == OP_RESET) {
          if (outcome) {
            // X on q: the surviving |1> half moves to |0>
            for (uint32_t g = threadIdx.x; g < size / 2; g += blockDim.x) {
              const uint32_t i0 = ((g >> q) << (q + 1)) | (g & ((1u << q) - 1u));
              const uint32_t i1 = i0 | (1u << q);
              const V a = s[i1];
              s[i1] = s[i0];
              s[i0] = a;
            }
          }
        } else {
          int o = outcome;
          if (op.aux >= 0) {  // readout error: reals[aux] = p01, reals[aux+1] = p10
            const double pr = o ? reals[op.aux + 1] : reals[op.aux];
            if (dr[op.draw + 1] < pr) o ^= 1;
          }
          if (op.clbit >= 0) clbits = (clbits & ~(1ull << op.clbit)) | ((unsigned long long)o << op.clbit);
        }
        break;
      }
      case OP_COND: {
        // condition: (clbits & mask) == pattern, stored as two 64-bit words in the op
        const bool holds = ((clbits & op.cmask) == op.cpattern) != (op.negate != 0);
        if (holds)
          flags |= 1u << op.flag;
        else
          flags &= ~(1u << op.flag);
        break;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && clbits_out != nullptr) clbits_out[b] = clbits;
  if (probs_out != nullptr)
    for (uint32_t i = threadIdx.x; i < size; i += blockDim.x) {
      const V a = s[i];
      probs_out[(size_t)b * size + i] = (double)a.x * a.x + (double)a.y * a.y;
    }
  if (state_out != nullptr)
    for (uint32_t i = threadIdx.x; i < size; i += blockDim.x) state_out[(size_t)b * size + i] = s[i];
  for (int tt = 0; tt < nterms; ++tt) {
    const uint64_t X = pauli_x[(size_t)b * nterms + tt], Z = pauli_z[(size_t)b * nterms + tt];
    const int w = __popcll(X & Z) & 3;
    double v[1] = {0};
    for (uint32_t j = threadIdx.x; j < size; j += blockDim.x) {
      const V a = s[j], pp = s[j ^ X];
      const double zx = (double)pp.x * a.x + (double)pp.y * a.y;
      const double zy = (double)pp.x * a.y - (double)pp.y * a.x;
      double re;
      switch (w) {
        case 0: re = zx; break;
        case 1: re = -zy; break;
        case 2: re = -zx; break;
        default: re = zy; break;
      }
      v[0] += (__popcll(j & Z) & 1) ? -re : re;
    }
    wg_sum<1>(v, sh.red);
    if (threadIdx.x == 0) expect_out[(size_t)b * nterms + tt] = v[0];
  }
}

template <typename R>
static size_t lds_bytes(int n) {
  return ((sizeof(ProgShared) + 15) & ~size_t(15)) + (sizeof(typename Cx<R>::V) << n);
}

}  // namespace lsim

using namespace lsim;

extern "C" {

int lsim_program_max_qubits(int dtype) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  int limit = 0, optin = 0;
  if (hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return 0;
  if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) == hipSuccess &&
      optin > limit)
    limit = optin;
  if (limit > 160 * 1024) limit = 160 * 1024;  // gfx950: 160 KiB of LDS per CU
  for (int n = 16; n >= 1; --n) {
    const size_t need = dtype == LSIM_C64 ? lds_bytes<float>(n) : lds_bytes<double>(n);
    if (need <= (size_t)limit) return n;
  }
  return 0;
}

int lsim_program_run(int dtype, int n, long long instances, const lsim_prog_op* ops,
                     const int* op_start, const int* op_count, const void* mats,
                     const double* reals, const double* draws, const int* draw_start,
                     unsigned long long* clbits, double* probs, void* state, int nterms,
                     const unsigned long long* pauli_x, const unsigned long long* pauli_z,
                     double* expect, double* decisions, void* stream) {
  LSIM_REQUIRE(dtype == LSIM_C128 || dtype == LSIM_C64, "dtype must be LSIM_C128 or LSIM_C64");
  LSIM_REQUIRE(n >= 1 && n <= 16, "program states have 1 to 16 qubits, not %d", n);
  LSIM_REQUIRE(instances >= 1 && instances <= 0x7fffffff, "instances must be positive");
  LSIM_REQUIRE(ops != nullptr && op_start != nullptr && op_count != nullptr && mats != nullptr,
               "program arrays are NULL");
  LSIM_REQUIRE(nterms >= 0 && nterms <= LSIM_MAX_TERMS, "nterms must be between 0 and %d",
               LSIM_MAX_TERMS);
  LSIM_REQUIRE(nterms == 0 || (pauli_x && pauli_z && expect), "Pauli arrays are NULL");
  LSIM_REQUIRE(decisions == nullptr || (draws && draw_start),
               "decisions are recorded per draw: give the draws too");
  const size_t bytes = dtype == LSIM_C64 ? lds_bytes<float>(n) : lds_bytes<double>(n);
  hipStream_t s = as_stream(stream);
  if (dtype == LSIM_C128) {
    if (hipFuncSetAttribute((const void*)k_program<double>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
      return fail(LSIM_EHIP, "LDS program: %zu bytes of LDS are not available", bytes);
    k_program<double><<<(unsigned)instances, kProgThreads, bytes, s>>>(
        n, ops, op_start, op_count, (const double2*)mats, reals, draws, draw_start, clbits, probs,
        (double2*)state, nterms, pauli_x, pauli_z, expect, decisions);
  } else {
    if (hipFuncSetAttribute((const void*)k_program<float>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess)
      return fail(LSIM_EHIP, "LDS program: %zu bytes of LDS are not available", bytes);
    k_program<float><<<(unsigned)instances, kProgThreads, bytes, s>>>(
        n, ops, op_start, op_count, (const double2*)mats, reals, draws, draw_start, clbits, probs,
        (float2*)state, nterms, pauli_x, pauli_z, expect, decisions);
  }
  return check_launch("LDS program");
}

}  // extern "C"
