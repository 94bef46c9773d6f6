// Tile passes: many small operators applied in one read and one write of the state.
//
// A fused dense pass (lsim_sv.hip) applies one operator of up to 5 qubits per sweep of HBM. A
// tile pass picks kTileBits state bits (the "tile bits", always including the lowest ones, so
// tiles are read in runs of contiguous amplitudes). A tile is the 2^kTileBits amplitudes that
// agree on all other bits; every tile is read once, a list of 1- and 2-qubit operators on tile
// bits is applied to it, and it is written back. Any gates whose qubits are all tile bits can run
// in one pass, so a circuit needs far fewer sweeps of HBM (the host planner is plan.plan_tiles).
//
// Register blocking. One workgroup of 256 threads owns one tile at a time; each thread holds 16
// of its amplitudes in registers. In a layout R (4 of the 12 tile positions) thread t holds the
// entries whose other 8 positions spell t (ascending) and whose R positions spell k = 0..15, in
// register k. A dense operator whose positions lie in R mixes registers of one thread only: no
// LDS, no barrier. A diagonal operator runs in any layout. Only when the next dense operator
// needs positions outside R does the tile go through LDS into the next layout. The launcher
// splits the operator list into such segments, each with the 4 positions its operators use
// first. The layouts the tile is read and written in keep positions 0..3 (the lowest state bits)
// on the thread bits, so 16 consecutive lanes move 16 consecutive amplitudes.
//
// Operators come from a host-built table in the kernel arguments: wavefront-uniform, so each
// matrix is read with scalar loads. LDS holds one tile (64 KiB for complex128): two workgroups
// share a CU, and one reads or writes HBM while the other computes.
#include <string.h>

#include "lsim_internal.h"

namespace lsim {

constexpr int kTileBits = 12;
constexpr int kTileSize = 1 << kTileBits;
constexpr int kRegs = kTileSize / kBlock;  // 16 amplitudes per thread: 4 register positions
constexpr int kMaxTileOps = LSIM_MAX_TILE_OPS;
constexpr int kMaxSegs = kMaxTileOps + 2;
// Workgroups of a tile launch (per state): four per CU of the 256 on MI355X (two resident at a
// time), each working through its share of the tiles with the next one in flight.
constexpr uint64_t kTileGrid = 1024;

// Every field is 32-bit: the kernel indexes the arrays with wavefront-uniform indices, and
// 32-bit kernel-argument words are read with scalar loads (narrower ones would take vector
// loads, whose waits would also wait for the prefetched amplitudes).
struct TileProgram {
  uint32_t bits[kTileBits];       // state bit of tile position i, ascending
  uint32_t layout[kMaxSegs];      // segment s: mask of its 4 register positions
  uint32_t first[kMaxSegs + 1];   // segment s runs operators first[s] .. first[s+1]-1
  uint32_t kind[kMaxTileOps];     // LSIM_TILE_DENSE or LSIM_TILE_DIAG
  uint32_t k[kMaxTileOps];        // 1 or 2
  uint32_t a[kMaxTileOps];        // dense: register slot of the most significant matrix bit;
  uint32_t b[kMaxTileOps];        //   diagonal: tile positions; b unused for k = 1
  uint32_t mat[kMaxTileOps];      // element offset of the operator in the complex128 pool
  int nsegs;
};

// acc + u * a in four fused multiply-adds
template <typename R, typename V>
__device__ __forceinline__ V mac(V acc, double2 u, V a) {
  const R ux = (R)u.x, uy = (R)u.y;
  acc.x = fma(ux, a.x, acc.x);
  acc.x = fma(-uy, a.y, acc.x);
  acc.y = fma(ux, a.y, acc.y);
  acc.y = fma(uy, a.x, acc.y);
  return acc;
}

// Tile index of register k of thread t in layout R: t's bits on the other positions (ascending),
// k's on R's.
__device__ __forceinline__ uint32_t entry(uint32_t t, uint32_t k, uint32_t R) {
  uint32_t e = 0;
#pragma unroll
  for (int p = 0; p < kTileBits; ++p) {
    if ((R >> p) & 1u) {
      e |= (k & 1u) << p;
      k >>= 1;
    } else {
      e |= (t & 1u) << p;
      t >>= 1;
    }
  }
  return e;
}

// entry(t, 0, R) for the thread's own t, cheaper: the positions outside R are found on the
// scalar unit, the wavefront's bits of t (6 and 7) are uniform, and only the 6 lane bits are
// deposited per lane.
__device__ __forceinline__ uint32_t thread_part(uint32_t t, uint32_t R) {
  uint32_t free = ~R & ((1u << kTileBits) - 1u);
  uint32_t e = 0;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const uint32_t p = (uint32_t)__builtin_ctz(free);
    free &= free - 1u;
    e |= ((t >> i) & 1u) << p;
  }
  const uint32_t w = __builtin_amdgcn_readfirstlane(t >> 6);
  uint32_t we = 0;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const uint32_t p = (uint32_t)__builtin_ctz(free);
    free &= free - 1u;
    we |= ((w >> i) & 1u) << p;
  }
  return e | we;
}

// State offset of tile index e: bit i of e at state bit bits[i].
__device__ __forceinline__ uint32_t deposit(uint32_t e, const uint32_t* bits) {
  uint32_t o = 0;
#pragma unroll
  for (int i = 0; i < kTileBits; ++i) o |= ((e >> i) & 1u) << bits[i];
  return o;
}

// A 2 x 2 operator on register slot J: registers k and k | 2^J for every k with bit J clear.
template <int J, typename R, typename V>
__device__ __forceinline__ void regs1(V (&a)[kRegs], const double2* __restrict__ M) {
  const double2 m00 = M[0], m01 = M[1], m10 = M[2], m11 = M[3];
#pragma unroll
  for (int k = 0; k < kRegs; ++k) {
    if (k & (1 << J)) continue;
    const V x0 = a[k], x1 = a[k | (1 << J)];
    V y0 = {0, 0}, y1 = {0, 0};
    a[k] = mac<R>(mac<R>(y0, m00, x0), m01, x1);
    a[k | (1 << J)] = mac<R>(mac<R>(y1, m10, x0), m11, x1);
  }
}

// A 4 x 4 operator on register slots H (the most significant matrix bit) and L, H > L: quad by
// quad, in place; the matrix is read from the scalar cache for each quad.
template <int H, int L, typename R, typename V>
__device__ __forceinline__ void regs2(V (&a)[kRegs], const double2* __restrict__ M) {
#pragma unroll
  for (int k = 0; k < kRegs; ++k) {
    if (k & ((1 << H) | (1 << L))) continue;
    const int i0 = k, i1 = k | (1 << L), i2 = k | (1 << H), i3 = k | (1 << H) | (1 << L);
    const V x0 = a[i0], x1 = a[i1], x2 = a[i2], x3 = a[i3];
    V y[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      V acc = {0, 0};
      acc = mac<R>(acc, M[r * 4], x0);
      acc = mac<R>(acc, M[r * 4 + 1], x1);
      acc = mac<R>(acc, M[r * 4 + 2], x2);
      acc = mac<R>(acc, M[r * 4 + 3], x3);
      y[r] = acc;
    }
    a[i0] = y[0];
    a[i1] = y[1];
    a[i2] = y[2];
    a[i3] = y[3];
  }
}

template <typename R, typename V>
__device__ __forceinline__ void dense_op(V (&a)[kRegs], int k, int h, int l, const double2* M) {
  if (k == 1) {
    switch (h) {
      case 0: regs1<0, R>(a, M); break;
      case 1: regs1<1, R>(a, M); break;
      case 2: regs1<2, R>(a, M); break;
      default: regs1<3, R>(a, M); break;
    }
    return;
  }
  // the launcher lists the higher slot first (h > l)
  switch (h * 4 + l) {
    case 4: regs2<1, 0, R>(a, M); break;
    case 8: regs2<2, 0, R>(a, M); break;
    case 9: regs2<2, 1, R>(a, M); break;
    case 12: regs2<3, 0, R>(a, M); break;
    case 13: regs2<3, 1, R>(a, M); break;
    default: regs2<3, 2, R>(a, M); break;
  }
}

// The register part of a tile index in a layout: register k's bits at the layout's 4 positions,
// given as the one-bit masks P[0] < P[1] < P[2] < P[3] (wavefront-uniform).
__device__ __forceinline__ uint32_t kpart(int k, const uint32_t (&P)[4]) {
  return ((k & 1) ? P[0] : 0u) | ((k & 2) ? P[1] : 0u) | ((k & 4) ? P[2] : 0u) | ((k & 8) ? P[3] : 0u);
}

__device__ __forceinline__ void layout_bits(uint32_t layout, uint32_t (&P)[4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    P[i] = layout & (~layout + 1u);  // lowest set bit
    layout ^= P[i];
  }
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// State offset of the register part of a tile index: the deposit of kpart(k, P) (wavefront-
// uniform). deposit() maps disjoint bit sets to disjoint bit sets, so an entry's offset is the
// thread part's offset OR this.
__device__ __forceinline__ uint32_t kofs(int k, const uint32_t (&P)[4], const uint32_t* bits) {
  return deposit(kpart(k, P), bits);
}

// The tile's state offset: the tile number q with zeros inserted at the tile bits.
__device__ __forceinline__ uint64_t tile_base(uint64_t q, const uint32_t* bits) {
#pragma unroll
  for (int i = 0; i < kTileBits; ++i) {
    const uint32_t p = bits[i];
    q = ((q >> p) << (p + 1)) | (q & ((1ull << p) - 1));
  }
  return q;
}

// Operator j of the program on the tile in registers (layout: thread part tpart, register
// positions P).
template <typename R, typename V>
__device__ __forceinline__ void apply_op(V (&a)[kRegs], const TileProgram& prog,
                                         const double2* __restrict__ mats, int j, uint32_t tpart,
                                         const uint32_t (&P)[4]) {
  const double2* M = mats + uniform(prog.mat[j]);
  const uint32_t kind = uniform(prog.kind[j]), kk = uniform(prog.k[j]);
  const uint32_t pa = uniform(prog.a[j]), pb = uniform(prog.b[j]);
  if (kind == LSIM_TILE_DENSE) {
    dense_op<R>(a, (int)kk, (int)pa, (int)pb, M);
    return;
  }
  // diagonal: the entry's own bits pick the factor
  const bool two = kk == 2;
  const double2 d0 = M[0], d1 = M[1];
  const double2 d2 = two ? M[2] : d0, d3 = two ? M[3] : d1;
#pragma unroll
  for (int k = 0; k < kRegs; ++k) {
    const uint32_t e = tpart | kpart(k, P);
    // factor index (hb, lb): selects written out, so no per-lane indexed array
    const bool hb = two && ((e >> pa) & 1u);
    const bool lb = two ? ((e >> pb) & 1u) : ((e >> pa) & 1u);
    double2 d;
    d.x = hb ? (lb ? d3.x : d2.x) : (lb ? d1.x : d0.x);
    d.y = hb ? (lb ? d3.y : d2.y) : (lb ? d1.y : d0.y);
    V y = {0, 0};
    a[k] = mac<R>(y, d, a[k]);
  }
}

// Every segment of the program on one tile held in registers (a, in the load layout on entry),
// regrouping through LDS between segments; on return (tpart, P) describe the final layout.
template <typename R, typename V>
__device__ __forceinline__ void run_program(V (&a)[kRegs], V* tile, const TileProgram& prog,
                                            const double2* __restrict__ mats, uint32_t t,
                                            uint32_t& tpart, uint32_t (&P)[4]) {
  uint32_t layout = uniform(prog.layout[0]);
  tpart = thread_part(t, layout);
  layout_bits(layout, P);
  const int nsegs = (int)uniform((uint32_t)prog.nsegs);
  for (int s = 0; s < nsegs; ++s) {
    if (s > 0) {
      // into segment s's layout (the planner starts a segment only where the layout changes)
      // through LDS; the thread and register parts of an index have disjoint bits, so an
      // entry's element offset is their sum
      // A wave in an exchange holds up its workgroup at the barriers, while the other
      // workgroup on the compute unit has arithmetic to fill the gaps: the exchanging waves issue
      // first until their reads are out (2% per pass at 26 qubits).
      __builtin_amdgcn_s_setprio(1);
      __syncthreads();  // the previous exchange (or the previous tile's) has been read
      {
        V* mine = tile + tpart;
#pragma unroll
        for (int k = 0; k < kRegs; ++k) mine[kpart(k, P)] = a[k];
      }
      layout = uniform(prog.layout[s]);
      tpart = thread_part(t, layout);
      layout_bits(layout, P);
      __syncthreads();
      {
        const V* mine = tile + tpart;
#pragma unroll
        for (int k = 0; k < kRegs; ++k) a[k] = mine[kpart(k, P)];
      }
      __builtin_amdgcn_s_setprio(0);
    }
    const int j1 = (int)uniform(prog.first[s + 1]);
    // two operators per iteration: the register allocator can then let the first write where the
    // second reads, instead of moving every amplitude back at the end of each operator
    int j = (int)uniform(prog.first[s]);
    for (; j + 1 < j1; j += 2) {
      apply_op<R>(a, prog, mats, j, tpart, P);
      apply_op<R>(a, prog, mats, j + 1, tpart, P);
    }
    if (j < j1) apply_op<R>(a, prog, mats, j, tpart, P);
  }
}

// Persistent workgroups: workgroup w takes tiles w, w + G, w + 2G, ... (G = gridDim.x). While one
// tile is computed in registers, the next one's amplitudes are already being loaded into a second
// register set, so HBM traffic overlaps the FP64 and LDS work instead of alternating with it.
// (The barriers of the LDS exchanges wait only for LDS: the prefetch stays in flight.)
template <typename R, bool NT>
__global__ void __launch_bounds__(kBlock)
    k_tile(typename Cx<R>::V* __restrict__ states, uint32_t nbits, TileProgram prog,
           const double2* __restrict__ mats, uint32_t batch_base) {
  using V = typename Cx<R>::V;
  __shared__ V tile[kTileSize];
  V* __restrict__ st = states + ((size_t)(batch_base + blockIdx.y) << nbits);
  static_assert(kMaxBits <= 32, "tile offsets are 32-bit");
  const uint64_t tiles = 1ull << (nbits - kTileBits);
  const uint32_t t = threadIdx.x;
  // offsets of this thread's entries in the layouts tiles are read and written in: a per-thread
  // part and a wavefront-uniform part per register. The last segment's layout is the one every
  // tile is written back in (an exchange runs exactly where the layout changes).
  const uint32_t load_layout = uniform(prog.layout[0]);
  const uint32_t store_layout = uniform(prog.layout[uniform((uint32_t)prog.nsegs) - 1]);
  uint32_t P0[4], PS[4];
  layout_bits(load_layout, P0);
  layout_bits(store_layout, PS);
  const uint32_t toff = deposit(entry(t, 0, load_layout), prog.bits);
  const uint32_t tsoff = deposit(entry(t, 0, store_layout), prog.bits);
  uint32_t lofs[kRegs], sofs[kRegs];
#pragma unroll
  for (int k = 0; k < kRegs; ++k) {
    lofs[k] = uniform(kofs(k, P0, prog.bits));
    sofs[k] = uniform(kofs(k, PS, prog.bits));
  }
  uint64_t q = blockIdx.x;
  if (q >= tiles) return;
  V a[kRegs], b[kRegs];
  auto load = [&](V (&dst)[kRegs], uint64_t tq) {
    const V* src = st + (tile_base(tq, prog.bits) | toff);
#pragma unroll
    for (int k = 0; k < kRegs; ++k) dst[k] = load_amp<NT>(src + lofs[k]);
  };
  auto finish = [&](V (&v)[kRegs], uint64_t tq) {
    uint32_t tpart, P[4];
    run_program<R>(v, tile, prog, mats, t, tpart, P);
    V* dst = st + (tile_base(tq, prog.bits) | tsoff);
#pragma unroll
    for (int k = 0; k < kRegs; ++k) store_amp<NT>(dst + sofs[k], v[k]);
  };
  // Vector memory operations complete in order of issue. After a tile's stores are issued, the
  // wait until at most kRegs operations are outstanding (those stores) means the next tile's
  // loads, issued before them, have landed; the copy into `a` keeps one register set live across
  // the loop, so the compiler's own waits know `a` is complete and the LDS exchanges and the
  // arithmetic wait for nothing in flight. (s_waitcnt immediates: vmcnt in bits 3:0 and 15:14;
  // expcnt 7 and lgkmcnt 15 wait for nothing.)
  constexpr int kWaitLoads = ((kRegs & 15) | ((kRegs >> 4) << 14)) | (7 << 4) | (15 << 8);
  constexpr int kWaitAll = (7 << 4) | (15 << 8);
  load(a, q);
  __builtin_amdgcn_s_waitcnt(kWaitAll);
  for (;;) {
    const uint64_t q1 = q + gridDim.x;
    const bool more = q1 < tiles;
    if (more) load(b, q1);
    finish(a, q);
    if (!more) break;
    __builtin_amdgcn_s_waitcnt(kWaitLoads);
    for (int k = 0; k < kRegs; ++k) a[k] = b[k];
    q = q1;
  }
}

// Splits the operators into segments of one register layout each. A layout is 4 tile positions:
// those of the dense operators that come next, in order, as many as fit; the first and last
// layouts avoid positions 0..kLowPositions-1 (reading and writing HBM in contiguous runs).
constexpr int kLowPositions = 4;

// `avoid`: positions a layout that reads or writes HBM leaves on the thread bits -- the tile's
// run of lowest state bits (at least kLowPositions of them, at most 8), so consecutive lanes
// move consecutive amplitudes.
static uint32_t fill_layout(const lsim_tile_op* ops, int from, int nops, uint32_t need, uint32_t avoid) {
  uint32_t R = need;
  for (int j = from; j < nops && __builtin_popcount(R) < 4; ++j) {
    if (ops[j].kind != LSIM_TILE_DENSE) continue;
    uint32_t want = 1u << ops[j].t[0];
    if (ops[j].k == 2) want |= 1u << ops[j].t[1];
    if (want & avoid) break;
    if (__builtin_popcount(R | want) > 4) break;
    R |= want;
  }
  for (int p = kTileBits - 1; p >= 0 && __builtin_popcount(R) < 4; --p)
    if (!((R >> p) & 1u) && !((avoid >> p) & 1u)) R |= 1u << p;
  return R;
}

static uint32_t op_mask(const lsim_tile_op& o) {
  return (1u << o.t[0]) | (o.k == 2 ? 1u << o.t[1] : 0u);
}

// Reorders the operators into runs that each fit one register layout, so that fewer LDS
// exchanges are needed. Repeatedly, over the sets of 4 tile positions (in the first round, those
// off `hbm` when one lets any operator run), take the set that lets the most remaining operators
// run, preferring sets off `hbm` on ties, and emit those operators in their original order. An
// operator runs when no earlier remaining operator that does not run shares a position with it,
// and, if it is dense, when its positions lie in the set (diagonal ones run in any layout).
// Operators on disjoint positions commute, so the product is unchanged up to rounding.
static int runnable(const uint32_t* masks, const bool* dense, const int* left, int nleft, uint32_t R,
                    int* taken) {
  uint32_t blocked = 0;
  int count = 0;
  for (int i = 0; i < nleft; ++i) {
    const int j = left[i];
    const uint32_t m = masks[j];
    if ((m & blocked) || (dense[j] && (m & ~R))) {
      blocked |= m;
    } else {
      if (taken) taken[count] = j;
      ++count;
    }
  }
  return count;
}

static void schedule_ops(int nops, const lsim_tile_op* ops, uint32_t hbm, lsim_tile_op* out) {
  uint32_t masks[kMaxTileOps];
  bool dense[kMaxTileOps];
  int left[kMaxTileOps], taken[kMaxTileOps], nleft = nops, nout = 0;
  for (int j = 0; j < nops; ++j) {
    masks[j] = (1u << ops[j].t[0]) | (ops[j].k == 2 ? 1u << ops[j].t[1] : 0u);
    dense[j] = ops[j].kind == LSIM_TILE_DENSE;
    left[j] = j;
  }
  bool first = true;
  while (nleft > 0) {
    uint32_t best = 0;
    int best_count = 0;
    bool best_off = false;
    for (int pass = 0; pass < 2 && best_count == 0; ++pass) {
      for (uint32_t R = 0; R < (1u << kTileBits); ++R) {
        if (__builtin_popcount(R) != 4) continue;
        if (pass == 0 && first && (R & hbm)) continue;
        const int count = runnable(masks, dense, left, nleft, R, nullptr);
        const bool off = !(R & hbm);
        if (count > best_count || (count == best_count && count > 0 && off && !best_off)) {
          best = R;
          best_count = count;
          best_off = off;
        }
      }
    }
    int n = runnable(masks, dense, left, nleft, best, taken);
    if (n == 0) taken[n++] = left[0];  // unreachable: a set holding the first operator's positions runs it
    // emit the taken operators and keep the rest in order
    int keep = 0, t = 0;
    for (int i = 0; i < nleft; ++i) {
      if (t < n && left[i] == taken[t]) {
        out[nout++] = ops[taken[t++]];
      } else {
        left[keep++] = left[i];
      }
    }
    nleft = keep;
    first = false;
  }
}

static int make_program(const int* bits, int nops, const lsim_tile_op* ops, TileProgram* prog) {
  memset(prog, 0, sizeof *prog);
  for (int i = 0; i < kTileBits; ++i) prog->bits[i] = (uint8_t)bits[i];
  int run = 0;
  while (run < 8 && bits[run] == run) ++run;
  const uint32_t hbm = (1u << run) - 1u;  // positions kept on the thread bits for HBM access
  int s = 0;
  uint32_t R = fill_layout(ops, 0, nops, 0, hbm);
  prog->layout[0] = (uint16_t)R;
  prog->first[0] = 0;
  for (int j = 0; j < nops; ++j) {
    const lsim_tile_op& o = ops[j];
    if (o.kind == LSIM_TILE_DENSE) {
      uint32_t want = 1u << o.t[0];
      if (o.k == 2) want |= 1u << o.t[1];
      if ((R & want) != want) {
        R = fill_layout(ops, j, nops, want, 0);
        ++s;
        prog->layout[s] = (uint16_t)R;
        prog->first[s] = (uint8_t)j;
      }
      // register slots: the rank of each position among the layout's
      auto slot = [R](int p) { return __builtin_popcount(R & ((1u << p) - 1u)); };
      prog->a[j] = (uint8_t)slot(o.t[0]);
      prog->b[j] = (uint8_t)(o.k == 2 ? slot(o.t[1]) : 0);
      if (o.k == 2 && prog->a[j] < prog->b[j]) return -2;  // the caller lists the higher bit first
    } else {
      prog->a[j] = (uint8_t)o.t[0];
      prog->b[j] = (uint8_t)(o.k == 2 ? o.t[1] : 0);
    }
    prog->kind[j] = (uint8_t)o.kind;
    prog->k[j] = (uint8_t)o.k;
    prog->mat[j] = (uint32_t)o.mat;
  }
  if (R & hbm) {  // written back in a layout off the lowest positions
    ++s;
    prog->layout[s] = (uint16_t)fill_layout(ops, nops, nops, 0, hbm);
    prog->first[s] = (uint8_t)nops;
  }
  prog->nsegs = s + 1;
  prog->first[s + 1] = (uint8_t)nops;
  for (int i = 0; i < prog->nsegs; ++i) {
    if (__builtin_popcount(prog->layout[i]) != 4 || (prog->layout[i] >> kTileBits)) return -1;
    if (i > 0 && prog->layout[i] == prog->layout[i - 1]) return -1;  // the kernel exchanges at every segment
  }
  if ((prog->layout[0] & hbm) || (prog->layout[prog->nsegs - 1] & hbm)) return -1;
  return 0;
}

}  // namespace lsim

using namespace lsim;

extern "C" {

int lsim_sv_apply_tile(void* states, int dtype, int nbits, long long batch, const int* bits,
                       int nops, const lsim_tile_op* ops, const void* mats, void* stream) {
  LSIM_REQUIRE(states != nullptr && mats != nullptr, "the state or matrix pointer is NULL");
  LSIM_REQUIRE(dtype == LSIM_C128 || dtype == LSIM_C64, "dtype must be LSIM_C128 or LSIM_C64");
  LSIM_REQUIRE(nbits >= kTileBits && nbits <= kMaxBits, "tile passes need %d to %d bits, not %d",
               kTileBits, kMaxBits, nbits);
  LSIM_REQUIRE(batch >= 1 && batch <= (1ll << 31), "batch must be between 1 and 2^31");
  LSIM_REQUIRE(bits != nullptr, "the tile bits are NULL");
  LSIM_REQUIRE(nops >= 0 && nops <= kMaxTileOps, "a tile pass takes 0 to %d operators, not %d",
               kMaxTileOps, nops);
  LSIM_REQUIRE(nops == 0 || ops != nullptr, "the operator table is NULL");
  for (int i = 0; i < kTileBits; ++i) {
    LSIM_REQUIRE(bits[i] >= 0 && bits[i] < nbits, "tile bit %d is outside 0..%d", bits[i], nbits - 1);
    LSIM_REQUIRE(i == 0 || bits[i] > bits[i - 1], "tile bits must ascend");
  }
  for (int i = 0; i < kLowPositions; ++i)
    LSIM_REQUIRE(bits[i] == i, "the tile bits must include state bits 0..%d", kLowPositions - 1);
  for (int j = 0; j < nops; ++j) {
    const lsim_tile_op& o = ops[j];
    LSIM_REQUIRE(o.kind == LSIM_TILE_DENSE || o.kind == LSIM_TILE_DIAG, "operator %d: unknown kind %d", j, o.kind);
    LSIM_REQUIRE(o.k == 1 || o.k == 2, "operator %d acts on 1 or 2 tile bits, not %d", j, o.k);
    LSIM_REQUIRE(o.t[0] >= 0 && o.t[0] < kTileBits, "operator %d: tile bit %d is outside 0..%d", j, o.t[0], kTileBits - 1);
    if (o.k == 2)
      LSIM_REQUIRE(o.t[1] >= 0 && o.t[1] < kTileBits && o.t[1] != o.t[0], "operator %d: bad second tile bit %d", j, o.t[1]);
    LSIM_REQUIRE(o.mat >= 0 && o.mat <= 0x7fffffffll, "operator %d: matrix offset %lld", j, o.mat);
  }
  for (int j = 0; j < nops; ++j)
    if (ops[j].kind == LSIM_TILE_DENSE && ops[j].k == 2)
      LSIM_REQUIRE(ops[j].t[0] > ops[j].t[1], "operator %d: a dense 2-bit operator lists its higher tile bit first", j);
  lsim_tile_op ordered[kMaxTileOps];
  {
    int run = 0;
    while (run < 8 && bits[run] == run) ++run;
    schedule_ops(nops, ops, (1u << run) - 1u, ordered);
  }
  TileProgram prog;
  if (make_program(bits, nops, ordered, &prog) != 0)
    return fail(LSIM_EINVAL, "tile pass: the layout plan failed its own checks");
  hipStream_t s = as_stream(stream);
  const uint64_t tiles = 1ull << (nbits - kTileBits);
  const uint32_t grid = (uint32_t)(tiles < kTileGrid ? tiles : kTileGrid);
  const bool nt = dense_streams((uint64_t)batch * (dtype_bytes(dtype) << nbits));
  for (long long b0 = 0; b0 < batch; b0 += kMaxGridY) {
    const uint32_t gy = (uint32_t)(batch - b0 < kMaxGridY ? batch - b0 : kMaxGridY);
    const dim3 g(grid, gy);
    if (dtype == LSIM_C128) {
      if (nt)
        k_tile<double, true><<<g, kBlock, 0, s>>>((double2*)states, nbits, prog, (const double2*)mats, (uint32_t)b0);
      else
        k_tile<double, false><<<g, kBlock, 0, s>>>((double2*)states, nbits, prog, (const double2*)mats, (uint32_t)b0);
    } else {
      if (nt)
        k_tile<float, true><<<g, kBlock, 0, s>>>((float2*)states, nbits, prog, (const double2*)mats, (uint32_t)b0);
      else
        k_tile<float, false><<<g, kBlock, 0, s>>>((float2*)states, nbits, prog, (const double2*)mats, (uint32_t)b0);
    }
  }
  return check_launch("tile pass");
}

}  // extern "C"
