// Host-side pass planning (no device code): the fusion of a gate sequence into operators of at
// most `width` qubits, and the choice of tile bits for tile passes. Planning runs here, in C++,
// so that it costs less host time per gate than the device takes to apply the gates; the Python
// layer only wraps the result (tests/plan_reference.py restates the same steps on NumPy).
//
// Fusion (the passes of a statevector plan):
//   1. one-qubit runs: consecutive one-qubit gates on a qubit with no other gate on that qubit in
//      between are multiplied into one 2x2 operator, placed where the run's first gate was (the
//      later gates commute with everything in between, which acts on other qubits);
//   2. greedy: each operator joins the latest block touching one of its qubits while the union
//      has at most `width` qubits (at most `low_width` when the union holds a qubit below
//      `low_bits`), else starts a block; new qubits of a block become its least significant
//      matrix positions;
//   3. merge: a block moves forward to a later block (or backward to an earlier one) when their
//      union fits and no block between them shares a qubit with it; repeated until nothing moves;
//   4. each block's targets are listed from the highest bit down, and a block whose matrix has no
//      off-diagonal entry becomes a diagonal operator.
// Tile passes: each pass takes the pending gates that can run next on one set of `tile_bits` bits
// (the `low` lowest, then bits added one at a time -- two when one alone lets no more gates run
// -- each time the bit that lets the most pending gates run, then single swaps that let more
// run), cut at the longest prefix whose fused operators (fusion at width 2) cost at most
// `budget` and number at most `max_ops`. A gate runs on a set when its qubits lie in the set and
// no earlier pending gate that cannot run shares a qubit with it. Gates on more than two qubits
// run as their own operator pass.
#include <algorithm>
#include <vector>

#include "lsim_internal.h"

struct lsim_plan {
  struct Op {
    int kind;      // LSIM_PLAN_DENSE or LSIM_PLAN_DIAG
    int k;         // targets
    int bits[5];   // targets, highest first (tile passes: tile positions)
    int gates;     // gates of the circuit the operator stands for
    long long at;  // offset of its values (complex entries) in `values`
  };
  struct Pass {
    unsigned long long tile;  // tile bits; 0: a single-operator pass on the state's own bits
    int first, count;         // its operators
  };
  std::vector<Pass> passes;
  std::vector<Op> ops;
  std::vector<double> values;  // interleaved re, im
};

namespace {

constexpr int kMaxQ = 5;

struct Cd {
  double re, im;
};
inline Cd mul(Cd a, Cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline Cd add(Cd a, Cd b) { return {a.re + b.re, a.im + b.im}; }

// An operator on `k` qubits; q[0] is its most significant matrix position. `m` is empty when only
// the structure is wanted.
struct Block {
  int k = 0;
  int q[kMaxQ] = {};
  std::vector<Cd> m;
  int gates = 0;
  bool diag = true;  // every gate in it is diagonal (the tile cost model)

  int index(int qubit) const {
    for (int i = 0; i < k; ++i)
      if (q[i] == qubit) return i;
    return -1;
  }
};

struct Gate {
  int k;
  int q[kMaxQ];
  const Cd* m;  // 2^k x 2^k
  bool diag;
};

bool off_diagonal_zero(const Cd* m, int d) {
  for (int r = 0; r < d; ++r)
    for (int c = 0; c < d; ++c)
      if (r != c && (m[r * d + c].re != 0.0 || m[r * d + c].im != 0.0)) return false;
  return true;
}

// kron(m, I): the qubits of `qs` missing from the block become its least significant positions.
void extend(Block& b, const int* qs, int nq, bool matrices) {
  int add_q[kMaxQ], na = 0;
  for (int i = 0; i < nq; ++i)
    if (b.index(qs[i]) < 0) add_q[na++] = qs[i];
  if (na == 0) return;
  if (matrices) {
    const int d = 1 << b.k, e = 1 << na, nd = d * e;
    std::vector<Cd> out((size_t)nd * nd, Cd{0.0, 0.0});
    for (int r = 0; r < d; ++r)
      for (int c = 0; c < d; ++c)
        for (int a = 0; a < e; ++a) out[(size_t)(r * e + a) * nd + (c * e + a)] = b.m[(size_t)r * d + c];
    b.m.swap(out);
  }
  for (int i = 0; i < na; ++i) b.q[b.k++] = add_q[i];
}

// m <- G m for a one-qubit G on row bit `bit`, in place: each pair of rows (bit clear, bit set)
// is mixed column by column.
void apply1(Block& b, const Cd* g, int bit) {
  const int d = 1 << b.k, h = 1 << bit;
  Cd* m = b.m.data();
  for (int r0 = 0; r0 < d; ++r0) {
    if (r0 & h) continue;
    Cd* x = m + (size_t)r0 * d;
    Cd* y = m + (size_t)(r0 | h) * d;
    for (int c = 0; c < d; ++c) {
      const Cd u = x[c], v = y[c];
      x[c] = add(mul(g[0], u), mul(g[1], v));
      y[c] = add(mul(g[2], u), mul(g[3], v));
    }
  }
}

// m <- G m for a two-qubit G on row bits hi (its most significant bit) and lo, in place.
void apply2(Block& b, const Cd* g, int hi, int lo) {
  const int d = 1 << b.k, bh = 1 << hi, bl = 1 << lo;
  Cd* m = b.m.data();
  for (int r0 = 0; r0 < d; ++r0) {
    if (r0 & (bh | bl)) continue;
    Cd* row[4] = {m + (size_t)r0 * d, m + (size_t)(r0 | bl) * d, m + (size_t)(r0 | bh) * d,
                  m + (size_t)(r0 | bh | bl) * d};
    for (int c = 0; c < d; ++c) {
      const Cd v[4] = {row[0][c], row[1][c], row[2][c], row[3][c]};
      for (int i = 0; i < 4; ++i) {
        Cd acc = mul(g[4 * i], v[0]);
        for (int j = 1; j < 4; ++j) acc = add(acc, mul(g[4 * i + j], v[j]));
        row[i][c] = acc;
      }
    }
  }
}

// m <- G m, G acting on the block positions `pos` (pos[0] = G's most significant bit).
void apply(Block& b, const Cd* g, const int* pos, int m) {
  if (m == 1) return apply1(b, g, b.k - 1 - pos[0]);
  if (m == 2) return apply2(b, g, b.k - 1 - pos[0], b.k - 1 - pos[1]);
  const int k = b.k, d = 1 << k, dm = 1 << m;
  int mask = 0, shift[kMaxQ];
  for (int i = 0; i < m; ++i) {
    shift[i] = k - 1 - pos[i];
    mask |= 1 << shift[i];
  }
  std::vector<Cd> out((size_t)d * d);
  int rows[1 << kMaxQ];
  for (int r = 0; r < d; ++r) {
    int s = 0;
    for (int i = 0; i < m; ++i) s |= ((r >> shift[i]) & 1) << (m - 1 - i);
    const int rest = r & ~mask;
    for (int j = 0; j < dm; ++j) {
      int src = rest;
      for (int i = 0; i < m; ++i) src |= ((j >> (m - 1 - i)) & 1) << shift[i];
      rows[j] = src;
    }
    for (int c = 0; c < d; ++c) {
      Cd acc = mul(g[s * dm], b.m[(size_t)rows[0] * d + c]);
      for (int j = 1; j < dm; ++j) acc = add(acc, mul(g[s * dm + j], b.m[(size_t)rows[j] * d + c]));
      out[(size_t)r * d + c] = acc;
    }
  }
  b.m.swap(out);
}

void absorb(Block& b, const Cd* g, const int* qs, int nq, int count, bool gdiag, bool matrices) {
  extend(b, qs, nq, matrices);
  if (matrices) {
    int pos[kMaxQ];
    for (int i = 0; i < nq; ++i) pos[i] = b.index(qs[i]);
    apply(b, g, pos, nq);
  }
  b.gates += count;
  b.diag = b.diag && gdiag;
}

Block start(const int* qs, int nq, bool matrices) {
  Block b;
  b.k = nq;
  for (int i = 0; i < nq; ++i) b.q[i] = qs[i];
  if (matrices) {
    const int d = 1 << nq;
    b.m.assign((size_t)d * d, Cd{0.0, 0.0});
    for (int i = 0; i < d; ++i) b.m[(size_t)i * d + i] = Cd{1.0, 0.0};
  }
  return b;
}

// Step 1: runs of one-qubit gates multiplied into one operator at the run's first gate.
struct Item {
  int k;
  int q[kMaxQ];
  std::vector<Cd> m;  // empty: the gate's own matrix (`g`)
  const Cd* g;
  int count;
  bool diag;
  const Cd* matrix() const { return m.empty() ? g : m.data(); }
};

std::vector<Item> one_qubit_runs(const std::vector<const Gate*>& gates, bool matrices) {
  std::vector<Item> out;
  out.reserve(gates.size());
  std::vector<int> open;  // qubit -> index in out of its open run, or -1
  for (const Gate* gt : gates) {
    for (int i = 0; i < gt->k; ++i)
      if (gt->q[i] >= (int)open.size()) open.resize(gt->q[i] + 1, -1);
    if (gt->k == 1) {
      const int at = open[gt->q[0]];
      if (at < 0) {
        open[gt->q[0]] = (int)out.size();
        Item it;
        it.k = 1;
        it.q[0] = gt->q[0];
        it.g = gt->m;
        it.count = 1;
        it.diag = gt->diag;
        out.push_back(std::move(it));
      } else {
        Item& it = out[at];
        if (matrices) {
          const Cd* p = it.matrix();
          const Cd* g = gt->m;
          std::vector<Cd> prod(4);
          for (int r = 0; r < 2; ++r)
            for (int c = 0; c < 2; ++c) prod[r * 2 + c] = add(mul(g[r * 2], p[c]), mul(g[r * 2 + 1], p[2 + c]));
          it.m.swap(prod);
        }
        it.count += 1;
        it.diag = it.diag && gt->diag;
      }
      continue;
    }
    for (int i = 0; i < gt->k; ++i) open[gt->q[i]] = -1;
    Item it;
    it.k = gt->k;
    for (int i = 0; i < gt->k; ++i) it.q[i] = gt->q[i];
    it.g = gt->m;
    it.count = 1;
    it.diag = gt->diag;
    out.push_back(std::move(it));
  }
  return out;
}

bool fits(const Block& a, const int* qs, int nq, int width, int low_bits, int low_width, int* size) {
  int u = a.k, lowest = a.k ? a.q[0] : 1 << 30;
  for (int i = 0; i < a.k; ++i) lowest = std::min(lowest, a.q[i]);
  for (int i = 0; i < nq; ++i) {
    lowest = std::min(lowest, qs[i]);
    if (a.index(qs[i]) < 0) ++u;
  }
  *size = u;
  const int limit = (low_width && low_bits && lowest < low_bits) ? low_width : width;
  return u <= limit;
}

// Step 2.
std::vector<Block> greedy(const std::vector<Item>& items, int width, int low_bits, int low_width,
                          bool matrices) {
  std::vector<Block> blocks;
  std::vector<int> last;  // qubit -> latest block touching it, or -1
  for (const Item& it : items) {
    int j = -1;
    for (int i = 0; i < it.k; ++i) {
      if (it.q[i] >= (int)last.size()) last.resize(it.q[i] + 1, -1);
      j = std::max(j, last[it.q[i]]);
    }
    int size = 0;
    if (j >= 0 && fits(blocks[j], it.q, it.k, width, low_bits, low_width, &size)) {
      absorb(blocks[j], it.matrix(), it.q, it.k, it.count, it.diag, matrices);
    } else {
      Block b = start(it.q, it.k, matrices);
      absorb(b, it.matrix(), it.q, it.k, it.count, it.diag, matrices);
      blocks.push_back(std::move(b));
      j = (int)blocks.size() - 1;
    }
    for (int i = 0; i < it.k; ++i) last[it.q[i]] = j;
  }
  return blocks;
}

bool share(const Block& a, const Block& b) {
  for (int i = 0; i < a.k; ++i)
    if (b.index(a.q[i]) >= 0) return true;
  return false;
}

// `first` then `second` as one block (first's qubits lead).
Block join(const Block& first, const Block& second, bool matrices) {
  Block b = start(first.q, first.k, matrices);
  if (matrices) b.m = first.m;
  b.gates = first.gates;
  b.diag = first.diag;
  absorb(b, matrices ? second.m.data() : nullptr, second.q, second.k, second.gates, second.diag, matrices);
  return b;
}

// Step 3.
void merge(std::vector<Block>& blocks, int width, int low_bits, int low_width, bool matrices) {
  while (true) {
    const size_t before = blocks.size();
    for (int backward = 0; backward < 2; ++backward) {
      const int nb = (int)blocks.size();
      std::vector<char> alive(nb, 1);
      for (int o = 0; o < nb; ++o) {
        const int j = backward ? nb - 1 - o : o;
        if (!alive[j]) continue;
        for (int j2 = backward ? j - 1 : j + 1; backward ? j2 >= 0 : j2 < nb; j2 += backward ? -1 : 1) {
          if (!alive[j2]) continue;
          int size = 0;
          if (fits(blocks[j], blocks[j2].q, blocks[j2].k, width, low_bits, low_width, &size)) {
            const Block& first = backward ? blocks[j2] : blocks[j];
            const Block& second = backward ? blocks[j] : blocks[j2];
            Block joined = join(first, second, matrices);
            blocks[j2] = std::move(joined);
            alive[j] = 0;
            break;
          }
          if (share(blocks[j], blocks[j2])) break;
        }
      }
      std::vector<Block> kept;
      kept.reserve(nb);
      for (int j = 0; j < nb; ++j)
        if (alive[j]) kept.push_back(std::move(blocks[j]));
      blocks.swap(kept);
    }
    if (blocks.size() == before) return;
  }
}

std::vector<Block> fuse(const std::vector<const Gate*>& gates, int width, int low_bits,
                        int low_width, bool do_merge, bool matrices) {
  std::vector<Item> items = one_qubit_runs(gates, matrices);
  std::vector<Block> blocks = greedy(items, width, low_bits, low_width, matrices);
  if (do_merge) merge(blocks, width, low_bits, low_width, matrices);
  return blocks;
}

// Step 4: targets from the highest bit down, then dense or diagonal. `local` maps qubits to
// the bits the operator is recorded on (nullptr: the qubits themselves).
void emit(lsim_plan* p, const Block& b, const int* local) {
  const int k = b.k, d = 1 << k;
  int order[kMaxQ];
  for (int i = 0; i < k; ++i) order[i] = i;
  std::sort(order, order + k, [&](int x, int y) { return b.q[x] > b.q[y]; });
  int old[1 << kMaxQ];
  for (int idx = 0; idx < d; ++idx) {
    int o = 0;
    for (int j = 0; j < k; ++j) o |= ((idx >> (k - 1 - j)) & 1) << (k - 1 - order[j]);
    old[idx] = o;
  }
  std::vector<Cd> m((size_t)d * d);
  for (int r = 0; r < d; ++r)
    for (int c = 0; c < d; ++c) m[(size_t)r * d + c] = b.m[(size_t)old[r] * d + old[c]];
  lsim_plan::Op op;
  op.k = k;
  for (int j = 0; j < kMaxQ; ++j) op.bits[j] = 0;
  for (int j = 0; j < k; ++j) op.bits[j] = local ? local[b.q[order[j]]] : b.q[order[j]];
  op.gates = b.gates;
  op.at = (long long)(p->values.size() / 2);
  if (off_diagonal_zero(m.data(), d)) {
    op.kind = LSIM_PLAN_DIAG;
    for (int i = 0; i < d; ++i) {
      p->values.push_back(m[(size_t)i * d + i].re);
      p->values.push_back(m[(size_t)i * d + i].im);
    }
  } else {
    op.kind = LSIM_PLAN_DENSE;
    for (const Cd& v : m) {
      p->values.push_back(v.re);
      p->values.push_back(v.im);
    }
  }
  p->ops.push_back(op);
}

int read_gates(int count, const int* nq, const int* qubits, const double* mats, int nbits,
               std::vector<Gate>* gates) {
  LSIM_REQUIRE(count >= 0, "negative gate count");
  LSIM_REQUIRE(count == 0 || (nq != nullptr && qubits != nullptr && mats != nullptr), "a NULL array");
  gates->resize(count);
  const Cd* at = reinterpret_cast<const Cd*>(mats);
  for (int i = 0; i < count; ++i) {
    Gate& g = (*gates)[i];
    g.k = nq[i];
    LSIM_REQUIRE(g.k >= 1 && g.k <= kMaxQ, "gate %d acts on %d qubits (1 to %d)", i, g.k, kMaxQ);
    for (int j = 0; j < kMaxQ; ++j) g.q[j] = 0;
    for (int j = 0; j < g.k; ++j) {
      g.q[j] = qubits[i * kMaxQ + j];
      LSIM_REQUIRE(g.q[j] >= 0 && g.q[j] < nbits, "gate %d: qubit %d is outside 0..%d", i, g.q[j], nbits - 1);
      for (int l = 0; l < j; ++l) LSIM_REQUIRE(g.q[l] != g.q[j], "gate %d repeats qubit %d", i, g.q[j]);
    }
    g.m = at;
    g.diag = off_diagonal_zero(at, 1 << g.k);
    at += (size_t)1 << (2 * g.k);
  }
  return LSIM_OK;
}

// The tile cost of the operators fusion at width 2 forms from `gates`, and their number.
void tile_structure(const std::vector<const Gate*>& gates, double* cost, int* count) {
  std::vector<Block> blocks = fuse(gates, 2, 0, 0, true, false);
  double c = 0.0;
  for (const Block& b : blocks) c += b.diag ? 0.5 * b.k : (b.k == 1 ? 1.0 : 2.5);
  *cost = c;
  *count = (int)blocks.size();
}

int runnable(const std::vector<unsigned long long>& masks, unsigned long long s,
             unsigned long long full) {
  unsigned long long blocked = 0;
  int runs = 0;
  for (unsigned long long m : masks) {
    if ((m & blocked) || (m & ~s)) {
      blocked |= m;
      if (blocked == full) break;
    } else {
      ++runs;
    }
  }
  return runs;
}

int popcount(unsigned long long v) { return __builtin_popcountll(v); }

unsigned long long choose(const std::vector<unsigned long long>& wm, unsigned long long first_mask,
                          int n, int tile_bits, int low, unsigned long long full) {
  const unsigned long long low_mask = (1ull << low) - 1;
  unsigned long long s = low_mask | first_mask;
  while (popcount(s) < tile_bits) {
    const int base = runnable(wm, s, full);
    unsigned long long best = 0;
    int value = -1;
    for (int q = 0; q < n; ++q)
      if (!((s >> q) & 1)) {
        const int v = runnable(wm, s | (1ull << q), full);
        if (v > value) {
          value = v;
          best = s | (1ull << q);
        }
      }
    if (value == base && popcount(s) <= tile_bits - 2) {
      unsigned long long pick = 0;
      int pv = -1;
      for (int a = 0; a < n; ++a) {
        if ((s >> a) & 1) continue;
        for (int b = a + 1; b < n; ++b) {
          if ((s >> b) & 1) continue;
          const unsigned long long t = s | (1ull << a) | (1ull << b);
          const int v = runnable(wm, t, full);
          if (v > pv) {
            pv = v;
            pick = t;
          }
        }
      }
      s = pick;
    } else {
      s = best;
    }
  }
  int cur = runnable(wm, s, full);
  for (int round = 0; round < 3; ++round) {
    bool better = false;
    int cands[64], nc = 0;
    for (int q = low; q < n; ++q)
      if (((s >> q) & 1) && !((first_mask >> q) & 1)) cands[nc++] = q;
    for (int i = 0; i < nc; ++i) {
      const int a = cands[i];
      const unsigned long long from = s;
      for (int b = 0; b < n; ++b) {
        if ((from >> b) & 1) continue;
        const unsigned long long t = (from & ~(1ull << a)) | (1ull << b);
        const int v = runnable(wm, t, full);
        if (v > cur) {
          s = t;
          cur = v;
          better = true;
          break;
        }
      }
    }
    if (!better) break;
  }
  return s;
}

int finish(lsim_plan* p, lsim_plan** out) {
  *out = p;
  return LSIM_OK;
}

}  // namespace

extern "C" {

int lsim_host_plan_fused(int count, const int* nq, const int* qubits, const double* mats,
                         int nbits, int width, int low_bits, int low_width, int merge_blocks,
                         lsim_plan** plan) {
  LSIM_REQUIRE(plan != nullptr, "the plan pointer is NULL");
  *plan = nullptr;
  LSIM_REQUIRE(nbits >= 1 && nbits <= 63, "nbits must be between 1 and 63, not %d", nbits);
  LSIM_REQUIRE(width >= 1 && width <= kMaxQ, "fusion width must be between 1 and %d", kMaxQ);
  LSIM_REQUIRE(low_width >= 0 && low_width <= kMaxQ && low_bits >= 0, "bad low-qubit limits");
  std::vector<Gate> gates;
  if (int rc = read_gates(count, nq, qubits, mats, nbits, &gates)) return rc;
  std::vector<const Gate*> ptrs(count);
  for (int i = 0; i < count; ++i) ptrs[i] = &gates[i];
  lsim_plan* p = new lsim_plan;
  for (const Block& b : fuse(ptrs, width, low_bits, low_width, merge_blocks != 0, true)) {
    p->passes.push_back({0ull, (int)p->ops.size(), 1});
    emit(p, b, nullptr);
  }
  return finish(p, plan);
}

int lsim_host_plan_tiles(int count, const int* nq, const int* qubits, const double* mats,
                         int nbits, double budget, int tile_bits, int low, int max_ops,
                         lsim_plan** plan) {
  LSIM_REQUIRE(plan != nullptr, "the plan pointer is NULL");
  *plan = nullptr;
  LSIM_REQUIRE(nbits >= 1 && nbits <= 63, "nbits must be between 1 and 63, not %d", nbits);
  LSIM_REQUIRE(tile_bits >= 2 && tile_bits <= nbits, "tile bits must be between 2 and nbits");
  LSIM_REQUIRE(low >= 0 && low <= tile_bits - 2, "the low bits must leave room for a gate's two");
  LSIM_REQUIRE(max_ops >= 1 && budget > 0, "a positive budget and operator limit are needed");
  std::vector<Gate> gates;
  if (int rc = read_gates(count, nq, qubits, mats, nbits, &gates)) return rc;
  std::vector<unsigned long long> masks(count);
  for (int i = 0; i < count; ++i) {
    masks[i] = 0;
    for (int j = 0; j < gates[i].k; ++j) masks[i] |= 1ull << gates[i].q[j];
  }
  const unsigned long long full = nbits == 64 ? ~0ull : (1ull << nbits) - 1;
  const size_t window = (size_t)64 * nbits;  // pending gates looked at when choosing bits
  lsim_plan* p = new lsim_plan;
  std::vector<int> pending(count);
  for (int i = 0; i < count; ++i) pending[i] = i;
  std::vector<unsigned long long> wm;
  std::vector<int> taken;
  std::vector<const Gate*> prefix;
  std::vector<char> used(count, 0);
  while (!pending.empty()) {
    const int first = pending[0];
    if (gates[first].k > 2) {
      Block b = start(gates[first].q, gates[first].k, true);
      absorb(b, gates[first].m, gates[first].q, gates[first].k, 1, gates[first].diag, true);
      p->passes.push_back({0ull, (int)p->ops.size(), 1});
      emit(p, b, nullptr);
      pending.erase(pending.begin());
      continue;
    }
    wm.clear();
    for (size_t i = 0; i < pending.size() && i < window; ++i) wm.push_back(masks[pending[i]]);
    const unsigned long long s = choose(wm, masks[first], nbits, tile_bits, low, full);
    unsigned long long blocked = 0;
    taken.clear();
    for (int gi : pending) {
      const unsigned long long m = masks[gi];
      if ((m & blocked) || (m & ~s) || gates[gi].k > 2)
        blocked |= m;
      else
        taken.push_back(gi);
    }
    // the longest prefix of the runnable gates whose operators fit the budget and the table
    int lo = 1, hi = (int)taken.size();
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      prefix.assign(mid, nullptr);
      for (int i = 0; i < mid; ++i) prefix[i] = &gates[taken[i]];
      double cost = 0;
      int nops = 0;
      tile_structure(prefix, &cost, &nops);
      if (cost <= budget && nops <= max_ops)
        lo = mid;
      else
        hi = mid - 1;
    }
    prefix.assign(lo, nullptr);
    for (int i = 0; i < lo; ++i) {
      prefix[i] = &gates[taken[i]];
      used[taken[i]] = 1;
    }
    int local[64];
    for (int q = 0, at = 0; q < nbits; ++q) local[q] = ((s >> q) & 1) ? at++ : -1;
    const int op0 = (int)p->ops.size();
    for (const Block& b : fuse(prefix, 2, 0, 0, true, true)) emit(p, b, local);
    p->passes.push_back({s, op0, (int)p->ops.size() - op0});
    size_t w = 0;
    for (size_t i = 0; i < pending.size(); ++i)
      if (!used[pending[i]]) pending[w++] = pending[i];
    pending.resize(w);
  }
  return finish(p, plan);
}

int lsim_host_plan_sizes(const lsim_plan* plan, int* passes, int* ops, long long* values) {
  LSIM_REQUIRE(plan != nullptr && passes != nullptr && ops != nullptr && values != nullptr,
               "a NULL pointer");
  *passes = (int)plan->passes.size();
  *ops = (int)plan->ops.size();
  *values = (long long)(plan->values.size() / 2);
  return LSIM_OK;
}

int lsim_host_plan_read(const lsim_plan* plan, unsigned long long* tiles, int* pass_ops,
                        int* op_info, long long* op_at, double* values) {
  LSIM_REQUIRE(plan != nullptr, "the plan pointer is NULL");
  LSIM_REQUIRE((plan->passes.empty() || (tiles != nullptr && pass_ops != nullptr)) &&
                   (plan->ops.empty() || (op_info != nullptr && op_at != nullptr)) &&
                   (plan->values.empty() || values != nullptr),
               "a NULL array");
  for (size_t i = 0; i < plan->passes.size(); ++i) {
    tiles[i] = plan->passes[i].tile;
    pass_ops[2 * i] = plan->passes[i].first;
    pass_ops[2 * i + 1] = plan->passes[i].count;
  }
  for (size_t i = 0; i < plan->ops.size(); ++i) {
    const lsim_plan::Op& op = plan->ops[i];
    int* o = op_info + 8 * i;
    o[0] = op.kind;
    o[1] = op.k;
    for (int j = 0; j < kMaxQ; ++j) o[2 + j] = op.bits[j];
    o[7] = op.gates;
    op_at[i] = op.at;
  }
  std::copy(plan->values.begin(), plan->values.end(), values);
  return LSIM_OK;
}

void lsim_host_plan_free(lsim_plan* plan) { delete plan; }

}  // extern "C"
