/* lucivant-sim MI355X engine: the C ABI of liblsim.so.
 *
 * This is the drop-in boundary underneath the Python engine classes `Statevector` and
 * `DensityMatrix` (lucivant-sim 0.3.0, src/lucivant_sim/engines.py:138-264) and the drivers that
 * call them (src/lucivant_sim/simulator.py:229-317, 404-527). The reference has no FFI of its
 * own: each entry point below names the NumPy routine of the reference whose arithmetic it
 * replaces, and `lucivant_sim/_native.py` binds them with ctypes.
 *
 * Conventions for every entry point:
 *  - the return value is a status: LSIM_OK (0) or an LSIM_E* code; lsim_last_error() then holds
 *    a message for the calling thread. No C++ exception crosses this boundary.
 *  - `states` is device memory owned by the caller: `batch` states of 2^nbits amplitudes each,
 *    stored back to back; amplitude index bit q is qubit q (little-endian, engines.py:8-9).
 *    `dtype` is LSIM_C128 (interleaved float64 re/im) or LSIM_C64 (interleaved float32).
 *  - a density matrix of n qubits is passed as a state of nbits = 2n: flat index row*2^n + col,
 *    so vector bit n+q is the row bit and bit q the column bit of qubit q (engines.py:215-231).
 *  - `targets` is a HOST array of k bit positions; the first listed is the most significant bit
 *    of the operator's matrix index (gates.py:1-7, engines.py:83-89).
 *  - matrices, diagonals, selectors, outputs and workspaces are DEVICE pointers.
 *  - `stream` is a hipStream_t (NULL = the null stream); every call only enqueues work.
 */
#ifndef LSIM_H
#define LSIM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LSIM_OK 0
#define LSIM_EINVAL 1 /* bad argument: sizes, bit positions, pointers */
#define LSIM_EHIP 2   /* the HIP runtime reported an error */

#define LSIM_C128 0
#define LSIM_C64 1

#define LSIM_MAX_K 5        /* widest dense operator: 32 x 32 */
#define LSIM_MAX_TERMS 64   /* Pauli strings per expectation call */

/* Message of the last failed call on this thread ("" if none). */
const char* lsim_last_error(void);

/* ABI version of this library (bumped on any signature change). */
int lsim_abi_version(void);

/* Device 0 facts: compute units, HBM bytes, gfx architecture name. */
int lsim_device_info(int* cu_count, long long* hbm_bytes, char* arch, int arch_len);

/* HOST arrays: out[i] = in[0] + ... + in[i], added strictly left to right in float64 -- the
 * numbers numpy.cumsum gives, bit for bit. Replaces the cumulative distribution of _sample
 * (simulator.py:267); needs no GPU. */
int lsim_host_cumsum(const double* in, double* out, size_t n);

/* HOST pass planning (no GPU). A plan is a list of passes, each a list of operators; every
 * operator is LSIM_PLAN_DENSE (a 2^k x 2^k complex128 matrix, row-major) or LSIM_PLAN_DIAG (its
 * 2^k diagonal entries), on k targets listed from the highest bit down. The input is the static
 * circuit's gates in program order: gate i acts on nq[i] (1..5) qubits qubits[5i .. 5i+nq[i]-1]
 * (the first is its matrix's most significant bit), with its 4^nq[i] complex128 entries next in
 * `mats` (interleaved re, im). The reference applies one NumPy sweep per gate and fuses only
 * one-qubit runs (simulator.py:196-251); these plans fuse further (results change by rounding). */
#define LSIM_PLAN_DENSE 0
#define LSIM_PLAN_DIAG 1
typedef struct lsim_plan lsim_plan;

/* Fused operators of at most `width` qubits (at most `low_width`, when non-zero, for those that
 * touch a qubit below `low_bits`): one-qubit runs multiplied, then greedy blocks, then (merge_blocks
 * non-zero) blocks joined across others that commute with them. One operator per pass, on the
 * state's own bits. */
int lsim_host_plan_fused(int count, const int* nq, const int* qubits, const double* mats,
                         int nbits, int width, int low_bits, int low_width, int merge_blocks,
                         lsim_plan** plan);

/* Tile passes for lsim_sv_apply_tile: each pass a set of `tile_bits` state bits (the `low` lowest
 * always among them) and the operators (at most two targets, recorded as tile positions, the
 * i-th lowest tile bit being position i) of the longest run of pending gates that can run on it
 * whose cost stays within `budget` (dense one-qubit 1, dense two-qubit 2.5, diagonal 0.5 per
 * target) and count within `max_ops`; a gate on more than two qubits is a pass of its own. */
int lsim_host_plan_tiles(int count, const int* nq, const int* qubits, const double* mats,
                         int nbits, double budget, int tile_bits, int low, int max_ops,
                         lsim_plan** plan);

/* Sizes of a plan: passes, operators, complex values. */
int lsim_host_plan_sizes(const lsim_plan* plan, int* passes, int* ops, long long* values);

/* Copies a plan out: per pass its tile bits (0: an operator pass on the state's own bits) and
 * (first operator, operator count) in pass_ops; per operator 8 ints (kind, k, 5 targets, gates it
 * stands for) and the offset of its values in `values` (complex entries). */
int lsim_host_plan_read(const lsim_plan* plan, unsigned long long* tiles, int* pass_ops,
                        int* op_info, long long* op_at, double* values);

/* Releases a plan (NULL is ignored). */
void lsim_host_plan_free(lsim_plan* plan);

/* |index> in every state: replaces Statevector.__init__ / DensityMatrix.__init__
 * (engines.py:139-148, 216-225; for a density matrix pass nbits=2n, index=0). */
int lsim_sv_set_basis(void* states, int dtype, int nbits, long long batch, long long index,
                      void* stream);

/* State dst[j] <- state src[j] for j < count, within one batch (src and dst: DEVICE int arrays of
 * `count` indices; an index outside 0..batch-1 skips its copy; no state may be both a source and
 * a destination). No reference counterpart: the reference runs one fresh state per trajectory
 * (simulator.py:364-371); the batched trajectories keep shots with equal histories in one state
 * and copy it where their choices part. */
int lsim_sv_copy_states(void* states, int dtype, int nbits, long long batch, long long count,
                        const int* src, const int* dst, void* stream);

/* psi <- M psi on k <= 5 target bits, in place: replaces _apply (engines.py:38-42) and
 * _apply_slices (engines.py:45-119), and for a density matrix DensityMatrix.apply and
 * apply_channel (engines.py:233-250) given the superoperator U (x) conj(U) or sum K (x) conj(K)
 * on targets [rows..., cols...].
 * `mats`: M matrices of 2^k x 2^k, row-major, in `dtype`. `sel`: NULL (every state uses
 * mats[0]) or `batch` ints, state b uses mats[sel[b]] and is left untouched if sel[b] < 0. */
int lsim_sv_apply_dense(void* states, int dtype, int nbits, long long batch, int k,
                        const int* targets, const void* mats, const int* sel, void* stream);

/* Select the dense kernel variant: with tiling on (the default), operators whose targets include
 * low index bits run through LDS-staged tiles; off, every placement uses the direct kernel.
 * Returns the previous setting; a negative argument only reads it. No reference counterpart (a
 * tuning switch). This and the next two switches are process-wide atomics: a change applies to
 * the launches issued after it, from any thread. */
int lsim_set_dense_tiling(int enabled);

/* Select which complex128 operator widths run on the FP64 matrix cores (v_mfma_f64_16x16x4f64,
 * 16 groups per wavefront step) instead of the vector kernel: bit k of `widths` for k = 3, 4, 5
 * (states of at least k + 5 bits; other bits are ignored), on every placement of the targets.
 * The default is all three bits. Returns the
 * previous mask; a negative argument only reads it. Results agree with the vector kernel to
 * rounding. No reference counterpart (a tuning switch the autotuner sets). */
int lsim_set_dense_mfma(int widths);

/* Streaming (non-temporal) loads and stores in the dense, diagonal and scale kernels: 0 never;
 * 1 (the default) unless the launch's states together fit the 256 MiB last-level cache but not
 * the 32 MiB of L2 (cached accesses win there); 2 at every size. The dense kernels use them only
 * for placements whose load instructions cover whole 128-byte lines (every target at bit 3 or
 * above). Returns the previous setting; a negative
 * argument only reads it. Results are identical either way. No reference counterpart. */
int lsim_set_dense_streaming(int enabled);

/* ------------------------------------------------------------------------------------------
 * Tile passes: a run of small operators in one read and one write of each state.
 *
 * The amplitudes that agree on every bit except the LSIM_TILE_BITS tile bits `bits` (HOST ints,
 * ascending; tile bit i is state bit bits[i]) form a tile. Each tile is staged in LDS, the `nops`
 * operators of `ops` (HOST array) are applied to it in order, and it is written back. Operator j
 * acts on k = 1 or 2 tile bits t[0], t[1] (t[0] the most significant matrix bit) with the
 * complex128 matrix at element offset `mat` of the DEVICE array `mats` (also for LSIM_C64
 * states): 2^k x 2^k row-major for LSIM_TILE_DENSE, the 2^k diagonal entries for LSIM_TILE_DIAG.
 * A dense operator on two tile bits lists the higher one first (t[0] > t[1]).
 * Replaces, for a run of gates whose qubits are all tile bits, one NumPy sweep per gate of
 * _apply / _apply_slices (engines.py:38-119, driven by simulator.py:229-251).
 */
#define LSIM_TILE_BITS 12
#define LSIM_MAX_TILE_OPS 64
#define LSIM_TILE_DENSE 0
#define LSIM_TILE_DIAG 1

typedef struct lsim_tile_op {
  int kind, k;
  int t[2];
  long long mat;
} lsim_tile_op;

int lsim_sv_apply_tile(void* states, int dtype, int nbits, long long batch, const int* bits,
                       int nops, const lsim_tile_op* ops, const void* mats, void* stream);

/* psi <- diag(d) psi on k <= 5 target bits: the diagonal case of _apply_slices
 * (engines.py:91-104; rz, s, t, cz, cp). `diags`: M diagonals of 2^k entries; `sel` as above. */
int lsim_sv_apply_diag(void* states, int dtype, int nbits, long long batch, int k,
                       const int* targets, const void* diags, const int* sel, void* stream);

/* A permutation with phases on k <= 5 target bits: row r of the operator has its one non-zero
 * entry phases[r] in column perm[r] (perm: HOST ints, a permutation of 0..2^k-1; phases: DEVICE,
 * 2^k complex values in dtype per selector row). Bit r of unit_rows says phases[r] is exactly 1
 * (a plain move, no multiply). sel (DEVICE, optional) picks each state's phase row, or skips the
 * state when negative. The reference's slice moves for x, y, cx, swap and their phased forms
 * (engines.py:91-104): one read and one write of every amplitude, no matrix arithmetic. */
int lsim_sv_apply_perm(void* states, int dtype, int nbits, long long batch, int k,
                       const int* targets, const int* perm, unsigned int unit_rows,
                       const void* phases, const int* sel, void* stream);

/* psi <- c psi in every state: Statevector.scale (engines.py:156-157). */
int lsim_sv_scale(void* states, int dtype, int nbits, long long batch, double re, double im,
                  void* stream);

/* y <- y + (re + i im) x over `batch` states: the accumulation of DensityMatrix.apply_channel's
 * Kraus sum (engines.py:245-250) for channels too wide for one superoperator pass. */
int lsim_sv_axpy(void* y, const void* x, int dtype, int nbits, long long batch, double re,
                 double im, void* stream);

/* probs[b][i] = |psi_b[i]|^2 as float64: Statevector.probabilities (engines.py:207-209). */
int lsim_sv_probabilities(const void* states, int dtype, int nbits, long long batch,
                          double* probs, void* stream);

/* Bytes of device workspace the reductions below need for these sizes. `width` is k for
 * lsim_sv_reduced_dm and the number of terms for the expectation calls. */
size_t lsim_reduce_workspace_bytes(int nbits, long long batch, int width);

/* Reduced density matrix of each state on k <= 2 target bits, out[b] = 2^k x 2^k complex128,
 * row-major, index bits as for matrices. Deterministic (fixed summation order). It yields
 * probability_of_one (engines.py:185-187, the [1][1] entry for k = 1) and every Kraus branch
 * weight ||K psi||^2 = Tr(K rho_T K^dagger) of Statevector.apply_channel (engines.py:171-183)
 * from one read of the state. */
int lsim_sv_reduced_dm(const void* states, int dtype, int nbits, long long batch, int k,
                       const int* targets, double* out, double* workspace, void* stream);

/* out[b][t] = <psi_b| P_t |psi_b> for nterms <= 64 Pauli strings P_t given by HOST bit masks
 * (x: qubits with X or Y, z: qubits with Z or Y): the per-term probe loop of expectation
 * (simulator.py:517-527), P|j> = i^{nY} (-1)^{popcount(j & z)} |j ^ x>. */
int lsim_sv_pauli_expect(const void* states, int dtype, int nbits, long long batch, int nterms,
                         const unsigned long long* xmask, const unsigned long long* zmask,
                         double* out, double* workspace, void* stream);

/* lsim_sv_pauli_expect without a workspace argument, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No reference counterpart. It runs the same kernels, in the same summation order (bit-identical
 * results), with partials in a 4 MiB scratch array the library holds on each GPU until the
 * process ends, and calls no allocator. It refuses (LSIM_EINVAL) sizes whose partials exceed the
 * scratch: batch x nterms x (one partial per workgroup, at most 4096) x 8 bytes must stay
 * within 4 MiB, so one state of any size takes up to 64 strings, and 128 states of 20 bits or
 * fewer take up to 16 strings each (see lsim_reduce_workspace_bytes). Every call on one GPU shares
 * that GPU's scratch: calls that overlap on one GPU (on different streams, or from different
 * threads) race and must be serialised by the caller. */
int lsim_sv_pauli_expect_noalloc(const void* states, int dtype, int nbits, long long batch,
                                 int nterms, const unsigned long long* xmask,
                                 const unsigned long long* zmask, double* out, void* stream);

/* Bytes of device workspace lsim_marginal needs for these sizes (0 when qubit 0 is kept). */
size_t lsim_marginal_workspace_bytes(int n, long long batch, int nkeep, const int* keep);

/* out[b][j] = the probability that the `nkeep` qubits `keep` (HOST ints, ascending) read j
 * (bit i of j = keep[i]), from float64 probabilities probs[b][2^n] on the device: the marginal of
 * engines.py:267-275, summed over the other qubits in exactly NumPy's order for
 * probs.reshape([2]*n).sum(axis=dropped) (pieces of at most 8192 contiguous entries, each
 * pairwise-summed, added in increasing index order), so sampling from it matches the host
 * marginal bit for bit and only 2^nkeep values leave the device. `workspace` (device) holds
 * lsim_marginal_workspace_bytes(n, batch, nkeep, keep) bytes; it may be NULL when that is 0. */
int lsim_marginal(const double* probs, int n, long long batch, int nkeep, const int* keep,
                  double* out, double* workspace, void* stream);

/* probs[b][i] = max(Re rho_b[i][i], 0) for n-qubit density matrices: DensityMatrix.probabilities
 * (engines.py:257-260). */
int lsim_dm_probabilities(const void* rho, int dtype, int n, long long batch, double* probs,
                          void* stream);

/* out[b][t] = Re Tr(P_t rho_b): the noisy branch of expectation (simulator.py:505-516, 530-533). */
int lsim_dm_pauli_expect(const void* rho, int dtype, int n, long long batch, int nterms,
                         const unsigned long long* xmask, const unsigned long long* zmask,
                         double* out, double* workspace, void* stream);

/* ------------------------------------------------------------------------------------------
 * Whole-circuit programs on LDS-resident states (one workgroup per circuit instance).
 *
 * Replaces, for states of up to lsim_program_max_qubits() qubits (13 for complex128 on
 * gfx950's 160 KiB of LDS), the per-shot loop of the trajectory method (simulator.py:287-317,
 * 364-371, with engines.py:159-205 per operation) and the per-operation loop of static
 * evolution (simulator.py:229-251) for a batch of circuits. Every instance starts in |0...0>.
 *
 * An op runs only when (flags & guard_mask) == guard_value; COND sets flag bit `flag` to
 * ((clbits & cmask) == cpattern) != negate. Matrices are complex128 in `mats` (also for
 * LSIM_C64 states); `reals` holds float64 side data (see kinds below).
 */
#define LSIM_OP_DENSE 0   /* k <= 4; mats[mat]: 2^k x 2^k matrix */
#define LSIM_OP_DIAG 1    /* k <= 5; mats[mat]: 2^k diagonal */
#define LSIM_OP_MIX 2     /* k <= 4; reals[aux..aux+nb): cumulative probabilities; reals[aux+nb+j]:
                             the mats offset of branch j or -1 for the identity; one draw */
#define LSIM_OP_KRAUS 3   /* k <= 2; mats[mat + j*4^k]: the nb Kraus operators; one draw */
#define LSIM_OP_MEASURE 4 /* qubit t[0] into clbit (or none if clbit < 0); one draw, plus one more
                             for a readout flip when aux >= 0 (reals[aux] = p01, reals[aux+1] = p10) */
#define LSIM_OP_RESET 5   /* qubit t[0]; one draw */
#define LSIM_OP_COND 6    /* sets flag bit `flag` */
#define LSIM_OP_PHASE 7   /* mats[mat]: global phase factor */

typedef struct lsim_prog_op {
  unsigned char kind, k, flag, negate;
  unsigned char t[5]; /* target qubits, first = most significant matrix bit */
  unsigned char pad[3];
  int mat, nb, aux, draw, clbit;
  unsigned int guard_mask, guard_value;
  unsigned long long cmask, cpattern;
} lsim_prog_op;

/* Widest state the program kernel holds in one workgroup's LDS for `dtype` (0 on error). */
int lsim_program_max_qubits(int dtype);

/* Run `instances` programs: instance b executes ops[op_start[b] .. op_start[b]+op_count[b])
 * with draws from draws[draw_start[b] ..] (draws/draw_start may be NULL when no op draws).
 * Optional outputs (NULL to skip): clbits[b]; probs[b][2^n] = |psi|^2; state[b][2^n] in dtype;
 * expect[b][t] = <psi_b|P_bt|psi_b> for the nterms Pauli strings of each instance, given by the
 * DEVICE arrays pauli_x[b][t] / pauli_z[b][t] (bit masks as for lsim_sv_pauli_expect);
 * decisions, laid out as the draws with two doubles per draw: for the draw of each MEASURE or
 * RESET, {p1, outcome}, and of each KRAUS, {running weight at the branch taken, branch index}
 * (entries of other draws are left as they are). The record states the error of a complex64
 * run against complex128 per decision; it has no reference counterpart. */
int lsim_program_run(int dtype, int n, long long instances, const lsim_prog_op* ops,
                     const int* op_start, const int* op_count, const void* mats,
                     const double* reals, const double* draws, const int* draw_start,
                     unsigned long long* clbits, double* probs, void* state, int nterms,
                     const unsigned long long* pauli_x, const unsigned long long* pauli_z,
                     double* expect, double* decisions, void* stream);

/* ------------------------------------------------------------------------------------------
 * Library-owned HIP graphs (beyond the reference: its engines have no launch overhead to cut).
 * A caller records a sequence of calls to this library on one stream and replays it with one
 * launch: lsim_graph_capture_begin(stream); the calls (each only enqueues kernels on `stream`);
 * lsim_graph_capture_end(stream, &g); then lsim_graph_launch(g, stream) as often as wanted.
 * Capture is thread-local: other threads may use the GPU meanwhile. The graph keeps the
 * pointers and arguments of the recorded calls; what they point to may change between replays.
 * `stream` must not be the null stream. No call during capture may allocate or synchronise.
 */
typedef struct lsim_graph lsim_graph;
int lsim_graph_capture_begin(void* stream);
int lsim_graph_capture_end(void* stream, lsim_graph** out);
/* Kernel nodes the graph holds (-1 for NULL). */
long long lsim_graph_nodes(const lsim_graph* g);
int lsim_graph_launch(lsim_graph* g, void* stream);
int lsim_graph_destroy(lsim_graph* g);

#ifdef __cplusplus
}
#endif

#endif /* LSIM_H */
