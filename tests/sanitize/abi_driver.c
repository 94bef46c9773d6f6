/* Host-side checks of liblsim's C ABI under AddressSanitizer and UndefinedBehaviorSanitizer:
 * every entry point refuses bad arguments with LSIM_EINVAL before any device work, and the host
 * cumulative sum adds left to right. Needs no GPU. Built and run by tests/test_host_sanitizers.py
 * against a copy of the library whose host code is sanitized. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "lsim.h"
static int fails = 0;
#define EXPECT(call, code) do { int rc_ = (call); if (rc_ != (code)) { printf("FAIL %s -> %d: %s\n", #call, rc_, lsim_last_error()); fails++; } } while (0)
int main(void) {
  void* fake = (void*)0x1000;
  int t2[2] = {0, 0}, t1[1] = {0}, t6[6] = {0};
  unsigned long long x1[1] = {1ull << 5}, z1[1] = {0}, x64[64], z64[64];
  for (int i = 0; i < 64; ++i) { x64[i] = 1; z64[i] = 0; }
  EXPECT(lsim_sv_apply_dense(NULL, 0, 4, 1, 1, t1, fake, NULL, NULL), LSIM_EINVAL);
  EXPECT(lsim_sv_apply_dense(fake, 0, 4, 1, 6, t6, fake, NULL, NULL), LSIM_EINVAL);
  EXPECT(lsim_sv_apply_dense(fake, 0, 4, 1, 2, t2, fake, NULL, NULL), LSIM_EINVAL);
  EXPECT(lsim_sv_apply_diag(fake, 7, 4, 1, 1, t1, fake, NULL, NULL), LSIM_EINVAL);
  EXPECT(lsim_sv_apply_dense(fake, 0, 40, 1, 1, t1, fake, NULL, NULL), LSIM_EINVAL);
  EXPECT(lsim_sv_reduced_dm(fake, 0, 4, 1, 3, t6, fake, fake, NULL), LSIM_EINVAL);
  EXPECT(lsim_sv_pauli_expect(fake, 0, 4, 1, 1, x1, z1, fake, fake, NULL), LSIM_EINVAL);
  EXPECT(lsim_sv_pauli_expect(fake, 0, 4, 1, 65, x64, z64, fake, fake, NULL), LSIM_EINVAL);
  EXPECT(lsim_sv_pauli_expect_noalloc(fake, 0, 20, 1024, 64, x64, z64, fake, NULL), LSIM_EINVAL);
  EXPECT(lsim_dm_pauli_expect(fake, 0, 2, 1, 1, x1, z1, fake, fake, NULL), LSIM_EINVAL);
  EXPECT(lsim_sv_set_basis(fake, 0, 4, 1, 16, NULL), LSIM_EINVAL);
  EXPECT(lsim_sv_set_basis(fake, 0, 4, 0, 0, NULL), LSIM_EINVAL);
  EXPECT(lsim_sv_scale(NULL, 0, 4, 1, 1.0, 0.0, NULL), LSIM_EINVAL);
  EXPECT(lsim_sv_axpy(fake, NULL, 0, 4, 1, 1.0, 0.0, NULL), LSIM_EINVAL);
  EXPECT(lsim_sv_probabilities(fake, 0, 4, 1, NULL, NULL), LSIM_EINVAL);
  EXPECT(lsim_dm_probabilities(fake, 0, 2, 1, NULL, NULL), LSIM_EINVAL);
  EXPECT(lsim_sv_copy_states(fake, 0, 4, 2, -1, NULL, NULL, NULL), LSIM_EINVAL);
  EXPECT(lsim_program_run(0, 0, 1, fake, fake, fake, fake, NULL, NULL, NULL, NULL, NULL, NULL, 0, NULL, NULL, NULL, NULL, NULL), LSIM_EINVAL);
  EXPECT(lsim_program_run(0, 5, 1, fake, fake, fake, fake, NULL, NULL, NULL, NULL, NULL, NULL, 65, NULL, NULL, NULL, NULL, NULL), LSIM_EINVAL);
  EXPECT(lsim_program_run(0, 5, 1, fake, fake, fake, fake, NULL, NULL, NULL, NULL, NULL, NULL, 0, NULL, NULL, NULL, fake, NULL), LSIM_EINVAL);
  if (lsim_reduce_workspace_bytes(26, 1, 1) == 0) { printf("FAIL workspace bytes\n"); fails++; }
  if (lsim_reduce_workspace_bytes(0, 1, 1) != 0) { printf("FAIL workspace bytes of 0 bits\n"); fails++; }
  /* the host cumulative sum, on sizes around its blocking */
  for (size_t n = 0; n < 5000; n += 37) {
    double* in = (double*)malloc((n + 1) * sizeof(double));
    double* out = (double*)malloc((n + 1) * sizeof(double));
    double acc = 0;
    for (size_t i = 0; i < n; ++i) in[i] = (double)((i * 2654435761u) % 1000) / 997.0;
    EXPECT(lsim_host_cumsum(in, out, n), LSIM_OK);
    for (size_t i = 0; i < n; ++i) { acc += in[i]; if (out[i] != acc) { printf("FAIL cumsum %zu\n", i); fails++; break; } }
    free(in); free(out);
  }
  EXPECT(lsim_host_cumsum(NULL, NULL, 4), LSIM_EINVAL);
  /* the host planners: bad arguments, then a 14-qubit chain of h, rz and cx planned both ways */
  {
    enum { G = 60 };
    int nq[G], qb[5 * G];
    double* m = (double*)calloc(16 * 2 * G, sizeof(double));
    double* at = m;
    memset(qb, 0, sizeof qb);
    for (int i = 0; i < G; ++i) {
      const double s = 0.7071067811865476;
      if (i % 3 == 2) {
        nq[i] = 2; qb[5 * i] = i % 14; qb[5 * i + 1] = (i * 5 + 1) % 14;
        if (qb[5 * i + 1] == qb[5 * i]) qb[5 * i + 1] = (qb[5 * i] + 1) % 14;
        double cx[32] = {1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0, 0, 0};
        memcpy(at, cx, sizeof cx); at += 32;
      } else if (i % 3 == 1) {
        nq[i] = 1; qb[5 * i] = (i * 7) % 14;
        double rz[8] = {0.8, -0.6, 0, 0, 0, 0, 0.8, 0.6};
        memcpy(at, rz, sizeof rz); at += 8;
      } else {
        nq[i] = 1; qb[5 * i] = (i * 3) % 14;
        double h[8] = {s, 0, s, 0, s, 0, -s, 0};
        memcpy(at, h, sizeof h); at += 8;
      }
    }
    lsim_plan* plan = NULL;
    EXPECT(lsim_host_plan_fused(G, nq, qb, m, 14, 6, 0, 0, 1, &plan), LSIM_EINVAL);
    EXPECT(lsim_host_plan_fused(G, nq, qb, m, 10, 4, 0, 0, 1, &plan), LSIM_EINVAL);
    EXPECT(lsim_host_plan_fused(G, NULL, qb, m, 14, 4, 0, 0, 1, &plan), LSIM_EINVAL);
    EXPECT(lsim_host_plan_tiles(G, nq, qb, m, 14, 24.0, 12, 11, 64, &plan), LSIM_EINVAL);
    EXPECT(lsim_host_plan_tiles(G, nq, qb, m, 14, 24.0, 15, 4, 64, &plan), LSIM_EINVAL);
    EXPECT(lsim_host_plan_sizes(NULL, NULL, NULL, NULL), LSIM_EINVAL);
    for (int kind = 0; kind < 2; ++kind) {
      plan = NULL;
      if (kind == 0)
        EXPECT(lsim_host_plan_fused(G, nq, qb, m, 14, 4, 3, 2, 1, &plan), LSIM_OK);
      else
        EXPECT(lsim_host_plan_tiles(G, nq, qb, m, 14, 6.0, 12, 4, 64, &plan), LSIM_OK);
      if (!plan) { printf("FAIL no plan\n"); fails++; continue; }
      int np = 0, no = 0;
      long long nv = 0;
      EXPECT(lsim_host_plan_sizes(plan, &np, &no, &nv), LSIM_OK);
      unsigned long long* tiles = (unsigned long long*)malloc((size_t)(np + 1) * sizeof *tiles);
      int* pass_ops = (int*)malloc((size_t)(2 * np + 1) * sizeof(int));
      int* info = (int*)malloc((size_t)(8 * no + 1) * sizeof(int));
      long long* offs = (long long*)malloc((size_t)(no + 1) * sizeof(long long));
      double* vals = (double*)malloc((size_t)(2 * nv + 2) * sizeof(double));
      EXPECT(lsim_host_plan_read(plan, tiles, pass_ops, info, offs, vals), LSIM_OK);
      int gates = 0;
      for (int i = 0; i < no; ++i) gates += info[8 * i + 7];
      if (gates != G || np < 1) { printf("FAIL plan %d: %d passes, %d gates\n", kind, np, gates); fails++; }
      free(tiles); free(pass_ops); free(info); free(offs); free(vals);
      lsim_host_plan_free(plan);
    }
    lsim_host_plan_free(NULL);
    free(m);
  }
  printf("%s: %d failures\n", fails ? "FAILED" : "ok", fails);
  return fails != 0;
}
