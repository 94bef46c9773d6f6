"""The oracle: a CPU restatement in NumPy of lucivant-sim 0.3.0's engine arithmetic.

TEST INFRASTRUCTURE ONLY. Nothing in the product imports this package; only ``tests/``, the
smoke test in ``__graft_entry__.py`` and ``bench.py``'s CPU-baseline leg use it: to compare the
device results with (or, for the baseline, as the CPU implementation that is timed). Each
function cites the reference file:line whose behaviour it restates.

How each part is pinned (see ``tests/test_oracle.py``):
- gate application: brute force against the full 2^n x 2^n Kronecker operator on every target
  placement (n <= 4), norm preservation, and the reference's own counts fixtures;
- density matrix and channels: rho' = U rho U^dagger and sum K rho K^dagger with dense matrices,
  trace/Hermiticity/positivity, the depolarising composition law and p = 1 limit, the analytic
  values <XX> = 0.8 and P(1) = exp(-0.3) of the reference tests, and rho = |psi><psi| for ideal
  circuits;
- expectations: the dense Pauli matrix at small n, Bell <ZZ> = 1;
- sampling and trajectories: counts fixtures written by the reference itself
  (``tests/golden/``), the README Bell counts, and TVD against the exact distribution.
"""
