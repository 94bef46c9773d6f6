"""lucivant_sim on AMD Instinct MI355X: the Lucivant Orrery API with HBM-resident engines.

    import lucivant_sim as ls

    result = ls.run('''
        OPENQASM 3;
        include "stdgates.inc";
        qubit[2] q; bit[2] c;
        h q[0]; cx q[0], q[1];
        c = measure q;
    ''', shots=1000, seed=7)
    print(result.counts)

The names exported here are those of lucivant-sim 0.3.0 (src/lucivant_sim/__init__.py:49-76).
Statevector, density-matrix and trajectory simulations run in hand-written gfx950 kernels
(liblsim.so); seeds, sampling, digests and fingerprints match the reference, so saved results
verify on either implementation. ``__version__`` is the reference version the results are
compatible with (it enters every fingerprint).
"""

__version__ = "0.3.0"

from lucivant_sim.circuit import Circuit
from lucivant_sim.clifford import is_clifford
from lucivant_sim.engines import SimulationError
from lucivant_sim.noise import (
    Channel,
    NoiseModel,
    ReadoutError,
    amplitude_damping,
    bit_flip,
    depolarizing,
    from_gate_error,
    pauli_channel,
    phase_damping,
    phase_flip,
    thermal_relaxation,
)
from lucivant_sim.qasm import QasmError, load, parse
from lucivant_sim.simulator import (
    ENGINE,
    Result,
    density_matrix,
    expectation,
    outcome_probabilities,
    probabilities,
    run,
    statevector,
    verify,
)

__all__ = [
    "ENGINE",
    "Channel",
    "Circuit",
    "NoiseModel",
    "QasmError",
    "ReadoutError",
    "Result",
    "SimulationError",
    "__version__",
    "amplitude_damping",
    "bit_flip",
    "density_matrix",
    "depolarizing",
    "expectation",
    "from_gate_error",
    "is_clifford",
    "load",
    "outcome_probabilities",
    "parse",
    "pauli_channel",
    "phase_damping",
    "phase_flip",
    "probabilities",
    "run",
    "statevector",
    "thermal_relaxation",
    "verify",
]
