"""A small dynamic circuit and noise model shared by the trajectory tests: strong noise (many
non-identity branches), gates and global phases inside if/else arms, a reset, Kraus channels
right after their gates, readout errors, and a qubit measured again after a gate from outside
(the host must drop its kept reduced density matrix there)."""

from __future__ import annotations

import lucivant_sim as ls

DYNAMIC = """OPENQASM 3;
include "stdgates.inc";
qubit[6] q;
bit[5] c;
h q[0]; cx q[0], q[1]; rz(0.3) q[2]; gphase(0.7); ry(1.1) q[3]; cz q[2], q[3]; cx q[3], q[4];
c[0] = measure q[1];
if (c[0]) { x q[4]; cx q[4], q[5]; } else { ry(0.4) q[4]; cx q[4], q[5]; }
reset q[1];
h q[1]; t q[1]; cx q[1], q[2];
c[1] = measure q[4];
if (c[1]) { gphase(0.2); z q[0]; } else { s q[0]; }
h q[3];
c[2] = measure q[0];
cx q[3], q[0];
c[3] = measure q[0];
c[4] = measure q[5];
"""


def dynamic_noise() -> ls.NoiseModel:
    return (
        ls.NoiseModel.uniform(single=0.05, two=0.1)
        .add_gate_error(ls.amplitude_damping(0.2), gates="h")
        .add_readout_error(ls.ReadoutError(p01=0.02, p10=0.05))
    )
