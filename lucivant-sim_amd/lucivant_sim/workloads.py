"""Seeded synthetic circuits of the benchmark configurations, as OpenQASM 3 text.

Every angle is printed with ``repr`` so the text round-trips exactly, and every generator takes
its randomness from ``numpy.random.default_rng(seed)`` in a fixed order, so the same call gives
byte-identical programs on any machine (the golden fixtures store the text as well).

- :func:`random_layers`: layers of ``u3`` on every qubit (theta, phi, lambda uniform in
  [0, 2 pi)) followed by two-qubit gates on a random disjoint pairing; the 12-qubit, 26-qubit
  and 13-qubit configurations use it with seeds 1001, 1002 and 1004.
- :func:`teleport_chain`: hops of teleportation with mid-circuit measurements and if/else
  corrections whose arms consume the same number of draws, on a wide register of entangled
  spectators (the 20-qubit trajectory configuration).
- :func:`hea_ansatz`: hardware-efficient RY/RZ layers with a CX ladder (serving requests).
- :func:`pauli_strings`: random Pauli labels (qubit 0 rightmost).
"""

from __future__ import annotations

import math

import numpy as np

HEADER = 'OPENQASM 3;\ninclude "stdgates.inc";\n'


def random_layers(
    num_qubits: int,
    layers: int,
    seed: int,
    *,
    two_qubit: tuple[str, ...] = ("cx",),
    measure: bool = True,
) -> str:
    """``layers`` x (u3 on every qubit, then a 2q gate on each pair of a random pairing).

    Per layer the generator draws 3n angles (qubit by qubit: theta, phi, lambda), then one
    permutation of the qubits, then, when more than one 2q gate name is allowed, one choice per
    pair. Pairs are consecutive entries of the permutation; an odd qubit out stays idle."""
    rng = np.random.default_rng(seed)
    lines = [HEADER, f"qubit[{num_qubits}] q;\n"]
    if measure:
        lines.append(f"bit[{num_qubits}] c;\n")
    for _ in range(layers):
        angles = rng.uniform(0.0, 2 * math.pi, size=(num_qubits, 3))
        for q in range(num_qubits):
            t, p, lam = (float(a) for a in angles[q])
            lines.append(f"u3({t!r}, {p!r}, {lam!r}) q[{q}];\n")
        order = rng.permutation(num_qubits)
        pairs = [(int(order[2 * i]), int(order[2 * i + 1])) for i in range(num_qubits // 2)]
        names = [two_qubit[0]] * len(pairs)
        if len(two_qubit) > 1:
            names = [two_qubit[int(i)] for i in rng.integers(0, len(two_qubit), size=len(pairs))]
        for (a, b), name in zip(pairs, names):
            lines.append(f"{name} q[{a}], q[{b}];\n")
    if measure:
        lines.append("c = measure q;\n")
    return "".join(lines)


def teleport_chain(num_qubits: int = 20, hops: int = 3, theta: float = 1.234) -> str:
    """Teleport ry(theta)|0> along ``hops`` hops (qubits 0 -> 2 -> 4 ...), two mid-circuit
    measurements per hop with X/Z corrections conditioned on them; the else arms apply
    ``u(0, 0, 0)`` so both arms run one one-qubit gate (equal draws under a noise model that
    covers one-qubit gates). The remaining qubits hold a GHZ chain built with h/cx. The last
    teleport target and the spectator ends are measured into ``out``."""
    hop_qubits = 2 * hops + 1
    if num_qubits < hop_qubits + 2:
        raise ValueError(f"{hops} hops need at least {hop_qubits + 2} qubits")
    lines = [HEADER, f"qubit[{num_qubits}] q;\n", f"bit[{2 * hops}] m;\n", "bit[3] out;\n"]
    lines.append(f"ry({theta!r}) q[0];\n")
    spect = list(range(hop_qubits, num_qubits))
    lines.append(f"h q[{spect[0]}];\n")
    for a, b in zip(spect, spect[1:]):
        lines.append(f"cx q[{a}], q[{b}];\n")
    for h in range(hops):
        src, mid, dst = 2 * h, 2 * h + 1, 2 * h + 2
        lines.append(f"h q[{mid}];\ncx q[{mid}], q[{dst}];\ncx q[{src}], q[{mid}];\nh q[{src}];\n")
        lines.append(f"m[{2 * h}] = measure q[{src}];\nm[{2 * h + 1}] = measure q[{mid}];\n")
        lines.append(f"if (m[{2 * h + 1}]) {{ x q[{dst}]; }} else {{ u(0, 0, 0) q[{dst}]; }}\n")
        lines.append(f"if (m[{2 * h}]) {{ z q[{dst}]; }} else {{ u(0, 0, 0) q[{dst}]; }}\n")
    last = 2 * hops
    lines.append(f"out[0] = measure q[{last}];\nout[1] = measure q[{spect[0]}];\nout[2] = measure q[{spect[-1]}];\n")
    return "".join(lines)


def hea_ansatz(num_qubits: int, layers: int, seed: int, *, measure: bool = False) -> str:
    """Per layer: ry(theta) and rz(theta') on every qubit, then cx q[i], q[i+1] down the line."""
    rng = np.random.default_rng(seed)
    lines = [HEADER, f"qubit[{num_qubits}] q;\n"]
    if measure:
        lines.append(f"bit[{num_qubits}] c;\n")
    for _ in range(layers):
        angles = rng.uniform(0.0, 2 * math.pi, size=(num_qubits, 2))
        for q in range(num_qubits):
            lines.append(f"ry({float(angles[q, 0])!r}) q[{q}];\nrz({float(angles[q, 1])!r}) q[{q}];\n")
        for q in range(num_qubits - 1):
            lines.append(f"cx q[{q}], q[{q + 1}];\n")
    if measure:
        lines.append("c = measure q;\n")
    return "".join(lines)


def pauli_strings(num_qubits: int, count: int, seed: int) -> list[str]:
    """``count`` random labels over IXYZ (distinct; qubit 0 is the rightmost letter)."""
    rng = np.random.default_rng(seed)
    out: list[str] = []
    while len(out) < count:
        label = "".join("IXYZ"[int(i)] for i in rng.integers(0, 4, size=num_qubits))
        if label not in out:
            out.append(label)
    return out


def config_c1() -> str:
    """12 qubits, 20 layers of u3 + cx, all measured (seed 1001)."""
    return random_layers(12, 20, 1001)


def config_c2() -> tuple[str, list[str]]:
    """26 qubits, 40 layers of u3 + cx/cz (seed 1002), and 8 Pauli strings (seed 1003)."""
    return random_layers(26, 40, 1002, two_qubit=("cx", "cz"), measure=False), pauli_strings(26, 8, 1003)


def config_c3(num_qubits: int = 13, depth: int = 30) -> str:
    """13 qubits, depth 30 of u3 + cx/cz (seed 1004), all measured."""
    return random_layers(num_qubits, depth, 1004, two_qubit=("cx", "cz"))


def noise_c3():  # type: ignore[no-untyped-def]
    """Uniform depolarising 1e-3 (1q) / 1e-2 (2q), thermal relaxation (T1=120, T2=90, t=0.05) on
    u3 -- which, being the more specific rule, replaces the 1q depolariser on u3 -- and readout
    p01=0.01, p10=0.03."""
    from lucivant_sim.noise import NoiseModel, ReadoutError, thermal_relaxation

    return (
        NoiseModel.uniform(single=1e-3, two=1e-2)
        .add_gate_error(thermal_relaxation(120, 90, 0.05), gates="u3")
        .add_readout_error(ReadoutError(p01=0.01, p10=0.03))
    )


def noise_c4():  # type: ignore[no-untyped-def]
    """Uniform depolarising 1e-3 / 1e-2 plus amplitude damping 0.01 on h (replacing h's
    depolariser, the more specific rule)."""
    from lucivant_sim.noise import NoiseModel, amplitude_damping

    return NoiseModel.uniform(single=1e-3, two=1e-2).add_gate_error(amplitude_damping(0.01), gates="h")
