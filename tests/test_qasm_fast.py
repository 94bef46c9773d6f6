"""The statement fast path of the OpenQASM reader: every program reads to the same circuit, and
every error carries the same message, line and column, with the fast path on and off."""

from __future__ import annotations

import json

import numpy as np
import pytest

from lucivant_sim import qasm
from lucivant_sim import workloads as wl
from lucivant_sim.circuit import Gate


def _both(src: str):  # type: ignore[no-untyped-def]
    out = []
    for fast in (True, False):
        qasm.FAST_STATEMENTS = fast
        try:
            out.append(("ok", qasm.parse(src)))
        except qasm.QasmError as exc:
            out.append(("err", (str(exc), exc.line, exc.column)))
        finally:
            qasm.FAST_STATEMENTS = True
    return out


def _same_circuit(a, b) -> None:  # type: ignore[no-untyped-def]
    assert json.dumps(a.canonical(), sort_keys=True) == json.dumps(b.canonical(), sort_keys=True)
    assert a.digest() == b.digest()
    for x, y in zip(a.operations, b.operations):
        assert type(x) is type(y)
        if isinstance(x, Gate):
            assert x.name == y.name and x.qubits == y.qubits and x.params == y.params
            assert np.array_equal(np.asarray(x.matrix), np.asarray(y.matrix))


HEAD2 = 'OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[4];\nqreg r[2];\ncreg c[4];\n'
HEAD3 = 'OPENQASM 3;\ninclude "stdgates.inc";\nqubit[4] q;\nbit[4] c;\n'

PROGRAMS = [
    wl.random_layers(7, 4, 11),
    wl.hea_ansatz(6, 3, 2),
    wl.teleport_chain(9, 3),
    HEAD2 + "u3(0.1, -2e-3, 3) q[0];\ncx q[0], r[1];\nrz(.5)q[2];\nh q [ 1 ] ;\ncx q[0] ,\n r[0];\nmeasure q -> c;\n",
    HEAD2 + "u3(pi/2, 0, -pi) q[0];\nU(1,2,3) q[1];\nbarrier q[0], q[1];\nreset r[0];\nccx q[0],q[1],q[2];\n",
    HEAD2 + "gate h a { U(0.1,0,0) a; }\nh q[0];\ncx q[1],q[2];\n",  # user gate shadows a built-in
    HEAD2 + "gate g(t) a, b { cx a, b; rz(t) b; }\ng(0.25) q[0], q[3];\nif (c==1) x q[0];\n",
    HEAD3 + "ctrl @ x q[0], q[1];\ninv @ s q[2];\nrx(1e-3) q[3];\nc[0] = measure q[0];\n",
    HEAD3 + "// comment\nh q[0]; /* block */ cx q[0], q[1];\n#pragma keep\nswap q[2], q[3];\n",
    HEAD2 + "rccx q[0], q[1], q[2];\nrc3x q[0], q[1], q[2], q[3];\n",
    HEAD2 + "u3(1_000, 0x1F, 0b11) q[0];\nrz(-0.0) q[1];\n",
]

ERRORS = [
    HEAD2 + "h q[4];\n",  # index out of range
    HEAD2 + "h s[0];\n",  # unknown register
    HEAD2 + "cx q[1], q[1];\n",  # repeated qubit
    HEAD2 + "cx q[1];\n",  # too few qubits
    HEAD2 + "u3(0.1, 0.2) q[0];\n",  # too few parameters
    HEAD2 + "foo q[0];\n",  # unknown gate
    HEAD2 + "h q[0];\n  nope q[1] ,\n  q[2];\n",  # error on a later line of a statement
    HEAD2 + "h q[0];\nh q[0]",  # missing semicolon at the end
    HEAD2 + "x q[0] ,\n r[0];\n",  # arity error in a statement over two lines
]


@pytest.mark.parametrize("src", PROGRAMS)
def test_fast_path_reads_the_same_circuit(src: str) -> None:
    fast, slow = _both(src)
    assert fast[0] == slow[0] == "ok", (fast, slow)
    _same_circuit(fast[1], slow[1])


@pytest.mark.parametrize("src", ERRORS)
def test_fast_path_keeps_errors(src: str) -> None:
    fast, slow = _both(src)
    assert fast[0] == slow[0] == "err"
    assert fast[1] == slow[1]


def test_simple_tokens_are_used() -> None:
    toks = qasm.tokenize(HEAD2 + "u3(0.1,0.2,0.3) q[0];\ncx q[0], q[1];\nmeasure q[0] -> c[0];\n")
    kinds = [t.kind for t in toks]
    assert kinds.count("simple") == 2
    qasm.FAST_STATEMENTS = False
    try:
        assert "simple" not in [t.kind for t in qasm.tokenize(HEAD2 + "h q[0];\n")]
    finally:
        qasm.FAST_STATEMENTS = True
