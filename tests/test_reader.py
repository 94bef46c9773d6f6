"""The OpenQASM reader against lucivant-sim 0.3.0's own reader: circuits (digests, registers,
inputs, operation labels) for a corpus of valid programs, and message, line and column for a
corpus of invalid ones (tests/golden/reader_cases.json, written by
tests/golden/make_reader_cases.py). The reader's one-expression path for plain gate statements
must read every program as the general path does."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

import lucivant_sim as ls
from lucivant_sim import qasm
from lucivant_sim import workloads as wl
from lucivant_sim.circuit import Gate

CASES = json.loads((Path(__file__).parent / "golden" / "reader_cases.json").read_text())


def _ops(c: ls.Circuit) -> list[str]:
    return [repr(op.__class__.__name__) + " " + str(getattr(op, "name", "")) for op in c.operations]


@pytest.mark.parametrize("name", sorted(CASES["valid"]))
def test_programs_read_as_the_reference_reads_them(name: str) -> None:
    rec = CASES["valid"][name]
    c = ls.parse(rec["source"], rec["params"])
    assert c.digest() == rec["digest"]
    assert (c.num_qubits, c.num_clbits) == (rec["num_qubits"], rec["num_clbits"])
    assert [[r.name, r.offset, r.size] for r in c.qregs] == rec["qregs"]
    assert [[r.name, r.offset, r.size] for r in c.cregs] == rec["cregs"]
    assert dict(c.inputs) == rec["inputs"]
    assert _ops(c) == rec["ops"]


@pytest.mark.parametrize("name", sorted(CASES["errors"]))
def test_errors_carry_the_reference_message_and_position(name: str) -> None:
    rec = CASES["errors"][name]
    with pytest.raises(qasm.QasmError) as err:
        ls.parse(rec["source"], rec["params"])
    assert (str(err.value), err.value.line, err.value.column) == (rec["message"], rec["line"], rec["column"])


def _both(src: str) -> list:
    out = []
    for plain in (True, False):
        qasm.PLAIN_STATEMENTS = plain
        try:
            out.append(("ok", qasm.parse(src)))
        except qasm.QasmError as exc:
            out.append(("err", (str(exc), exc.line, exc.column)))
        finally:
            qasm.PLAIN_STATEMENTS = True
    return out


HEAD2 = 'OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[4];\nqreg r[2];\ncreg c[4];\n'
PLAIN = [
    wl.random_layers(7, 4, 11),
    wl.hea_ansatz(6, 3, 2),
    wl.teleport_chain(9, 3),
    HEAD2 + "u3(0.1, -2e-3, 3) q[0];\ncx q[0], r[1];\nrz(.5)q[2];\nh q [ 1 ] ;\ncx q[0] ,\n r[0];\nmeasure q -> c;\n",
    HEAD2 + "u3(+1.5, -0, 1e-300) q[0];\nU(1,2,3) q[1];\nh q[0]; gate h a { U(0.1,0,0) a; } h q[0];\n",
    HEAD2 + "u3(1_000, 0x1F, 0b11) q[0];\nrz(-0.0) q[1];\n",
]
PLAIN_ERRORS = [
    HEAD2 + "h q[4];\n",
    HEAD2 + "h s[0];\n",
    HEAD2 + "cx q[1], q[1];\n",
    HEAD2 + "cx q[1];\n",
    HEAD2 + "u3(0.1, 0.2) q[0];\n",
    HEAD2 + "foo q[0];\n",
    HEAD2 + "measure q[0];\nreset q[0];\nh c[0];\n",
    HEAD2 + "h q[0];\nh q[0]",
    HEAD2 + "h q[0];\nhq[0];\n",  # one name, not h applied to q[0]
]


@pytest.mark.parametrize("src", PLAIN)
def test_plain_statements_read_as_the_general_path_reads_them(src: str) -> None:
    fast, slow = _both(src)
    assert fast[0] == slow[0] == "ok", (fast, slow)
    a, b = fast[1], slow[1]
    assert json.dumps(a.canonical(), sort_keys=True) == json.dumps(b.canonical(), sort_keys=True)
    for x, y in zip(a.operations, b.operations):
        assert type(x) is type(y)
        if isinstance(x, Gate):
            assert x.name == y.name and x.qubits == y.qubits and x.params == y.params
            assert np.array_equal(np.asarray(x.matrix), np.asarray(y.matrix))


@pytest.mark.parametrize("src", PLAIN_ERRORS)
def test_plain_statements_keep_the_general_path_errors(src: str) -> None:
    fast, slow = _both(src)
    assert fast[0] == slow[0] == "err"
    assert fast[1] == slow[1]


def test_include_reads_a_header_and_reports_positions_in_its_file(tmp_path) -> None:  # type: ignore[no-untyped-def]
    (tmp_path / "defs.inc").write_text('OPENQASM 3;\ngate flip a { x a; }\ngate bad a { y a;\n')
    (tmp_path / "ok.inc").write_text('OPENQASM 3;\ngate flip a { x a; }\n')
    (tmp_path / "main.qasm").write_text('OPENQASM 3;\ninclude "ok.inc";\nqubit q;\nflip q;\n')
    c = ls.load(tmp_path / "main.qasm")
    assert [op.name for op in c.operations] == ["flip"]
    (tmp_path / "bad.qasm").write_text('OPENQASM 3;\ninclude "defs.inc";\nqubit q;\n')
    with pytest.raises(qasm.QasmError) as err:
        ls.load(tmp_path / "bad.qasm")
    assert err.value.line == 3 and "never closed" in str(err.value)
