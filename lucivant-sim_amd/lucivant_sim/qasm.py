"""An OpenQASM 2 / 3 reader producing :class:`~lucivant_sim.circuit.Circuit`.

The accepted language and the messages of :class:`QasmError` follow lucivant-sim 0.3.0
src/lucivant_sim/qasm.py (its docstring, lines 1-9, and ``parse``/``load`` at 842-885):
register and ``qubit``/``bit`` declarations, the standard gates, user ``gate`` definitions, the
``ctrl @``/``negctrl @``/``inv @``/``pow(k) @`` modifiers, ``input``/``const`` values,
broadcasting over registers, slices and index sets, both ``measure`` spellings, ``reset``,
``barrier``, ``gphase``, ``box``, ``delay`` and ``if``/``else`` on classical bits. Loops,
subroutines and calibration blocks are refused by name.

The reader is written from that description: a regular-expression tokenizer and a
recursive-descent parser with one token of lookahead (a few more for modifiers).
"""

from __future__ import annotations

import math
import re
from collections.abc import Callable, Mapping
from dataclasses import dataclass
from pathlib import Path
from typing import Any, NamedTuple

import numpy as np

from lucivant_sim.circuit import (
    Barrier,
    Circuit,
    Condition,
    Gate,
    GlobalPhase,
    IfElse,
    Measure,
    Operation,
    Register,
    Reset,
)
from lucivant_sim.gates import GATES, Matrix, canonical_name, controlled, gate_matrix, matrix_power, negcontrolled

#: Declared qubits beyond this are refused before anything is allocated.
MAX_QUBITS = 2048

STANDARD_INCLUDES = frozenset({"qelib1.inc", "stdgates.inc"})

#: qelib1.inc gates given there as decompositions rather than matrices.
QELIB_DECOMPOSED = """
gate rccx a, b, c { u2(0, pi) c; u1(pi/4) c; cx b, c; u1(-pi/4) c; cx a, c; u1(pi/4) c; cx b, c;
  u1(-pi/4) c; u2(0, pi) c; }
gate rc3x a, b, c, d { u2(0, pi) d; u1(pi/4) d; cx c, d; u1(-pi/4) d; u2(0, pi) d; cx a, d;
  u1(pi/4) d; cx b, d; u1(-pi/4) d; cx a, d; u1(pi/4) d; cx b, d; u1(-pi/4) d; u2(0, pi) d;
  u1(pi/4) d; cx c, d; u1(-pi/4) d; u2(0, pi) d; }
"""

UNSUPPORTED = frozenset({"for", "while", "def", "defcal", "cal", "extern", "opaque", "return", "break", "continue"})


class QasmError(ValueError):
    """The program can't be read; ``line`` and ``column`` locate the problem (1-based)."""

    def __init__(self, message: str, line: int = 0, column: int = 0) -> None:
        self.line = line
        self.column = column
        self.message = message
        super().__init__(f"line {line}, column {column}: {message}" if line else message)

    def __reduce__(self) -> tuple:
        # keeps line and column when the error crosses a process boundary
        return (type(self), (self.message, self.line, self.column))


# ----------------------------------------------------------------------------------- lexing


class Tok(NamedTuple):
    kind: str  # id, num, str, phys, op, eof
    text: str
    line: int
    col: int


_LEX = re.compile(
    r"(?P<ws>\s+)|(?P<line_comment>//[^\n]*)|(?P<block_comment>/\*.*?\*/)"
    r"|(?P<num>0[bB][01_]+|0[xX][0-9a-fA-F_]+|0[oO][0-7_]+|(?:\d[\d_]*\.?[\d_]*|\.\d[\d_]*)(?:[eE][+-]?\d+)?)"
    r"|(?P<str>\"[^\"\n]*\"|'[^'\n]*')|(?P<phys>\$\d+)|(?P<id>[^\W\d]\w*)"
    r"|(?P<op>->|==|!=|<=|>=|&&|\|\||\*\*|<<|>>|[\[\](){};,:=+\-*/^%<>!~@&|.])",
    re.DOTALL,
)
_PRAGMA = re.compile(r"(?m)^[ \t]*(?:#?pragma|@)[^\n]*")
_SKIP = frozenset({"ws", "line_comment", "block_comment"})

# A whole gate-call statement with literal-number arguments and indexed qubits, e.g.
# ``u3(0.1, -2e-3, 3) q[0];`` or ``cx q[0], q[11];``: lexed as one ``simple`` token when it starts a
# statement, so the common case costs one regex match instead of a dozen tokens. Statement
# keywords are excluded; the parser reads such a token straight into a Gate, or lexes its text
# into ordinary tokens when it is anything else (a user-defined gate, an error).
_SNUM = r"-?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?"
_SARG = r"([A-Za-z_]\w*)\s*\[\s*(\d+)\s*\]"
_SARG_NC = r"[A-Za-z_]\w*\s*\[\s*\d+\s*\]"
_SIMPLE = re.compile(
    r"(?!(?:measure|reset|barrier|qreg|creg|qubit|bit|include|gate|opaque|if|else|for|while|box|delay"
    r"|gphase|input|output|const|float|angle|int|uint|bool|def|defcal|cal|return|let|extern|ctrl|negctrl"
    r"|inv|pow)\b)"
    rf"([A-Za-z_]\w*)(?:\s*\(\s*({_SNUM}(?:\s*,\s*{_SNUM})*)\s*\)\s*|\s+)"
    rf"({_SARG_NC}(?:\s*,\s*{_SARG_NC})*)\s*;"
)
_SIMPLE_ARG = re.compile(_SARG)


class _SimpleTok(NamedTuple):
    """A ``simple`` token with its regex groups (name, parameter text, qubit text)."""

    kind: str
    text: str
    line: int
    col: int
    groups: tuple[str, str | None, str]
_BOUNDARY = frozenset({";", "{", "}"})
#: lex simple gate calls as single tokens (False: every statement takes the general path)
FAST_STATEMENTS = True


def _lex_span(text: str, line: int, col: int) -> list[Tok]:
    """Ordinary tokens of ``text``, which starts at ``line``/``col`` of the source."""
    out: list[Tok] = []
    line_start = -(col - 1)
    for m in _LEX.finditer(text):
        kind = m.lastgroup
        start = m.start()
        if kind in _SKIP:
            part = m.group()
            nl = part.count("\n")
            if nl:
                line += nl
                line_start = start + part.rfind("\n") + 1
            continue
        out.append(Tok(kind, m.group(), line, start - line_start + 1))  # type: ignore[arg-type]
    return out


def tokenize(source: str) -> list[Tok]:
    """Tokens of ``source`` with 1-based line/column, ending in one ``eof`` token. Pragma and
    annotation lines are dropped (they don't change the circuit). Simple gate calls at the start
    of a statement come as single ``simple`` tokens (see ``_SIMPLE``)."""
    source = _PRAGMA.sub("", source)
    out: list[Tok] = []
    append = out.append
    lex = _LEX.match
    simple = _SIMPLE.match
    line, line_start, pos, end = 1, 0, 0, len(source)
    boundary = FAST_STATEMENTS
    while pos < end:
        if boundary:
            m = simple(source, pos)
            if m is not None:
                text = m.group()
                append(_SimpleTok("simple", text, line, pos - line_start + 1, m.groups()))
                nl = text.count("\n")
                if nl:
                    line += nl
                    line_start = pos + text.rfind("\n") + 1
                pos = m.end()
                continue
        m = lex(source, pos)
        if m is None:
            raise QasmError(f"unexpected character {source[pos]!r}", line, pos - line_start + 1)
        kind = m.lastgroup
        start, pos = pos, m.end()
        text = m.group()
        if kind in _SKIP:
            nl = text.count("\n")
            if nl:
                line += nl
                line_start = start + text.rfind("\n") + 1
            continue
        append(Tok(kind, text, line, start - line_start + 1))  # type: ignore[arg-type]
        boundary = FAST_STATEMENTS and kind == "op" and text in _BOUNDARY
    append(Tok("eof", "", line, pos - line_start + 1))
    return out


def _number(text: str) -> float:
    t = text.replace("_", "").lower()
    for prefix, base in (("0b", 2), ("0x", 16), ("0o", 8)):
        if t.startswith(prefix):
            return float(int(t[2:], base))
    return float(t)


# ------------------------------------------------------------------------------ expressions

_CONSTANTS = {"pi": math.pi, "π": math.pi, "tau": math.tau, "τ": math.tau, "euler": math.e, "ℇ": math.e}
_FUNCTIONS: dict[str, Callable[[float], float]] = {
    "sin": math.sin,
    "cos": math.cos,
    "tan": math.tan,
    "arcsin": math.asin,
    "asin": math.asin,
    "arccos": math.acos,
    "acos": math.acos,
    "arctan": math.atan,
    "atan": math.atan,
    "exp": math.exp,
    "ln": math.log,
    "log": math.log,
    "sqrt": math.sqrt,
}

# An expression tree: ("num", value) | ("var", name, tok) | ("neg", e) | ("call", fn, e, tok)
# | ("bin", op, a, b, tok)
Expr = tuple[Any, ...]


@dataclass
class _Call:
    """One gate call inside a ``gate`` body, kept unevaluated until the gate is used."""

    modifiers: list[tuple[str, Expr | None]]
    name: Tok
    params: list[Expr]
    args: list[str]


@dataclass
class _GateDef:
    params: list[str]
    qubits: list[str]
    body: list[_Call | Expr]  # an Expr entry is a gphase angle


# ---------------------------------------------------------------------------------- parsing


class _Reader:
    def __init__(self, source: str, inputs: Mapping[str, float], include: Callable[[str], str] | None) -> None:
        self.toks = tokenize(source)
        self.i = 0
        self.inputs = dict(inputs)
        self.include = include
        # shared with nested includes
        self.qregs: dict[str, Register] = {}
        self.cregs: dict[str, Register] = {}
        self.values: dict[str, float] = {}
        self.used_inputs: dict[str, float] = {}
        self.gates: dict[str, _GateDef] = {}
        self.user_defined: set[str] = set()
        self.num_qubits = 0
        self.num_clbits = 0
        self.physical = -1

    # -- token helpers ---------------------------------------------------------------------

    def peek(self, ahead: int = 0) -> Tok:
        j = self.i + ahead
        return self.toks[j] if j < len(self.toks) else self.toks[-1]

    def take(self) -> Tok:
        t = self.peek()
        if t.kind != "eof":
            self.i += 1
        return t

    def err(self, message: str, tok: Tok | None = None) -> QasmError:
        t = tok or self.peek()
        return QasmError(message, t.line, t.col)

    def is_(self, text: str, ahead: int = 0) -> bool:
        t = self.peek(ahead)
        return t.text == text and t.kind in ("op", "id")

    def accept(self, text: str) -> bool:
        if self.is_(text):
            self.i += 1
            return True
        return False

    def expect(self, text: str) -> Tok:
        t = self.peek()
        if not self.is_(text):
            found = "the end of the program" if t.kind == "eof" else repr(t.text)
            raise self.err(f"expected {text!r} but found {found}", t)
        self.i += 1
        return t

    def name(self) -> Tok:
        t = self.take()
        if t.kind != "id":
            raise self.err(f"expected a name but found {t.text!r}", t)
        return t

    # -- expressions -----------------------------------------------------------------------

    def expr(self) -> Expr:
        node = self.product()
        while self.peek().kind == "op" and self.peek().text in ("+", "-"):
            op = self.take()
            node = ("bin", op.text, node, self.product(), op)
        return node

    def product(self) -> Expr:
        node = self.signed()
        while self.peek().kind == "op" and self.peek().text in ("*", "/", "%"):
            op = self.take()
            node = ("bin", op.text, node, self.signed(), op)
        return node

    def signed(self) -> Expr:
        if self.peek().kind == "op" and self.peek().text in ("+", "-"):
            negative = self.take().text == "-"
            inner = self.signed()
            return ("neg", inner) if negative else inner
        base = self.atom()
        if self.peek().kind == "op" and self.peek().text in ("**", "^"):
            op = self.take()
            return ("bin", "**", base, self.signed(), op)
        return base

    def atom(self) -> Expr:
        t = self.take()
        if t.kind == "num":
            return ("num", _number(t.text))
        if t.kind == "op" and t.text == "(":
            inner = self.expr()
            self.expect(")")
            return inner
        if t.kind == "id":
            if t.text in _FUNCTIONS and self.is_("("):
                self.take()
                arg = self.expr()
                self.expect(")")
                return ("call", t.text, arg, t)
            return ("var", t.text, t)
        raise self.err(f"expected a number or an expression but found {t.text!r}", t)

    def value(self, e: Expr, env: Mapping[str, float] | None = None) -> float:
        kind = e[0]
        if kind == "num":
            return float(e[1])
        if kind == "var":
            name = e[1]
            if env is not None and name in env:
                return env[name]
            if name in self.values:
                return self.values[name]
            if name in _CONSTANTS:
                return _CONSTANTS[name]
            raise self.err(f"unknown name {name!r}", e[2])
        if kind == "neg":
            return -self.value(e[1], env)
        if kind == "call":
            try:
                return _FUNCTIONS[e[1]](self.value(e[2], env))
            except ValueError as exc:
                raise self.err(f"{e[1]}: {exc}", e[3]) from exc
        op, a, b = e[1], self.value(e[2], env), self.value(e[3], env)
        try:
            if op == "+":
                return a + b
            if op == "-":
                return a - b
            if op == "*":
                return a * b
            if op == "/":
                return a / b
            if op == "%":
                return math.fmod(a, b)
            return float(a**b)
        except (ZeroDivisionError, OverflowError, TypeError) as exc:
            raise self.err(f"can't evaluate the expression: {exc}", e[4]) from exc

    def whole(self, e: Expr, what: str) -> int:
        v = self.value(e)
        if not math.isfinite(v) or abs(v - round(v)) > 1e-9:
            raise self.err(f"{what} must be a whole number")
        return round(v)

    # -- program ---------------------------------------------------------------------------

    def program(self) -> list[Operation]:
        if self.is_("OPENQASM"):
            self.take()
            version = self.take()
            if version.kind != "num" or version.text.split(".")[0] not in ("2", "3"):
                raise self.err("only OpenQASM 2 and 3 are supported", version)
            self.expect(";")
        ops: list[Operation] = []
        while self.peek().kind != "eof":
            ops.extend(self.statement())
        return ops

    def body(self) -> list[Operation]:
        """A braced block or a single statement."""
        if not self.accept("{"):
            return self.statement()
        ops: list[Operation] = []
        while not self.accept("}"):
            if self.peek().kind == "eof":
                raise self.err("this block is never closed with '}'")
            ops.extend(self.statement())
        return ops

    def statement(self) -> list[Operation]:
        t = self.peek()
        if t.kind == "simple":
            return self.simple_statement(t)
        word = t.text if t.kind == "id" else ""
        if t.kind == "op" and t.text == ";":
            self.take()
            return []
        if word == "include":
            return self.include_statement()
        if word in ("qreg", "creg"):
            self.take()
            reg = self.name()
            self.expect("[")
            size = self.whole(self.expr(), "a register size")
            self.expect("]")
            self.expect(";")
            self.declare(reg, size, quantum=word == "qreg")
            return []
        if word in ("qubit", "bit"):
            self.take()
            size = 1
            if self.accept("["):
                size = self.whole(self.expr(), "a register size")
                self.expect("]")
            reg = self.name()
            self.declare(reg, size, quantum=word == "qubit")
            if word == "bit" and self.accept("="):
                self.expect("measure")
                qubits = self.qubits()
                self.expect(";")
                return self.measures(qubits, self.whole_creg(reg.text))
            self.expect(";")
            return []
        if word == "gate":
            self.gate_definition()
            return []
        if word in ("input", "const", "float", "angle", "int", "uint", "bool"):
            self.classical()
            return []
        if word == "measure":
            self.take()
            qubits = self.qubits()
            if self.accept("->"):
                bits = self.bits()
                self.expect(";")
                return self.measures(qubits, bits)
            self.expect(";")
            return [Measure(q, None) for q in qubits[0]]
        if word == "reset":
            self.take()
            qubits = self.qubits()
            self.expect(";")
            return [Reset(q) for q in qubits[0]]
        if word == "barrier":
            self.take()
            listed: list[int] = []
            while not self.is_(";"):
                listed.extend(self.qubits()[0])
                if not self.accept(","):
                    break
            self.expect(";")
            return [Barrier(tuple(listed) or tuple(range(self.num_qubits)))]
        if word == "if":
            return [self.if_statement()]
        if word == "box":
            self.take()
            if self.accept("["):
                self.skip_to("]")
            return self.body()
        if word == "delay":
            self.take()
            self.expect("[")
            self.skip_to("]")
            while not self.is_(";") and self.peek().kind != "eof":
                self.take()
            self.expect(";")
            return []
        if word in UNSUPPORTED:
            raise self.err(f"'{word}' isn't supported by Lucivant Orrery yet; unroll it or define a gate instead")
        if word == "gphase":
            self.take()
            self.expect("(")
            angle = self.value(self.expr())
            self.expect(")")
            self.expect(";")
            return [GlobalPhase(angle)]
        if t.kind == "id" and t.text in self.cregs and (self.is_("=", 1) or self.is_("[", 1)):
            bits = self.bits()
            self.expect("=")
            if not self.accept("measure"):
                raise self.err("only measurement results can be assigned to bits")
            qubits = self.qubits()
            self.expect(";")
            return self.measures(qubits, bits)
        if t.kind == "id":
            fast = self.simple_gate_statement()
            if fast is not None:
                return fast
            return self.gate_statement()
        raise self.err(f"unexpected {t.text!r}" if t.kind != "eof" else "unexpected end of program")

    def simple_statement(self, tok: Tok) -> list[Operation]:
        """A ``simple`` token: a built-in gate on in-range, distinct qubits becomes its Gate
        directly (the same Gate :meth:`gate_statement` makes); anything else is lexed into
        ordinary tokens and read by the general path, with its usual errors."""
        name, params_text, args_text = tok.groups  # type: ignore[attr-defined]
        spec = GATES.get(name)
        if spec is not None and name not in self.gates:
            # _SNUM admits only plain decimal literals, which float() reads exactly as _number does
            params = [float(x) for x in params_text.split(",")] if params_text else []
            qubits: list[int] = []
            for reg_name, index_text in _SIMPLE_ARG.findall(args_text):
                reg = self.qregs.get(reg_name)
                index = int(index_text)
                if reg is None or index >= reg.size:
                    qubits = []
                    break
                qubits.append(reg.offset + index)
            if qubits and len(params) == spec.params and len(qubits) == spec.qubits and len(set(qubits)) == len(qubits):
                self.i += 1
                pt = tuple(params)
                return [Gate(canonical_name(name), tuple(qubits), gate_matrix(name, pt), pt)]
        # the general path over this statement's ordinary tokens
        sub = _lex_span(tok.text, tok.line, tok.col)
        last = sub[-1]
        saved, saved_i = self.toks, self.i
        self.toks, self.i = sub + [Tok("eof", "", last.line, last.col + len(last.text))], 0
        try:
            ops = self.statement()
            if self.peek().kind != "eof":
                raise self.err(f"unexpected {self.peek().text!r}")
        finally:
            self.toks, self.i = saved, saved_i
        self.i += 1
        return ops

    def simple_gate_statement(self) -> list[Operation] | None:
        """``name(num, -num, ...) reg[int], reg[int], ...;`` for a built-in gate, read straight
        from the token list; None (nothing consumed) for anything else, which then takes the
        general path and its error messages. Produces exactly what :meth:`gate_statement`
        produces for such a statement."""
        toks = self.toks
        j = self.i
        name = toks[j].text
        spec = GATES.get(name)
        if spec is None or name in self.gates:
            return None
        j += 1
        params: list[float] = []
        tok = toks[j]
        if tok.text == "(" and tok.kind == "op":
            j += 1
            while True:
                tok = toks[j]
                negative = tok.text == "-" and tok.kind == "op"
                if negative:
                    j += 1
                    tok = toks[j]
                if tok.kind != "num":
                    return None
                v = _number(tok.text)
                params.append(-v if negative else v)
                j += 1
                sep = toks[j]
                j += 1
                if sep.text == ")" and sep.kind == "op":
                    break
                if sep.text != "," or sep.kind != "op":
                    return None
        qubits: list[int] = []
        while True:
            reg_tok = toks[j]
            reg = self.qregs.get(reg_tok.text) if reg_tok.kind == "id" else None
            if reg is None or toks[j + 1].text != "[" or toks[j + 2].kind != "num" or toks[j + 3].text != "]":
                return None
            text = toks[j + 2].text
            if not text.isdigit():
                return None
            index = int(text)
            if index >= reg.size:
                return None
            qubits.append(reg.offset + index)
            sep = toks[j + 4]
            j += 5
            if sep.kind != "op":
                return None
            if sep.text == ";":
                break
            if sep.text != ",":
                return None
        if len(params) != spec.params or len(qubits) != spec.qubits or len(set(qubits)) != len(qubits):
            return None
        self.i = j
        pt = tuple(params)
        return [Gate(canonical_name(name), tuple(qubits), gate_matrix(name, pt), pt)]

    def include_statement(self) -> list[Operation]:
        self.take()
        file = self.take()
        if file.kind != "str":
            raise self.err("include needs a file name in quotes", file)
        self.expect(";")
        name = file.text[1:-1]
        if name in STANDARD_INCLUDES:
            return []
        if self.include is None:
            raise self.err(f"can't include {name!r}: only qelib1.inc and stdgates.inc are built in", file)
        nested = _Reader(self.include(name), self.inputs, self.include)
        for attr in ("qregs", "cregs", "values", "gates", "user_defined", "used_inputs"):
            setattr(nested, attr, getattr(self, attr))
        nested.num_qubits, nested.num_clbits, nested.physical = self.num_qubits, self.num_clbits, self.physical
        ops = nested.program()
        self.num_qubits, self.num_clbits, self.physical = nested.num_qubits, nested.num_clbits, nested.physical
        return ops

    def skip_to(self, closing: str) -> None:
        depth = 0
        while True:
            t = self.take()
            if t.kind == "eof":
                raise self.err(f"expected {closing!r}")
            if t.text == "[":
                depth += 1
            elif t.text == closing:
                if depth == 0:
                    return
                depth -= 1

    # -- declarations ----------------------------------------------------------------------

    def declare(self, reg: Tok, size: int, *, quantum: bool) -> None:
        if reg.text in self.qregs or reg.text in self.cregs or reg.text in self.values:
            raise self.err(f"{reg.text!r} is already declared", reg)
        if size < 1:
            raise self.err("a register needs at least one bit", reg)
        if quantum:
            if self.physical >= 0:
                raise self.err("a program can't mix declared qubits with physical qubits ($n)", reg)
            if self.num_qubits + size > MAX_QUBITS:
                raise self.err(f"more than {MAX_QUBITS} qubits", reg)
            self.qregs[reg.text] = Register(reg.text, self.num_qubits, size)
            self.num_qubits += size
        else:
            self.cregs[reg.text] = Register(reg.text, self.num_clbits, size)
            self.num_clbits += size

    def classical(self) -> None:
        is_input = self.accept("input")
        self.accept("const")
        kind = self.name()
        if kind.text not in ("float", "angle", "int", "uint", "bool"):
            raise self.err(f"'{kind.text}' parameters aren't supported; use float or angle", kind)
        if self.accept("["):
            self.skip_to("]")
        var = self.name()
        if var.text in self.values:
            raise self.err(f"{var.text!r} is already declared", var)
        if is_input:
            if var.text not in self.inputs:
                raise self.err(f"the program needs a value for its input {var.text!r}", var)
            v = float(self.inputs[var.text])
            self.used_inputs[var.text] = v
        else:
            self.expect("=")
            v = self.value(self.expr())
        self.expect(";")
        self.values[var.text] = v

    # -- operands --------------------------------------------------------------------------

    def index_list(self, size: int, reg: Tok) -> list[int]:
        """The inside of ``name[...]`` after the opening bracket, resolved to positions."""
        if self.accept("{"):
            picked = [self.whole(self.expr(), "an index")]
            while self.accept(","):
                picked.append(self.whole(self.expr(), "an index"))
            self.expect("}")
        else:
            start = self.whole(self.expr(), "an index")
            if self.accept(":"):
                second = self.whole(self.expr(), "an index")
                if self.accept(":"):
                    step, stop = second, self.whole(self.expr(), "an index")
                else:
                    step, stop = 1, second
                if step == 0:
                    raise self.err("a range's step can't be 0", reg)
                picked = list(range(start, stop + (1 if step > 0 else -1), step))
            else:
                picked = [start]
        self.expect("]")
        out = []
        for i in picked:
            j = i + size if i < 0 else i
            if not 0 <= j < size:
                raise self.err(f"index {i} is outside {reg.text!r}, which has {size}", reg)
            out.append(j)
        return out

    def qubits(self, local: Mapping[str, int] | None = None) -> tuple[list[int], bool]:
        """Qubits an operand names, and whether it was a whole register (broadcast)."""
        t = self.take()
        if t.kind == "phys":
            if self.qregs:
                raise self.err("a program can't mix declared qubits with physical qubits ($n)", t)
            q = int(t.text[1:])
            if q >= MAX_QUBITS:
                raise self.err(f"more than {MAX_QUBITS} qubits", t)
            self.physical = max(self.physical, q)
            self.num_qubits = self.physical + 1
            return [q], False
        if t.kind != "id":
            raise self.err(f"expected a qubit but found {t.text!r}", t)
        if local is not None:
            if t.text not in local:
                raise self.err(f"{t.text!r} isn't an argument of this gate", t)
            return [local[t.text]], False
        reg = self.qregs.get(t.text)
        if reg is None:
            raise self.err(f"unknown qubit register {t.text!r}", t)
        if self.accept("["):
            return [reg.offset + i for i in self.index_list(reg.size, t)], False
        return list(range(reg.offset, reg.offset + reg.size)), True

    def bits(self) -> tuple[list[int], bool]:
        t = self.name()
        reg = self.cregs.get(t.text)
        if reg is None:
            raise self.err(f"unknown classical register {t.text!r}", t)
        if self.accept("["):
            return [reg.offset + i for i in self.index_list(reg.size, t)], False
        return list(range(reg.offset, reg.offset + reg.size)), True

    def whole_creg(self, name: str) -> tuple[list[int], bool]:
        reg = self.cregs[name]
        return list(range(reg.offset, reg.offset + reg.size)), True

    def measures(self, qubits: tuple[list[int], bool], bits: tuple[list[int], bool]) -> list[Operation]:
        if len(qubits[0]) != len(bits[0]):
            raise self.err(f"can't measure {len(qubits[0])} qubits into {len(bits[0])} bits")
        return [Measure(q, c) for q, c in zip(qubits[0], bits[0])]

    # -- conditions ------------------------------------------------------------------------

    def if_statement(self) -> IfElse:
        self.expect("if")
        self.expect("(")
        negate = self.accept("!")
        reg = self.name()
        if reg.text not in self.cregs:
            raise self.err(f"unknown classical register {reg.text!r}", reg)
        self.i -= 1
        bits = self.bits()[0]
        value = 1
        if self.is_("==") or self.is_("!="):
            if self.take().text == "!=":
                negate = not negate
            v = self.take()
            if v.kind == "num":
                value = int(_number(v.text))
            elif v.text in ("true", "false"):
                value = int(v.text == "true")
            else:
                raise self.err("a condition compares with a whole number", v)
        elif len(bits) != 1:
            raise self.err("compare a register with a value, e.g. if (c == 1)", reg)
        self.expect(")")
        then = self.body()
        otherwise: list[Operation] = []
        if self.accept("else"):
            otherwise = self.body()
        return IfElse(Condition(tuple(bits), value, negate), tuple(then), tuple(otherwise))

    # -- gates -----------------------------------------------------------------------------

    def modifiers(self) -> list[tuple[str, Expr | None]]:
        mods: list[tuple[str, Expr | None]] = []
        while self.peek().kind == "id" and self.peek().text in ("ctrl", "negctrl", "inv", "pow"):
            # only a modifier when an '@' follows (after an optional parenthesised argument)
            ahead = 1
            if self.is_("(", 1):
                depth = 0
                while True:
                    t = self.peek(ahead)
                    if t.kind == "eof":
                        break
                    if t.text == "(":
                        depth += 1
                    elif t.text == ")":
                        depth -= 1
                        if depth == 0:
                            break
                    ahead += 1
                ahead += 1
            if not self.is_("@", ahead):
                break
            word = self.take().text
            arg = None
            if self.accept("("):
                arg = self.expr()
                self.expect(")")
            self.expect("@")
            mods.append((word, arg))
        return mods

    def call_head(self) -> tuple[list[tuple[str, Expr | None]], Tok, list[Expr]]:
        mods = self.modifiers()
        name = self.name()
        params: list[Expr] = []
        if self.accept("(") and not self.accept(")"):
            params.append(self.expr())
            while self.accept(","):
                params.append(self.expr())
            self.expect(")")
        return mods, name, params

    def gate_statement(self) -> list[Operation]:
        mods, name, exprs = self.call_head()
        params = tuple(self.value(e) for e in exprs)
        operands: list[tuple[list[int], bool]] = []
        while not self.is_(";"):
            operands.append(self.qubits())
            if not self.accept(","):
                break
        self.expect(";")
        if name.text == "gphase" and not operands:
            if len(params) != 1:
                raise self.err("gphase takes one angle", name)
            return [GlobalPhase(params[0])]
        matrix, arity = self.resolve(mods, name, params)
        if len(operands) != arity:
            raise self.err(f"{name.text} acts on {arity} qubit(s) but was given {len(operands)}", name)
        widths = {len(qs) for qs, whole in operands if whole or len(qs) > 1}
        if len(widths) > 1:
            raise self.err("registers of different sizes can't be combined in one gate", name)
        times = widths.pop() if widths else 1
        label = "".join(f"{m}@" for m, _ in mods) + canonical_name(name.text)
        ops: list[Operation] = []
        for i in range(times):
            qs = tuple(q[i] if len(q) > 1 else q[0] for q, _ in operands)
            if len(set(qs)) != len(qs):
                raise self.err(f"{name.text} is given the same qubit twice", name)
            ops.append(Gate(label, qs, matrix, params))
        return ops

    def resolve(
        self, mods: list[tuple[str, Expr | None]], name: Tok, params: tuple[float, ...], env: Mapping[str, float] | None = None
    ) -> tuple[Matrix, int]:
        """The matrix and arity of ``mods name(params)``; the modifier nearest the name applies
        first."""
        matrix, arity = self.base(name, params)
        for word, arg in reversed(mods):
            if word == "inv":
                matrix = matrix.conj().T
            elif word == "pow":
                if arg is None:
                    raise self.err("pow needs an exponent, e.g. pow(2) @", name)
                matrix = matrix_power(matrix, self.value(arg, env))
            else:
                count = 1 if arg is None else self.whole(arg, "the number of controls")
                for _ in range(count):
                    matrix = controlled(matrix) if word == "ctrl" else negcontrolled(matrix)
                arity += count
        return matrix, arity

    def base(self, name: Tok, params: tuple[float, ...]) -> tuple[Matrix, int]:
        user = self.gates.get(name.text)
        if user is not None:
            if len(params) != len(user.params):
                raise self.err(f"{name.text} takes {len(user.params)} parameter(s) but was given {len(params)}", name)
            return self.user_unitary(user, params), len(user.qubits)
        spec = GATES.get(name.text)
        if spec is None:
            raise self.err(f"unknown gate {name.text!r}", name)
        if len(params) != spec.params:
            raise self.err(f"{name.text} takes {spec.params} parameter(s) but was given {len(params)}", name)
        return gate_matrix(name.text, params), spec.qubits

    def user_unitary(self, gate: _GateDef, params: tuple[float, ...]) -> Matrix:
        """The unitary of a user gate: its body applied in order to the identity."""
        k = len(gate.qubits)
        env = dict(zip(gate.params, params))
        where = {q: i for i, q in enumerate(gate.qubits)}
        op = np.eye(2**k, dtype=np.complex128)
        for item in gate.body:
            if not isinstance(item, _Call):
                op = op * np.exp(1j * self.value(item, env))
                continue
            values = tuple(self.value(p, env) for p in item.params)
            sub, arity = self.resolve(item.modifiers, item.name, values, env)
            if len(item.args) != arity:
                raise self.err(f"{item.name.text} acts on {arity} qubit(s) but was given {len(item.args)}", item.name)
            positions = [where[a] for a in item.args]
            if len(set(positions)) != len(positions):
                raise self.err(f"{item.name.text} is given the same qubit twice", item.name)
            op = embed(sub, positions, k) @ op
        return op

    def gate_definition(self) -> None:
        self.expect("gate")
        name = self.name()
        if name.text in self.user_defined:
            raise self.err(f"gate {name.text!r} is already defined", name)
        params: list[str] = []
        if self.accept("(") and not self.accept(")"):
            params.append(self.name().text)
            while self.accept(","):
                params.append(self.name().text)
            self.expect(")")
        qubits = [self.name().text]
        while self.accept(","):
            qubits.append(self.name().text)
        if len(set(qubits)) != len(qubits):
            raise self.err("a gate's qubit arguments must have different names", name)
        self.expect("{")
        body: list[_Call | Expr] = []
        while not self.accept("}"):
            if self.peek().kind == "eof":
                raise self.err(f"gate {name.text!r} is never closed with '}}'", name)
            if self.is_("barrier"):
                while self.take().text != ";":
                    pass
                continue
            mods, callee, exprs = self.call_head()
            if callee.text == "gphase":
                self.expect(";")
                if len(exprs) != 1:
                    raise self.err("gphase takes one angle", callee)
                body.append(exprs[0])
                continue
            args: list[str] = []
            while not self.is_(";"):
                a = self.name()
                if a.text not in qubits:
                    raise self.err(f"{a.text!r} isn't an argument of gate {name.text!r}", a)
                args.append(a.text)
                if not self.accept(","):
                    break
            self.expect(";")
            if callee.text == name.text:
                raise self.err("a gate can't call itself", callee)
            body.append(_Call(mods, callee, exprs, args))
        self.gates[name.text] = _GateDef(params, qubits, body)
        self.user_defined.add(name.text)

    def define_prelude(self) -> None:
        """The relative-phase Toffolis qelib1.inc defines by decomposition (not user gates, so a
        program may redefine them); parsed once per process."""
        global _PRELUDE_GATES
        if _PRELUDE_GATES is None:
            saved = (self.toks, self.i)
            self.toks, self.i = tokenize(QELIB_DECOMPOSED), 0
            while self.peek().kind != "eof":
                self.gate_definition()
            self.toks, self.i = saved
            self.user_defined.clear()
            _PRELUDE_GATES = dict(self.gates)
        self.gates.update(_PRELUDE_GATES)


def embed(sub: Matrix, positions: list[int], k: int) -> Matrix:
    """``sub`` acting on local qubits ``positions`` (first = its most significant bit) of a
    k-qubit register whose local qubit j is bit k-1-j of the index, as a 2^k x 2^k matrix."""
    m = len(positions)
    d = 2**k
    rows = np.arange(d)
    # index bits of the positions, most significant first
    sub_index = np.zeros(d, dtype=np.int64)
    for i, p in enumerate(positions):
        sub_index |= ((rows >> (k - 1 - p)) & 1) << (m - 1 - i)
    mask = 0
    for p in positions:
        mask |= 1 << (k - 1 - p)
    rest = rows & ~mask
    same_rest = rest[:, None] == rest[None, :]
    return np.where(same_rest, sub[sub_index[:, None], sub_index[None, :]], 0).astype(np.complex128)


_PRELUDE_GATES: dict[str, _GateDef] | None = None

# --------------------------------------------------------------------------------------- API


def parse(source: str, params: Mapping[str, float] | None = None, *, include_dir: str | Path | None = None) -> Circuit:
    """Read OpenQASM 2 or 3 source into a :class:`Circuit`. ``params`` gives the values of the
    program's ``input`` parameters; ``include_dir`` permits includes from inside that folder."""
    include: Callable[[str], str] | None = None
    if include_dir is not None:
        root = Path(include_dir).resolve()

        def include(name: str) -> str:
            path = (root / name).resolve()
            if root not in path.parents:
                raise QasmError(f"include {name!r} is outside {root}")
            return path.read_text(encoding="utf-8")

    reader = _Reader(source, params or {}, include)
    reader.define_prelude()
    ops = reader.program()
    unused = set(params or {}) - set(reader.used_inputs)
    if unused:
        raise QasmError(f"the program has no input named {', '.join(sorted(unused))}")
    qregs = tuple(reader.qregs.values())
    if reader.physical >= 0:
        qregs = (Register("$", 0, reader.num_qubits),)
    return Circuit(
        num_qubits=reader.num_qubits,
        num_clbits=reader.num_clbits,
        operations=tuple(ops),
        qregs=qregs,
        cregs=tuple(reader.cregs.values()),
        inputs=dict(reader.used_inputs),
    )


def load(path: str | Path, params: Mapping[str, float] | None = None) -> Circuit:
    """Read an OpenQASM file; its includes resolve beside it."""
    p = Path(path)
    return parse(p.read_text(encoding="utf-8"), params, include_dir=p.parent)
