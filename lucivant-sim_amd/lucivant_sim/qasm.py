"""OpenQASM 2.0 and 3 reader: the dialect lucivant-sim 0.3.0 accepts (src/lucivant_sim/qasm.py).

Accepted: ``qreg``/``creg`` and ``qubit``/``bit`` declarations, physical qubits ``$n``, the
built-in gate table (qelib1.inc and stdgates.inc, which are never read from disk), user ``gate``
blocks, the ``ctrl @``/``negctrl @``/``inv @``/``pow(k) @`` modifiers, ``input`` parameters and
classical constants, broadcasting over registers, index lists and ranges, ``measure`` in both
spellings, ``reset``, ``barrier``, ``gphase``, ``box``, ``delay`` and ``if``/``else`` on classical
bits. Loops, subroutines, calibrations and classical arithmetic are refused with a message.
Results, messages and error positions are those of the reference reader
(tests/golden/reader_cases.json records them).

How this reader is built:

- *Source.* Pragma and annotation lines are blanked, then the whole text is checked for lexical
  errors by two regular-expression passes (comments and strings masked, then a search for a
  character no token may start with), so a bad character is reported before any later error.
  Tokens carry their source offset; lines and columns are computed only when an error is raised.
- *Stream.* Tokens are lexed lazily, as the parser looks ahead.
- *Plain statements.* At a statement boundary one regular expression tries the most common form,
  a built-in gate with literal parameters on indexed qubits (``u3(0.1, 2, -3) q[0];``). When it
  matches and every check passes, the gate is made without tokenising the statement; otherwise
  the statement is read by the general path, which then reports any error.
- *Expressions* are parsed by precedence climbing into closures that evaluate under a name
  environment (gate parameters, then declared values, then the constants).
- *Statements* are dispatched on their first word through a table of handlers.
"""

from __future__ import annotations

import math
import re
from collections.abc import Callable, Mapping
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from lucivant_sim.circuit import Barrier, Circuit, Condition, Gate, GlobalPhase, IfElse, Measure, Operation, Register, Reset
from lucivant_sim.gates import GATES, Matrix, canonical_name, controlled, gate_matrix, matrix_power, negcontrolled

#: Declared qubits beyond this are refused before anything is allocated (the reference's limit).
MAX_QUBITS = 2048

#: Read plain gate statements with one regular expression (off: every statement is tokenised).
PLAIN_STATEMENTS = True


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


# ----------------------------------------------------------------------------------- source

_PRAGMA = re.compile(r"(?m)^[ \t]*(#?pragma|@)[^\n]*")
_MASK = re.compile(r"//[^\n]*|/\*.*?\*/|\"[^\"\n]*\"|'[^'\n]*'", re.DOTALL)
# characters no token starts with, a '$' not followed by a digit, and quotes left unmatched
_BAD = re.compile(r"[^ \t\n\r\f\v\w/.$\[\](){};,:=+\-*^%<>!~@&|]|\$(?!\d)|[\"']")


class _Source:
    """Program text (pragmas blanked) and line/column lookup for offsets."""

    __slots__ = ("text",)

    def __init__(self, text: str) -> None:
        self.text = _PRAGMA.sub("", text)

    def where(self, offset: int) -> tuple[int, int]:
        line = self.text.count("\n", 0, offset) + 1
        return line, offset - (self.text.rfind("\n", 0, offset) + 1) + 1

    def first_bad(self) -> int:
        """Offset of the first character the lexer can't read, or -1."""
        text = self.text
        masked = _MASK.sub(lambda m: " " * (m.end() - m.start()), text) if ('"' in text or "'" in text or "/" in text) else text
        m = _BAD.search(masked)
        return -1 if m is None else m.start()


# Token: (kind, text, offset, source); kinds id, num, str, phys, op, eof
Token = tuple


_SKIP = re.compile(r"(?:[ \t\r\f\v\n]+|//[^\n]*|/\*.*?\*/)*", re.DOTALL)
_LEX = re.compile(
    r"(0[bB][01_]+|0[xX][0-9a-fA-F_]+|0[oO][0-7_]+|(?:\d[\d_]*\.?[\d_]*|\.\d[\d_]*)(?:[eE][+-]?\d+)?)"
    r"|(\"[^\"\n]*\"|'[^'\n]*')"
    r"|(\$\d+)"
    r"|([^\W\d]\w*)"
    r"|(->|==|!=|<=|>=|&&|\|\||\*\*|<<|>>|[\[\](){};,:=+\-*/^%<>!~@&|.])"
)
_KINDS = (None, "num", "str", "phys", "id", "op")


class _Stream:
    """Tokens of one source, lexed on demand."""

    __slots__ = ("src", "text", "pos", "ahead")

    def __init__(self, src: _Source) -> None:
        self.src = src
        self.text = src.text
        self.pos = 0  # offset of the first character not yet lexed
        self.ahead: list[Token] = []  # lexed, not yet consumed

    def _lex(self) -> Token:
        text = self.text
        pos = _SKIP.match(text, self.pos).end()
        if pos >= len(text):
            self.pos = pos
            return ("eof", "", len(text), self.src)
        m = _LEX.match(text, pos)
        if m is None:  # excluded by the check of the whole source; kept as a guard
            raise _error_at(self.src, pos, f"unexpected character {text[pos]!r}")
        self.pos = m.end()
        return (_KINDS[m.lastindex], m.group(), pos, self.src)

    def peek(self, k: int = 0) -> Token:
        ahead = self.ahead
        while len(ahead) <= k:
            if ahead and ahead[-1][0] == "eof":
                return ahead[-1]
            ahead.append(self._lex())
        return ahead[k]

    def take(self) -> Token:
        tok = self.peek()
        if tok[0] != "eof":
            self.ahead.pop(0)
        return tok

    def at_boundary(self) -> bool:
        return not self.ahead


def _error_at(src: _Source, offset: int, message: str) -> QasmError:
    line, col = src.where(offset)
    return QasmError(message, line, col)


def _tok_error(tok: Token, message: str) -> QasmError:
    return _error_at(tok[3], tok[2], message)


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

Env = Mapping[str, float]
Expr = Callable[[Env, "_Reader"], float]


class _Unevaluable(Exception):
    """An arithmetic error; reported at the token the reader is on when it evaluates."""


def _literal(text: str) -> float:
    t = text.replace("_", "").lower()
    if t[:2] in ("0b", "0x", "0o"):
        return float(int(t[2:], {"0b": 2, "0x": 16, "0o": 8}[t[:2]]))
    return float(t)


def _arith(op: str, a: float, b: float) -> float:
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
        raise _Unevaluable(f"can't evaluate the expression: {exc}") from exc


def _const(v: float) -> Expr:
    return lambda env, rd: v


def _name(tok: Token) -> Expr:
    name = tok[1]

    def look(env: Env, rd: _Reader) -> float:
        if env and name in env:
            return env[name]
        if name in rd.values:
            return rd.values[name]
        if name in _CONSTANTS:
            return _CONSTANTS[name]
        raise _tok_error(tok, f"unknown name {name!r}")

    return look


def _call(tok: Token, arg: Expr) -> Expr:
    fn = _FUNCTIONS[tok[1]]

    def apply(env: Env, rd: _Reader) -> float:
        # an error anywhere in the argument is reported as the function's, at its name, as the
        # reference does (its argument is evaluated inside the handler, qasm.py:319-323)
        try:
            return fn(arg(env, rd))
        except _Unevaluable as exc:
            raise _tok_error(tok, f"{tok[1]}: {rd.err(str(exc))}") from exc
        except ValueError as exc:
            raise _tok_error(tok, f"{tok[1]}: {exc}") from exc

    return apply


def _binary(op: str, left: Expr, right: Expr) -> Expr:
    def apply(env: Env, rd: _Reader) -> float:
        a = left(env, rd)
        return _arith(op, a, right(env, rd))

    return apply


def _negate(inner: Expr) -> Expr:
    return lambda env, rd: -inner(env, rd)


# binding powers: sums < products < signs < powers (right-associative, may take a sign)
_INFIX = {"+": 10, "-": 10, "*": 20, "/": 20, "%": 20, "^": 30, "**": 30}
_SIGN_BP = 25
_POWER_RIGHT_BP = 24


# ------------------------------------------------------------------------------ gate blocks


@dataclass(frozen=True)
class _BodyCall:
    modifiers: tuple[tuple[str, Expr | None], ...]
    token: Token
    params: tuple[Expr, ...]
    args: tuple[str, ...]


@dataclass(frozen=True)
class _UserGate:
    name: str
    params: tuple[str, ...]
    qubits: tuple[str, ...]
    body: tuple[_BodyCall | Expr, ...]  # an Expr entry is a gphase angle


def _embed_after(op: Matrix, sub: Matrix, positions: tuple[int, ...], k: int) -> Matrix:
    """``sub`` applied after ``op`` on local qubits ``positions`` of a k-qubit block (local qubit
    j is matrix bit k-1-j; the first position is ``sub``'s most significant bit): the rows of
    ``op`` are regrouped so each column of ``sub`` meets the rows it mixes."""
    m = len(positions)
    rest = [j for j in range(k) if j not in positions]
    order = list(positions) + rest
    t = op.reshape([2] * k + [1 << k]).transpose(order + [k])
    t = (sub @ t.reshape(1 << m, -1)).reshape([2] * k + [1 << k])
    back = np.argsort(order).tolist()
    return np.ascontiguousarray(t.transpose(back + [k]).reshape(1 << k, 1 << k))


# ----------------------------------------------------------------------- plain statements

_Q_NUM = r"[+-]?(?:[0-9]+(?:\.[0-9]*)?|\.[0-9]+)(?:[eE][+-]?[0-9]+)?"
_Q_OPER = r"[A-Za-z_]\w*[ \t]*\[[ \t]*[0-9]+[ \t]*\]"
_QUICK = re.compile(
    rf"([A-Za-z_]\w*)(?:[ \t]*\([ \t]*({_Q_NUM}(?:[ \t]*,[ \t]*{_Q_NUM})*)[ \t]*\)[ \t]*|[ \t]+)"
    rf"({_Q_OPER}(?:[ \t]*,[ \t]*{_Q_OPER})*)[ \t]*;"
)
_Q_OPERAND = re.compile(r"([A-Za-z_]\w*)[ \t]*\[[ \t]*([0-9]+)")
#: decimal places of the parameters that key the matrix cache (one per parameter)
_DIGITS = (14,) * 16

_REFUSED = ("for", "while", "def", "defcal", "cal", "extern", "opaque", "return", "break", "continue")
_CLASSICAL = ("input", "const", "float", "angle", "int", "uint", "bool")
_STANDARD_INCLUDES = ("qelib1.inc", "stdgates.inc")

# qelib1.inc's relative-phase Toffolis, defined by decomposition rather than as matrices
_QELIB_BLOCKS = """
gate rccx a,b,c { u2(0,pi) c; u1(pi/4) c; cx b,c; u1(-pi/4) c; cx a,c; u1(pi/4) c; cx b,c;
  u1(-pi/4) c; u2(0,pi) c; }
gate rc3x a,b,c,d { u2(0,pi) d; u1(pi/4) d; cx c,d; u1(-pi/4) d; u2(0,pi) d; cx a,d; u1(pi/4) d;
  cx b,d; u1(-pi/4) d; cx a,d; u1(pi/4) d; cx b,d; u1(-pi/4) d; u2(0,pi) d; u1(pi/4) d; cx c,d;
  u1(-pi/4) d; u2(0,pi) d; }
"""
_QELIB_GATES: dict[str, _UserGate] | None = None


# ------------------------------------------------------------------------------------ reader

Operand = tuple[list[int], bool]  # qubits (or clbits), and whether a whole register was named


class _Reader:
    """One program's state: registers, declared values, gate definitions, and the stream being
    read (an ``include`` reads another stream into the same state)."""

    def __init__(self, inputs: Mapping[str, float], include: Callable[[str], str] | None) -> None:
        self.inputs = dict(inputs)
        self.used_inputs: dict[str, float] = {}
        self.include = include
        self.qregs: dict[str, Register] = {}
        self.cregs: dict[str, Register] = {}
        self.values: dict[str, float] = {}
        self.num_qubits = 0
        self.num_clbits = 0
        self.physical = -1  # highest $n named
        self.gates: dict[str, _UserGate] = dict(_qelib_gates())
        self.user_defined: set[str] = set()
        self.matrices: dict[tuple, Matrix] = {}
        self.s: _Stream | None = None

    # -- token helpers --------------------------------------------------------------------

    def err(self, message: str, tok: Token | None = None) -> QasmError:
        return _tok_error(tok if tok is not None else self.s.peek(), message)  # type: ignore[union-attr]

    def accept(self, text: str) -> bool:
        tok = self.s.peek()  # type: ignore[union-attr]
        if tok[1] == text and tok[0] in ("op", "id"):
            self.s.take()  # type: ignore[union-attr]
            return True
        return False

    def expect(self, text: str) -> Token:
        tok = self.s.peek()  # type: ignore[union-attr]
        if tok[1] != text or tok[0] not in ("op", "id"):
            found = "the end of the program" if tok[0] == "eof" else repr(tok[1])
            raise self.err(f"expected {text!r} but found {found}", tok)
        return self.s.take()  # type: ignore[union-attr]

    def word(self) -> Token:
        tok = self.s.take()  # type: ignore[union-attr]
        if tok[0] != "id":
            raise self.err(f"expected a name but found {tok[1]!r}", tok)
        return tok

    def is_op(self, *texts: str) -> bool:
        tok = self.s.peek()  # type: ignore[union-attr]
        return tok[0] == "op" and tok[1] in texts

    # -- expressions ----------------------------------------------------------------------

    def expr(self, rbp: int = 0) -> Expr:
        s = self.s
        assert s is not None
        tok = s.take()
        kind, text = tok[0], tok[1]
        if kind == "op" and text in ("-", "+"):
            inner = self.expr(_SIGN_BP)
            left = _negate(inner) if text == "-" else inner
        elif kind == "num":
            left = _const(_literal(text))
        elif kind == "op" and text == "(":
            left = self.expr()
            self.expect(")")
        elif kind == "id":
            if text in _FUNCTIONS and s.peek()[1] == "(":
                self.expect("(")
                arg = self.expr()
                self.expect(")")
                left = _call(tok, arg)
            else:
                left = _name(tok)
        else:
            raise self.err(f"expected a number or an expression but found {text!r}", tok)
        while True:
            nxt = s.peek()
            bp = _INFIX.get(nxt[1], 0) if nxt[0] == "op" else 0
            if bp <= rbp:
                return left
            s.take()
            if bp == 30:  # the exponent is a signed operand and may itself be a power
                left = _binary("**", left, self.expr(_POWER_RIGHT_BP))
            else:
                left = _binary(nxt[1], left, self.expr(bp))

    def value(self, e: Expr, env: Env | None = None) -> float:
        try:
            return e(env or {}, self)
        except _Unevaluable as exc:
            raise self.err(str(exc)) from None

    def whole(self, e: Expr, what: str) -> int:
        v = self.value(e)
        if not math.isfinite(v) or abs(v - round(v)) > 1e-9:
            raise self.err(f"{what} must be a whole number")
        return round(v)

    # -- program --------------------------------------------------------------------------

    def read(self, src: _Source) -> list[Operation]:
        bad = src.first_bad()
        if bad >= 0:
            raise _error_at(src, bad, f"unexpected character {src.text[bad]!r}")
        outer, self.s = self.s, _Stream(src)
        try:
            if self.s.peek()[1] == "OPENQASM":
                self.s.take()
                tok = self.s.take()
                if tok[0] != "num" or tok[1].split(".")[0] not in ("2", "3"):
                    raise self.err("only OpenQASM 2 and 3 are supported", tok)
                self.expect(";")
            ops: list[Operation] = []
            stream = self.s
            while True:
                if stream.at_boundary() and self.plain_gates(ops):
                    continue
                if stream.peek()[0] == "eof":
                    return ops
                ops += self.statement(plain_tried=True)
        finally:
            self.s = outer

    def statement(self, plain_tried: bool = False) -> list[Operation]:
        s = self.s
        assert s is not None
        if not plain_tried and s.at_boundary():
            plain = self.plain_gate()
            if plain is not None:
                return plain
        tok = s.peek()
        kind, text = tok[0], tok[1]
        if kind == "id":
            handler = _HANDLERS.get(text)
            if handler is not None:
                return handler(self, tok)
            if s.peek(1)[1] in ("=", "[") and text in self.cregs:
                return self.assignment()
            return self.gate_statement()
        if kind == "op" and text == ";":
            s.take()
            return []
        raise self.err(f"unexpected {text!r}" if kind != "eof" else "unexpected end of program", tok)

    def block(self) -> list[Operation]:
        """A ``{ ... }`` block or a single statement."""
        if not self.accept("{"):
            return self.statement()
        ops: list[Operation] = []
        while not self.accept("}"):
            if self.s.peek()[0] == "eof":  # type: ignore[union-attr]
                raise self.err("this block is never closed with '}'")
            ops += self.statement()
        return ops

    def skip_past(self, closing: str) -> None:
        depth = 0
        while True:
            tok = self.s.take()  # type: ignore[union-attr]
            if tok[0] == "eof":
                raise self.err(f"expected {closing!r}")
            if tok[1] == "[":
                depth += 1
            elif tok[1] == closing:
                if depth == 0:
                    return
                depth -= 1

    # -- the plain form -------------------------------------------------------------------

    def plain_gate(self) -> list[Operation] | None:
        """A built-in gate with literal parameters on indexed qubits, read by one match; None
        when the statement has another form or fails a check (the general path then reads it
        and reports the error)."""
        s = self.s
        assert s is not None
        if not PLAIN_STATEMENTS:
            return None
        text = s.text
        start = _SKIP.match(text, s.pos).end()
        m = _QUICK.match(text, start)
        if m is None:
            return None
        name, args, operands = m.group(1, 2, 3)
        spec = GATES.get(name)
        if spec is None or name in _HANDLERS or name in self.gates:
            return None
        params = tuple(float(a) for a in args.split(",")) if args else ()
        if len(params) != spec.params:
            return None
        qubits = []
        for reg_name, idx in _Q_OPERAND.findall(operands):
            reg = self.qregs.get(reg_name)
            i = int(idx)
            if reg is None or i >= reg.size:
                return None
            qubits.append(reg.offset + i)
        if len(qubits) != spec.qubits or len(set(qubits)) != len(qubits):
            return None
        s.pos = m.end()
        return [Gate(canonical_name(name), tuple(qubits), self.builtin(name, params), params)]

    def plain_gates(self, ops: list[Operation]) -> bool:
        """:meth:`plain_gate` over a run of consecutive statements, appending to ``ops``; True if it
        read at least one. (One call per run keeps the per-statement work to the match itself.)"""
        s = self.s
        assert s is not None
        if not PLAIN_STATEMENTS:
            return False
        text, pos = s.text, s.pos
        skip, quick, operand = _SKIP.match, _QUICK.match, _Q_OPERAND.findall
        qregs, user, builtin, append = self.qregs, self.gates, self.builtin, ops.append
        took = False
        while True:
            m = quick(text, skip(text, pos).end())
            if m is None:
                break
            name, args, operands = m.group(1, 2, 3)
            spec = GATES.get(name)
            if spec is None or name in _HANDLERS or name in user:
                break
            params = tuple(map(float, args.split(","))) if args else ()
            if len(params) != spec.params:
                break
            qubits = []
            for reg_name, idx in operand(operands):
                reg = qregs.get(reg_name)
                i = int(idx)
                if reg is None or i >= reg.size:
                    qubits = None
                    break
                qubits.append(reg.offset + i)
            if qubits is None or len(qubits) != spec.qubits or len(set(qubits)) != len(qubits):
                break
            pos = m.end()
            append(Gate(canonical_name(name), tuple(qubits), builtin(name, params), params))
            took = True
        if took:
            s.pos = pos
        return took

    def builtin(self, name: str, params: tuple[float, ...]) -> Matrix:
        key = (name, tuple(map(round, params, _DIGITS)))
        mat = self.matrices.get(key)
        if mat is None:
            mat = self.matrices[key] = gate_matrix(name, params)
        return mat

    # -- declarations ---------------------------------------------------------------------

    def declare(self, name: Token, size: int, quantum: bool) -> None:
        n = name[1]
        if n in self.qregs or n in self.cregs or n in self.values:
            raise self.err(f"{n!r} is already declared", name)
        if size < 1:
            raise self.err("a register needs at least one bit", name)
        if quantum:
            if self.physical >= 0:
                raise self.err("a program can't mix declared qubits with physical qubits ($n)", name)
            if self.num_qubits + size > MAX_QUBITS:
                raise self.err(f"more than {MAX_QUBITS} qubits", name)
            self.qregs[n] = Register(n, self.num_qubits, size)
            self.num_qubits += size
        else:
            self.cregs[n] = Register(n, self.num_clbits, size)
            self.num_clbits += size

    def st_register2(self, tok: Token) -> list[Operation]:
        """``qreg q[n];`` / ``creg c[n];``"""
        self.s.take()  # type: ignore[union-attr]
        name = self.word()
        self.expect("[")
        size = self.whole(self.expr(), "a register size")
        self.expect("]")
        self.expect(";")
        self.declare(name, size, tok[1] == "qreg")
        return []

    def st_register3(self, tok: Token) -> list[Operation]:
        """``qubit[n] q;`` / ``bit[n] c;`` / ``bit[n] c = measure q;``"""
        self.s.take()  # type: ignore[union-attr]
        size = 1
        if self.accept("["):
            size = self.whole(self.expr(), "a register size")
            self.expect("]")
        name = self.word()
        self.declare(name, size, tok[1] == "qubit")
        if tok[1] == "bit" and self.accept("="):
            self.expect("measure")
            qubits = self.qubit_operand()
            self.expect(";")
            return self.measures(qubits, self.whole_creg(name[1]))
        self.expect(";")
        return []

    def st_classical(self, tok: Token) -> list[Operation]:
        """``input float x;``, ``const float k = ...;``, ``angle a = ...;`` and the like."""
        is_input = self.accept("input")
        self.accept("const")
        kind = self.word()
        if kind[1] not in ("float", "angle", "int", "uint", "bool"):
            raise self.err(f"'{kind[1]}' parameters aren't supported; use float or angle", kind)
        if self.accept("["):
            self.skip_past("]")
        name = self.word()
        if name[1] in self.values:
            raise self.err(f"{name[1]!r} is already declared", name)
        if is_input:
            if name[1] not in self.inputs:
                raise self.err(f"the program needs a value for its input {name[1]!r}", name)
            v = float(self.inputs[name[1]])
            self.used_inputs[name[1]] = v
        else:
            self.expect("=")
            v = self.value(self.expr())
        self.expect(";")
        self.values[name[1]] = v
        return []

    def st_include(self, tok: Token) -> list[Operation]:
        self.s.take()  # type: ignore[union-attr]
        name = self.s.take()  # type: ignore[union-attr]
        if name[0] != "str":
            raise self.err("include needs a file name in quotes", name)
        self.expect(";")
        file = name[1][1:-1]
        if file in _STANDARD_INCLUDES:
            return []
        if self.include is None:
            raise self.err(f"can't include {file!r}: only qelib1.inc and stdgates.inc are built in", name)
        return self.read(_Source(self.include(file)))

    # -- operands -------------------------------------------------------------------------

    def index_list(self, size: int, reg: Token) -> list[int]:
        """The part of ``reg[...]`` after the bracket: an index, a range or a ``{...}`` set."""
        if self.accept("{"):
            picked = [self.whole(self.expr(), "an index")]
            while self.accept(","):
                picked.append(self.whole(self.expr(), "an index"))
            self.expect("}")
        else:
            first = self.whole(self.expr(), "an index")
            if self.accept(":"):
                bound = self.whole(self.expr(), "an index")
                step, stop = 1, bound
                if self.accept(":"):
                    step, stop = bound, self.whole(self.expr(), "an index")
                if step == 0:
                    raise self.err("a range's step can't be 0", reg)
                picked = list(range(first, stop + (1 if step > 0 else -1), step))
            else:
                picked = [first]
        self.expect("]")
        out = []
        for i in picked:
            j = i + size if i < 0 else i
            if not 0 <= j < size:
                raise self.err(f"index {i} is outside {reg[1]!r}, which has {size}", reg)
            out.append(j)
        return out

    def qubit_operand(self) -> Operand:
        tok = self.s.take()  # type: ignore[union-attr]
        if tok[0] == "phys":
            if self.qregs:
                raise self.err("a program can't mix declared qubits with physical qubits ($n)", tok)
            n = int(tok[1][1:])
            if n >= MAX_QUBITS:
                raise self.err(f"more than {MAX_QUBITS} qubits", tok)
            self.physical = max(self.physical, n)
            self.num_qubits = self.physical + 1
            return [n], False
        if tok[0] != "id":
            raise self.err(f"expected a qubit but found {tok[1]!r}", tok)
        reg = self.qregs.get(tok[1])
        if reg is None:
            raise self.err(f"unknown qubit register {tok[1]!r}", tok)
        if self.accept("["):
            return [reg.offset + i for i in self.index_list(reg.size, tok)], False
        return list(range(reg.offset, reg.offset + reg.size)), True

    def creg_operand(self, tok: Token) -> Operand:
        reg = self.cregs.get(tok[1])
        if reg is None:
            raise self.err(f"unknown classical register {tok[1]!r}", tok)
        if self.accept("["):
            return [reg.offset + i for i in self.index_list(reg.size, tok)], False
        return list(range(reg.offset, reg.offset + reg.size)), True

    def whole_creg(self, name: str) -> Operand:
        reg = self.cregs.get(name)
        if reg is None:
            raise self.err(f"unknown classical register {name!r}")
        return list(range(reg.offset, reg.offset + reg.size)), True

    def measures(self, qubits: Operand, bits: Operand) -> list[Operation]:
        if len(qubits[0]) != len(bits[0]):
            raise self.err(f"can't measure {len(qubits[0])} qubits into {len(bits[0])} bits")
        return [Measure(q, c) for q, c in zip(qubits[0], bits[0])]

    # -- measurement, reset, barrier, timing ----------------------------------------------

    def st_measure(self, tok: Token) -> list[Operation]:
        self.s.take()  # type: ignore[union-attr]
        qubits = self.qubit_operand()
        if self.accept("->"):
            bits = self.creg_operand(self.word())
            self.expect(";")
            return self.measures(qubits, bits)
        self.expect(";")
        return [Measure(q, None) for q in qubits[0]]

    def assignment(self) -> list[Operation]:
        """``c = measure q;`` / ``c[0] = measure q[0];``"""
        bits = self.creg_operand(self.word())
        self.expect("=")
        if not self.accept("measure"):
            raise self.err("only measurement results can be assigned to bits")
        qubits = self.qubit_operand()
        self.expect(";")
        return self.measures(qubits, bits)

    def st_reset(self, tok: Token) -> list[Operation]:
        self.s.take()  # type: ignore[union-attr]
        qubits = self.qubit_operand()
        self.expect(";")
        return [Reset(q) for q in qubits[0]]

    def st_barrier(self, tok: Token) -> list[Operation]:
        s = self.s
        assert s is not None
        s.take()
        qubits: list[int] = []
        while s.peek()[1] != ";":
            qubits += self.qubit_operand()[0]
            if not self.accept(","):
                break
        self.expect(";")
        return [Barrier(tuple(qubits) or tuple(range(self.num_qubits)))]

    def st_box(self, tok: Token) -> list[Operation]:
        self.s.take()  # type: ignore[union-attr]
        if self.accept("["):
            self.skip_past("]")
        return self.block()

    def st_delay(self, tok: Token) -> list[Operation]:
        s = self.s
        assert s is not None
        s.take()
        self.expect("[")
        self.skip_past("]")
        while s.peek()[1] != ";" and s.peek()[0] != "eof":
            s.take()
        self.expect(";")
        return []

    def st_refused(self, tok: Token) -> list[Operation]:
        raise self.err(f"'{tok[1]}' isn't supported by Lucivant Orrery yet; unroll it or define a gate instead", tok)

    def st_gphase(self, tok: Token) -> list[Operation]:
        self.s.take()  # type: ignore[union-attr]
        self.expect("(")
        angle = self.value(self.expr())
        self.expect(")")
        self.expect(";")
        return [GlobalPhase(angle)]

    # -- conditions -----------------------------------------------------------------------

    def st_if(self, tok: Token) -> list[Operation]:
        s = self.s
        assert s is not None
        self.expect("if")
        self.expect("(")
        negate = self.accept("!")
        reg = self.word()
        bits, _ = self.creg_operand(reg)
        value = 1
        if s.peek()[1] in ("==", "!="):
            if s.take()[1] == "!=":
                negate = not negate
            rhs = s.take()
            if rhs[0] == "num":
                value = int(_literal(rhs[1]))
            elif rhs[1] in ("true", "false"):
                value = int(rhs[1] == "true")
            else:
                raise self.err("a condition compares with a whole number", rhs)
        elif len(bits) != 1:
            raise self.err("compare a register with a value, e.g. if (c == 1)", reg)
        self.expect(")")
        then = self.block()
        otherwise: list[Operation] = []
        if self.accept("else"):
            otherwise = self.block()
        return [IfElse(Condition(tuple(bits), value, negate), tuple(then), tuple(otherwise))]

    # -- gate applications ----------------------------------------------------------------

    def modifier_list(self) -> list[tuple[str, Expr | None]]:
        s = self.s
        assert s is not None
        mods: list[tuple[str, Expr | None]] = []
        while True:
            head = s.peek()
            if head[0] != "id" or head[1] not in ("ctrl", "negctrl", "inv", "pow"):
                return mods
            # a modifier, with its optional (argument), is followed by '@'
            k = 1
            if s.peek(1)[1] == "(":
                depth = 0
                while True:
                    t = s.peek(k)
                    if t[1] == "(":
                        depth += 1
                    elif t[1] == ")":
                        depth -= 1
                        if depth == 0:
                            break
                    elif t[0] == "eof":
                        break
                    k += 1
                k += 1
            if s.peek(k)[1] != "@":
                return mods
            s.take()
            arg = None
            if self.accept("("):
                arg = self.expr()
                self.expect(")")
            self.expect("@")
            mods.append((head[1], arg))

    def call_head(self) -> tuple[list[tuple[str, Expr | None]], Token, list[Expr]]:
        mods = self.modifier_list()
        name = self.word()
        params: list[Expr] = []
        if self.accept("(") and not self.accept(")"):
            params.append(self.expr())
            while self.accept(","):
                params.append(self.expr())
            self.expect(")")
        return mods, name, params

    def gate_statement(self) -> list[Operation]:
        s = self.s
        assert s is not None
        mods, name, exprs = self.call_head()
        params = tuple(self.value(e) for e in exprs)
        operands: list[Operand] = []
        while s.peek()[1] != ";":
            operands.append(self.qubit_operand())
            if not self.accept(","):
                break
        self.expect(";")
        if name[1] == "gphase" and not operands:
            if len(params) != 1:
                raise self.err("gphase takes one angle", name)
            return [GlobalPhase(params[0])]
        matrix, arity = self.resolve(mods, name, params)
        if len(operands) != arity:
            raise self.err(f"{name[1]} acts on {arity} qubit(s) but was given {len(operands)}", name)
        sizes = {len(qs) for qs, whole in operands if whole or len(qs) > 1}
        if len(sizes) > 1:
            raise self.err("registers of different sizes can't be combined in one gate", name)
        times = sizes.pop() if sizes else 1
        label = "".join(f"{m}@" for m, _ in mods) + canonical_name(name[1])
        out: list[Operation] = []
        for i in range(times):
            qubits = tuple(qs[i] if len(qs) > 1 else qs[0] for qs, _ in operands)
            if len(set(qubits)) != len(qubits):
                raise self.err(f"{name[1]} is given the same qubit twice", name)
            out.append(Gate(label, qubits, matrix, params))
        return out

    def resolve(
        self, mods: list[tuple[str, Expr | None]] | tuple, name: Token, params: tuple[float, ...], env: Env | None = None
    ) -> tuple[Matrix, int]:
        """The matrix and width of a gate with its modifiers, applied innermost first."""
        matrix, arity = self.base_gate(name, params)
        for word, arg in reversed(mods):
            if word == "inv":
                matrix = matrix.conj().T
            elif word == "pow":
                if arg is None:
                    raise self.err("pow needs an exponent, e.g. pow(2) @", name)
                matrix = matrix_power(matrix, self.value(arg, env))
            else:
                if arg is None:
                    count = 1
                else:
                    # evaluated among the declared values only, not the gate's parameters, as
                    # the reference does (qasm.py:742)
                    count = self.whole(arg, "the number of controls")
                grow = controlled if word == "ctrl" else negcontrolled
                for _ in range(count):
                    matrix = grow(matrix)
                arity += count
        return matrix, arity

    def base_gate(self, name: Token, params: tuple[float, ...]) -> tuple[Matrix, int]:
        n = name[1]
        user = self.gates.get(n)
        if user is not None and (n in self.user_defined or n not in GATES):
            if len(params) != len(user.params):
                raise self.err(f"{n} takes {len(user.params)} parameter(s) but was given {len(params)}", name)
            key = (n, tuple(round(p, 14) for p in params))
            mat = self.matrices.get(key)
            if mat is None:
                mat = self.matrices[key] = self.block_unitary(user, params)
            return mat, len(user.qubits)
        spec = GATES.get(n)
        if spec is None:
            raise self.err(f"unknown gate {n!r}", name)
        if len(params) != spec.params:
            raise self.err(f"{n} takes {spec.params} parameter(s) but was given {len(params)}", name)
        return self.builtin(n, params), spec.qubits

    def block_unitary(self, gate: _UserGate, params: tuple[float, ...]) -> Matrix:
        k = len(gate.qubits)
        env = dict(zip(gate.params, params))
        local = {q: i for i, q in enumerate(gate.qubits)}
        op = np.eye(1 << k, dtype=np.complex128)
        for item in gate.body:
            if not isinstance(item, _BodyCall):
                op = op * np.exp(1j * self.value(item, env))
                continue
            values = tuple(self.value(p, env) for p in item.params)
            sub, arity = self.resolve(item.modifiers, item.token, values, env)
            if len(item.args) != arity:
                raise self.err(f"{item.token[1]} acts on {arity} qubit(s) but was given {len(item.args)}", item.token)
            positions = tuple(local[a] for a in item.args)
            if len(set(positions)) != len(positions):
                raise self.err(f"{item.token[1]} is given the same qubit twice", item.token)
            op = _embed_after(op, sub, positions, k)
        return op

    def st_gate(self, tok: Token) -> list[Operation]:
        """A ``gate name(params) a, b { ... }`` definition."""
        s = self.s
        assert s is not None
        self.expect("gate")
        name = self.word()
        if name[1] in self.user_defined:
            raise self.err(f"gate {name[1]!r} is already defined", name)
        params: list[str] = []
        if self.accept("(") and not self.accept(")"):
            params.append(self.word()[1])
            while self.accept(","):
                params.append(self.word()[1])
            self.expect(")")
        qubits = [self.word()[1]]
        while self.accept(","):
            qubits.append(self.word()[1])
        if len(set(qubits)) != len(qubits):
            raise self.err("a gate's qubit arguments must have different names", name)
        self.expect("{")
        body: list[_BodyCall | Expr] = []
        while not self.accept("}"):
            head = s.peek()
            if head[0] == "eof":
                raise self.err(f"gate {name[1]!r} is never closed with '}}'", name)
            if head[1] == "barrier":
                while s.take()[1] != ";":
                    if s.peek()[0] == "eof":
                        raise self.err(f"gate {name[1]!r} is never closed with '}}'", name)
                continue
            mods, call, exprs = self.call_head()
            if call[1] == "gphase":
                self.expect(";")
                if len(exprs) != 1:
                    raise self.err("gphase takes one angle", call)
                body.append(exprs[0])
                continue
            args: list[str] = []
            while s.peek()[1] != ";":
                arg = self.word()
                if arg[1] not in qubits:
                    raise self.err(f"{arg[1]!r} isn't an argument of gate {name[1]!r}", arg)
                args.append(arg[1])
                if not self.accept(","):
                    break
            self.expect(";")
            if call[1] == name[1]:
                raise self.err("a gate can't call itself", call)
            body.append(_BodyCall(tuple(mods), call, tuple(exprs), tuple(args)))
        self.gates[name[1]] = _UserGate(name[1], tuple(params), tuple(qubits), tuple(body))
        self.user_defined.add(name[1])
        return []


_HANDLERS: dict[str, Callable[[_Reader, Token], list[Operation]]] = {
    "include": _Reader.st_include,
    "qreg": _Reader.st_register2,
    "creg": _Reader.st_register2,
    "qubit": _Reader.st_register3,
    "bit": _Reader.st_register3,
    "gate": _Reader.st_gate,
    "measure": _Reader.st_measure,
    "reset": _Reader.st_reset,
    "barrier": _Reader.st_barrier,
    "if": _Reader.st_if,
    "box": _Reader.st_box,
    "delay": _Reader.st_delay,
    "gphase": _Reader.st_gphase,
    **{w: _Reader.st_classical for w in _CLASSICAL},
    **{w: _Reader.st_refused for w in _REFUSED},
}


def _qelib_gates() -> dict[str, _UserGate]:
    """qelib1.inc's decomposed gates, read once per process. They are not user definitions:
    a program may define gates of the same names."""
    global _QELIB_GATES
    if _QELIB_GATES is None:
        _QELIB_GATES = {}
        rd = _Reader({}, None)
        rd.read(_Source(_QELIB_BLOCKS))
        _QELIB_GATES = dict(rd.gates)
    return _QELIB_GATES


# --------------------------------------------------------------------------------------- API


def parse(source: str, params: Mapping[str, float] | None = None, *, include_dir: str | Path | None = None) -> Circuit:
    """Read OpenQASM 2 or 3 source into a :class:`Circuit` (qasm.py:842-879). ``params`` gives
    the values of the program's ``input`` parameters; ``include_dir`` permits includes of files
    inside that folder (never outside it)."""
    include: Callable[[str], str] | None = None
    if include_dir is not None:
        root = Path(include_dir).resolve()

        def include(name: str) -> str:
            path = (root / name).resolve()
            if root not in path.parents:
                raise QasmError(f"include {name!r} is outside {root}")
            return path.read_text(encoding="utf-8")

    rd = _Reader(params or {}, include)
    ops = rd.read(_Source(source))
    unused = set(params or {}) - set(rd.used_inputs)
    if unused:
        raise QasmError(f"the program has no input named {', '.join(sorted(unused))}")
    qregs = tuple(rd.qregs.values())
    if rd.physical >= 0:
        qregs = (Register("$", 0, rd.num_qubits),)
    return Circuit(
        num_qubits=rd.num_qubits,
        num_clbits=rd.num_clbits,
        operations=tuple(ops),
        qregs=qregs,
        cregs=tuple(rd.cregs.values()),
        inputs=dict(rd.used_inputs),
    )


def load(path: str | Path, params: Mapping[str, float] | None = None) -> Circuit:
    """Read an OpenQASM file; includes resolve beside it (qasm.py:882-885)."""
    p = Path(path)
    return parse(p.read_text(encoding="utf-8"), params, include_dir=p.parent)


__all__ = ["MAX_QUBITS", "QasmError", "load", "parse"]

