"""Write tests/golden/reader_cases.json: what lucivant-sim 0.3.0's OpenQASM reader makes of a
corpus of programs, valid and invalid.

    python tests/golden/make_reader_cases.py [--reference /root/reference/src]

The reference is imported from its checkout in this container only, to record data: for each
valid program its circuit digest (which covers every operation, qubit, parameter and matrix to 12
decimals), sizes, registers and operation labels; for each invalid one the error's message, line
and column. The library's reader is checked against these records (tests/test_reader.py).
"""

from __future__ import annotations

import argparse
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent

H2 = 'OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[4];\nqreg r[2];\ncreg c[4];\n'
H3 = 'OPENQASM 3;\ninclude "stdgates.inc";\nqubit[4] q;\nbit[4] c;\n'

VALID = {
    "exprs": H3 + "rx(-pi/2 + 3*0.25 - 2**-1) q[0];\nry(2^3^0.5) q[1];\nrz(-2**2) q[2];\nu(sin(0.3), cos(pi)*exp(1), ln(2)/sqrt(3)) q[3];\nrx(7 % 3 - -1) q[0];\nrx(tau/4 + euler) q[1];\nrx((1+2)*(3-4)/5) q[2];\n",
    "numbers": H3 + "rx(1_000.5) q[0];\nrx(0x1F) q[1];\nrx(0b101) q[2];\nrx(0o17) q[3];\nrx(.5e1) q[0];\nrx(3.) q[1];\nrx(1E-2) q[2];\nrx(+4) q[3];\n",
    "functions": H3 + "rx(arcsin(0.5)) q[0]; rx(asin(0.5)) q[1]; rx(arccos(0.1)) q[2]; rx(acos(0.1)) q[3];\nrx(arctan(2)) q[0]; rx(atan(2)) q[1]; rx(tan(0.4)) q[2]; rx(log(5)) q[3];\n",
    "modifiers": H3 + "ctrl @ h q[0], q[1];\nctrl(2) @ x q[0], q[1], q[2];\nnegctrl @ rx(0.3) q[3], q[0];\ninv @ t q[1];\npow(0.5) @ x q[2];\npow(3) @ s q[3];\ninv @ ctrl @ ry(0.2) q[1], q[2];\nctrl @ negctrl @ x q[0], q[1], q[2];\n",
    "user_gates": H3 + "gate g(a, b) x, y { rx(a) x; cx x, y; rz(b / 2) y; gphase(a); }\ngate h2 p { h p; barrier p; h p; }\ngate outer(t) u, v, w { g(t, 2*t) u, w; ctrl @ h2 v, u; }\nouter(0.7) q[0], q[1], q[2];\ng(0.1, 0.2) q[3], q[0];\nh2 q[1];\ninv @ g(0.3, 0.4) q[2], q[3];\n",
    "gate_no_params": H3 + "gate bell a, b { h a; cx a, b; }\ngate e() a { x a; }\nbell q[0], q[1];\ne q[2];\ne() q[3];\n",
    "shadow_builtin": H2 + "gate h a { U(0.1,0,0) a; }\nh q[0];\ncx q[1],q[2];\n",
    "broadcast2": H2 + "h q;\ncx q[0], r;\ncx r, q[1];\nmeasure q -> c;\nreset r;\nbarrier q, r[0];\nbarrier;\n",
    "slices": H3 + "x q[0:2:3];\nx q[3:-1:0];\nh q[-2:-1];\nz q[{3, 1}];\ny q[1:1];\n",
    "measure_forms": H3 + "bit b;\nbit[2] d;\nmeasure q[0] -> c[0];\nc[1] = measure q[1];\nd = measure q[2:3];\nb = measure q[0];\nmeasure q[3];\nmeasure q;\nbit[4] e = measure q;\n",
    "if_forms": H3 + "c[0] = measure q[0];\nif (c[0]) x q[1];\nif (!c[0]) x q[2];\nif (c == 5) { h q[0]; h q[1]; } else { x q[3]; }\nif (c != 0) z q[0];\nif (c[1] == true) y q[1];\nif (c[2] == false) y q[2];\nif (!c[3] == 1) s q[3];\nif (c[0] == 1) if (c[1] == 0) x q[0]; else x q[1];\n",
    "box_delay": H3 + "box { h q[0]; x q[1]; }\nbox[100ns] h q[2];\ndelay[50ns] q[0];\ndelay[dt] q[1], q[2];\nx q[3];\n",
    "classical": 'OPENQASM 3;\ninclude "stdgates.inc";\ninput float theta;\ninput angle[32] phi;\nconst float k = theta * 2;\nfloat[64] m = k + phi;\nangle a = pi/3;\nint n = 3;\nuint u = 2;\nbool f = 1;\nqubit[2] q;\nrx(theta) q[0];\nry(phi) q[1];\nrz(k + m + a + n + u + f) q[0];\n',
    "comments": H3 + "// a comment\nh q[0]; /* a\nblock\ncomment */ cx q[0], q[1];\n  #pragma keep this\n@annotation here\nswap q[2], q[3];\npragma also\n",
    "physical3": "OPENQASM 3;\ninclude \"stdgates.inc\";\nbit[2] c;\nh $0;\ncx $0, $3;\nc[0] = measure $3;\nc[1] = measure $1;\n",
    "no_header": 'include "stdgates.inc";\nqubit[2] q;\nh q[0];\ncx q[0], q[1];\n',
    "qasm2_gates": H2 + "u1(0.1) q[0];\nu2(0.2, 0.3) q[1];\nu3(0.4, 0.5, 0.6) q[2];\nU(0.7, 0.8, 0.9) q[3];\nCX q[0], q[1];\ncu1(0.3) q[0], q[1];\ncu3(0.1, 0.2, 0.3) q[1], q[2];\ncrx(0.4) q[2], q[3];\ncry(0.5) q[3], q[0];\ncrz(0.6) q[0], q[2];\nch q[1], q[3];\ncy q[0], q[3];\ncz q[1], q[2];\nid q[0];\nsdg q[1];\ntdg q[2];\nsxdg q[3];\nrxx(0.2) q[0], q[1];\nryy(0.3) q[1], q[2];\nrzz(0.4) q[2], q[3];\nc3sqrtx q[0], q[1], q[2], q[3];\np(0.5) q[0];\ncp(0.6) q[0], q[1];\ncswap q[0], q[1], q[2];\n",
    "stdgates3": H3 + "phase(0.3) q[0];\ncphase(0.2) q[0], q[1];\nu(0.1, 0.2, 0.3) q[2];\ncu(0.1, 0.2, 0.3, 0.4) q[2], q[3];\nsx q[0];\nswap q[1], q[2];\nccx q[0], q[1], q[3];\n",
    "gphase_top": H3 + "gphase(pi/4);\nh q[0];\ngphase(-1.5);\n",
    "semicolons": H3 + ";;\nh q[0];;\n;x q[1];\n",
    "include_twice": 'OPENQASM 2.0;\ninclude "qelib1.inc";\ninclude "stdgates.inc";\nqreg q[1];\nx q[0];\n',
    "versions": "OPENQASM 3.0;\nqubit q;\nU(1, 2, 3) q;\n",
    "qasm2_if": H2 + "measure q[0] -> c[0];\nif(c==1) x q[1];\nif (c==3) cx q[0], q[1];\n",
    "reset_static": H3 + "reset q[1];\nh q[0];\n",
    "unicode_consts": H3 + "rx(π/2) q[0];\nrx(τ/8) q[1];\nrx(ℇ) q[2];\n",
}

ERRORS = {
    "unknown_gate": "OPENQASM 3; qubit q; foo q;",
    "same_qubit": "OPENQASM 3;\nqubit[2] q;\ncx q[0], q[0];",
    "index_out": "OPENQASM 3; qubit[2] q; h q[2];",
    "negative_out": "OPENQASM 3; qubit[2] q; h q[-3];",
    "for_loop": "OPENQASM 3; qubit q; for uint i in [0:3] { h q; }",
    "while_loop": "OPENQASM 3; qubit q;\n  while (true) { h q; }",
    "def_sub": "OPENQASM 3; def f() { }",
    "missing_param": "OPENQASM 3; qubit q; rx q;",
    "too_many_params": "OPENQASM 3; qubit q; h(0.1) q;",
    "sizes_differ": "OPENQASM 3; qubit[2] q; qubit[3] r; cx q, r;",
    "include_other": 'OPENQASM 2.0; include "mine.inc"; qreg q[1];',
    "include_no_quotes": "OPENQASM 2.0; include qelib1;",
    "version4": "OPENQASM 4; qubit q;",
    "version_word": "OPENQASM three; qubit q;",
    "missing_semicolon": "OPENQASM 3; qubit q; h q",
    "missing_semicolon_mid": "OPENQASM 3; qubit q;\nh q\nx q;",
    "unknown_qreg": H2 + "h s[0];\n",
    "unknown_creg": H2 + "measure q[0] -> d[0];\n",
    "arity_few": H2 + "cx q[1];\n",
    "arity_many": H2 + "x q[1], q[2];\n",
    "later_line": H2 + "h q[0];\n  nope q[1] ,\n  q[2];\n",
    "unexpected_char": "OPENQASM 3; qubit q; h q; ?",
    "unexpected_char2": "OPENQASM 3;\nqubit q;\n  x q; $",
    "redeclare": "OPENQASM 3; qubit q; bit q;",
    "redeclare_value": "OPENQASM 3; const float a = 1; qubit[2] a;",
    "empty_register": "OPENQASM 3; qubit[0] q;",
    "fractional_size": "OPENQASM 3; qubit[1.5] q;",
    "mix_physical": "OPENQASM 3; qubit q; x $1;",
    "mix_physical2": "OPENQASM 3; x $1; qubit q;",
    "too_many_qubits": "OPENQASM 3; qubit[3000] q;",
    "too_many_physical": "OPENQASM 3; x $5000;",
    "unknown_name": "OPENQASM 3; qubit q; rx(alpha) q;",
    "div_zero": "OPENQASM 3; qubit q; rx(1/0) q;",
    "bad_function": "OPENQASM 3; qubit q; rx(sqrt(-1)) q;",
    "log_zero": "OPENQASM 3; qubit q; rx(ln(0)) q;",
    "bad_atom": "OPENQASM 3; qubit q; rx(*) q;",
    "unclosed_paren": "OPENQASM 3; qubit q; rx((1+2) q;",
    "input_missing": "OPENQASM 3; input float theta; qubit q; rx(theta) q;",
    "input_type": "OPENQASM 3; input complex z; qubit q;",
    "const_redeclare": "OPENQASM 3; const float a = 1; const float a = 2;",
    "step_zero": "OPENQASM 3; qubit[4] q; x q[0:0:3];",
    "bad_index": "OPENQASM 3; qubit[4] q; x q[0.5];",
    "assign_non_measure": "OPENQASM 3; qubit q; bit c; c = 1;",
    "bit_assign_other": "OPENQASM 3; qubit q; bit c = 3;",
    "measure_sizes": "OPENQASM 3; qubit[2] q; bit[3] c; c = measure q;",
    "measure_arrow_sizes": H2 + "measure q -> r;",
    "if_register_no_value": "OPENQASM 3; qubit q; bit[2] c; if (c) x q;",
    "if_compare_name": "OPENQASM 3; qubit q; bit[2] c; if (c == x) x q;",
    "if_unknown": "OPENQASM 3; qubit q; if (d == 1) x q;",
    "block_unclosed": "OPENQASM 3; qubit q; bit c; if (c) { x q;",
    "gate_unclosed": "OPENQASM 3; gate g a { x a;",
    "gate_redefined": "OPENQASM 3; gate g a { x a; } gate g a { y a; }",
    "gate_dup_args": "OPENQASM 3; gate g a, a { x a; }",
    "gate_unknown_arg": "OPENQASM 3; gate g a { cx a, b; }",
    "gate_recursive": "OPENQASM 3; gate g a { g a; }",
    "gate_bad_call": "OPENQASM 3; gate g a { cx a; } qubit q; g q;",
    "gate_same_arg": "OPENQASM 3; gate g a, b { cx a, a; } qubit[2] q; g q[0], q[1];",
    "gate_param_count": "OPENQASM 3; gate g(t) a { rx(t) a; } qubit q; g q;",
    "gate_unknown_param": "OPENQASM 3; gate g(t) a { rx(s) a; } qubit q; g(1) q;",
    "gate_gphase_args": "OPENQASM 3; gate g a { gphase(1, 2); } qubit q; g q;",
    "gphase_args": "OPENQASM 3; qubit q; gphase(1, 2);",
    "pow_no_arg": "OPENQASM 3; qubit q; pow @ x q;",
    "ctrl_fraction": "OPENQASM 3; qubit[3] q; ctrl(1.5) @ x q[0], q[1];",
    "ctrl_arity": "OPENQASM 3; qubit[3] q; ctrl @ x q[0];",
    "unexpected_token": "OPENQASM 3; qubit q; ) x q;",
    "unexpected_number": "OPENQASM 3; qubit q; 3 x q;",
    "expected_qubit": "OPENQASM 3; qubit q; x 3;",
    "not_an_argument": "OPENQASM 3; qubit q; x q[;",
    "gphase_no_paren": "OPENQASM 3; qubit q; gphase pi/2;",
    "overflow": "OPENQASM 3; qubit q; rx(10.0 ** 400) q;",
    "unknown_input_param": "OPENQASM 3; qubit q; h q;",
    "extern": "OPENQASM 3; extern f(int) -> int;",
    "opaque": "OPENQASM 2.0; opaque g a;",
    "eof_in_block": "OPENQASM 3; qubit q; box { h q;",
    "bad_skip": "OPENQASM 3; qubit q; box [ h q;",
    "string_operand": 'OPENQASM 3; qubit q; x "q";',
    "bad_bit_name": "OPENQASM 3; qubit q; measure q -> 3;",
    "classical_target_index": "OPENQASM 3; qubit[2] q; bit[2] c; c[2] = measure q[0];",
    "function_arg_unknown": "OPENQASM 3; qubit q; rx(sqrt(-q1)) q;",
    "function_arg_div_zero": "OPENQASM 3; qubit q; rx(sin(1/0)) q;",
    "function_arg_fmod": "OPENQASM 3; qubit q; rx(cos(5 % 0)) q;",
    "body_div_zero": "OPENQASM 3; gate g(t) a { rx(t/0) a; } qubit q; g(1) q;\nh q;",
    "body_ctrl_param": "OPENQASM 3; gate g(n) a, b { ctrl(n) @ x a, b; } qubit[2] q; g(1) q[0], q[1];",
    "power_complex": "OPENQASM 3; qubit q; rx((-8)**(1/3)) q;",
    "zero_negative_power": "OPENQASM 3; qubit q; rx(0.0 ** -1) q;",
    "modifier_eof": "OPENQASM 3; qubit q; ctrl(",
    "if_eof": "OPENQASM 3; qubit q; bit c; if (c",
}

PARAMS = {"classical": {"theta": 0.3, "phi": 1.1}, "unknown_input_param": {"nope": 1.0}}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference/src")
    args = ap.parse_args()
    sys.path.insert(0, args.reference)
    import lucivant_sim as ref  # the reference implementation

    valid = {}
    for name, src in VALID.items():
        c = ref.parse(src, PARAMS.get(name))
        valid[name] = {
            "source": src,
            "params": PARAMS.get(name),
            "digest": c.digest(),
            "num_qubits": c.num_qubits,
            "num_clbits": c.num_clbits,
            "qregs": [[r.name, r.offset, r.size] for r in c.qregs],
            "cregs": [[r.name, r.offset, r.size] for r in c.cregs],
            "inputs": dict(c.inputs),
            "ops": [repr(op.__class__.__name__) + " " + str(getattr(op, "name", "")) for op in c.operations],
        }
    errors = {}
    for name, src in ERRORS.items():
        try:
            ref.parse(src, PARAMS.get(name))
        except ref.QasmError as exc:
            errors[name] = {"source": src, "params": PARAMS.get(name), "message": str(exc), "line": exc.line, "column": exc.column}
        else:
            raise SystemExit(f"{name} was read without an error")
    out = {"reference_version": ref.__version__, "valid": valid, "errors": errors}
    (HERE / "reader_cases.json").write_text(json.dumps(out, indent=1, sort_keys=True, ensure_ascii=False))
    print(f"wrote {len(valid)} programs and {len(errors)} errors")


if __name__ == "__main__":
    main()
