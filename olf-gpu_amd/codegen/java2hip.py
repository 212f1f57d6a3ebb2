"""java2hip: renders the device fill's term functions as HIP from OLF's own equation Java expressions.

Input: a family's formula manifest, the value and hand-Jacobian expression bodies of OLF's term classes as the
reference's java2cuda parser extracts them (codegen/golden/<family>.golden.json; kept here as fixtures under
tests/golden/): per method its name, parameters and one Java expression over doubles, which may call sibling methods.
Output: olf-gpu_amd/csrc/generated/<family>_terms.hpp, one __host__ __device__ function per method whose body is the
Java expression itself (sibling calls inlined, the side-2 ratio R2 = 1 of the stable-row packing folded in as the
reference's own constant), plus family helpers:

* closed branch: all 28 outputs in the scatter order of the reference's stamp (p1, its derivatives by v1 v2 ph1 ph2 a1
  r1, then q1, p2, q2), with theta1 = ksi - a1 + ph2 - ph1 and theta2 = ksi + a1 + ph1 - ph2 (A2 = 0); and the
  own-bus / far-bus form the batched fill gathers, p = A vo^2 - B vo vf sin(theta), q = C vo^2 - B vo vf cos(theta),
  whose coefficients are read off the rendered p/q functions at unit points (the generator checks the form holds);
* shunt: p, q and their voltage derivatives.

The steps follow the reference's pipeline (codegen/java2cuda/README.md:87-105) in own code: a recursive-descent parser
of the expression subset, inlining of sibling calls, rendering. Its proofs are tests/test_java2hip.py: parse fidelity
(parsed == manifest), a finite-difference audit of the hand Jacobian, rendered == committed (drift), rendered ==
manifest (a host build of the header).

    python olf-gpu_amd/codegen/java2hip.py            # regenerate the headers
    python olf-gpu_amd/codegen/java2hip.py --check    # exit 1 when a committed header differs from its rendering
"""
from __future__ import annotations

import argparse
import json
import math
import os
import random
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
MANIFESTS = os.path.join(REPO, "tests", "golden")
OUT_DIR = os.path.join(os.path.dirname(HERE), "csrc", "generated")
CONSTANTS = {"R2": 1.0}    # the side-2 ratio of the stable-row packing (the reference's R2 constexpr)

# ---- parser: expr := term (('+'|'-') term)*; term := unary (('*'|'/') unary)*; unary := '-' unary | primary;
#      primary := number | name | name '(' args ')' | '(' expr ')'
_TOKEN = re.compile(r"\s*(?:(\d+\.\d*(?:[eE][-+]?\d+)?|\d+(?:[eE][-+]?\d+)?|\.\d+)|([A-Za-z_]\w*)|(.))")


def tokenize(text):
    out = []
    pos = 0
    text = text.strip()
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise SyntaxError(f"cannot read {text[pos:]!r}")
        num, name, op = m.groups()
        if num is not None:
            out.append(("num", num))
        elif name is not None:
            out.append(("name", name))
        elif op.strip():
            if op not in "+-*/(),":
                raise SyntaxError(f"operator {op!r} is outside the expression subset")
            out.append(("op", op))
        pos = m.end()
    return out


class Parser:
    def __init__(self, text):
        self.toks = tokenize(text)
        self.i = 0

    def peek(self):
        return self.toks[self.i] if self.i < len(self.toks) else (None, None)

    def take(self, kind=None, value=None):
        t = self.peek()
        if t[0] is None or (kind and t[0] != kind) or (value and t[1] != value):
            raise SyntaxError(f"expected {value or kind}, found {t[1]!r}")
        self.i += 1
        return t

    def parse(self):
        e = self.expr()
        if self.peek()[0] is not None:
            raise SyntaxError(f"trailing {self.peek()[1]!r}")
        return e

    def expr(self):
        e = self.term()
        while self.peek() in (("op", "+"), ("op", "-")):
            op = self.take()[1]
            e = ("bin", op, e, self.term())
        return e

    def term(self):
        e = self.unary()
        while self.peek() in (("op", "*"), ("op", "/")):
            op = self.take()[1]
            e = ("bin", op, e, self.unary())
        return e

    def unary(self):
        if self.peek() == ("op", "-"):
            self.take()
            return ("neg", self.unary())
        return self.primary()

    def primary(self):
        kind, val = self.peek()
        if kind == "num":
            self.take()
            return ("num", val)
        if kind == "name":
            self.take()
            if self.peek() == ("op", "("):
                self.take()
                args = []
                if self.peek() != ("op", ")"):
                    args.append(self.expr())
                    while self.peek() == ("op", ","):
                        self.take()
                        args.append(self.expr())
                self.take("op", ")")
                return ("call", val, args)
            return ("var", val)
        if (kind, val) == ("op", "("):
            self.take()
            e = self.expr()
            self.take("op", ")")
            return ("paren", e)
        raise SyntaxError(f"unexpected {val!r}")


def parse_manifest(path):
    """{method: (params, ast)} of a manifest, in its order."""
    with open(path) as f:
        man = json.load(f)
    return man["java_class"], {m["name"]: (list(m["params"]), Parser(m["body"]).parse()) for m in man["methods"]}


def inline(ast, methods, env=None):
    """The expression with sibling calls replaced by the callee's body (its parameters bound to the argument
    expressions, parenthesised) and named constants folded."""
    env = env or {}
    kind = ast[0]
    if kind == "var":
        if ast[1] in env:
            return ("paren", env[ast[1]])
        if ast[1] in CONSTANTS:
            return ("num", repr(CONSTANTS[ast[1]]))
        return ast
    if kind == "num":
        return ast
    if kind == "neg":
        return ("neg", inline(ast[1], methods, env))
    if kind == "paren":
        return ("paren", inline(ast[1], methods, env))
    if kind == "bin":
        return ("bin", ast[1], inline(ast[2], methods, env), inline(ast[3], methods, env))
    if kind == "call":
        name, args = ast[1], [inline(a, methods, env) for a in ast[2]]
        if name not in methods:
            raise NameError(f"call of {name}, not a method of the manifest")
        params, body = methods[name]
        if len(params) != len(args):
            raise TypeError(f"{name} takes {len(params)} arguments")
        return ("paren", inline(body, methods, dict(zip(params, args))))
    raise ValueError(kind)


def render(ast):
    """C text of an expression (its parentheses kept as written; doubles throughout)."""
    kind = ast[0]
    if kind == "num":
        v = ast[1]
        return v if ("." in v or "e" in v or "E" in v) else v + ".0"
    if kind == "var":
        return ast[1]
    if kind == "neg":
        return "-" + render(ast[1])
    if kind == "paren":
        return "(" + render(ast[1]) + ")"
    if kind == "bin":
        return render(ast[2]) + " " + ast[1] + " " + render(ast[3])
    if kind == "call":
        return ast[1] + "(" + ", ".join(render(a) for a in ast[2]) + ")"
    raise ValueError(kind)


def evaluate(ast, values):
    """The expression's value for a dict of parameter values (after inline)."""
    kind = ast[0]
    if kind == "num":
        return float(ast[1])
    if kind == "var":
        return values[ast[1]]
    if kind == "neg":
        return -evaluate(ast[1], values)
    if kind == "paren":
        return evaluate(ast[1], values)
    a, b = evaluate(ast[2], values), evaluate(ast[3], values)
    op = ast[1]
    if op == "+":
        return a + b
    if op == "-":
        return a - b
    return a * b if op == "*" else a / b


def free_vars(ast, acc=None):
    acc = set() if acc is None else acc
    if ast[0] == "var":
        acc.add(ast[1])
    for a in ast[1:]:
        if isinstance(a, tuple):
            free_vars(a, acc)
        elif isinstance(a, list):
            for x in a:
                free_vars(x, acc)
    return acc


# ---- families
CLOSED_BRANCH_ORDER = ["p1", "dp1dv1", "dp1dv2", "dp1dph1", "dp1dph2", "dp1da1", "dp1dr1",
                       "q1", "dq1dv1", "dq1dv2", "dq1dph1", "dq1dph2", "dq1da1", "dq1dr1",
                       "p2", "dp2dv1", "dp2dv2", "dp2dph1", "dp2dph2", "dp2da1", "dp2dr1",
                       "q2", "dq2dv1", "dq2dv2", "dq2dph1", "dq2dph2", "dq2da1", "dq2dr1"]
SHUNT_ORDER = ["p", "dpdv", "q", "dqdv", "dqdb"]


def _cname(prefix, name):
    return prefix + name[0].upper() + name[1:]


def _functions(prefix, methods, order):
    """One inline function per method, parameters in the manifest's order."""
    lines = []
    for name in order:
        params, body = methods[name]
        expr = render(inline(body, methods))
        args = ", ".join(f"double {p}" for p in params)
        lines.append(f"OLFHIP_GEN_FN double {_cname(prefix, name)}({args}) {{ return {expr}; }}")
    return lines


def _call(prefix, name, params, bind):
    return f"{_cname(prefix, name)}(" + ", ".join(bind[p] for p in params) + ")"


def render_closed_branch(java_class, methods):
    if sorted(methods) != sorted(CLOSED_BRANCH_ORDER):
        raise ValueError("the closed-branch manifest must define the 28 scatter outputs")
    out = _header(java_class, "closed_branch")
    out += _functions("cb", methods, CLOSED_BRANCH_ORDER)
    # the 28 outputs at a state: side-1 outputs read theta1, side-2 outputs theta2
    body = ["", "// The 28 outputs of closed branch p8 = (y g1 g2 b1 b2 ksi r1 a1) at (v1, ph1, v2, ph2), in the scatter order.",
            "OLFHIP_GEN_FN void closedBranchOutputs(const double *p8, double v1, double ph1, double v2, double ph2,",
            "                                       double *out28) {",
            "    const double y = p8[0], g1 = p8[1], g2 = p8[2], b1 = p8[3], b2 = p8[4], ksi = p8[5], r1 = p8[6];",
            "    const double a1 = p8[7];",
            "    const double sinKsi = sin(ksi), cosKsi = cos(ksi);",
            "    const double theta1 = ksi - a1 + ph2 - ph1, theta2 = ksi + a1 + ph1 - ph2;",
            "    const double sinTheta1 = sin(theta1), cosTheta1 = cos(theta1);",
            "    const double sinTheta2 = sin(theta2), cosTheta2 = cos(theta2);"]
    for i, name in enumerate(CLOSED_BRANCH_ORDER):
        params = methods[name][0]
        side = "2" if name in ("p2", "q2") or name.startswith(("dp2", "dq2")) else "1"
        bind = {p: p for p in params}
        bind["sinTheta"] = "sinTheta" + side
        bind["cosTheta"] = "cosTheta" + side
        body.append(f"    out28[{i}] = {_call('cb', name, params, bind)};")
    body.append("}")
    # own-bus / far-bus coefficients read off p and q at unit points
    def at(name, side, vo, vf, trig):
        params = methods[name][0]
        bind = {p: p for p in params}
        own, far = ("v1", "v2") if side == "1" else ("v2", "v1")
        bind[own] = vo
        bind[far] = vf
        bind["sinTheta"] = trig
        bind["cosTheta"] = trig
        return _call("cb", name, params, bind)
    body += ["",
             "// A closed branch seen from one end: p = A vo^2 - B vo vf sin(theta), q = C vo^2 - B vo vf cos(theta),",
             "// theta = psi - phi_own + phi_far (vo, phi_own at this end). A, C: p and q with vo = 1, vf = 0; B: minus",
             "// the derivative of p by the far voltage with vo = 1 and sin(theta) = 1.",
             "struct HalfBranchForm {",
             "    double A, B, C, psi;",
             "};",
             "OLFHIP_GEN_FN void closedBranchHalfForms(const double *p8, HalfBranchForm *side1, HalfBranchForm *side2) {",
             "    const double y = p8[0], g1 = p8[1], g2 = p8[2], b1 = p8[3], b2 = p8[4], ksi = p8[5], r1 = p8[6];",
             "    const double a1 = p8[7];",
             "    const double sinKsi = sin(ksi), cosKsi = cos(ksi);",
             f"    side1->A = {at('p1', '1', '1.0', '0.0', '0.0')};",
             f"    side1->B = -{at('dp1dv2', '1', '1.0', '1.0', '1.0')};",
             f"    side1->C = {at('q1', '1', '1.0', '0.0', '0.0')};",
             "    side1->psi = ksi - a1;",
             f"    side2->A = {at('p2', '2', '1.0', '0.0', '0.0')};",
             f"    side2->B = -{at('dp2dv1', '2', '1.0', '1.0', '1.0')};",
             f"    side2->C = {at('q2', '2', '1.0', '0.0', '0.0')};",
             "    side2->psi = ksi + a1;",
             "}"]
    _check_half_forms(methods)
    return "\n".join(out + body + _footer()) + "\n"


def _check_half_forms(methods):
    """The generator's own guard: p and q of both sides have the own/far form the fill relies on."""
    rng = random.Random(5)
    for _ in range(50):
        v = dict(y=rng.uniform(1, 50), g1=rng.uniform(-0.1, 0.1), g2=rng.uniform(-0.1, 0.1),
                 b1=rng.uniform(-0.2, 0.2), b2=rng.uniform(-0.2, 0.2), r1=rng.uniform(0.9, 1.1))
        ksi = rng.uniform(-0.5, 0.5)
        v.update(sinKsi=math.sin(ksi), cosKsi=math.cos(ksi))
        for side, (p, q) in (("1", ("p1", "q1")), ("2", ("p2", "q2"))):
            own, far = ("v1", "v2") if side == "1" else ("v2", "v1")

            def f(name, vo, vf, s, c):
                vals = dict(v)
                vals.update({own: vo, far: vf, "sinTheta": s, "cosTheta": c})
                return evaluate(inline(methods[name][1], methods), vals)
            A, C = f(p, 1, 0, 0, 0), f(q, 1, 0, 0, 0)
            B = -f("dp1dv2" if side == "1" else "dp2dv1", 1, 1, 1, 1)
            vo, vf, th = rng.uniform(0.9, 1.1), rng.uniform(0.9, 1.1), rng.uniform(-1, 1)
            s, c = math.sin(th), math.cos(th)
            for name, want in ((p, A * vo * vo - B * vo * vf * s), (q, C * vo * vo - B * vo * vf * c)):
                got = f(name, vo, vf, s, c)
                if abs(got - want) > 1e-9 * max(1.0, abs(want)):
                    raise ValueError(f"{name} is not of the own/far form the fill gathers")


def render_shunt(java_class, methods):
    if sorted(methods) != sorted(SHUNT_ORDER):
        raise ValueError("the shunt manifest must define p, dpdv, q, dqdv, dqdb")
    return "\n".join(_header(java_class, "shunt") + _functions("sh", methods, SHUNT_ORDER) + _footer()) + "\n"


def _header(java_class, family):
    return ["// GENERATED by olf-gpu_amd/codegen/java2hip.py from OLF's " + java_class + " expressions",
            f"// (manifest tests/golden/{family}.golden.json). Do not edit: regenerate with",
            "// python olf-gpu_amd/codegen/java2hip.py",
            "#pragma once",
            "",
            "#include <cmath>",
            "",
            "// a host-only translation unit defines OLFHIP_GEN_FN (e.g. as inline) before including this header",
            "#ifndef OLFHIP_GEN_FN",
            "#include <hip/hip_runtime.h>",
            "#define OLFHIP_GEN_FN __host__ __device__ inline",
            "#endif",
            "",
            "namespace olfhip {",
            "namespace gen {",
            ""]


def _footer():
    return ["", "}  // namespace gen", "}  // namespace olfhip"]


FAMILIES = {"closed_branch": render_closed_branch, "shunt": render_shunt}


def rendered(family):
    java_class, methods = parse_manifest(os.path.join(MANIFESTS, f"{family}.golden.json"))
    return FAMILIES[family](java_class, methods)


def header_path(family):
    return os.path.join(OUT_DIR, f"{family}_terms.hpp")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--check", action="store_true", help="compare with the committed headers instead of writing")
    args = ap.parse_args(argv)
    stale = []
    for fam in FAMILIES:
        text = rendered(fam)
        path = header_path(fam)
        if args.check:
            if not os.path.exists(path) or open(path).read() != text:
                stale.append(path)
        else:
            os.makedirs(OUT_DIR, exist_ok=True)
            with open(path, "w") as f:
                f.write(text)
    if stale:
        print("out of date: " + ", ".join(stale), file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
