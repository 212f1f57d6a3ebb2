"""Proofs of the java2hip codegen (olf-gpu_amd/codegen/java2hip.py), the reference's four (codegen/java2cuda/
README.md:87-105) restated for this build:

A. parse fidelity: every parsed manifest expression evaluates as the manifest's own text does (the Java text of
   these bodies is also valid Python arithmetic; sibling calls bound to Python functions), 1e-12, and re-renders to
   the same token sequence;
B. finite-difference audit: OLF's hand Jacobian (the derivative outputs) equals central differences of OLF's value
   outputs in v1, v2, ph1, ph2, a1 and r1;
C. drift: the committed headers are exactly what the generator renders from the manifests now;
D. rendered == manifest: a host build of the committed closed-branch header (g++) gives the manifest's values, and the
   reference header's own outputs (tests/golden/closed_branch_vectors.npz), 1e-12; its own/far coefficients reproduce
   p and q; the shunt header gives g v^2, -b v^2 and their derivatives.
"""
import importlib.util
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")


def _gen():
    spec = importlib.util.spec_from_file_location("java2hip", os.path.join(REPO, "olf-gpu_amd", "codegen",
                                                                          "java2hip.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


G = _gen()


def manifest(family):
    with open(os.path.join(GOLDEN, f"{family}.golden.json")) as f:
        return json.load(f)


def python_methods(family):
    """The manifest's bodies as Python functions (their own text), sibling calls resolved among them."""
    ns = {"R2": 1.0}
    for m in manifest(family)["methods"]:
        exec(f"def {m['name']}({', '.join(m['params'])}):\n    return {m['body']}\n", ns)
    return ns


def random_point(rng):
    ksi = rng.uniform(-0.6, 0.6)
    th = rng.uniform(-1, 1)
    return dict(y=rng.uniform(1, 60), g1=rng.uniform(-0.05, 0.05), g2=rng.uniform(-0.05, 0.05),
                b1=rng.uniform(-0.1, 0.1), b2=rng.uniform(-0.1, 0.1), r1=rng.uniform(0.9, 1.1),
                v1=rng.uniform(0.85, 1.15), v2=rng.uniform(0.85, 1.15), sinKsi=math.sin(ksi), cosKsi=math.cos(ksi),
                sinTheta=math.sin(th), cosTheta=math.cos(th), v=rng.uniform(0.85, 1.15), g=rng.uniform(-1, 1),
                b=rng.uniform(-1, 1))


@pytest.mark.parametrize("family", ["closed_branch", "shunt"])
def test_a_parse_fidelity(family):
    java_class, methods = G.parse_manifest(os.path.join(GOLDEN, f"{family}.golden.json"))
    py = python_methods(family)
    rng = np.random.default_rng(1)
    for m in manifest(family)["methods"]:
        params, ast = methods[m["name"]]
        assert params == m["params"]
        inl = G.inline(ast, methods)
        for _ in range(20):
            pt = random_point(rng)
            args = [pt[p] for p in params]
            want = py[m["name"]](*args)
            got = G.evaluate(inl, dict(zip(params, args)))
            assert abs(got - want) <= 1e-12 * max(1.0, abs(want)), m["name"]
        # the parse keeps the text: rendering the un-inlined tree gives back the body's tokens
        assert _strip_float_suffix(G.tokenize(G.render(ast))) == G.tokenize(m["body"])


def _strip_float_suffix(tokens):
    """render writes integer literals as doubles (2 -> 2.0); the manifest text has them bare"""
    out = []
    for kind, val in tokens:
        if kind == "num" and val.endswith(".0") and val[:-2].isdigit():
            val = val[:-2]
        out.append((kind, val))
    return out


def outputs_from_manifest(p8, v1, ph1, v2, ph2):
    """The 28 closed-branch outputs from the manifest's own text, with the reference's angles."""
    py = python_methods("closed_branch")
    y, g1, g2, b1, b2, ksi, r1, a1 = p8
    t1, t2 = ksi - a1 + ph2 - ph1, ksi + a1 + ph1 - ph2
    env = dict(y=y, g1=g1, g2=g2, b1=b1, b2=b2, r1=r1, v1=v1, v2=v2, sinKsi=math.sin(ksi), cosKsi=math.cos(ksi))
    out = []
    for name in G.CLOSED_BRANCH_ORDER:
        params = next(m["params"] for m in manifest("closed_branch")["methods"] if m["name"] == name)
        side2 = name in ("p2", "q2") or name.startswith(("dp2", "dq2"))
        e = dict(env, sinTheta=math.sin(t2 if side2 else t1), cosTheta=math.cos(t2 if side2 else t1))
        out.append(py[name](*[e[p] for p in params]))
    return np.array(out)


def test_b_finite_difference_audit():
    vec = np.load(os.path.join(GOLDEN, "closed_branch_vectors.npz"))
    h = 1e-6
    for i in range(0, 256, 8):
        p8 = vec["p8"][i].copy()
        st = vec["state"][i].copy()     # v1 ph1 v2 ph2

        def f(p8_, st_):
            return outputs_from_manifest(p8_, *st_)
        base = f(p8, st)
        for var, (where, idx) in {"v1": ("s", 0), "ph1": ("s", 1), "v2": ("s", 2), "ph2": ("s", 3),
                                  "a1": ("p", 7), "r1": ("p", 6)}.items():
            def bump(d):
                p, s = p8.copy(), st.copy()
                (s if where == "s" else p)[idx] += d
                return f(p, s)
            fd = (bump(h) - bump(-h)) / (2 * h)
            for q in range(4):                       # p1, q1, p2, q2 and their derivative by var
                deriv_slot = 7 * q + 1 + ["v1", "v2", "ph1", "ph2", "a1", "r1"].index(var)
                assert abs(base[deriv_slot] - fd[7 * q]) <= 1e-6 * max(1.0, abs(fd[7 * q])), (i, var, q)


def test_c_committed_headers_are_the_rendering():
    assert G.main(["--check"]) == 0
    for fam in G.FAMILIES:
        assert open(G.header_path(fam)).read() == G.rendered(fam)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_d_host_build_gives_the_manifest_and_the_reference_outputs(tmp_path):
    src = tmp_path / "probe.cpp"
    src.write_text("""
#define OLFHIP_GEN_FN inline
#include <cstdio>
#include "closed_branch_terms.hpp"
#include "shunt_terms.hpp"
using namespace olfhip::gen;
int main() {
    double p8[8], s[4], o[28];
    while (std::scanf("%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", p8, p8 + 1, p8 + 2, p8 + 3, p8 + 4, p8 + 5,
                      p8 + 6, p8 + 7, s, s + 1, s + 2, s + 3) == 12) {
        closedBranchOutputs(p8, s[0], s[1], s[2], s[3], o);
        HalfBranchForm h1, h2;
        closedBranchHalfForms(p8, &h1, &h2);
        for (int i = 0; i < 28; i++) std::printf("%.17g ", o[i]);
        std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g ", h1.A, h1.B, h1.C, h1.psi, h2.A, h2.B, h2.C,
                    h2.psi);
        std::printf("%.17g %.17g %.17g %.17g %.17g\\n", shP(s[0], p8[1]), shDpdv(s[0], p8[1]), shQ(s[0], p8[3]),
                    shDqdv(s[0], p8[3]), shDqdb(s[0]));
    }
    return 0;
}
""")
    exe = tmp_path / "probe"
    subprocess.run(["g++", "-std=c++17", "-O0", "-Wall", "-Wextra", "-Werror", "-I",
                    os.path.join(REPO, "olf-gpu_amd", "csrc", "generated"), str(src), "-o", str(exe)], check=True)
    vec = np.load(os.path.join(GOLDEN, "closed_branch_vectors.npz"))
    lines = "\n".join(" ".join(f"{x:.17g}" for x in np.concatenate([vec["p8"][i], vec["state"][i]]))
                      for i in range(len(vec["p8"])))
    res = subprocess.run([str(exe)], input=lines, capture_output=True, text=True, check=True).stdout.split("\n")
    rows = np.array([[float(x) for x in r.split()] for r in res if r.strip()])
    out, forms, shunt = rows[:, :28], rows[:, 28:36], rows[:, 36:]
    # the reference header's own outputs
    np.testing.assert_allclose(out, vec["out"], rtol=1e-12, atol=1e-12)
    for i in range(0, 256, 16):
        p8, st = vec["p8"][i], vec["state"][i]
        np.testing.assert_allclose(out[i], outputs_from_manifest(p8, *st), rtol=1e-12, atol=1e-12)
        v1, ph1, v2, ph2 = st
        A1, B1, C1, psi1, A2, B2, C2, psi2 = forms[i]
        t1, t2 = psi1 - ph1 + ph2, psi2 - ph2 + ph1
        assert abs(out[i][0] - (A1 * v1 * v1 - B1 * v1 * v2 * math.sin(t1))) < 1e-12
        assert abs(out[i][7] - (C1 * v1 * v1 - B1 * v1 * v2 * math.cos(t1))) < 1e-12
        assert abs(out[i][14] - (A2 * v2 * v2 - B2 * v1 * v2 * math.sin(t2))) < 1e-12
        assert abs(out[i][21] - (C2 * v2 * v2 - B2 * v1 * v2 * math.cos(t2))) < 1e-12
        assert abs(B1 - B2) < 1e-12
        v, g, b = st[0], p8[1], p8[3]
        np.testing.assert_allclose(shunt[i], [g * v * v, 2 * g * v, -b * v * v, -2 * b * v, -v * v], rtol=1e-14)
