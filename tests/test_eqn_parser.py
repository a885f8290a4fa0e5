"""gpue_amd's C++ equation parser (host code behind the C ABI; no GPU needed)
against GPUE's parser semantics (src/dynamic.cu:187-603)."""
import itertools
import math
import os

import numpy as np
import pytest

import gpue_amd as G
from oracle import eqn_oracle as EO




def test_reference_kat():
    # src/unit_test.cu:399-466: evaluates to 3x + 14406 exactly at x = 0..63
    nodes, td = G.eqn_parse("(((3*x)+7)+(5-7)+cos(0)*1)+pow(120,2)")
    assert not td
    for i in range(64):
        assert G.eqn_eval(nodes, x=float(i)) == (((3 * i) + 7) + (5 - 7) + math.cos(0) * 1) + math.pow(120, 2)
    n = EO.parse("(((3*x)+7)+(5-7)+cos(0)*1)+pow(120,2)")
    assert len(nodes) == EO.count_nodes(n)
    for i in range(64):
        assert EO.evaluate(n, x=float(i)) == 3 * i + 14406


def test_example_cfg():
    # src/example.cfg: rad = 10, omegaR = 7071, V = rad*omegaR*x*y*z; the leftmost
    # split makes products right-associative: rad*(omegaR*(x*(y*z)))
    nodes, _ = G.eqn_parse("rad*omegaR*x*y*z", {"rad": 10.0, "omegaR": 7071.0})
    for (x, y, z) in [(0.1, 0.2, 0.3), (-1.0, 2.0, 0.5)]:
        assert G.eqn_eval(nodes, x, y, z) == 10.0 * (7071.0 * (x * (y * z)))


def test_gauge_cfg():
    # src/gauge.cfg: Ax = y*y*omega*omegaX*30000
    d = {"omega": 0.5, "omegaX": 6.283, "omegaY": 6.283}
    nodes, _ = G.eqn_parse("y * y * omega * omegaX * 30000".replace(" ", ""), d)
    assert G.eqn_eval(nodes, 0.0, 0.3) == 0.3 * (0.3 * (0.5 * (6.283 * 30000)))


@pytest.mark.parametrize("expr,value", [
    ("1-2+3", 1 - (2 + 3)),         # leftmost '-' splits first
    ("10-4-3", 10 - (4 - 3)),
    ("8/2*2", 8 / (2 * 2)),         # '/' before '*'
    ("2^3^2", 2 ** (3 ** 2)),
    ("(1-2)+3", (1 - 2) + 3),
    ("[2+3]*4", 20.0),
    ("2*pow(3,2)", 18.0),
    ("sqrt(16)+exp(0)", 5.0),
])
def test_precedence_quirks(expr, value):
    nodes, _ = G.eqn_parse(expr)
    assert G.eqn_eval(nodes) == value
    assert EO.evaluate(EO.parse(expr)) == value


def test_bruteforce_plus_times_pow():
    # on +, *, ^ the leftmost-split rule equals ordinary precedence: enumerate
    rng = np.random.default_rng(0)
    leaves = ["x", "y", "2", "3", "0.5", "1.25"]
    ops = ["+", "*", "^"]
    for _ in range(400):
        k = int(rng.integers(1, 5))
        toks = [leaves[int(rng.integers(len(leaves)))]]
        for _ in range(k):
            op = ops[int(rng.integers(3))]
            rhs = "2" if op == "^" else leaves[int(rng.integers(len(leaves)))]
            toks += [op, rhs]
        expr = "".join(toks)
        x, y = 0.7, -1.3
        py = expr.replace("^", "**")
        want = eval(py, {"x": x, "y": y})
        nodes, _ = G.eqn_parse(expr)
        got = G.eqn_eval(nodes, x, y)
        assert math.isclose(got, want, rel_tol=1e-14, abs_tol=1e-14), expr


def has_t(n):
    if n.op == 0:
        return n.var == "t"
    return has_t(n.left) or (n.op < 6 and has_t(n.right))


def test_corpus_matches_python_restatement():
    corpus = ["x*x+y*y", "0.5*(x*x+y*y)", "exp(0-x*x)", "cos(x)*sin(y)", "sqrt(x*x+1)", "x-y-t",
              "(x+1)/(y+2)/3", "tan(0.1*x)+pow(y,3)", "[x+y]*[x-y]", "x/y*2", "1+2*3-4/5+6", "t*t*x"]
    for expr in corpus:
        nodes, td = G.eqn_parse(expr)
        ref = EO.parse(expr)
        assert len(nodes) == EO.count_nodes(ref), expr
        assert td == has_t(ref), expr
        for (x, y, z, t) in [(0.3, -0.7, 1.1, 0.25), (1.5, 2.0, -0.5, 3.0)]:
            a = G.eqn_eval(nodes, x, y, z, t)
            b = EO.evaluate(ref, x, y, z, t)
            assert a == b or math.isclose(a, b, rel_tol=1e-15), expr


@pytest.mark.parametrize("bad", ["", "(x+1", "x+1)", "foo*2", "k2n(1,2)", "x+", "exp(-x*x)"])
def test_rejects(bad):
    with pytest.raises(G.GpueError):
        G.eqn_parse(bad)
