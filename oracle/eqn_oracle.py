"""Pure-Python restatement of GPUE's equation-string parser and evaluator.

TEST INFRASTRUCTURE ONLY (the checker of gpue_amd's C++ parser).
Follows src/dynamic.cu:227-470 (parse_eqn), :472-507 (tree_to_array) and
:527-603 (evaluate_eqn_gpu): bracket pairs recorded in closing order, a pair
spanning the whole string stripped, split at the leftmost '-', else '+', '/',
'*', '^' outside brackets, digits/'.' literals, x y z t leaves, one-argument
functions (two-argument pow), names resolved to earlier equations or Grid
doubles; function nodes keep their left child only.

Pinned by the reference KAT (src/unit_test.cu:399-466), src/example.cfg and a
brute-force equivalence with ordinary precedence on the operators where the
leftmost-split rule coincides with it (+, *, ^ with integer exponents).
"""
import math

FUNCS = {"cos": 6, "sin": 7, "tan": 8, "sqrt": 9, "exp": 10, "pow": 5, "jn": 11, "yn": 12, "k2n": 13}
OPS = {"+": 1, "-": 2, "*": 3, "/": 4, "^": 5}


class Node:
    def __init__(self):
        self.val = 0.0
        self.var = None
        self.op = 0
        self.left = None
        self.right = None


def parse(s, doubles=None, asts=None):
    doubles = doubles or {}
    asts = asts or {}
    if not s:
        raise ValueError("empty expression")
    stack, ign = [], []
    for i, ch in enumerate(s):
        if ch in "([":
            stack.append(i)
        if ch in "])":
            ign += [stack.pop(), i]
    tmp = s
    if ign:
        if ign[-1] == len(s) - 1 and ign[-2] == 0:
            ign = ign[:-2]
            s = s[1:-1]
            ign = [q - 1 for q in ign]
        tmp = s
        tp = list(ign)
        for i in range(0, len(tp), 2):
            a, b = tp[i], tp[i + 1]
            tmp = tmp[:a] + tmp[b + 1:]
            for j in range(i + 2, len(tp)):
                if tp[j] > a:
                    tp[j] -= b - a + 1
    n = Node()
    if all(ch in "0123456789." for ch in tmp):
        n.val = float(tmp) if tmp not in ("", ".") else 0.0
        return n
    if len(tmp) == 1 and tmp in "xyzt":
        n.var = tmp
        return n
    mop = -1
    for m in "-+/*^":
        f = tmp.find(m)
        if f >= 0:
            mop = f
            break
    if mop < 0:
        if tmp in FUNCS:
            mp = len(tmp) - 1
            chk = s[mp + 1:]
            n.op = FUNCS[tmp]
            c = chk.find(",")
            if c >= 0:
                n.left = parse(chk[1:c], doubles, asts)
                n.right = parse(chk[c + 1:len(chk) - 1], doubles, asts)
            else:
                n.left = parse(chk, doubles, asts)
            return n
        if tmp in asts:
            return asts[tmp]
        if tmp in doubles:
            n.val = float(doubles[tmp])
            return n
        raise ValueError(f"no value {tmp!r}")
    if ign:
        count = 0
        for _ in range(mop + 1):
            for j in range(0, len(ign), 2):
                if ign[j] == count:
                    count += ign[j + 1] - ign[j] + 1
            count += 1
        mop = count - 1
    n.op = OPS[s[mop]]
    n.left = parse(s[:mop], doubles, asts)
    n.right = parse(s[mop + 1:], doubles, asts)
    return n


def evaluate(n, x=0.0, y=0.0, z=0.0, t=0.0):
    if n.op == 0:
        return {"x": x, "y": y, "z": z, "t": t}[n.var] if n.var else n.val
    a = evaluate(n.left, x, y, z, t) if n.left is not None else 0.0
    b = evaluate(n.right, x, y, z, t) if (n.right is not None and n.op < 6) else 0.0
    op = n.op
    if op == 1:
        return a + b
    if op == 2:
        return a - b
    if op == 3:
        return a * b
    if op == 4:
        return a / b
    if op == 5:
        return math.pow(a, int(b))
    return {6: math.cos, 7: math.sin, 8: math.tan, 9: math.sqrt, 10: math.exp}[op](a)


def count_nodes(n):
    if n.op == 0:
        return 1
    return 1 + count_nodes(n.left) + (count_nodes(n.right) if n.op < 6 else 0)
