"""Write golden vectors of the reference's NRG coefficient tables.

The reference ships its Clebsch-Gordan / fermion-sign coefficient tables as
expression fragments under coefnew/<sym>/<sym>-<n>ch-*.dat (for example
coefnew/qs/qs-1ch-offdiag.dat:1-4, coefnew/qs/qs-1ch-doubletp.dat:1-5,
coefnew/qsz/qsz-3ch-spinupa.dat).  This script reads them as text, evaluates
every coefficient numerically at SS (or SSZ) = 1..SSMAX and writes the numbers
as JSON fixtures next to itself.  The fixtures are data: the oracle builds its
tables from them, and the engine's own derivation of the tables is checked
against them.

Run in a container where the reference tree is mounted:
    python tests/golden/make_table_fixtures.py /root/reference
"""
import ast
import json
import math
import os
import re
import sys

SSMAX = 64
HERE = os.path.dirname(os.path.abspath(__file__))


def S(ss):
    return (ss - 1) / 2.0


def Power(a, b):
    return a ** b


def _sqrt(x):
    if x < 0:
        raise ValueError("sqrt of negative")
    return math.sqrt(x)


_FUNCS = {"S": S, "Power": Power, "sqrt": _sqrt}
_BINOPS = {ast.Add: lambda a, b: a + b, ast.Sub: lambda a, b: a - b, ast.Mult: lambda a, b: a * b,
           ast.Div: lambda a, b: a / b, ast.Pow: lambda a, b: a ** b}


def _eval_node(node, env):
    """Arithmetic over numbers, the variables in `env`, + - * / ** and calls
    of S / Power / sqrt; any other syntax is refused (the expression text
    comes from an untrusted tree and is never executed)."""
    if isinstance(node, ast.Expression):
        return _eval_node(node.body, env)
    if isinstance(node, ast.Constant) and type(node.value) in (int, float):
        return node.value
    if isinstance(node, ast.Name) and node.id in env:
        return env[node.id]
    if isinstance(node, ast.BinOp) and type(node.op) in _BINOPS:
        return _BINOPS[type(node.op)](_eval_node(node.left, env), _eval_node(node.right, env))
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        v = _eval_node(node.operand, env)
        return -v if isinstance(node.op, ast.USub) else v
    if (isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and node.func.id in _FUNCS
            and not node.keywords):
        return _FUNCS[node.func.id](*[_eval_node(a, env) for a in node.args])
    raise ValueError("unsupported expression element: %s" % ast.dump(node)[:80])


def evaluate(expr, env):
    try:
        v = float(_eval_node(ast.parse(expr, mode="eval"), env))
        if not math.isfinite(v):
            return None
        return v
    except (ZeroDivisionError, ValueError, OverflowError):
        return None


def parse_invar(text):
    m = re.fullmatch(r"\s*Invar\((.*),(.*)\)\s*", text)
    if not m:
        raise ValueError(text)
    return m.group(1).strip(), m.group(2).strip()


def offset_of(expr, var):
    """Invar components are 'k + var' / 'var' / '-1 + var'; return k."""
    e = expr.replace(" ", "")
    if e == var:
        return 0
    m = re.fullmatch(r"(-?\d+)\+" + var, e)
    if m:
        return int(m.group(1))
    raise ValueError(expr)


def split_top(s):
    """Split on commas not nested inside parentheses."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    out.append(cur)
    return [x.strip() for x in out]


def read_lines(path):
    with open(path) as f:
        return [l.strip() for l in f if l.strip()]


def table_invars(path, name):
    res = {}
    for l in read_lines(path):
        m = re.fullmatch(name + r"\[(\d+)\]\s*=\s*Invar\((-?\d+),\s*(-?\d+)\);", l)
        res[int(m.group(1))] = [int(m.group(2)), int(m.group(3))]
    return [res[k] for k in sorted(res)]


def macro_table(path, macro, spinvar):
    rows = []
    for l in read_lines(path):
        m = re.fullmatch(macro + r"\((.*)\);", l)
        if not m:
            continue
        args = split_top(m.group(1))
        head = [int(a) for a in args[:-1]]
        expr = args[-1]
        vals = [evaluate(expr, {spinvar: ss}) for ss in range(1, SSMAX + 1)]
        rows.append({"idx": head, "values": vals})
    return rows


def recalc_f_table(path, spinvar):
    rows = []
    for l in read_lines(path):
        m = re.fullmatch(r"\{(.*)\},?", l)
        args = split_top(m.group(1))
        i1, ip, expr = int(args[0]), int(args[1]), args[2]
        vals = [evaluate(expr, {spinvar: ss}) for ss in range(1, SSMAX + 1)]
        rows.append({"i1": i1, "ip": ip, "values": vals})
    return rows


def recalc_table(path, spinvar):
    rows = []
    for l in read_lines(path):
        m = re.fullmatch(r"\{(.*)\},?", l)
        args = split_top(m.group(1))
        i1, ip = int(args[0]), int(args[1])
        q1, s1 = parse_invar(args[2])
        qp, sp = parse_invar(args[3])
        expr = args[4]
        vals = [evaluate(expr, {spinvar: ss}) for ss in range(1, SSMAX + 1)]
        rows.append({
            "i1": i1, "ip": ip,
            "IN1": [offset_of(q1, "q1"), offset_of(s1, spinvar)],
            "INp": [offset_of(qp, "q1"), offset_of(sp, spinvar)],
            "values": vals,
        })
    return rows


def main(ref):
    out_all = {}
    for sym, nchs in (("qs", (1, 2)), ("qsz", (1, 3))):
        spin1 = "ss1" if sym == "qs" else "ssz1"
        spinp = "ssp" if sym == "qs" else "sszp"
        spin = "ss" if sym == "qs" else "ssz"
        for nch in nchs:
            d = os.path.join(ref, "coefnew", sym)
            pre = os.path.join(d, "%s-%dch-" % (sym, nch))
            fx = {"sym": sym.upper(), "channels": nch, "ssmax": SSMAX}
            fx["In"] = table_invars(pre + "In2.dat", "In")
            fx["QN"] = table_invars(pre + "QN.dat", "QN")
            fx["offdiag"] = macro_table(pre + "offdiag.dat", "OFFDIAG", spin)
            fx["diag"] = macro_table(pre + "diag.dat", "DIAG", spin)
            if sym == "qsz":
                fx["spinz"] = macro_table(pre + "spinz.dat", "SPINZ", spin)
                fx["diag_up"] = macro_table(pre + "diag-UP.dat", "DIAG_UP", spin)
            letters = "abc"[:nch]
            fx["spinup"] = [recalc_f_table(pre + "spinup%s.dat" % c, spinp) for c in letters]
            fx["spindown"] = [recalc_f_table(pre + "spindown%s.dat" % c, spinp) for c in letters]
            fx["doubletp"] = recalc_table(pre + "doubletp.dat", spin1)
            fx["doubletm"] = recalc_table(pre + "doubletm.dat", spin1)
            fx["singlet"] = recalc_table(pre + "singlet.dat", spin1)
            name = "tables_%s_%dch.json" % (sym, nch)
            with open(os.path.join(HERE, name), "w") as f:
                json.dump(fx, f, separators=(",", ":"))
            out_all[name] = len(fx["offdiag"])
    print(out_all)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
