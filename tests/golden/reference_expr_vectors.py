"""Write reference_expr_vectors.json: input/output vectors of the reference's generated per-point
kernels, evaluated from their expression text.

Reads, as text, the generated plug-in headers of the reference checkout:
  ChorinOps.H  vam_pressure_blocks, vam_pressure_residual, vam_corrector        (:100-138)
  Numerics.H   numerical_flux, numerical_fluctuations, local_max_abs_eigenvalue (:32-107)
  Model.H      source, source_jacobian_wrt_variables, source_jacobian_wrt_aux_variables,
               update_variables, update_aux_variables, eigenvalues, boundary_conditions,
               initial_condition                                              (:94-274, 356-375)
Nothing of the reference is compiled, imported or handed to eval(): every statement of a
function body is parsed with Python's ast module and evaluated by the small interpreter below,
which accepts only + - * / and unary minus, the calls pow / max / min / fabs / sqrt, array
subscripts by integer literals, local scalars and numeric literals.  Evaluation is IEEE double in
the order the C++ states it (no contraction), with the C library's pow for std::pow and the
std::max / std::min tie rules, so each output is what the compiled header computes.

The inputs are seeded (numpy PCG64) and include dry, near-dry and negative depths.  Run in a
checkout that has the reference at REFERENCE (default /root/reference):
    python tests/golden/reference_expr_vectors.py [REFERENCE]
The JSON it writes is committed; the tests read only the JSON.
"""
import ast
import json
import math
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "reference_expr_vectors.json")


# ----------------------------------------------------------------- C++ text -> statements
def strip_comments(src):
    return re.sub(r"//[^\n]*", "", src)


def function_body(src, name, must_contain=None):
    """Parameter list and body text of the first definition of `name` (a signature containing
    `must_contain`, when given)."""
    for m in re.finditer(r"\b%s\s*\(" % re.escape(name), src):
        i = m.end()
        depth = 1
        while depth:
            depth += {"(": 1, ")": -1}.get(src[i], 0)
            i += 1
        params = src[m.end():i - 1]
        j = i
        while src[j].isspace():
            j += 1
        if src[j] != "{":
            continue          # a call, not a definition
        if must_contain and must_contain not in params:
            continue
        depth, k = 0, j
        while True:
            depth += {"{": 1, "}": -1}.get(src[k], 0)
            k += 1
            if depth == 0:
                break
        return params, src[j + 1:k - 1]
    raise KeyError("no definition of %s" % name)


def statements(body):
    """Split a body into statements and if-blocks: ('stmt', text) / ('if', cond, [items])."""
    items, i, n = [], 0, len(body)
    while i < n:
        while i < n and body[i].isspace():
            i += 1
        if i >= n:
            break
        if body.startswith("if", i) and re.match(r"if\s*\(", body[i:]):
            p = body.index("(", i)
            depth, q = 1, p + 1
            while depth:
                depth += {"(": 1, ")": -1}.get(body[q], 0)
                q += 1
            cond = body[p + 1:q - 1]
            b0 = body.index("{", q)
            depth, r = 1, b0 + 1
            while depth:
                depth += {"{": 1, "}": -1}.get(body[r], 0)
                r += 1
            items.append(("if", cond, statements(body[b0 + 1:r - 1])))
            i = r
            continue
        j = body.index(";", i)
        items.append(("stmt", body[i:j].strip()))
        i = j + 1
    return items


def to_python(expr):
    e = expr.replace("std::", "")
    e = re.sub(r"\(\s*T\s*\)", "", e)
    return e.strip()


# ----------------------------------------------------------------- whitelisted evaluator
def _cxx_max(a, b):
    return b if a < b else a


def _cxx_min(a, b):
    return b if b < a else a


FUNCS = {"pow": math.pow, "max": _cxx_max, "min": _cxx_min, "fabs": math.fabs, "sqrt": math.sqrt}


class Interp:
    def __init__(self, env):
        self.env = env

    def ev(self, node):
        if isinstance(node, ast.Expression):
            return self.ev(node.body)
        if isinstance(node, ast.Constant) and type(node.value) in (int, float):
            return node.value
        if isinstance(node, ast.Name):
            v = self.env[node.id]
            if isinstance(v, list):
                raise ValueError("array used as a scalar: " + node.id)
            return v
        if isinstance(node, ast.Subscript):
            arr = self.env[node.value.id]
            idx = node.slice
            if not (isinstance(idx, ast.Constant) and type(idx.value) is int):
                raise ValueError("non-literal subscript")
            return arr[idx.value]
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
            v = self.ev(node.operand)
            return -v if isinstance(node.op, ast.USub) else +v
        if isinstance(node, ast.BinOp) and isinstance(node.op, (ast.Add, ast.Sub, ast.Mult, ast.Div)):
            a, b = self.ev(node.left), self.ev(node.right)
            if isinstance(node.op, ast.Add):
                return a + b
            if isinstance(node.op, ast.Sub):
                return a - b
            if isinstance(node.op, ast.Mult):
                return a * b
            if type(a) is int and type(b) is int:
                raise ValueError("integer division in the expression text")
            return a / b
        if isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and node.func.id in FUNCS \
                and not node.keywords:
            return FUNCS[node.func.id](*[float(self.ev(a)) for a in node.args])
        if isinstance(node, ast.Compare) and len(node.ops) == 1 and isinstance(node.ops[0], ast.Eq):
            return self.ev(node.left) == self.ev(node.comparators[0])
        raise ValueError("construct outside the whitelist: " + ast.dump(node)[:80])

    def expr(self, text):
        return self.ev(ast.parse(to_python(text), mode="eval"))


def run_body(items, env, out_name=None):
    """Execute statements; returns the returned array (or env[out_name] for void kernels)."""
    it = Interp(env)
    for item in items:
        if item[0] == "if":
            if it.expr(item[1]):
                r = run_body(item[2], env, out_name)
                if r is not None:
                    return r
            continue
        s = item[1]
        m = re.fullmatch(r"return\s+(\w+)", s)
        if m:
            return list(env[m.group(1)])
        m = re.fullmatch(r"SimpleArray<\s*T\s*,\s*(\d+)\s*>\s+(\w+)(\s*=\s*\{\s*0\s*\})?", s)
        if m:
            env[m.group(2)] = [0.0] * int(m.group(1)) if m.group(3) else [None] * int(m.group(1))
            continue
        m = re.fullmatch(r"T\s+(\w+)\s*=\s*(.+)", s, re.S)
        if m:
            env[m.group(1)] = float(it.expr(m.group(2)))
            continue
        m = re.fullmatch(r"(\w+)\[(\d+)\]\s*=\s*(.+)", s, re.S)
        if m:
            env[m.group(1)][int(m.group(2))] = float(it.expr(m.group(3)))
            continue
        raise ValueError("statement outside the whitelist: " + s[:80])
    return list(env[out_name]) if out_name else None


def evaluate(src, name, args, n_out=None, must_contain=None):
    params, body = function_body(src, name, must_contain)
    names = [re.findall(r"(\w+)\s*$", p.strip())[0] for p in params.split(",")]
    env = {k: (list(v) if isinstance(v, (list, tuple, np.ndarray)) else v)
           for k, v in zip(names, args)}
    out_name = None
    if n_out is not None:              # void kernel writing into its last pointer argument
        out_name = names[-1]
        env[out_name] = [None] * n_out
    res = run_body(statements(body), env, out_name)
    if any(v is None for v in res):
        raise ValueError("%s left an output unassigned" % name)
    return res


# ----------------------------------------------------------------- seeded inputs
def fl(a):
    return [float(x) for x in a]


def vam_cases(rng, n=48):
    cases = []
    for _ in range(n):
        Q = np.concatenate([rng.uniform(0, 0.1, 1), rng.uniform(0.3, 1.5, 1), rng.uniform(-0.3, 0.3, 4),
                            rng.uniform(-0.1, 0.1, 2), rng.uniform(-1, 1, 2)])
        A = rng.uniform(-1, 1, 20) * np.array([1, 3, 1, 3] * 2 + [0.3, 1, 0.3, 1] + [0.5, 2, 0.5, 2] +
                                              [0.5] * 4)
        p = [9.81, float(rng.uniform(0.8, 1.2)), 1e-3, 0.1]
        dt = float(rng.uniform(1e-4, 1e-2))
        cases.append({"Q": fl(Q), "Qaux": fl(A), "p": p, "dt": dt})
    return cases


def sme1d_state(rng):
    h = rng.choice([rng.uniform(0.02, 1.5), rng.uniform(0.005, 0.02), 0.0, 1e-9, -1e-3],
                   p=[0.7, 0.1, 0.07, 0.07, 0.06])
    return [float(rng.uniform(0, 0.3)), float(h), float(rng.uniform(-1, 1))]


def sme1d_cases(rng, n=64):
    cases = []
    for k in range(n):
        qm, qp = sme1d_state(rng), sme1d_state(rng)
        am = fl(np.concatenate([rng.uniform(-1, 1, 3), rng.uniform(0, 3, 1)]))
        ap = fl(np.concatenate([rng.uniform(-1, 1, 3), rng.uniform(0, 3, 1)]))
        p = [9.81, 1.0, 1.0, 0.5, 0.0] if k % 2 == 0 else \
            [float(rng.uniform(5, 12)), float(rng.uniform(0.8, 1.2)), float(rng.uniform(0.1, 2)),
             float(rng.uniform(0.1, 1)), float(rng.uniform(-0.2, 0.2))]
        nrm = [float(rng.choice([1.0, -1.0, rng.uniform(-1.5, 1.5)]))]
        cases.append({"Q_minus": qm, "Q_plus": qp, "Qaux_minus": am, "Qaux_plus": ap, "p": p, "n": nrm,
                      "X": [float(rng.uniform(0, 1)), 0.0, 0.0], "time": float(rng.uniform(0, 1)),
                      "dt": float(rng.uniform(1e-4, 1e-1))})
    return cases


# ----------------------------------------------------------------- main
def main(ref):
    read = lambda f: strip_comments(open(os.path.join(ref, f)).read())
    chorin, numerics, model = read("ChorinOps.H"), read("Numerics.H"), read("Model.H")
    rng = np.random.default_rng(20261204)
    out = {"source": "expression text of ChorinOps.H, Numerics.H and Model.H, evaluated by "
                     "tests/golden/reference_expr_vectors.py",
           "vam": [], "sme0_1d_numerics": [], "sme0_1d_model": []}
    for c in vam_cases(rng):
        args = (c["Q"], c["Qaux"], c["p"], c["dt"])
        out["vam"].append(dict(c, blocks=evaluate(chorin, "vam_pressure_blocks", args + (None,), 20),
                               residual=evaluate(chorin, "vam_pressure_residual", args + (None,), 2),
                               corrector=evaluate(chorin, "vam_corrector", args + (None,), 6)))
    for c in sme1d_cases(rng):
        a6 = (c["Q_minus"], c["Q_plus"], c["Qaux_minus"], c["Qaux_plus"], c["p"], c["n"])
        out["sme0_1d_numerics"].append(dict(
            c, flux=evaluate(numerics, "numerical_flux", a6),
            fluctuations=evaluate(numerics, "numerical_fluctuations", a6),
            lmax_minus=evaluate(numerics, "local_max_abs_eigenvalue",
                                (c["Q_minus"], c["Qaux_minus"], c["p"], c["n"]))))
        Q, A, p = c["Q_minus"], c["Qaux_minus"], c["p"]
        out["sme0_1d_model"].append(dict(
            Q=Q, Qaux=A, p=p, n=c["n"], X=c["X"], time=c["time"], dt=c["dt"],
            source=evaluate(model, "source", (Q, A, p, c["time"], c["dt"], c["X"])),
            source_jacobian_wrt_variables=evaluate(model, "source_jacobian_wrt_variables", (Q, A, p)),
            source_jacobian_wrt_aux_variables=evaluate(model, "source_jacobian_wrt_aux_variables", (Q, A, p)),
            update_variables=evaluate(model, "update_variables", (Q, A, p, c["dt"])),
            update_aux_variables=evaluate(model, "update_aux_variables", (Q, A, p, c["time"], c["X"]),
                                          must_contain="const T time"),
            eigenvalues=evaluate(model, "eigenvalues", (Q, A, p, c["n"])),
            boundary_conditions=evaluate(model, "boundary_conditions",
                                         (0, Q, A, p, c["n"], c["X"], c["time"], 0.0),
                                         must_contain="bc_idx"),
            initial_condition=evaluate(model, "initial_condition", (c["X"], p))))
    with open(OUT, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", OUT, {k: len(v) for k, v in out.items() if isinstance(v, list)})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
