"""Writes duffy_map_vectors.json: input/output vectors of the reference's
Duffy region maps, _vertex_map (the corner-relative part) and _edge_map
(src/DuffyKernel.jl:288-314, 375-408), for the oracle's parity tests.

No closed form fixes the edge maps (they are not a valid change of variables),
and the reference cannot be run here (no Julia).  This script therefore reads
the two functions' region branches from the reference's source as text, checks
that every statement is a plain assignment of +, -, * arithmetic over
rho, eta1..eta3 and numeric literals (anything else is refused), and evaluates
those expressions itself, in Float64, left to right as Julia does, at fixed
points.  Only the numbers are written; nothing of the source is kept.

Run in a container that holds the reference, from the repository root:

    python tests/golden/make_duffy_map_vectors.py [/root/reference]
"""
import ast
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "duffy_map_vectors.json")
NAMES = {"ρ": 0, "η1": 1, "η2": 2, "η3": 3}


def value(node, args, env):
    """Evaluate a restricted arithmetic expression tree (no calls, no
    attributes, no other operators)."""
    if isinstance(node, ast.Constant) and type(node.value) in (int, float):
        return node.value
    if isinstance(node, ast.Name):
        if node.id in NAMES:
            return args[NAMES[node.id]]
        if node.id in env:
            return env[node.id]
        raise ValueError(f"unknown name {node.id}")
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
        return -value(node.operand, args, env)
    if isinstance(node, ast.BinOp) and type(node.op) in (ast.Add, ast.Sub, ast.Mult):
        a, b = value(node.left, args, env), value(node.right, args, env)
        # Julia promotes Int to Float64 before each operation; Python does the same
        return {ast.Add: lambda: a + b, ast.Sub: lambda: a - b, ast.Mult: lambda: a * b}[type(node.op)]()
    raise ValueError(f"refused expression node {ast.dump(node)[:60]}")


def branches(src, func, targets, last_region):
    """{region: [statement ast, ...]} of the if/elseif/else chain on `region`
    at the top of `func`."""
    m = re.search(r"function " + func + r"\(", src)
    body = src[m.end():].split("\n", 1)[1].splitlines()  # after the signature line
    out, cur = {}, None
    for line in body:
        code = line.split("#", 1)[0].strip()
        if not code:
            continue
        h = re.fullmatch(r"(?:if|elseif) region == (\d+)", code)
        if h:
            cur = int(h.group(1)); out[cur] = []
            continue
        if code == "else":
            cur = last_region; out[cur] = []
            continue
        if code == "end":
            break
        tree = ast.parse(code, mode="exec").body
        if len(tree) != 1 or not isinstance(tree[0], ast.Assign):
            raise ValueError(f"refused statement in {func}")
        out[cur].append(tree[0])
    for r, stmts in out.items():
        for st in stmts:  # the target names must be the map's outputs
            tg = st.targets[0]
            names = [e.id for e in tg.elts] if isinstance(tg, ast.Tuple) else [tg.id]
            if not set(names) <= set(targets):
                raise ValueError(f"unexpected target {names}")
    return out


def run(stmts, targets, args):
    env = {}
    for st in stmts:
        tg = st.targets[0]
        if isinstance(tg, ast.Tuple):
            vals = [value(e, args, env) for e in st.value.elts]
            for t, v in zip(tg.elts, vals):
                env[t.id] = v
        else:
            env[tg.id] = value(st.value, args, env)
    return [float(env[t]) for t in targets]


def points():
    """Fixed inputs in [0,1]^4: the 3-point Gauss nodes mapped to [0,1], the
    end points, and seeded uniform draws."""
    g = (np.polynomial.legendre.leggauss(3)[0] + 1.0) / 2.0
    fixed = [[g[0], g[1], g[2], g[1]], [g[2], g[2], g[0], g[2]], [1.0, 1.0, 1.0, 1.0], [1.0, 0.0, 1.0, 0.0],
             [0.5, 0.25, 0.75, 0.125]]
    rng = np.random.default_rng(20240611)
    return [list(map(float, p)) for p in fixed] + rng.uniform(0, 1, (19, 4)).tolist()


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with open(os.path.join(ref, "src", "DuffyKernel.jl"), encoding="utf-8") as f:
        src = f.read()
    pts = points()
    data = {"source": "src/DuffyKernel.jl:288-314 (_vertex_map, corner-relative part), 375-408 (_edge_map)",
            "inputs": pts, "input_names": ["rho", "eta1", "eta2", "eta3"], "maps": {}}
    for kind, func, targets, last in (("vertex", "_vertex_map", ["du", "dv", "ds", "dt"], 8),
                                      ("edge", "_edge_map", ["u", "v", "s", "t"], 5)):
        br = branches(src, func, targets, last)
        assert sorted(br) == list(range(1, last + 1)), sorted(br)
        data["maps"][kind] = {str(r): [run(br[r], targets, p) for p in pts] for r in sorted(br)}
    with open(OUT, "w") as f:
        json.dump(data, f)
        f.write("\n")
    print(f"wrote {OUT}: {sum(len(v) for v in data['maps'].values())} regions x {len(pts)} points")


if __name__ == "__main__":
    main()
