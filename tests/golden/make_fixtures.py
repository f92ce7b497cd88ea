"""Extract reference-held numeric tables into small JSON fixtures.

Reads (as text, no execution) the orthonormal polynomial coefficient tables
and the 1D Gauss rules that BoSSS ships:
  - L2-foundation/BoSSS.Foundation/Square.cs:192-6495 (2D, [-1,1]^2)
  - L2-foundation/BoSSS.Foundation/Cube_polynomials.cs:13-13846 (3D, [-1,1]^3)
  - L2-foundation/BoSSS.Foundation/LineQuadRules/Line{n}.txt
Only the modes up to total degree 4 and the rules with n <= 12 nodes are kept.
Run from this container (the reference tree is not present at test time):
    python tests/golden/make_fixtures.py
"""
import json
import os
import re

REF = "/root/reference/src/L2-foundation/BoSSS.Foundation"
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_polys(path, D, maxdeg):
    txt = open(path, encoding="utf-8-sig").read()
    parts = re.split(r"OrthonormalPolynomials\[(\d+)\]\s*=\s*p;", txt)
    out = {}
    for i in range(1, len(parts) - 1, 2):
        idx = int(parts[i])
        body = parts[i + 1].split("new Polynomial")[0]
        mons = re.findall(r"AddCoeff\(([-+0-9.eE]+),\s*new int\[\]\s*\{\s*([0-9, ]+)\}", body)
        terms = [[float(c), [int(x) for x in e.split(",")]] for c, e in mons]
        deg = max(sum(t[1]) for t in terms)
        if deg <= maxdeg:
            out[idx] = terms
    n = 0
    while n in out:
        n += 1
    return [out[i] for i in range(n)]


def parse_rules(nmax):
    rules = {}
    for n in range(1, nmax + 1):
        lines = open(os.path.join(REF, "LineQuadRules", f"Line{n}.txt")).read().split("\n")
        order = int(lines[0].strip().replace("order", ""))
        nodes, weights = [], []
        for ln in lines[1:]:
            if ln.strip():
                a, b = ln.split()
                nodes.append(float(a))
                weights.append(float(b))
        rules[n] = {"order": order, "nodes": nodes, "weights": weights}
    return rules


def main():
    fx = {
        "source": "BoSSS Square.cs / Cube_polynomials.cs / LineQuadRules (reference snapshot 16a8e2e)",
        "square": parse_polys(os.path.join(REF, "Square.cs"), 2, 4),
        "cube": parse_polys(os.path.join(REF, "Cube_polynomials.cs"), 3, 4),
        "line_rules": parse_rules(12),
    }
    with open(os.path.join(HERE, "ref_tables.json"), "w") as f:
        json.dump(fx, f, indent=0)
    print(len(fx["square"]), len(fx["cube"]), len(fx["line_rules"]))


if __name__ == "__main__":
    main()
