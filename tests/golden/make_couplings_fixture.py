"""Write couplings_const.json: the reference's exchange-matrix table
couplings_const (Pyrochlore.cu:169-184) evaluated at a few (J1, J2, J3, J4).

Run once in a container that holds the reference sources; the fixture holds
numbers only. The table is a brace initialiser C[b][b'][9] (row-major 3x3 per
basis pair) whose entries are 0, J_k or -J_k; this script reads those 144
tokens from the #define and evaluates them, nothing else.

    python tests/golden/make_couplings_fixture.py /root/reference/Pyrochlore.cu
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))

J_SETS = [
    [1.0, 0.0, 0.0, 0.0], [0.0, 1.0, 0.0, 0.0], [0.0, 0.0, 1.0, 0.0], [0.0, 0.0, 0.0, 1.0],
    [1.0, 1.0, 0.0, 0.0],                      # Heisenberg antiferromagnet
    [-0.09 / 0.29, -0.22 / 0.29, -1.0, 0.0],   # the reference's defaults (Pyrochlore.cu:139-142)
    [0.37, -0.81, 0.53, -0.26],                # generic
]


def table_tokens(path):
    lines = open(path).read().splitlines()
    start = next(i for i, ln in enumerate(lines) if ln.lstrip().startswith("#define couplings_const"))
    text, i = "", start
    while True:
        ln = lines[i]
        text += ln.rstrip("\\") + " "
        if not ln.rstrip().endswith("\\"):
            break
        i += 1
    body = text.split("couplings_const", 1)[1]
    toks = re.findall(r"-?\s*J[1-4]|-?\d+(?:\.\d+)?", body)
    toks = [t.replace(" ", "") for t in toks]
    if len(toks) != 144:
        raise SystemExit(f"expected 144 entries, found {len(toks)}")
    return toks, (start + 1, i + 1)


def evaluate(tok, J):
    sign = -1.0 if tok.startswith("-") else 1.0
    t = tok.lstrip("-")
    return sign * J[int(t[1]) - 1] if t.startswith("J") else sign * float(t)


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/Pyrochlore.cu"
    toks, (a, b) = table_tokens(src)
    out = {
        "source": f"Pyrochlore.cu:{a}-{b} (couplings_const), evaluated by "
                  "tests/golden/make_couplings_fixture.py",
        "layout": "C[4 b + b'][3 i + j], 144 values per J set",
        "J_sets": J_SETS,
        "C": [[evaluate(t, J) for t in toks] for J in J_SETS],
    }
    with open(os.path.join(HERE, "couplings_const.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
